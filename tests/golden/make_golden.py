#!/usr/bin/env python3
"""Writes the golden fixtures of tests/golden/ (run from the repository root):

  raster_tiny.npz   brute-force rasteriser known answers: a few random triangles in a small
                    crop, keys (orderable depth << 32 | triangle) and barycentrics computed
                    here by numpy over every (pixel, triangle) pair with the FP32 operation
                    sequence of the coverage test (funcs.cu:940-979), with and without the
                    4-pixel per-triangle cap (funcs.cu:980), and with a crop offset.
  small_scene.npz   FP64 oracle outputs on a small seeded scene (render keys, blurred image,
                    objective, normal equations at the start point, parameters after 6
                    Gauss-Newton iterations) with a checksum of the generated model, so that
                    a change of the oracle or of the scene generator shows up as a diff.

blur_taps_reference.json is not generated: it holds the four 15-tap tables hard-coded in
the reference (video_fitter.cu:53-56), transcribed as data.

The brute-force rasteriser below is imported by tests/test_oracle.py as well.
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def orderable(z: np.ndarray) -> np.ndarray:
    u = np.asarray(z, np.float32).view(np.uint32).astype(np.uint64)
    neg = (u & 0x80000000) != 0
    return np.where(neg, (~u) & 0xFFFFFFFF, u | 0x80000000).astype(np.uint64)


def brute_raster(tris, xp, yp, z, x0, y0, W, H, ntmp=0):
    """Every pixel sample of the crop against every triangle, in float32.

    A triangle (b, c, a) = tris[t] covers integer sample (x, y) inside its bounding box
    [floor(min), ceil(max)) when alpha > 0, beta > 0 and alpha + beta <= 1, with alpha and
    beta the FP32 barycentrics of the b and c corners; the visible triangle of a pixel is
    the covering one with the smallest (depth of its first corner, triangle id).  With
    ntmp > 0 a triangle only claims its first ntmp covered samples in x-major order."""
    f32 = np.float32
    xp = np.asarray(xp, f32)
    yp = np.asarray(yp, f32)
    keys = np.full(W * H, np.iinfo(np.uint64).max, np.uint64)
    bary = np.zeros(2 * W * H, f32)
    X, Y = np.meshgrid(np.arange(x0, x0 + W), np.arange(y0, y0 + H), indexing="ij")   # x-major
    X = X.ravel()
    Y = Y.ravel()
    pix = (X - x0) * H + (Y - y0)
    cov = []
    with np.errstate(all="ignore"):
        for t, (b, c, a) in enumerate(np.asarray(tris)):
            bx, cx, ax = xp[b], xp[c], xp[a]
            by, cy, ay = yp[b], yp[c], yp[a]
            xmin = np.floor(min(bx, cx, ax))
            xmax = np.ceil(max(bx, cx, ax))
            ymin = np.floor(min(by, cy, ay))
            ymax = np.ceil(max(by, cy, ay))
            inbox = (X >= xmin) & (X < xmax) & (Y >= ymin) & (Y < ymax)
            v0x, v1x, v2x = bx - ax, cx - ax, X.astype(f32) - ax
            v0y, v1y, v2y = by - ay, cy - ay, Y.astype(f32) - ay
            d00 = v0x * v0x + v0y * v0y
            d01 = v0x * v1x + v0y * v1y
            d11 = v1x * v1x + v1y * v1y
            d20 = v0x * v2x + v0y * v2y
            d21 = v1x * v2x + v1y * v2y
            den = d00 * d11 - d01 * d01
            al = (d11 * d20 - d01 * d21) / den
            be = (d00 * d21 - d01 * d20) / den
            inside = inbox & (al > 0) & (be > 0) & (al + be <= 1)
            if ntmp > 0:
                inside &= np.cumsum(inside) <= ntmp
            key = (int(orderable(np.float32(z[b]))) << 32) | t
            cov.append((np.nonzero(inside)[0], np.uint64(key), al, be))
            sel = np.nonzero(inside)[0]
            keys[pix[sel]] = np.minimum(keys[pix[sel]], np.uint64(key))
    for sel, key, al, be in cov:
        win = sel[keys[pix[sel]] == key]
        bary[2 * pix[win]] = al[win]
        bary[2 * pix[win] + 1] = be[win]
    return keys, bary


def tiny_triangles(seed: int, n_tris: int, W: int, H: int, x0: int, y0: int):
    """Random triangles around a crop (some crossing its border), random depths with ties."""
    rng = np.random.default_rng(seed)
    V = 3 * n_tris
    xp = rng.uniform(x0 - 4, x0 + W + 4, V).astype(np.float32)
    yp = rng.uniform(y0 - 4, y0 + H + 4, V).astype(np.float32)
    # shrink every triangle towards its centroid so that sizes vary from ~1 px to ~W/2
    for t in range(n_tris):
        s = rng.uniform(0.05, 0.5)
        for arr in (xp, yp):
            c = arr[3 * t:3 * t + 3].mean()
            arr[3 * t:3 * t + 3] = (c + s * (arr[3 * t:3 * t + 3] - c)).astype(np.float32)
    z = rng.choice(np.array([400.0, 410.0, 420.0, -5.0], np.float32), V)   # repeated depths -> ties
    # a few vertices exactly on integer samples
    xp[::7] = np.round(xp[::7])
    yp[::5] = np.round(yp[::5])
    tris = np.arange(V, dtype=np.int32).reshape(n_tris, 3)
    return tris, xp, yp, z


RASTER_CASES = [  # (seed, n_tris, W, H, x0, y0, ntmp)
    (1, 40, 24, 20, 0, 0, 0),
    (2, 60, 31, 17, 5, 9, 0),
    (3, 60, 31, 17, 5, 9, 4),
    (4, 25, 16, 16, 100, 50, 4),
]


def make_raster():
    out = {}
    for i, (seed, nt, W, H, x0, y0, ntmp) in enumerate(RASTER_CASES):
        tris, xp, yp, z = tiny_triangles(seed, nt, W, H, x0, y0)
        keys, bary = brute_raster(tris, xp, yp, z, x0, y0, W, H, ntmp)
        out[f"c{i}_meta"] = np.array([seed, nt, W, H, x0, y0, ntmp], np.int64)
        out[f"c{i}_tris"] = tris
        out[f"c{i}_xp"] = xp
        out[f"c{i}_yp"] = yp
        out[f"c{i}_z"] = z
        out[f"c{i}_keys"] = keys
        out[f"c{i}_bary"] = bary
    np.savez_compressed(os.path.join(HERE, "raster_tiny.npz"), **out)


SMALL_SCENE = dict(grid=28, k_exp=6, k_id=4, k_tex=4, fit_id=True, fit_tex=True, fit_light=True,
                   crops=(64,), pose_deg=12.0, noise=0.02, iters=6, seed=5, cfg=21)


def model_checksum(m) -> str:
    h = hashlib.sha256()
    for a in (m.mean_shape, m.exp_basis, m.id_basis, m.tex_mean, m.tex_basis, m.triangles, m.landmark_ids):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def make_small_scene():
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from scenes import make_scene
    sc = make_scene(**SMALL_SCENE)
    orc = sc.oracle()
    th0 = sc.init()[0]
    r = orc.render(0, th0)
    ne = orc.normal_equations(0, th0)
    thf, reps = orc.fit(sc.init())
    fr = sc.frames[0]
    np.savez_compressed(
        os.path.join(HERE, "small_scene.npz"),
        model_sha256=np.array(model_checksum(sc.model)), theta0=th0, obs=fr.obs,
        keys=r["keys"], img=r["img"], obj=r["obj"], npix=np.array(r["npix"]),
        H=ne["H"], g=ne["g"], theta_fit=thf[0], stop=np.array(reps[0].stop), iters=np.array(reps[0].iters))


if __name__ == "__main__":
    make_raster()
    make_small_scene()
    print("wrote", sorted(f for f in os.listdir(HERE) if f.endswith(".npz")))
