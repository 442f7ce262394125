"""GPU parity of the frame preparation (tdi_prepare_crops) against the oracle: grey crops of
resized 8-bit BGR frames, bit for bit.  Both sides perform the same FP32 operations in the same
order (the device code is built without multiply-add contraction for this kernel), so every
crop pixel must be identical.  Cases: random coefficients from 0.3 to 3 (down- and upscaling),
the copy case (destination size equal to the source size), coefficient 1, the exact 2x downscale, crops that reach past
the resized frame into the zero border, one-pixel crops, one-pixel-wide and one-pixel-high
sources, a crop entirely past the resized frame, random sources of 1-97 pixels a side at scales
0.2-6, and the host-layer chain (landmark geometry, then crops) on synthetic frames.
"""
import importlib
import sys

import numpy as np
import pytest

import pyoracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tdi = importlib.import_module("3di_amd.tdi")
video = importlib.import_module("3di_amd.video")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case(rng, n, sh, sw):
    frames = rng.integers(0, 256, (n, sh, sw, 3), dtype=np.uint8)
    coef = rng.uniform(0.3, 3.0, n).astype(np.float32)
    coef[0], coef[1] = 1.0, np.float32(1.0 + 0.2 / sw)      # coefficient 1, and an equal-size copy
    coef[3] = 0.5                                           # the exact 2x downscale (2x2 area average)
    crop = np.zeros((n, 4), np.int32)
    for i in range(n):
        dw, dh = O.resized_size(sw, coef[i]), O.resized_size(sh, coef[i])
        crop[i, 0], crop[i, 1] = rng.integers(0, max(dw - 8, 1)), rng.integers(0, max(dh - 8, 1))
        crop[i, 2], crop[i, 3] = rng.integers(1, 400), rng.integers(1, 400)
    crop[2, 2:] = 1
    return frames, coef, crop


@pytest.mark.parametrize("seed,n,sh,sw", [(1, 24, 120, 160), (2, 16, 481, 643)])
def test_prepare_crops_bitexact(seed, n, sh, sw):
    rng = np.random.default_rng(seed)
    frames, coef, crop = _case(rng, n, sh, sw)
    want = O.prepare_crops(frames, coef, crop)
    _say(f"case: prepare_crops seed {seed}, {n} frames {sw}x{sh}, {want.size} crop pixels")
    got = tdi.prepare_crops(torch.from_numpy(frames).cuda(), coef, crop)
    torch.cuda.synchronize()
    got = got.cpu().numpy()
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, f"{bad.size} pixels differ, first at {bad[0]}: {got[bad[0]]} vs {want[bad[0]]}"


def test_prepare_crops_edge_cases():
    # one-pixel-wide and one-pixel-high sources, a crop entirely past the resized frame, and a
    # large upscale
    rng = np.random.default_rng(9)
    for sh, sw in ((1, 37), (29, 1), (1, 1)):
        frames = rng.integers(0, 256, (3, sh, sw, 3), dtype=np.uint8)
        coef = np.array([2.5, 7.0, 1.0], np.float32)
        crop = np.array([[0, 0, 20, 9], [3, 2, 40, 30], [sw + 2, sh + 2, 5, 4]], np.int32)
        want = O.prepare_crops(frames, coef, crop)
        _say(f"case: prepare_crops edge source {sw}x{sh}")
        got = tdi.prepare_crops(torch.from_numpy(frames).cuda(), coef, crop)
        torch.cuda.synchronize()
        got = got.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
        assert (got[-20:] == 0).all()


def test_prepare_crops_rejects_bad_arguments():
    frames = torch.zeros((2, 8, 8, 3), dtype=torch.uint8, device="cuda")
    with pytest.raises(tdi.TdiError):
        tdi.prepare_crops(frames, [1.0, 0.0], [[0, 0, 4, 4], [0, 0, 4, 4]])      # coefficient 0
    with pytest.raises(tdi.TdiError):
        tdi.prepare_crops(frames, [1.0, 1.0], [[0, 0, 4, 4], [-1, 0, 4, 4]])     # negative origin
    with pytest.raises(tdi.TdiError):
        tdi.prepare_crops(frames, [1.0, 1e-3], [[0, 0, 4, 4], [0, 0, 4, 4]])     # empty resized frame


def test_host_layer_prepare_frames():
    rng = np.random.default_rng(5)
    n, sh, sw = 12, 360, 480
    frames = rng.integers(0, 256, (n, sh, sw, 3), dtype=np.uint8)
    cx, cy = rng.uniform(150, 330, n), rng.uniform(120, 240, n)
    size = rng.uniform(60, 140, n)
    base = rng.uniform(-0.5, 0.5, (51, 2))
    lx = (cx[:, None] + size[:, None] * base[None, :, 0]).astype(np.float32)
    ly = (cy[:, None] + size[:, None] * base[None, :, 1]).astype(np.float32)
    lx[3, 0] = -5.0                                          # one frame with a landmark outside
    cams = np.tile(np.array([[900, 900, sw / 2, sh / 2]], np.float32), (n, 1))
    _say(f"case: host-layer prepare_frames, {n} frames {sw}x{sh}")
    g = video.prepare_frames(torch.from_numpy(frames).cuda(), lx, ly, cams, 100.0, pad=True, crop_size=(192, 160))
    torch.cuda.synchronize()
    keep = g["frames"]
    assert list(np.flatnonzero(g["status"])) == [3] and keep.size == n - 1
    want = O.prepare_crops(frames[keep], g["coef"][keep], g["crop"][keep])
    assert np.array_equal(g["obs"].cpu().numpy().view(np.uint32), want.view(np.uint32))


def test_batch_accepts_device_observations():
    """tdi_batch_create reads tdi_frame_desc.obs from device memory as well as from host memory
    (the crops of tdi_prepare_crops need not leave the GPU): both batches give the same bits."""
    import ctypes as C
    from scenes import make_scene
    abi = importlib.import_module("3di_amd.abi")
    sc = make_scene(grid=48, k_exp=12, crops=(96, 128, 80), seed=11)
    _say("case: batch from device observations, 3 frames")
    ctx = tdi.Context(sc.model, sc.opts)
    try:
        host = tdi.Batch(ctx, sc.frames)
        dev_obs = [torch.from_numpy(np.ascontiguousarray(f.obs, np.float32).reshape(-1)).cuda() for f in sc.frames]
        fd = abi.FrameDescs(sc.frames)
        for i, t in enumerate(dev_obs):
            fd.arr[i].obs = C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_float))
        h = C.c_void_p()
        rc = tdi.lib().tdi_batch_create(ctx.h, len(sc.frames), C.addressof(fd.arr), C.byref(h))
        assert rc == 0, tdi.lib().tdi_last_error()
        dev = object.__new__(tdi.Batch)
        dev.ctx, dev.frames, dev.fd, dev.h = ctx, list(sc.frames), fd, h
        dev.areas, dev.offsets = host.areas, host.offsets
        th = sc.init().astype(np.float32)
        _, obj_h, npix_h = host.render(th)
        _, obj_d, npix_d = dev.render(th)
        torch.cuda.synchronize()
        assert np.array_equal(obj_h.cpu().numpy(), obj_d.cpu().numpy())
        assert np.array_equal(npix_h.cpu().numpy(), npix_d.cpu().numpy())
        H_h, g_h, _ = host.normal_equations(th)[:3]
        H_d, g_d, _ = dev.normal_equations(th)[:3]
        torch.cuda.synchronize()
        assert np.array_equal(g_h.cpu().numpy(), g_d.cpu().numpy())
        dev.close()
        host.close()
    finally:
        ctx.close()


@pytest.mark.parametrize("seed", range(4))
def test_prepare_crops_random_sources_bitexact(seed):
    # sources from 1 to 97 pixels a side, scales 0.2-6, crops anywhere (also past the frame)
    rng = np.random.default_rng(100 + seed)
    sh, sw = (int(v) for v in rng.integers(1, 98, 2))
    n = 8
    frames = rng.integers(0, 256, (n, sh, sw, 3), dtype=np.uint8)
    coef = np.exp(rng.uniform(np.log(0.2), np.log(6.0), n)).astype(np.float32)
    coef = np.where([O.resized_size(sw, c) >= 1 and O.resized_size(sh, c) >= 1 for c in coef], coef,
                    np.float32(1.0)).astype(np.float32)
    crop = np.stack([rng.integers(0, 3 * sw, n), rng.integers(0, 3 * sh, n), rng.integers(1, 90, n),
                     rng.integers(1, 90, n)], 1).astype(np.int32)
    want = O.prepare_crops(frames, coef, crop)
    _say(f"case: prepare_crops random seed {seed}, {n} frames {sw}x{sh}")
    got = tdi.prepare_crops(torch.from_numpy(frames).cuda(), coef, crop)
    torch.cuda.synchronize()
    assert np.array_equal(got.cpu().numpy().view(np.uint32), want.view(np.uint32))


def _mouths(rng, n, w, h):
    """Random faces' landmarks: a mouth polygon around a random centre (some near or past the frame
    edges), random scale, lip noise and opening (some closed), landmarks 0 and 1 setting the face
    size."""
    base = np.array([(-20, 0), (-10, -10), (0, -10), (10, -10), (20, 0), (10, 10), (0, 10), (-10, 10)], np.float64)
    X = np.zeros((n, 51), np.float32)
    Y = np.zeros((n, 51), np.float32)
    for f in range(n):
        s = rng.uniform(0.3, 3.0)
        c = np.array([rng.uniform(-20, w + 20), rng.uniform(-20, h + 20)])
        P = base * s
        P[5:8, 1] *= rng.uniform(0.0, 2.0)        # opening of the lower lip
        P += rng.normal(0, 0.8 * s, P.shape)
        x = np.full(51, c[0]) + rng.normal(0, 5 * s, 51)
        y = np.full(51, c[1]) + rng.normal(0, 5 * s, 51)
        x[43:51], y[43:51] = P[:, 0] + c[0], P[:, 1] + c[1]
        x[0], x[1] = c[0] - 43 * s * rng.uniform(1, 1.3), c[0] + 43 * s * rng.uniform(1, 1.3)
        X[f], Y[f] = x, y
    X[3, 40] = np.nan
    return X, Y


@pytest.mark.parametrize("seed,n,h,w", [(3, 48, 240, 320), (4, 12, 481, 643)])
def test_paint_innermouth_black_bitexact(seed, n, h, w):
    rng = np.random.default_rng(seed)
    frames = rng.integers(1, 256, (n, h, w, 3), dtype=np.uint8)
    X, Y = _mouths(rng, n, w, h)
    want = O.paint_innermouth_black(frames, X, Y, 86.0)
    painted = int((want[..., 0] == 0).sum())
    _say(f"case: paint_innermouth_black seed {seed}, {n} frames {w}x{h}, {painted} pixels blackened")
    assert painted > 100
    dev = torch.from_numpy(frames).cuda()
    tdi.paint_innermouth_black(dev, X, Y, 86.0)
    torch.cuda.synchronize()
    got = dev.cpu().numpy()
    bad = np.argwhere(got != want)
    assert bad.size == 0, f"{len(bad)} bytes differ, first {bad[:5].tolist()}"


def test_paint_innermouth_black_rejects_bad_arguments():
    dev = torch.zeros((2, 16, 16, 3), dtype=torch.uint8, device="cuda")
    X = np.zeros((2, 51), np.float32)
    with pytest.raises(Exception):
        tdi.paint_innermouth_black(dev, X, X, 0.0)
    with pytest.raises(ValueError):
        tdi.paint_innermouth_black(dev, X[:1], X[:1], 86.0)
