"""Synthetic benchmark workloads, rendered by the HIP library itself.

The batched per-frame stage: identity and albedo drawn once on the 231x231-grid
model (53,361 vertices, 105,800 triangles, K_id = K_tex = 199, K_exp = 79) and folded into
the means; every frame has its own crop side S ~ U{96..384}, focal length 800*S/192, pose
within +-30 deg, tz in 400..600 mm, expression ~ N(0,1) clipped to +-3, and its own
light, fitted jointly: the point light (lambda, L) and the five second-order terms
(6 pose + 79 expression + 9 light = 94 unknowns).
Observations are the GPU render of the truth plus N(0, 0.02) noise; starting points are
perturbed truths made landmark-feasible.
"""
from __future__ import annotations

import dataclasses
import importlib
import time

import numpy as np

synth = importlib.import_module("3di_amd.synth")
abi = importlib.import_module("3di_amd.abi")
tdi = importlib.import_module("3di_amd.tdi")


@dataclasses.dataclass
class Workload:
    model: object
    frames: list
    truths: list
    opts: object
    lay: dict
    name: str

    def init(self):
        return np.stack([t.init for t in self.truths]).astype(np.float32)

    def truth(self):
        return np.stack([t.theta for t in self.truths])


_cfg1_cache = {}


_base_cache = {}


def base_model(grid: int = 231, k_exp: int = 79, k_id: int = 199, k_tex: int = 199):
    """The model with its identity and albedo bases (generated once per process)."""
    key = (grid, k_exp, k_id, k_tex)
    if key not in _base_cache:
        _base_cache[key] = synth.make_model(grid, k_id, k_tex, k_exp, synth.BASE_SEED + 1)
    return _base_cache[key]


def folded_identity(k_id: int = 199, k_tex: int = 199):
    """The identity (alpha) and albedo (beta) coefficients folded into the per-frame stage's model."""
    rng = np.random.default_rng(synth.BASE_SEED + 1)
    alpha = np.clip(rng.standard_normal(k_id), -3, 3)
    beta = np.clip(rng.standard_normal(k_tex), -3, 3)
    return alpha, beta


def cfg2_model(grid: int = 231, k_exp: int = 79, k_id: int = 199, k_tex: int = 199):
    key = (grid, k_exp, k_id, k_tex)
    if key not in _cfg1_cache:
        base = base_model(grid, k_exp, k_id, k_tex)
        alpha, beta = folded_identity(k_id, k_tex)
        _cfg1_cache[key] = synth.fold_identity(base, alpha, beta)
    return _cfg1_cache[key]


def cfg2(n_frames: int = 4096, iters: int = 12, grid: int = 231, k_exp: int = 79, fit_light: int = 2,
         noise: float = 0.02, log=print, crop: int = 0, config: int = 2) -> Workload:
    """crop > 0: every crop crop x crop (the differentiation configuration: 256); config: the seed
    stream of the frames."""
    t0 = time.time()
    m = cfg2_model(grid, k_exp) if grid == 231 and k_exp == 79 else cfg2_model(grid, k_exp, 0, 0)
    log(f"[workload] model {m.n_vertices} vertices, {m.n_triangles} triangles ({time.time() - t0:.1f}s)")
    lay = synth.layout(m, False, False, True, fit_light)
    opts = abi.default_options(max_iters=iters, fit_light=int(fit_light))
    crops = synth.cfg2_crop_sizes(n_frames) if crop <= 0 else np.full(n_frames, crop)
    frames, truths = [], []
    for i in range(n_frames):
        rng = np.random.default_rng(synth.frame_seed(config, i))
        S = int(crops[i])
        tr = synth.draw_truth(m, rng, lay, 30.0, fit_light=fit_light)
        fr = synth.make_frame(m, tr, lay, rng, synth.cfg2_focal(S), S, S)
        frames.append(fr)
        truths.append(tr)
    log(f"[workload] {n_frames} frame descriptors ({time.time() - t0:.1f}s)")
    # observations: GPU render of the truth + noise
    import torch
    ctx = tdi.Context(m, opts)
    b = tdi.Batch(ctx, frames)
    th = torch.from_numpy(np.stack([t.theta for t in truths]).astype(np.float32)).cuda()
    img, _, npix = b.render(th)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(synth.BASE_SEED + 2)
    img = img + noise * torch.randn(img.shape, device="cuda", generator=gen)
    host = img.cpu().numpy()
    b.close()
    ctx.close()
    off = 0
    for fr in frames:
        A = fr.width * fr.height
        fr.obs = np.ascontiguousarray(host[off:off + A])
        off += A
    for fr, tr in zip(frames, truths):
        synth.make_feasible(m, tr, lay, fr)
    log(f"[workload] observations rendered, mean {float(npix.float().mean()):.0f} face pixels/frame "
        f"({time.time() - t0:.1f}s)")
    return Workload(model=m, frames=frames, truths=truths, opts=opts, lay=lay,
                    name=f"{'batched per-frame fit' if crop <= 0 else 'render loss and adjoint'}: {n_frames} frames, {grid}x{grid} grid ({m.n_vertices} V, {m.n_triangles} T), "
                         f"K={lay['K']} (6 pose + {k_exp} expression + {synth.light_unknowns(fit_light)} light), "
                         + (f"crops U{{96..384}}, {iters} GN iterations" if crop <= 0 else f"{crop}x{crop} crops"))


@dataclasses.dataclass
class LongVideo:
    """A synthetic long video: per-frame truths, landmarks and cameras on the host; the frames
    (8-bit BGR, width x height) rendered on the GPU, handed out chunk by chunk by `source`."""
    model: object
    opts: object
    lay: dict
    theta: np.ndarray          # [n, K] truths
    lmk_x: np.ndarray          # [n, 51] image pixels (1 px noise)
    lmk_y: np.ndarray
    cams: np.ndarray           # [n, 4]
    width: int
    height: int
    render_ms: list            # time spent rendering frames, per rendered chunk
    call_times: list = dataclasses.field(default_factory=list)   # host clock at each source call
    source: object = None

    @property
    def n(self):
        return self.theta.shape[0]


def long_video(n_frames: int = 2048, width: int = 1920, height: int = 1080, face_px=(280.0, 420.0),
               iters: int = 12, crop: int = 896, noise: float = 0.02, seed: int = 0x3D1 + 4,
               resident: bool = True) -> LongVideo:
    """A face video: the pose follows a random walk (0.5 deg and 2 mm per frame per axis, angles
    within +-30 deg), the expression an AR(1) process (rho = 0.95), and the steps are redrawn until
    the 51 landmarks stay 20 px inside the frame with an extent within face_px; the light is fixed
    at the reference's initial values.  The fit's unknowns: 6 pose + 79 expression (the per-frame
    stage fits the light in its own sub-fits).  The frames are the render of the truth (crop x crop
    around the face) over a grey N(0.5, noise) background, + N(0, noise), as 8-bit BGR.
    resident: every frame rendered into one device buffer here (n x height x width x 3 bytes), the
    source handing out views of it; else the source renders each chunk when it is asked for it."""
    import torch
    m = cfg2_model(231, 79)
    lay = synth.layout(m, False, False, True, 0)
    opts = abi.default_options(max_iters=iters, fit_light=0)
    K = lay["K"]
    rng = np.random.default_rng(seed)
    ids = m.landmark_ids.astype(np.int64)
    f = 1800.0
    cx, cy = width / 2.0, height / 2.0
    ne = lay["pose"] - lay["exp"]
    ub = m.exp_ub[:ne].astype(np.float64)
    margin = 20.0

    P0 = m.mean_shape[ids].astype(np.float64)
    E = m.exp_basis.reshape(-1, 3, m.exp_basis.shape[1])[ids].astype(np.float64)      # [51, 3, K_exp]
    po = lay["pose"]

    def landmarks(th):
        P = P0 + E @ th[lay["exp"]:po]
        u = th[po + 3:po + 6]
        t = np.linalg.norm(u)
        n = u / t
        N = np.array([[0, -n[2], n[1]], [n[2], 0, -n[0]], [-n[1], n[0], 0]])
        q = P @ (np.eye(3) + np.sin(t) * N + (1 - np.cos(t)) * N @ N).T + th[po:po + 3]
        return np.stack([f * q[:, 0] / q[:, 2] + cx, f * q[:, 1] / q[:, 2] + cy], 1)

    def inside(lm):
        ext = max(np.ptp(lm[:, 0]), np.ptp(lm[:, 1]))
        return (face_px[0] <= ext <= face_px[1] and lm[:, 0].min() > margin and lm[:, 1].min() > margin
                and lm[:, 0].max() < width - margin and lm[:, 1].max() < height - margin)

    theta = np.zeros((n_frames, K))
    e = np.clip(rng.standard_normal(ne), -3, 3)
    ang = np.zeros(3)
    th = np.zeros(K)
    th[lay["exp"]:lay["pose"]] = e * (ub / 3.5)
    th[lay["pose"] + 3:lay["pose"] + 6] = synth.rotvec_from_euler(*ang)
    # the depth that puts the first face in the middle of face_px
    th[lay["pose"] + 2] = 1000.0
    ext = max(np.ptp(landmarks(th), axis=0))
    th[lay["pose"] + 2] = 1000.0 * ext / (0.5 * (face_px[0] + face_px[1]))
    rho = 0.95
    for i in range(n_frames):
        if i:
            # an expression and pose step that keeps the face inside the frame at the wanted
            # size (redrawn otherwise; the frame repeats the previous one if no draw fits)
            for _ in range(32):
                e2 = rho * e + np.sqrt(1 - rho * rho) * rng.standard_normal(ne)
                a2 = np.clip(ang + np.deg2rad(rng.normal(0, 0.5, 3)), -np.deg2rad(30), np.deg2rad(30))
                cand = th.copy()
                cand[lay["exp"]:po] = np.clip(e2, -3, 3) * (ub / 3.5)
                cand[po:po + 3] += rng.normal(0, 2.0, 3)
                cand[po + 3:po + 6] = synth.rotvec_from_euler(*a2)
                if inside(landmarks(cand)):
                    th, ang, e = cand, a2, e2
                    break
        theta[i] = th
    lx = np.zeros((n_frames, 51), np.float32)
    ly = np.zeros((n_frames, 51), np.float32)
    for i in range(n_frames):
        lm = landmarks(theta[i]) + rng.normal(0, 1.0, (51, 2))
        lx[i], ly[i] = lm[:, 0], lm[:, 1]
    cams = np.tile(np.array([[f, f, cx, cy]], np.float32), (n_frames, 1))
    lv = LongVideo(model=m, opts=opts, lay=lay, theta=theta, lmk_x=lx, lmk_y=ly, cams=cams, width=width,
                   height=height, render_ms=[])
    state = {}
    zero_obs = np.zeros(crop * crop, np.float32)      # the renderer ignores the observation
    frng = np.random.default_rng(seed + 1)

    def render(first: int, count: int):
        if "ctx" not in state:
            state["ctx"] = tdi.Context(m, opts)
        frames = []
        for i in range(first, first + count):
            # a crop x crop window around the face, clipped to the frame
            mx, my = float(lx[i].mean()), float(ly[i].mean())
            x0 = int(np.clip(round(mx - crop / 2), 0, width - crop))
            y0 = int(np.clip(round(my - crop / 2), 0, height - crop))
            fr = synth.make_frame(m, synth.Truth(theta=theta[i], init=theta[i], lam=synth.LAMBDA0,
                                                 Lint=synth.L0), lay, frng, f, crop, crop)
            fr.cx, fr.cy, fr.x0, fr.y0 = cx, cy, x0, y0
            fr.lmk_x, fr.lmk_y, fr.obs = lx[i], ly[i], zero_obs
            frames.append(fr)
        b = tdi.Batch(state["ctx"], frames)
        th = torch.from_numpy(theta[first:first + count].astype(np.float32)).cuda()
        img, _, _ = b.render(th)
        b.close()
        gen = torch.Generator(device="cuda")
        gen.manual_seed(seed + first)
        grey = (0.5 + noise * torch.randn((count, height, width), device="cuda", generator=gen))
        face = img.view(count, crop, crop).transpose(1, 2)        # [count, y, x]
        noisy = face + noise * torch.randn(face.shape, device="cuda", generator=gen)
        for j, fr in enumerate(frames):
            win = grey[j, fr.y0:fr.y0 + crop, fr.x0:fr.x0 + crop]
            win.copy_(torch.where(face[j] != 0, noisy[j], win))
        return (grey.clamp(0, 1) * 255.0 + 0.5).to(torch.uint8)

    if resident:
        allf = torch.empty((n_frames, height, width, 3), dtype=torch.uint8, device="cuda")
        step = 256
        for first in range(0, n_frames, step):
            t0 = time.perf_counter()
            count = min(step, n_frames - first)
            allf[first:first + count].copy_(render(first, count).unsqueeze(-1).expand(count, height, width, 3))
            torch.cuda.synchronize()
            lv.render_ms.append(1e3 * (time.perf_counter() - t0))
        state.pop("ctx").close()

        def source(first: int, count: int):
            lv.call_times.append(time.perf_counter())
            return allf[first:first + count]
    else:
        def source(first: int, count: int):
            t0 = time.perf_counter()
            lv.call_times.append(t0)
            buf = state.get("buf")
            if buf is None or buf.shape[0] < count:
                buf = state["buf"] = torch.empty((count, height, width, 3), dtype=torch.uint8, device="cuda")
            out = buf[:count]
            out.copy_(render(first, count).unsqueeze(-1).expand(count, height, width, 3))
            torch.cuda.synchronize()
            lv.render_ms.append(1e3 * (time.perf_counter() - t0))
            return out

    lv.source = source
    return lv


def identity_stage(n_frames: int = 10, S: int = 512, iters: int = 30, noise: float = 0.02, coef_pert: float = 0.05,
                   log=print) -> Workload:
    """The identity stage at reference scale: the 231x231-grid model with its bases (K_id = K_tex =
    199, K_exp = 79), n_frames frames of S x S sharing one albedo and identity draw, each with its
    own pose (+-30 deg), expression and light (nine light unknowns): 398 + n_frames * 94 joint
    unknowns.  Observations are the GPU render of the truth plus N(0, noise); the start shares one
    perturbed albedo/identity block and is pulled towards the truth until every frame's
    landmark-box constraints hold."""
    import torch
    t0 = time.time()
    m = base_model(231, 79, 199, 199)
    lay = synth.layout(m, True, True, True, 2)
    opts = abi.default_options(max_iters=iters, fit_light=2, fit_id=1, fit_tex=1)
    Ks = lay["exp"]
    rng0 = np.random.default_rng(synth.frame_seed(1, 100000))
    shared = np.clip(rng0.standard_normal(Ks), -3, 3)
    shared_init = np.clip(shared + rng0.normal(0, coef_pert, Ks), -3.2, 3.2)
    frames, truths = [], []
    for i in range(n_frames):
        rng = np.random.default_rng(synth.frame_seed(1, i))
        tr = synth.draw_truth(m, rng, lay, 30.0, fit_light=2, coef_pert=coef_pert)
        tr.theta[:Ks] = shared
        tr.init[:Ks] = shared_init
        fr = synth.make_frame(m, tr, lay, rng, synth.cfg2_focal(S), S, S)
        frames.append(fr)
        truths.append(tr)
    ctx = tdi.Context(m, opts)
    b = tdi.Batch(ctx, frames)
    img, _, npix = b.render(torch.from_numpy(np.stack([t.theta for t in truths]).astype(np.float32)).cuda())
    gen = torch.Generator(device="cuda")
    gen.manual_seed(synth.BASE_SEED + 1)
    host = (img + noise * torch.randn(img.shape, device="cuda", generator=gen)).cpu().numpy()
    b.close()
    ctx.close()
    off = 0
    for fr in frames:
        fr.obs = np.ascontiguousarray(host[off:off + S * S])
        off += S * S
    for _ in range(30):
        if all(synth.landmark_slack(m, tr.init, lay, fr) > 0.5 for tr, fr in zip(truths, frames)):
            break
        for tr in truths:
            tr.init = tr.theta + 0.7 * (tr.init - tr.theta)
    log(f"[workload] identity stage: {n_frames} frames {S}x{S}, {float(npix.float().mean()):.0f} face pixels/frame "
        f"({time.time() - t0:.1f}s)")
    return Workload(model=m, frames=frames, truths=truths, opts=opts, lay=lay,
                    name=f"identity stage: {n_frames} frames {S}x{S} sharing albedo and identity, 231x231 grid "
                         f"({m.n_vertices} V, {m.n_triangles} T), K_id = K_tex = 199, K_exp = 79, nine light unknowns, "
                         f"{Ks + n_frames * (lay['K'] - Ks)} joint unknowns, {iters} GN iterations")
