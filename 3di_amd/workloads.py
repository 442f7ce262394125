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


def cfg2_model(grid: int = 231, k_exp: int = 79, k_id: int = 199, k_tex: int = 199):
    key = (grid, k_exp, k_id, k_tex)
    if key not in _cfg1_cache:
        base = synth.make_model(grid, k_id, k_tex, k_exp, synth.BASE_SEED + 1)
        rng = np.random.default_rng(synth.BASE_SEED + 1)
        alpha = np.clip(rng.standard_normal(k_id), -3, 3)
        beta = np.clip(rng.standard_normal(k_tex), -3, 3)
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
