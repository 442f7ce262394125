"""The render loss and its adjoint on the GPU (tdi_render_loss_grad, 3di_amd/autograd.py), against
the FP64 oracle:
  - the loss agrees with the oracle's (FP32 render against FP64) to 1e-4 relative;
  - directional derivatives g . d of the adjoint agree with FP64 central differences of the
    oracle's loss with the visibility held fixed (the GPU render's triangle of each pixel), along
    seeded random directions (over all of pose, expression, light and per-vertex offsets, or over
    one of these blocks at a time, so that each block is checked on its own), to 1e-3 of
    sum |g_k d_k|: at the small scene's sizes and at the differentiation configuration's (the
    231x231-grid model, 79 expressions, nine light unknowns, 256x256 crops; 64 directions);
  - against the reference's analytic J^T r (the fit's data gradient, tdi_normal_equations minus the
    barrier's; newfuncs.cu:173-463): the adjoint's data part is (2/w) times the exact derivative of
    the data objective, the reference's gradient an approximation of it by design: its normalised-
    gradient Jacobian damps every pixel by h^3 / (0.001 + h^3) (c = 1/(0.001 + h^3), h = 2|g|, so
    a pixel with |g| = 0.05 counts half), takes neighbour differences of the unblurred dI at
    compacted neighbour positions, and uses the image-gradient (optical-flow) form for the
    geometric columns.  The blocks are compared (cosine, relative norm, printed); the expression
    block, 79 of the 94 unknowns, must point the same way (cosine > 0.9), the pose block broadly
    (> 0.5); the light block is reported only (the damping and the missing blur dominate there:
    its columns are smooth over the face, so their neighbour differences are small);
  - the torch.autograd.Function returns the library's gradients scaled by the upstream weights.
"""
import importlib
import sys

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from scenes import make_scene

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tdi = importlib.import_module("3di_amd.tdi")
synth = importlib.import_module("3di_amd.synth")
abi = importlib.import_module("3di_amd.abi")
workloads = importlib.import_module("3di_amd.workloads")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _small():
    return make_scene(grid=48, k_exp=12, crops=(128, 160, 200), pose_deg=15.0, fit_light=2)


def _cfg3(n=2, S=256):
    """Frames of the differentiation configuration: the 231x231-grid model (identity folded in),
    79 expressions, nine light unknowns, S x S crops; observations rendered by the oracle."""
    import pyoracle as O
    from scenes import Scene
    m = workloads.cfg2_model(231, 79)
    lay = synth.layout(m, False, False, True, 2)
    opts = abi.default_options(max_iters=12, fit_light=2)
    frames, truths = [], []
    for i in range(n):
        rng = np.random.default_rng(synth.frame_seed(3, i))
        tr = synth.draw_truth(m, rng, lay, 30.0, fit_light=2)
        fr = synth.make_frame(m, tr, lay, rng, synth.cfg2_focal(S), S, S)
        raw = O.Scene(m, [fr], opts).render(0, tr.theta)["raw"].astype(np.float32)
        fr.obs = synth.add_noise(raw, rng, 0.02)
        frames.append(fr)
        truths.append(tr)
    return Scene(model=m, frames=frames, truths=truths, opts=opts, lay=lay)


BLOCKS = ("all", "pose", "expression", "light", "vertices")


def _direction(sc, rng, K, which):
    """A seeded random direction over every block, or over one block only (each block's gradient
    is then checked on its own)."""
    lay = sc.lay
    d = np.zeros(K)
    po = lay["pose"]
    if which in ("all", "pose"):
        d[po:po + 3] = rng.normal(0, 1.0, 3)                  # mm
        d[po + 3:po + 6] = rng.normal(0, 0.01, 3)             # rad
    if which in ("all", "expression"):
        d[lay["exp"]:po] = rng.normal(0, 0.1, po - lay["exp"])
    if which in ("all", "light"):
        d[lay["light"]:K] = rng.normal(0, 0.01, K - lay["light"])
    dv = np.zeros((sc.model.n_vertices, 3))
    if which in ("all", "vertices"):
        dv = rng.normal(0, 0.05, (sc.model.n_vertices, 3))    # mm
    return d, dv


def _check_directions(sc, th, n_dirs, seed):
    ors = sc.oracle()
    b = tdi.Batch(tdi.Context(sc.model, sc.opts), sc.frames)
    n, K, V = len(sc.frames), sc.K, sc.model.n_vertices
    loss, gth, gv = b.render_loss(torch.from_numpy(th.astype(np.float32)).cuda())
    torch.cuda.synchronize()
    loss, gth, gv = loss.cpu().numpy(), gth.cpu().numpy().astype(np.float64), gv.cpu().numpy().astype(np.float64)
    rng = np.random.default_rng(seed)
    worst = 0.0
    for i in range(n):
        fr = sc.frames[i]
        thi = th[i].astype(np.float32).astype(np.float64)
        tri = np.zeros(fr.width * fr.height, np.int32)
        want = ors.render_loss(i, thi, tri=tri)
        el = abs(loss[i] - want) / want
        _say(f"  frame {i}: loss {loss[i]:.6f} vs oracle {want:.6f} ({el:.1e}), {(tri >= 0).sum()} visible pixels")
        assert el < 1e-4
        for j in range(n_dirs):
            d, dv = _direction(sc, rng, K, BLOCKS[j % len(BLOCKS)])
            pred = float(gth[i] @ d + (gv[i] * dv).sum())
            scale = float(np.abs(gth[i] * d).sum() + np.abs(gv[i] * dv).sum())
            h = 1e-5
            fd = (ors.render_loss(i, thi + h * d, dverts=h * dv, tri=tri, frozen=True)
                  - ors.render_loss(i, thi - h * d, dverts=-h * dv, tri=tri, frozen=True)) / (2 * h)
            err = abs(pred - fd) / scale
            worst = max(worst, err)
            assert err < 1e-3, (i, j, BLOCKS[j % len(BLOCKS)], pred, fd, scale)
    _say(f"  worst directional error {worst:.1e} of sum |g_k d_k|")
    return b, gth


def test_loss_and_directions_small():
    _say("case: adjoint, small scene (48x48 grid, 12 expressions, nine light unknowns)")
    sc = _small()
    _check_directions(sc, sc.init(), 12, 31)


def test_directions_differentiation_config():
    _say("case: adjoint at the differentiation configuration's sizes (231x231 grid, 256x256 crops), 64 directions")
    sc = _cfg3(n=2)
    _check_directions(sc, sc.init(), 32, 32)


def _landmark_grad(sc, th):
    """Gradient of the landmark term alone (numpy central differences of the numpy term)."""
    m = sc.model
    po, e0 = sc.lay["pose"], sc.lay["exp"]
    ids = m.landmark_ids
    E = m.exp_basis.reshape(-1, 3, m.exp_basis.shape[1])[ids].astype(np.float64)

    def term(t, fr):
        P0 = m.mean_shape[ids].astype(np.float64) + E @ t[e0:po]
        C = P0 @ Rotation.from_rotvec(t[po + 3:po + 6]).as_matrix().T + t[po:po + 3]
        xp = float(np.float32(fr.phix)) * C[:, 0] / C[:, 2] + float(np.float32(fr.cx))
        yp = float(np.float32(fr.phiy)) * C[:, 1] / C[:, 2] + float(np.float32(fr.cy))
        return ((xp - fr.lmk_x) ** 2 + (yp - fr.lmk_y) ** 2).sum()

    out = np.zeros_like(th)
    for i, fr in enumerate(sc.frames):
        for k in range(sc.K):
            hk = 1e-6 * max(1.0, abs(th[i, k]))
            tp, tm = th[i].copy(), th[i].copy()
            tp[k] += hk
            tm[k] -= hk
            out[i, k] = (term(tp, fr) - term(tm, fr)) / (2 * hk)
    return out


def test_against_the_analytic_jacobian():
    _say("case: adjoint data gradient vs the reference's analytic J^T r")
    sc = _small()
    th = sc.init().astype(np.float32)
    b = tdi.Batch(tdi.Context(sc.model, sc.opts), sc.frames)
    dth = torch.from_numpy(th).cuda()
    _, gth, _ = b.render_loss(dth)
    _, g_ne, _ = b.normal_equations(dth)
    _, g_bar, _ = b.barrier_terms(dth)
    torch.cuda.synchronize()
    w = sc.opts.data_weight
    adj = gth.cpu().numpy().astype(np.float64) - _landmark_grad(sc, th.astype(np.float64))
    ref = (2.0 / w) * (g_ne.cpu().numpy().astype(np.float64) - g_bar.cpu().numpy())
    lay = sc.lay
    blocks = {"expression": slice(lay["exp"], lay["pose"]), "pose": slice(lay["pose"], lay["pose"] + 6),
              "light": slice(lay["light"], sc.K)}
    for name, sl in blocks.items():
        a, r = adj[:, sl].ravel(), ref[:, sl].ravel()
        cos = float(a @ r / (np.linalg.norm(a) * np.linalg.norm(r)))
        rel = float(np.linalg.norm(a - r) / np.linalg.norm(r))
        _say(f"  {name:10s}: cosine {cos:.4f}, |adjoint - J^T r| / |J^T r| = {rel:.3f}")
        if name == "expression":
            assert cos > 0.9
        elif name == "pose":
            assert cos > 0.5


def test_autograd_function():
    _say("case: torch.autograd.Function of the render loss")
    ag = importlib.import_module("3di_amd.autograd")
    sc = _small()
    b = tdi.Batch(tdi.Context(sc.model, sc.opts), sc.frames)
    n, V = len(sc.frames), sc.model.n_vertices
    th = torch.from_numpy(sc.init().astype(np.float32)).cuda().requires_grad_(True)
    dv = (0.01 * torch.randn((n, V, 3), device="cuda", generator=torch.Generator(device="cuda").manual_seed(5))
          ).requires_grad_(True)
    wts = torch.tensor([0.5, 1.0, 2.0][:n], device="cuda", dtype=torch.float64)
    loss = ag.render_loss(b, th, dv)
    (loss * wts).sum().backward()
    l2, gth, gv = b.render_loss(th.detach(), dv.detach(), gloss=wts)
    torch.cuda.synchronize()
    assert torch.allclose(loss.detach(), l2, rtol=1e-12)
    assert torch.allclose(th.grad, gth, rtol=1e-5, atol=1e-6 * float(gth.abs().max()))
    assert torch.allclose(dv.grad, gv, rtol=1e-5, atol=1e-6 * float(gv.abs().max()))
    assert float(th.grad.abs().sum()) > 0 and float(dv.grad.abs().sum()) > 0
