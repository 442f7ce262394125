"""GPU parity at the configurations' own sizes, through the C ABI, against the FP64 oracle.

cfg2 (the batched per-frame workload): the 231 x 231-grid model (53,361 vertices, 105,800
triangles), K_exp = 79, nine light unknowns (K = 94), eight frames with crops from 96 to 384
pixels a side.  cfg0 (the plain small fit): 100 x 100 grid, identity, albedo, expression, pose
and point light fitted (K = 110), 192 x 192 crops.

Integer buffers, bit for bit, as the fit computes them (tdi_fit_buffers):
  * the projected vertices the rasteriser reads, from the synthesis the batched fit uses (the
    FP64-MFMA tile when only expressions are fitted), equal to the oracle's FP64 projection
    rounded to FP32;
  * the rendered pixels' neighbour positions (left, right, above, below in sorted order) and the
    pixel counts, equal to the oracle's.
FP64 barrier terms (tdi_barrier_terms): H within 1e-10 relative of the oracle's (an FP32 barrier
misses by ~1e-6), g within 1e-6 (the fits keep g in FP32), the objective within 1e-12.
cfg0 normal equations and one Gauss-Newton iteration at K = 110, on the K > 96 J^T J tiles.
"""
import importlib
import sys

import numpy as np
import pytest

import pyoracle as O
import scenes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tdi = importlib.import_module("3di_amd.tdi")
abi = importlib.import_module("3di_amd.abi")

EMPTY = np.iinfo(np.uint64).max


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module", params=["cfg2", "cfg0"])
def cfg(request):
    _say(f"configuration: {request.param}")
    sc = scenes.cfg2_scene() if request.param == "cfg2" else scenes.cfg0_scene()
    ctx = tdi.Context(sc.model, sc.opts)
    b = tdi.Batch(ctx, sc.frames)
    yield request.param, sc, b
    b.close()
    ctx.close()


def _start(sc):
    return np.asarray(sc.init(), np.float32).astype(np.float64)


def test_fit_vertices_and_neighbours_bitexact(cfg):
    name, sc, b = cfg
    _say(f"case: fit buffers bit-exact ({name})")
    th = _start(sc)
    vproj, nbr, npix = b.fit_buffers(th, mixed_flags=sc.opts.mixed_flags)
    vproj, nbr, npix = vproj.cpu().numpy(), nbr.cpu().numpy(), npix.cpu().numpy()
    orc = sc.oracle()
    for i, f in enumerate(sc.frames):
        xyz, xyp, _, _ = orc.geometry(i, th[i])
        want = np.stack([xyp[:, 0].astype(np.float32), xyp[:, 1].astype(np.float32), xyz[:, 2].astype(np.float32)], 1)
        got = vproj[i, :, :3]
        bad = np.sum(np.any(got.view(np.uint32) != want.view(np.uint32), axis=1))
        assert bad == 0, f"frame {i}: {bad} of {len(got)} projected vertices differ from the oracle's FP32 rounding"
        ne = orc.normal_equations(i, th[i], want_dI=True)
        n = ne["npix"]
        assert int(npix[i]) == n, (i, int(npix[i]), n)
        o0 = int(b.offsets[i])
        g = nbr[o0:o0 + n]
        diff = np.sum(np.any(g != ne["nbr"], axis=1))
        _say(f"  frame {i}: {f.width}x{f.height} crop, {n} rendered pixels, vertices and neighbours bit-exact")
        assert diff == 0, f"frame {i}: {diff} of {n} neighbour rows differ"


def test_barrier_terms_fp64(cfg):
    name, sc, b = cfg
    _say(f"case: barrier terms in FP64 ({name})")
    th = _start(sc)
    H, g, obj = b.barrier_terms(th)
    H, g, obj = H.cpu().numpy(), g.cpu().numpy(), obj.cpu().numpy()
    orc = sc.oracle()
    for i in range(len(sc.frames)):
        ob, go, Ho = orc.barrier(i, th[i])
        eH = np.linalg.norm(H[i] - Ho) / np.linalg.norm(Ho)
        eg = np.linalg.norm(g[i] - go) / np.linalg.norm(go)
        eo = abs(obj[i] - ob) / abs(ob)
        _say(f"  frame {i}: |dH|/|H| {eH:.1e}, |dg|/|g| {eg:.1e}, objective {eo:.1e}")
        assert eH < 1e-10 and eg < 1e-6 and eo < 1e-12


def test_normal_equations_at_size(cfg):
    """H and g of the data term plus barrier against the oracle: the FP32 per-pixel Jacobian
    bounds the agreement at about 1e-7 (observed); asserted within 10x of that."""
    name, sc, b = cfg
    _say(f"case: normal equations ({name})")
    th = _start(sc)
    orc = sc.oracle()
    for mixed in (True, False):
        H, g, _ = b.normal_equations(th, mixed=mixed)
        H, g = H.cpu().numpy().astype(np.float64), g.cpu().numpy().astype(np.float64)
        for i in range(len(sc.frames)):
            r = orc.normal_equations(i, th[i])
            eH = np.linalg.norm(H[i] - r["H"]) / np.linalg.norm(r["H"])
            eg = np.linalg.norm(g[i] - r["g"]) / np.linalg.norm(r["g"])
            _say(f"  frame {i} mixed={mixed}: |dH|/|H| {eH:.1e}, |dg|/|g| {eg:.1e}")
            assert eH < 2e-6 and eg < 5e-6


def test_one_iteration_at_size(cfg):
    """One Gauss-Newton iteration of the batched path from the same FP32 start: the same
    line-search decisions as the oracle, and the step difference below 1e-3 of the step in the
    energy norm of the oracle's H.  A trial that leaves a landmark box is shrunk without counting
    as a line-search evaluation (model_fitter.cu:276-282), so the two can stop one shrink apart
    when the box edge crosses the line within the FP32 step's difference from the FP64 one: such a
    frame (at most one here) must still have the oracle's step direction to 1e-3.  cfg0 fits
    identity and albedo too (110 unknowns) and its normal matrix is worse conditioned, so the FP32
    per-pixel Jacobian moves its step further: 5e-3 there (scaled condition numbers printed)."""
    name, sc, b = cfg
    _say(f"case: one GN iteration ({name})")
    opts = abi.default_options(**{f[0]: getattr(sc.opts, f[0]) for f in abi.tdi_fit_options._fields_})
    opts.max_iters = 1
    th0 = _start(sc)
    orc = O.Scene(sc.model, sc.frames, opts)
    tho, repo = orc.fit(th0)
    ctx = tdi.Context(sc.model, opts)
    b1 = tdi.Batch(ctx, sc.frames)
    thg, repg = b1.fit_batched(torch.from_numpy(th0.astype(np.float32)).cuda())
    thg = thg.cpu().numpy().astype(np.float64)
    b1.close()
    ctx.close()
    assert [r.ls_evals for r in repg] == [r.ls_evals for r in repo]
    flips = 0
    for i in range(len(sc.frames)):
        Hm = sc.oracle().normal_equations(i, th0[i])["H"]
        s_, o_ = thg[i] - th0[i], tho[i] - th0[i]
        lam = float(s_ @ Hm @ o_) / float(o_ @ Hm @ o_)       # the GPU's step length over the oracle's
        if abs(lam - 1.0) > 1e-2:
            assert min(abs(lam - 0.8), abs(lam - 1.25)) < 1e-2, f"frame {i}: step length ratio {lam}"
            flips += 1
            s_ = s_ / lam
        d = s_ - o_
        en = np.sqrt(max(d @ Hm @ d, 0.0)) / max(np.sqrt(max(o_ @ Hm @ o_, 0.0)), 1e-30)
        dg = np.sqrt(np.diag(Hm))
        kappa = np.linalg.cond(Hm / np.outer(dg, dg))
        _say(f"  frame {i}: step length ratio {lam:.4f}, energy-norm rel {en:.1e}, scaled condition number {kappa:.1e}")
        assert en < (1e-3 if name == "cfg2" else 5e-3)
    assert flips <= 1
