"""GPU parity of the landmark-only fit with slack (tdi_fit_landmarks_slack,
tdi_landmark_slack_terms; Logbarrier_Initializer::set_minimal_slack and fit_model with use_slack,
logbarrier_initializer.cu:1041-1267) against the oracle.  Both sides compute in FP64; they differ
only in the order of sums and in log/sqrt/division rounding, so
  - the terms (objective, gradient, Gauss-Newton Hessian) agree to 1e-11 relative;
  - whole fits take the same decisions (status, Newton steps, rounds, accepted steps) and end at
    the same unknowns to 1e-6 relative (the device returns them in FP32); a fit that can never
    become feasible drives its slack to the bound 0, where steps differ in the last bits: there only
    the status, the rounds and the untouched theta are compared;
at the small-scene sizes (40x40 grid, 10 expressions; k_exp 10 and 6) and at the per-frame
workload's (231x231 grid, 79 expressions, k_exp 79: 86 unknowns).  Also: a frame already deep
inside its boxes (no Newton step), a zero-width box (never feasible, theta untouched), and the
argument checks.
"""
import importlib
import sys

import numpy as np
import pytest

from scenes import cfg2_scene, make_scene
from test_lmk_oracle import _constraints

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tdi = importlib.import_module("3di_amd.tdi")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _start(sc, seed, exp=0.7):
    K = sc.K
    po = sc.lay["pose"]
    rng = np.random.default_rng(seed)
    th = sc.truth().astype(np.float32).copy()
    th[:, po:po + 3] += rng.normal(0, 8, (len(th), 3)).astype(np.float32)
    th[:, po + 3:po + 6] += rng.normal(0, 0.08, (len(th), 3)).astype(np.float32)
    th[:, sc.lay["exp"]:po] = exp
    assert po + 6 <= K
    return th


def _zs(sc, Kl, seed):
    rng = np.random.default_rng(seed)
    po = sc.lay["pose"]
    zs = []
    for i in range(len(sc.frames)):
        z = np.zeros(Kl + 7)
        half = np.minimum(-sc.model.exp_lb[:Kl], sc.model.exp_ub[:Kl]).astype(np.float64)
        z[:Kl] = rng.uniform(-0.5, 0.5, Kl) * half          # inside the expression bounds
        z[Kl:Kl + 6] = sc.truth()[i, po:po + 6] + np.concatenate([rng.normal(0, 2, 3), rng.normal(0, 0.02, 3)])
        z[Kl + 6] = 2.0 - _constraints_pose(sc, i, z, Kl).min()
        zs.append(z)
    return np.stack(zs)


def _constraints_pose(sc, i, z, Kl):
    return _constraints(sc, i, z, Kl)


def _check_terms(sc, Kl, tau, seed):
    ors = sc.oracle()
    zs = _zs(sc, Kl, seed)
    th = sc.truth().astype(np.float32)
    b = tdi.Batch(tdi.Context(sc.model, sc.opts), sc.frames)
    H, g, obj = b.landmark_slack_terms(th, Kl, zs, tau)
    torch.cuda.synchronize()
    H, g, obj = H.cpu().numpy(), g.cpu().numpy(), obj.cpu().numpy()
    for i in range(len(sc.frames)):
        o, gw, Hw = ors.landmark_slack_terms(i, th[i].astype(np.float64), Kl, zs[i], tau)
        eo = abs(obj[i] - o) / max(1.0, abs(o))
        eg = np.abs(g[i] - gw).max() / max(1.0, np.abs(gw).max())
        eH = np.abs(H[i] - Hw).max() / np.abs(Hw).max()
        _say(f"  frame {i}: obj {eo:.1e}, g {eg:.1e}, H {eH:.1e}")
        assert np.isfinite(o) and eo < 1e-11 and eg < 1e-11 and eH < 1e-11


@pytest.mark.parametrize("Kl", [10, 6])
def test_terms_small(Kl):
    _say(f"case: landmark slack terms, small scene, k_exp {Kl}")
    sc = make_scene(grid=40, k_exp=10, crops=(96, 150, 220, 300), pose_deg=15.0)
    _check_terms(sc, Kl, 100.0, 11)


def _check_fit(sc, Kl, th0):
    ors = sc.oracle()
    want, wreps = ors.fit_landmarks_slack(th0.astype(np.float64), Kl)
    b = tdi.Batch(tdi.Context(sc.model, sc.opts), sc.frames)
    dth = torch.from_numpy(th0.copy()).cuda()
    dth, reps = b.fit_landmarks_slack(dth, Kl)
    torch.cuda.synchronize()
    got = dth.cpu().numpy()
    for i, (r, w) in enumerate(zip(reps, wreps)):
        _say(f"  frame {i}: status {r.status}/{w.status} steps {r.iterations}/{w.iterations} rounds "
             f"{r.rounds}/{w.rounds} accepted {r.accepted}/{w.accepted} slack {r.slack:.4f}/{w.slack:.4f}")
        assert (r.status, r.rounds) == (w.status, w.rounds)
        assert abs(r.slack0 - w.slack0) <= 1e-6 * max(1.0, abs(w.slack0))
        if w.status != 0:
            # never feasible: the slack sinks towards its bound 0, where the objective's steps
            # differ in the last bits and the step counts may part; theta stays as it was
            assert np.array_equal(got[i], th0[i]) and abs(r.slack) < 1e-3
            continue
        assert (r.iterations, r.accepted) == (w.iterations, w.accepted)
        assert abs(r.slack - w.slack) <= 1e-5 * max(1.0, abs(w.slack))
        err = np.abs(got[i] - want[i]) / np.maximum(1.0, np.abs(want[i]))
        assert err.max() < 1e-6, (i, err.argmax(), got[i, err.argmax()], want[i, err.argmax()])
    return reps


@pytest.mark.parametrize("Kl", [10, 6])
def test_fit_small(Kl):
    _say(f"case: landmark slack fit, small scene, k_exp {Kl}")
    sc = make_scene(grid=40, k_exp=10, crops=(96, 150, 220, 300, 130, 260), pose_deg=15.0)
    reps = _check_fit(sc, Kl, _start(sc, 21))
    assert all(r.status == 0 for r in reps)


def test_workload_sizes():
    _say("case: landmark slack terms and fit at the per-frame workload's sizes, k_exp 79")
    sc = cfg2_scene(crops=(96, 180, 260, 384))
    _check_terms(sc, 79, 1000.0, 12)
    reps = _check_fit(sc, 79, _start(sc, 22))
    assert all(r.status == 0 for r in reps)


def test_deep_inside_and_zero_width_box():
    _say("case: landmark slack fit, start deep inside the boxes and a zero-width box")
    sc = make_scene(grid=40, k_exp=10, crops=(150, 150), pose_deg=15.0)
    f0, f1 = sc.frames
    for b_ in (f0.box_lx, f0.box_ly):
        b_[:] = -0.6
    for b_ in (f0.box_ux, f0.box_uy):
        b_[:] = 0.6
    f1.box_lx[7] = 0.0
    f1.box_ux[7] = 0.0
    th0 = sc.truth().astype(np.float32)
    reps = _check_fit(sc, 10, th0)
    assert reps[0].status == 0 and reps[0].iterations == 0
    assert reps[1].status == 8 and reps[1].rounds == 5


def test_rejects_bad_arguments():
    sc = make_scene(grid=40, k_exp=10, crops=(96,), pose_deg=15.0)
    b = tdi.Batch(tdi.Context(sc.model, sc.opts), sc.frames)
    th = torch.from_numpy(sc.truth().astype(np.float32)).cuda()
    for Kl in (0, 11):
        with pytest.raises(tdi.TdiError):
            b.fit_landmarks_slack(th, Kl)
