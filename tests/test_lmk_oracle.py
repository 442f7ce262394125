"""The landmark-only log-barrier fit with slack (Logbarrier_Initializer::set_minimal_slack and
fit_model with use_slack, logbarrier_initializer.cu:1041-1267; terms of
compute_gradient_and_hessian / evaluate_objective_function, 1272-1842, 1921-2081): the oracle's
restatement pinned by
  - its objective against an independent numpy form of the reference's definitions (compute_Gs,
    458-518: G = s + the landmark's signed distance inside each side of its box; scipy's rotation
    vector for R; the expression-bound barrier), at random feasible points;
  - its gradient against central differences of that objective (every unknown, the slack too);
  - its Gauss-Newton Hessian against sum_c dG_c dG_c^T / G_c^2 over the 204 constraints, with dG
    from central differences of the numpy constraint values (the slack's own diagonal takes the
    reference's 1/(1e-6 + G)^2), plus the bound diagonal;
  - known answers of the loop: the start s = 5 - min G at s = 0; a frame whose landmarks already
    lie deep inside their boxes succeeds at once (no Newton step; expressions reset to 0, pose
    kept); a zero-width box cannot be met (s >= 0 at best), so all five rounds run and theta is
    left as it is; on success every landmark lies more than 0.01 px inside its box at s = 0.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from scenes import make_scene

NL = 51


def _scene(k_exp=10, crops=(96, 150, 220, 300)):
    return make_scene(grid=40, k_exp=k_exp, crops=crops, pose_deg=15.0)


def _constraints(sc, i, z, Kl):
    """The 204 constraint values G (numpy, from the reference's definitions) at z."""
    m, fr = sc.model, sc.frames[i]
    ids = m.landmark_ids
    E = m.exp_basis.reshape(-1, 3, m.exp_basis.shape[1])[ids][:, :, :Kl].astype(np.float64)
    P = m.mean_shape[ids].astype(np.float64) + E @ z[:Kl]
    R = Rotation.from_rotvec(z[Kl + 3:Kl + 6]).as_matrix()
    C = P @ R.T + z[Kl:Kl + 3]
    phix, phiy, cx, cy = (float(np.float32(v)) for v in (fr.phix, fr.phiy, fr.cx, fr.cy))
    xp = phix * C[:, 0] / C[:, 2] + cx
    yp = phiy * C[:, 1] / C[:, 2] + cy
    fs, s = float(np.float32(fr.face_size)), z[Kl + 6]
    lx, ly = fr.lmk_x.astype(np.float64), fr.lmk_y.astype(np.float64)
    blx, bux, bly, buy = (np.asarray(b, np.float32).astype(np.float64) for b in
                          (fr.box_lx, fr.box_ux, fr.box_ly, fr.box_uy))
    return np.concatenate([s + xp - (lx + blx * fs), s + (lx + bux * fs) - xp,
                           s + yp - (ly + bly * fs), s + (ly + buy * fs) - yp])


def _objective(sc, i, z, Kl, tau):
    m = sc.model
    G = _constraints(sc, i, z, Kl)
    e = z[:Kl]
    ub, lb = m.exp_ub[:Kl].astype(np.float64), m.exp_lb[:Kl].astype(np.float64)
    return tau * z[Kl + 6] - np.log(G).sum() - (np.log(ub - e) + np.log(e - lb)).sum()


def _feasible_point(sc, i, Kl, rng):
    K = sc.K
    th = sc.truth()[i]
    z = np.zeros(Kl + 7)
    z[:Kl] = rng.normal(0, 0.3, Kl)
    z[Kl:Kl + 6] = th[K - 6:] + np.concatenate([rng.normal(0, 2, 3), rng.normal(0, 0.02, 3)])
    G = _constraints(sc, i, z, Kl)
    z[Kl + 6] = 2.0 - G.min()           # every constraint at least 2
    return z


@pytest.mark.parametrize("Kl", [10, 6])
def test_objective_gradient_hessian(Kl):
    sc = _scene()
    ors = sc.oracle()
    rng = np.random.default_rng(5)
    tau = 100.0
    for i in range(len(sc.frames)):
        z = _feasible_point(sc, i, Kl, rng)
        th = sc.truth()[i]
        obj, g, H = ors.landmark_slack_terms(i, th, Kl, z, tau)
        want = _objective(sc, i, z, Kl, tau)
        assert abs(obj - want) <= 1e-10 * max(1.0, abs(want))
        # gradient: central differences of the numpy objective
        n = Kl + 7
        fd = np.zeros(n)
        for j in range(n):
            h = 1e-6 * max(1.0, abs(z[j]))
            zp, zm = z.copy(), z.copy()
            zp[j] += h
            zm[j] -= h
            fd[j] = (_objective(sc, i, zp, Kl, tau) - _objective(sc, i, zm, Kl, tau)) / (2 * h)
        assert np.abs(g - fd).max() <= 1e-5 * max(1.0, np.abs(fd).max()), (np.abs(g - fd).max(), np.abs(fd).max())
        # Gauss-Newton Hessian from the constraints' Jacobian
        G = _constraints(sc, i, z, Kl)
        D = np.zeros((4 * NL, n))
        for j in range(n):
            h = 1e-6 * max(1.0, abs(z[j]))
            zp, zm = z.copy(), z.copy()
            zp[j] += h
            zm[j] -= h
            D[:, j] = (_constraints(sc, i, zp, Kl, ) - _constraints(sc, i, zm, Kl)) / (2 * h)
        Hw = (D / G[:, None]).T @ (D / G[:, None])
        Hw[n - 1, n - 1] = (1.0 / (1e-6 + G) ** 2).sum()
        e = z[:Kl]
        ub, lb = sc.model.exp_ub[:Kl].astype(np.float64), sc.model.exp_lb[:Kl].astype(np.float64)
        Hw[np.arange(Kl), np.arange(Kl)] += 1 / (ub - e + 1e-12) ** 2 + 1 / (e - lb + 1e-12) ** 2
        assert np.allclose(H, H.T)
        assert np.abs(H - Hw).max() <= 1e-5 * np.abs(Hw).max()


def test_fit_known_answers():
    Kl = 10
    sc = _scene()
    ors = sc.oracle()
    K = sc.K
    rng = np.random.default_rng(9)
    th0 = sc.truth().copy()
    th0[:, K - 6:K - 3] += rng.normal(0, 8, (len(th0), 3))
    th0[:, K - 3:] += rng.normal(0, 0.08, (len(th0), 3))
    th0[:, :K - 6] = 0.7
    th, reps = ors.fit_landmarks_slack(th0, Kl)
    for i, r in enumerate(reps):
        # start: s0 = 5 - min G at eps = 0, s = 0
        z0 = np.zeros(Kl + 7)
        z0[Kl:Kl + 6] = th0[i, K - 6:]
        assert abs(r.slack0 - (5.0 - _constraints(sc, i, z0, Kl).min())) < 1e-5 * max(1.0, abs(r.slack0))
        assert r.status == 0 and r.rounds >= 1 and r.iterations >= 1 and r.slack < -0.01
        # every landmark at least 0.01 px inside its box at the result (s = 0)
        z = np.zeros(Kl + 7)
        z[:Kl] = th[i, :Kl]
        z[Kl:Kl + 6] = th[i, K - 6:]
        assert _constraints(sc, i, z, Kl).min() > 0.01
        assert (th[i, :Kl] > sc.model.exp_lb[:Kl]).all() and (th[i, :Kl] < sc.model.exp_ub[:Kl]).all()


def test_fit_deep_inside_succeeds_at_once():
    Kl = 10
    sc = _scene(crops=(150,))
    fr = sc.frames[0]
    for b in (fr.box_lx, fr.box_ly):
        b[:] = -0.6
    for b in (fr.box_ux, fr.box_uy):
        b[:] = 0.6
    K = sc.K
    th0 = sc.truth().copy()
    th0[:, :Kl] = 0.5
    ors = sc.oracle()
    th, (r,) = ors.fit_landmarks_slack(th0, Kl)
    assert r.status == 0 and r.iterations == 0 and r.accepted == 0 and r.rounds == 1
    assert r.slack0 < -0.01 and r.slack == r.slack0
    assert (th[0, :Kl] == 0).all() and np.array_equal(th[0, K - 6:], th0[0, K - 6:])


def test_fit_zero_width_box_is_infeasible():
    Kl = 10
    sc = _scene(crops=(150,))
    fr = sc.frames[0]
    fr.box_lx[7] = 0.0
    fr.box_ux[7] = 0.0
    th0 = sc.truth().copy()
    ors = sc.oracle()
    th, (r,) = ors.fit_landmarks_slack(th0, Kl)
    assert r.status == 8 and r.rounds == 5 and r.slack > -0.01
    assert 5 <= r.iterations <= 500
    assert np.array_equal(th, th0)


def test_fit_rejects_bad_expression_count():
    sc = _scene(crops=(96,))
    ors = sc.oracle()
    with pytest.raises(ValueError):
        ors.fit_landmarks_slack(sc.truth(), 0)
    with pytest.raises(ValueError):
        ors.fit_landmarks_slack(sc.truth(), sc.model.exp_basis.shape[1] + 1)
