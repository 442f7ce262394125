"""The joint landmark-only fit over a group of frames (Logbarrier_Initializer::fit_model over T
frames with identity unknowns: fit_to_multi_images_landmarks_only, model_fitter.cu:1205-1247, and
the two landmark fits at the start of fit_to_multi_images, 1291-1330; terms of
compute_gradient_and_hessian, logbarrier_initializer.cu:1272-1842).  The oracle's restatement
(oracle/tdi_oracle.cpp, lmk_joint_*) pinned by
  - one frame without identity unknowns: exactly the single-frame fit's objective, gradient and
    Hessian (pinned on their own in test_lmk_oracle.py), with the slack moved to the front;
  - its objective against an independent numpy form of the reference's definitions over all
    frames (shared identity columns in every frame's landmark positions, identity bounds and
    the slack's tau_b s once, each frame's expression bounds), at random feasible points, with
    and without the slack;
  - its gradient against central differences of that objective (every unknown);
  - its Gauss-Newton Hessian against sum_c dG_c dG_c^T / G_c^2 over every frame's 204
    constraints, dG from central differences of the numpy constraint values (the slack's
    diagonal with the reference's 1/(1e-6 + G)^2), plus the bound diagonals: zero between two
    frames' own blocks;
  - the fit: from zero identity and expressions and perturbed poses the fit with the slack
    succeeds (s < -0.01), after which every landmark of every frame lies inside its box at s = 0;
    the fit without the slack from there lowers its objective and keeps every landmark inside.
"""
import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from scenes import joint_scene

NL = 51


def _scene(crops=(128, 160, 144)):
    return joint_scene(crops=crops)


def _unpack(sc, z, Ka, Ke, slack):
    n = len(sc.frames)
    alpha = z[:Ka]
    s = z[Ka] if slack else 0.0
    o = Ka + (1 if slack else 0)
    blocks = [z[o + t * (Ke + 6):o + (t + 1) * (Ke + 6)] for t in range(n)]
    return alpha, s, blocks


def _constraints(sc, t, z, Ka, Ke, slack, theta0):
    """The 204 constraint values G of frame t (numpy, from the reference's definitions)."""
    m, fr, lay = sc.model, sc.frames[t], sc.lay
    alpha, s, blocks = _unpack(sc, z, Ka, Ke, slack)
    b = blocks[t]
    ids = m.landmark_ids
    A = m.id_basis.reshape(-1, 3, m.id_basis.shape[1])[ids].astype(np.float64)
    E = m.exp_basis.reshape(-1, 3, m.exp_basis.shape[1])[ids][:, :, :Ke].astype(np.float64)
    nid = lay["exp"] - lay["id"]
    fixed = theta0[lay["id"] + Ka:lay["id"] + nid]
    P = m.mean_shape[ids].astype(np.float64) + A[:, :, :Ka] @ alpha + A[:, :, Ka:nid] @ fixed + E @ b[:Ke]
    R = Rotation.from_rotvec(b[Ke + 3:Ke + 6]).as_matrix()
    C = P @ R.T + b[Ke:Ke + 3]
    phix, phiy, cx, cy = (float(np.float32(v)) for v in (fr.phix, fr.phiy, fr.cx, fr.cy))
    xp = phix * C[:, 0] / C[:, 2] + cx
    yp = phiy * C[:, 1] / C[:, 2] + cy
    fs = float(np.float32(fr.face_size))
    lx, ly = fr.lmk_x.astype(np.float64), fr.lmk_y.astype(np.float64)
    blx, bux, bly, buy = (np.asarray(v, np.float32).astype(np.float64) for v in
                          (fr.box_lx, fr.box_ux, fr.box_ly, fr.box_uy))
    return np.concatenate([s + xp - (lx + blx * fs), s + (lx + bux * fs) - xp,
                           s + yp - (ly + bly * fs), s + (ly + buy * fs) - yp])


def _objective(sc, z, Ka, Ke, slack, tau, theta0):
    m = sc.model
    alpha, s, blocks = _unpack(sc, z, Ka, Ke, slack)
    obj = tau * s if slack else 0.0
    iub, ilb = m.id_ub[:Ka].astype(np.float64), m.id_lb[:Ka].astype(np.float64)
    obj -= (np.log(iub - alpha) + np.log(alpha - ilb)).sum()
    eub, elb = m.exp_ub[:Ke].astype(np.float64), m.exp_lb[:Ke].astype(np.float64)
    for t, b in enumerate(blocks):
        obj -= np.log(_constraints(sc, t, z, Ka, Ke, slack, theta0)).sum()
        obj -= (np.log(eub - b[:Ke]) + np.log(b[:Ke] - elb)).sum()
    return obj


def _feasible_point(sc, Ka, Ke, slack, rng):
    lay = sc.lay
    tr = sc.truth()
    n = len(sc.frames)
    nz = Ka + (1 if slack else 0) + n * (Ke + 6)
    z = np.zeros(nz)
    z[:Ka] = tr[0, lay["id"]:lay["id"] + Ka] + rng.normal(0, 0.05, Ka)
    o = Ka + (1 if slack else 0)
    for t in range(n):
        b = z[o + t * (Ke + 6):o + (t + 1) * (Ke + 6)]
        b[:Ke] = tr[t, lay["exp"]:lay["exp"] + Ke] * 0.5
        b[Ke:Ke + 6] = tr[t, lay["pose"]:lay["pose"] + 6] + np.concatenate([rng.normal(0, 1, 3), rng.normal(0, 0.01, 3)])
    if slack:
        G = np.concatenate([_constraints(sc, t, z, Ka, Ke, slack, tr[0]) for t in range(n)])
        z[Ka] = 2.0 - G.min()
    return z


def test_one_frame_is_the_single_frame_fit():
    sc = _scene(crops=(150,))
    ors = sc.oracle()
    lay = sc.lay
    Ke = 10
    rng = np.random.default_rng(3)
    th = sc.truth()[0].copy()
    zs = np.zeros(Ke + 7)
    zs[:Ke] = rng.normal(0, 0.2, Ke)
    zs[Ke:Ke + 6] = th[lay["pose"]:lay["pose"] + 6]
    zs[Ke + 6] = 30.0
    tau = 1000.0
    o1, g1, H1 = ors.landmark_slack_terms(0, th, Ke, zs, tau)
    # joint order [s | eps | tau | u]; identity past Ka = 0 from theta, as the single-frame fit
    # takes theta's identity
    perm = np.r_[Ke + 6, np.arange(Ke + 6)]
    oj, gj, Hj = ors.landmark_joint_terms(th[None], 0, Ke, True, zs[perm], tau)
    assert abs(oj - o1) <= 1e-12 * abs(o1)
    assert np.allclose(gj, g1[perm], rtol=1e-12, atol=1e-12 * np.abs(g1).max())
    assert np.allclose(Hj, H1[np.ix_(perm, perm)], rtol=1e-12, atol=1e-12 * np.abs(H1).max())


@pytest.mark.parametrize("slack", [True, False])
def test_objective_gradient_hessian(slack):
    sc = _scene()
    ors = sc.oracle()
    Ka, Ke, tau = 12, 10, 100.0
    n = len(sc.frames)
    rng = np.random.default_rng(11 + slack)
    th0 = sc.truth()
    z = _feasible_point(sc, Ka, Ke, slack, rng)
    if not slack:
        G = np.concatenate([_constraints(sc, t, z, Ka, Ke, slack, th0[0]) for t in range(n)])
        assert G.min() > 0.0
    obj, g, H = ors.landmark_joint_terms(th0, Ka, Ke, slack, z, tau)
    want = _objective(sc, z, Ka, Ke, slack, tau, th0[0])
    assert abs(obj - want) <= 1e-10 * max(1.0, abs(want))
    nz = z.size
    fd = np.zeros(nz)
    D = []
    for j in range(nz):
        h = 1e-6 * max(1.0, abs(z[j]))
        zp, zm = z.copy(), z.copy()
        zp[j] += h
        zm[j] -= h
        fd[j] = (_objective(sc, zp, Ka, Ke, slack, tau, th0[0]) - _objective(sc, zm, Ka, Ke, slack, tau, th0[0])) / (2 * h)
        D.append(np.concatenate([(_constraints(sc, t, zp, Ka, Ke, slack, th0[0])
                                  - _constraints(sc, t, zm, Ka, Ke, slack, th0[0])) / (2 * h) for t in range(n)]))
    assert np.abs(g - fd).max() <= 1e-5 * max(1.0, np.abs(fd).max()), (np.abs(g - fd).max(), np.abs(fd).max())
    D = np.stack(D, axis=1)
    G = np.concatenate([_constraints(sc, t, z, Ka, Ke, slack, th0[0]) for t in range(n)])
    Hw = (D / G[:, None]).T @ (D / G[:, None])
    if slack:
        Hw[Ka, Ka] = (1.0 / (1e-6 + G) ** 2).sum()
    a = z[:Ka]
    iub, ilb = sc.model.id_ub[:Ka].astype(np.float64), sc.model.id_lb[:Ka].astype(np.float64)
    Hw[np.arange(Ka), np.arange(Ka)] += 1 / (iub - a + 1e-12) ** 2 + 1 / (a - ilb + 1e-12) ** 2
    o = Ka + (1 if slack else 0)
    eub, elb = sc.model.exp_ub[:Ke].astype(np.float64), sc.model.exp_lb[:Ke].astype(np.float64)
    for t in range(n):
        k = o + t * (Ke + 6) + np.arange(Ke)
        Hw[k, k] += 1 / (eub - z[k] + 1e-12) ** 2 + 1 / (z[k] - elb + 1e-12) ** 2
    assert np.allclose(H, H.T)
    assert np.abs(H - Hw).max() <= 1e-5 * np.abs(Hw).max()
    # two frames' own blocks do not meet
    b0 = slice(o, o + Ke + 6)
    b1 = slice(o + Ke + 6, o + 2 * (Ke + 6))
    assert not H[b0, b1].any()


def test_fit_with_and_without_slack():
    sc = _scene()
    ors = sc.oracle()
    lay = sc.lay
    Ka, Ke = 12, 10
    n = len(sc.frames)
    rng = np.random.default_rng(21)
    tr = sc.truth()
    th0 = tr.copy()
    th0[:, lay["id"]:lay["exp"]] = 0.0
    th0[:, lay["exp"]:lay["pose"]] = 0.0
    th0[:, lay["pose"]:lay["pose"] + 3] += rng.normal(0, 6, (n, 3))
    th0[:, lay["pose"] + 3:lay["pose"] + 6] += rng.normal(0, 0.05, (n, 3))
    th1, r1 = ors.fit_landmarks_joint(th0, Ka, Ke, True)
    assert r1.status == 0 and r1.rounds >= 1 and r1.iterations >= 1 and r1.slack < -0.01
    assert r1.slack0 > 0
    # the shared identity is written to every frame; every landmark inside its box at s = 0
    assert all(np.array_equal(th1[t, lay["id"]:lay["id"] + Ka], th1[0, lay["id"]:lay["id"] + Ka]) for t in range(n))
    z1 = np.concatenate([th1[0, lay["id"]:lay["id"] + Ka]] +
                        [np.r_[th1[t, lay["exp"]:lay["exp"] + Ke], th1[t, lay["pose"]:lay["pose"] + 6]] for t in range(n)])
    G = np.concatenate([_constraints(sc, t, z1, Ka, Ke, False, th1[0]) for t in range(n)])
    assert G.min() > 0.01
    # the fit without the slack from there: lower objective at its last barrier weight, landmarks inside
    o_start = _objective(sc, z1, Ka, Ke, False, 0.0, th1[0])
    th2, r2 = ors.fit_landmarks_joint(th1, Ka, Ke, False)
    assert r2.status == 0 and r2.rounds == 5 and r2.iterations >= 1
    z2 = np.concatenate([th2[0, lay["id"]:lay["id"] + Ka]] +
                        [np.r_[th2[t, lay["exp"]:lay["exp"] + Ke], th2[t, lay["pose"]:lay["pose"] + 6]] for t in range(n)])
    assert _objective(sc, z2, Ka, Ke, False, 0.0, th2[0]) < o_start
    G2 = np.concatenate([_constraints(sc, t, z2, Ka, Ke, False, th2[0]) for t in range(n)])
    assert G2.min() > 0.0
