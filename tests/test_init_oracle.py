"""The oracle's landmark-driven initialisation: the orthographic 7-parameter fit
(OrthographicInitializer::fit_model, orthographic_initializer.cu:243-377), its angle helpers
(compute_euler_angles, rotation_header.cu:421-449; eulerAnglesToRotationMatrix,
newfuncs.cu:720-747, with the cv::Rodrigues log map of logbarrier_initializer.cu:1883) and the
pose seeding of initialize_kernel (logbarrier_initializer.cu:31-46).

How each function is pinned (no reference test or fixture covers them):
  - ortho_residuals: the Jacobian against central differences of the residuals, column by
    column, and the residuals at a noise-free truth are zero;
  - ortho_fit: recovers a noise-free orthographic truth;
  - rotvec_from_euler: scipy's extrinsic x-y-z Euler rotation (Rz Ry Rx) and its rotation
    vector, including the theta -> pi branch;
  - euler_angles: scipy's Z-Y-X decomposition of the same matrices;
  - the pose seed: returns the translation and depth of a weak-perspective camera;
  - the mirrored-pose refit: runs on mirrored landmarks.
"""
from __future__ import annotations

import importlib

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

import pyoracle as O

synth = importlib.import_module("3di_amd.synth")

_cache = {}


def _model():
    if "m" not in _cache:
        _cache["m"] = synth.make_model(231, 0, 0, 79, 0x3D1 + 2)
    return _cache["m"]


def _p0L(m):
    return m.mean_shape.astype(np.float64)[m.landmark_ids.astype(np.int64)]


def _ortho_landmarks(p0L, x, offset):
    """Landmarks of the orthographic model at x = (tau, sigma, u), shifted by offset."""
    R = Rotation.from_rotvec(x[4:7]).as_matrix()
    Rp = p0L @ R.T
    return np.stack([x[2] * (Rp[:, 0] + x[0]) + offset[0], x[3] * (Rp[:, 1] + x[1]) + offset[1]])


def _perspective_frames(n, seed=2):
    m = _model()
    lay = synth.layout(m, False, False, True, True)
    frames, truths = [], []
    for i in range(n):
        rng = np.random.default_rng(synth.frame_seed(seed, i))
        tr = synth.draw_truth(m, rng, lay, 30.0, fit_light=True)
        S = int(rng.integers(96, 385))
        frames.append(synth.make_frame(m, tr, lay, rng, synth.cfg2_focal(S), S, S))
        truths.append(tr)
    lmk = np.stack([np.stack([f.lmk_x, f.lmk_y]) for f in frames]).astype(np.float32)
    cams = np.array([[f.phix, f.phiy, f.cx, f.cy] for f in frames], np.float32)
    return m, lay, lmk, cams, truths


def test_ortho_jacobian_matches_central_differences():
    m = _model()
    p0L = _p0L(m)
    rng = np.random.default_rng(5)
    xlb, ylb = rng.normal(0, 40, 51), rng.normal(0, 40, 51)
    for _ in range(4):
        x = np.r_[rng.normal(0, 5, 2), rng.uniform(1, 4, 2), rng.normal(0, 0.4, 3)]
        e, J = O.ortho_residuals(x, xlb, ylb, p0L)
        for k in range(7):
            h = 1e-6 * max(1.0, abs(x[k]))
            xp, xm = x.copy(), x.copy()
            xp[k] += h
            xm[k] -= h
            fd = (O.ortho_residuals(xp, xlb, ylb, p0L, False) - O.ortho_residuals(xm, xlb, ylb, p0L, False)) / (2 * h)
            assert np.allclose(J[:, k], fd, rtol=1e-6, atol=1e-6 * np.abs(fd).max()), k


def test_ortho_fit_recovers_noise_free_truth():
    m = _model()
    p0L = _p0L(m)
    rng = np.random.default_rng(11)
    lmk, truths = [], []
    for _ in range(8):
        u = rng.uniform(-0.3, 0.3, 3)
        s = rng.uniform(1.0, 4.0)
        R = Rotation.from_rotvec(u).as_matrix()
        Rp = p0L @ R.T
        # tau centres the model landmarks, as the fit sees mean-centred detections
        x = np.r_[-Rp[:, 0].mean(), -Rp[:, 1].mean(), s, s * rng.uniform(0.95, 1.05), u]
        lmk.append(_ortho_landmarks(p0L, x, rng.uniform(100, 600, 2)))
        truths.append(x)
        e = O.ortho_residuals(x, lmk[-1][0] - lmk[-1][0].mean(), lmk[-1][1] - lmk[-1][1].mean(), p0L, False)
        assert np.abs(e).max() < 1e-9
    cams = np.tile(np.array([[800, 800, 320, 240]], np.float32), (8, 1))
    out = O.ortho_init(np.stack(lmk), cams, p0L)
    assert (out["info"][:, 3] == 0).all() and (out["info"][:, 2] == 0).all()
    truths = np.stack(truths)
    # float32 landmarks limit the recovery to ~1e-6 of the scale
    assert np.allclose(out["x_orth"], truths, rtol=1e-4, atol=2e-4)


def test_rotvec_from_euler_matches_scipy():
    rng = np.random.default_rng(3)
    for _ in range(20):
        e = rng.uniform(-1.2, 1.2, 3)
        want = Rotation.from_euler("xyz", e).as_rotvec()     # extrinsic: Rz(e2) Ry(e1) Rx(e0)
        assert np.allclose(O.rotvec_from_euler(e), want, atol=1e-12)
    # theta -> 0 and theta -> pi branches: compare the rotations themselves
    for e in ([0.0, 0.0, 0.0], [np.pi, 0.0, 0.0], [0.0, np.pi, 0.0], [np.pi - 1e-9, 0.0, 0.0],
              [0.3, np.pi, -0.2]):
        M = Rotation.from_euler("xyz", e).as_matrix()
        r = O.rotvec_from_euler(np.array(e))
        assert np.allclose(Rotation.from_rotvec(r).as_matrix(), M, atol=1e-6), e


def test_euler_angles_are_the_zyx_decomposition():
    rng = np.random.default_rng(4)
    for _ in range(20):
        pitch, yaw, roll = rng.uniform(-1.2, 1.2), rng.uniform(-1.2, 1.2), rng.uniform(-1.2, 1.2)
        M = Rotation.from_euler("xyz", [pitch, yaw, roll]).as_matrix()   # Rz(roll) Ry(yaw) Rx(pitch)
        assert np.allclose(O.euler_angles(M), [yaw, pitch, roll], atol=1e-12)


def test_pose_seed_recovers_a_weak_perspective_pose():
    # Landmarks of a weak-perspective camera: sigma = phi / Z, centroid at the projection of the
    # translation.  The seed must return that translation and depth.
    m = _model()
    p0L = _p0L(m)
    rng = np.random.default_rng(13)
    lmk, cams, want = [], [], []
    for _ in range(8):
        phi, cx, cy = rng.uniform(500, 1500), rng.uniform(100, 400), rng.uniform(100, 400)
        Z, tx, ty = rng.uniform(300, 900), rng.uniform(-40, 40), rng.uniform(-40, 40)
        u = rng.uniform(-0.3, 0.3, 3)
        R = Rotation.from_rotvec(u).as_matrix()
        Rp = p0L @ R.T
        s = phi / Z
        x = np.r_[-Rp[:, 0].mean(), -Rp[:, 1].mean(), s, s, u]
        lmk.append(_ortho_landmarks(p0L, x, (cx + phi * tx / Z, cy + phi * ty / Z)))
        cams.append([phi, phi, cx, cy])
        want.append(np.r_[tx, ty, Z, u])
    out = O.ortho_init(np.stack(lmk), np.array(cams, np.float32), p0L)
    want = np.stack(want)
    assert np.allclose(out["pose"][:, 2], want[:, 2], rtol=1e-4)
    assert np.allclose(out["pose"][:, :2], want[:, :2], atol=1e-3 * want[:, 2:3])
    assert np.allclose(out["pose"][:, 3:], want[:, 3:], atol=1e-4)


def test_mirrored_landmarks_take_the_refit():
    m, lay, lmk, cams, _ = _perspective_frames(16)
    mirrored = lmk.copy()
    mirrored[:, 0, :] = 2 * cams[:, None, 2] - mirrored[:, 0, :]
    out = O.ortho_init(mirrored, cams, _p0L(m))
    assert (out["info"][:, 2] == 1).all() and (out["info"][:, 1] > 0).all()
    plain = O.ortho_init(lmk, cams, _p0L(m))
    assert (plain["info"][:, 2] == 0).all() and (plain["info"][:, 1] == 0).all()


@pytest.mark.parametrize("bad", ["ctx"])
def test_ortho_init_rejects_a_null_context(bad):
    tdi = importlib.import_module("3di_amd.tdi")
    import ctypes as C
    rc = tdi.lib().tdi_ortho_init(None, 1, C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), None,
                                  C.c_void_p(1), None)
    assert rc == 1   # TDI_ERR_ARG, before any device call
