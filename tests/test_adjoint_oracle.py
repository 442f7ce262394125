"""The oracle's render loss, the function the adjoint differentiates (tdi_render_loss_grad):
over the visible pixels, |g^_r - g^_s|^2 of the normalised gradients of the blurred render and
of the blurred observation (fill_grads, funcs.cu:842-897), plus the squared image-pixel distance
of the 51 projected landmark vertices to the detected landmarks.  Pinned by
  - a numpy evaluation of the same sum from the oracle's own render (its blurred image and
    visibility, which other tests pin) and the numpy blur of the observation, with the landmark
    projection through scipy's rotation;
  - the relation to the fit's objective: sum g^_r.g^_s = -obj_data / w, so the loss's data part is
    sum (|g^_r|^2 + |g^_s|^2) + (2 / w) obj_data;
  - per-vertex offsets: a common offset d of every vertex equals the translation R d, and the
    loss moves with them (the offset path is not a no-op);
  - the frozen-visibility evaluation (the triangle of each pixel fixed, exact FP64 barycentrics)
    agrees with the rasterised one at the same point to FP32 barycentric rounding, and is smooth:
    central differences at steps 1e-6 and 1e-7 agree.
"""
import numpy as np
from scipy.spatial.transform import Rotation

import pyoracle as O
from scenes import make_scene


def _scene():
    return make_scene(grid=48, k_exp=12, crops=(128, 160), pose_deg=15.0, fit_light=2)


def _numpy_loss(sc, i, th):
    ors = sc.oracle()
    fr = sc.frames[i]
    W, H = fr.width, fr.height
    r = ors.render(i, th)
    vis = r["keys"] != np.uint64(0xFFFFFFFFFFFFFFFF)
    B = r["img"]
    S = O.blur(fr.obs.astype(np.float64), W, H, sc.opts.blur_sigma)
    k = np.arange(W * H)
    x, y = k // H, k % H
    kr = np.where(x < W - 1, k + H, 0)
    kl = np.where(x > 0, k - H, 0)
    ka = np.where(y > 0, k - 1, 0)
    kb = np.where(y < H - 1, k + 1, 0)

    def nrm(I):
        gx, gy = 0.5 * (I[kr] - I[kl]), 0.5 * (I[kb] - I[ka])
        h = 1e-16 + np.sqrt(4 * gx * gx + 4 * gy * gy)
        return 2 * gx / h, 2 * gy / h

    rx, ry = nrm(B)
    sx, sy = nrm(S)
    data = ((rx - sx) ** 2 + (ry - sy) ** 2)[vis].sum()
    norms = (rx ** 2 + ry ** 2 + sx ** 2 + sy ** 2)[vis].sum()
    # landmarks
    m, K = sc.model, sc.K
    po = sc.lay["pose"]
    ids = m.landmark_ids
    P0 = m.mean_shape[ids].astype(np.float64) + (m.exp_basis.reshape(-1, 3, m.exp_basis.shape[1])[ids].astype(np.float64)
                                               @ th[sc.lay["exp"]:po])
    C = P0 @ Rotation.from_rotvec(th[po + 3:po + 6]).as_matrix().T + th[po:po + 3]
    xp = float(np.float32(fr.phix)) * C[:, 0] / C[:, 2] + float(np.float32(fr.cx))
    yp = float(np.float32(fr.phiy)) * C[:, 1] / C[:, 2] + float(np.float32(fr.cy))
    lmk = ((xp - fr.lmk_x) ** 2 + (yp - fr.lmk_y) ** 2).sum()
    return data, lmk, norms, r["obj"][0]


def test_loss_matches_numpy_and_the_objective():
    sc = _scene()
    ors = sc.oracle()
    for i in range(len(sc.frames)):
        th = sc.init()[i]
        got = ors.render_loss(i, th)
        data, lmk, norms, obj_data = _numpy_loss(sc, i, th)
        assert abs(got - (data + lmk)) <= 1e-9 * (data + lmk)
        w = sc.opts.data_weight
        assert abs(data - (norms + 2.0 / w * obj_data)) <= 1e-9 * data


def test_vertex_offsets():
    sc = _scene()
    ors = sc.oracle()
    V = sc.model.n_vertices
    po = sc.lay["pose"]
    th = sc.init()[0]
    d = np.array([1.5, -2.0, 0.7])
    R = Rotation.from_rotvec(th[po + 3:po + 6]).as_matrix()
    th2 = th.copy()
    th2[po:po + 3] += R @ d
    a = ors.render_loss(0, th, dverts=np.tile(d, (V, 1)))
    b = ors.render_loss(0, th2)
    assert abs(a - b) <= 1e-9 * abs(b)
    rng = np.random.default_rng(3)
    assert abs(ors.render_loss(0, th, dverts=rng.normal(0, 0.3, (V, 3))) - ors.render_loss(0, th)) > 1e-6


def test_frozen_visibility_is_smooth_and_consistent():
    sc = _scene()
    ors = sc.oracle()
    fr = sc.frames[0]
    th = sc.init()[0]
    tri = np.zeros(fr.width * fr.height, np.int32)
    L0 = ors.render_loss(0, th, tri=tri)
    assert (tri >= 0).sum() > 500
    L1 = ors.render_loss(0, th, tri=tri, frozen=True)
    assert abs(L0 - L1) <= 1e-4 * L0
    rng = np.random.default_rng(4)
    for _ in range(3):
        d = rng.normal(size=th.size)
        d /= np.linalg.norm(d)
        fd = []
        for h in (1e-6, 1e-7):
            fd.append((ors.render_loss(0, th + h * d, tri=tri, frozen=True)
                       - ors.render_loss(0, th - h * d, tri=tri, frozen=True)) / (2 * h))
        assert abs(fd[0] - fd[1]) <= 1e-4 * max(1.0, abs(fd[1]))
