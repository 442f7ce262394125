"""The oracle's joint system of the identity stage (pyoracle.JointScene): frames that share the
albedo and identity coefficients, every frame with its own expression, pose and light
(fit_to_multi_images' shape fit, model_fitter.cu:1252-1479 and 121-375; the frames' products
summed as in get_obj_hessian_and_gradient_multiframe, 21-116).  Pinned by
  - one frame: the joint system is the frame's own normal equations (same matrix, gradient and
    objective: the shared bounds counted once is the single-frame barrier);
  - the block structure: symmetric, zero between different frames' blocks, the shared block the
    sum of the frames' shared blocks plus the bound terms, the coupling blocks the frames' own;
  - the shared bound terms against their closed form (-log(ub - x) - log(x - lb), the gradient
    and the Gauss-Newton diagonal of logbarrier_initializer.cu's coefficient bounds);
  - the objective: the sum of the frames' objectives with the shared bounds counted once, and its
    directional derivative along the joint gradient's own sign (a descent direction decreases it);
  - the joint fit: from a perturbed start it lowers the objective and moves the shared block
    towards the truth.
"""
import numpy as np

import pyoracle as O
from scenes import joint_scene


def _scene(**kw):
    sc = joint_scene(**kw)
    return sc, O.JointScene(sc.model, sc.frames, sc.opts)


def test_one_frame_is_the_frame_system():
    sc, js = _scene(crops=(128,))
    th = sc.init()[0]
    H, g, obj = js.normal_equations(js.join(th[None]))
    ne = sc.oracle().normal_equations(0, th)
    assert np.allclose(H, ne["H"], rtol=1e-12, atol=1e-12 * np.abs(ne["H"]).max())
    assert np.allclose(g, ne["g"], rtol=1e-12, atol=1e-12 * np.abs(ne["g"]).max())
    assert abs(obj - ne["obj"][2]) <= 1e-12 * abs(ne["obj"][2])


def test_block_structure():
    sc, js = _scene()
    th = sc.init()
    x = js.join(th)
    assert np.array_equal(js.split(x)[0, :js.Ks], th[0, :js.Ks])
    H, g, obj = js.normal_equations(x)
    assert np.allclose(H, H.T, rtol=0, atol=1e-9 * np.abs(H).max())
    Ks, Kf = js.Ks, js.Kf
    th_j = js.split(x)
    frames = [js.free.normal_equations(t, th_j[t]) for t in range(js.n)]
    ob, gb, hb = js.shared_bounds(th_j[0])
    want = sum(f["H"][:Ks, :Ks] for f in frames) + np.diag(hb)
    assert np.allclose(H[:Ks, :Ks], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    for t in range(js.n):
        sl = slice(Ks + t * Kf, Ks + (t + 1) * Kf)
        assert np.array_equal(H[sl, sl], frames[t]["H"][Ks:, Ks:])
        assert np.array_equal(H[:Ks, sl], frames[t]["H"][:Ks, Ks:])
        assert np.array_equal(g[sl], frames[t]["g"][Ks:])
        for u in range(js.n):
            if u != t:
                assert not H[sl, Ks + u * Kf:Ks + (u + 1) * Kf].any()
    assert abs(obj - (sum(f["obj"][2] for f in frames) + ob)) <= 1e-12 * abs(obj)


def test_shared_bounds_closed_form():
    sc, js = _scene()
    th0 = sc.init()[0]
    Ks = js.Ks
    lb = np.concatenate([sc.model.tex_lb, sc.model.id_lb]).astype(np.float64)
    ub = np.concatenate([sc.model.tex_ub, sc.model.id_ub]).astype(np.float64)
    x = th0[:Ks]
    ob, gb, hb = js.shared_bounds(th0)
    fub, flb = x - ub, lb - x
    assert abs(ob - np.sum(-np.log(-fub) - np.log(-flb))) <= 1e-9 * abs(ob)
    assert np.allclose(gb, -1 / fub + 1 / flb, rtol=1e-9)
    assert np.allclose(hb, (1 / (fub + 1e-12)) ** 2 + (1 / (flb + 1e-12)) ** 2, rtol=1e-9)


def test_objective_and_descent():
    sc, js = _scene()
    x = js.join(sc.init())
    H, g, obj = js.normal_equations(x)
    assert abs(js.objective(x) - obj) <= 1e-9 * abs(obj)
    d = -np.linalg.solve(H, g)
    assert d @ g < 0
    assert js.objective(x + 1e-3 * d) < obj


def test_joint_fit_moves_towards_truth():
    sc, js = _scene(noise=0.0)
    x0 = js.join(sc.init())
    xt = js.join(sc.truth())
    x, performed, stop = js.fit(x0, max_iters=6)
    assert performed >= 1
    assert js.objective(x) < js.objective(x0)
    Ks = js.Ks
    e0 = np.linalg.norm(x0[:Ks] - xt[:Ks])
    e1 = np.linalg.norm(x[:Ks] - xt[:Ks])
    assert e1 < e0
