"""Known answers for the line-search branches of fit_3DMM_shape_rigid (model_fitter.cu:121-375),
derived by hand and checked on the FP64 oracle (CPU).  The frames are built so that the branch
is taken for certain (tests/scenes.py, branch_scene):

  nan_observation a NaN pixel in the observation makes g, and so the search direction, NaN: the
                 fit terminates before any trial (206-213): stop = NaN, status TDI_NAN, theta
                 unchanged, one outer iteration;
  face_too_large every trial fails the face-size test (276-282), so the step shrinks by BETA
                 without counting an evaluation (324-327) until t < t_min terminates the fit
                 (241-244): stop = step, no evaluation, theta unchanged, one outer iteration;
  no_pixels      nothing rendered, the light rows of H are zero: the Cholesky fails; the
                 reference ignores a failed potrf (194), the oracle and the library stop the
                 frame with TDI_NOT_PD, theta unchanged;
  decrement      with decrement_tol = +inf the decrement test after the first evaluated trial
                 terminates (336-341): stop = decrement, one evaluation, theta unchanged.
  ordinary frames in the same batch fit as they do alone (one frame's branch changes nothing
                 for the others).
"""
import copy
import importlib

import numpy as np

import pyoracle as O
import scenes
from scenes import make_scene

abi = importlib.import_module("3di_amd.abi")



def _fit(sc, opts=None):
    orc = O.Scene(sc.model, sc.frames, opts or sc.opts)
    th0 = np.asarray(sc.init(), np.float32).astype(np.float64)
    th, reps = orc.fit(th0)
    return th0, th, reps


def test_forced_branches_known_answers():
    sc = scenes.branch_scene()
    th0, th, reps = _fit(sc)
    for i, kind in enumerate(scenes.BRANCH_FRAMES):
        r = reps[i]
        if kind == "face_too_large":
            assert r.stop == 4 and r.ls_evals == 0 and r.iters == 1 and r.status == 0, (kind, r.stop, r.ls_evals)
            assert np.array_equal(th[i], th0[i])
        elif kind == "nan_observation":
            assert r.stop == 8 and r.status == abi.TDI_NAN and r.ls_evals == 0 and r.iters == 1, (r.stop, r.status)
            assert np.array_equal(th[i], th0[i])
        elif kind == "no_pixels":
            assert r.stop == 16 and r.status == abi.TDI_NOT_PD and r.iters == 1 and r.n_pixels == 0
            assert np.array_equal(th[i], th0[i])
        else:
            assert r.stop in (1, 2, 4) and r.ls_evals >= 1 and r.status == 0
            assert not np.array_equal(th[i], th0[i])


def test_ordinary_frames_unaffected_by_forced_neighbours():
    sc = scenes.branch_scene()
    th0, th, _ = _fit(sc)
    normal = [i for i, k in enumerate(scenes.BRANCH_FRAMES) if k == "normal"]
    alone = copy.copy(sc)
    alone.frames = [sc.frames[i] for i in normal]
    alone.truths = [sc.truths[i] for i in normal]
    _, th_a, _ = _fit(alone)
    assert np.array_equal(th[normal], th_a)


def test_decrement_test_on_the_first_trial():
    sc = scenes.branch_scene()
    normal = [i for i, k in enumerate(scenes.BRANCH_FRAMES) if k == "normal"]
    sc.frames = [sc.frames[i] for i in normal]
    sc.truths = [sc.truths[i] for i in normal]
    opts = abi.default_options(**{f[0]: getattr(sc.opts, f[0]) for f in abi.tdi_fit_options._fields_})
    opts.decrement_tol = float("inf")
    th0, th, reps = _fit(sc, opts)
    for i, r in enumerate(reps):
        assert r.stop == 2 and r.ls_evals == 1 and r.iters == 1, (r.stop, r.ls_evals, r.iters)
        assert np.array_equal(th[i], th0[i])


def test_finetune_holds_the_pose_and_widens_the_expression_bounds():
    """fit_3DMM_epsilon_finetune (model_fitter.cu:397-661): trials move every unknown but the pose
    (set_xtmp over Ktotal - 6, 502), so the pose is returned bit for bit, and the expression
    bounds are FINETUNE_COEF times wider (409-410).  A start whose first expression coefficient
    lies at 1.2 times its regular bound (landmark barrier off, so the bounds alone decide
    feasibility) is outside the shape fit's feasible set -- its objective is NaN, no trial is
    ever accepted and theta comes back unchanged -- and inside the finetune fit's, which moves
    the expressions and keeps them within the widened bounds."""
    sc = make_scene(grid=40, k_exp=10, fit_light=0, crops=(96, 96, 96), pose_deg=15.0, noise=0.02, iters=6,
                    seed=7, cfg=43, use_landmarks=0)
    lay = sc.lay
    e0, p = lay["exp"], lay["pose"]
    ub = sc.model.exp_ub.astype(np.float64)
    th0 = np.asarray(sc.init(), np.float32).astype(np.float64)
    th0[:, e0] = np.float32(1.2 * ub[0])
    orc = O.Scene(sc.model, sc.frames, sc.opts)
    th_r, rep_r = orc.fit(th0)
    assert np.array_equal(th_r, th0)
    th, reps = orc.fit(th0, finetune=True)
    assert np.array_equal(th[:, p:p + 6], th0[:, p:p + 6])
    assert all(r.ls_evals > 0 for r in reps)
    assert not np.array_equal(th[:, e0:p], th0[:, e0:p])
    assert np.all(np.abs(th[:, e0:p]) < sc.opts.finetune_coef * ub[:p - e0])
