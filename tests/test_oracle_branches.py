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
