"""The line-search branches forced frame by frame (tests/scenes.py branch_scene), through the C
ABI on the GPU, batched and plain, against the oracle's known answers (test_oracle_branches.py):
a NaN observation stops with TDI_STOP_NAN, a crop with nothing rendered with TDI_STOP_NOT_PD, a
face too large for every trial with TDI_STOP_STEP and no evaluation, each with theta unchanged;
the ordinary frames of the same launch take the oracle's Gauss-Newton step (one iteration,
energy norm 1e-3, with at most one frame a feasibility shrink apart) and fit as they do in a
launch without the forced frames."""
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


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


EXPECT = {"nan_observation": (8, abi.TDI_NAN), "no_pixels": (16, abi.TDI_NOT_PD), "face_too_large": (4, abi.TDI_OK)}


def _check_forced(sc, th0, th, reps, label):
    for i, kind in enumerate(scenes.BRANCH_FRAMES):
        r = reps[i]
        _say(f"  {label} frame {i} ({kind}): stop {r.stop} status {r.status} iters {r.iters} ls {r.ls_evals}")
        if kind in EXPECT:
            stop, status = EXPECT[kind]
            assert (r.stop, r.status, r.iters) == (stop, status, 1), kind
            assert r.ls_evals == 0
            assert np.array_equal(th[i].astype(np.float32), th0[i].astype(np.float32)), kind
        else:
            assert r.status == abi.TDI_OK and r.ls_evals >= 1


@pytest.mark.parametrize("path", ["batched", "plain"])
def test_forced_branches(path):
    _say(f"case: forced branches ({path})")
    sc = scenes.branch_scene(iters=1)
    th0 = np.asarray(sc.init(), np.float32)
    ctx = tdi.Context(sc.model, sc.opts)
    b = tdi.Batch(ctx, sc.frames)
    if path == "batched":
        thg, reps = b.fit_batched(torch.from_numpy(th0.copy()).cuda())
        thg = thg.cpu().numpy()
    else:
        thg, reps = b.fit_plain(th0.copy())
    b.close()
    ctx.close()
    _check_forced(sc, th0, thg, reps, path)
    # the ordinary frames take the oracle's step
    tho, repo = O.Scene(sc.model, sc.frames, sc.opts).fit(th0.astype(np.float64))
    orc = sc.oracle()
    flips = 0
    for i, kind in enumerate(scenes.BRANCH_FRAMES):
        if kind != "normal":
            assert (repo[i].stop, repo[i].status) == EXPECT[kind]
            continue
        assert reps[i].ls_evals == repo[i].ls_evals
        Hm = orc.normal_equations(i, th0[i].astype(np.float64))["H"]
        s_, o_ = thg[i].astype(np.float64) - th0[i], tho[i] - th0[i]
        lam = float(s_ @ Hm @ o_) / float(o_ @ Hm @ o_)
        if abs(lam - 1.0) > 1e-2:
            assert min(abs(lam - 0.8), abs(lam - 1.25)) < 1e-2
            flips += 1
            s_ = s_ / lam
        d = s_ - o_
        en = np.sqrt(max(d @ Hm @ d, 0.0)) / np.sqrt(o_ @ Hm @ o_)
        _say(f"  frame {i}: step length ratio {lam:.4f}, energy-norm rel {en:.1e}")
        assert en < 1e-3
    assert flips <= 1


def test_ordinary_frames_unaffected_by_forced_neighbours():
    """The ordinary frames of the launch with forced frames against a launch of the ordinary
    frames alone: the same stop reasons and evaluation counts, parameters within 1e-5 (the
    per-frame work is independent; only the objective partial sums may be split differently
    when the launch has fewer frames)."""
    _say("case: forced neighbours do not disturb the others")
    sc = scenes.branch_scene(iters=3)
    normal = [i for i, k in enumerate(scenes.BRANCH_FRAMES) if k == "normal"]
    th0 = np.asarray(sc.init(), np.float32)
    ctx = tdi.Context(sc.model, sc.opts)
    b = tdi.Batch(ctx, sc.frames)
    th_all, rep_all = b.fit_batched(torch.from_numpy(th0.copy()).cuda())
    b.close()
    b = tdi.Batch(ctx, [sc.frames[i] for i in normal])
    th_n, rep_n = b.fit_batched(torch.from_numpy(th0[normal].copy()).cuda())
    b.close()
    ctx.close()
    th_all, th_n = th_all.cpu().numpy(), th_n.cpu().numpy()
    for j, i in enumerate(normal):
        d = np.abs(th_all[i].astype(np.float64) - th_n[j]).max()
        _say(f"  frame {i}: stop {rep_all[i].stop}/{rep_n[j].stop}, ls {rep_all[i].ls_evals}/{rep_n[j].ls_evals}, "
             f"max |dtheta| {d:.1e}")
        assert (rep_all[i].stop, rep_all[i].ls_evals) == (rep_n[j].stop, rep_n[j].ls_evals)
        assert d < 1e-5


@pytest.mark.parametrize("path", ["batched", "plain"])
def test_finetune_vs_oracle(path):
    """fit_3DMM_epsilon_finetune on the GPU from a start outside the regular expression bounds
    and inside the widened ones (as test_oracle_branches): the pose comes back bit for bit, the
    first iteration takes the oracle's step (energy norm 1e-3) with the same evaluation count,
    and the regular shape fit leaves theta unchanged, as the oracle's does."""
    _say(f"case: expression finetune ({path})")
    sc = scenes.make_scene(grid=40, k_exp=10, fit_light=0, crops=(96, 120, 80), pose_deg=15.0, noise=0.02, iters=1,
                           seed=7, cfg=43, use_landmarks=0)
    lay = sc.lay
    e0, p = lay["exp"], lay["pose"]
    th0 = np.asarray(sc.init(), np.float32)
    th0[:, e0] = np.float32(1.2 * sc.model.exp_ub[0])
    ctx = tdi.Context(sc.model, sc.opts)
    b = tdi.Batch(ctx, sc.frames)

    def run(finetune):
        if path == "batched":
            t, r = b.fit_batched(torch.from_numpy(th0.copy()).cuda(), finetune=finetune)
            return t.cpu().numpy().astype(np.float64), r
        t, r = b.fit_plain(th0.copy(), finetune=finetune)
        return t.astype(np.float64), r

    th_r, _ = run(False)
    thg, repg = run(True)
    b.close()
    ctx.close()
    assert np.array_equal(th_r, th0.astype(np.float64))
    assert np.array_equal(thg[:, p:p + 6], th0[:, p:p + 6].astype(np.float64))
    orc = sc.oracle()
    tho, repo = orc.fit(th0.astype(np.float64), finetune=True)
    assert [r.ls_evals for r in repg] == [r.ls_evals for r in repo]
    for i in range(len(sc.frames)):
        Hm = orc.normal_equations(i, th0[i].astype(np.float64))["H"]
        s_, o_ = thg[i] - th0[i], tho[i] - th0[i]
        d = s_ - o_
        en = np.sqrt(max(d @ Hm @ d, 0.0)) / np.sqrt(o_ @ Hm @ o_)
        _say(f"  frame {i}: ls {repg[i].ls_evals}, energy-norm rel {en:.1e}")
        assert en < 1e-3
