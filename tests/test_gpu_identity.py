"""The identity stage on the GPU against the FP64 oracle (pyoracle.JointScene):
  - a frame's normal equations above 111 unknowns (albedo + identity + expression + pose + light):
    J^T J and J^T r on MFMA over 64 x 64 tiles with BF16 hi/mid/lo parts, the barrier's rows in
    global scratch; the same bars as the per-frame path's (|dH|/|H| < 2e-5, |dg|/|g| < 3e-5);
  - the joint system of frames sharing the albedo and identity (tdi_normal_equations_joint):
    matrix, gradient and objective, at a size under and a size over 111 unknowns per frame;
    the joint step (FP64 blocked Cholesky of the joint matrix) against the oracle's;
  - the joint fit (tdi_fit_multiframe): the oracle's fit from the same start reaches an objective
    the GPU's matches to 1% and the GPU's shared block is as close to the truth (within 1.5x +
    0.05) as the oracle's; the shared block is written to every frame;
  - the frame selection and learn_identity on a synthetic video (a yaw sweep of one face): the
    selection is the binning of the angles the landmark fits give, and learn_identity's identity
    and albedo equal the average of the same steps made one by one through the Python bindings
    (geometry, masking, crops, orthographic start, landmark fit, the light sub-fits, the joint fit),
    bit for bit.
"""
import importlib
import sys

import numpy as np
import pytest

import pyoracle as O
from scenes import joint_scene, make_scene

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tdi = importlib.import_module("3di_amd.tdi")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_normal_equations_above_111_unknowns():
    _say("case: per-frame normal equations, K = 60 + 60 + 20 + 6 + 9 = 155")
    sc = make_scene(grid=48, k_id=60, k_tex=60, k_exp=20, fit_id=True, fit_tex=True, fit_light=2,
                    crops=(128, 176), pose_deg=15.0)
    assert sc.K == 155
    b = tdi.Batch(tdi.Context(sc.model, sc.opts), sc.frames)
    th = sc.init()
    H, g, obj = b.normal_equations(th)
    H = H.cpu().numpy().astype(np.float64)
    g = g.cpu().numpy().astype(np.float64)
    orc = sc.oracle()
    for i in range(len(sc.frames)):
        r = orc.normal_equations(i, th[i])
        eH = np.linalg.norm(H[i] - r["H"]) / np.linalg.norm(r["H"])
        eg = np.linalg.norm(g[i] - r["g"]) / np.linalg.norm(r["g"])
        _say(f"  frame {i}: {r['npix']} px, |dH|/|H| = {eH:.2e}  |dg|/|g| = {eg:.2e}")
        assert eH < 2e-5 and eg < 3e-5
        assert abs(float(obj[i]) - r["obj"][2]) <= 1e-5 * abs(r["obj"][2])
    b.close()


@pytest.mark.parametrize("kid", [12, 60])
def test_joint_normal_equations_and_step(kid):
    _say(f"case: joint normal equations, K_id = K_tex = {kid}, 3 frames")
    sc = joint_scene(k_id=kid, k_tex=kid, k_exp=10 if kid < 60 else 20)
    js = O.JointScene(sc.model, sc.frames, sc.opts)
    b = tdi.Batch(tdi.Context(sc.model, sc.opts), sc.frames)
    assert b.joint_unknowns() == js.KJ
    th = sc.init().astype(np.float32)
    HJ, gJ, obj = b.joint_normal_equations(th)
    HJ, gJ, obj = HJ.cpu().numpy(), gJ.cpu().numpy(), float(obj.cpu()[0])
    Ho, go, oo = js.normal_equations(js.join(th.astype(np.float64)))
    eH = np.linalg.norm(HJ - Ho) / np.linalg.norm(Ho)
    eg = np.linalg.norm(gJ - go) / np.linalg.norm(go)
    _say(f"  KJ = {js.KJ}: |dH|/|H| = {eH:.2e}  |dg|/|g| = {eg:.2e}  obj {obj:.6f} vs {oo:.6f}")
    assert eH < 2e-5 and eg < 3e-5
    assert abs(obj - oo) <= 1e-5 * abs(oo)
    # blocks of different frames are exactly zero
    Ks, Kf = js.Ks, js.Kf
    assert not HJ[Ks:Ks + Kf, Ks + Kf:].any()
    # the joint step: blocked FP64 Cholesky of the GPU's matrix against the oracle's solve
    d_o = -np.linalg.solve(Ho, go)
    d_g, st = tdi.cholesky_solve(torch.from_numpy(HJ[None].astype(np.float32)).cuda(),
                                 torch.from_numpy(gJ[None].astype(np.float32)).cuda(), 0)
    assert int(st.cpu()[0]) == 0
    d_g = d_g.cpu().numpy()[0].astype(np.float64)
    en = np.sqrt((d_g - d_o) @ Ho @ (d_g - d_o) / (d_o @ Ho @ d_o))
    _say(f"  joint step energy-norm difference {en:.2e}")
    assert en < 1e-2
    b.close()


@pytest.mark.parametrize("kid", [12, 60])
def test_fit_multiframe_vs_oracle(kid):
    _say(f"case: joint fit, K_id = K_tex = {kid}, 3 frames, 6 iterations")
    sc = joint_scene(k_id=kid, k_tex=kid, k_exp=10 if kid < 60 else 20, iters=6)
    js = O.JointScene(sc.model, sc.frames, sc.opts)
    b = tdi.Batch(tdi.Context(sc.model, sc.opts), sc.frames)
    th0 = sc.init().astype(np.float32)
    thg, rep = b.fit_multiframe(th0)
    b.close()
    Ks = js.Ks
    assert all(np.array_equal(thg[t, :Ks], thg[0, :Ks]) for t in range(len(sc.frames)))
    x0 = js.join(th0.astype(np.float64))
    xo, performed, stop = js.fit(x0)
    xg = js.join(thg.astype(np.float64))
    xt = js.join(sc.truth())
    og, oo = js.objective(xg), js.objective(xo)
    eg, eo = np.linalg.norm(xg[:Ks] - xt[:Ks]), np.linalg.norm(xo[:Ks] - xt[:Ks])
    _say(f"  gpu: {rep.iters} iterations, stop {rep.stop}, objective {og:.3f} (reported {rep.objective:.3f}); "
         f"oracle: {performed} iterations, stop {stop}, objective {oo:.3f}; shared error vs truth gpu {eg:.4f} "
         f"oracle {eo:.4f} (start {np.linalg.norm(x0[:Ks] - xt[:Ks]):.4f})")
    assert rep.status == 0
    assert og < oo + 0.01 * abs(oo)
    assert eg < 1.5 * eo + 0.05
    assert abs(rep.objective - og) <= 1e-4 * abs(og)


def _identity_video(n=12, W=320, H=240, f=700.0):
    """n BGR8 frames of one synthetic face (a shared albedo and identity, a yaw sweep of +-35 deg),
    rendered by the oracle over the whole image, their landmarks (1 px noise) and the truths."""
    from scenes import model as scene_model
    synth = importlib.import_module("3di_amd.synth")
    abi = importlib.import_module("3di_amd.abi")
    m = scene_model(48, 12, 12, 10, 11)
    lay = synth.layout(m, True, True, True, 0)
    # a joint fit counts when it runs more than NITERS_SUCCESS iterations (model_fitter.cu:374);
    # the landmark start on the synthetic cap gives poor poses and lights, from which the joint
    # fits stop after a few steps, so the threshold is 0 here: every group's fit counts and the
    # average is the one to compare
    opts = abi.default_options(max_iters=20, fit_light=0, fit_id=1, fit_tex=1, niters_success=0)
    Ks = lay["exp"]
    rng0 = np.random.default_rng(77)
    shared = np.clip(rng0.standard_normal(Ks), -3, 3)
    frames, lx, ly, truths = [], [], [], []
    for i in range(n):
        rng = np.random.default_rng(1000 + i)
        yaw = -35.0 + 70.0 * i / max(1, n - 1)
        pose = np.concatenate([[rng.uniform(-5, 5), rng.uniform(-5, 5), 480.0],
                               synth.rotvec_from_euler(np.radians(yaw), np.radians(rng.uniform(-8, 8)), 0.0)])
        tr = synth.draw_truth(m, rng, lay, 0.0, pose=pose)
        tr.theta[:Ks] = shared
        fr = synth.make_frame(m, tr, lay, rng, f, W, H)
        raw = O.Scene(m, [fr], opts).render(0, tr.theta)["raw"]
        grey = np.clip(np.rint(raw.reshape(W, H).T * 255.0), 0, 255).astype(np.uint8)
        frames.append(np.repeat(grey[..., None], 3, axis=2))
        lx.append(fr.lmk_x)
        ly.append(fr.lmk_y)
        truths.append(tr)
    cam = np.array([f, f, W / 2.0, H / 2.0], np.float32)
    return m, lay, opts, np.stack(frames), np.stack(lx), np.stack(ly), cam, truths


def test_select_and_learn_identity():
    _say("case: identity frame selection and learn_identity, 12 frames, groups of 3")
    video = importlib.import_module("3di_amd.video")
    m, lay, opts, frames, lx, ly, cam, truths = _identity_video()
    n, H, W = frames.shape[:3]
    ctx = tdi.Context(m, opts)
    p = video.identity_params(nframes=3, nframes_per_angle_bin=3, nperms=2, ntot_reconstrs=2, crop_w=128,
                              crop_h=128)
    sel, ang, coef = video.select_identity_frames(ctx, p, lx, ly, W, H, cam)
    _say(f"  resize coefficient {coef:.4f}, yaw (deg) {np.round(np.degrees(ang[:, 0]), 1).tolist()}, selected {sel.tolist()}")
    assert sel.size == 6
    # the kept groups are groups of the binning, in its order (a group is kept when the joint
    # landmark-only fit of its frames succeeds; that fit is checked against the oracle in
    # test_gpu_lmk_joint.py)
    every = video.bin_identity_frames(video.identity_params(nframes=3, nframes_per_angle_bin=3, nperms=2,
                                                            ntot_reconstrs=1000, crop_w=128, crop_h=128),
                                      ang[:, 0], ang[:, 1])
    groups = [tuple(every[i:i + 3]) for i in range(0, every.size, 3)]
    kept = [tuple(sel[i:i + 3]) for i in range(0, sel.size, 3)]
    pos = [groups.index(gk) for gk in kept]
    assert pos == sorted(pos)
    assert np.isfinite(ang[sel]).all()
    dframes = torch.from_numpy(np.ascontiguousarray(frames[sel])).cuda()
    alpha, beta, used = video.learn_identity(ctx, p, dframes, lx[sel], ly[sel], cam)
    _say(f"  groups used {used}")
    # the same steps one by one through the Python bindings, on the frames learn_identity masked
    synth = importlib.import_module("3di_amd.synth")
    K, Ks, off_id = lay["K"], lay["exp"], lay["id"]
    sums_a, sums_b, used2 = None, None, 0
    for g in range(2):
        idx = np.arange(3 * g, 3 * g + 3)
        geo = video.prepare_frames(dframes[torch.from_numpy(idx).cuda()].contiguous(), lx[sel][idx], ly[sel][idx],
                                   np.tile(cam, (3, 1)), p.ref_face_size, pad=True, crop_size=(128, 128))
        assert (geo["status"] == 0).all()
        obs = geo["obs"].cpu().numpy()
        fds, off = [], 0
        for i in range(3):
            w_, h_ = int(geo["crop"][i, 2]), int(geo["crop"][i, 3])
            fds.append(synth.Frame(phix=geo["cams"][i, 0], phiy=geo["cams"][i, 1], cx=geo["cams"][i, 2],
                                   cy=geo["cams"][i, 3], x0=int(geo["crop"][i, 0]), y0=int(geo["crop"][i, 1]),
                                   width=w_, height=h_, face_size=float(geo["face_size"][i]), obs=obs[off:off + w_ * h_],
                                   lmk_x=geo["lmk_x"][i], lmk_y=geo["lmk_y"][i],
                                   box_lx=np.full(51, -0.15, np.float32), box_ux=np.full(51, 0.15, np.float32),
                                   box_ly=np.full(51, -0.15, np.float32), box_uy=np.full(51, 0.15, np.float32),
                                   lam=synth.LAMBDA0, Lint=0.005))
            off += w_ * h_
        b = tdi.Batch(ctx, fds)
        pose, _, _ = video.initialize_with_orthographic(ctx, geo["lmk_x"], geo["lmk_y"], geo["cams"])
        th = np.zeros((3, K), np.float32)
        th[:, lay["pose"]:lay["pose"] + 6] = pose
        kl = min(p.lmk_k_exp, lay["pose"] - lay["exp"])
        nid = lay["exp"] - lay["id"]
        th, lrep = b.fit_landmarks_joint(th, min(p.lmk_k_id, nid), kl, True)
        if lrep.status != 0:
            b.close()
            _say(f"  group {g}: joint landmark fit with the slack failed (slack {lrep.slack:.2f})")
            continue
        th, _ = b.fit_landmarks_joint(th, nid, kl, False)
        lam = np.tile(synth.LAMBDA0, (3, 1)).astype(np.float32)
        Li = np.full(3, 0.005, np.float32)
        b.set_light(lam, Li)
        for _ in range(2):
            lam, Li, _ = b.fit_lintensity(th, lam, Li)
            b.set_light(lam, Li)
            lam, Li, _ = b.fit_lambdas(th, lam, Li)
            b.set_light(lam, Li)
        thj, rep = b.fit_multiframe(th)
        b.close()
        _say(f"  group {g}: landmark fit slack {lrep.slack:.2f}, light {Li.tolist()}, joint fit {rep.iters} "
             f"iterations, stop {rep.stop}, status {rep.status}, success {rep.success}")
        if rep.success and rep.status == 0:
            used2 += 1
            a_, b_ = thj[0, off_id:Ks].copy(), thj[0, :off_id].copy()
            sums_a = a_ if sums_a is None else sums_a + a_
            sums_b = b_ if sums_b is None else sums_b + b_
    ctx.close()
    assert used >= 1
    assert used2 == used
    w = np.float32(1.0) / np.float32(used)
    assert np.array_equal(alpha, sums_a * w) and np.array_equal(beta, sums_b * w)
