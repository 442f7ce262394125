"""The per-frame stage of a video through the host layer (tdi_fit_video_frames_nres) on the GPU:
landmark-driven starting poses, NRES_COEFS reference face sizes (update_shape_single_ref_size,
video_fitter.cu:355-393), each with the frames resized, greyed and cropped on the device, the
light + shape fits of fit_to_video_frame and the expression finetune, the results added to a
VideoOutput.  Checked against the same steps made one by one through the Python bindings (bit for
bit: the driver adds nothing of its own), the VideoOutput files against the per-size results
averaged by hand, and the fitted poses against the synthetic truth."""
import importlib
import os
import sys

import numpy as np
import pytest

import pyoracle as O
from scenes import model as scene_model

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tdi = importlib.import_module("3di_amd.tdi")
abi = importlib.import_module("3di_amd.abi")
synth = importlib.import_module("3di_amd.synth")
video = importlib.import_module("3di_amd.video")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _video(n=3, W=320, H=240, f=700.0):
    """n BGR8 frames of the synthetic face rendered by the oracle over the whole image, their
    landmarks (1 px noise) and cameras, and the truths."""
    m = scene_model(40, 0, 0, 10, 7)
    lay = synth.layout(m, False, False, True, 0)
    opts = abi.default_options(max_iters=6, fit_light=0)
    frames, lx, ly, truths = [], [], [], []
    for i in range(n):
        rng = np.random.default_rng(synth.frame_seed(45, i))
        tr = synth.draw_truth(m, rng, lay, 12.0)
        fr = synth.make_frame(m, tr, lay, rng, f, W, H)
        raw = O.Scene(m, [fr], opts).render(0, tr.theta)["raw"]
        grey = np.clip(np.rint(raw.reshape(W, H).T * 255.0), 0, 255).astype(np.uint8)   # [H, W]
        frames.append(np.repeat(grey[..., None], 3, axis=2))
        lx.append(fr.lmk_x)
        ly.append(fr.lmk_y)
        truths.append(tr)
    cams = np.tile(np.array([[f, f, W / 2.0, H / 2.0]], np.float32), (n, 1))
    return m, lay, opts, np.stack(frames), np.stack(lx), np.stack(ly), cams, truths


def _steps(ctx, dframes, lx, ly, cams, th_start, lam0, sizes, finetune, init=None):
    """The stage's steps one by one through the Python bindings; returns the results per size.
    init: None (start from theta), "ortho" (the orthographic pose of each size's landmarks) or
    "slack" (that pose, then the landmark-only fit with slack; frames it fails keep theta)."""
    n = len(th_start)
    t = th_start.copy()
    la = lam0.copy()
    per_size = []
    for r in sizes:
        g = video.prepare_frames(dframes, lx, ly, cams, r, pad=True, crop_size=(128, 128))
        assert (g["status"] == 0).all()
        fds = []
        off = 0
        obs = g["obs"].cpu().numpy()
        for i in range(n):
            w_, h_ = int(g["crop"][i, 2]), int(g["crop"][i, 3])
            fds.append(synth.Frame(phix=g["cams"][i, 0], phiy=g["cams"][i, 1], cx=g["cams"][i, 2], cy=g["cams"][i, 3],
                                   x0=int(g["crop"][i, 0]), y0=int(g["crop"][i, 1]), width=w_, height=h_,
                                   face_size=float(g["face_size"][i]), obs=obs[off:off + w_ * h_],
                                   lmk_x=g["lmk_x"][i], lmk_y=g["lmk_y"][i],
                                   box_lx=np.full(51, -0.15, np.float32), box_ux=np.full(51, 0.15, np.float32),
                                   box_ly=np.full(51, -0.15, np.float32), box_uy=np.full(51, 0.15, np.float32),
                                   lam=la[i], Lint=0.005))
            off += w_ * h_
        b = tdi.Batch(ctx, fds)
        if init is not None:
            K = t.shape[1]
            pose, _, _ = video.initialize_with_orthographic(ctx, g["lmk_x"], g["lmk_y"], g["cams"])
            tt = t.copy()
            tt[:, K - 6:] = pose
            if init == "ortho":
                t = tt
            else:
                dt, lreps = b.fit_landmarks_slack(torch.from_numpy(tt).cuda(), min(79, K - 6))
                dt = dt.cpu().numpy()
                _say(f"  landmark fit: status {[r.status for r in lreps]}, steps {[r.iterations for r in lreps]}")
                for i, r in enumerate(lreps):
                    if r.status == 0:
                        t[i] = dt[i]
        t, la, L_, _ = video.fit_video_frames(b, t, la, twostage_illum=False, batched=True)
        if finetune:
            dt = torch.from_numpy(t.copy()).cuda()
            dt, _ = b.fit_batched(dt, finetune=True)
            t = dt.cpu().numpy()
        b.close()
        per_size.append((t.copy(), la.copy(), L_.copy()))
    return t, la, per_size


def test_video_stage_matches_its_steps(tmp_path):
    _say("case: per-frame video stage, NRES sizes")
    m, lay, opts, frames, lx, ly, cams, truths = _video()
    n, K = len(truths), lay["K"]
    ctx = tdi.Context(m, opts)
    dframes = torch.from_numpy(frames).cuda()
    lam0 = np.tile(synth.LAMBDA0, (n, 1)).astype(np.float32)
    sizes = (86.0, 86.0 + 5.0 + 5.0)
    # from perturbed truths (the fit's own start), two sizes, the expression finetune
    params = video.video_params(nres_coefs=2, twostage_illum=0, finetune_expressions=1, batched=1, init_pose=0,
                                crop_w=128, crop_h=128, ref_face_size=86.0)
    th0 = np.stack([tr.init for tr in truths]).astype(np.float32)
    vo = video.VideoOutput(K - 6, 2)
    th, lam, Li, st, reps = video.fit_video_frames_nres(ctx, params, dframes, lx, ly, cams, th0, lam0, vo=vo)
    _say(f"  status {st.tolist()}, stops {[r.stop for r in reps]}")
    assert (st == 0).all()
    # the frames were painted in place first (PAINT_INNERMOUTH_BLACK, default 1), as the oracle paints them
    assert params.paint_innermouth_black == 1
    assert np.array_equal(dframes.cpu().numpy(), O.paint_innermouth_black(frames, lx, ly, 86.0))
    t, la, per_size = _steps(ctx, dframes, lx, ly, cams, th0, lam0, sizes, True)
    assert np.array_equal(th, t) and np.array_equal(lam, la)
    # VideoOutput: expressions averaged over the two sizes (frames t < max(frame id) written)
    vo.save_expressions(str(tmp_path / "e.txt"))
    rows = np.loadtxt(tmp_path / "e.txt", ndmin=2)
    want = np.mean([ps[0][:, :K - 6].astype(np.float64) for ps in per_size], axis=0)
    assert rows.shape[0] == n - 1
    assert np.abs(rows - want[:n - 1]).max() < 1e-5 * max(1.0, np.abs(want).max())
    # the fitted poses stay near the synthetic truth (faces of ~90 px after resizing, light fitted
    # by the sub-fits: a few mm and degrees; a broken sequence lands far away)
    for i, tr in enumerate(truths):
        dpos = np.abs(th[i, K - 6:K - 3] - tr.theta[K - 6:K - 3]).max()
        drot = np.degrees(np.abs(th[i, K - 3:] - tr.theta[K - 3:]).max())
        _say(f"  frame {i}: |dtau| {dpos:.2f} mm, |du| {drot:.3f} deg")
        assert dpos < 20.0 and drot < 10.0
    # with the landmark-driven start: at each size the orthographic pose of the resized landmarks,
    # then the landmark-only fit with slack (the synthetic landmarks lie on a smooth cap, so the
    # orthographic start is far from the truth in pitch; only the orchestration is compared)
    params.init_pose = 1
    params.nres_coefs = 2
    params.finetune_expressions = 0
    assert params.lmk_k_exp == 79
    th1, _, _, st1, _ = video.fit_video_frames_nres(ctx, params, dframes, lx, ly, cams, th0, lam0)
    t1, _, _ = _steps(ctx, dframes, lx, ly, cams, th0, lam0, sizes, False, init="slack")
    assert np.array_equal(th1, t1)
    # without the landmark fit: the orthographic pose alone
    params.lmk_k_exp = 0
    params.nres_coefs = 1
    th2, _, _, _, _ = video.fit_video_frames_nres(ctx, params, dframes, lx, ly, cams, th0, lam0)
    t2, _, _ = _steps(ctx, dframes, lx, ly, cams, th0, lam0, sizes[:1], False, init="ortho")
    ctx.close()
    assert np.array_equal(th2, t2)


def test_long_video_chunks_and_resume(tmp_path):
    _say("case: long-video driver, 6 frames in chunks of 2, stop and resume")
    m, lay, opts, frames, lx, ly, cams, truths = _video(n=6)
    n, K = len(truths), lay["K"]
    H, W = frames.shape[1:3]
    ctx = tdi.Context(m, opts)
    dframes = torch.from_numpy(frames).cuda()

    def source(first, count):
        return dframes[first:first + count]

    params = video.video_params(nres_coefs=2, twostage_illum=0, batched=1, crop_w=128, crop_h=128,
                                ref_face_size=86.0)
    th0 = truths[0].init.astype(np.float32)
    lam0 = np.asarray(synth.LAMBDA0, np.float32)
    # uninterrupted
    vo_a = video.VideoOutput(K - 6, 2)
    th_a, la_a, L_a, st_a, done = video.fit_video_long(ctx, params, n, W, H, source, lx, ly, cams, th0, lam0, vo_a,
                                                       chunk_frames=2, state_path=str(tmp_path / "a.state"))
    _say(f"  uninterrupted: status {st_a.tolist()}, frames done {done}")
    assert done == n and (st_a == 0).all()
    # one chunk, then a resumed run with a fresh VideoOutput
    vo_b = video.VideoOutput(K - 6, 2)
    _, _, _, st_b, done_b = video.fit_video_long(ctx, params, n, W, H, source, lx, ly, cams, th0, lam0, vo_b,
                                                 chunk_frames=2, max_chunks=1, state_path=str(tmp_path / "b.state"))
    assert done_b == 2 and (st_b[:2] == 0).all() and (st_b[2:] == -1).all()
    vo_c = video.VideoOutput(K - 6, 2)
    th_c, la_c, L_c, st_c, done_c = video.fit_video_long(ctx, params, n, W, H, source, lx, ly, cams, th0, lam0, vo_c,
                                                         chunk_frames=2, state_path=str(tmp_path / "b.state"),
                                                         resume=True)
    assert done_c == n and (st_c[:2] == -1).all() and (st_c[2:] == 0).all()
    assert np.array_equal(th_c, th_a) and np.array_equal(la_c, la_a) and L_c == L_a
    # the chunks are tdi_fit_video_frames_nres calls from the carried start
    vo_d = video.VideoOutput(K - 6, 2)
    carry = (th0.copy(), lam0.copy(), 0.005)
    for c in range(0, n, 2):
        sl = slice(c, c + 2)
        th, la, Li, st, _ = video.fit_video_frames_nres(
            ctx, params, dframes[sl], lx[sl], ly[sl], cams[sl], np.tile(carry[0], (2, 1)), np.tile(carry[1], (2, 1)),
            frame_ids=np.arange(c, c + 2), vo=vo_d, Lint=np.full(2, carry[2], np.float32))
        j = max(i for i in range(2) if st[i] == 0)
        carry = (th[j].copy(), la[j].copy(), float(Li[j]))
    assert np.array_equal(carry[0], th_a)
    for v, tag in ((vo_a, "a"), (vo_c, "c"), (vo_d, "d")):
        v.save_expressions_all(str(tmp_path / f"{tag}.ea"))
        v.save_poses(str(tmp_path / f"{tag}.p"))
        v.save_illums(str(tmp_path / f"{tag}.i"))
    for ext in ("ea", "p", "i"):
        a = (tmp_path / f"a.{ext}").read_bytes()
        assert a == (tmp_path / f"c.{ext}").read_bytes() and a == (tmp_path / f"d.{ext}").read_bytes()
    ctx.close()
