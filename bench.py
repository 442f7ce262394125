#!/usr/bin/env python3
"""Benchmark of the batched mixed-precision per-frame 3DMM fit on one MI355X.

One step = a complete fit (12 Gauss-Newton iterations with Armijo line search) of every
frame of the cfg2 workload (4096 synthetic frames, crops U{96..384}, 89 unknowns per frame),
inputs resident in HBM, through the batched path (device-resident loop in a hipGraph).

Also measured in the same run:
  * the plain path (one frame at a time, host-driven loop, FP32 J^T J, FP64 Cholesky) on a
    prefix of the same frames -> frames/s and the speed-up of the batched path;
  * the FP64 OpenMP oracle on a prefix of the same frames on the host cores (cpu_baseline),
    which also gives the batched path's parameter error against FP64;
  * per-stage times with HIP events on the first workspace chunk -> the dominant kernel's
    roofline entry;
  * the landmark-driven initialisation (tdi_ortho_init) of every frame;
  * the frame preparation (tdi_prepare_crops): 8-bit BGR 1280x720 frames in HBM -> the
    workload's grey crops, against the oracle on a prefix (time and bit-exactness).

Usage: python bench.py [--steps K] [--warmup W] [--frames N] [--gpus 1]
Prints one JSON line.
"""
from __future__ import annotations

import argparse
import importlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "frames/s per-frame fit, batched+mixed vs plain, at <=1e-4 param RMS vs FP64"
PEAK_HBM_GBS = 8000.0
PEAK_TFLOPS = {"bf16": 2500.0, "fp64": 78.6, "fp32": 157.3}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def stage_work(w, K, npix, n):
    """Algorithmic work of each timed stage over the first workspace chunk.

    Returns a list of (name, bound, amount, unit) with amount in bytes or FLOPs."""
    m = w.model
    V, NT, Ke = m.n_vertices, m.n_triangles, m.k_exp
    A = np.array([f.width * f.height for f in w.frames[:n]], np.float64)
    P = npix[:n].astype(np.float64)
    synth_flops = 2.0 * 3 * V * Ke * n
    raster_bytes = float(np.sum(8 * A + 8 * A + 4 * A + NT * (12 + 3 * 16) + 8 * P + 12 * P))
    shade_bytes = float(np.sum(4 * A + P * (4 + 8 + 8 + 3 * 12 + 3 * 4 + 4) + 16 * A + P * (4 + 4 * 4 + 8 + 16)))
    barrier_flops = 2.0 * 51 * 3 * Ke * n
    # dI at the HBM level: per rendered pixel its vertex ids, barycentrics and image gradient
    # (16 + 8 + 16 B) and the K-float row written; per frame the vertex positions once; per
    # launch the expression basis once (it stays in the caches).  The 9 basis rows each pixel
    # gathers (9 * Ke * 4 B) go through L2/MALL and are reported apart as gather_GBs.
    dI_bytes = float(np.sum(P * (16 + 8 + 16 + 4 * K)) + n * 12.0 * V + 12.0 * V * Ke)
    dI_gather = float(np.sum(P * 9 * Ke * 4))
    jtj_flops = float(np.sum(4.0 * K * K * P))
    chol_flops = n * K ** 3 / 3.0
    return [("synthesis", "mfma", synth_flops, "fp64"), ("raster", "hbm", raster_bytes, None),
            ("shade+blur+grads", "hbm", shade_bytes, None), ("barrier", "mfma", barrier_flops, "fp32"),
            ("jacobian dI", "hbm", dI_bytes, None, dI_gather), ("JtJ (BF16 split MFMA)", "mfma", jtj_flops, "bf16"),
            ("cholesky", "mfma", chol_flops, "fp32")]


def _per_frame_rms(a, b, sl):
    d = np.asarray(a, np.float64)[:, sl] - np.asarray(b, np.float64)[:, sl]
    return np.sqrt(np.mean(d * d, axis=1))


def sample_accuracy(w, O, sc, init, nc, th_o, ro, th_batched, plain, threads, tdi, ctx):
    """Parameter agreement with the FP64 oracle on the first nc frames.

    A whole 12-iteration fit is a chain of discrete decisions (Armijo accept/reject, stop
    tests, pixels entering or leaving the rendered set), so two correct implementations
    agree to round-off only while every decision agrees.  Reported side by side:
      * whole fit: batched (and plain) GPU vs oracle, per-frame RMS of pose + expression;
      * floor: the oracle against itself started 1e-6 away -- the agreement any
        implementation can reach on these frames;
      * one iteration: the first GN step (render, J^T J, Cholesky, line search) of the
        batched mixed path vs the oracle's, from the same FP32 start;
      * visibility: triangle-id buffers of the GPU raster vs the oracle's at the start point.
    """
    import torch
    abi = importlib.import_module("3di_amd.abi")
    lay = w.lay
    sl = slice(lay["exp"], lay["pose"] + 6)
    th0 = init[:nc].astype(np.float64)
    acc = {"frames": nc}
    per = _per_frame_rms(th_batched[:nc], th_o, sl)
    acc["whole_fit"] = dict(rms=float(np.sqrt(np.mean(per ** 2))), median_frame_rms=float(np.median(per)),
                            frames_within_1e4=int(np.sum(per <= 1e-4)))
    th_p, _ = sc.fit(th0 + 1e-6, n_threads=threads)
    fl = _per_frame_rms(th_p, th_o, sl)
    acc["oracle_floor"] = dict(median_frame_rms=float(np.median(fl)), frames_within_1e4=int(np.sum(fl <= 1e-4)))
    if plain is not None:
        npc = min(nc, plain["frames"])
        pp = _per_frame_rms(plain["theta"][:npc], th_o[:npc], sl)
        acc["plain_whole_fit"] = dict(frames=int(npc), median_frame_rms=float(np.median(pp)),
                                      frames_within_1e4=int(np.sum(pp <= 1e-4)))
    # one GN iteration from the same start
    o1 = abi.default_options(**{f[0]: getattr(w.opts, f[0]) for f in abi.tdi_fit_options._fields_})
    o1.max_iters = 1
    sc1 = O.Scene(w.model, w.frames[:nc], o1)
    th_o1, _ = sc1.fit(th0, n_threads=threads)
    c1 = tdi.Context(w.model, o1)
    b1 = tdi.Batch(c1, w.frames[:nc])
    g1 = torch.from_numpy(init[:nc].astype(np.float32)).cuda()
    b1.fit_batched(g1)
    g1 = g1.cpu().numpy()
    step = _per_frame_rms(th_o1, th0, sl)
    one = _per_frame_rms(g1, th_o1, sl)
    # the same difference in the energy norm of the oracle's H at the start, relative to the
    # step: it weighs each direction by what the images determine
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=threads) as ex:
        Hs = list(ex.map(lambda i: sc.normal_equations(i, th0[i])["H"], range(nc)))
    en = []
    for i in range(nc):
        d = g1[i].astype(np.float64) - th_o1[i]
        st = th_o1[i] - th0[i]
        en.append(float(np.sqrt(max(d @ Hs[i] @ d, 0.0)) / max(np.sqrt(max(st @ Hs[i] @ st, 0.0)), 1e-30)))
    en = np.array(en)
    acc["one_iteration"] = dict(median_step_rms=float(np.median(step)), median_frame_rms=float(np.median(one)),
                                max_frame_rms=float(one.max()), frames_within_1e4=int(np.sum(one <= 1e-4)),
                                median_energy_rel=float(np.median(en)), max_energy_rel=float(en.max()))
    # visibility buffers at the start point
    xyz, xyp = b1.synthesize(init[:nc])
    keys, _ = b1.rasterize(xyz, xyp)
    keys = keys.cpu().numpy().view(np.uint64)
    exact = 0
    for i in range(nc):
        r = sc.render(i, th0[i])
        exact += int(np.array_equal(keys[b1.offsets[i]:b1.offsets[i + 1]], r["keys"]))
    acc["visibility_bitexact_frames"] = exact
    b1.close()
    c1.close()
    stops = {}
    for r in ro:
        stops[int(r.stop)] = stops.get(int(r.stop), 0) + 1
    acc["oracle_stop_flags"] = stops
    log(f"[bench] accuracy on {nc} frames: whole fit {acc['whole_fit']['frames_within_1e4']}/{nc} within 1e-4 "
        f"(oracle floor {acc['oracle_floor']['frames_within_1e4']}/{nc}), one iteration "
        f"{acc['one_iteration']['frames_within_1e4']}/{nc} (median {acc['one_iteration']['median_frame_rms']:.1e}, "
        f"energy-norm median {acc['one_iteration']['median_energy_rel']:.1e}), "
        f"visibility bit-exact {exact}/{nc}")
    return acc


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--frames", type=int, default=4096)
    ap.add_argument("--plain-frames", type=int, default=64)
    ap.add_argument("--cpu-frames", type=int, default=32)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a 1 GiB device copy (reference kernel for scaling rocprof FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--traffic-json", default=os.path.join(ROOT, "profiles", "traffic.json"),
                    help="HBM bytes per launch of the dominant kernel, measured with rocprof counters")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the fitting step shards by frames with no exchange, "
              "so --gpus N > 1 is not run here; frames are independent, and more GPUs would each fit a "
              "contiguous range of the frames with their own model copy.", file=sys.stderr)
        sys.exit(2)

    import torch
    tdi = importlib.import_module("3di_amd.tdi")
    workloads = importlib.import_module("3di_amd.workloads")

    w = workloads.cfg2(n_frames=args.frames, log=log)
    K = w.lay["K"]
    init = w.init()

    ctx = tdi.Context(w.model, w.opts)
    batch = tdi.Batch(ctx, w.frames)
    th_init = torch.from_numpy(init).cuda()
    th = th_init.clone()
    log(f"[bench] workspace chunk = {batch.workspace_frames()} frames")

    for i in range(args.warmup):
        t0 = time.perf_counter()
        th.copy_(th_init)
        _, reps = batch.fit_batched(th)
        torch.cuda.synchronize()
        log(f"[bench] warmup {i}: {time.perf_counter() - t0:.2f}s")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        th.copy_(th_init)
        _, reps = batch.fit_batched(th)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    ms_per_step = 1e3 * dt / args.steps
    fps = args.frames * args.steps / dt
    log(f"[bench] batched: {ms_per_step:.1f} ms/step, {fps:.1f} frames/s")
    th_batched = th.cpu().numpy().astype(np.float64)
    iters = np.array([r.iters for r in reps])

    # plain path on a prefix of the same frames
    plain = None
    if not args.no_plain and args.plain_frames > 0:
        npf = min(args.plain_frames, args.frames)
        ctxp = tdi.Context(w.model, w.opts)
        bp = tdi.Batch(ctxp, w.frames[:npf])
        t0 = time.perf_counter()
        th_plain, rp = bp.fit_plain(init[:npf])
        dtp = time.perf_counter() - t0
        plain = dict(frames=npf, seconds=dtp, fps=npf / dtp, theta=th_plain)
        log(f"[bench] plain: {npf} frames in {dtp:.2f}s = {npf / dtp:.2f} frames/s")
        bp.close()
        ctxp.close()

    # stage times on the first chunk (live HIP events on the batch's stream)
    n_chunk = min(args.frames, batch.workspace_frames())
    ms = batch.stage_times(th_init, mixed_flags=w.opts.mixed_flags, reps=3)
    npix = batch.pixel_counts()
    work = stage_work(w, K, npix, n_chunk)
    stages = []
    for wk, t in zip(work, ms):
        name, bound, amount, prec = wk[:4]
        if bound == "hbm":
            ach = amount / (t * 1e-3) / 1e9
            peak = PEAK_HBM_GBS
            unit = "GB/s"
        else:
            ach = amount / (t * 1e-3) / 1e12
            peak = PEAK_TFLOPS[prec]
            unit = "TFLOP/s"
        stages.append(dict(stage=name, ms=round(float(t), 4), bound=bound, achieved=round(float(ach), 3), peak=peak,
                           unit=unit, frac=round(float(ach / peak), 5), frames=int(n_chunk)))
        extra = ""
        if len(wk) > 4:
            stages[-1]["gather_GBs"] = round(float(wk[4] / (t * 1e-3) / 1e9), 1)
            extra = f", basis gathers {stages[-1]['gather_GBs']:.0f} GB/s through L2/MALL"
        log(f"[bench] stage {name:24s} {t:8.3f} ms  {ach:9.2f} {unit} ({100 * ach / peak:.1f}% of peak{extra})")
    dom = max(stages, key=lambda s: s["ms"])
    traffic = None
    if os.path.exists(args.traffic_json):
        # measured HBM bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated) of the
        # kernels profiled at this launch size
        for tj in json.load(open(args.traffic_json)).get("kernels", []):
            if tj.get("stage") == dom["stage"] and tj.get("frames_per_launch") == int(n_chunk):
                traffic = tj.get("bytes_per_launch")
    roofline = dict(kernel=dom["stage"], bound=dom["bound"], achieved=dom["achieved"], peak=dom["peak"],
                    unit=dom["unit"], frac=dom["frac"], traffic=traffic,
                    launch_ms=dom["ms"], frames_per_launch=int(n_chunk))
    if args.copy_probe:
        a = torch.empty(1 << 28, dtype=torch.float32, device="cuda").fill_(1.0)
        b_ = torch.empty_like(a)
        torch.cuda.synchronize()
        b_.copy_(a)
        torch.cuda.synchronize()
        log("[bench] copy probe: 1 GiB read + 1 GiB written")
        del a, b_

    # landmark-driven initialisation of every frame (orthographic fit + pose seed, tdi_ortho_init)
    lmk = torch.from_numpy(np.stack([np.stack([f.lmk_x, f.lmk_y]) for f in w.frames]).astype(np.float32)).cuda()
    cams = torch.from_numpy(np.array([[f.phix, f.phiy, f.cx, f.cy] for f in w.frames], np.float32)).cuda()
    tdi.ortho_init(ctx, lmk, cams)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        init_out = tdi.ortho_init(ctx, lmk, cams)
    e1.record()
    torch.cuda.synchronize()
    init_ms = e0.elapsed_time(e1) / 5
    landmark_init = dict(frames=args.frames, ms=round(init_ms, 4), frames_per_s=round(args.frames / (init_ms * 1e-3), 1),
                         mean_outer_iterations=round(float(init_out["info"][:, 0].float().mean()), 2))
    log(f"[bench] landmark initialisation: {args.frames} frames in {init_ms:.3f} ms")

    # frame preparation (tdi_prepare_crops): 8-bit BGR 1280x720 frames resident in HBM -> the
    # workload's crop sizes, resize coefficients U(0.5, 1.5)
    sw, sh = 1280, 720
    rng = np.random.default_rng(0x3D1 + 7)
    pcoef = rng.uniform(0.5, 1.5, args.frames).astype(np.float32)
    pcrop = np.zeros((args.frames, 4), np.int32)
    for i, f in enumerate(w.frames):
        dw, dh = int(np.rint(sw * float(pcoef[i]))), int(np.rint(sh * float(pcoef[i])))
        pcrop[i] = (rng.integers(0, max(dw - f.width, 1)), rng.integers(0, max(dh - f.height, 1)), f.width, f.height)
    src = torch.randint(0, 256, (args.frames, sh, sw, 3), dtype=torch.uint8, device="cuda")
    tdi.prepare_crops(src, pcoef, pcrop)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(5):
        crops = tdi.prepare_crops(src, pcoef, pcrop)
    e1.record()
    torch.cuda.synchronize()
    prep_ms = e0.elapsed_time(e1) / 5
    # algorithmic bytes: the crops written plus the source footprint of each crop read once
    k = pcoef.astype(np.float64)
    fx0 = np.clip(np.floor((pcrop[:, 0] + 0.5) / k - 0.5), 0, sw - 1)
    fx1 = np.clip(np.floor((pcrop[:, 0] + pcrop[:, 2] - 0.5) / k - 0.5) + 1, 0, sw - 1)
    fy0 = np.clip(np.floor((pcrop[:, 1] + 0.5) / k - 0.5), 0, sh - 1)
    fy1 = np.clip(np.floor((pcrop[:, 1] + pcrop[:, 3] - 0.5) / k - 0.5) + 1, 0, sh - 1)
    prep_bytes = float((4.0 * pcrop[:, 2] * pcrop[:, 3]).sum() + (3.0 * (fx1 - fx0 + 1) * (fy1 - fy0 + 1)).sum())
    frame_prep = dict(frames=args.frames, source=f"{sw}x{sh} BGR8", ms=round(prep_ms, 4),
                      frames_per_s=round(args.frames / (prep_ms * 1e-3), 1), algorithmic_bytes=int(prep_bytes),
                      achieved_GBs=round(prep_bytes / (prep_ms * 1e-3) / 1e9, 1), peak_GBs=8000.0,
                      frac=round(prep_bytes / (prep_ms * 1e-3) / 8e12, 4),
                      note="ms per call includes the per-call descriptor upload and stream synchronisation")
    log(f"[bench] frame preparation: {args.frames} frames in {prep_ms:.3f} ms")

    # FP64 oracle on a prefix: host baseline and parameter error of the batched fit
    cpu = None
    accuracy = None
    if not args.no_cpu and args.cpu_frames > 0:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O
        nc = min(args.cpu_frames, args.frames)
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        sc = O.Scene(w.model, w.frames[:nc], w.opts)
        t0 = time.perf_counter()
        th_o, ro = sc.fit(init[:nc].astype(np.float64), n_threads=threads)
        dtc = time.perf_counter() - t0
        cpu = dict(value=round(nc / dtc, 3), unit="frames/s", cores=threads, kind="port",
                   sample=f"first {nc} frames of the workload, FP64 OpenMP oracle, one frame per thread, "
                          f"{w.opts.max_iters} GN iterations, {dtc:.1f} s")
        log(f"[bench] oracle: {nc} frames in {dtc:.1f}s on {threads} threads")
        accuracy = sample_accuracy(w, O, sc, init, nc, th_o, ro, th_batched, plain, threads, tdi, ctx)
        p0L = w.model.mean_shape.astype(np.float64)[w.model.landmark_ids.astype(np.int64)]
        t0 = time.perf_counter()
        O.ortho_init(lmk.cpu().numpy(), cams.cpu().numpy(), p0L)
        landmark_init["cpu_frames_per_s"] = round(args.frames / (time.perf_counter() - t0), 1)
        landmark_init["cpu_cores"] = threads
        npf = min(256, args.frames)
        host_src = src[:npf].cpu().numpy()
        t0 = time.perf_counter()
        want = O.prepare_crops(host_src, pcoef[:npf], pcrop[:npf])
        frame_prep["cpu_frames_per_s"] = round(npf / (time.perf_counter() - t0), 1)
        frame_prep["cpu_sample"] = f"first {npf} frames, oracle on {os.environ.get('OMP_NUM_THREADS', 'all')} OpenMP threads"
        got = crops[:want.size].cpu().numpy()
        frame_prep["bitexact_vs_oracle"] = bool(np.array_equal(got.view(np.uint32), want.view(np.uint32)))
    del src, crops

    out = {
        "metric": METRIC,
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": w.name, "frames": args.frames, "unknowns_per_frame": K,
                   "gn_iterations": int(w.opts.max_iters), "mean_iterations_used": float(iters.mean()),
                   "precision": "FP64 vertex synthesis (f64 MFMA), BF16 hi/mid/lo J^T J on MFMA with FP32 "
                                "accumulation, FP32 J^T r, equilibrated FP32 Cholesky + 1 FP64 refinement"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_gpu": ({k: v for k, v in plain.items() if k != "theta"} if plain else None),
        "speedup_vs_plain": (round(fps / plain["fps"], 2) if plain else None),
        "batched_stop_flags": {str(k): int(v) for k, v in zip(*np.unique([r.stop for r in reps], return_counts=True))},
        "accuracy_vs_fp64": accuracy,
        "batched_stop_flags_note": "1 max iterations, 2 decrement test, 4 step below t_min, 8 NaN, 16 not PD",
        "stages": stages,
        "landmark_init": landmark_init,
        "frame_prep": frame_prep,
    }
    print(json.dumps(out), flush=True)
    batch.close()
    ctx.close()


if __name__ == "__main__":
    main()
