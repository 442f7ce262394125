"""The launch-shape autotuner (tdi_batch_autotune): it times the per-pixel Jacobian kernel's grid
sizes and the J^T J split counts on the batch's first chunk, keeps the fastest, and a later batch
of the same shape starts from that choice.  The chosen shapes change only the grid and the order
of the partial sums: the normal equations stay within FP32 summation-order noise of the default
shapes' (1e-5 relative), and a fit with the tuned shapes succeeds on every frame (its distance to
the default shapes' fit is printed: summation order alone can part whole-fit trajectories).
"""
import importlib
import sys

import numpy as np
import pytest

from scenes import make_scene

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tdi = importlib.import_module("3di_amd.tdi")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_autotune_choice_and_results():
    _say("case: autotune on a 12-frame small scene")
    sc = make_scene(grid=48, k_exp=12, crops=(96, 128, 160, 200, 240, 110, 150, 190, 230, 270, 100, 300),
                    pose_deg=15.0, fit_light=2)
    ctx = tdi.Context(sc.model, sc.opts)
    b = tdi.Batch(ctx, sc.frames)
    th = torch.from_numpy(sc.init().astype(np.float32)).cuda()
    H0, g0, _ = b.normal_equations(th, mixed=True)
    f0, r0 = b.fit_batched(th.clone())
    shape0 = b.launch_shape()
    res = b.autotune(th, reps=2)
    _say(f"  default {shape0}, chosen {res}")
    assert res["default_splits"] == shape0[0]
    assert b.launch_shape() == (res["jtj_splits"], res["dI_area"])
    assert all(t > 0 for t in res["dI_ms"].values()) and len(res["jtj_ms"]) >= 1
    assert min(res["jtj_ms"].values()) == res["jtj_ms"][str(res["jtj_splits"])]
    H1, g1, _ = b.normal_equations(th, mixed=True)
    torch.cuda.synchronize()
    eH = float((H1 - H0).abs().max() / H0.abs().max())
    eg = float((g1 - g0).abs().max() / g0.abs().max())
    _say(f"  normal equations, tuned vs default: H {eH:.1e}, g {eg:.1e}")
    assert eH < 1e-5 and eg < 1e-5
    f1, r1 = b.fit_batched(th.clone())
    torch.cuda.synchronize()
    assert all(r.status == 0 for r in r1)
    d = (f1 - f0).abs().max().item()
    _say(f"  fit, tuned vs default shapes: max |dtheta| {d:.2e}")
    # a second batch of the same frames starts from the tuned shape
    b2 = tdi.Batch(ctx, sc.frames)
    assert b2.launch_shape() == (res["jtj_splits"], res["dI_area"])
    b2.close()
    b.close()
    ctx.close()
