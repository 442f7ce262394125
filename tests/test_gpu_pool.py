"""The context's device block pool: a batch created after others were destroyed takes its
workspace from their cached blocks, which still hold their data.  A fit must not depend on what
those blocks held: the same frames fitted in a fresh context's first batch, and again after a
larger batch of the same context was fitted and destroyed, give the same parameters and reports
bit for bit."""
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


def test_pooled_workspace_gives_the_same_fit():
    _say("case: batch workspaces from the context's block pool")
    sc = make_scene(grid=48, k_exp=12, crops=(96, 128, 160, 200, 240), pose_deg=15.0, fit_light=2)
    ctx = tdi.Context(sc.model, sc.opts)
    th0 = torch.from_numpy(sc.init().astype(np.float32)).cuda()
    b = tdi.Batch(ctx, sc.frames)
    f0, r0 = b.fit_batched(th0.clone())
    torch.cuda.synchronize()
    b.close()
    # a larger batch of the same context: more frames, so other block sizes; its blocks go back
    # to the pool holding its data
    b2 = tdi.Batch(ctx, sc.frames + sc.frames[::-1])
    f2, r2 = b2.fit_batched(torch.cat([th0, th0.flip(0)]).clone())
    torch.cuda.synchronize()
    b2.close()
    for rep in range(2):
        b3 = tdi.Batch(ctx, sc.frames)
        f3, r3 = b3.fit_batched(th0.clone())
        torch.cuda.synchronize()
        b3.close()
        d = (f3 - f0).abs().max().item()
        _say(f"  pass {rep}: max |dtheta| against the first batch {d:.1e}")
        assert torch.equal(f3, f0)
        assert [(r.status, r.iters, r.stop, r.ls_evals) for r in r3] == [(r.status, r.iters, r.stop, r.ls_evals) for r in r0]
    # the kept blocks freed: the next batch allocates afresh, with the same fit
    ctx.release_cached()
    b4 = tdi.Batch(ctx, sc.frames)
    f4, _ = b4.fit_batched(th0.clone())
    torch.cuda.synchronize()
    b4.close()
    assert torch.equal(f4, f0)
    ctx.close()
