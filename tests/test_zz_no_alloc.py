"""GPU check of tdi_prepare_crops_noalloc (the frame preparation with no device allocation)
through the C ABI: bit-equal to tdi_prepare_crops on two calls in a row, the second with fewer
frames and smaller crops, so that what the first call left in the library's fixed descriptor
array cannot leak into the second call's result."""
import importlib
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
tdi = importlib.import_module("3di_amd.tdi")


def _case(rng, n, sw, sh, smax):
    frames = torch.from_numpy(rng.integers(0, 256, (n, sh, sw, 3), dtype=np.uint8)).cuda()
    coef = rng.uniform(0.4, 2.5, n).astype(np.float32)
    crop = np.zeros((n, 4), np.int32)
    for i in range(n):
        w, h = (int(v) for v in rng.integers(1, smax + 1, 2))
        crop[i] = (rng.integers(0, max(1, int(sw * coef[i]))), rng.integers(0, max(1, int(sh * coef[i]))), w, h)
    return frames, coef, crop


def test_prepare_crops_noalloc_matches_the_allocating_entry_point():
    rng = np.random.default_rng(2024)
    for n, smax in ((48, 160), (5, 40)):      # the second call is smaller
        print(f"[no-alloc] {n} frames, crops up to {smax} px", flush=True)
        sys.stdout.flush()
        frames, coef, crop = _case(rng, n, 200, 150, smax)
        want = tdi.prepare_crops(frames, coef, crop).cpu().numpy()
        got = torch.full((want.size,), float("nan"), device="cuda", dtype=torch.float32)
        tdi.prepare_crops(frames, coef, crop, out=got, noalloc=True)
        got = got.cpu().numpy()
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), f"{n} frames: results differ"
