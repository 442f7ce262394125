"""The identity stage's frame binning (tdi_bin_identity_frames, fit_video_frames_landmarks_sparse,
video_fitter.cu:2380-2550), host arithmetic, on synthetic angles:
  - the selection is groups of NFRAMES, at most NTOT_RECONSTRS of them, of frames that have an
    angle; no bin contributes more than NFRAMES_PER_ANGLE_BIN distinct frames (bins: nearest of the
    13 x 13 yaw/pitch centroids -55 .. 65 degrees, recomputed here with numpy);
  - with one pass (NPERMS = 1) and room for every kept frame, the groups hold distinct frames and
    every bin with fewer frames than its cap is kept whole;
  - the result is a pure function of the angles (the shuffles use a fixed seed, 1907, as the
    reference's std::default_random_engine);
  - frames without an angle (NaN) are never selected; no angle at all selects nothing.
The shuffles are libstdc++'s std::shuffle with minstd_rand0 seeded 1907 in the reference's order
of calls; no output of a reference run exists to pin the exact permutation (parity unpinned there).
"""
import importlib

import numpy as np

video = importlib.import_module("3di_amd.video")

CENTROIDS = np.arange(-60, 61, 10) + 5


def _bins(yaw, pitch):
    yd, pd = np.degrees(yaw), np.degrees(pitch)
    yi = np.argmin(np.abs(CENTROIDS[None, :] - yd[:, None]), axis=1)
    pi = np.argmin(np.abs(CENTROIDS[None, :] - pd[:, None]), axis=1)
    return yi * 100 + pi


def _angles(n=300, seed=5, nan_every=7):
    rng = np.random.default_rng(seed)
    yaw = np.radians(rng.uniform(-50, 50, n)).astype(np.float32)
    pitch = np.radians(rng.uniform(-25, 25, n)).astype(np.float32)
    yaw[::nan_every] = np.nan
    return yaw, pitch


def test_groups_bins_and_caps():
    yaw, pitch = _angles()
    p = video.identity_params(nframes=4, nframes_per_angle_bin=2, nperms=3, ntot_reconstrs=6)
    sel = video.bin_identity_frames(p, yaw, pitch)
    assert sel.size % 4 == 0 and 0 < sel.size <= 24
    assert np.isfinite(yaw[sel]).all()
    b = _bins(yaw[sel], pitch[sel])
    for key in np.unique(b):
        assert np.unique(sel[b == key]).size <= 2
    assert np.array_equal(sel, video.bin_identity_frames(p, yaw, pitch))


def test_single_pass_keeps_small_bins_whole():
    yaw, pitch = _angles(n=60, seed=9)
    p = video.identity_params(nframes=3, nframes_per_angle_bin=1000, nperms=1, ntot_reconstrs=1000)
    sel = video.bin_identity_frames(p, yaw, pitch, cap=1000)
    finite = np.flatnonzero(np.isfinite(yaw))
    assert sel.size == (finite.size // 3) * 3
    assert np.unique(sel).size == sel.size
    assert set(sel.tolist()) <= set(finite.tolist())


def test_nothing_without_angles():
    p = video.identity_params(nframes=3)
    yaw = np.full(20, np.nan, np.float32)
    assert video.bin_identity_frames(p, yaw, np.zeros(20, np.float32)).size == 0
