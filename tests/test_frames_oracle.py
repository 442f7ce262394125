"""Frame preparation: the oracle's grey crops of resized 8-bit BGR frames
(InputData::get_resized_frame, input_data.cu:44-53; the crop and transpose of
video_fitter.cu:537-543) and the host layer's landmark reader and per-frame geometry.

The pixel steps are OpenCV 4 calls (cvtColor BGR2GRAY, convertTo with scale 1/255, resize
INTER_LINEAR, copyMakeBorder), a dependency absent here, so the oracle restates OpenCV's published
semantics.  How each part is pinned:
  - grey conversion: OpenCV's known outputs for pure blue, green and red (29, 150, 76) and
    grey(v, v, v) = v for every v;
  - destination size: round-half-even of src * coef (2.5 -> 2, 3.5 -> 4);
  - resize: the known 1x2 -> 1x4 upsampling [0, 1] -> [0, .25, .75, 1] in both directions; the
    copy when the destination size equals the source size; linear images reproduced inside the
    frame (bilinear interpolation is exact on them); an independent float64 bilinear with
    half-pixel centres and clamped edges agrees to 1e-6 on random frames (sources of 1-97 pixels
    a side, scales 0.2-6, crops reaching past the frame);
  - crop layout: column-major, origin (x0, y0), zero beyond the resized frame.
  - the exact 2x downscale (coef 0.5), for which OpenCV switches INTER_LINEAR to its area-averaging
    path: the linear weights are then exactly 1/2, so the linear passes give
    RN(RN(a + b) + RN(c + d)) / 4, the value of the vectorised area average ((a + b) + (c + d)) * 0.25
    over each 2x2 block; checked bit for bit on even and odd source sizes.
Left open ("parity unpinned" against OpenCV itself): OpenCV builds whose SIMD paths fuse the
multiply-adds of the resize round differently in the last bit, and the summation order of the
area path's scalar tail (columns past the last full vector), which this build does not restate.
"""
import importlib
import os

import numpy as np
import pytest

import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

f32 = np.float32


def _bgr(grey):
    return np.repeat(np.asarray(grey, np.uint8)[..., None], 3, axis=-1)


def _grey(fr):
    """OpenCV's 8-bit BGR -> grey, as integers (fr: [..., 3] uint8)."""
    b, g, r = (fr[..., i].astype(np.int64) for i in range(3))
    return (1868 * b + 9617 * g + 4899 * r + 8192) >> 14


def _crops(frames, coef, crop):
    out = O.prepare_crops(frames, coef, crop)
    res, at = [], 0
    for x0, y0, w, h in np.asarray(crop):
        res.append(out[at:at + w * h].reshape(w, h).T)   # column-major -> [row, col]
        at += w * h
    return res


def test_grey_conversion_known_answers():
    px = np.array([[[255, 0, 0], [0, 255, 0], [0, 0, 255]]], np.uint8)[None]   # B, G, R
    got = _crops(px, [1.0], [[0, 0, 3, 1]])[0][0]
    want = np.array([29, 150, 76], f32) * f32(1.0 / 255.0)
    assert np.array_equal(got, want)
    v = np.arange(256, dtype=np.uint8).reshape(1, 16, 16)
    got = _crops(_bgr(v), [1.0], [[0, 0, 16, 16]])[0]
    assert np.array_equal(got, v.astype(f32)[0] * f32(1.0 / 255.0))


def test_destination_size_rounds_half_to_even():
    assert [O.resized_size(n, 0.5) for n in (3, 5, 7, 9)] == [2, 2, 4, 4]
    assert O.resized_size(640, 1.0001) == 640 and O.resized_size(100, 1.37) == 137


def test_known_two_pixel_upsampling():
    one = np.array([[0, 255]], np.uint8)
    got = _crops(_bgr(one)[None], [2.0], [[0, 0, 4, 2]])[0]
    assert np.allclose(got, [[0, .25, .75, 1], [0, .25, .75, 1]], atol=1e-7)
    got = _crops(_bgr(one.T)[None], [2.0], [[0, 0, 2, 4]])[0]
    assert np.allclose(got.T, [[0, .25, .75, 1], [0, .25, .75, 1]], atol=1e-7)


def test_equal_size_is_a_copy_and_layout_is_column_major():
    rng = np.random.default_rng(1)
    fr = rng.integers(0, 256, (1, 40, 64, 3), dtype=np.uint8)
    grey = _grey(fr[0]).astype(f32) * f32(1.0 / 255.0)
    a, b = _crops(np.concatenate([fr, fr]), [1.0, 1.004], [[5, 7, 30, 20], [5, 7, 30, 20]])
    assert O.resized_size(64, 1.004) == 64 and O.resized_size(40, 1.004) == 40
    assert np.array_equal(a, grey[7:27, 5:35]) and np.array_equal(b, a)


@pytest.mark.parametrize("sh,sw", [(48, 64), (49, 65), (720, 1280)])
def test_exact_half_scale_is_the_two_by_two_area_average(sh, sw):
    rng = np.random.default_rng(sh + sw)
    fr = rng.integers(0, 256, (1, sh, sw, 3), dtype=np.uint8)
    v = _grey(fr[0]).astype(f32) * f32(1.0 / 255.0)
    dw, dh = O.resized_size(sw, 0.5), O.resized_size(sh, 0.5)
    assert 2 * dw <= sw and 2 * dh <= sh
    got = _crops(fr, [0.5], [[0, 0, dw, dh]])[0]
    a, b = v[0:2 * dh:2, 0:2 * dw:2], v[0:2 * dh:2, 1:2 * dw:2]
    c, d = v[1:2 * dh:2, 0:2 * dw:2], v[1:2 * dh:2, 1:2 * dw:2]
    want = ((a + b) + (c + d)) * f32(0.25)
    assert want.dtype == f32
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_linear_image_is_reproduced_and_border_is_zero():
    # grey = column index: bilinear interpolation returns the source coordinate itself
    sw, sh = 200, 120
    fr = _bgr(np.tile(np.arange(sw, dtype=np.uint8), (sh, 1)))[None]
    for coef in (0.63, 1.71):
        dw, dh = O.resized_size(sw, coef), O.resized_size(sh, coef)
        x0, y0 = 10, 5
        w, h = dw - x0 + 7, dh - y0 + 9            # reaches past the resized frame
        got = _crops(fr, [coef], [[x0, y0, w, h]])[0]
        X = x0 + np.arange(w)
        src = np.clip(((X + 0.5) / coef - 0.5), 0, sw - 1)
        inside = got[: dh - y0, : dw - x0]
        assert np.allclose(inside, np.broadcast_to(src[: dw - x0] / 255, inside.shape), atol=2e-6)
        assert (got[dh - y0:, :] == 0).all() and (got[:, dw - x0:] == 0).all()


def _bilinear64(grey, coef, x0, y0, w, h):
    sh, sw = grey.shape
    dw, dh = O.resized_size(sw, coef), O.resized_size(sh, coef)
    X, Y = x0 + np.arange(w), y0 + np.arange(h)
    fx = (X + 0.5) / coef - 0.5
    sx = np.floor(fx)
    wx = fx - sx
    edge = (sx < 0) | (sx >= sw - 1)
    sx = np.clip(sx, 0, sw - 1).astype(int)
    wx[edge] = 0
    fy = (Y + 0.5) / coef - 0.5
    sy = np.floor(fy)
    wy = fy - sy
    r0, r1 = np.clip(sy, 0, sh - 1).astype(int), np.clip(sy + 1, 0, sh - 1).astype(int)
    c1 = np.minimum(sx + 1, sw - 1)
    top = grey[r0][:, sx] * (1 - wx) + grey[r0][:, c1] * wx
    bot = grey[r1][:, sx] * (1 - wx) + grey[r1][:, c1] * wx
    v = top * (1 - wy)[:, None] + bot * wy[:, None]
    v[(Y >= dh)] = 0
    v[:, (X >= dw)] = 0
    return v


def test_random_frames_agree_with_a_float64_bilinear():
    rng = np.random.default_rng(7)
    n, sh, sw = 6, 90, 130
    fr = rng.integers(0, 256, (n, sh, sw, 3), dtype=np.uint8)
    coef = np.array([0.37, 0.8, 1.3, 2.2, 3.0, 1.05], f32)
    crop = np.array([[0, 0, 40, 30], [3, 9, 60, 70], [20, 11, 150, 120], [100, 80, 90, 60], [0, 0, 64, 64],
                     [50, 40, 90, 60]], np.int32)
    got = _crops(fr, coef, crop)
    for i in range(n):
        grey = _grey(fr[i]) / 255.0
        want = _bilinear64(grey, float(coef[i]), *crop[i])
        assert np.allclose(got[i], want, atol=1e-6), i


# ------------------------------------------------------------------ host layer (no GPU calls)

video = importlib.import_module("3di_amd.video")


def _host_or_skip():
    try:
        return video.lib()
    except Exception as e:  # pragma: no cover - the library is built by the test session
        pytest.skip(f"host layer not loadable: {e}")


def test_read_landmarks(tmp_path):
    _host_or_skip()
    rng = np.random.default_rng(3)
    rows = rng.uniform(0, 500, (3, 102)).astype(f32)
    lines = [" ".join(f"{v:.4f}" for v in r) for r in rows]
    toks = lines[1].split()
    toks[6] = "nan"                     # x of landmark 3 in frame 1
    lines[1] = " ".join(toks)
    p = tmp_path / "lmk.txt"
    p.write_text("\n".join(lines) + "\n")
    lx, ly = video.read_landmarks(str(p))
    want = np.round(rows.astype(np.float64), 4).astype(f32)
    assert lx.shape == (3, 51)
    assert np.isnan(lx[1, 3])
    mask = np.ones_like(lx, bool)
    mask[1, 3] = False
    assert np.array_equal(lx[mask], want[:, 0::2][mask]) and np.array_equal(ly, want[:, 1::2])
    (tmp_path / "short.txt").write_text("1 2 3\n")
    with pytest.raises(Exception):
        video.read_landmarks(str(tmp_path / "short.txt"))


def test_frame_geometry_known_answers():
    _host_or_skip()
    lx = np.linspace(100.6, 180.2, 51, dtype=f32)[None]      # extents: x 100..180 -> 80
    ly = np.linspace(50.3, 150.9, 51, dtype=f32)[None]       # y 50..150 -> 100 (int truncation)
    cams = np.array([[1000, 1000, 320, 240]], f32)
    g = video.frame_geometry(lx, ly, cams, 640, 480, 200.0, pad=True, crop_size=(256, 256))
    k = f32(200.0) / f32(100.0)
    assert g["status"][0] == 0 and g["coef"][0] == k
    assert np.array_equal(g["cams"][0], cams[0] * k)
    assert np.array_equal(g["lmk_x"][0], k * lx[0]) and np.array_equal(g["lmk_y"][0], k * ly[0])
    rx, ry = k * lx[0], k * ly[0]
    sz = max(rx.max() - rx.min(), ry.max() - ry.min())
    assert g["face_size"][0] == f32(sz)
    assert tuple(g["crop"][0]) == (int(max(f32(rx.min() - sz / 1.5), 0)), int(max(f32(ry.min() - sz / 1.5), 0)), 256,
                                   256)
    g = video.frame_geometry(lx, ly, cams, 640, 480, 200.0, pad=False)
    assert tuple(g["crop"][0][:2]) == (int(rx.min() - 20), int(ry.min() - 20))
    # a landmark outside the frame: the frame is skipped
    g = video.frame_geometry(lx, ly, cams, 170, 480, 200.0)
    assert g["status"][0] == 6 and g["coef"][0] == 0


def test_prepare_crops_rejects_bad_arguments_before_any_device_call():
    tdi = importlib.import_module("3di_amd.tdi")
    import ctypes as C
    L = tdi.lib()
    fake = C.c_void_p(1)            # never dereferenced: validation fails first
    good_crop = np.array([[0, 0, 8, 8]], np.int32)
    cases = [([0.0], good_crop, 16), ([np.inf], good_crop, 16),          # coefficient
             ([1.0], np.array([[-1, 0, 8, 8]], np.int32), 16),                 # crop origin
             ([1.0], np.array([[0, 0, 0, 8]], np.int32), 16),                  # crop size
             ([1.0], np.array([[2**31 - 8, 0, 8, 8]], np.int32), 16),          # x0 + w past 2^31 - 1
             ([1.0], np.array([[0, 2**31 - 2, 8, 8]], np.int32), 16),          # y0 + h past 2^31 - 1
             ([1.0], good_crop, 0),                                            # source width
             ([1e-4], good_crop, 16)]                                          # empty resized frame
    for coef, crop, sw in cases:
        c = np.array(coef, np.float32)
        rc = L.tdi_prepare_crops(1, fake, sw, 16, 3 * 16 * 16, c.ctypes.data, crop.ctypes.data, fake, None)
        assert rc == 1, (coef, crop, sw)    # TDI_ERR_ARG
    assert L.tdi_prepare_crops(0, None, 16, 16, 768, None, None, None, None) == 0
    # the allocation-free variant checks the same arguments, and its frame limit, first
    for coef, crop, sw in cases:
        c = np.array(coef, np.float32)
        rc = L.tdi_prepare_crops_noalloc(1, fake, sw, 16, 3 * 16 * 16, c.ctypes.data, crop.ctypes.data, fake, None)
        assert rc == 1, (coef, crop, sw)
    nmax = 16384                    # TDI_PREP_NOALLOC_MAX_FRAMES
    assert "TDI_PREP_NOALLOC_MAX_FRAMES %d" % nmax in open(os.path.join(ROOT, "include", "tdi_api.h")).read()
    c = np.ones(nmax + 1, np.float32)
    crops = np.tile(good_crop, (nmax + 1, 1))
    rc = L.tdi_prepare_crops_noalloc(nmax + 1, fake, 16, 16, 768, c.ctypes.data, crops.ctypes.data, fake, None)
    assert rc == 1 and b"at most" in L.tdi_last_error()
    assert L.tdi_prepare_crops_noalloc(0, None, 16, 16, 768, None, None, None, None) == 0


def test_host_frame_functions_reject_bad_arguments(tmp_path):
    _host_or_skip()
    with pytest.raises(Exception):
        video.read_landmarks(str(tmp_path / "missing.txt"))
    with pytest.raises(Exception):
        video.frame_geometry(np.zeros((1, 51)), np.zeros((1, 51)), np.zeros((1, 4)), 640, 480, 0.0)


@pytest.mark.parametrize("seed", range(6))
def test_random_sources_and_scales_agree_with_a_float64_bilinear(seed):
    # sources from 1 pixel to 97 pixels a side, scales 0.2-6, crops anywhere (also past the frame)
    rng = np.random.default_rng(100 + seed)
    sh, sw = (int(v) for v in rng.integers(1, 98, 2))
    n = 4
    fr = rng.integers(0, 256, (n, sh, sw, 3), dtype=np.uint8)
    coef = np.exp(rng.uniform(np.log(0.2), np.log(6.0), n)).astype(f32)
    coef = np.where([O.resized_size(sw, c) >= 1 and O.resized_size(sh, c) >= 1 for c in coef], coef, f32(1.0))
    crop = np.stack([rng.integers(0, 3 * sw, n), rng.integers(0, 3 * sh, n), rng.integers(1, 90, n),
                     rng.integers(1, 90, n)], 1).astype(np.int32)
    got = _crops(fr, coef, crop)
    for i in range(n):
        dw, dh = O.resized_size(sw, coef[i]), O.resized_size(sh, coef[i])
        grey = _grey(fr[i]) / 255.0
        if dw == sw and dh == sh:           # equal sizes: a copy
            want = np.zeros((crop[i, 3], crop[i, 2]))
            x0, y0, w, h = crop[i]
            ys, xs = slice(y0, min(y0 + h, sh)), slice(x0, min(x0 + w, sw))
            part = grey[ys, xs]
            want[: part.shape[0], : part.shape[1]] = part
        else:
            want = _bilinear64(grey, float(coef[i]), *crop[i])
        assert np.allclose(got[i], want, atol=1e-6), (seed, i)
