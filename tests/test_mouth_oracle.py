"""Inner-mouth masking (paint_innermouth_black, preprocessing.cu:40-103): the oracle against
known answers on an eight-vertex mouth polygon.

The reference strokes the inner-lip polygon with 3-pixel lines, traces the outer contour with
OpenCV and blackens the mouth-box pixels whose signed distance to that contour exceeds
u = face size / REF_FACE_SIZE.  The restatement takes the contour as the polygon moved one pixel
outwards, so a pixel is blackened when it is inside the polygon by more than u - 1 (anywhere inside
when u < 1) or outside it by less than 1 - u.  Pinned here against an independent computation on a
convex polygon (half-planes: a lattice point's depth inside a convex polygon is its least distance
to the edge lines) for u = 1, 2 and 0.5, the lattice counts in closed form for u = 1 and 0.5, the
open-mouth threshold (mean lip distance in REF_FACE_SIZE units > 3), the clipping of the box to the
frame and the landmarks 43 and 47 moving 3u inwards.  Left open ("parity unpinned" against
OpenCV's rasterised contour itself, a dependency absent here): pixels within about one pixel of the
contour, where the traced polyline and its geometric offset differ.
"""
import numpy as np
import pytest

import pyoracle as O

REF = 86.0
# the shifted polygon: 43 .. 50 (43 and 47 after their 3u shift)
POLY = np.array([(100, 110), (110, 100), (120, 100), (130, 100), (140, 110), (130, 120), (120, 120), (110, 120)],
                np.float64)


def _landmarks(u, dx=0.0, dy=0.0, lower=None):
    """51 landmarks whose face size is 86u, with the mouth polygon above (43 and 47 placed 3u
    outwards so that their shift brings them onto it); the rest sit at the mouth's centre except
    landmarks 0 and 1, which fix the face size."""
    x = np.full(51, 120.0)
    y = np.full(51, 110.0)
    x[43:51] = POLY[:, 0]
    y[43:51] = POLY[:, 1]
    if lower is not None:
        y[48:51] = lower
    x[43] -= 3 * u
    x[47] += 3 * u
    span = x[43:51].max() - x[43:51].min()
    extra = (86.0 * u - span) / 2
    assert extra >= 0
    x[0] = x[43:51].min() - extra
    x[1] = x[43:51].max() + extra
    return (x + dx).astype(np.float32), (y + dy).astype(np.float32)


def _paint(x, y, w=256, h=256, fill=200):
    fr = np.full((1, h, w, 3), fill, np.uint8)
    return O.paint_innermouth_black(fr, x[None], y[None], REF)[0]


def _painted(fr):
    z = (fr == 0)
    assert np.array_equal(z[..., 0], z[..., 1]) and np.array_equal(z[..., 0], z[..., 2])
    return z[..., 0]


def _convex_expectation(poly, m, box, w, h):
    """Pixels (rows i, columns j) of the box that the rule blackens, from the half-plane form of a
    convex polygon.  Outside points are only ever blackened within |m| < 1 of the polygon; the
    caller's shapes keep every outside lattice point at least 0.7 from every edge line that it lies
    beyond, so the line distance decides them."""
    x0, x1, y0, y1 = box
    out = np.zeros((h, w), bool)
    P = poly
    area = np.sum(P[:, 0] * np.roll(P[:, 1], -1) - np.roll(P[:, 0], -1) * P[:, 1])
    sgn = 1.0 if area > 0 else -1.0
    for i in range(y0, y1):
        for j in range(x0, x1):
            d = []
            for k in range(len(P)):
                a, b = P[k], P[(k + 1) % len(P)]
                e = b - a
                cross = e[0] * (i - a[1]) - e[1] * (j - a[0])
                d.append(sgn * cross / np.hypot(*e))
            depth = min(d)
            if depth > 0:
                out[i, j] = m < 0 or depth > m
            elif depth == 0:
                out[i, j] = m < 0
            else:
                worst = -depth
                assert worst >= 0.7 or worst > abs(m) + 0.05
                out[i, j] = m < 0 and worst < -m
    return out


@pytest.mark.parametrize("u", [1.0, 2.0, 0.5])
def test_open_mouth_matches_the_half_plane_form(u):
    x, y = _landmarks(u)
    fr = _paint(x, y)
    got = _painted(fr)
    box = (int(x[31:].min()), int(x[31:].max()), int(y[31:].min()), int(y[31:].max()))
    want = _convex_expectation(POLY, u - 1.0, box, 256, 256)
    assert np.array_equal(got, want)
    # pixels outside the mouth box keep their value
    keep = np.ones((256, 256), bool)
    keep[box[2]:box[3], box[0]:box[1]] = False
    assert (fr[keep] == 200).all()


def test_closed_form_counts():
    # u = 1: the lattice points strictly inside the polygon, 2 * sum_{d=1..9} (19 + 2d) + 39
    assert _painted(_paint(*_landmarks(1.0))).sum() == 561
    # u = 0.5: inside and on the boundary, rows 100..119 of the box (row 120 is its excluded end)
    assert _painted(_paint(*_landmarks(0.5))).sum() == 620
    # u = 2: inside by more than one pixel
    assert _painted(_paint(*_landmarks(2.0))).sum() == _convex_expectation(
        POLY, 1.0, (0, 256, 0, 256), 256, 256).sum()


def test_closed_mouth_is_left_alone():
    # lower lip one pixel below the upper: opening 1 (REF units at u = 1) is not above 3
    x, y = _landmarks(1.0, lower=101.0)
    assert _painted(_paint(x, y)).sum() == 0
    # opening exactly 3 is not above 3; a little more is
    x, y = _landmarks(1.0, lower=103.0)
    assert _painted(_paint(x, y)).sum() == 0
    x, y = _landmarks(1.0, lower=103.5)
    assert _painted(_paint(x, y)).sum() > 0


def test_box_is_clipped_to_the_frame():
    # the face moved 115 px left: the mouth straddles the frame's left edge
    x, y = _landmarks(1.0, dx=-115.0)
    fr = _paint(x, y, w=64, h=160)
    got = _painted(fr)
    poly = POLY + np.array([-115.0, 0.0])
    want = _convex_expectation(poly, 0.0, (0, int(x[31:].max()), 100, 120), 64, 160)
    assert got.sum() > 0 and np.array_equal(got, want)
    # moved down past the bottom edge: rows beyond the frame are never touched
    x, y = _landmarks(1.0, dy=45.0)
    fr = _paint(x, y, w=256, h=160)
    got = _painted(fr)
    want = _convex_expectation(POLY + np.array([0.0, 45.0]), 0.0, (97, 143, 145, 160), 256, 160)
    assert got.sum() > 0 and np.array_equal(got, want)
    # entirely outside the frame: nothing
    x, y = _landmarks(1.0, dx=-400.0)
    assert _painted(_paint(x, y)).sum() == 0


def test_non_finite_landmarks_leave_the_frame():
    x, y = _landmarks(1.0)
    x[35] = np.nan
    assert _painted(_paint(x, y)).sum() == 0
    x, y = _landmarks(1.0)
    y[44] = np.inf
    assert _painted(_paint(x, y)).sum() == 0


def test_frames_are_independent():
    xs, ys = zip(*[_landmarks(u, dx=dx) for u, dx in [(1.0, 0.0), (2.0, 10.0), (1.0, -115.0)]])
    fr = np.full((3, 256, 256, 3), 77, np.uint8)
    many = O.paint_innermouth_black(fr, np.stack(xs), np.stack(ys), REF)
    for f in range(3):
        one = O.paint_innermouth_black(fr[f:f + 1], xs[f][None], ys[f][None], REF)[0]
        assert np.array_equal(many[f], one)
