"""The rasteriser scans only part of a triangle's [floor, ceil) box for small, well-shaped
triangles: the integers within its vertex box widened by 1e-3 px.  That keeps every decision of
the reference's FP32 coverage test (funcs.cu:950-979, cover_rn) only if no sample outside the
widened box can pass it.  This checks the claim by brute force on the FP32 operation sequence
(each product and sum rounded, as in cover_rn), including near-degenerate shapes, and checks that
the shape test is what separates the triangles for which it fails.
"""
import numpy as np

f32 = np.float32


def _cover(B, C, A, px, py):
    v0x, v1x, v2x = f32(B[:, 0] - A[:, 0]), f32(C[:, 0] - A[:, 0]), f32(px - A[:, 0])
    v0y, v1y, v2y = f32(B[:, 1] - A[:, 1]), f32(C[:, 1] - A[:, 1]), f32(py - A[:, 1])
    d00 = f32(f32(v0x * v0x) + f32(v0y * v0y))
    d01 = f32(f32(v0x * v1x) + f32(v0y * v1y))
    d11 = f32(f32(v1x * v1x) + f32(v1y * v1y))
    d20 = f32(f32(v0x * v2x) + f32(v0y * v2y))
    d21 = f32(f32(v1x * v2x) + f32(v1y * v2y))
    den = f32(f32(d00 * d11) - f32(d01 * d01))
    with np.errstate(all="ignore"):
        na = f32(f32(d11 * d20) - f32(d01 * d21))
        nb = f32(f32(d00 * d21) - f32(d01 * d20))
        a, b = f32(na / den), f32(nb / den)
        cov = (a > 0) & (b > 0) & (f32(a + b) <= 1)
    return cov, d00, d11, den


def _triangles(rng, n):
    base = rng.uniform(0, 2048, (n, 2))
    size = 10 ** rng.uniform(-3, np.log10(4), n)
    ang = rng.uniform(0, 2 * np.pi, (n, 3))
    rad = size[:, None] * rng.uniform(0.1, 1.0, (n, 3))
    skinny = rng.random(n) < 0.5
    ang[skinny, 1] = ang[skinny, 0] + np.pi + 10 ** rng.uniform(-6, -1, skinny.sum())
    ang[skinny, 2] = ang[skinny, 0] + 10 ** rng.uniform(-6, -1, skinny.sum())
    V = (base[:, None, :] + rad[..., None] * np.stack([np.cos(ang), np.sin(ang)], -1)).astype(f32)
    return V[:, 0], V[:, 1], V[:, 2]


def test_no_sample_outside_the_widened_box_is_covered():
    rng = np.random.default_rng(2024)
    outside_well, outside_degenerate = 0, 0
    for _ in range(3):
        B, C, A = _triangles(rng, 400_000)
        xs = np.stack([B[:, 0], C[:, 0], A[:, 0]], 1)
        ys = np.stack([B[:, 1], C[:, 1], A[:, 1]], 1)
        xmin, xmax, ymin, ymax = xs.min(1), xs.max(1), ys.min(1), ys.max(1)
        x0, y0 = np.floor(xmin), np.floor(ymin)
        for dx in range(6):
            for dy in range(6):
                px, py = (x0 + dx).astype(f32), (y0 + dy).astype(f32)
                inbox = (px < np.ceil(xmax)) & (py < np.ceil(ymax))
                cov, d00, d11, den = _cover(B, C, A, px, py)
                cov &= inbox
                well = (xmax - xmin <= 4) & (ymax - ymin <= 4) & (d00 > 0) & (d11 > 0) & \
                       (den >= f32(1e-4) * f32(d00 * d11))
                out = (px < xmin - f32(1e-3)) | (px > xmax + f32(1e-3)) | (py < ymin - f32(1e-3)) | \
                      (py > ymax + f32(1e-3))
                outside_well += int((cov & out & well).sum())
                outside_degenerate += int((cov & out & ~well).sum())
    assert outside_well == 0
    # the shape test matters: degenerate triangles do cover samples outside their vertex box
    assert outside_degenerate > 0
