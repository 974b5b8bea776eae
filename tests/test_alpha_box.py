"""The rasterizers' alpha box (k_raster.hip, alpha_box) is conservative.

At every pixel centre outside the box, the alpha that the reference rasterizer
computes (rasterizer_cpu.cpp:159-169, float32 without contraction) is below
1/255. So skipping the Gaussian there changes no result.

The box is restated here with each approximate operation (log2, rcp, sqrt)
pushed the wrong way by more than its error. The conics go up to the
anisotropy limit tr^2/det <= 1000. The check is shown to be sensitive by two
plausible mistakes that it must catch: swapped half-extents, and the factor
2 missing from r^2."""
import numpy as np

f32 = np.float32


def _conics(rng, n):
    s1 = 10 ** rng.uniform(-0.5, 2.7, n)
    s2 = s1 * 10 ** rng.uniform(-1.6, 0.0, n)
    th = rng.uniform(0.0, np.pi, n)
    ct, st = np.cos(th), np.sin(th)
    # covariance R diag(s1^2, s2^2) R^T plus the 0.3 px^2 dilation, inverted
    cxx = ct * ct * s1 * s1 + st * st * s2 * s2 + 0.3
    cyy = st * st * s1 * s1 + ct * ct * s2 * s2 + 0.3
    cxy = ct * st * (s1 * s1 - s2 * s2)
    d = cxx * cyy - cxy * cxy
    a, b, c = (cyy / d).astype(f32), (-cxy / d).astype(f32), (cxx / d).astype(f32)
    pick = rng.integers(0, 3, n)
    o = np.where(pick == 0, (1 / 255) * (1 + 10 ** rng.uniform(-7, 0, n)),
                 np.where(pick == 1, rng.uniform(0.004, 1.0, n), 1.0)).astype(f32)
    det = (a * c - b * b).astype(f32)
    tr = (a + c).astype(f32)
    ok = (o >= f32(1 / 255)) & (det > 0) & (tr * tr <= f32(1000) * det)
    return a[ok], b[ok], c[ok], o[ok], det[ok]


def _box(a, c, o, det, mx, my, r2_factor=2.0, swap=False):
    """alpha_box with every approximation pessimised (smaller box)."""
    ln2x = np.log2(255.0 * o.astype(np.float64)) * (1 - 2e-6)
    r2 = (f32(r2_factor * np.log(2)) * ln2x.astype(f32) * f32(1.001) + f32(1e-3)).astype(f32)
    inv = (1.0 / det.astype(np.float64) * (1 + 2e-7)).astype(f32)
    hx, hy = (a, c) if swap else (c, a)
    ex = (np.sqrt((r2 * hx * inv).astype(np.float64)) * (1 - 2e-7)).astype(f32) + f32(1)
    ey = (np.sqrt((r2 * hy * inv).astype(np.float64)) * (1 - 2e-7)).astype(f32) + f32(1)
    return mx - ex, mx + ex, my - ey, my + ey


def _passing_outside(a, b, c, o, box, mx, my):
    """Counts pixel centres just outside the box (61 along each side) whose
    reference alpha passes the 1/255 cut."""
    x0, x1, y0, y1 = box
    u = np.linspace(0.0, 1.0, 61)[None, :]
    fx = np.floor(x0[:, None] + (x1 - x0)[:, None] * u)
    fy = np.floor(y0[:, None] + (y1 - y0)[:, None] * u)
    left = np.floor(x0)[:, None] - (np.floor(x0) == x0)[:, None]
    right = np.floor(x1)[:, None] + 1
    top = np.floor(y0)[:, None] - (np.floor(y0) == y0)[:, None]
    bottom = np.floor(y1)[:, None] + 1
    px = np.concatenate([np.broadcast_to(left, fy.shape), np.broadcast_to(right, fy.shape), fx, fx], 1).astype(f32)
    py = np.concatenate([fy, fy, np.broadcast_to(top, fx.shape), np.broadcast_to(bottom, fx.shape)], 1).astype(f32)
    assert ((px < x0[:, None]) | (px > x1[:, None]) | (py < y0[:, None]) | (py > y1[:, None])).all()
    dx = (mx[:, None] - px).astype(f32)
    dy = (my[:, None] - py).astype(f32)
    A, B, C = a[:, None], b[:, None], c[:, None]
    quad = ((A * dx * dx).astype(f32) + (C * dy * dy).astype(f32)).astype(f32)
    power = ((f32(-0.5) * quad).astype(f32) - (B * dx * dy).astype(f32)).astype(f32)
    alpha = np.minimum(f32(0.99), (o[:, None] * np.exp(power.astype(np.float64)).astype(f32)).astype(f32))
    return int(((power <= 0) & (alpha >= f32(1 / 255))).sum())


def test_alpha_box_is_conservative():
    rng = np.random.default_rng(11)
    a, b, c, o, det = _conics(rng, 20000)
    assert len(a) > 18000
    mx = rng.uniform(0, 2000, len(a)).astype(f32)
    my = rng.uniform(0, 2000, len(a)).astype(f32)
    assert _passing_outside(a, b, c, o, _box(a, c, o, det, mx, my), mx, my) == 0
    # the same check fails for a wrong box
    assert _passing_outside(a, b, c, o, _box(a, c, o, det, mx, my, swap=True), mx, my) > 0
    assert _passing_outside(a, b, c, o, _box(a, c, o, det, mx, my, r2_factor=1.0), mx, my) > 0
