"""The rasterizers' block test (k_raster.hip, strip_hits) is conservative.

For a staged Gaussian and one pixel block of a tile (a 16 x 4 strip in the
backward, an 8 x 8 quadrant in the forward), the test clears the block's bit
only when the minimum of Q = a u^2 + 2 b u v + c v^2 over the block's
rectangle (grown by half a pixel) exceeds r2 = 2 ln(255 o), widened.  A cleared bit must mean that at every pixel centre of the strip
the alpha the reference rasterizer computes (rasterizer_cpu.cpp:159-169,
float32 without contraction) is below 1/255, so skipping the Gaussian there
changes no result.

The test is restated here in float32, operation by operation, with the
approximate log2 pushed the wrong way (smaller r2) and the approximate
reciprocals perturbed by more than their error.  Conics go up to the
anisotropy limit tr^2/det <= 1000, opacities down to just above 1/255, and
the tiles are placed so that strips straddle the ellipse's edge.  Three
plausible mistakes must be caught: the factor 2 missing from the cross term,
the two edge slopes swapped, and r2 = ln(255 o) instead of 2 ln(255 o)."""
import numpy as np
import pytest

f32 = np.float32


def _conics(rng, n):
    s1 = 10 ** rng.uniform(-0.5, 2.0, n)
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
    return a[ok], b[ok], c[ok], o[ok], np.sqrt(np.maximum(cxx, cyy))[ok]


def _quad(a, b, c, u, v, cross=2.0):
    t1 = (a * u).astype(f32) * u
    t2 = (c * v).astype(f32) * v
    t3 = ((f32(cross) * b).astype(f32) * u).astype(f32) * v
    return ((t1.astype(f32) + t2.astype(f32)).astype(f32) + t3.astype(f32)).astype(f32)


def _blocks(quad):
    """(x offset, y offset, width - 1, height - 1) of the four blocks."""
    if quad:
        return [(8 * (w & 1), 8 * (w >> 1), 7, 7) for w in range(4)]
    return [(0, 4 * w, 15, 3) for w in range(4)]


def _strip_bits(a, b, c, o, mx, my, x0, y0, rng, quad=False, cross=2.0, swap=False, r2_factor=2.0):
    """strip_hits in float32, with log2 pessimised and rcp perturbed."""
    ln2x = (np.log2(255.0 * o.astype(np.float64)) * (1 - 2e-6)).astype(f32)
    r2 = ((f32(r2_factor * np.log(2)) * ln2x).astype(f32) * f32(1.001) + f32(1e-3)).astype(f32)
    G = f32(0.5)

    def rcp(x):
        return (1.0 / x.astype(np.float64) * (1 + rng.choice([-3e-7, 3e-7], len(x)))).astype(f32)

    ka = (-b * rcp(a)).astype(f32)
    kc = (-b * rcp(c)).astype(f32)
    if swap:
        ka, kc = kc, ka
    bits = np.zeros((len(a), 4), bool)
    for w, (ox, oy, bw, bh) in enumerate(_blocks(quad)):
        bx = (x0 + f32(ox)).astype(f32)
        by = (y0 + f32(oy)).astype(f32)
        u0 = ((bx - G).astype(f32) - mx).astype(f32)
        u1 = ((bx + f32(bw + 0.5)).astype(f32) - mx).astype(f32)
        vl, vr = (kc * u0).astype(f32), (kc * u1).astype(f32)
        v0 = ((by - G).astype(f32) - my).astype(f32)
        v1 = ((by + f32(bh + 0.5)).astype(f32) - my).astype(f32)
        q = np.where((u0 <= 0) & (u1 >= 0) & (v0 <= 0) & (v1 >= 0), f32(0), f32(3.0e38))
        q = np.minimum(q, _quad(a, b, c, np.minimum(np.maximum((ka * v0).astype(f32), u0), u1), v0, cross))
        q = np.minimum(q, _quad(a, b, c, np.minimum(np.maximum((ka * v1).astype(f32), u0), u1), v1, cross))
        q = np.minimum(q, _quad(a, b, c, u0, np.minimum(np.maximum(vl, v0), v1), cross))
        q = np.minimum(q, _quad(a, b, c, u1, np.minimum(np.maximum(vr, v0), v1), cross))
        bits[:, w] = q <= r2
    return bits


def _passing_in_cleared(a, b, c, o, mx, my, x0, y0, bits, quad=False):
    """Pixel centres of cleared blocks where the reference alpha passes the cut."""
    bad = 0
    A, B, C, O = a[:, None], b[:, None], c[:, None], o[:, None]
    for w, (ox, oy, bw, bh) in enumerate(_blocks(quad)):
        cleared = ~bits[:, w]
        W, H = bw + 1, bh + 1
        px = (x0[:, None] + f32(ox) + np.tile(np.arange(W, dtype=f32), H)[None, :]).astype(f32)
        py = (y0[:, None] + f32(oy) + np.repeat(np.arange(H, dtype=f32), W)[None, :]).astype(f32)
        dx = (mx[:, None] - px).astype(f32)
        dy = (my[:, None] - py).astype(f32)
        quad = ((A * dx * dx).astype(f32) + (C * dy * dy).astype(f32)).astype(f32)
        power = ((f32(-0.5) * quad).astype(f32) - (B * dx * dy).astype(f32)).astype(f32)
        alpha = np.minimum(f32(0.99), (O * np.exp(power.astype(np.float64)).astype(f32)).astype(f32))
        passing = (power <= 0) & (alpha >= f32(1 / 255))
        bad += int((passing & cleared[:, None]).sum())
    return bad


@pytest.mark.parametrize("quad", [False, True])
def test_strip_test_is_conservative(quad):
    rng = np.random.default_rng(12)
    a, b, c, o, ext = _conics(rng, 40000)
    assert len(a) > 35000
    # tiles whose strips straddle the ellipse's edge: mean within about the
    # ellipse's 3-sigma extent (plus a tile) of the tile
    x0 = np.floor(rng.uniform(0, 4000, len(a))).astype(f32)
    y0 = np.floor(rng.uniform(0, 4000, len(a))).astype(f32)
    reach = (2.5 * ext + 12).astype(np.float64)
    mx = (x0 + 8 + rng.uniform(-1, 1, len(a)) * reach).astype(f32)
    my = (y0 + 8 + rng.uniform(-1, 1, len(a)) * reach).astype(f32)
    bits = _strip_bits(a, b, c, o, mx, my, x0, y0, rng, quad=quad)
    cleared = int((~bits).sum())
    assert cleared > 20000 and bits.sum() > 20000  # both outcomes well represented
    assert _passing_in_cleared(a, b, c, o, mx, my, x0, y0, bits, quad) == 0
    # the same check fails for plausible mistakes
    for kw in ({"cross": 1.0}, {"swap": True}, {"r2_factor": 1.0}):
        wrong = _strip_bits(a, b, c, o, mx, my, x0, y0, rng, quad=quad, **kw)
        assert _passing_in_cleared(a, b, c, o, mx, my, x0, y0, wrong, quad) > 0, kw
