"""The reference's DensityController (density_controller.cpp, compiled from
source into oracle/_ref) driven through the test drivers, against the
expectations of its own unit tests (tests/test_density_controller.cpp:73-287).
These pin the drivers that the GPU density tests compare the device
implementation with."""
import numpy as np
import pytest

from density_helpers import inv_sigmoid, make_test_params, reference_side
from oracle.oracle import RefLib

pytestmark = pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (no reference tree)")


def _run(n, d_means2D, extent=100.0, setup=None, **cfg):
    R = reference_side()
    arrs = list(make_test_params(n))
    if setup:
        setup(arrs)
    p = R.new_params(*arrs, 0)
    dc = R.dc_new()
    R.reset(dc, n)
    R.accumulate(dc, np.asarray(d_means2D, np.float32))
    new_n = R.densify(dc, p, extent, **cfg)
    out = R.get_params(p, 1)
    R.dc_free(dc)
    R.params_free(p)
    return new_n, out


def test_reference_clone_small():
    d = np.zeros(10, np.float32)
    d[[0, 1, 4, 5]] = 0.1
    n, (pos, *_rest) = _run(5, d, min_opacity=0.001)
    assert n == 7
    assert pos[5, 0] == pos[0, 0] and pos[6, 0] == pos[2, 0]


def test_reference_split_large():
    def big(a):
        a[1][1] = np.log(np.float32(5.0))
    d = np.zeros(6, np.float32)
    d[2:4] = 0.5
    n, (pos, sc, *_r) = _run(3, d, setup=big, min_opacity=0.001)
    assert n == 4
    # the children (last two) have the parent's log-scale minus log(1.6)
    np.testing.assert_allclose(sc[2:], np.log(np.float32(5.0)) - np.log(np.float32(1.6)), rtol=1e-6)


def test_reference_prune_and_empty():
    def low(a):
        a[4][1] = inv_sigmoid(0.001)
        a[4][3] = inv_sigmoid(0.002)
    n, _ = _run(5, np.zeros(10, np.float32), setup=low)
    assert n == 3

    def all_low(a):
        a[4][:] = inv_sigmoid(0.001)
    n, _ = _run(3, np.zeros(6, np.float32), setup=all_low)
    assert n == 0


def test_reference_clone_and_split_same_step():
    def big(a):
        a[1][1] = np.log(np.float32(5.0))
    n, _ = _run(3, [0.5, 0.5, 0.5, 0.5, 0.0, 0.0], setup=big, min_opacity=0.001)
    assert n == 5


def test_reference_opacity_reset():
    R = reference_side()
    p = R.new_params(*make_test_params(3), 0)
    R.reset_opacity(p)
    op = R.get_params(p, 1)[4]
    R.params_free(p)
    assert np.allclose(1.0 / (1.0 + np.exp(-op.astype(np.float64))), 0.01, atol=1e-4)


def _random_params(n, seed, k=16):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-2, 2, (n, 3)).astype(np.float32)
    sc = rng.normal(-1.5, 1.2, (n, 3)).astype(np.float32)  # some above 0.01 * extent, some below
    rot = rng.normal(0, 1, (n, 4)).astype(np.float32)
    sh = rng.normal(0, 0.5, (n, k, 3)).astype(np.float32)
    op = rng.normal(-1.0, 3.0, n).astype(np.float32)  # some below the opacity cut
    return [pos, sc, rot, sh, op]


def _both(n, seed, cycles, extent, k=16, **cfg):
    """The oracle and the reference through `cycles` accumulate + densify
    rounds on the same random inputs; returns the parameter lists per round."""
    from oracle.oracle import OracleDensity
    R = reference_side()
    O = OracleDensity()
    params = _random_params(n, seed, k)
    deg = {1: 0, 4: 1, 9: 2, 16: 3}[k]
    h = R.new_params(*params, deg)
    dc = R.dc_new()
    rng = np.random.default_rng(seed + 1)
    out = []
    cur = [a.copy() for a in params]
    for c in range(cycles):
        m = len(cur[4])
        for _ in range(2):
            d = (rng.normal(0, 1, 2 * m) * np.where(rng.random(2 * m) < 0.5, 1e-3, 1e-5)).astype(np.float32)
            radii = rng.integers(0, 30, m).astype(np.int32)
            R.accumulate(dc, d, radii)
            O.accumulate(d, radii)
        nr = R.densify(dc, h, extent, **cfg)
        cur = O.densify(cur, extent, **cfg)
        out.append((nr, R.get_params(h, k), cur))
    R.dc_free(dc)
    R.params_free(h)
    return out


@pytest.mark.parametrize("n,seed,extent", [(8, 3, 10.0), (10, 4, 30.0), (12, 5, 3.0)])
def test_oracle_densify_equals_reference_with_splits(n, seed, extent):
    """Clones, splits (with the generator's noise), opacity, screen-size and
    world-size pruning over two rounds, bit for bit; small enough that every
    split child lies within the reference's remove mask's first 64-bit word."""
    rounds = _both(n, seed, 2, extent, min_opacity=0.05)
    splits_seen = 0
    for nr, ref, ora in rounds:
        assert nr == len(ora[4])
        for a, b in zip(ref, ora):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
        splits_seen += nr
    assert splits_seen > 0


def test_oracle_densify_equals_reference_clone_and_prune_at_scale():
    """3000 Gaussians, no splits (every Gaussian is below the size cut): clones
    and the opacity and screen-size prune rules at scale, bit for bit."""
    rounds = _both(3000, 7, 2, 1000.0, min_opacity=0.05)
    for nr, ref, ora in rounds:
        assert nr == len(ora[4]) and nr > 0
        for a, b in zip(ref, ora):
            np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
