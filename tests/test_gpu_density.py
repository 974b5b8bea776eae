"""Adaptive density control on the device (gs_trainer_densify_and_prune and
the C++ DensityControllerHIP / TrainerHIP::step_with_densify) against the
reference's own DensityController and Trainer (compiled from source into
oracle/_ref) and against the oracle's restatement:
- the reference's unit-test cases (tests/test_density_controller.cpp:73-287)
  and random rounds with clones, splits and all prune rules: every parameter
  bit for bit equal to the reference's;
- 40k Gaussians over two rounds (thousands of splits, so the reference's own
  compaction would read its mask out of bounds): bit for bit equal to the
  oracle, which equals the reference wherever the reference is defined;
- Trainer::step_with_densify over several iterations (SGD): the same counts at
  every step and parameters within float rounding of the reference's."""
import numpy as np
import pytest

from density_helpers import hip_side, inv_sigmoid, make_test_params, reference_side, train_arrays
from oracle.oracle import RefLib

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (no reference tree)")]


@pytest.fixture(scope="module")
def shim_loaded(hip):
    import os
    from density_helpers import SHIM
    if not os.path.exists(SHIM):
        pytest.fail(f"{SHIM} not built (make -C tests/cpp, or __graft_entry__.build())")
    return True


def _same(a, b, what):
    assert len(a) == len(b), what
    for x, y in zip(a, b):
        assert x.shape == y.shape, (what, x.shape, y.shape)
        np.testing.assert_array_equal(x.view(np.uint32), y.view(np.uint32), err_msg=what)


def _case(sides, n, d_means2D, setup=None, extent=100.0, radii=None, **cfg):
    outs = []
    for S in sides:
        arrs = list(make_test_params(n))
        if setup:
            setup(arrs)
        p = S.new_params(*arrs, 0)
        dc = S.dc_new()
        S.reset(dc, n)
        S.accumulate(dc, np.asarray(d_means2D, np.float32), radii)
        m = S.densify(dc, p, extent, **cfg)
        outs.append((m, S.get_params(p, 1)))
        S.dc_free(dc)
        S.params_free(p)
    return outs


def test_density_unit_cases_match_reference(shim_loaded):
    R, H = reference_side(), hip_side()
    big = lambda a: a[1].__setitem__(1, np.log(np.float32(5.0)))
    cases = [
        ("clone", 5, [0.1, 0.1, 0, 0, 0.1, 0.1, 0, 0, 0, 0], None, dict(min_opacity=0.001), 7),
        ("split", 3, [0, 0, 0.5, 0.5, 0, 0], big, dict(min_opacity=0.001), 4),
        ("single", 1, [0.5, 0.5], None, dict(min_opacity=0.001), 2),
        ("clone+split", 3, [0.5, 0.5, 0.5, 0.5, 0, 0], big, dict(min_opacity=0.001), 5),
    ]
    for name, n, d, setup, cfg, expect in cases:
        print("[case] density", name, flush=True)
        (mr, pr), (mh, ph) = _case([R, H], n, d, setup, **cfg)
        assert mr == mh == expect, (name, mr, mh)
        _same(pr, ph, name)

    def low(a):
        a[4][1] = inv_sigmoid(0.001)
        a[4][3] = inv_sigmoid(0.002)
    (mr, pr), (mh, ph) = _case([R, H], 5, np.zeros(10), low)
    assert mr == mh == 3
    _same(pr, ph, "prune")

    def all_low(a):
        a[4][:] = inv_sigmoid(0.001)
    (mr, _), (mh, _) = _case([R, H], 3, np.zeros(6), all_low)
    assert mr == mh == 0
    # no Gaussians at all
    (mr, pr), (mh, ph) = _case([R, H], 0, np.zeros(0), None, min_opacity=0.001)
    assert mr == mh == 0
    _same(pr, ph, "empty")
    # the screen-size rule reads the statistics' largest radius
    (mr, pr), (mh, ph) = _case([R, H], 5, np.zeros(10), None, radii=np.array([0, 25, 3, 40, 20], np.int32))
    assert mr == mh == 3
    _same(pr, ph, "screen size")


def test_density_multiple_cycles_match_reference(shim_loaded):
    """MultipleDensifyCycles (test_density_controller.cpp:214-246): 2 -> 4 -> 8."""
    R, H = reference_side(), hip_side()
    res = []
    for S in (R, H):
        p = S.new_params(*make_test_params(2), 0)
        dc = S.dc_new()
        S.reset(dc, 2)
        S.accumulate(dc, np.full(4, 0.5, np.float32))
        n1 = S.densify(dc, p, 100.0, min_opacity=0.001)
        S.accumulate(dc, np.full(2 * n1, 0.5, np.float32))
        n2 = S.densify(dc, p, 100.0, min_opacity=0.001)
        res.append((n1, n2, S.get_params(p, 1)))
        S.dc_free(dc)
        S.params_free(p)
    assert res[0][:2] == res[1][:2] == (4, 8)
    _same(res[0][2], res[1][2], "cycles")


def _random_params(n, seed, k=16):
    rng = np.random.default_rng(seed)
    return [rng.uniform(-2, 2, (n, 3)).astype(np.float32), rng.normal(-1.5, 1.2, (n, 3)).astype(np.float32),
            rng.normal(0, 1, (n, 4)).astype(np.float32), rng.normal(0, 0.5, (n, k, 3)).astype(np.float32),
            rng.normal(-1.0, 3.0, n).astype(np.float32)]


def _rounds(S, params, seed, cycles, extent, k, oracle=False, **cfg):
    """accumulate x2 + densify, `cycles` times, on one implementation."""
    rng = np.random.default_rng(seed + 1)
    if oracle:
        from oracle.oracle import OracleDensity
        O = OracleDensity()
        cur = [a.copy() for a in params]
    else:
        deg = {1: 0, 4: 1, 9: 2, 16: 3}[k]
        h = S.new_params(*params, deg)
        dc = S.dc_new()
    out = []
    m = params[4].shape[0]
    for _ in range(cycles):
        for _ in range(2):
            d = (rng.normal(0, 1, 2 * m) * np.where(rng.random(2 * m) < 0.5, 1e-3, 1e-5)).astype(np.float32)
            radii = rng.integers(0, 30, m).astype(np.int32)
            if oracle:
                O.accumulate(d, radii)
            else:
                S.accumulate(dc, d, radii)
        if oracle:
            cur = O.densify(cur, extent, **cfg)
            m = len(cur[4])
            out.append([a.copy() for a in cur])
        else:
            m = S.densify(dc, h, extent, **cfg)
            out.append(S.get_params(h, k))
    if not oracle:
        S.dc_free(dc)
        S.params_free(h)
    return out


@pytest.mark.parametrize("n,seed,extent", [(8, 3, 10.0), (10, 4, 30.0), (12, 5, 3.0)])
def test_density_random_rounds_match_reference(shim_loaded, n, seed, extent):
    print("[case] density random", n, seed, extent, flush=True)
    params = _random_params(n, seed)
    ref = _rounds(reference_side(), params, seed, 2, extent, 16, min_opacity=0.05)
    hip = _rounds(hip_side(), params, seed, 2, extent, 16, min_opacity=0.05)
    for r, (a, b) in enumerate(zip(ref, hip)):
        _same(a, b, f"round {r}")


def test_density_large_rounds_match_oracle(shim_loaded):
    """40k Gaussians (SH3), two rounds: thousands of clones, splits and prunes;
    the device's prefix-sum placement and the generator's noise against the
    oracle's sequential restatement, bit for bit."""
    print("[case] density 40k vs oracle", flush=True)
    params = _random_params(40_000, 11)
    ora = _rounds(None, params, 11, 2, 8.0, 16, oracle=True, min_opacity=0.05)
    hip = _rounds(hip_side(), params, 11, 2, 8.0, 16, min_opacity=0.05)
    for r, (a, b) in enumerate(zip(ora, hip)):
        n0 = len(params[4]) if r == 0 else len(ora[r - 1][4])
        print(f"  round {r}: {n0} -> {len(a[4])} Gaussians", flush=True)
        _same(a, b, f"round {r}")
    assert len(ora[0][4]) != len(params[4])


def test_step_with_densify_matches_reference(shim_loaded):
    """Trainer::step_with_densify (SGD, densify every 2 iterations from 0,
    opacity reset at 4) against the reference's Trainer on the same scene:
    the same Gaussian count after every step, parameters within float rounding
    (the gradient sums run in another order on the device).  The extent puts
    every Gaussian under the size cut, so densification clones: the
    reference's compaction after splits reads its mask out of bounds once the
    children pass its last 64-bit word (splits are compared above)."""
    from gs_scenes import RenderCfg, random_small_scene
    print("[case] step_with_densify", flush=True)
    scene, cam, gt = random_small_scene(400, 3, 91, width=64, height=48)
    cfg = RenderCfg(sh_degree=3, training=True)
    dcfg = dict(enabled=True, densify_from_iter=0, densify_until_iter=100, densify_interval=2,
                opacity_reset_interval=4, grad_threshold=2e-5, min_opacity=0.005)
    tc = train_arrays(lr=1e-3, use_adam=False)
    R, H = reference_side(), hip_side()
    ph = [R.new_params(*scene.arrays(), scene.sh_degree), H.new_params(*scene.arrays(), scene.sh_degree)]
    tr = [R.new_trainer(), H.new_trainer()]
    counts = []
    for it in range(6):
        res = [S.step_densify(t, p, cam, cfg, gt, tc, 100.0, it, **dcfg) for S, t, p in zip((R, H), tr, ph)]
        counts.append((res[0][1], res[1][1]))
        assert res[0][1] == res[1][1], (it, res)
        assert abs(res[0][0] - res[1][0]) <= 1e-5, (it, res)
    assert counts[-1][0] != scene.n  # densification changed the count
    a, b = R.get_params(ph[0], scene.max_coeffs), H.get_params(ph[1], scene.max_coeffs)
    for x, y in zip(a, b):
        assert np.abs(x - y).max() <= 1e-4 * max(1.0, float(np.abs(x).max())), np.abs(x - y).max()
    R.trainer_free(tr[0])
    H.trainer_free(tr[1])
    R.params_free(ph[0])
    H.params_free(ph[1])
