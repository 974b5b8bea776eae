"""Antialiasing (RenderConfig::antialiasing; preprocessor_cpu.cpp:213-216,250):
the 2D opacity is scaled by h = sqrt(max(2.5e-5, det_cov / det)), the ratio of
the 2D covariance's determinant before and after the 0.3 filter.

- The oracle's forward with antialiasing equals the reference's, bit for bit.
- The reference's backward has no term for h.  The oracle's does (into the
  opacity and, through both determinants, into the 2D covariance and on to
  position, scale and rotation); it agrees with FP64 central differences of
  the forward within 1e-4 of each gradient group's largest entry, while the
  reference's own backward misses the opacity gradient by more than 1%."""
import numpy as np
import pytest

from gs_scenes import RenderCfg, random_small_scene
from oracle.oracle import RefLib

GROUPS = lambda K: {"pos": (0, 3), "scale": (3, 3), "rot": (6, 4), "sh": (10, 3 * K), "op": (10 + 3 * K, 1)}


def _scene(deg):
    return random_small_scene(40, deg, 60 + deg, width=64, height=48, logscale=(-2.2, 0.5), op=(0.0, 0.8),
                              spread=0.7)


def _rows(grads, idx):
    n = len(idx)
    return np.concatenate([grads["pos"][idx], grads["scale"][idx], grads["rot"][idx],
                           grads["sh"][idx].reshape(n, -1), grads["op"][idx][:, None]], 1).astype(np.float64)


def _group_errors(A, G, ok, K):
    out = {}
    for nm, (a, l) in GROUPS(K).items():
        m = ok[:, a:a + l]
        out[nm] = np.abs(A[:, a:a + l] - G[:, a:a + l])[m].max() / np.abs(G[:, a:a + l][m]).max()
    return out


@pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("deg", [0, 3])
def test_antialiased_forward_equals_reference(oracle, deg):
    scene, cam, gt = _scene(deg)
    cfg = RenderCfg(sh_degree=3, training=True, antialiasing=True)
    ref = RefLib().forward_backward(scene, cam, cfg, gt)
    f = oracle.forward(scene, cam, cfg)
    np.testing.assert_array_equal(f["img"].view(np.uint32), ref["img"].view(np.uint32))
    np.testing.assert_array_equal(f["pre"]["op2d"].view(np.uint32), ref["op2d"].view(np.uint32))
    plain = oracle.forward(scene, cam, RenderCfg(sh_degree=3, training=True))
    assert not np.array_equal(plain["img"], f["img"])  # the scaling is in effect


@pytest.mark.parametrize("deg", [0, 3])
def test_antialiased_backward_against_fp64_differences(oracle, deg):
    from oracle.oracle import fd_grads_local
    scene, cam, gt = _scene(deg)
    cfg = RenderCfg(sh_degree=3, training=True, antialiasing=True)
    f = oracle.forward_backward(scene, cam, cfg, gt)
    idx = np.nonzero(np.abs(f["grads"]["op"]) > 0)[0][:20]
    G, steps = fd_grads_local(scene, cam, cfg, gt, idx, threads=8)
    ok = steps > 0
    assert ok.mean() > 0.9
    e = _group_errors(_rows(f["grads"], idx), G, ok, scene.max_coeffs)
    assert max(e.values()) <= 1e-4, e
    if RefLib.available():
        # the reference's backward (no term for h) against the same differences
        ref = RefLib().forward_backward(scene, cam, cfg, gt)
        rg = {k: ref["g_" + k] for k in ["pos", "scale", "rot", "sh", "op"]}
        eref = _group_errors(_rows(rg, idx), G, ok, scene.max_coeffs)
        assert eref["op"] > 1e-2, eref
