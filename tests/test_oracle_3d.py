"""The eval_3D forward oracle (gso_preprocess3d / gso_duplicate3d /
gso_rasterize3d), pinned against the reference's own CPU classes compiled
from source (oracle/_ref: PreprocessorCPU, TileBinnerCPU and RasterizerCPU
with RenderConfig::eval_3D) and against closed forms."""
import ctypes

import numpy as np
import pytest

from helpers import eval3d_cases
from oracle.oracle import Oracle3D, RefLib

CASES = eval3d_cases()
IDS = [c["name"] for c in CASES]


@pytest.fixture(scope="module")
def o3():
    return Oracle3D()


def _visible(d, r):
    return {k: (v[r > 0] if v.shape[:1] == r.shape else v) for k, v in d.items() if isinstance(v, np.ndarray)}


@pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_oracle3d_bitexact_vs_reference(o3, case):
    c = case
    ref = RefLib().render_3d(c["g"], c["n"], c["deg"], c["K"], c["cam"], c["cfg"], c["filter3d"])
    got = o3.render(c["g"], c["n"], c["deg"], c["K"], c["cam"], c["cfg"], c["filter3d"])
    assert ref["K"] == len(got["keys"])
    assert ref["K"] > 0
    np.testing.assert_array_equal(got["radii"], ref["radii"])
    np.testing.assert_array_equal(got["tiles"], ref["tiles"])
    for k in ("means2D", "depths", "op2d", "rgb", "g2s", "cov3D_inv", "mean_offset"):
        np.testing.assert_array_equal(got[k].view(np.uint32), ref[k].view(np.uint32), err_msg=k)
    for k in ("keys", "values", "values_sorted", "ranges", "n_contrib"):
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)
    for k in ("img", "T_final"):
        np.testing.assert_array_equal(got[k].view(np.uint32), ref[k].view(np.uint32), err_msg=k)


def test_eval3d_cases_reach_every_branch(o3):
    """The cases exercise the branches a plausible mistake would hide in: the
    view-space fallback bounds, per-tile depths differing from the view depth,
    and k-buffer overflow (more than 16 blended entries in a pixel)."""
    L = o3.lib
    fallback = 0
    for c in CASES:
        pre = o3.preprocess(c["g"], c["n"], c["deg"], c["K"], c["cam"], c["cfg"], c["filter3d"])
        for i in np.nonzero(pre["radii"] > 0)[0]:
            cutoff = np.float32(min(np.float32(11.11), np.float32(2.0) * np.float32(
                np.log(np.float64(pre["op2d"][i] / np.float32(1.0 / 255.0))))))
            m, e = np.zeros(2, np.float32), np.zeros(2, np.float32)
            g = np.ascontiguousarray(pre["g2s"][i])
            if not L.gso_aabb_screen(g.ctypes.data_as(ctypes.c_void_p), ctypes.c_float(cutoff),
                                     m.ctypes.data_as(ctypes.c_void_p), e.ctypes.data_as(ctypes.c_void_p)):
                fallback += 1
    assert fallback > 0
    c = CASES[3]
    out = o3.render(c["g"], c["n"], c["deg"], c["K"], c["cam"], c["cfg"], c["filter3d"])
    assert out["n_contrib"].max() > 16
    keys = out["keys"]
    vis = out["radii"] > 0
    own = np.repeat(out["depths"][vis].view(np.uint32), out["tiles"][vis])
    assert np.mean((keys & 0xFFFFFFFF).astype(np.uint32) != own) > 0.5


def test_ray_contrib_closed_form(o3):
    """maxContribRay = squared distance from the origin to the line where the
    planes a.(x, 1) = 0 and b.(x, 1) = 0 meet (min-norm point, FP64)."""
    rng = np.random.default_rng(5)
    for _ in range(200):
        a = rng.normal(size=4).astype(np.float32)
        b = rng.normal(size=4).astype(np.float32)
        got, mp = o3.ray_contrib(a, b)
        A = np.stack([a[:3], b[:3]]).astype(np.float64)
        w = -np.array([a[3], b[3]], np.float64)
        p0 = np.linalg.pinv(A) @ w
        assert got == pytest.approx(float(p0 @ p0), rel=2e-4, abs=1e-6)
        np.testing.assert_allclose(mp[:3], p0, rtol=2e-3, atol=1e-4)


def test_invert4_matches_numpy(o3):
    rng = np.random.default_rng(6)
    for _ in range(50):
        M = rng.normal(size=(4, 4)) + 3 * np.eye(4)
        ok, inv = o3.invert4(M.T.reshape(-1).astype(np.float32))  # column-major
        assert ok
        np.testing.assert_allclose(inv.reshape(4, 4).T, np.linalg.inv(M.astype(np.float32).astype(np.float64)),
                                   rtol=1e-4, atol=1e-5)
    ok, _ = o3.invert4(np.zeros(16, np.float32))
    assert not ok


@pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("n,deg,seed", [(300, 1, 41), (3000, 3, 42)])
def test_inv_depth_oracle_bitexact_vs_reference(oracle, n, deg, seed):
    """gso_rasterize_inv_depth against RasterizerCPU::rasterize's out_depth
    (rasterizer_cpu.cpp:170-172,197), both on the reference's own 2D forward."""
    from gs_scenes import RenderCfg, random_small_scene
    from oracle.oracle import inv_depth_oracle
    scene, cam, _ = random_small_scene(n, deg, seed, width=90, height=70)
    cfg = RenderCfg(sh_degree=3, training=False, bg_color=(0.1, 0.2, 0.3))
    g = oracle.activate(scene)
    rimg, rinv = RefLib().inv_depth(g, n, scene.sh_degree, scene.max_coeffs, cam, cfg)
    f = oracle.forward(scene, cam, cfg)
    np.testing.assert_array_equal(f["img"].view(np.uint32), rimg.view(np.uint32))
    inv = inv_depth_oracle(f["pre"], f["values_sorted"], f["ranges"], cam.width, cam.height)
    assert rinv.max() > 0
    np.testing.assert_array_equal(inv.view(np.uint32), rinv.view(np.uint32))
