"""The C++ drop-in classes (PreprocessorHIP, TileBinnerHIP, SorterHIP,
RasterizerHIP, Renderer, RasterizerBackwardHIP, PreprocessorBackwardHIP,
AdamOptimizerHIP, TrainerHIP) driven through tests/cpp/stage_shim.cpp on host
arrays, as a user of the reference's C++ API would, against the oracle."""
import ctypes
import os

import numpy as np
import pytest

from gs_scenes import RenderCfg, make_workload, random_small_scene
from helpers import assert_grad_close, assert_image_close

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM = os.path.join(HERE, "cpp", "_build", "libgs_stage_shim.so")
_vp = ctypes.c_void_p


@pytest.fixture(scope="module")
def shim(hip):
    if not os.path.exists(SHIM):
        pytest.fail(f"{SHIM} not built (make -C tests/cpp, or __graft_entry__.build())")
    lib = ctypes.CDLL(SHIM)
    lib.shim_error.restype = ctypes.c_char_p
    return lib


def _p(a):
    return a.ctypes.data_as(_vp)


def _check(lib, st, what):
    assert st == 0, f"{what}: {lib.shim_error().decode()}"


def _activated(oracle, scene):
    g = oracle.activate(scene)
    return [np.ascontiguousarray(g[k], np.float32) for k in ["pos", "sh", "scale", "rot", "op"]]


@pytest.mark.parametrize("n,deg,seed", [(400, 3, 81), (2500, 1, 82), (0, 3, 83)])
def test_cpp_renderer_matches_oracle(shim, oracle, n, deg, seed):
    print("[case] cpp renderer", n, deg, seed, flush=True)
    scene, cam, gt = random_small_scene(n, deg, seed, width=120, height=80)
    cfg = RenderCfg(sh_degree=3, training=False, bg_color=(0.1, 0.3, 0.2))
    act = _activated(oracle, scene)
    img = np.zeros((cam.height, cam.width, 3), np.float32)
    c, f = cam.to_c(), cfg.to_c()
    _check(shim, shim.shim_render(n, scene.sh_degree, scene.max_coeffs, *[_p(a) for a in act], ctypes.byref(c),
                                  ctypes.byref(f), _p(img)), "shim_render")
    ref = oracle.forward(scene, cam, cfg)
    assert_image_close(img, ref["img"])


def test_cpp_renderer_inverse_depth(shim, oracle):
    """Renderer::render(..., output_depth) through RasterizerHIP: the image and
    the inverse depth against the oracle (the depth replay restates the
    reference's exact alpha, so it is bit-exact wherever the forward stopped at
    the reference's last contributor)."""
    from oracle.oracle import inv_depth_oracle
    print("[case] cpp renderer inverse depth", flush=True)
    scene, cam, gt = random_small_scene(1500, 3, 84, width=120, height=80)
    cfg = RenderCfg(sh_degree=3, training=False)
    act = _activated(oracle, scene)
    img = np.zeros((cam.height, cam.width, 3), np.float32)
    depth = np.zeros((cam.height, cam.width), np.float32)
    c, f = cam.to_c(), cfg.to_c()
    _check(shim, shim.shim_render_depth(scene.n, scene.sh_degree, scene.max_coeffs, *[_p(a) for a in act],
                                        ctypes.byref(c), ctypes.byref(f), _p(img), _p(depth)), "shim_render_depth")
    ref = oracle.forward(scene, cam, cfg)
    assert_image_close(img, ref["img"])
    inv = inv_depth_oracle(ref["pre"], ref["values_sorted"], ref["ranges"], cam.width, cam.height)
    assert inv.max() > 0
    assert (depth == inv).mean() > 0.99
    assert np.abs(depth - inv).max() <= 1e-4 * max(1.0, float(inv.max()))


def test_cpp_l1_loss_fixtures(shim):
    """l1_loss / l1_loss_double (loss.h:6-10) on host arrays against the
    reference's l1_verify fixtures: d_image bit-exact, the float loss within
    1e-6 relative of the reference's serial float sum, the double loss within
    1e-12 relative of an FP64 sum of the float differences (loss.cpp:18-30)."""
    import json
    from helpers import GOLDEN
    base = os.path.join(GOLDEN, "l1_verify")
    for case in json.load(open(os.path.join(base, "manifest.json"))):
        print("[case] cpp l1", case["name"], flush=True)
        H, W = case["H"], case["W"]
        rd = lambda k: np.fromfile(os.path.join(base, f"{case['name']}_{k}.bin"), np.float32).reshape(H, W, 3)
        r, g, want = rd("rendered"), rd("gt"), rd("d_image")
        d1 = np.zeros_like(r)
        d2 = np.zeros_like(r)
        lf = ctypes.c_float()
        ld = ctypes.c_double()
        _check(shim, shim.shim_l1(_p(r), _p(g), H, W, _p(d1), _p(d2), ctypes.byref(lf), ctypes.byref(ld)), "shim_l1")
        np.testing.assert_array_equal(d1, want)
        np.testing.assert_array_equal(d2, want)
        assert abs(lf.value - case["loss_serial"]) <= 1e-6 * abs(case["loss_serial"])
        exact = np.abs((r - g).astype(np.float64)).sum() / r.size  # float differences, FP64 sum
        assert abs(ld.value - exact) <= 1e-12 * exact


def test_cpp_stages_outputs_match_oracle(shim, oracle):
    print("[case] cpp stages", flush=True)
    scene, cams, _ = make_workload("10k_256")
    cam = cams[0]
    n = scene.n
    cfg = RenderCfg(sh_degree=3, training=True)
    act = _activated(oracle, scene)
    ref = oracle.forward(scene, cam, cfg)
    cap = len(ref["keys"]) + 16
    out = dict(means2D=np.zeros((n, 2), np.float32), depths=np.zeros(n, np.float32), conics=np.zeros((n, 3), np.float32),
               op2d=np.zeros(n, np.float32), rgb=np.zeros((n, 3), np.float32), radii=np.zeros(n, np.int32),
               tiles=np.zeros(n, np.int32))
    ks = np.zeros(cap, np.uint64)
    vs = np.zeros(cap, np.uint32)
    rg = np.zeros(cam.num_tiles * 2, np.uint32)
    img = np.zeros((cam.height, cam.width, 3), np.float32)
    T = np.zeros((cam.height, cam.width), np.float32)
    nc = np.zeros((cam.height, cam.width), np.int32)
    K = ctypes.c_longlong()
    c, f = cam.to_c(), cfg.to_c()
    _check(shim, shim.shim_stages(n, scene.sh_degree, scene.max_coeffs, *[_p(a) for a in act], ctypes.byref(c),
                                  ctypes.byref(f), *[_p(out[k]) for k in ["means2D", "depths", "conics", "op2d", "rgb",
                                                                          "radii", "tiles"]],
                                  ctypes.c_longlong(cap), ctypes.byref(K), _p(ks), _p(vs), _p(rg), _p(img), _p(T), _p(nc)),
           "shim_stages")
    pre = ref["pre"]
    assert np.array_equal(out["radii"], pre["radii"]) and np.array_equal(out["tiles"], pre["tiles"])
    for k in ["means2D", "depths", "conics", "op2d", "rgb"]:
        assert np.abs(out[k] - pre[k]).max() <= 1e-4, k
    assert K.value == len(ref["keys"])
    assert np.array_equal(ks[:K.value], ref["keys_sorted"])
    assert np.array_equal(vs[:K.value], ref["values_sorted"])
    assert np.array_equal(rg, ref["ranges"])
    assert_image_close(img, ref["img"])
    assert np.array_equal(nc, ref["n_contrib"])
    assert np.abs(T - ref["T_final"]).max() <= 1e-5


def test_cpp_stages_accept_host_inputs(shim, oracle):
    """Outputs of a CPU preprocessor (host arrays, no device data) through the
    HIP binner, sorter and rasterizer."""
    print("[case] cpp from host", flush=True)
    scene, cam, gt = random_small_scene(1500, 2, 83, width=100, height=90)
    cfg = RenderCfg(sh_degree=3, training=True)
    ref = oracle.forward(scene, cam, cfg)
    pre = {k: np.ascontiguousarray(v) for k, v in ref["pre"].items()}
    img = np.zeros((cam.height, cam.width, 3), np.float32)
    c, f = cam.to_c(), cfg.to_c()
    _check(shim, shim.shim_from_host(scene.n, *[_p(pre[k]) for k in ["means2D", "depths", "conics", "op2d", "rgb",
                                                                     "radii", "tiles"]],
                                     ctypes.byref(c), ctypes.byref(f), _p(img)), "shim_from_host")
    assert_image_close(img, ref["img"])


@pytest.mark.parametrize("n,deg,seed", [(600, 3, 84), (2000, 0, 85)])
def test_cpp_backward_classes_match_oracle(shim, oracle, n, deg, seed):
    print("[case] cpp backward", n, deg, seed, flush=True)
    scene, cam, gt = random_small_scene(n, deg, seed, width=110, height=70)
    cfg = RenderCfg(sh_degree=3, training=True)
    ref = oracle.forward_backward(scene, cam, cfg, gt)
    raw = [np.ascontiguousarray(a, np.float32).copy() for a in scene.arrays()]
    d_img = np.ascontiguousarray(ref["d_image"], np.float32)
    rg = dict(d_means2D=np.zeros((n, 2), np.float32), d_conics=np.zeros((n, 3), np.float32),
              d_rgb=np.zeros((n, 3), np.float32), d_op2d=np.zeros(n, np.float32))
    gr = dict(pos=np.zeros((n, 3), np.float32), scale=np.zeros((n, 3), np.float32), rot=np.zeros((n, 4), np.float32),
              sh=np.zeros((n, scene.max_coeffs, 3), np.float32), op=np.zeros(n, np.float32))
    c, f = cam.to_c(), cfg.to_c()
    _check(shim, shim.shim_backward(n, scene.sh_degree, scene.max_coeffs, *[_p(a) for a in raw], ctypes.byref(c),
                                    ctypes.byref(f), _p(d_img), *[_p(rg[k]) for k in rg], *[_p(gr[k]) for k in gr]),
           "shim_backward")
    for k in rg:
        assert_grad_close(rg[k], ref["rgrad"][k], k)
    for k in gr:
        assert_grad_close(gr[k], ref["grads"][k], "grad_" + k)


def test_cpp_adam_matches_oracle(shim, oracle):
    print("[case] cpp adam", flush=True)
    rng = np.random.default_rng(9)
    n, mc = 777, 4
    shapes = [(n, 3), (n, 3), (n, 4), (n, mc, 3), (n,)]
    p = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    g = [rng.standard_normal(s).astype(np.float32) for s in shapes]
    pg = [a.copy() for a in p]
    _check(shim, shim.shim_adam(n, 1, mc, *[_p(a) for a in pg], *[_p(a) for a in g], 3), "shim_adam")
    m = [np.zeros_like(a) for a in p]
    v = [np.zeros_like(a) for a in p]
    lr = [oracle.lr_schedule(0.00016, 0.0000016, s, 30000) for s in range(3)]
    for s in range(3):
        bc1, bc2 = oracle.bias_corrections(0.9, 0.999, s + 1)
        lrs = [lr[s], 0.005, 0.001, 0.0025, 0.025]
        for k in range(5):
            oracle.adam_update(p[k], g[k], m[k], v[k], lrs[k], 0.9, 0.999, 1e-8, bc1, bc2)
    for a, b in zip(pg, p):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("use_adam", [1, 0])
def test_cpp_trainer_first_step_matches_oracle(shim, oracle, use_adam):
    print("[case] cpp trainer", use_adam, flush=True)
    scene, cam, gt = random_small_scene(900, 3, 86, width=96, height=72)
    cfg = RenderCfg(sh_degree=3, training=True)
    raw = [np.ascontiguousarray(a, np.float32).copy() for a in scene.arrays()]
    losses = np.zeros(3, np.float32)
    img = np.zeros((cam.height, cam.width, 3), np.float32)
    c, f = cam.to_c(), cfg.to_c()
    _check(shim, shim.shim_trainer(scene.n, scene.sh_degree, scene.max_coeffs, *[_p(a) for a in raw], ctypes.byref(c),
                                   ctypes.byref(f), _p(np.ascontiguousarray(gt, np.float32)), use_adam, 3,
                                   _p(losses), _p(img)), "shim_trainer")
    ref = oracle.forward_backward(scene, cam, cfg, gt)
    assert abs(losses[0] - ref["loss"]) <= 1e-6
    assert losses[2] <= losses[0] + 1e-6  # the loss does not increase over steps


def test_cpp_stages_eval3d_match_oracle(shim):
    """RenderConfig::eval_3D through PreprocessorHIP, TileBinnerHIP, SorterHIP
    and RasterizerHIP, against the oracle's 3D forward."""
    from helpers import eval3d_cases
    from oracle.oracle import Oracle3D
    o3 = Oracle3D()
    for c in eval3d_cases()[1:3]:
        print("[case] cpp eval3d", c["name"], flush=True)
        g = c["g"]
        arr = [np.ascontiguousarray(g[k], np.float32) for k in ["pos", "sh", "scale", "rot", "op"]]
        flt = None if c["filter3d"] is None else np.ascontiguousarray(c["filter3d"], np.float32)
        H, W = c["cam"].height, c["cam"].width
        img = np.zeros((H, W, 3), np.float32)
        T = np.zeros((H, W), np.float32)
        nc = np.zeros((H, W), np.int32)
        radii = np.zeros(c["n"], np.int32)
        depth = np.full((H, W), 7.0, np.float32)
        cam, cfg = c["cam"].to_c(), c["cfg"].to_c()
        _check(shim, shim.shim_render_3d(c["n"], c["deg"], c["K"], *[_p(a) for a in arr],
                                         _p(flt) if flt is not None else None, ctypes.byref(cam), ctypes.byref(cfg),
                                         _p(img), _p(T), _p(nc), _p(radii), _p(depth)), "shim_render_3d")
        # rasterizer_cpu.cpp:53,197: the 3D branch writes an inverse depth of 0
        assert not depth.any()
        ref = o3.render(g, c["n"], c["deg"], c["K"], c["cam"], c["cfg"], c["filter3d"])
        np.testing.assert_array_equal(radii, ref["radii"])
        np.testing.assert_array_equal(nc, ref["n_contrib"])
        assert_image_close(img, ref["img"])
