"""GPU parity of the eval_3D forward (gs_preprocess_fwd_3d, gs_duplicate_keys_3d,
gs_sort_pairs on 3D keys, gs_raster_fwd_3d, and the fast binning
gs_bin_tiles_3d) against the oracle's restatement of the reference's CPU 3D
path, stage by stage on identical inputs.

Bars: radii, tiles, keys, values, sorted values, tile ranges and n_contrib
bit-exact; the per-Gaussian floats bit-exact (the same IEEE operations in the
same order); image and T_final within 1e-4 max-abs.
"""
import sys

import numpy as np
import pytest
import torch

from gs_scenes import RenderCfg, gaussians_normal, make_camera
from helpers import IMAGE_TOL, eval3d_cases

pytestmark = pytest.mark.gpu

CASES = eval3d_cases(scale=1) + [c for c in eval3d_cases(scale=3) if c["name"] != "reference_eval3d_scene"]
IDS = [f'{c["name"]}_{c["n"]}' for c in CASES]


def _say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def _np(t):
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def o3():
    from oracle.oracle import Oracle3D
    return Oracle3D()


def _upload(g, n, deg, K, dev):
    f = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(dev)
    return dict(pos=f(g["pos"]), scale=f(g["scale"]), rot=f(g["rot"]), sh=f(g["sh"]), op=f(g["op"]),
                sh_degree=deg, max_coeffs=K, n=n)


@pytest.mark.parametrize("case", CASES, ids=IDS)
def test_eval3d_stages_match_oracle(hip, o3, case):
    c = case
    _say("eval3d", c["name"], c["n"], c["cam"].width, c["cam"].height)
    p = _upload(c["g"], c["n"], c["deg"], c["K"], hip.dev)
    flt = None if c["filter3d"] is None else torch.from_numpy(c["filter3d"]).to(hip.dev)
    pre, pre3d = hip.preprocess_3d(p, c["cam"], c["cfg"], raw=False, filter3d=flt)
    hip.sync()
    ref = o3.preprocess(c["g"], c["n"], c["deg"], c["K"], c["cam"], c["cfg"], c["filter3d"])
    got = dict(means2D=_np(pre["means2D"]), depths=_np(pre["depths"]), op2d=_np(pre["op2d"]), rgb=_np(pre["rgb"]),
               radii=_np(pre["radii"]), tiles=_np(pre["tiles"]), g2s=_np(pre3d["g2s"]),
               cov3D_inv=_np(pre3d["cov3D_inv"]), mean_offset=_np(pre3d["mean_offset"]))
    np.testing.assert_array_equal(got["radii"], ref["radii"])
    np.testing.assert_array_equal(got["tiles"], ref["tiles"])
    for k in ("means2D", "depths", "op2d", "rgb", "g2s", "cov3D_inv", "mean_offset"):
        np.testing.assert_array_equal(got[k].view(np.uint32), ref[k].view(np.uint32), err_msg=k)

    # binning on the device's own preprocess output
    offsets, K = hip.scan_tiles(pre["tiles"])
    assert K == int(ref["tiles"].sum()) and K > 0
    keys, vals = hip.duplicate_3d(pre, pre3d, offsets, K, c["cam"])
    rk, rv = o3.duplicate(got, c["n"], c["cam"])
    np.testing.assert_array_equal(_np(keys).view(np.uint64), rk)
    np.testing.assert_array_equal(_np(vals).view(np.uint32), rv)
    gx, gy = (c["cam"].width + 15) // 16, (c["cam"].height + 15) // 16
    ks, vs, ranges = hip.sort_pairs(keys, vals, gx * gy)
    sk, sv, sr = o3.sort(rk, rv, gx * gy)
    np.testing.assert_array_equal(_np(vs).view(np.uint32), sv)
    np.testing.assert_array_equal(_np(ranges).view(np.uint32).reshape(-1, 2), sr)
    # the fast path (direct placement + per-tile sort) gives the same pairs, keys and ranges
    fv, fk, fr, fK = hip.bin_tiles_3d(pre, pre3d, c["cam"])
    assert fK == K
    np.testing.assert_array_equal(_np(fv).view(np.uint32), sv)
    np.testing.assert_array_equal(_np(fk).view(np.uint64), sk)
    np.testing.assert_array_equal(_np(fr).view(np.uint32).reshape(-1, 2), sr)

    img, T, nc = hip.raster_fwd_3d(pre, pre3d, vs, ranges, c["cam"], tuple(c["cfg"].bg_color))
    hip.sync()
    rimg, rT, rnc = o3.rasterize(got, sv, sr, c["cam"], c["cfg"].bg_color)
    np.testing.assert_array_equal(_np(nc), rnc)
    assert np.abs(_np(img) - rimg).max() <= IMAGE_TOL
    assert np.abs(_np(T) - rT).max() <= IMAGE_TOL


def test_eval3d_raw_parameters_and_whole_chain(hip, o3):
    """Raw (pre-activation) parameters through the whole chain against the
    oracle's activation + 3D forward."""
    from oracle.oracle import Oracle
    scene = gaussians_normal(6000, 21, 2, sigma=1.2, logscale=(-3.8, 0.5), op=(0.5, 1.0))
    cam = make_camera((0.5, -0.4, -3.0), 200, 150, 60.0)
    cfg = RenderCfg(bg_color=(0.2, 0.1, 0.0), training=False, sh_degree=2)
    _say("eval3d raw chain", scene.n)
    out = hip.render_3d(hip.upload_scene(scene), cam, cfg, raw=True)
    hip.sync()
    g = Oracle().activate(scene)
    ref = o3.render(g, scene.n, scene.sh_degree, scene.max_coeffs, cam, cfg)
    np.testing.assert_array_equal(_np(out["pre"]["radii"]), ref["radii"])
    np.testing.assert_array_equal(_np(out["values_sorted"]).view(np.uint32), ref["values_sorted"])
    np.testing.assert_array_equal(_np(out["n_contrib"]), ref["n_contrib"])
    assert np.abs(_np(out["img"]) - ref["img"]).max() <= IMAGE_TOL


def test_eval3d_nothing_visible(hip):
    """Every Gaussian behind the camera: no pairs, the image is the background."""
    scene = gaussians_normal(500, 22, 0, sigma=0.2)
    scene.pos[:, 2] -= 10.0
    cam = make_camera((0.0, 0.0, -3.0), 48, 40, 60.0)
    cfg = RenderCfg(bg_color=(0.25, 0.5, 0.75), training=False, sh_degree=0)
    _say("eval3d empty")
    out = hip.render_3d(hip.upload_scene(scene), cam, cfg, raw=True)
    hip.sync()
    assert out["K"] == 0
    assert int(_np(out["pre"]["radii"]).max()) == 0
    np.testing.assert_array_equal(_np(out["img"]), np.broadcast_to(np.float32([0.25, 0.5, 0.75]), (40, 48, 3)))
    assert int(_np(out["n_contrib"]).max()) == 0


@pytest.mark.parametrize("n,w,h", [(40_000, 64, 64), (60_000, 160, 96)])
def test_eval3d_fast_binning_long_tile_lists(hip, o3, n, w, h):
    """Tiles holding more pairs than one workgroup sorts in LDS (16384): a
    dense cluster in front of the camera.  The fast binning's long lists go
    through the radix sort; values, keys and ranges equal the oracle's
    duplicate + stable sort bit for bit, and the image the stage path's."""
    _say("eval3d long lists", n, w, h)
    scene = gaussians_normal(n, 31, 1, sigma=0.25, logscale=(-1.8, 0.3), op=(0.0, 1.0))
    cam = make_camera((0.0, 0.0, -3.0), w, h, 60.0)
    cfg = RenderCfg(training=False, sh_degree=1)
    p = hip.upload_scene(scene)
    pre, pre3d = hip.preprocess_3d(p, cam, cfg, raw=True)
    fv, fk, fr, K = hip.bin_tiles_3d(pre, pre3d, cam)
    hip.sync()
    per_tile = np.diff(_np(fr).view(np.uint32).reshape(-1, 2), axis=1).ravel()
    assert per_tile.max() > 16384, per_tile.max()
    got = dict(means2D=_np(pre["means2D"]), depths=_np(pre["depths"]), radii=_np(pre["radii"]),
               tiles=_np(pre["tiles"]), cov3D_inv=_np(pre3d["cov3D_inv"]),
               mean_offset=_np(pre3d["mean_offset"]))
    rk, rv = o3.duplicate(got, n, cam)
    gx, gy = (w + 15) // 16, (h + 15) // 16
    sk, sv, sr = o3.sort(rk, rv, gx * gy)
    assert K == len(sv)
    np.testing.assert_array_equal(_np(fv).view(np.uint32), sv)
    np.testing.assert_array_equal(_np(fk).view(np.uint64), sk)
    np.testing.assert_array_equal(_np(fr).view(np.uint32).reshape(-1, 2), sr)
    a = hip.render_3d(p, cam, cfg, raw=True)
    b = hip.render_3d(p, cam, cfg, raw=True, fast=True)
    hip.sync()
    assert torch.equal(a["img"], b["img"]) and torch.equal(a["n_contrib"], b["n_contrib"])


def test_eval3d_fast_binning_full_size(hip):
    """1M Gaussians (SH3), 1920x1080: the fast binning equals the stage path
    (gs_duplicate_keys_3d + gs_sort_pairs, bit-exact against the oracle in the
    tests above) bit for bit: values, keys and ranges."""
    from gs_scenes import make_workload
    _say("eval3d fast binning 1m 1080p")
    scene, cams, _ = make_workload("1m_1080p")
    cam = cams[0]
    cfg = RenderCfg(training=False, sh_degree=3)
    p = hip.upload_scene(scene)
    pre, pre3d = hip.preprocess_3d(p, cam, cfg, raw=True)
    offsets, K = hip.scan_tiles(pre["tiles"])
    keys, vals = hip.duplicate_3d(pre, pre3d, offsets, K, cam)
    ks, vs, ranges = hip.sort_pairs(keys, vals, ((cam.width + 15) // 16) * ((cam.height + 15) // 16))
    del keys, vals
    fv, fk, fr, fK = hip.bin_tiles_3d(pre, pre3d, cam)
    hip.sync()
    assert fK == K and K > 40_000_000
    assert torch.equal(fv, vs)
    assert torch.equal(fk, ks)
    assert torch.equal(fr, ranges)


def test_eval3d_fast_binning_errors(hip):
    """gs_bin_tiles_3d refuses a pair buffer smaller than the pair count
    (GS_ERR_CAPACITY, the count still reported) and views wider than 255 tiles
    (GS_ERR_UNSUPPORTED: the stage path serves those)."""
    from gs_hip import GsError
    scene = gaussians_normal(3000, 41, 0, sigma=0.8, logscale=(-3.0, 0.4))
    cam = make_camera((0.0, 0.0, -3.0), 96, 64, 60.0)
    cfg = RenderCfg(training=False, sh_degree=0)
    p = hip.upload_scene(scene)
    pre, pre3d = hip.preprocess_3d(p, cam, cfg, raw=True)
    K = int(pre["tiles"].to(torch.int64).sum().item())
    assert K > 8
    with pytest.raises(GsError, match="capacity"):
        hip.bin_tiles_3d(pre, pre3d, cam, capacity=K - 1)
    vs, ks, rg, K2 = hip.bin_tiles_3d(pre, pre3d, cam, capacity=K)
    assert K2 == K
    wide = make_camera((0.0, 0.0, -3.0), 4100, 64, 60.0)
    pre_w, pre3d_w = hip.preprocess_3d(p, wide, cfg, raw=True)
    with pytest.raises(GsError, match="255"):
        hip.bin_tiles_3d(pre_w, pre3d_w, wide, capacity=1 << 20)
