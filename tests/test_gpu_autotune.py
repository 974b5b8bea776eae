"""The autotuners (gs_autotune, gs_batch_autotune): the chosen knobs are
valid candidates, the parameters are untouched, and the path run with the
chosen knobs still matches the oracle."""
import numpy as np
import pytest
import torch

from gs_scenes import RenderCfg, make_camera, random_small_scene, sphere_eye, target_image
from helpers import assert_grad_close, assert_image_close

pytestmark = pytest.mark.gpu


def _valid(t):
    assert 4 <= t["radix_bits"] <= 8
    assert t["raster_batch"] in (64, 128, 256, 512)
    assert t["raster_bwd_batch"] in (64, 128, 256)
    assert t["bin_tiles_per_lane"] in (1, 2, 4)
    assert t["bin_group"] >= 256 and t["bin_group"] & (t["bin_group"] - 1) == 0
    assert t["binning_ms"] > 0 and t["raster_fwd_ms"] > 0 and t["raster_bwd_ms"] > 0


def test_autotune_keeps_parameters_and_parity(oracle):
    from gs_hip import GsHip, GsTrainer, train_cfg
    print("[case] autotune single view", flush=True)
    hip = GsHip(0)  # own context: the tuned knobs stay local to this test
    scene, cam, gt = random_small_scene(3000, 3, 91, width=200, height=150)
    cfg = RenderCfg(sh_degree=3, training=True)
    tr = GsTrainer(hip, scene)
    before = tr.params()
    t = tr.autotune(cam, cfg, torch.tensor(gt, device=hip.dev), reps=2)
    _valid(t)
    assert hip.tuning()["raster_batch"] == t["raster_batch"]
    after = tr.params()
    for a, b in zip(before, after):
        assert np.array_equal(a, b)
    # one step with the tuned knobs: the gradients match the oracle
    tr.step(cam, cfg, torch.tensor(gt, device=hip.dev), train_cfg(lr=1e-3), 0)
    ref = oracle.forward_backward(scene, cam, cfg, gt)
    n = scene.n
    shapes = dict(pos=(n, 3), scale=(n, 3), rot=(n, 4), sh=(n, scene.max_coeffs, 3), op=(n,))
    for k in shapes:
        assert_grad_close(tr.buffer("grad_" + k, shapes[k]).cpu().numpy(), ref["grads"][k], k)
    assert_image_close(tr.buffer("image", (cam.height, cam.width, 3)).cpu().numpy(), ref["img"])


def test_tuning_roundtrip():
    from gs_hip import GsHip
    hip = GsHip(0)
    t = hip.tuning()
    t.update(raster_batch=128, raster_bwd_batch=64, radix_bits=5, bin_group=4096, bin_tiles_per_lane=4,
             persistent_per_cu=3)
    hip.apply_tuning(t)
    got = hip.tuning()
    for k in ["raster_batch", "raster_bwd_batch", "radix_bits", "bin_group", "bin_tiles_per_lane", "persistent_per_cu"]:
        assert got[k] == t[k]


def test_batch_autotune_parity(oracle):
    from gs_hip import GsBatch, GsHip, GsTrainer
    print("[case] autotune batch", flush=True)
    hip = GsHip(0)
    scene, _, _ = random_small_scene(2000, 3, 92)
    rng = np.random.Generator(np.random.Philox(99))
    cams = [make_camera(sphere_eye(rng, 4.0), w, h, 60.0) for (w, h) in [(96, 64), (160, 120), (64, 48)]]
    gts = [target_image(c.width, c.height, 300 + i) for i, c in enumerate(cams)]
    cfg = RenderCfg(sh_degree=3, training=True)
    tr = GsTrainer(hip, scene)
    before = tr.params()
    b = GsBatch(tr, cams, cfg, gts, persistent=True, fused_adam=True, use_graph=True)
    t = b.autotune(reps=2)
    assert t["bin_tiles_per_lane"] in (1, 2, 4) and t["persistent_per_cu"] in (0, 2, 4, 6, 8)
    assert t["raster_batch"] in (128, 256, 512) and t["raster_bwd_batch"] in (128, 256)
    for a, c in zip(before, tr.params()):
        assert np.array_equal(a, c)
    b.render()
    for v, cam in enumerate(cams):
        ref = oracle.forward(scene, cam, cfg)
        assert_image_close(b.images()[v].cpu().numpy(), ref["img"], f"view {v}")
