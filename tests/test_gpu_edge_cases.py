"""Degenerate inputs through the native training step (gs_trainer_step) and
the batched step (gs_batch_step), against the oracle: an empty scene and a
scene wholly behind the camera (no pairs: the image is the background,
every gradient is zero, and Adam leaves every parameter as it was), a single
Gaussian, views of 1 x 1 and 17 x 3 pixels (one partial tile), and a batch
in which one view sees no Gaussian."""
import numpy as np
import pytest
import torch

from gs_scenes import RawScene, RenderCfg, make_camera, random_small_scene, sphere_eye, target_image
from helpers import assert_grad_close, assert_image_close

pytestmark = pytest.mark.gpu

NAMES = ["pos", "scale", "rot", "sh", "op"]
BG = (0.2, 0.4, 0.6)


def _shapes(scene):
    n = scene.n
    return dict(pos=(n, 3), scale=(n, 3), rot=(n, 4), sh=(n, scene.max_coeffs, 3), op=(n,))


def _step_against_oracle(hip, oracle, scene, cam, gt, cfg, steps=2):
    """Runs `steps` Adam steps; the first one's loss, pair count, image and
    gradients against the oracle; returns the trainer and the results."""
    from gs_hip import GsTrainer, train_cfg
    tr = GsTrainer(hip, scene)
    gt_t = torch.tensor(gt, device=hip.dev)
    res = tr.step(cam, cfg, gt_t, train_cfg(lr=1e-3), 0)
    torch.cuda.synchronize()
    ref = oracle.forward_backward(scene, cam, cfg, gt)
    assert abs(res.loss - ref["loss"]) <= 1e-6, (res.loss, ref["loss"])
    assert res.pairs == len(ref["keys"])
    img = tr.buffer("image", (cam.height, cam.width, 3)).cpu().numpy()
    assert_image_close(img, ref["img"])
    sh = _shapes(scene)
    grads = {k: tr.buffer("grad_" + k, sh[k]).cpu().numpy() for k in NAMES}
    if scene.n:
        for k in NAMES:
            assert_grad_close(grads[k], ref["grads"][k], "grad_" + k)
    for i in range(1, steps):
        tr.step(cam, cfg, gt_t, train_cfg(lr=1e-3), i)
    torch.cuda.synchronize()
    return tr, res, ref, img, grads


def _behind(scene):
    """The scene moved behind the camera at (0, 0, -4), which looks along +z."""
    s = scene.copy()
    s.pos[:, 2] = np.float32(-8.0) + np.float32(0.1) * s.pos[:, 2]
    return s


@pytest.mark.parametrize("n", [0, 300])
def test_step_without_pairs_leaves_parameters_unchanged(hip, oracle, n):
    print("[case] no pairs", n, flush=True)
    scene, cam, gt = random_small_scene(max(n, 1), 3, 901, width=40, height=30)
    if n == 0:
        scene = RawScene(scene.pos[:0], scene.log_scale[:0], scene.rot[:0], scene.sh[:0], scene.opacity_logit[:0], 3)
    else:
        scene = _behind(scene)
    cfg = RenderCfg(sh_degree=3, training=True, bg_color=BG)
    tr, res, ref, img, grads = _step_against_oracle(hip, oracle, scene, cam, gt, cfg, steps=3)
    assert res.pairs == 0
    assert np.array_equal(img, np.broadcast_to(np.array(BG, np.float32), img.shape))
    for k in NAMES:
        assert not np.any(grads[k]), k
    # zero gradients: Adam's moments stay zero and its update is 0 / (0 + eps)
    for a, b in zip(tr.params(), scene.arrays()):
        assert np.array_equal(a.reshape(b.shape), b)


def test_step_with_one_gaussian(hip, oracle):
    print("[case] one Gaussian", flush=True)
    scene, cam, gt = random_small_scene(1, 3, 902, width=48, height=40)
    scene.pos[:] = 0.0
    # anisotropic: an isotropic Gaussian's rotation gradient is zero up to rounding
    scene.log_scale[:] = np.array([-1.0, -1.6, -0.7], np.float32)
    cfg = RenderCfg(sh_degree=3, training=True, bg_color=BG)
    tr, res, ref, _, _ = _step_against_oracle(hip, oracle, scene, cam, gt, cfg)
    assert res.pairs > 0


@pytest.mark.parametrize("w,h", [(1, 1), (3, 3), (17, 3)])
def test_step_on_views_smaller_than_a_tile(hip, oracle, w, h):
    """At 1 x 1 and 2 x 2 every radius (at least 2 px here) exceeds the view's
    larger side, which the reference culls (preprocessor_cpu.cpp:228): no
    pairs, the background; from 3 x 3 on the Gaussians are drawn."""
    print("[case] small view", w, h, flush=True)
    scene, _, _ = random_small_scene(400, 3, 903, spread=0.3, logscale=(-1.0, 0.3))
    cam = make_camera((0.0, 0.0, -4.0), w, h, 60.0)
    gt = target_image(w, h, 904)
    cfg = RenderCfg(sh_degree=3, training=True, bg_color=BG)
    _, res, _, img, _ = _step_against_oracle(hip, oracle, scene, cam, gt, cfg)
    if w == 1:
        assert res.pairs == 0 and np.array_equal(img, np.broadcast_to(np.array(BG, np.float32), img.shape))
    else:
        assert res.pairs > 0


def test_batch_with_a_view_that_sees_nothing(hip, oracle):
    from gs_hip import GsBatch, GsTrainer, train_cfg
    print("[case] batch with an empty view", flush=True)
    scene, _, _ = random_small_scene(1200, 3, 905)
    rng = np.random.Generator(np.random.Philox(906))
    cams = [make_camera(sphere_eye(rng, 4.0), 96, 64, 60.0),
            # looking away from the scene (which lies around the origin)
            make_camera((0.0, 0.0, -4.0), 80, 48, 60.0, center=(0.0, 0.0, -10.0)),
            make_camera(sphere_eye(rng, 4.0), 64, 48, 60.0)]
    gts = [target_image(c.width, c.height, 907 + i) for i, c in enumerate(cams)]
    cfg = RenderCfg(sh_degree=3, training=True, bg_color=BG)
    refs = [oracle.forward_backward(scene, c, cfg, g) for c, g in zip(cams, gts)]
    assert len(refs[1]["keys"]) == 0 and len(refs[0]["keys"]) > 0
    for use_graph in (False, True):
        tr = GsTrainer(hip, scene)
        b = GsBatch(tr, cams, cfg, gts, persistent=True, fused_adam=False, use_graph=use_graph)
        b.step(train_cfg(lr=1e-3), 0, sync=True)
        _, losses = b.result()
        for v in range(len(cams)):
            assert abs(losses[v] - refs[v]["loss"]) <= 1e-6, (v, losses[v], refs[v]["loss"])
            assert_image_close(b.images()[v].cpu().numpy(), refs[v]["img"], f"view {v}")
        sh = _shapes(scene)
        for k in NAMES:
            want = sum(r["grads"][k] for r in refs)
            assert_grad_close(tr.buffer("grad_" + k, sh[k]).cpu().numpy(), want, "grad_" + k)
        del b, tr


def test_batch_on_an_empty_scene(hip, oracle):
    """No Gaussians at all in a batched, graph-captured step: every view is the
    background and its loss the oracle's."""
    from gs_hip import GsBatch, GsTrainer, train_cfg
    print("[case] batch, empty scene", flush=True)
    s, _, _ = random_small_scene(1, 3, 908)
    scene = RawScene(s.pos[:0], s.log_scale[:0], s.rot[:0], s.sh[:0], s.opacity_logit[:0], 3)
    rng = np.random.Generator(np.random.Philox(909))
    cams = [make_camera(sphere_eye(rng, 4.0), w, h, 60.0) for (w, h) in [(40, 30), (17, 9)]]
    gts = [target_image(c.width, c.height, 910 + i) for i, c in enumerate(cams)]
    cfg = RenderCfg(sh_degree=3, training=True, bg_color=BG)
    tr = GsTrainer(hip, scene)
    b = GsBatch(tr, cams, cfg, gts, persistent=True, fused_adam=True, use_graph=True)
    for it in range(2):
        b.step(train_cfg(lr=1e-3), it, sync=True)
    _, losses = b.result()
    for v, (c, g) in enumerate(zip(cams, gts)):
        ref = oracle.forward_backward(scene, c, cfg, g)
        assert abs(losses[v] - ref["loss"]) <= 1e-6, (v, losses[v], ref["loss"])
        img = b.images()[v].cpu().numpy()
        assert np.array_equal(img, np.broadcast_to(np.array(BG, np.float32), img.shape)), v


@pytest.mark.parametrize("fast", [False, True])
def test_eval3d_on_an_empty_scene(hip, fast):
    """The eval_3D forward with no Gaussians: no pairs, the background."""
    print("[case] eval3d empty scene", fast, flush=True)
    s, cam, _ = random_small_scene(1, 3, 911, width=40, height=24)
    scene = RawScene(s.pos[:0], s.log_scale[:0], s.rot[:0], s.sh[:0], s.opacity_logit[:0], 3)
    cfg = RenderCfg(sh_degree=3, training=False, bg_color=BG)
    out = hip.render_3d(hip.upload_scene(scene), cam, cfg, raw=True, fast=fast)
    torch.cuda.synchronize()
    assert out["K"] == 0
    img = out["img"].cpu().numpy()
    assert np.array_equal(img, np.broadcast_to(np.array(BG, np.float32), img.shape))
