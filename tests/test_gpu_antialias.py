"""Antialiasing on the device (RenderConfig::antialiasing,
preprocessor_cpu.cpp:213-216,250) against the oracle: the forward (opacity
scaled by h = sqrt(max(2.5e-5, det_cov / det)), image within 1e-4) and the
backward with the adjoint of h, which the oracle carries and which is pinned
against FP64 central differences on the CPU (tests/test_antialias.py):
the C-ABI stages, the training driver's step and the batched step's summed
gradients, all within 1e-3 relative."""
import numpy as np
import pytest
import torch

from gs_scenes import RenderCfg, make_camera, random_small_scene, sphere_eye, target_image
from helpers import assert_grad_close, assert_image_close

pytestmark = pytest.mark.gpu

NAMES = ["pos", "scale", "rot", "sh", "op"]
AA = RenderCfg(sh_degree=3, training=True, antialiasing=True)


@pytest.mark.parametrize("n,deg,seed", [(700, 3, 201), (2500, 1, 202)])
def test_antialiased_forward_backward_matches_oracle(hip, oracle, n, deg, seed):
    print("[case] antialiased stages", n, deg, seed, flush=True)
    scene, cam, gt = random_small_scene(n, deg, seed, width=120, height=90, logscale=(-2.6, 0.5))
    f = hip.forward_backward_fast(hip.upload_scene(scene), cam, AA, torch.tensor(gt, device=hip.dev))
    torch.cuda.synchronize()
    ref = oracle.forward_backward(scene, cam, AA, gt)
    np.testing.assert_array_equal(f["pre"]["radii"].cpu().numpy(), ref["pre"]["radii"])
    assert np.abs(f["pre"]["op2d"].cpu().numpy() - ref["pre"]["op2d"]).max() <= 1e-6
    assert_image_close(f["img"].cpu().numpy(), ref["img"])
    for k in NAMES:
        assert_grad_close(f["grads"][k].cpu().numpy(), ref["grads"][k], "grad_" + k)


def test_antialiased_trainer_step_matches_oracle(hip, oracle):
    from gs_hip import GsTrainer, train_cfg
    print("[case] antialiased trainer step", flush=True)
    scene, cam, gt = random_small_scene(1200, 3, 203, width=128, height=96)
    tr = GsTrainer(hip, scene)
    res = tr.step(cam, AA, torch.tensor(gt, device=hip.dev), train_cfg(), 3)
    torch.cuda.synchronize()
    ref = oracle.forward_backward(scene, cam, AA, gt)
    assert abs(res.loss - ref["loss"]) <= 1e-6
    n = scene.n
    shapes = dict(pos=(n, 3), scale=(n, 3), rot=(n, 4), sh=(n, scene.max_coeffs, 3), op=(n,))
    for k in NAMES:
        assert_grad_close(tr.buffer("grad_" + k, shapes[k]).cpu().numpy(), ref["grads"][k], "grad_" + k)


def test_antialiased_batched_gradients_are_sum_of_single_views(hip, oracle):
    from gs_hip import GsBatch, GsTrainer, train_cfg
    print("[case] antialiased batch", flush=True)
    scene, _, _ = random_small_scene(1500, 3, 204)
    rng = np.random.Generator(np.random.Philox(211))
    cams = [make_camera(sphere_eye(rng, 4.0), w, h, 60.0) for (w, h) in [(96, 64), (64, 48), (130, 90)]]
    gts = [target_image(c.width, c.height, 300 + i) for i, c in enumerate(cams)]
    tr = GsTrainer(hip, scene)
    b = GsBatch(tr, cams, AA, gts, persistent=True, fused_adam=False, use_graph=False)
    b.step(train_cfg(lr=1e-3), 0, sync=True)
    n = scene.n
    shapes = dict(pos=(n, 3), scale=(n, 3), rot=(n, 4), sh=(n, scene.max_coeffs, 3), op=(n,))
    got = {k: tr.buffer("grad_" + k, shapes[k]).cpu().numpy() for k in NAMES}
    ref = {k: np.zeros(shapes[k], np.float32) for k in NAMES}
    for cam, gt in zip(cams, gts):
        f = oracle.forward_backward(scene, cam, AA, gt)
        for k in NAMES:
            ref[k] += f["grads"][k]
    for k in NAMES:
        assert_grad_close(got[k], ref[k], "grad_" + k)
