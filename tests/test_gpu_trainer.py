"""The native training driver (gs_trainer_step): one full step against the
oracle (Trainer::step restated), and the optimizer update against the oracle
Adam applied to the driver's own gradients (bit-exact)."""
import numpy as np
import pytest
import torch

from gs_scenes import RenderCfg, make_workload, random_small_scene
from helpers import assert_grad_close, assert_image_close

pytestmark = pytest.mark.gpu

NAMES = ["pos", "scale", "rot", "sh", "op"]


@pytest.mark.parametrize("use_adam", [True, False])
@pytest.mark.parametrize("n,deg,seed", [(800, 3, 31), (3000, 1, 32)])
def test_trainer_step_matches_oracle(hip, oracle, n, deg, seed, use_adam):
    from gs_hip import GsTrainer, train_cfg
    print("[case] trainer step", n, deg, seed, use_adam, flush=True)
    scene, cam, gt = random_small_scene(n, deg, seed, width=128, height=96)
    cfg = RenderCfg(sh_degree=3, training=True)
    tc = train_cfg(use_adam=use_adam)
    tr = GsTrainer(hip, scene)
    gt_t = torch.tensor(gt, device=hip.dev)
    res = tr.step(cam, cfg, gt_t, tc, 7)
    torch.cuda.synchronize()
    ref = oracle.forward_backward(scene, cam, cfg, gt)
    assert abs(res.loss - ref["loss"]) <= 1e-6
    assert res.pairs == len(ref["keys"])
    assert_image_close(tr.buffer("image", (cam.height, cam.width, 3)).cpu().numpy(), ref["img"])
    shapes = dict(pos=(n, 3), scale=(n, 3), rot=(n, 4), sh=(n, scene.max_coeffs, 3), op=(n,))
    grads = {k: tr.buffer("grad_" + k, shapes[k]).cpu().numpy() for k in NAMES}
    for k in NAMES:
        assert_grad_close(grads[k], ref["grads"][k], "grad_" + k)
    # optimizer: the oracle's update applied to the driver's own gradients
    after = tr.params()
    lr_pos = oracle.lr_schedule(tc.lr_position_init, tc.lr_position_final, 7, tc.max_steps)
    lrs = [lr_pos, tc.lr_scaling, tc.lr_rotation, tc.lr_feature, tc.lr_opacity]
    raw = [a.copy() for a in scene.arrays()]
    if use_adam:
        bc1, bc2 = oracle.bias_corrections(0.9, 0.999, 1)
        for i, k in enumerate(NAMES):
            m = np.zeros_like(raw[i])
            v = np.zeros_like(raw[i])
            oracle.adam_update(raw[i], grads[k].reshape(raw[i].shape), m, v, lrs[i], 0.9, 0.999, 1e-8, bc1, bc2)
    else:
        for i, k in enumerate(NAMES):
            raw[i] = raw[i] - np.float32(lrs[i]) * grads[k].reshape(raw[i].shape)
    for i, k in enumerate(NAMES):
        assert np.array_equal(after[i].reshape(raw[i].shape), raw[i]), k


def test_trainer_loss_decreases_on_c1_workload(hip):
    """test_trainer_gpu.cpp's criterion: the loss does not increase over steps
    (10k Gaussians, 256x256, Adam)."""
    from gs_hip import GsTrainer, train_cfg
    print("[case] trainer c1 convergence", flush=True)
    scene, cams, gts = make_workload("10k_256")
    cfg = RenderCfg(sh_degree=3, training=True)
    tr = GsTrainer(hip, scene)
    gt = torch.tensor(gts[0], device=hip.dev)
    tc = train_cfg(lr=1e-2)
    losses = [tr.step(cams[0], cfg, gt, tc, i).loss for i in range(30)]
    assert losses[-1] < losses[0] * 0.97, losses


def test_trainer_fused_adam_matches_separate_adam(hip):
    """gs_trainer_set_options(fused_adam): one step with Adam inside the
    preprocess backward against the separate kernels.  The screen-space
    gradients are summed with float atomics, so two runs differ in the last
    bits; where a near-zero gradient changes sign, Adam's first step moves the
    parameter by +-lr instead of -+lr; the rest is bit-identical almost
    everywhere."""
    from gs_hip import GsTrainer, train_cfg
    print("[case] trainer fused adam", flush=True)
    scene, cam, gt = random_small_scene(2500, 3, 33, width=140, height=100)
    cfg = RenderCfg(sh_degree=3, training=True)
    gt_t = torch.tensor(gt, device=hip.dev)
    lr = 1e-3
    out = []
    for fused in (False, True):
        tr = GsTrainer(hip, scene)
        tr.set_options(fused_adam=fused)
        loss = tr.step(cam, cfg, gt_t, train_cfg(lr=lr), 0).loss
        out.append((tr.params(), loss))
    assert out[0][1] == out[1][1]  # same forward
    total = same = 0
    for a, b in zip(out[0][0], out[1][0]):
        assert np.abs(a - b).max() <= 2 * lr * 1.0001
        same += int((a.view(np.uint32) == b.view(np.uint32)).sum())
        total += a.size
    assert same >= 0.99 * total, (same, total)


def test_trainer_step_redoes_when_pairs_outgrow_buffer(hip, oracle):
    """The pair buffer is sized by the first step (far camera); the second
    step's view is much closer, so its binning overflows the buffer while the
    host is not waiting for the count.  The step must detect it, grow the
    buffer, redo the forward and backward, and give the oracle's result for
    the parameters after the first step."""
    from gs_hip import GsTrainer, train_cfg
    from gs_scenes import RawScene, make_camera
    print("[case] trainer pair-buffer overflow redo", flush=True)
    scene, _, gt = random_small_scene(1500, 3, 34, width=128, height=96)
    cfg = RenderCfg(sh_degree=3, training=True)
    far = make_camera((0.0, 0.0, -9.0), 128, 96, 60.0)
    near = make_camera((0.0, 0.0, -2.2), 128, 96, 60.0)
    tc = train_cfg(use_adam=False)
    tr = GsTrainer(hip, scene)
    gt_t = torch.tensor(gt, device=hip.dev)
    r1 = tr.step(far, cfg, gt_t, tc, 0)
    p1 = tr.params()
    r2 = tr.step(near, cfg, gt_t, tc, 1)
    torch.cuda.synchronize()
    assert r2.pairs > (r1.pairs * 9) // 8 + 1024, (r1.pairs, r2.pairs)  # the buffer did overflow
    mid = RawScene(p1[0], p1[1], p1[2], p1[3].reshape(scene.sh.shape), p1[4], scene.sh_degree)
    ref = oracle.forward_backward(mid, near, cfg, gt)
    assert abs(r2.loss - ref["loss"]) <= 1e-6
    assert r2.pairs == len(ref["keys"])
    assert_image_close(tr.buffer("image", (96, 128, 3)).cpu().numpy(), ref["img"])
    n = scene.n
    shapes = dict(pos=(n, 3), scale=(n, 3), rot=(n, 4), sh=(n, scene.max_coeffs, 3), op=(n,))
    for k in NAMES:
        assert_grad_close(tr.buffer("grad_" + k, shapes[k]).cpu().numpy(), ref["grads"][k], "grad_" + k)


def test_trainer_stage_timing(hip):
    """gs_trainer_set_timing / gs_trainer_stage_times: every timed stage gets a
    positive mean, untimed ones read 0, and the stages fit inside the step."""
    import time
    from gs_hip import GsTrainer, train_cfg
    print("[case] trainer stage timing", flush=True)
    scene, cam, gt = random_small_scene(3000, 3, 35, width=256, height=192)
    cfg = RenderCfg(sh_degree=3, training=True)
    tr = GsTrainer(hip, scene)
    tr.set_options(fused_adam=True)
    gt_t = torch.tensor(gt, device=hip.dev)
    tc = train_cfg()
    tr.step(cam, cfg, gt_t, tc, 0)
    tr.set_timing(None)
    t0 = time.perf_counter()
    for i in range(4):
        tr.step(cam, cfg, gt_t, tc, 1 + i)
    wall_ms = (time.perf_counter() - t0) * 1e3 / 4
    ms, steps = tr.stage_times()
    assert steps == 4
    assert all(v > 0 for v in ms.values()), ms
    assert sum(ms.values()) < wall_ms, (ms, wall_ms)
    tr.set_timing(["raster_bwd"])
    tr.step(cam, cfg, gt_t, tc, 5)
    ms, steps = tr.stage_times()
    assert steps == 1 and ms["raster_bwd"] > 0
    assert all(v == 0 for k, v in ms.items() if k != "raster_bwd"), ms
    tr.set_timing([])


def _render_state(tr, cam):
    H, W = cam.height, cam.width
    return [tr.buffer("image", (H, W, 3)).cpu().numpy(),
            tr.buffer("T_final", (H, W)).cpu().numpy(),
            tr.buffer("n_contrib", (H, W), torch.int32).cpu().numpy(),
            tr.buffer("last_contrib", (H, W), torch.int32).cpu().numpy()]


@pytest.mark.parametrize("split", [1, 3, 20])
def test_trainer_two_round_forward_equals_one_round(hip, oracle, split):
    """The trainer's two-round forward (the nearest `split` depth groups placed
    and rendered first, the rest placed only for bins with unfinished tiles)
    gives the one-round forward's image and pixel state bit for bit, and the
    step then matches the oracle.  A dense scene (most tiles finish in round
    1) with small groups (256 Gaussians, 80 groups)."""
    from gs_hip import GsTrainer, train_cfg
    print("[case] two-round forward", split, flush=True)
    scene, cam, gt = random_small_scene(20000, 3, 41, width=160, height=128, logscale=(-2.0, 0.4), op=(0.6, 1.0))
    cfg = RenderCfg(sh_degree=3, training=True)
    gt_t = torch.tensor(gt, device=hip.dev)
    hip.set_binning(256, 1)
    try:
        states = []
        for sp in (0, split):
            hip.set_split(sp)
            tr = GsTrainer(hip, scene)
            K = tr.render(cam, cfg)
            K = tr.render(cam, cfg)  # the second render uses the asynchronous pair count
            torch.cuda.synchronize()
            states.append((K, _render_state(tr, cam)))
        assert states[0][0] == states[1][0]
        for a, b in zip(states[0][1], states[1][1]):
            assert np.array_equal(a, b)
        # an SGD step with the split (the second one: its pair count is
        # asynchronous): the oracle's loss, image and gradients
        hip.set_split(split)
        tr2 = GsTrainer(hip, scene)
        tr2.step(cam, cfg, gt_t, train_cfg(use_adam=False), 0)
        torch.cuda.synchronize()
        p1 = tr2.params()
        res = tr2.step(cam, cfg, gt_t, train_cfg(use_adam=False), 0)
        torch.cuda.synchronize()
    finally:
        hip.set_binning(4096, 1)
        hip.set_split(64)  # the context default
    n = scene.n
    from gs_scenes import RawScene
    mid = RawScene(p1[0], p1[1], p1[2], p1[3].reshape(scene.sh.shape), p1[4], scene.sh_degree)
    ref = oracle.forward_backward(mid, cam, cfg, gt)
    assert abs(res.loss - ref["loss"]) <= 1e-6
    assert res.pairs == len(ref["keys"])
    assert_image_close(tr2.buffer("image", (cam.height, cam.width, 3)).cpu().numpy(), ref["img"])
    shapes = dict(pos=(n, 3), scale=(n, 3), rot=(n, 4), sh=(n, scene.max_coeffs, 3), op=(n,))
    for k in NAMES:
        assert_grad_close(tr2.buffer("grad_" + k, shapes[k]).cpu().numpy(), ref["grads"][k], "grad_" + k)
