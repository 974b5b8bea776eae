"""The batched multi-view step (gs_batch_*): per-view images against the
oracle's single-view renders, summed gradients against the sum of the
oracle's per-view gradients, fused Adam against the separate Adam kernel,
persistent raster against one group per tile (bit-exact), and hipGraph
replays against eager runs of the same step."""
import numpy as np
import pytest
import torch

from gs_scenes import RenderCfg, make_camera, random_small_scene, sphere_eye, target_image
from helpers import assert_grad_close, assert_image_close

pytestmark = pytest.mark.gpu

NAMES = ["pos", "scale", "rot", "sh", "op"]
SIZES = [(96, 64), (64, 48), (130, 90)]


def _setup(n=1500, deg=3, seed=51):
    scene, _, _ = random_small_scene(n, deg, seed)
    rng = np.random.Generator(np.random.Philox(seed + 7))
    cams = [make_camera(sphere_eye(rng, 4.0), w, h, 60.0) for (w, h) in SIZES]
    gts = [target_image(c.width, c.height, seed + 100 + i) for i, c in enumerate(cams)]
    return scene, cams, gts


def _shapes(scene):
    n = scene.n
    return dict(pos=(n, 3), scale=(n, 3), rot=(n, 4), sh=(n, scene.max_coeffs, 3), op=(n,))


def test_batched_render_matches_oracle(hip, oracle):
    from gs_hip import GsBatch, GsTrainer
    print("[case] batch render", flush=True)
    scene, cams, gts = _setup()
    cfg = RenderCfg(sh_degree=3, training=True)
    tr = GsTrainer(hip, scene)
    imgs = {}
    for persistent in (True, False):
        b = GsBatch(tr, cams, cfg, gts, persistent=persistent, fused_adam=True, use_graph=False)
        b.render()
        imgs[persistent] = [im.cpu().numpy() for im in b.images()]
    for v, cam in enumerate(cams):
        ref = oracle.forward(scene, cam, cfg)
        assert_image_close(imgs[True][v], ref["img"], f"view {v}")
        assert np.array_equal(imgs[True][v], imgs[False][v])  # forward has no atomics: bit-exact


def test_batched_gradients_are_sum_of_single_views(hip, oracle):
    from gs_hip import GsBatch, GsTrainer, train_cfg
    print("[case] batch gradients", flush=True)
    scene, cams, gts = _setup()
    cfg = RenderCfg(sh_degree=3, training=True)
    tr = GsTrainer(hip, scene)
    b = GsBatch(tr, cams, cfg, gts, persistent=True, fused_adam=False, use_graph=False)
    res = b.step(train_cfg(lr=1e-3), 0, sync=True)
    _, losses = b.result()
    shapes = _shapes(scene)
    grads = {k: tr.buffer("grad_" + k, shapes[k]).cpu().numpy() for k in NAMES}
    ref_sum = {k: np.zeros(shapes[k], np.float32) for k in NAMES}
    ref_losses = []
    for cam, gt in zip(cams, gts):
        f = oracle.forward_backward(scene, cam, cfg, gt)
        ref_losses.append(f["loss"])
        for k in NAMES:
            ref_sum[k] += f["grads"][k]
    assert np.allclose(losses, ref_losses, rtol=0, atol=1e-6), (losses, ref_losses)
    assert abs(res.loss - np.mean(ref_losses)) <= 1e-6
    for k in NAMES:
        assert_grad_close(grads[k], ref_sum[k], "grad_" + k)


def _run(hip, scene, cams, gts, cfg, steps, **opts):
    from gs_hip import GsBatch, GsTrainer, train_cfg
    tr = GsTrainer(hip, scene)
    b = GsBatch(tr, cams, cfg, gts, **opts)
    tc = train_cfg(lr=1e-3)
    out = []
    for s in range(steps):
        r = b.step(tc, s, sync=True)
        torch.cuda.synchronize()  # eager work between replays
        out.append((r.loss, [p.copy() for p in tr.params()], [im.cpu().numpy() for im in b.images()]))
    return out


def test_fused_adam_matches_separate_adam(hip):
    print("[case] fused adam", flush=True)
    scene, cams, gts = _setup()
    cfg = RenderCfg(sh_degree=3, training=True)
    a = _run(hip, scene, cams, gts, cfg, 3, persistent=True, fused_adam=True, use_graph=False)
    b = _run(hip, scene, cams, gts, cfg, 3, persistent=True, fused_adam=False, use_graph=False)
    for (la, pa, ia), (lb, pb, ib) in zip(a, b):
        assert abs(la - lb) <= 1e-5 * abs(lb) + 1e-7
        # gradient sums differ only by atomic ordering; an Adam step moves a
        # parameter by about lr, so a sign flip of a near-zero gradient costs
        # at most 2 lr per step
        for x, y in zip(pa, pb):
            assert np.abs(x - y).max() <= 6e-3, np.abs(x - y).max()
        for x, y in zip(ia, ib):
            assert np.abs(x - y).max() <= 5e-3


def test_graph_replays_match_eager(hip):
    print("[case] graph replay", flush=True)
    scene, cams, gts = _setup()
    cfg = RenderCfg(sh_degree=3, training=True)
    g = _run(hip, scene, cams, gts, cfg, 4, persistent=True, fused_adam=True, use_graph=True)
    e = _run(hip, scene, cams, gts, cfg, 4, persistent=True, fused_adam=True, use_graph=False)
    for step, ((lg, pg, ig), (le, pe, ie)) in enumerate(zip(g, e)):
        assert abs(lg - le) <= 1e-5 * abs(le) + 1e-7, (step, lg, le)
        for x, y in zip(pg, pe):
            assert np.abs(x - y).max() <= 8e-3
        for x, y in zip(ig, ie):
            assert np.abs(x - y).max() <= 5e-3


def test_batched_loss_decreases(hip):
    from gs_hip import GsBatch, GsTrainer, train_cfg
    print("[case] batch convergence", flush=True)
    scene, cams, gts = _setup(n=3000, seed=61)
    cfg = RenderCfg(sh_degree=3, training=True)
    tr = GsTrainer(hip, scene)
    b = GsBatch(tr, cams, cfg, gts)
    tc = train_cfg(lr=1e-2)
    first = b.step(tc, 0, sync=True).loss
    for s in range(1, 25):
        b.step(tc, s)
    last, _ = b.result()
    assert last.loss < first * 0.97, (first, last.loss)


@pytest.mark.parametrize("fused", [True, False])
def test_overflowing_step_changes_nothing_and_redo_is_one_clean_step(hip, fused):
    """A batched (graph-captured) step whose pairs outgrow the pre-sized
    buffers: the step reports the overflow, leaves parameters and Adam
    moments untouched, hands its Adam step number back, and regrows; the redo
    then equals one clean step from the same parameters (a fresh trainer and
    batch sized for them).  Exactly equal almost everywhere: the gradient
    sums use float atomics, so a near-zero gradient can change sign, moving
    that parameter by -+lr instead of +-lr."""
    from gs_hip import GsBatch, GsError, GsTrainer, train_cfg
    print("[case] batch overflow redo, fused", fused, flush=True)
    scene, cams, gts = _setup()
    cfg = RenderCfg(sh_degree=3, training=True)
    big = scene.copy()
    big.log_scale = big.log_scale + np.float32(2.0)  # ~50x the pairs
    tc = train_cfg(lr=1e-3)
    tr = GsTrainer(hip, scene)
    b = GsBatch(tr, cams, cfg, gts, persistent=True, fused_adam=fused, use_graph=True)
    b.render()  # sizes the pair buffers for the small Gaussians
    _, _, cap = b.info()
    tr.set_params(big)
    with pytest.raises(GsError, match="not applied"):
        b.step(tc, 0, sync=True)
    for x, y in zip(tr.params(), big.arrays()):
        assert np.array_equal(x, y.reshape(x.shape))
    for k, shp in _shapes(scene).items():
        assert not tr.buffer("adam_m_" + k, shp).any() and not tr.buffer("adam_v_" + k, shp).any()
    assert b.info()[2] > cap
    res = b.step(tc, 0, sync=True)  # the redo
    redo = tr.params()
    tr2 = GsTrainer(hip, big)
    b2 = GsBatch(tr2, cams, cfg, gts, persistent=True, fused_adam=fused, use_graph=True)
    res2 = b2.step(tc, 0, sync=True)
    assert res.pairs == res2.pairs > cap
    assert abs(res.loss - res2.loss) <= 1e-6
    for x, y in zip(redo, tr2.params()):
        assert np.abs(x - y).max() <= 2.5e-3
        assert (x == y).mean() >= 0.99, (x == y).mean()


def test_batched_word_sort_path_with_views_wider_than_255_tiles(hip):
    """A view 4160 px wide (260 tiles) takes the whole batch off the direct
    placement onto the word sort: (tile | index) words sorted by the radix
    sort over the pre-sized capacity with the live pair count on the device.
    With ~28M pairs that sort runs the one-sweep passes.  Graph-captured and
    eager batches render every view as the single-view path does, bit for bit,
    and one step's per-view losses equal the single-view losses."""
    from gs_hip import GsBatch, GsTrainer, train_cfg
    from gs_scenes import gaussians_normal
    print("[case] batch word-sort path 4160x2400 + 1280x720", flush=True)
    scene = gaussians_normal(300_000, 91, 1, sigma=1.5, logscale=(-4.0, 0.6), op=(-1.0, 1.5))
    rng = np.random.Generator(np.random.Philox(92))
    cams = [make_camera(sphere_eye(rng, 4.0), 4160, 2400, 60.0), make_camera(sphere_eye(rng, 4.0), 1280, 720, 50.0)]
    gts = [target_image(c.width, c.height, 93 + i) for i, c in enumerate(cams)]
    cfg = RenderCfg(sh_degree=1, training=True)
    tr = GsTrainer(hip, scene)
    singles, single_losses = [], []
    for cam, gt in zip(cams, gts):
        tr.render(cam, cfg)
        singles.append(tr.buffer("image", (cam.height, cam.width, 3)))
        # a fresh trainer per view: its step's loss is the loss at the scene's parameters
        # (a zero learning rate is no no-op: the position schedule's log-interpolation of
        # 0 and 0 is NaN at step 0, in the reference as here)
        single_losses.append(GsTrainer(hip, scene).step(cam, cfg, torch.tensor(gt, device=hip.dev),
                                                        train_cfg(lr=1e-3), 0).loss)
    for graph in (True, False):
        b = GsBatch(GsTrainer(hip, scene), cams, cfg, gts, persistent=True, fused_adam=False, use_graph=graph)
        K = b.render()
        assert K > 4_000_000, K
        for v, im in enumerate(b.images()):
            assert torch.equal(im, singles[v]), (graph, v)
        b.step(train_cfg(lr=1e-3), 0, sync=True)
        _, losses = b.result()
        assert np.allclose(losses, single_losses, rtol=0, atol=1e-6), (losses, single_losses)
        del b
