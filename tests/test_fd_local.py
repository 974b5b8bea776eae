"""The tile-local FP64 finite differences (oracle.fd_grads_local, gsd_fd_local)
that check gradients on scenes too large for whole-image differences.

Pinned two ways on the CPU:
- against whole-image central differences of the FP64 forward (the same
  function evaluated everywhere): equal to 1e-8 of the largest derivative, so
  the subset, the tile region and the insertion of the perturbed Gaussian into
  each tile's depth order lose nothing;
- against the float oracle's analytic backward on a 60k-Gaussian scene, with
  a planted sign mistake and a planted swap of two parameters that the same
  comparison must reject."""
import numpy as np
import pytest

from gs_scenes import RenderCfg, gaussians_uniform_cube, make_camera, random_small_scene, target_image
from test_oracle import _fd_case

GROUPS = lambda K: {"pos": (0, 3), "scale": (3, 3), "rot": (6, 4), "sh": (10, 3 * K), "op": (10 + 3 * K, 1)}


def _analytic(grads, idx):
    n = len(idx)
    return np.concatenate([grads["pos"][idx], grads["scale"][idx], grads["rot"][idx],
                           grads["sh"][idx].reshape(n, -1), grads["op"][idx][:, None]], 1).astype(np.float64)


@pytest.mark.parametrize("deg", [0, 3])
def test_fd_local_equals_whole_image_differences(oracle64, deg):
    from oracle.oracle import fd_grads_local
    scene, cam, gt = random_small_scene(12, deg, 40 + deg, width=48, height=40, logscale=(-1.6, 0.2),
                                        op=(0.0, 0.5), spread=0.6)
    cfg = RenderCfg(sh_degree=3, training=True)
    G, steps = fd_grads_local(scene, cam, cfg, gt, list(range(12)), fallback=())
    assert (steps == 1e-4).all()
    worst, top = 0.0, 0.0
    for t in range(12):
        for nm, (a, l) in GROUPS(scene.max_coeffs).items():
            for k in range(l):
                fd = _fd_case(oracle64, scene, cam, cfg, gt, nm, t * l + k)
                worst = max(worst, abs(fd - G[t, a + k]))
                top = max(top, abs(fd))
    assert top > 1e-4
    assert worst <= 1e-8 * top, (worst, top)


def test_fd_local_against_oracle_backward(oracle):
    """60k Gaussians, 480x270: 24 Gaussians that reach a pixel, every raw
    parameter.  The float oracle's analytic gradients agree within 1e-4 of
    each group's largest derivative (measured: ~1e-5); a flipped sign or two
    swapped parameters fail the same bar."""
    from oracle.oracle import fd_grads_local
    scene = gaussians_uniform_cube(60000, 2, 3)
    cam = make_camera((0.0, 0.0, -3.0), 480, 270, 60.0)
    gt = target_image(480, 270, 2002)
    cfg = RenderCfg(sh_degree=3, training=True)
    f = oracle.forward_backward(scene, cam, cfg, gt)
    live = np.nonzero(np.abs(f["grads"]["op"]) > 0)[0]
    idx = np.sort(np.random.default_rng(5).choice(live, 24, replace=False))
    G, steps = fd_grads_local(scene, cam, cfg, gt, idx, threads=8)
    ok = steps > 0
    assert ok.mean() >= 0.9, ok.mean()
    A = _analytic(f["grads"], idx)

    def err(a):
        out = {}
        for nm, (s, l) in GROUPS(scene.max_coeffs).items():
            m = ok[:, s:s + l]
            b = G[:, s:s + l]
            out[nm] = np.abs(a[:, s:s + l] - b)[m].max() / np.abs(b[m]).max()
        return out

    e = err(A)
    assert max(e.values()) <= 1e-4, e
    # planted mistakes: a dropped sign on the position gradient, swapped scale axes
    bad = A.copy()
    bad[:, 0:3] *= -1
    assert err(bad)["pos"] > 0.5
    bad = A.copy()
    bad[:, [3, 4]] = bad[:, [4, 3]]
    assert err(bad)["scale"] > 1e-2
