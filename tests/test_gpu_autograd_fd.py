"""Test-harness autograd wrapper over the C ABI, checked against FP64 central
finite differences (h = 1e-4) of the double-precision oracle forward.

The wrapper lives here, in tests/, on purpose: it is a checking device for the
hand-written adjoint kernels, not part of the product.
"""
import sys

import numpy as np
import pytest
import torch

from gs_scenes import RawScene, RenderCfg, random_small_scene

pytestmark = pytest.mark.gpu


class RenderFn(torch.autograd.Function):
    """image = render(raw params) through gs_preprocess_fwd .. gs_raster_fwd;
    backward through gs_raster_bwd and gs_preprocess_bwd."""

    @staticmethod
    def forward(ctx, pos, log_scale, rot, sh, op, hip, meta, cam, cfg):
        p = dict(pos=pos.contiguous(), scale=log_scale.contiguous(), rot=rot.contiguous(), sh=sh.contiguous(),
                 op=op.contiguous(), **meta)
        f = hip.forward(p, cam, cfg)
        ctx.hip, ctx.p, ctx.cam, ctx.cfg, ctx.f = hip, p, cam, cfg, f
        return f["img"]

    @staticmethod
    def backward(ctx, d_img):
        hip, p, cam, cfg, f = ctx.hip, ctx.p, ctx.cam, ctx.cfg, ctx.f
        rg = hip.raster_bwd(f["pre"], f["values_sorted"], f["ranges"], cam.width, cam.height, f["px"],
                            d_img.contiguous(), cfg.bg_color)
        g = hip.preprocess_bwd(p, cam, cfg, f["pre"]["radii"], rg)
        return g["pos"], g["scale"], g["rot"], g["sh"], g["op"], None, None, None, None


def _fd(oracle64, scene, cam, cfg, gt, param, idx, h=1e-4):
    attr = {"pos": "pos", "scale": "log_scale", "rot": "rot", "sh": "sh", "op": "opacity_logit"}[param]

    def loss(delta):
        s64 = RawScene(*(a.astype(np.float64) for a in scene.arrays()), scene.sh_degree)
        a = getattr(s64, attr).reshape(-1).copy()
        a[idx] += delta
        setattr(s64, attr, a.reshape(getattr(s64, attr).shape))
        return oracle64.loss(s64, cam, cfg, gt)

    return (loss(h) - loss(-h)) / (2 * h)


@pytest.mark.parametrize("deg,seed", [(0, 71), (1, 72), (3, 73)])
def test_autograd_gradients_match_fp64_finite_differences(hip, oracle64, deg, seed):
    print("[case] autograd fd", deg, seed, flush=True)
    scene, cam, gt = random_small_scene(12, deg, seed, width=48, height=40, logscale=(-1.6, 0.2), op=(0.0, 0.5),
                                        spread=0.6)
    cfg = RenderCfg(sh_degree=3, training=True)
    d = hip.dev
    leaves = [torch.tensor(a, device=d, requires_grad=True) for a in scene.arrays()]
    meta = dict(n=scene.n, sh_degree=scene.sh_degree, max_coeffs=scene.max_coeffs)
    img = RenderFn.apply(*leaves, hip, meta, cam, cfg)
    loss = (img - torch.tensor(gt, device=d)).abs().mean()
    loss.backward()
    rng = np.random.default_rng(seed)
    worst = 0.0
    for li, param in enumerate(["pos", "scale", "rot", "sh", "op"]):
        g = leaves[li].grad.detach().cpu().numpy().reshape(-1)
        scale = max(np.abs(g).max(), 1e-12)
        for idx in rng.choice(g.size, size=min(6, g.size), replace=False):
            fd = _fd(oracle64, scene, cam, cfg, gt, param, int(idx))
            err = abs(fd - g[idx]) / scale
            worst = max(worst, err)
            assert abs(fd - g[idx]) <= 2e-2 * scale + 1e-7, (param, idx, fd, g[idx])
    print("worst FD discrepancy (fraction of max |grad|):", worst, flush=True)
