"""Test-harness autograd wrapper over the C ABI, checked against FP64 central
finite differences (h = 1e-4, smaller where 1e-4 spans a discontinuity) of
the double-precision oracle forward, every parameter of a small scene.

The wrapper lives here, in tests/, on purpose: it is a checking device for the
hand-written adjoint kernels, not part of the product.
"""
import numpy as np
import pytest
import torch

from gs_scenes import RenderCfg, random_small_scene

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


@pytest.mark.parametrize("deg,seed", [(0, 71), (1, 72), (3, 73)])
def test_autograd_gradients_match_fp64_finite_differences(hip, deg, seed):
    """Every raw parameter of every Gaussian of a small scene: the gradients
    through the wrapper against FP64 central differences of the oracle's
    double forward (oracle.fd_grads_local: h = 1e-4, smaller steps where 1e-4
    spans a discontinuity).  Bars: within each parameter group 1e-3 of its
    largest derivative; each element 1e-2 of itself plus 1e-4 of that largest."""
    from helpers import fd_mismatches
    from oracle.oracle import fd_grads_local
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
    grads = {k: leaves[li].grad.detach().cpu().numpy() for li, k in enumerate(["pos", "scale", "rot", "sh", "op"])}
    idx = np.arange(scene.n)
    fd, steps = fd_grads_local(scene, cam, cfg, gt, idx, h=1e-4)
    fails, report = fd_mismatches(grads, idx, fd, steps, scene.max_coeffs, 1e-3, 1e-2, 1e-4)
    print("\n".join(report), flush=True)
    assert not fails, fails
