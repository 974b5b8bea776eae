"""Shared helpers for the parity tests: golden-fixture loading and the
comparison rules (bit-exact for integer work, 1e-4 max-abs for images,
1e-3 relative for gradients)."""
import json
import os

import numpy as np

from gs_scenes import Camera, RawScene, RenderCfg

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
SCENES = os.path.join(GOLDEN, "scenes")

IMAGE_TOL = 1e-4  # max-abs per channel on [0, 1]
GRAD_RTOL = 1e-3  # relative, per gradient array


def scene_dirs():
    return sorted(d for d in os.listdir(SCENES) if os.path.isdir(os.path.join(SCENES, d)))


def load_golden(name):
    """One of the reference's golden scene directories (train/verify_grads_ext/*,
    train/verify_forward) as (scene, camera, cfg, gt, expected dict)."""
    p = os.path.join(SCENES, name)
    m = json.load(open(os.path.join(p, "meta.json")))
    N, W, H, deg, K = m["N"], m["W"], m["H"], m["sh_degree"], m["max_coeffs"]

    def rd(f, *shape, dt=np.float32):
        return np.fromfile(os.path.join(p, f + ".bin"), dtype=dt).reshape(*shape)

    scene = RawScene(rd("raw_positions", N, 3), rd("raw_scales", N, 3), rd("raw_rotations", N, 4),
                     rd("raw_sh_coeffs", N, K, 3), rd("raw_opacities", N), deg, name)
    cam = Camera(rd("view_matrix", 16), rd("viewproj_matrix", 16), np.zeros(3, np.float32),
                 float(np.float32(m["tan_fovx"])), float(np.float32(m["tan_fovy"])), W, H)
    # verify_grads_main.cpp:117-126: bg 0, training mode, cfg degree = model degree
    cfg = RenderCfg(bg_color=(0.0, 0.0, 0.0), training=True, sh_degree=deg)
    gt = rd("gt_image", H, W, 3)
    exp = dict(img=rd("cpp_rendered", H, W, 3))
    if os.path.exists(os.path.join(p, "cpp_loss.bin")):
        exp["loss"] = float(np.fromfile(os.path.join(p, "cpp_loss.bin"), np.float64)[0])
    for key, fn, shape in [("pos", "positions", (N, 3)), ("scale", "scales", (N, 3)), ("rot", "rotations", (N, 4)),
                           ("sh", "sh_coeffs", (N, K, 3)), ("op", "opacities", (N,))]:
        f = os.path.join(p, "cpp_grad_raw_" + fn + ".bin")
        if os.path.exists(f):
            exp["grad_" + key] = np.fromfile(f, np.float32).reshape(shape)
    return scene, cam, cfg, gt, exp


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.abs(ref).max()
    if den == 0:
        return float(np.abs(got).max())
    return float(np.abs(got - ref).max() / den)


def assert_grad_close(got, ref, name, rtol=GRAD_RTOL):
    """Relative error of a gradient array, normalised by its largest entry."""
    e = rel_err(got, ref)
    assert e <= rtol, f"{name}: relative error {e:.3e} > {rtol:.0e}"


def assert_image_close(got, ref, name="image", tol=IMAGE_TOL):
    e = float(np.abs(np.asarray(got, np.float64) - np.asarray(ref, np.float64)).max())
    assert e <= tol, f"{name}: max-abs {e:.3e} > {tol:.0e}"


def u32(t):
    """torch int32/int64 tensor holding unsigned data -> numpy unsigned view."""
    a = t.cpu().numpy()
    return a.view(np.uint32) if a.dtype == np.int32 else a.view(np.uint64)
