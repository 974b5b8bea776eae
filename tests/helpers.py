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


# ---- eval_3D forward cases (activated parameters) ---------------------------
def _activated(scene):
    """RawGaussianParams::activate (train_types.cpp:6-36) via the oracle."""
    from oracle.oracle import Oracle
    return Oracle().activate(scene)


def eval3d_reference_scene():
    """PreprocessorEval3DTestScene of the reference's
    tests/test_preprocessor_gpu.cpp:247-338: 3 Gaussians, 32x32, identity
    view, w = z projection, SH degree 0."""
    n, W, H = 3, 32, 32
    pos = np.array([[i * 0.5 - 0.5, i * 0.3 - 0.3, 3.0 + i * 0.5] for i in range(n)], np.float32)
    g = dict(pos=pos, scale=np.full((n, 3), 0.1, np.float32),
             rot=np.tile(np.array([1, 0, 0, 0], np.float32), (n, 1)), sh=np.full((n, 1, 3), 0.5, np.float32),
             op=np.full(n, 0.8, np.float32))
    view = np.zeros(16, np.float32)
    view[[0, 5, 10, 15]] = 1.0
    fx = np.float32(W) / (np.float32(2.0) * np.float32(0.5))
    fy = np.float32(H) / (np.float32(2.0) * np.float32(0.5))
    vp = np.zeros(16, np.float32)
    vp[0], vp[5], vp[10], vp[11] = fx, fy, 1.0, 1.0
    cam = Camera(view, vp, np.zeros(3, np.float32), 0.5, 0.5, W, H)
    cfg = RenderCfg(bg_color=(0.0, 0.0, 0.0), training=True, sh_degree=0)
    return dict(name="reference_eval3d_scene", g=g, n=n, deg=0, K=1, cam=cam, cfg=cfg,
                filter3d=np.zeros(n, np.float32))


def eval3d_cases(scale=1):
    """Seeded eval_3D workloads: a cube seen from outside, a cloud with the
    camera inside it (near-plane fallback bounds, camera-inside culls), and
    large overlapping Gaussians (k-buffer overflow, > 16 blended per pixel)."""
    from gs_scenes import gaussians_normal, gaussians_uniform_cube, make_camera
    out = [eval3d_reference_scene()]
    rng = np.random.Generator(np.random.Philox(77))
    s = gaussians_uniform_cube(3000 * scale, 11, 1)
    out.append(dict(name="cube_sh1", g=_activated(s), n=s.n, deg=1, K=s.max_coeffs,
                    cam=make_camera((0.0, 0.0, -3.0), 160 * scale, 120 * scale, 60.0),
                    cfg=RenderCfg(bg_color=(0.1, 0.2, 0.3), training=False, sh_degree=1),
                    filter3d=rng.uniform(0.0, 0.02, s.n).astype(np.float32)))
    s = gaussians_normal(4000 * scale, 12, 3, sigma=1.5, logscale=(-3.5, 0.6), op=(0.0, 1.5))
    out.append(dict(name="inside_cloud_sh3", g=_activated(s), n=s.n, deg=3, K=s.max_coeffs,
                    cam=make_camera((0.3, 0.2, -1.2), 128 * scale, 96 * scale, 70.0),
                    cfg=RenderCfg(bg_color=(0.0, 0.0, 0.0), training=True, sh_degree=3), filter3d=None))
    s = gaussians_uniform_cube(2000 * scale, 13, 0, logscale=(-2.2, 0.3), op=(1.5, 0.5))
    out.append(dict(name="dense_kbuffer_sh0", g=_activated(s), n=s.n, deg=0, K=s.max_coeffs,
                    cam=make_camera((0.0, 0.5, -2.8), 96 * scale, 64 * scale, 55.0),
                    cfg=RenderCfg(bg_color=(1.0, 1.0, 1.0), training=False, sh_degree=0), filter3d=None))
    return out


# ---- 3DGS PLY files ---------------------------------------------------------
def write_ply(path, n, rest, seed, filter3d=False, normals=True, zero_rot_at=None):
    """A binary little-endian 3DGS PLY with `rest` f_rest_* properties (raw,
    pre-activation values), as the reference's loader reads (ply_loader.cpp:32-121).
    Returns the raw property table and its names."""
    rng = np.random.default_rng(seed)
    names = ["x", "y", "z"] + (["nx", "ny", "nz"] if normals else []) + ["f_dc_0", "f_dc_1", "f_dc_2"]
    names += [f"f_rest_{k}" for k in range(rest)] + ["opacity", "scale_0", "scale_1", "scale_2",
                                                      "rot_0", "rot_1", "rot_2", "rot_3"]
    if filter3d:
        names.append("filter_3D")
    raw = rng.normal(0.0, 1.0, (n, len(names))).astype(np.float32)
    if zero_rot_at is not None:
        for k in ("rot_0", "rot_1", "rot_2", "rot_3"):
            raw[zero_rot_at, names.index(k)] = 0.0
    hdr = ["ply", "format binary_little_endian 1.0", f"element vertex {n}"]
    hdr += [f"property float {nm}" for nm in names] + ["end_header"]
    with open(path, "wb") as f:
        f.write(("\n".join(hdr) + "\n").encode())
        f.write(raw.astype("<f4").tobytes())
    return raw, names


def fd_mismatches(grads, idx, fd, steps, max_coeffs, group_bar, rel_bar, abs_frac, min_clean=0.9):
    """Analytic gradients (dict pos/scale/rot/sh/op of full arrays) of the
    Gaussians `idx` against FP64 finite differences `fd` / `steps` from
    oracle.fd_grads_local (rows in idx order; a step of 0 marks a derivative
    whose every step spans a discontinuity of the forward, left out).  Within
    each parameter group: max |g - fd| <= group_bar * max |fd|; each element:
    |g - fd| <= rel_bar |fd| + abs_frac max |fd|; at least min_clean of the
    derivatives clean.  Returns (failures, report lines)."""
    n = len(idx)
    g = np.concatenate([np.asarray(grads["pos"])[idx], np.asarray(grads["scale"])[idx],
                        np.asarray(grads["rot"])[idx], np.asarray(grads["sh"])[idx].reshape(n, -1),
                        np.asarray(grads["op"])[idx].reshape(n, 1)], 1).astype(np.float64)
    ok = steps > 0
    fails, report = [], [f"clean derivatives: {ok.mean():.3f} of {ok.size}"]
    if ok.mean() < min_clean:
        fails.append(("clean", float(ok.mean())))
    K = max_coeffs
    groups = {"pos": (0, 3), "scale": (3, 3), "rot": (6, 4), "sh": (10, 3 * K), "op": (10 + 3 * K, 1)}
    for nm, (a, l) in groups.items():
        m = ok[:, a:a + l]
        gg, ff = g[:, a:a + l][m], fd[:, a:a + l][m]
        if ff.size == 0:
            continue
        top = max(np.abs(ff).max(), 1e-30)
        e_group = np.abs(gg - ff).max() / top
        excess = (np.abs(gg - ff) - (rel_bar * np.abs(ff) + abs_frac * top)).max()
        report.append(f"{nm}: {gg.size} values, max|g-fd|/max|fd| {e_group:.2e}, elementwise excess {excess:.2e}")
        if e_group > group_bar or excess > 0:
            fails.append((nm, e_group, excess))
    return fails, report

