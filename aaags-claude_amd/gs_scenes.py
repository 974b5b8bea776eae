"""Seeded synthetic scenes and pinhole cameras for the 3DGS render-and-train path.

The scene distributions are the benchmark workloads of this project (10k/200k/1M/6M
Gaussians); every random draw comes from numpy's Philox-4x32 counter-based
generator seeded with the workload's seed, so the CPU oracle and the HIP path
see identical float32 inputs.

Conventions follow the reference:
- raw (pre-activation) parameters, SoA, SH interleaved per basis
  (``sh[i, k, ch]``), ``harmonyos_3dgs/include/train_types.h:9-18``;
- column-major look-at view matrix and ``viewproj = proj * view`` with
  ``proj[11] = 1`` (w = view z), ``harmonyos_3dgs/src/main.cpp:28-85``;
- ``tan_fovy = tan_fovx * H / W`` for a horizontal field of view
  (``main.cpp:116-117``).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field

import numpy as np


class GsCamera(ctypes.Structure):
    """C layout of ``gs_camera`` (same field order as the reference ``Camera``,
    ``harmonyos_3dgs/include/types.h:48-54``)."""

    _fields_ = [
        ("view_matrix", ctypes.c_float * 16),
        ("viewproj_matrix", ctypes.c_float * 16),
        ("cam_pos", ctypes.c_float * 3),
        ("tan_fovx", ctypes.c_float),
        ("tan_fovy", ctypes.c_float),
        ("width", ctypes.c_int),
        ("height", ctypes.c_int),
    ]


class GsRenderCfg(ctypes.Structure):
    """C layout of ``gs_render_cfg`` (the 2D-path fields of the reference
    ``RenderConfig``, ``harmonyos_3dgs/include/types.h:57-66``)."""

    _fields_ = [
        ("bg_color", ctypes.c_float * 3),
        ("scale_modifier", ctypes.c_float),
        ("antialiasing", ctypes.c_int),
        ("training", ctypes.c_int),
        ("sh_degree", ctypes.c_int),
        ("tile_w", ctypes.c_int),
        ("tile_h", ctypes.c_int),
    ]


@dataclass
class Camera:
    view: np.ndarray  # (16,) float32, column-major
    viewproj: np.ndarray  # (16,) float32, column-major
    cam_pos: np.ndarray  # (3,) float32
    tan_fovx: float
    tan_fovy: float
    width: int
    height: int

    def to_c(self) -> GsCamera:
        c = GsCamera()
        for i in range(16):
            c.view_matrix[i] = float(self.view[i])
            c.viewproj_matrix[i] = float(self.viewproj[i])
        for i in range(3):
            c.cam_pos[i] = float(self.cam_pos[i])
        c.tan_fovx = float(np.float32(self.tan_fovx))
        c.tan_fovy = float(np.float32(self.tan_fovy))
        c.width = int(self.width)
        c.height = int(self.height)
        return c

    @property
    def num_tiles(self) -> int:
        return ((self.width + 15) // 16) * ((self.height + 15) // 16)


@dataclass
class RenderCfg:
    bg_color: tuple = (0.0, 0.0, 0.0)
    scale_modifier: float = 1.0
    antialiasing: bool = False
    training: bool = True
    sh_degree: int = 3
    tile_w: int = 16
    tile_h: int = 16

    def to_c(self) -> GsRenderCfg:
        c = GsRenderCfg()
        for i in range(3):
            c.bg_color[i] = float(self.bg_color[i])
        c.scale_modifier = float(self.scale_modifier)
        c.antialiasing = int(bool(self.antialiasing))
        c.training = int(bool(self.training))
        c.sh_degree = int(self.sh_degree)
        c.tile_w = int(self.tile_w)
        c.tile_h = int(self.tile_h)
        return c


@dataclass
class RawScene:
    """Raw (pre-activation) Gaussian parameters, float32 SoA."""

    pos: np.ndarray  # (N, 3)
    log_scale: np.ndarray  # (N, 3)
    rot: np.ndarray  # (N, 4) quaternion (r, x, y, z), not necessarily unit
    sh: np.ndarray  # (N, K, 3)
    opacity_logit: np.ndarray  # (N,)
    sh_degree: int = 0
    name: str = ""
    meta: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def max_coeffs(self) -> int:
        return int(self.sh.shape[1])

    def copy(self) -> "RawScene":
        return RawScene(self.pos.copy(), self.log_scale.copy(), self.rot.copy(), self.sh.copy(),
                        self.opacity_logit.copy(), self.sh_degree, self.name, dict(self.meta))

    def arrays(self):
        return [self.pos, self.log_scale, self.rot, self.sh, self.opacity_logit]


def _f32(x):
    return np.ascontiguousarray(x, dtype=np.float32)


def look_at(eye, center, up=(0.0, 1.0, 0.0)) -> np.ndarray:
    """Column-major world->camera matrix, camera looking down +z
    (``harmonyos_3dgs/src/main.cpp:28-63``)."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(center, np.float64) - eye
    f /= np.linalg.norm(f)
    up = np.asarray(up, np.float64)
    r = np.cross(f, up)
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    m = np.zeros(16)
    m[0], m[1], m[2] = r[0], u[0], f[0]
    m[4], m[5], m[6] = r[1], u[1], f[1]
    m[8], m[9], m[10] = r[2], u[2], f[2]
    m[12] = -r.dot(eye)
    m[13] = -u.dot(eye)
    m[14] = -f.dot(eye)
    m[15] = 1.0
    return m


def perspective(tan_fovx, tan_fovy, znear=0.01, zfar=100.0) -> np.ndarray:
    """Column-major projection with w = view z (``main.cpp:66-74``)."""
    p = np.zeros(16)
    p[0] = 1.0 / tan_fovx
    p[5] = 1.0 / tan_fovy
    p[10] = zfar / (zfar - znear)
    p[11] = 1.0
    p[14] = -(zfar * znear) / (zfar - znear)
    return p


def _mat4_mul_cm(a, b):
    A = a.reshape(4, 4).T  # column-major storage -> row-major matrix
    B = b.reshape(4, 4).T
    return (A @ B).T.reshape(16)


def make_camera(eye, width, height, fov_deg, center=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
                znear=0.01, zfar=100.0) -> Camera:
    tan_x = float(np.tan(np.radians(fov_deg) * 0.5))
    tan_y = tan_x * height / width
    view = look_at(eye, center, up)
    proj = perspective(tan_x, tan_y, znear, zfar)
    vp = _mat4_mul_cm(proj, view)
    return Camera(_f32(view), _f32(vp), _f32(eye), float(np.float32(tan_x)), float(np.float32(tan_y)),
                  int(width), int(height))


def sphere_eye(rng: np.random.Generator, radius: float) -> np.ndarray:
    """Uniform point on a sphere, away from the +-y poles (look-at up is +y)."""
    while True:
        v = rng.standard_normal(3)
        v /= np.linalg.norm(v)
        if abs(v[1]) < 0.95:
            return v * radius


def _quats(rng, n):
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    return q


def _sh(rng, n, degree, dc_sigma, rest_sigma):
    k = (degree + 1) ** 2
    sh = np.empty((n, k, 3), np.float32)
    sh[:, 0, :] = rng.standard_normal((n, 3), dtype=np.float32) * np.float32(dc_sigma)
    if k > 1:
        sh[:, 1:, :] = rng.standard_normal((n, k - 1, 3), dtype=np.float32) * np.float32(rest_sigma)
    return sh


def gaussians_uniform_cube(n, seed, degree, logscale=(-4.0, 0.5), op=(0.0, 1.0)) -> RawScene:
    rng = np.random.Generator(np.random.Philox(seed))
    pos = rng.uniform(-1.0, 1.0, (n, 3))
    ls = rng.normal(logscale[0], logscale[1], (n, 3))
    rot = _quats(rng, n)
    opl = rng.normal(op[0], op[1], n)
    sh = _sh(rng, n, degree, 0.5, 0.05)
    return RawScene(_f32(pos), _f32(ls), _f32(rot), sh, _f32(opl), degree)


def gaussians_normal(n, seed, degree, sigma=1.5, logscale=(-4.5, 0.7), op=(-1.0, 1.5)) -> RawScene:
    rng = np.random.Generator(np.random.Philox(seed))
    pos = rng.normal(0.0, sigma, (n, 3))
    ls = rng.normal(logscale[0], logscale[1], (n, 3))
    rot = _quats(rng, n)
    opl = rng.normal(op[0], op[1], n)
    sh = _sh(rng, n, degree, 0.5, 0.05)
    return RawScene(_f32(pos), _f32(ls), _f32(rot), sh, _f32(opl), degree)


def gaussians_clusters(n, seed, degree, clusters=32, sigma=0.3, extent=3.0, logscale=(-5.0, 0.8),
                       op=(-2.0, 2.0)) -> RawScene:
    rng = np.random.Generator(np.random.Philox(seed))
    centres = rng.uniform(-extent, extent, (clusters, 3))
    which = rng.integers(0, clusters, n)
    pos = centres[which] + rng.normal(0.0, sigma, (n, 3))
    ls = rng.normal(logscale[0], logscale[1], (n, 3))
    rot = _quats(rng, n)
    opl = rng.normal(op[0], op[1], n)
    sh = _sh(rng, n, degree, 0.5, 0.05)
    return RawScene(_f32(pos), _f32(ls), _f32(rot), sh, _f32(opl), degree)


def target_image(width, height, seed) -> np.ndarray:
    rng = np.random.Generator(np.random.Philox(seed))
    return rng.random((height, width, 3), dtype=np.float32)


# Workloads: (scene builder, cameras builder).  Seeds 0..4, camera seeds +1000,
# target images +2000 (+ view index).
WORKLOADS = {
    "10k_256": dict(seed=0, n=10_000, degree=0, kind="cube", views=[("fixed", 256, 256, 60.0, 3.0)]),
    "200k_720p": dict(seed=1, n=200_000, degree=3, kind="normal", views=[("sphere", 1280, 720, 50.0, 4.0)]),
    "1m_1080p": dict(seed=2, n=1_000_000, degree=3, kind="cube", views=[("fixed", 1920, 1080, 60.0, 3.0)]),
    "1m_16views": dict(seed=3, n=1_000_000, degree=3, kind="cube", views=[("mixed16", 0, 0, 60.0, 4.0)]),
    "6m_8x1080p": dict(seed=4, n=6_000_000, degree=3, kind="clusters", views=[("sphere8", 1920, 1080, 60.0, 8.0)]),
}

MIXED_RES = [(640, 480), (1280, 720), (1600, 900), (1920, 1080)]


def make_cameras(spec, seed) -> list:
    rng = np.random.Generator(np.random.Philox(seed + 1000))
    cams = []
    for kind, w, h, fov, dist in spec:
        if kind == "fixed":
            cams.append(make_camera((0.0, 0.0, -dist), w, h, fov))
        elif kind == "sphere":
            cams.append(make_camera(sphere_eye(rng, dist), w, h, fov))
        elif kind == "mixed16":
            for _ in range(16):
                ww, hh = MIXED_RES[int(rng.integers(0, len(MIXED_RES)))]
                cams.append(make_camera(sphere_eye(rng, dist), ww, hh, fov))
        elif kind == "sphere8":
            for _ in range(8):
                cams.append(make_camera(sphere_eye(rng, dist), w, h, fov))
        else:
            raise ValueError(kind)
    return cams


def make_workload(name: str, n: int | None = None):
    """Returns (scene, cameras, target images) for a named workload.  ``n``
    overrides the Gaussian count (same distributions) for reduced-size runs."""
    w = WORKLOADS[name]
    count = int(n if n is not None else w["n"])
    if w["kind"] == "cube":
        scene = gaussians_uniform_cube(count, w["seed"], w["degree"])
    elif w["kind"] == "normal":
        scene = gaussians_normal(count, w["seed"], w["degree"])
    else:
        scene = gaussians_clusters(count, w["seed"], w["degree"])
    scene.name = name
    cams = make_cameras(w["views"], w["seed"])
    gts = [target_image(c.width, c.height, w["seed"] + 2000 + i) for i, c in enumerate(cams)]
    return scene, cams, gts


def random_small_scene(n, degree, seed, width=64, height=48, spread=1.0, logscale=(-2.5, 0.4),
                       op=(0.5, 1.0), eye=(0.0, 0.0, -4.0), fov=60.0):
    """Small randomized scene for parity and finite-difference tests."""
    rng = np.random.Generator(np.random.Philox(seed))
    pos = rng.uniform(-spread, spread, (n, 3))
    ls = rng.normal(logscale[0], logscale[1], (n, 3))
    rot = rng.standard_normal((n, 4))
    rot[:, 0] += 1.0
    opl = rng.normal(op[0], op[1], n)
    sh = _sh(rng, n, degree, 0.5, 0.2)
    scene = RawScene(_f32(pos), _f32(ls), _f32(rot), sh, _f32(opl), degree, f"small{n}_{seed}")
    cam = make_camera(eye, width, height, fov)
    gt = target_image(width, height, seed + 2000)
    return scene, cam, gt
