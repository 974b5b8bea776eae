"""ctypes binding of the CPU oracle (oracle/gs_oracle.c) and of the reference
compiled from source (oracle/_ref/libgs_ref.so, when it was built).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
REF_LIB = os.path.join(HERE, "_ref", "libgs_ref.so")

_i32 = ctypes.c_int
_i64 = ctypes.c_longlong
_f32 = ctypes.c_float
_vp = ctypes.c_void_p


class _Cam(ctypes.Structure):
    _fields_ = [("view_matrix", ctypes.c_float * 16), ("viewproj_matrix", ctypes.c_float * 16),
                ("cam_pos", ctypes.c_float * 3), ("tan_fovx", ctypes.c_float), ("tan_fovy", ctypes.c_float),
                ("width", ctypes.c_int), ("height", ctypes.c_int)]


class _Cfg(ctypes.Structure):
    _fields_ = [("bg_color", ctypes.c_float * 3), ("scale_modifier", ctypes.c_float),
                ("antialiasing", ctypes.c_int), ("training", ctypes.c_int), ("sh_degree", ctypes.c_int),
                ("tile_w", ctypes.c_int), ("tile_h", ctypes.c_int)]


def cam_struct(cam) -> _Cam:
    c = _Cam()
    for i in range(16):
        c.view_matrix[i] = float(cam.view[i])
        c.viewproj_matrix[i] = float(cam.viewproj[i])
    for i in range(3):
        c.cam_pos[i] = float(cam.cam_pos[i])
    c.tan_fovx, c.tan_fovy = float(cam.tan_fovx), float(cam.tan_fovy)
    c.width, c.height = int(cam.width), int(cam.height)
    return c


def cfg_struct(cfg) -> _Cfg:
    c = _Cfg()
    for i in range(3):
        c.bg_color[i] = float(cfg.bg_color[i])
    c.scale_modifier = float(cfg.scale_modifier)
    c.antialiasing, c.training = int(cfg.antialiasing), int(cfg.training)
    c.sh_degree, c.tile_w, c.tile_h = int(cfg.sh_degree), int(cfg.tile_w), int(cfg.tile_h)
    return c


def build():
    """Compiles the oracle (and, where the reference tree is present, oracle/_ref)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference/harmonyos_3dgs/src"):
        subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile.ref")], check=True)


def _p(a):
    return a.ctypes.data_as(_vp) if a is not None else None


class Oracle:
    """The float (``double=False``) or FP64 (``double=True``) oracle."""

    def __init__(self, double: bool = False):
        name = "libgs_oracle64.so" if double else "libgs_oracle.so"
        path = os.path.join(BUILD, name)
        if not os.path.exists(path):
            build()
        self.lib = ctypes.CDLL(path)
        self.double = double
        self.dt = np.float64 if double else np.float32
        p = "gsd_" if double else "gso_"
        L = self.lib
        self._activate = getattr(L, p + "activate")
        self._preprocess = getattr(L, p + "preprocess")
        self._count = getattr(L, p + "count_pairs")
        self._count.restype = _i64
        self._dup = getattr(L, p + "duplicate")
        self._sort = getattr(L, p + "sort_pairs")
        self._sort.argtypes = [_i64, _vp, _vp, _vp, _vp, _i32, _vp]
        self._raster = getattr(L, p + "rasterize")
        self._l1 = getattr(L, p + "l1_loss")
        self._l1.restype = _f32
        self._l1d = getattr(L, p + "l1_loss_double")
        self._l1d.restype = ctypes.c_double
        self._rbwd = getattr(L, p + "raster_backward")
        self._pbwd = getattr(L, p + "preprocess_backward")
        self._adam = getattr(L, p + "adam_update")
        self._adam.argtypes = [_i64, _vp, _vp, _vp, _vp, _f32, _f32, _f32, _f32, _f32, _f32]
        self._bc = getattr(L, p + "adam_bias_corrections")
        self._bc.argtypes = [_f32, _f32, _i32, _vp, _vp]
        self._lrs = getattr(L, p + "lr_schedule")
        self._lrs.argtypes = [_f32, _f32, _i32, _i32]
        self._lrs.restype = _f32

    def _a(self, x, shape=None):
        a = np.ascontiguousarray(x, dtype=self.dt)
        return a if shape is None else a.reshape(shape)

    # --- stages -----------------------------------------------------------
    def activate(self, scene):
        n, k = scene.n, scene.max_coeffs
        raw = [self._a(a) for a in scene.arrays()]
        pos, sc, rot, sh, op = (np.empty_like(raw[0]), np.empty_like(raw[1]), np.empty_like(raw[2]),
                                np.empty_like(raw[3]), np.empty_like(raw[4]))
        self._activate(_i32(n), _i32(k), *[_p(a) for a in raw], _p(pos), _p(sc), _p(rot), _p(sh), _p(op))
        return dict(pos=pos, scale=sc, rot=rot, sh=sh, op=op)

    def preprocess(self, g, n, model_deg, max_coeffs, cam, cfg, cache=True):
        dt = self.dt
        out = dict(means2D=np.zeros((n, 2), dt), depths=np.zeros(n, dt), conics=np.zeros((n, 3), dt),
                   op2d=np.zeros(n, dt), rgb=np.zeros((n, 3), dt), radii=np.zeros(n, np.int32),
                   tiles=np.zeros(n, np.int32))
        if cache:
            out.update(cov3D=np.zeros((n, 6), dt), p_view=np.zeros((n, 3), dt), p_hom_w=np.zeros(n, dt),
                       cov2D=np.zeros((n, 3), dt), det=np.zeros(n, dt))
        c, f = cam_struct(cam), cfg_struct(cfg)
        self._preprocess(_i32(n), _i32(model_deg), _i32(max_coeffs), _p(g["pos"]), _p(g["sh"]),
                         _p(g["scale"]), _p(g["rot"]), _p(g["op"]), ctypes.byref(c), ctypes.byref(f),
                         _p(out["means2D"]), _p(out["depths"]), _p(out["conics"]), _p(out["op2d"]),
                         _p(out["rgb"]), _p(out["radii"]), _p(out["tiles"]),
                         _p(out.get("cov3D")), _p(out.get("p_view")), _p(out.get("p_hom_w")),
                         _p(out.get("cov2D")), _p(out.get("det")))
        return out

    def duplicate(self, pre, n, width, height, tile=16):
        tiles = np.ascontiguousarray(pre["tiles"], np.int32)
        K = int(self._count(_i32(n), _p(tiles)))
        keys = np.zeros(max(K, 1), np.uint64)
        vals = np.zeros(max(K, 1), np.uint32)
        self._dup(_i32(n), _p(self._a(pre["means2D"])), _p(self._a(pre["depths"])),
                  _p(np.ascontiguousarray(pre["radii"], np.int32)), _p(tiles), _i32(width), _i32(height),
                  _i32(tile), _i32(tile), _p(keys), _p(vals))
        return keys[:K], vals[:K]

    def sort_pairs(self, keys, vals, num_tiles):
        K = len(keys)
        keys = np.ascontiguousarray(keys, np.uint64)
        vals = np.ascontiguousarray(vals, np.uint32)
        ko = np.zeros(max(K, 1), np.uint64)
        vo = np.zeros(max(K, 1), np.uint32)
        ranges = np.zeros(num_tiles * 2, np.uint32)
        self._sort(K, _p(keys), _p(vals), _p(ko), _p(vo), num_tiles, _p(ranges))
        return ko[:K], vo[:K], ranges

    def rasterize(self, pre, values_sorted, ranges, width, height, bg=(0, 0, 0), tile=16):
        dt = self.dt
        img = np.zeros((height, width, 3), dt)
        T = np.zeros((height, width), dt)
        nc = np.zeros((height, width), np.int32)
        last = np.zeros((height, width), np.uint32)
        bgv = np.asarray(bg, np.float32)
        vs = np.ascontiguousarray(values_sorted, np.uint32)
        if vs.size == 0:
            vs = np.zeros(1, np.uint32)
        self._raster(_p(self._a(pre["means2D"])), _p(self._a(pre["conics"])), _p(self._a(pre["op2d"])),
                     _p(self._a(pre["rgb"])), _p(vs), _p(np.ascontiguousarray(ranges, np.uint32)),
                     _i32(width), _i32(height), _i32(tile), _i32(tile), _p(bgv), _p(img), _p(T), _p(nc),
                     _p(last))
        return img, T, nc, last

    def l1_loss(self, rendered, gt, double_sum=False):
        r = self._a(rendered)
        g = self._a(gt)
        h, w = r.shape[0], r.shape[1]
        d = np.zeros_like(r)
        fn = self._l1d if double_sum else self._l1
        loss = fn(_p(r), _p(g), _i32(h), _i32(w), _p(d))
        return loss, d

    def raster_backward(self, pre, values_sorted, ranges, width, height, T_final, n_contrib, d_image,
                        bg=(0, 0, 0), tile=16):
        n = pre["radii"].shape[0]
        dt = self.dt
        g = dict(d_means2D=np.zeros((n, 2), dt), d_conics=np.zeros((n, 3), dt), d_rgb=np.zeros((n, 3), dt),
                 d_op2d=np.zeros(n, dt))
        vs = np.ascontiguousarray(values_sorted, np.uint32)
        if vs.size == 0:
            vs = np.zeros(1, np.uint32)
        bgv = np.asarray(bg, np.float32)
        self._rbwd(_p(self._a(pre["means2D"])), _p(self._a(pre["conics"])), _p(self._a(pre["op2d"])),
                   _p(self._a(pre["rgb"])), _p(vs), _p(np.ascontiguousarray(ranges, np.uint32)), _i32(width),
                   _i32(height), _i32(tile), _i32(tile), _p(bgv), _p(self._a(T_final)),
                   _p(np.ascontiguousarray(n_contrib, np.int32)), _p(self._a(d_image)), _p(g["d_means2D"]),
                   _p(g["d_conics"]), _p(g["d_rgb"]), _p(g["d_op2d"]))
        return g

    def preprocess_backward(self, g, raw_rot, n, model_deg, max_coeffs, cam, cfg, pre, rg):
        dt = self.dt
        out = dict(pos=np.zeros((n, 3), dt), scale=np.zeros((n, 3), dt), rot=np.zeros((n, 4), dt),
                   sh=np.zeros((n, max_coeffs, 3), dt), op=np.zeros(n, dt))
        c, f = cam_struct(cam), cfg_struct(cfg)
        self._pbwd(_i32(n), _i32(model_deg), _i32(max_coeffs), _p(g["pos"]), _p(g["sh"]), _p(g["scale"]),
                   _p(g["rot"]), _p(g["op"]), _p(self._a(raw_rot)), ctypes.byref(c), ctypes.byref(f),
                   _p(np.ascontiguousarray(pre["radii"], np.int32)), _p(self._a(pre["means2D"])),
                   _p(self._a(pre["cov3D"])), _p(self._a(pre["p_view"])), _p(self._a(pre["p_hom_w"])),
                   _p(self._a(pre["cov2D"])), _p(self._a(pre["det"])), _p(self._a(rg["d_means2D"])),
                   _p(self._a(rg["d_conics"])), _p(self._a(rg["d_rgb"])), _p(self._a(rg["d_op2d"])),
                   _p(out["pos"]), _p(out["scale"]), _p(out["rot"]), _p(out["sh"]), _p(out["op"]))
        return out

    def adam_update(self, p, g, m, v, lr, b1, b2, eps, bc1, bc2):
        """In place on p, m, v (float32 contiguous arrays)."""
        self._adam(p.size, _p(p), _p(np.ascontiguousarray(g, self.dt)), _p(m), _p(v), lr, b1, b2, eps, bc1, bc2)

    def bias_corrections(self, b1, b2, t):
        a, b = _f32(), _f32()
        self._bc(b1, b2, t, ctypes.byref(a), ctypes.byref(b))
        return a.value, b.value

    def lr_schedule(self, lr_init, lr_final, step, max_steps):
        return self._lrs(lr_init, lr_final, step, max_steps)

    # --- whole path -------------------------------------------------------
    def forward(self, scene, cam, cfg, backward_cache=True):
        g = self.activate(scene)
        n = scene.n
        pre = self.preprocess(g, n, scene.sh_degree, scene.max_coeffs, cam, cfg, cache=backward_cache)
        keys, vals = self.duplicate(pre, n, cam.width, cam.height, cfg.tile_w)
        ks, vs, ranges = self.sort_pairs(keys, vals, cam.num_tiles)
        img, T, nc, last = self.rasterize(pre, vs, ranges, cam.width, cam.height, cfg.bg_color, cfg.tile_w)
        return dict(g=g, pre=pre, keys=keys, values=vals, keys_sorted=ks, values_sorted=vs, ranges=ranges,
                    img=img, T_final=T, n_contrib=nc, last_contrib=last)

    def forward_backward(self, scene, cam, cfg, gt, double_loss=True):
        f = self.forward(scene, cam, cfg)
        loss, d_img = self.l1_loss(f["img"], gt, double_sum=double_loss)
        rg = self.raster_backward(f["pre"], f["values_sorted"], f["ranges"], cam.width, cam.height,
                                  f["T_final"], f["n_contrib"], d_img, cfg.bg_color, cfg.tile_w)
        grads = self.preprocess_backward(f["g"], scene.rot, scene.n, scene.sh_degree, scene.max_coeffs, cam,
                                         cfg, f["pre"], rg)
        f.update(loss=loss, d_image=d_img, rgrad=rg, grads=grads)
        return f

    def loss(self, scene, cam, cfg, gt):
        """L1 loss of the full forward (double accumulation)."""
        f = self.forward(scene, cam, cfg, backward_cache=False)
        return self.l1_loss(f["img"], gt, double_sum=True)[0]


def _rects(means2D, radii, width, height, tile=16):
    """getRect (math_utils.cpp:21-28) for every Gaussian: (int) truncation
    toward zero, then the clamp to the tile grid.  Returns mn_x, mn_y, mx_x, mx_y."""
    gx, gy = (width + tile - 1) // tile, (height + tile - 1) // tile
    p = np.asarray(means2D, np.float64)
    r = np.asarray(radii, np.float64)
    out = []
    for v, g, add in ((p[:, 0] - r, gx, 0), (p[:, 1] - r, gy, 0), (p[:, 0] + r, gx, tile - 1),
                      (p[:, 1] + r, gy, tile - 1)):
        out.append(np.clip(np.trunc((v + add) / tile), 0, g).astype(np.int64))
    return out


def fd_grads_local(scene, cam, cfg, gt, targets, h=1e-4, fallback=(1e-5, 1e-6), threads=None, margin=1):
    """FP64 central finite differences (gsd_fd_local) of the mean L1 loss with
    respect to the raw parameters of the Gaussians ``targets``, on a scene of
    any size: each target is evaluated on the subset of Gaussians whose tile
    rectangles meet its own rectangle grown by ``margin`` tiles (grown further
    while a perturbed rectangle leaves it).

    Every parameter is first differenced with step ``h``.  Where that
    difference spans a discontinuity of the forward (a flipped alpha cut,
    transmittance stop, clamp, rectangle or loss sign in any pixel), it is
    taken again with the smaller steps of ``fallback`` in turn.  Returns
    (grads, steps): both [len(targets), P] with P = 11 + 3 * max_coeffs in the
    order pos, scale, rot, sh, opacity; steps holds the step each derivative
    was taken with, or 0 where every step spans a discontinuity."""
    from concurrent.futures import ThreadPoolExecutor
    o = Oracle(double=True)
    n, K = scene.n, scene.max_coeffs
    g = o.activate(scene)
    pre = o.preprocess(g, n, scene.sh_degree, K, cam, cfg, cache=False)
    radii = pre["radii"]
    mnx, mny, mxx, mxy = _rects(pre["means2D"], radii, cam.width, cam.height, cfg.tile_w)
    gx, gy = (cam.width + cfg.tile_w - 1) // cfg.tile_w, (cam.height + cfg.tile_h - 1) // cfg.tile_h
    raw = [np.ascontiguousarray(a, np.float64) for a in scene.arrays()]
    gt = np.ascontiguousarray(gt, np.float32)
    c, f = cam_struct(cam), cfg_struct(cfg)
    fn = o.lib.gsd_fd_local
    fn.restype = _i32
    fn.argtypes = [_i32, _i32, _i32, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _i32, _i32, _i32, _i32,
                   _vp, _vp, _vp]
    P = 11 + 3 * K
    live = radii > 0

    def one(t):
        m = margin
        gr = np.zeros(P, np.float64)
        used = np.zeros(P, np.float64)
        todo = np.ones(P, bool)
        for step in (h,) + tuple(fallback):
            while True:
                rx0, ry0 = max(0, int(mnx[t]) - m), max(0, int(mny[t]) - m)
                rx1, ry1 = min(gx, int(mxx[t]) + m), min(gy, int(mxy[t]) + m)
                sel = live & (mnx < rx1) & (mxx > rx0) & (mny < ry1) & (mxy > ry0)
                sel[t] = True
                idx = np.nonzero(sel)[0]
                sub = [np.ascontiguousarray(a[idx]) for a in raw]
                tl = int(np.searchsorted(idx, t))
                hs = np.where(todo, step, 0.0)
                g1 = np.zeros(P, np.float64)
                fl = np.zeros(P, np.int32)
                st = fn(len(idx), scene.sh_degree, K, *[_p(a) for a in sub], tl, ctypes.byref(c), ctypes.byref(f),
                        _p(gt), rx0, ry0, rx1, ry1, _p(hs), _p(g1), _p(fl))
                if st == 0:
                    break
                if st != 1 or (rx0 == 0 and ry0 == 0 and rx1 == gx and ry1 == gy):
                    raise RuntimeError(f"gsd_fd_local failed ({st}) for Gaussian {t}")
                m *= 2
            clean = todo & (fl == 0)
            gr[clean] = g1[clean]
            used[clean] = step
            todo &= ~clean
            if not todo.any():
                break
        return gr, used

    workers = threads or min(16, os.cpu_count() or 4)
    with ThreadPoolExecutor(workers) as ex:
        res = list(ex.map(one, [int(t) for t in targets]))
    return np.stack([r[0] for r in res]), np.stack([r[1] for r in res])


class OracleDensity:
    """The oracle's density controller (gso_dc_*, gso_densify_and_prune): the
    statistics, the generator state and the parameter arrays on the host."""

    def __init__(self):
        self.L = ctypes.CDLL(os.path.join(BUILD, "libgs_oracle.so"))
        self.L.gso_densify_and_prune.restype = _i32
        self.L.gso_densify_and_prune.argtypes = [_i32, _i32] + [_vp] * 8 + [_f32, _f32, _i32, _f32, _vp]
        self.L.gso_dc_accumulate.argtypes = [_i32, _vp, _vp, _vp, _vp, _vp]
        self.L.gso_dc_randn.restype = _f32
        self.L.gso_dc_randn.argtypes = [_vp]
        self.rng = np.array([12345], np.uint64)
        self.n = -1

    def reset(self, n):
        self.n = n
        self.accum = np.zeros(n, np.float32)
        self.denom = np.zeros(n, np.float32)
        self.maxr = np.zeros(n, np.int32)

    def accumulate(self, d_means2D, radii=None):
        d = np.ascontiguousarray(d_means2D, np.float32)
        n = d.size // 2
        if n != self.n:
            self.reset(n)
        r = None if radii is None else np.ascontiguousarray(radii, np.int32)
        self.L.gso_dc_accumulate(n, _p(d), _p(r), _p(self.accum), _p(self.denom), _p(self.maxr))

    def densify(self, params, extent, grad_threshold=0.0002, min_opacity=0.005, max_screen_size=20, **_):
        """params: [pos, scale, rot, sh, op] (raw); returns the new list."""
        n, k = params[0].shape[0], params[3].shape[1]
        if n != self.n:
            self.reset(n)
        cap = 3 * n + 1
        bufs = [np.zeros((cap,) + a.shape[1:], np.float32) for a in params]
        for b, a in zip(bufs, params):
            b[:n] = a
        m = self.L.gso_densify_and_prune(n, k, *[_p(b) for b in bufs], _p(self.accum), _p(self.denom), _p(self.maxr),
                                          grad_threshold, min_opacity, max_screen_size, extent, _p(self.rng))
        self.reset(m)
        return [b[:m].copy() for b in bufs]

    def randn(self, count):
        return np.array([self.L.gso_dc_randn(_p(self.rng)) for _ in range(count)], np.float32)


def inv_depth_oracle(pre, values_sorted, ranges, width, height):
    """gso_rasterize_inv_depth: the inverse-depth output of the 2D forward."""
    L = ctypes.CDLL(os.path.join(BUILD, "libgs_oracle.so"))
    f = lambda a: np.ascontiguousarray(a, np.float32)
    out = np.zeros((height, width), np.float32)
    vs = np.ascontiguousarray(values_sorted, np.uint32)
    if vs.size == 0:
        vs = np.zeros(1, np.uint32)
    L.gso_rasterize_inv_depth(_p(f(pre["means2D"])), _p(f(pre["conics"])), _p(f(pre["op2d"])), _p(f(pre["depths"])),
                              _p(vs), _p(np.ascontiguousarray(ranges, np.uint32)), _i32(width), _i32(height),
                              _i32(16), _i32(16), _p(out))
    return out


class Oracle3D:
    """The eval_3D forward of the float oracle (gso_preprocess3d & co.)."""

    def __init__(self):
        path = os.path.join(BUILD, "libgs_oracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        self.lib = L
        L.gso_frustum_contrib.restype = _f32
        L.gso_frustum_contrib.argtypes = [_f32, _f32, _f32, _f32, _vp, _vp]
        L.gso_ray_contrib.restype = _f32
        L.gso_gauss2screen.restype = _f32
        L.gso_gauss2screen.argtypes = [_vp, _vp, _vp, _f32, _vp, _vp, _vp, _f32, _f32, _f32, _i32, _i32, _vp, _vp]
        L.gso_aabb_screen.argtypes = [_vp, _f32, _vp, _vp]
        L.gso_aabb_view.argtypes = [_vp, _vp, _f32, _f32, _f32, _f32, _f32, _vp, _vp]
        L.gso_count_pairs.restype = _i64
        self._sort = L.gso_sort_pairs
        self._sort.argtypes = [_i64, _vp, _vp, _vp, _vp, _i32, _vp]

    @staticmethod
    def _f(x, shape=None):
        a = np.ascontiguousarray(x, dtype=np.float32)
        return a if shape is None else a.reshape(shape)

    def invert4(self, m):
        inv = np.zeros(16, np.float32)
        ok = self.lib.gso_invert4(_p(self._f(m)), _p(inv))
        return bool(ok), inv

    def ray_contrib(self, pa, pb):
        mp = np.zeros(4, np.float32)
        c = self.lib.gso_ray_contrib(_p(self._f(pa)), _p(self._f(pb)), _p(mp))
        return float(c), mp

    def preprocess(self, g, n, model_deg, max_coeffs, cam, cfg, filter3d=None):
        f32 = np.float32
        out = dict(means2D=np.zeros((n, 2), f32), depths=np.zeros(n, f32), op2d=np.zeros(n, f32),
                   rgb=np.zeros((n, 3), f32), radii=np.zeros(n, np.int32), tiles=np.zeros(n, np.int32),
                   g2s=np.zeros((n, 16), f32), cov3D_inv=np.zeros((n, 6), f32), mean_offset=np.zeros((n, 3), f32))
        c, cf = cam_struct(cam), cfg_struct(cfg)
        flt = None if filter3d is None else self._f(filter3d)
        self.lib.gso_preprocess3d(_i32(n), _i32(model_deg), _i32(max_coeffs), _p(self._f(g["pos"])),
                                  _p(self._f(g["sh"])), _p(self._f(g["scale"])), _p(self._f(g["rot"])),
                                  _p(self._f(g["op"])), _p(flt), ctypes.byref(c), ctypes.byref(cf),
                                  _p(out["means2D"]), _p(out["depths"]), _p(out["op2d"]), _p(out["rgb"]),
                                  _p(out["radii"]), _p(out["tiles"]), _p(out["g2s"]), _p(out["cov3D_inv"]),
                                  _p(out["mean_offset"]))
        return out

    def duplicate(self, pre, n, cam, tile=16):
        tiles = np.ascontiguousarray(pre["tiles"], np.int32)
        K = int(self.lib.gso_count_pairs(_i32(n), _p(tiles)))
        keys = np.zeros(max(K, 1), np.uint64)
        vals = np.zeros(max(K, 1), np.uint32)
        c = cam_struct(cam)
        self.lib.gso_duplicate3d(_i32(n), _p(self._f(pre["means2D"])), _p(self._f(pre["depths"])),
                                 _p(np.ascontiguousarray(pre["radii"], np.int32)), _p(tiles),
                                 _p(self._f(pre["cov3D_inv"])), _p(self._f(pre["mean_offset"])), ctypes.byref(c),
                                 _i32(tile), _i32(tile), _p(keys), _p(vals))
        return keys[:K], vals[:K]

    def sort(self, keys, vals, num_tiles):
        K = len(keys)
        ko = np.zeros(max(K, 1), np.uint64)
        vo = np.zeros(max(K, 1), np.uint32)
        ranges = np.zeros((num_tiles, 2), np.uint32)
        self._sort(_i64(K), _p(np.ascontiguousarray(keys, np.uint64)), _p(np.ascontiguousarray(vals, np.uint32)),
                   _p(ko), _p(vo), _i32(num_tiles), _p(ranges))
        return ko[:K], vo[:K], ranges

    def rasterize(self, pre, values_sorted, ranges, cam, bg, tile=16):
        H, W = cam.height, cam.width
        img = np.zeros((H, W, 3), np.float32)
        T = np.zeros((H, W), np.float32)
        nc = np.zeros((H, W), np.int32)
        c = cam_struct(cam)
        vs = np.ascontiguousarray(values_sorted, np.uint32)
        if vs.size == 0:
            vs = np.zeros(1, np.uint32)
        self.lib.gso_rasterize3d(_p(self._f(pre["depths"])), _p(self._f(pre["op2d"])), _p(self._f(pre["rgb"])),
                                 _p(self._f(pre["g2s"])), _p(self._f(pre["cov3D_inv"])),
                                 _p(self._f(pre["mean_offset"])), _p(vs),
                                 _p(np.ascontiguousarray(ranges, np.uint32)), ctypes.byref(c), _i32(tile),
                                 _i32(tile), _p(self._f(bg)), _p(img), _p(T), _p(nc))
        return img, T, nc

    def render(self, g, n, model_deg, max_coeffs, cam, cfg, filter3d=None):
        """The whole eval_3D forward: preprocess, per-tile keys, stable sort, k-buffer blend."""
        pre = self.preprocess(g, n, model_deg, max_coeffs, cam, cfg, filter3d)
        keys, vals = self.duplicate(pre, n, cam)
        gx, gy = (cam.width + 15) // 16, (cam.height + 15) // 16
        ks, vs, ranges = self.sort(keys, vals, gx * gy)
        img, T, nc = self.rasterize(pre, vs, ranges, cam, cfg.bg_color)
        pre.update(keys=keys, values=vals, keys_sorted=ks, values_sorted=vs, ranges=ranges, img=img,
                   T_final=T, n_contrib=nc)
        return pre


class RefLib:
    """The reference's own CPU classes compiled from source (oracle/_ref)."""

    def __init__(self):
        if not os.path.exists(REF_LIB):
            raise FileNotFoundError(REF_LIB)
        self.lib = ctypes.CDLL(REF_LIB)
        self.lib.ref_forward_backward.restype = _i64
        self.lib.ref_train_steps.restype = _f32

    @staticmethod
    def available() -> bool:
        return os.path.exists(REF_LIB)

    def forward_backward(self, scene, cam, cfg, gt, arena_bytes=None):
        n, k = scene.n, scene.max_coeffs
        raw = [np.ascontiguousarray(a, np.float32).copy() for a in scene.arrays()]
        H, W = cam.height, cam.width
        o = dict(img=np.zeros((H, W, 3), np.float32), T_final=np.zeros((H, W), np.float32),
                 n_contrib=np.zeros((H, W), np.int32), means2D=np.zeros((n, 2), np.float32),
                 depths=np.zeros(n, np.float32), conics=np.zeros((n, 3), np.float32),
                 op2d=np.zeros(n, np.float32), rgb=np.zeros((n, 3), np.float32), radii=np.zeros(n, np.int32),
                 tiles=np.zeros(n, np.int32), g_pos=np.zeros((n, 3), np.float32),
                 g_scale=np.zeros((n, 3), np.float32), g_rot=np.zeros((n, 4), np.float32),
                 g_sh=np.zeros((n, k, 3), np.float32), g_op=np.zeros(n, np.float32),
                 d_means2D=np.zeros((n, 2), np.float32), d_conics=np.zeros((n, 3), np.float32),
                 d_rgb=np.zeros((n, 3), np.float32), d_op2d=np.zeros(n, np.float32))
        loss = ctypes.c_double()
        c, f = cam_struct(cam), cfg_struct(cfg)
        if arena_bytes is None:
            arena_bytes = 64 * 1024 * 1024 + n * 2048 + H * W * 64
        K = self.lib.ref_forward_backward(
            _i32(n), _i32(scene.sh_degree), _i32(k), *[_p(a) for a in raw], ctypes.byref(c), ctypes.byref(f),
            _p(np.ascontiguousarray(gt, np.float32)), ctypes.c_size_t(arena_bytes), _p(o["img"]),
            ctypes.byref(loss), _p(o["T_final"]), _p(o["n_contrib"]), _p(o["means2D"]), _p(o["depths"]),
            _p(o["conics"]), _p(o["op2d"]), _p(o["rgb"]), _p(o["radii"]), _p(o["tiles"]), _p(o["g_pos"]),
            _p(o["g_scale"]), _p(o["g_rot"]), _p(o["g_sh"]), _p(o["g_op"]), _p(o["d_means2D"]),
            _p(o["d_conics"]), _p(o["d_rgb"]), _p(o["d_op2d"]))
        o["loss"] = loss.value
        o["K"] = int(K)
        return o

    def train_steps(self, scene, cam, cfg, gt, tc, first_iteration, steps, arena_bytes=None):
        """Runs Trainer::step with Adam ``steps`` times; returns (loss, raw arrays, image)."""
        n, k = scene.n, scene.max_coeffs
        raw = [np.ascontiguousarray(a, np.float32).copy() for a in scene.arrays()]
        img = np.zeros((cam.height, cam.width, 3), np.float32)
        c, f = cam_struct(cam), cfg_struct(cfg)
        if arena_bytes is None:
            arena_bytes = 64 * 1024 * 1024 + n * 2048 + cam.width * cam.height * 64
        loss = self.lib.ref_train_steps(
            _i32(n), _i32(scene.sh_degree), _i32(k), *[_p(a) for a in raw], ctypes.byref(c), ctypes.byref(f),
            _p(np.ascontiguousarray(gt, np.float32)), _f32(tc["lr_pos_init"]), _f32(tc["lr_pos_final"]),
            _f32(tc["lr_feature"]), _f32(tc["lr_opacity"]), _f32(tc["lr_scaling"]), _f32(tc["lr_rotation"]),
            _i32(tc["max_steps"]), _f32(tc["beta1"]), _f32(tc["beta2"]), _f32(tc["eps"]), _i32(first_iteration),
            _i32(steps), ctypes.c_size_t(arena_bytes), _p(img))
        return float(loss), raw, img

    def render_3d(self, g, n, model_deg, max_coeffs, cam, cfg, filter3d=None, max_pairs=None):
        """The reference's eval_3D forward on activated parameters ``g``."""
        f32 = np.float32
        H, W = cam.height, cam.width
        arr = [np.ascontiguousarray(g[k], f32).copy() for k in ("pos", "scale", "rot", "sh", "op")]
        flt = None if filter3d is None else np.ascontiguousarray(filter3d, f32).copy()
        if max_pairs is None:
            max_pairs = 64 * n + 1024
        gx, gy = (W + 15) // 16, (H + 15) // 16
        o = dict(means2D=np.zeros((n, 2), f32), depths=np.zeros(n, f32), op2d=np.zeros(n, f32),
                 rgb=np.zeros((n, 3), f32), radii=np.zeros(n, np.int32), tiles=np.zeros(n, np.int32),
                 g2s=np.zeros((n, 16), f32), cov3D_inv=np.zeros((n, 6), f32), mean_offset=np.zeros((n, 3), f32),
                 keys=np.zeros(max_pairs, np.uint64), values=np.zeros(max_pairs, np.uint32),
                 values_sorted=np.zeros(max_pairs, np.uint32), ranges=np.zeros((gx * gy, 2), np.uint32),
                 img=np.zeros((H, W, 3), f32), T_final=np.zeros((H, W), f32), n_contrib=np.zeros((H, W), np.int32))
        c, cf = cam_struct(cam), cfg_struct(cfg)
        fn = self.lib.ref_render_3d
        fn.restype = _i64
        arena = 256 * 1024 * 1024 + n * 4096 + H * W * 64 + max_pairs * 64
        K = fn(_i32(n), _i32(model_deg), _i32(max_coeffs), *[_p(a) for a in arr], _p(flt), ctypes.byref(c),
               ctypes.byref(cf), ctypes.c_size_t(arena), _i64(max_pairs), _p(o["means2D"]), _p(o["depths"]),
               _p(o["op2d"]), _p(o["rgb"]), _p(o["radii"]), _p(o["tiles"]), _p(o["g2s"]), _p(o["cov3D_inv"]),
               _p(o["mean_offset"]), _p(o["keys"]), _p(o["values"]), _p(o["values_sorted"]), _p(o["ranges"]),
               _p(o["img"]), _p(o["T_final"]), _p(o["n_contrib"]))
        if K < 0:
            raise ValueError("max_pairs too small")
        for k in ("keys", "values", "values_sorted"):
            o[k] = o[k][:K]
        o["K"] = int(K)
        return o

    def inv_depth(self, g, n, model_deg, max_coeffs, cam, cfg):
        """The reference's 2D forward with Rasterizer::rasterize's inverse-depth output."""
        f32 = np.float32
        arr = [np.ascontiguousarray(g[k], f32).copy() for k in ("pos", "scale", "rot", "sh", "op")]
        img = np.zeros((cam.height, cam.width, 3), f32)
        inv = np.zeros((cam.height, cam.width), f32)
        c, cf = cam_struct(cam), cfg_struct(cfg)
        arena = 256 * 1024 * 1024 + n * 4096 + cam.width * cam.height * 64
        self.lib.ref_inv_depth(_i32(n), _i32(model_deg), _i32(max_coeffs), *[_p(a) for a in arr], ctypes.byref(c),
                               ctypes.byref(cf), ctypes.c_size_t(arena), _p(img), _p(inv))
        return img, inv

    def load_ply(self, path):
        """The reference's loadPly (ply_loader.cpp:32-189)."""
        n, mc, hf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        fn = self.lib.ref_load_ply
        if fn(path.encode(), ctypes.byref(n), ctypes.byref(mc), ctypes.byref(hf), None, None, None, None, None,
              None) != 0:
            raise ValueError("loadPly failed: " + path)
        N, K = n.value, mc.value
        f32 = np.float32
        o = dict(pos=np.zeros((N, 3), f32), scale=np.zeros((N, 3), f32), rot=np.zeros((N, 4), f32),
                 sh=np.zeros((N, K, 3), f32), op=np.zeros(N, f32), filter3d=np.zeros(N, f32) if hf.value else None)
        fn(path.encode(), ctypes.byref(n), ctypes.byref(mc), ctypes.byref(hf), _p(o["pos"]), _p(o["scale"]),
           _p(o["rot"]), _p(o["sh"]), _p(o["op"]), _p(o["filter3d"]))
        o.update(n=N, max_coeffs=K)
        return o
