"""ctypes binding of libgs_hip.so (the C ABI in include/gs_hip.h).

Device memory and streams come from PyTorch (plumbing only); every computation
runs in the library's HIP kernels.  The library is required: if it has not been
built (``make -C aaags-claude_amd`` or ``__graft_entry__.build()``) importing
:class:`GsHip` raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import torch

from gs_scenes import GsCamera, GsRenderCfg

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgs_hip.so")

_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_u32 = ctypes.c_uint32


class _Gaussians(ctypes.Structure):
    _fields_ = [("count", _i32), ("sh_degree", _i32), ("max_coeffs", _i32), ("raw", _i32), ("positions", _vp),
                ("scales", _vp), ("rotations", _vp), ("sh_coeffs", _vp), ("opacities", _vp)]


class _Pre(ctypes.Structure):
    _fields_ = [("means2D", _vp), ("depths", _vp), ("conics", _vp), ("opacities", _vp), ("rgb", _vp),
                ("radii", _vp), ("tiles_touched", _vp)]


class _Pre3D(ctypes.Structure):
    _fields_ = [("gauss2screen", _vp), ("cov3D_inv", _vp), ("mean_offset", _vp)]


class _RGrads(ctypes.Structure):
    _fields_ = [("d_means2D", _vp), ("d_conics", _vp), ("d_rgb", _vp), ("d_opacities_2d", _vp)]


class _PGrads(ctypes.Structure):
    _fields_ = [("d_positions", _vp), ("d_scales", _vp), ("d_rotations", _vp), ("d_sh_coeffs", _vp),
                ("d_opacities", _vp)]


class _PixelState(ctypes.Structure):
    _fields_ = [("T_final", _vp), ("n_contrib", _vp), ("last_contrib", _vp)]


class _AdamHP(ctypes.Structure):
    _fields_ = [("lr", ctypes.c_float * 5), ("beta1", ctypes.c_float), ("beta2", ctypes.c_float),
                ("eps", ctypes.c_float), ("bc1", ctypes.c_float), ("bc2", ctypes.c_float)]


class _Tuning(ctypes.Structure):
    _fields_ = [("radix_bits", _i32), ("raster_batch", _i32), ("raster_bwd_batch", _i32), ("bin_group", _i32),
                ("bin_tiles_per_lane", _i32), ("persistent_per_cu", _i32), ("binning_ms", ctypes.c_float),
                ("raster_fwd_ms", ctypes.c_float), ("raster_bwd_ms", ctypes.c_float), ("split_groups", _i32),
                ("forward_ms", ctypes.c_float)]

    def as_dict(self):
        return {k: (round(getattr(self, k), 4) if isinstance(getattr(self, k), float) else getattr(self, k))
                for k, _ in self._fields_}


class _Peaks(ctypes.Structure):
    _fields_ = [("compute_units", _i32), ("clock_mhz", ctypes.c_double), ("hbm_copy_gbs", ctypes.c_double),
                ("lds_read_gbs", ctypes.c_double), ("valu_wave_insts_per_s", ctypes.c_double),
                ("exp_wave_insts_per_s", ctypes.c_double)]


class GsError(RuntimeError):
    pass


_LIB = None


def load_library(path: str = LIB_PATH):
    """Loads libgs_hip.so once; raises if it is missing (no fallback path)."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(path):
            raise GsError(f"{path} is not built; run `make -C {HERE}` or __graft_entry__.build()")
        _LIB = ctypes.CDLL(path)
        _LIB.gs_last_error.restype = ctypes.c_char_p
        _LIB.gs_lr_schedule.restype = ctypes.c_float
        _LIB.gs_lr_schedule.argtypes = [ctypes.c_float, ctypes.c_float, _i32, _i32]
        _LIB.gs_adam_bias_corrections.argtypes = [ctypes.c_float, ctypes.c_float, _i32, _vp, _vp]
        _LIB.gs_zero.argtypes = [_vp, _vp, ctypes.c_size_t]
        _LIB.gs_ctx_reserve.argtypes = [_vp, ctypes.c_size_t]
        _LIB.gs_radix_sort_kv.argtypes = [_vp, _u32, _vp, _vp, _vp, _vp, _i32, _i32]
        _LIB.gs_tile_ranges.argtypes = [_vp, _u32, _vp, _i32, _vp]
        _LIB.gs_sort_pairs.argtypes = [_vp, _u32, _vp, _vp, _vp, _vp, _i32, _vp]
        _LIB.gs_duplicate_keys.argtypes = [_vp, _i32, _vp, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _vp, _vp, _u32]
        _LIB.gs_scan_tiles.argtypes = [_vp, _i32, _vp, _vp, _vp, _vp]
        _LIB.gs_bin_sorted_words.argtypes = [_vp, _i32, _vp, _i32, _i32, _vp, _u32, _vp, _vp]
        _LIB.gs_bin_tiles.argtypes = [_vp, _i32, _vp, _i32, _i32, _vp, _u32, _vp, _vp]
        _LIB.gs_ctx_set_binning.argtypes = [_vp, _i32, _i32]
    return _LIB


_PLACEHOLDER = {}


def ptr(t) -> int | None:
    """Device address of a tensor for the C ABI.  The ABI takes no null array
    pointers, even for an empty scene or view, but torch reports address 0
    for every tensor without elements; those get a 256-byte allocation of
    their device (never read or written through a count of zero)."""
    if t is None:
        return None
    if t.numel() == 0:
        if t.device not in _PLACEHOLDER:
            _PLACEHOLDER[t.device] = torch.empty(64, dtype=torch.float32, device=t.device)
        return _PLACEHOLDER[t.device].data_ptr()
    return t.data_ptr()


def _f32(x, dev):
    if isinstance(x, torch.Tensor):
        return x.to(device=dev, dtype=torch.float32).contiguous()
    return torch.as_tensor(np.ascontiguousarray(x, dtype=np.float32), device=dev)


class GsHip:
    """One library context bound to a device and to torch's current stream."""

    def __init__(self, device: int = 0, radix_bits: int = 8, raster_batch: int = 256, raster_bwd_batch: int = 256):
        self.lib = load_library()
        self.dev = torch.device("cuda", device)
        self.ctx = _vp()
        stream = torch.cuda.current_stream(self.dev).cuda_stream
        self._check(self.lib.gs_ctx_create(_i32(device), _vp(stream), ctypes.byref(self.ctx)), "gs_ctx_create")
        self.set_tuning(radix_bits, raster_batch, raster_bwd_batch)

    def __del__(self):
        try:
            if self.ctx:
                self.lib.gs_ctx_destroy(self.ctx)
        except Exception:
            pass

    def _check(self, st, what):
        if st != 0:
            raise GsError(f"{what} failed ({st}): {self.lib.gs_last_error().decode()}")

    def set_tuning(self, radix_bits=0, raster_batch=0, raster_bwd_batch=0):
        self._check(self.lib.gs_ctx_set_tuning(self.ctx, _i32(radix_bits), _i32(raster_batch), _i32(raster_bwd_batch)),
                    "gs_ctx_set_tuning")

    def tuning(self) -> dict:
        t = _Tuning()
        self._check(self.lib.gs_ctx_get_tuning(self.ctx, ctypes.byref(t)), "gs_ctx_get_tuning")
        return t.as_dict()

    def apply_tuning(self, d: dict):
        t = _Tuning(**{k: d[k] for k, _ in _Tuning._fields_ if k in d})
        self._check(self.lib.gs_ctx_apply_tuning(self.ctx, ctypes.byref(t)), "gs_ctx_apply_tuning")

    def set_split(self, groups):
        """Trainer forward: depth groups placed before the first raster round (0: one round)."""
        d = self.tuning()
        d["split_groups"] = int(groups)
        self.apply_tuning(d)

    def set_binning(self, group=0, tiles_per_lane=0):
        """Direct-placement binning knobs (0 keeps the current value)."""
        self._check(self.lib.gs_ctx_set_binning(self.ctx, _i32(group), _i32(tiles_per_lane)), "gs_ctx_set_binning")

    def sync(self):
        self._check(self.lib.gs_ctx_synchronize(self.ctx), "gs_ctx_synchronize")

    # ------------------------------------------------------------ tensors --
    def upload_scene(self, scene) -> dict:
        d = self.dev
        return dict(pos=_f32(scene.pos, d), scale=_f32(scene.log_scale, d), rot=_f32(scene.rot, d),
                    sh=_f32(scene.sh, d), op=_f32(scene.opacity_logit, d), sh_degree=scene.sh_degree,
                    max_coeffs=scene.max_coeffs, n=scene.n)

    def _gauss(self, p, raw=True) -> _Gaussians:
        return _Gaussians(p["n"], p["sh_degree"], p["max_coeffs"], int(raw), ptr(p["pos"]), ptr(p["scale"]),
                          ptr(p["rot"]), ptr(p["sh"]), ptr(p["op"]))

    def alloc_pre(self, n):
        d = self.dev
        return dict(means2D=torch.empty((n, 2), device=d), depths=torch.empty(n, device=d),
                    conics=torch.empty((n, 3), device=d), op2d=torch.empty(n, device=d),
                    rgb=torch.empty((n, 3), device=d), radii=torch.empty(n, dtype=torch.int32, device=d),
                    tiles=torch.empty(n, dtype=torch.int32, device=d))

    @staticmethod
    def _pre_struct(pre) -> _Pre:
        return _Pre(ptr(pre["means2D"]), ptr(pre.get("depths")), ptr(pre["conics"]), ptr(pre["op2d"]),
                    ptr(pre["rgb"]), ptr(pre.get("radii")), ptr(pre.get("tiles")))

    # ------------------------------------------------------------- stages --
    def activate(self, p):
        n, k = p["n"], p["max_coeffs"]
        d = self.dev
        out = dict(pos=torch.empty((n, 3), device=d), scale=torch.empty((n, 3), device=d),
                   rot=torch.empty((n, 4), device=d), sh=torch.empty((n, k, 3), device=d),
                   op=torch.empty(n, device=d))
        g = self._gauss(p, raw=True)
        self._check(self.lib.gs_activate(self.ctx, ctypes.byref(g), _vp(ptr(out["pos"])), _vp(ptr(out["scale"])),
                                         _vp(ptr(out["rot"])), _vp(ptr(out["sh"])), _vp(ptr(out["op"]))),
                    "gs_activate")
        return out

    def preprocess(self, p, cam, cfg, raw=True, pre=None):
        pre = pre if pre is not None else self.alloc_pre(p["n"])
        g = self._gauss(p, raw)
        c, f = cam.to_c(), cfg.to_c()
        ps = self._pre_struct(pre)
        self._check(self.lib.gs_preprocess_fwd(self.ctx, ctypes.byref(g), ctypes.byref(c), ctypes.byref(f),
                                               ctypes.byref(ps)), "gs_preprocess_fwd")
        return pre

    def scan_tiles(self, tiles):
        n = tiles.numel()
        offsets = torch.empty(max(n, 1), dtype=torch.int32, device=self.dev)
        total_dev = torch.zeros(1, dtype=torch.int32, device=self.dev)
        total = ctypes.c_longlong()
        self._check(self.lib.gs_scan_tiles(self.ctx, n, ptr(tiles), ptr(offsets), ptr(total_dev), ctypes.byref(total)),
                    "gs_scan_tiles")
        return offsets[:n], int(total.value)

    def duplicate(self, pre, offsets, K, width, height, capacity=None):
        n = pre["radii"].numel()
        cap = K if capacity is None else capacity
        keys = torch.empty(max(cap, 1), dtype=torch.int64, device=self.dev)
        vals = torch.empty(max(cap, 1), dtype=torch.int32, device=self.dev)
        self._check(self.lib.gs_duplicate_keys(self.ctx, n, ptr(pre["means2D"]), ptr(pre["depths"]), ptr(pre["radii"]),
                                               ptr(offsets), width, height, 16, 16, ptr(keys), ptr(vals), cap),
                    "gs_duplicate_keys")
        return keys[:K], vals[:K]

    def check_overflow(self) -> bool:
        o = ctypes.c_int()
        self._check(self.lib.gs_check_overflow(self.ctx, ctypes.byref(o)), "gs_check_overflow")
        return bool(o.value)

    def radix_sort(self, keys, vals, key_bits=64, radix_bits=0):
        n = keys.numel()
        ko = torch.empty_like(keys)
        vo = torch.empty_like(vals)
        if n:
            self._check(self.lib.gs_radix_sort_kv(self.ctx, n, ptr(keys), ptr(vals), ptr(ko), ptr(vo), key_bits,
                                                  radix_bits), "gs_radix_sort_kv")
        return ko, vo

    def sort_pairs(self, keys, vals, num_tiles):
        n = keys.numel()
        ko = torch.empty(max(n, 1), dtype=torch.int64, device=self.dev)
        vo = torch.empty(max(n, 1), dtype=torch.int32, device=self.dev)
        ranges = torch.empty(num_tiles * 2, dtype=torch.int32, device=self.dev)
        self._check(self.lib.gs_sort_pairs(self.ctx, n, ptr(keys) if n else None, ptr(vals) if n else None, ptr(ko),
                                           ptr(vo), num_tiles, ptr(ranges)), "gs_sort_pairs")
        return ko[:n], vo[:n], ranges

    def tile_ranges(self, keys_sorted, num_tiles):
        ranges = torch.empty(num_tiles * 2, dtype=torch.int32, device=self.dev)
        n = keys_sorted.numel()
        self._check(self.lib.gs_tile_ranges(self.ctx, n, ptr(keys_sorted) if n else None, num_tiles, ptr(ranges)),
                    "gs_tile_ranges")
        return ranges

    def bin_sorted_words(self, pre, width, height, capacity=None):
        """Depth-ordered binning: (words[K] = tile << 32 | index, ranges, K)."""
        n = pre["radii"].numel()
        num_tiles = ((width + 15) // 16) * ((height + 15) // 16)
        ranges = torch.empty(num_tiles * 2, dtype=torch.int32, device=self.dev)
        K = ctypes.c_longlong()
        if capacity is None:
            # an upper bound: run once to learn K when it is unknown
            st = self.lib.gs_bin_sorted_words(self.ctx, n, ctypes.byref(self._pre_struct(pre)), width, height, None,
                                              _u32(0), _vp(ptr(ranges)), ctypes.byref(K))
            if st not in (0, 3):
                self._check(st, "gs_bin_sorted_words")
            capacity = max(int(K.value), 1)
        words = torch.empty(capacity, dtype=torch.int64, device=self.dev)
        self._check(self.lib.gs_bin_sorted_words(self.ctx, n, ctypes.byref(self._pre_struct(pre)), width, height,
                                                 _vp(ptr(words)), _u32(capacity), _vp(ptr(ranges)), ctypes.byref(K)),
                    "gs_bin_sorted_words")
        return words[:K.value], ranges, int(K.value)

    def bin_tiles(self, pre, width, height, capacity=None):
        """Direct-placement binning: (values[K] sorted per tile, ranges, K)."""
        n = pre["radii"].numel()
        num_tiles = ((width + 15) // 16) * ((height + 15) // 16)
        ranges = torch.empty(num_tiles * 2, dtype=torch.int32, device=self.dev)
        K = ctypes.c_longlong()
        if capacity is None:
            st = self.lib.gs_bin_tiles(self.ctx, n, ctypes.byref(self._pre_struct(pre)), width, height, None, _u32(0),
                                       _vp(ptr(ranges)), ctypes.byref(K))
            if st not in (0, 3):
                self._check(st, "gs_bin_tiles")
            capacity = max(int(K.value), 1)
        values = torch.empty(capacity, dtype=torch.int32, device=self.dev)
        self._check(self.lib.gs_bin_tiles(self.ctx, n, ctypes.byref(self._pre_struct(pre)), width, height,
                                          _vp(ptr(values)), _u32(capacity), _vp(ptr(ranges)), ctypes.byref(K)),
                    "gs_bin_tiles")
        return values[:K.value], ranges, int(K.value)

    def raster_fwd_words(self, pre, words, ranges, width, height, bg=(0.0, 0.0, 0.0)):
        d = self.dev
        img = torch.empty((height, width, 3), device=d)
        px = dict(T_final=torch.empty((height, width), device=d),
                  n_contrib=torch.empty((height, width), dtype=torch.int32, device=d),
                  last_contrib=torch.empty((height, width), dtype=torch.int32, device=d))
        st = _PixelState(ptr(px["T_final"]), ptr(px["n_contrib"]), ptr(px["last_contrib"]))
        bgv = (ctypes.c_float * 3)(*bg)
        w = words if words.numel() else torch.zeros(1, dtype=torch.int64, device=d)
        self._check(self.lib.gs_raster_fwd_words(self.ctx, width, height, ctypes.byref(self._pre_struct(pre)),
                                                 _vp(ptr(w)), _vp(ptr(ranges)), bgv, _vp(ptr(img)), ctypes.byref(st)),
                    "gs_raster_fwd_words")
        return img, px

    def raster_bwd_words(self, pre, words, ranges, width, height, px, d_image, bg=(0.0, 0.0, 0.0), rg=None):
        n = pre["op2d"].numel()
        rg = rg if rg is not None else self.alloc_rgrads(n)
        st = _PixelState(ptr(px["T_final"]), ptr(px.get("n_contrib")), ptr(px["last_contrib"]))
        gs = _RGrads(ptr(rg["d_means2D"]), ptr(rg["d_conics"]), ptr(rg["d_rgb"]), ptr(rg["d_op2d"]))
        bgv = (ctypes.c_float * 3)(*bg)
        w = words if words.numel() else torch.zeros(1, dtype=torch.int64, device=self.dev)
        self._check(self.lib.gs_raster_bwd_words(self.ctx, width, height, ctypes.byref(self._pre_struct(pre)),
                                                 _vp(ptr(w)), _vp(ptr(ranges)), bgv, ctypes.byref(st),
                                                 _vp(ptr(d_image)), ctypes.byref(gs)), "gs_raster_bwd_words")
        return rg

    def raster_fwd(self, pre, values_sorted, ranges, width, height, bg=(0.0, 0.0, 0.0)):
        d = self.dev
        img = torch.empty((height, width, 3), device=d)
        px = dict(T_final=torch.empty((height, width), device=d),
                  n_contrib=torch.empty((height, width), dtype=torch.int32, device=d),
                  last_contrib=torch.empty((height, width), dtype=torch.int32, device=d))
        ps = self._pre_struct(pre)
        st = _PixelState(ptr(px["T_final"]), ptr(px["n_contrib"]), ptr(px["last_contrib"]))
        bgv = (ctypes.c_float * 3)(*bg)
        vs = values_sorted if values_sorted.numel() else torch.zeros(1, dtype=torch.int32, device=d)
        self._check(self.lib.gs_raster_fwd(self.ctx, width, height, ctypes.byref(ps), _vp(ptr(vs)), _vp(ptr(ranges)),
                                           bgv, _vp(ptr(img)), ctypes.byref(st)), "gs_raster_fwd")
        return img, px

    # ------------------------------------------------ eval_3D forward --
    def alloc_pre3d(self, n):
        d = self.dev
        return dict(g2s=torch.empty((n, 16), device=d), cov3D_inv=torch.empty((n, 6), device=d),
                    mean_offset=torch.empty((n, 3), device=d))

    @staticmethod
    def _pre3d_struct(p3) -> _Pre3D:
        return _Pre3D(ptr(p3["g2s"]), ptr(p3["cov3D_inv"]), ptr(p3["mean_offset"]))

    def preprocess_3d(self, p, cam, cfg, raw=True, filter3d=None, pre=None, pre3d=None):
        pre = pre if pre is not None else self.alloc_pre(p["n"])
        pre3d = pre3d if pre3d is not None else self.alloc_pre3d(p["n"])
        g = self._gauss(p, raw)
        c, f = cam.to_c(), cfg.to_c()
        ps, p3 = self._pre_struct(pre), self._pre3d_struct(pre3d)
        self._check(self.lib.gs_preprocess_fwd_3d(self.ctx, ctypes.byref(g), _vp(ptr(filter3d)), ctypes.byref(c),
                                                  ctypes.byref(f), ctypes.byref(ps), ctypes.byref(p3)),
                    "gs_preprocess_fwd_3d")
        return pre, pre3d

    def duplicate_3d(self, pre, pre3d, offsets, K, cam, capacity=None):
        n = pre["radii"].numel()
        cap = K if capacity is None else capacity
        keys = torch.empty(max(cap, 1), dtype=torch.int64, device=self.dev)
        vals = torch.empty(max(cap, 1), dtype=torch.int32, device=self.dev)
        ps, p3, c = self._pre_struct(pre), self._pre3d_struct(pre3d), cam.to_c()
        self._check(self.lib.gs_duplicate_keys_3d(self.ctx, n, ctypes.byref(ps), ctypes.byref(p3), _vp(ptr(offsets)),
                                                  ctypes.byref(c), 16, 16, _vp(ptr(keys)), _vp(ptr(vals)),
                                                  ctypes.c_uint32(cap)), "gs_duplicate_keys_3d")
        return keys[:K], vals[:K]

    def raster_fwd_3d(self, pre, pre3d, values_sorted, ranges, cam, bg=(0.0, 0.0, 0.0), img=None, T_final=None,
                      n_contrib=None):
        d = self.dev
        H, W = cam.height, cam.width
        img = img if img is not None else torch.empty((H, W, 3), device=d)
        T_final = T_final if T_final is not None else torch.empty((H, W), device=d)
        n_contrib = n_contrib if n_contrib is not None else torch.empty((H, W), dtype=torch.int32, device=d)
        ps, p3, c = self._pre_struct(pre), self._pre3d_struct(pre3d), cam.to_c()
        bgv = (ctypes.c_float * 3)(*bg)
        vs = values_sorted if values_sorted.numel() else torch.zeros(1, dtype=torch.int32, device=d)
        self._check(self.lib.gs_raster_fwd_3d(self.ctx, ctypes.byref(c), ctypes.byref(ps), ctypes.byref(p3),
                                              _vp(ptr(vs)), _vp(ptr(ranges)), bgv, _vp(ptr(img)), _vp(ptr(T_final)),
                                              _vp(ptr(n_contrib))), "gs_raster_fwd_3d")
        return img, T_final, n_contrib

    def bin_tiles_3d(self, pre, pre3d, cam, capacity=None, keys=True, out=None):
        """gs_bin_tiles_3d: sorted pair values, keys (optional) and tile ranges
        of the eval_3D binning in one call; returns (values, keys, ranges, K).
        out: (values, keys, ranges) buffers to reuse (capacity from values)."""
        n = pre["radii"].numel()
        gx, gy = (cam.width + 15) // 16, (cam.height + 15) // 16
        if out is None:
            cap = capacity if capacity is not None else int(pre["tiles"].to(torch.int64).sum().item())
            vo = torch.empty(max(cap, 4), dtype=torch.int32, device=self.dev)
            ko = torch.empty(max(cap, 1), dtype=torch.int64, device=self.dev) if keys else None
            ranges = torch.empty(gx * gy * 2, dtype=torch.int32, device=self.dev)
        else:
            vo, ko, ranges = out
            cap = vo.numel()
        ps, p3, c = self._pre_struct(pre), self._pre3d_struct(pre3d), cam.to_c()
        K = ctypes.c_longlong()
        self._check(self.lib.gs_bin_tiles_3d(self.ctx, n, ctypes.byref(ps), ctypes.byref(p3), ctypes.byref(c),
                                             _vp(ptr(vo)), _vp(ptr(ko)), ctypes.c_uint32(cap), _vp(ptr(ranges)),
                                             ctypes.byref(K)), "gs_bin_tiles_3d")
        k = K.value
        return vo[:k], (ko[:k] if ko is not None else None), ranges, k

    def render_3d(self, p, cam, cfg, raw=True, filter3d=None, fast=False):
        """The eval_3D forward stage by stage: preprocess, scan (K read back),
        per-tile depth keys, stable pair sort, k-buffer raster.  fast: the
        binning (keys, sort, ranges) by gs_bin_tiles_3d instead."""
        pre, pre3d = self.preprocess_3d(p, cam, cfg, raw, filter3d)
        if fast:
            vs, ks, ranges, K = self.bin_tiles_3d(pre, pre3d, cam)
            keys = vals = None
        else:
            offsets, K = self.scan_tiles(pre["tiles"])
            keys, vals = self.duplicate_3d(pre, pre3d, offsets, K, cam)
            gx, gy = (cam.width + 15) // 16, (cam.height + 15) // 16
            ks, vs, ranges = self.sort_pairs(keys, vals, gx * gy)
        img, T, nc = self.raster_fwd_3d(pre, pre3d, vs, ranges, cam, tuple(cfg.bg_color))
        return dict(pre=pre, pre3d=pre3d, keys=keys, values=vals, keys_sorted=ks, values_sorted=vs, ranges=ranges,
                    img=img, T_final=T, n_contrib=nc, K=K)

    def raster_inv_depth(self, pre, values_sorted, ranges, width, height, px):
        """Inverse-depth output of the 2D forward, from its pixel state."""
        d = self.dev
        out = torch.empty((height, width), device=d)
        ps = self._pre_struct(pre)
        st = _PixelState(ptr(px["T_final"]), ptr(px["n_contrib"]), ptr(px["last_contrib"]))
        vs = values_sorted if values_sorted.numel() else torch.zeros(1, dtype=torch.int32, device=d)
        self._check(self.lib.gs_raster_inv_depth(self.ctx, width, height, ctypes.byref(ps), _vp(ptr(vs)),
                                                 _vp(ptr(ranges)), ctypes.byref(st), _vp(ptr(out))),
                    "gs_raster_inv_depth")
        return out

    def raster_work(self, pre, values_sorted, ranges, width, height, px):
        """gs_raster_work: the reference loops' pixel-Gaussian evaluations on
        this frame: dict(fwd_evals, bwd_evals, contributors)."""
        d = self.dev
        out = torch.zeros(3, dtype=torch.int64, device=d)
        ps = self._pre_struct(pre)
        st = _PixelState(ptr(px["T_final"]), ptr(px["n_contrib"]), ptr(px["last_contrib"]))
        vs = values_sorted if values_sorted.numel() else torch.zeros(1, dtype=torch.int32, device=d)
        self._check(self.lib.gs_raster_work(self.ctx, width, height, ctypes.byref(ps), _vp(ptr(vs)),
                                            _vp(ptr(ranges)), ctypes.byref(st), _vp(ptr(out))), "gs_raster_work")
        c = [int(x) for x in out.cpu().tolist()]
        return dict(fwd_evals=c[0], bwd_evals=c[1], contributors=c[2])

    def device_peaks(self) -> dict:
        """gs_device_peaks: measured HBM copy, LDS read and VALU issue ceilings."""
        p = _Peaks()
        self._check(self.lib.gs_device_peaks(self.ctx, ctypes.byref(p)), "gs_device_peaks")
        return {k: (round(getattr(p, k), 1) if isinstance(getattr(p, k), float) else getattr(p, k))
                for k, _ in p._fields_}

    # ---------------------------------------------------------- PLY input --
    def ply_info(self, path):
        n, deg, mc, hf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.gs_ply_info(path.encode(), ctypes.byref(n), ctypes.byref(deg), ctypes.byref(mc),
                                         ctypes.byref(hf)), "gs_ply_info")
        return dict(n=n.value, sh_degree=deg.value, max_coeffs=mc.value, has_filter_3d=bool(hf.value))

    def load_ply(self, path):
        """loadPly into device tensors (activated, SH interleaved)."""
        info = self.ply_info(path)
        n, k, d = info["n"], info["max_coeffs"], self.dev
        out = dict(pos=torch.empty((n, 3), device=d), scale=torch.empty((n, 3), device=d),
                   rot=torch.empty((n, 4), device=d), sh=torch.empty((n, k, 3), device=d), op=torch.empty(n, device=d),
                   filter3d=torch.empty(n, device=d) if info["has_filter_3d"] else None, **info)
        self._check(self.lib.gs_ply_load(self.ctx, path.encode(), _i32(n), _i32(k), _vp(ptr(out["pos"])),
                                         _vp(ptr(out["scale"])),
                                         _vp(ptr(out["rot"])), _vp(ptr(out["sh"])), _vp(ptr(out["op"])),
                                         _vp(ptr(out["filter3d"]))), "gs_ply_load")
        return out

    def load_ply_raw(self, path):
        """loadPly + initRawFromGaussianData into device tensors: the raw
        parameters a training run starts from (upload_scene's layout)."""
        info = self.ply_info(path)
        n, k, d = info["n"], info["max_coeffs"], self.dev
        out = dict(pos=torch.empty((n, 3), device=d), scale=torch.empty((n, 3), device=d),
                   rot=torch.empty((n, 4), device=d), sh=torch.empty((n, k, 3), device=d), op=torch.empty(n, device=d),
                   n=n, max_coeffs=k, sh_degree=info["sh_degree"])
        self._check(self.lib.gs_ply_load_raw(self.ctx, path.encode(), _i32(n), _i32(k), _vp(ptr(out["pos"])),
                                             _vp(ptr(out["scale"])), _vp(ptr(out["rot"])), _vp(ptr(out["sh"])),
                                             _vp(ptr(out["op"]))), "gs_ply_load_raw")
        return out

    def save_ply(self, path, p):
        """savePly of raw device parameters (upload_scene's dict)."""
        self._check(self.lib.gs_ply_save(self.ctx, path.encode(), p["n"], p["max_coeffs"], _vp(ptr(p["pos"])),
                                         _vp(ptr(p["scale"])), _vp(ptr(p["rot"])), _vp(ptr(p["sh"])),
                                         _vp(ptr(p["op"]))), "gs_ply_save")

    def l1_loss(self, img, gt):
        H, W = img.shape[0], img.shape[1]
        d_img = torch.empty_like(img)
        loss = torch.zeros(1, dtype=torch.float64, device=self.dev)
        self._check(self.lib.gs_l1_loss(self.ctx, W, H, _vp(ptr(img)), _vp(ptr(gt)), _vp(ptr(d_img)), _vp(ptr(loss))),
                    "gs_l1_loss")
        return loss, d_img

    def alloc_rgrads(self, n):
        d = self.dev
        return dict(d_means2D=torch.zeros((n, 2), device=d), d_conics=torch.zeros((n, 3), device=d),
                    d_rgb=torch.zeros((n, 3), device=d), d_op2d=torch.zeros(n, device=d))

    def raster_bwd(self, pre, values_sorted, ranges, width, height, px, d_image, bg=(0.0, 0.0, 0.0), rg=None):
        n = pre["radii"].numel() if "radii" in pre else pre["op2d"].numel()
        rg = rg if rg is not None else self.alloc_rgrads(n)
        ps = self._pre_struct(pre)
        st = _PixelState(ptr(px["T_final"]), ptr(px.get("n_contrib")), ptr(px["last_contrib"]))
        gs = _RGrads(ptr(rg["d_means2D"]), ptr(rg["d_conics"]), ptr(rg["d_rgb"]), ptr(rg["d_op2d"]))
        bgv = (ctypes.c_float * 3)(*bg)
        vs = values_sorted if values_sorted.numel() else torch.zeros(1, dtype=torch.int32, device=self.dev)
        self._check(self.lib.gs_raster_bwd(self.ctx, width, height, ctypes.byref(ps), _vp(ptr(vs)), _vp(ptr(ranges)),
                                           bgv, ctypes.byref(st), _vp(ptr(d_image)), ctypes.byref(gs)),
                    "gs_raster_bwd")
        return rg

    def alloc_pgrads(self, n, k):
        d = self.dev
        return dict(pos=torch.zeros((n, 3), device=d), scale=torch.zeros((n, 3), device=d),
                    rot=torch.zeros((n, 4), device=d), sh=torch.zeros((n, k, 3), device=d), op=torch.zeros(n, device=d))

    def preprocess_bwd(self, p, cam, cfg, radii, rg, out=None, accumulate=False):
        out = out if out is not None else self.alloc_pgrads(p["n"], p["max_coeffs"])
        g = self._gauss(p, raw=True)
        c, f = cam.to_c(), cfg.to_c()
        rgs = _RGrads(ptr(rg["d_means2D"]), ptr(rg["d_conics"]), ptr(rg["d_rgb"]), ptr(rg["d_op2d"]))
        pgs = _PGrads(ptr(out["pos"]), ptr(out["scale"]), ptr(out["rot"]), ptr(out["sh"]), ptr(out["op"]))
        self._check(self.lib.gs_preprocess_bwd(self.ctx, ctypes.byref(g), ctypes.byref(c), ctypes.byref(f),
                                               _vp(ptr(radii)), ctypes.byref(rgs), ctypes.byref(pgs),
                                               int(accumulate)), "gs_preprocess_bwd")
        return out

    def adam(self, p, grads, m, v, lrs, beta1, beta2, eps, step):
        bc1, bc2 = ctypes.c_float(), ctypes.c_float()
        self.lib.gs_adam_bias_corrections(ctypes.c_float(beta1), ctypes.c_float(beta2), step, ctypes.byref(bc1),
                                          ctypes.byref(bc2))
        hp = _AdamHP((ctypes.c_float * 5)(*lrs), beta1, beta2, eps, bc1.value, bc2.value)
        names = ["pos", "scale", "rot", "sh", "op"]
        P = (_vp * 5)(*[ptr(p[k]) for k in names])
        G = (_vp * 5)(*[ptr(grads[k]) for k in names])
        M = (_vp * 5)(*[ptr(m[k]) for k in names])
        V = (_vp * 5)(*[ptr(v[k]) for k in names])
        self._check(self.lib.gs_adam_step(self.ctx, p["n"], p["max_coeffs"], P, G, M, V, ctypes.byref(hp)),
                    "gs_adam_step")

    def lr_schedule(self, lr_init, lr_final, step, max_steps):
        return self.lib.gs_lr_schedule(lr_init, lr_final, step, max_steps)

    # -------------------------------------------------------- whole paths --
    def forward(self, p, cam, cfg):
        """Plain single-view forward (host readback of the pair count)."""
        pre = self.preprocess(p, cam, cfg)
        offsets, K = self.scan_tiles(pre["tiles"])
        keys, vals = self.duplicate(pre, offsets, K, cam.width, cam.height)
        ks, vs, ranges = self.sort_pairs(keys, vals, cam.num_tiles)
        img, px = self.raster_fwd(pre, vs, ranges, cam.width, cam.height, cfg.bg_color)
        return dict(pre=pre, K=K, keys=keys, values=vals, keys_sorted=ks, values_sorted=vs, ranges=ranges, img=img,
                    px=px)

    def forward_fast(self, p, cam, cfg):
        """Single-view forward with direct-placement binning (the trainer's path)."""
        pre = self.preprocess(p, cam, cfg)
        vals, ranges, K = self.bin_tiles(pre, cam.width, cam.height)
        img, px = self.raster_fwd(pre, vals, ranges, cam.width, cam.height, cfg.bg_color)
        return dict(pre=pre, K=K, values_sorted=vals, ranges=ranges, img=img, px=px)

    def forward_words(self, p, cam, cfg):
        """Single-view forward with the depth-ordered word sort."""
        pre = self.preprocess(p, cam, cfg)
        words, ranges, K = self.bin_sorted_words(pre, cam.width, cam.height)
        img, px = self.raster_fwd_words(pre, words, ranges, cam.width, cam.height, cfg.bg_color)
        return dict(pre=pre, K=K, words=words, ranges=ranges, img=img, px=px)

    def forward_backward_fast(self, p, cam, cfg, gt):
        f = self.forward_fast(p, cam, cfg)
        loss, d_img = self.l1_loss(f["img"], gt)
        rg = self.raster_bwd(f["pre"], f["values_sorted"], f["ranges"], cam.width, cam.height, f["px"], d_img,
                             cfg.bg_color)
        grads = self.preprocess_bwd(p, cam, cfg, f["pre"]["radii"], rg)
        f.update(loss=loss, d_image=d_img, rgrad=rg, grads=grads)
        return f

    def forward_backward(self, p, cam, cfg, gt):
        f = self.forward(p, cam, cfg)
        loss, d_img = self.l1_loss(f["img"], gt)
        rg = self.raster_bwd(f["pre"], f["values_sorted"], f["ranges"], cam.width, cam.height, f["px"], d_img,
                             cfg.bg_color)
        grads = self.preprocess_bwd(p, cam, cfg, f["pre"]["radii"], rg)
        f.update(loss=loss, d_image=d_img, rgrad=rg, grads=grads)
        return f


class _TrainCfg(ctypes.Structure):
    _fields_ = [("lr_position_init", ctypes.c_float), ("lr_position_final", ctypes.c_float),
                ("lr_feature", ctypes.c_float), ("lr_opacity", ctypes.c_float), ("lr_scaling", ctypes.c_float),
                ("lr_rotation", ctypes.c_float), ("max_steps", _i32), ("use_adam", _i32),
                ("adam_beta1", ctypes.c_float), ("adam_beta2", ctypes.c_float), ("adam_eps", ctypes.c_float)]


class _StepResult(ctypes.Structure):
    _fields_ = [("loss", ctypes.c_double), ("iteration", _i32), ("n_gaussians", _i32), ("pairs", ctypes.c_longlong)]


def train_cfg(lr=None, use_adam=True, max_steps=30000, beta1=0.9, beta2=0.999, eps=1e-8, **lrs) -> _TrainCfg:
    """TrainConfig (train_types.h:46-60).  ``lr`` sets every learning rate
    (position schedule constant); keyword overrides set individual ones."""
    d = dict(lr_position_init=0.00016, lr_position_final=0.0000016, lr_feature=0.0025, lr_opacity=0.025,
             lr_scaling=0.005, lr_rotation=0.001)
    if lr is not None:
        d = {k: lr for k in d}
    d.update(lrs)
    return _TrainCfg(d["lr_position_init"], d["lr_position_final"], d["lr_feature"], d["lr_opacity"],
                     d["lr_scaling"], d["lr_rotation"], max_steps, int(use_adam), beta1, beta2, eps)


class GsTrainer:
    """The native training driver (gs_trainer_*): parameters, gradients and
    optimizer state stay in HBM; each step is one call."""

    def __init__(self, hip: GsHip, scene):
        self.hip = hip
        self.lib = hip.lib
        self.n, self.k = scene.n, scene.max_coeffs
        self.tr = _vp()
        hip._check(self.lib.gs_trainer_create(hip.ctx, scene.n, scene.sh_degree, scene.max_coeffs,
                                              ctypes.byref(self.tr)), "gs_trainer_create")
        self.lib.gs_trainer_buffer.restype = _vp
        arrs = [np.ascontiguousarray(a, np.float32) for a in scene.arrays()]
        hip._check(self.lib.gs_trainer_set_params(self.tr, *[a.ctypes.data_as(_vp) for a in arrs], 1),
                   "gs_trainer_set_params")

    def __del__(self):
        try:
            if self.tr:
                self.lib.gs_trainer_destroy(self.tr)
        except Exception:
            pass

    def set_params(self, scene):
        """Replaces the raw parameters (same Gaussian count and SH size); the
        optimizer state is kept."""
        assert scene.n == self.n and scene.max_coeffs == self.k
        arrs = [np.ascontiguousarray(a, np.float32) for a in scene.arrays()]
        self.hip._check(self.lib.gs_trainer_set_params(self.tr, *[a.ctypes.data_as(_vp) for a in arrs], 1),
                        "gs_trainer_set_params")

    def step(self, cam, cfg, gt, tc, iteration, sync=True):
        """gs_trainer_step; sync=False passes no result (the step stays queued,
        the loss is not read back) and returns None."""
        res = _StepResult()
        c, f = cam.to_c(), cfg.to_c()
        self.hip._check(self.lib.gs_trainer_step(self.tr, ctypes.byref(c), ctypes.byref(f), _vp(ptr(gt)),
                                                 ctypes.byref(tc), iteration, ctypes.byref(res) if sync else None),
                        "gs_trainer_step")
        return res if sync else None

    def set_options(self, fused_adam=False):
        """fused_adam: Adam inside the preprocess backward (gradient buffers not written)."""
        self.hip._check(self.lib.gs_trainer_set_options(self.tr, _i32(int(fused_adam))), "gs_trainer_set_options")

    STAGES = ["preprocess", "binning", "raster_fwd", "l1", "zero_grads", "raster_bwd", "backward"]

    def set_timing(self, stages=None):
        """HIP-event timing of the named stages of each step (None: all; []: off);
        resets the totals."""
        names = self.STAGES if stages is None else stages
        mask = 0
        for nm in names:
            mask |= 1 << self.STAGES.index(nm)
        self.hip._check(self.lib.gs_trainer_set_timing(self.tr, ctypes.c_uint(mask)), "gs_trainer_set_timing")

    def stage_times(self):
        """Per-step mean device time (ms) of each timed stage since set_timing, and the step count."""
        ms = (ctypes.c_double * len(self.STAGES))()
        steps = ctypes.c_longlong()
        self.hip._check(self.lib.gs_trainer_stage_times(self.tr, ms, _i32(len(self.STAGES)), ctypes.byref(steps)),
                        "gs_trainer_stage_times")
        return {nm: float(ms[i]) for i, nm in enumerate(self.STAGES)}, int(steps.value)

    def autotune(self, cam, cfg, gt, reps=3) -> dict:
        """gs_autotune on this trainer's parameters and one view (no parameter update)."""
        c, f = cam.to_c(), cfg.to_c()
        t = _Tuning()
        self.hip._check(self.lib.gs_autotune(self.tr, ctypes.byref(c), ctypes.byref(f), _vp(ptr(gt)), _i32(reps),
                                             ctypes.byref(t)), "gs_autotune")
        return t.as_dict()

    def render(self, cam, cfg):
        k = ctypes.c_longlong()
        c, f = cam.to_c(), cfg.to_c()
        self.hip._check(self.lib.gs_trainer_render(self.tr, ctypes.byref(c), ctypes.byref(f), ctypes.byref(k)),
                        "gs_trainer_render")
        return int(k.value)

    def buffer(self, name, shape, dtype=torch.float32):
        """Wraps a trainer-owned device buffer as a torch tensor (a copy)."""
        p = self.lib.gs_trainer_buffer(self.tr, name.encode())
        if not p:
            raise GsError(f"no trainer buffer {name}")
        numel = int(np.prod(shape))
        out = torch.empty(numel, dtype=dtype, device=self.hip.dev)
        torch.cuda.synchronize(self.hip.dev)
        hipMemcpy = ctypes.CDLL("libamdhip64.so").hipMemcpy
        hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, _i32]
        st = hipMemcpy(_vp(out.data_ptr()), _vp(p), numel * out.element_size(), 3)  # device to device
        if st != 0:
            raise GsError(f"hipMemcpy failed ({st})")
        return out.view(*shape)

    def train_views(self, cams, cfg, targets, tc, iterations, seed=42, save_every=0, output=None,
                    first_iteration=0, log_every=0):
        """gs_trainer_train_views: the multi-view training loop (one random view
        per iteration, checkpoints); returns (losses, views)."""
        V = len(cams)
        carr = (GsCamera * V)(*[c.to_c() for c in cams])
        tg = [t if isinstance(t, torch.Tensor) else torch.tensor(np.asarray(t, np.float32), device=self.hip.dev)
              for t in targets]
        ptrs = (_vp * V)(*[_vp(ptr(t)) for t in tg])
        losses = np.zeros(max(iterations, 1), np.float64)
        views = np.zeros(max(iterations, 1), np.int32)
        f = cfg.to_c()
        self.lib.gs_trainer_train_views.argtypes = [_vp, _i32, _vp, _vp, _vp, _vp, _i32, _i32, ctypes.c_uint, _i32,
                                                    ctypes.c_char_p, _i32, _vp, _vp]
        self.hip._check(self.lib.gs_trainer_train_views(self.tr, V, carr, ptrs, ctypes.byref(f), ctypes.byref(tc),
                                                        first_iteration, iterations, seed, save_every,
                                                        output.encode() if output else None, log_every,
                                                        losses.ctypes.data_as(_vp), views.ctypes.data_as(_vp)),
                        "gs_trainer_train_views")
        return losses[:iterations], views[:iterations]

    def count(self):
        n = _i32()
        self.hip._check(self.lib.gs_trainer_count(self.tr, ctypes.byref(n)), "gs_trainer_count")
        return n.value

    def params(self):
        out = [np.empty((self.n, 3), np.float32), np.empty((self.n, 3), np.float32), np.empty((self.n, 4), np.float32),
               np.empty((self.n, self.k, 3), np.float32), np.empty(self.n, np.float32)]
        self.hip._check(self.lib.gs_trainer_get_params(self.tr, *[a.ctypes.data_as(_vp) for a in out], 1),
                        "gs_trainer_get_params")
        return out


class GsBatch:
    """Batched multi-view training step over a GsTrainer's parameters
    (gs_batch_*).  Views are fixed for the batch's lifetime."""

    def __init__(self, trainer: GsTrainer, cams, cfg, targets, persistent=True, fused_adam=True, use_graph=True):
        self.tr = trainer
        self.hip = trainer.hip
        self.lib = trainer.lib
        self.cams = list(cams)
        self.V = len(self.cams)
        self.b = _vp()
        carr = (GsCamera * self.V)(*[c.to_c() for c in self.cams])
        f = cfg.to_c()
        self.hip._check(self.lib.gs_batch_create(trainer.tr, self.V, carr, ctypes.byref(f), ctypes.byref(self.b)),
                        "gs_batch_create")
        self.lib.gs_batch_buffer.restype = _vp
        self.set_options(persistent, fused_adam, use_graph)
        arrs = [np.ascontiguousarray(t.detach().cpu().numpy() if isinstance(t, torch.Tensor) else t, np.float32)
                for t in targets]
        ptrs = (_vp * self.V)(*[a.ctypes.data_as(_vp) for a in arrs])
        self.hip._check(self.lib.gs_batch_set_targets(self.b, ptrs, 1), "gs_batch_set_targets")
        self.pix_off = [0]
        for c in self.cams:
            self.pix_off.append(self.pix_off[-1] + c.width * c.height)

    def __del__(self):
        try:
            if self.b:
                self.lib.gs_batch_destroy(self.b)
        except Exception:
            pass

    def autotune(self, reps=3) -> dict:
        """gs_batch_autotune on the batch's views and targets (no parameter update)."""
        t = _Tuning()
        self.hip._check(self.lib.gs_batch_autotune(self.b, _i32(reps), ctypes.byref(t)), "gs_batch_autotune")
        return t.as_dict()

    def set_options(self, persistent=True, fused_adam=True, use_graph=True):
        self.hip._check(self.lib.gs_batch_set_options(self.b, int(persistent), int(fused_adam), int(use_graph)),
                        "gs_batch_set_options")

    def step(self, tc, iteration, sync=False):
        res = _StepResult()
        st = self.lib.gs_batch_step(self.b, ctypes.byref(tc), iteration, ctypes.byref(res) if sync else None)
        self.hip._check(st, "gs_batch_step")
        return res if sync else None

    def result(self):
        res = _StepResult()
        losses = np.zeros(self.V, np.float64)
        self.hip._check(self.lib.gs_batch_result(self.b, ctypes.byref(res), losses.ctypes.data_as(_vp)),
                        "gs_batch_result")
        return res, losses

    def render(self):
        k = ctypes.c_longlong()
        self.hip._check(self.lib.gs_batch_render(self.b, ctypes.byref(k)), "gs_batch_render")
        return int(k.value)

    def info(self):
        t, p, c = _i32(), ctypes.c_longlong(), ctypes.c_longlong()
        self.hip._check(self.lib.gs_batch_info(self.b, ctypes.byref(t), ctypes.byref(p), ctypes.byref(c)),
                        "gs_batch_info")
        return int(t.value), int(p.value), int(c.value)

    def _buffer(self, name, numel, dtype):
        p = self.lib.gs_batch_buffer(self.b, name.encode())
        if not p:
            raise GsError(f"no batch buffer {name}")
        out = torch.empty(numel, dtype=dtype, device=self.hip.dev)
        torch.cuda.synchronize(self.hip.dev)
        hipMemcpy = ctypes.CDLL("libamdhip64.so").hipMemcpy
        hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, _i32]
        st = hipMemcpy(_vp(out.data_ptr()), _vp(p), numel * out.element_size(), 3)
        if st != 0:
            raise GsError(f"hipMemcpy failed ({st})")
        return out

    def images(self):
        """Per-view rendered images (copies), HWC."""
        flat = self._buffer("image", self.pix_off[-1] * 3, torch.float32)
        return [flat[self.pix_off[v] * 3:self.pix_off[v + 1] * 3].view(c.height, c.width, 3)
                for v, c in enumerate(self.cams)]
