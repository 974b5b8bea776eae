"""ctypes drivers for the density-control tests: the reference's own
DensityController / OwnedRawParams / Trainer::step_with_densify (compiled from
source into oracle/_ref) and this library's DensityControllerHIP /
TrainerHIP::step_with_densify (through tests/cpp/stage_shim.cpp), behind the
same small interface so a test runs one sequence of calls on both."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM = os.path.join(HERE, "cpp", "_build", "libgs_stage_shim.so")
_vp = ctypes.c_void_p
_i32 = ctypes.c_int
_f32 = ctypes.c_float


def _p(a):
    return a.ctypes.data_as(_vp)


def dcfg_arrays(grad_threshold=0.0002, min_opacity=0.005, percent_dense=0.01, densify_from_iter=500,
                densify_until_iter=15000, densify_interval=100, opacity_reset_interval=3000, max_screen_size=20,
                enabled=False):
    """DensifyConfig (train_types.h:71-81) as the two arrays the shims take."""
    ic = np.array([densify_from_iter, densify_until_iter, densify_interval, opacity_reset_interval, max_screen_size,
                   int(enabled)], np.int32)
    fc = np.array([grad_threshold, min_opacity, percent_dense], np.float32)
    return ic, fc


def train_arrays(lr=1e-3, use_adam=False, max_steps=30000, beta1=0.9, beta2=0.999, eps=1e-8):
    tcf = np.array([lr, lr, lr, lr, lr, lr, beta1, beta2, eps], np.float32)
    tci = np.array([max_steps, int(use_adam)], np.int32)
    return tcf, tci


class Side:
    """One implementation: prefix 'ref_' (the reference build) or 'shim_' (the
    HIP classes)."""

    def __init__(self, lib, prefix):
        self.L, self.pf = lib, prefix
        g = lambda n: getattr(lib, prefix + n)
        self.params_new, self.params_free, self.params_count, self.params_get = (
            g("params_new"), g("params_free"), g("params_count"), g("params_get"))
        self.params_new.restype = _vp
        self.params_new.argtypes = [_i32, _i32, _i32] + [_vp] * 5
        self.params_free.argtypes = [_vp]
        self.params_count.argtypes = [_vp]
        self.params_get.argtypes = [_vp] * 6
        self.dc_new, self.dc_free = g("dc_new"), g("dc_free")
        self.dc_new.restype = _vp
        self.dc_free.argtypes = [_vp]
        self.trainer_free = g("trainer_free")
        self.trainer_free.argtypes = [_vp]
        g("dc_reset_opacity").argtypes = [_vp]
        g("dc_reset").argtypes = [_vp, _i32]
        g("dc_accumulate").argtypes = [_vp, _vp, _vp, _i32]

    def _check(self, st, what):
        if self.pf == "shim_" and st != 0:
            self.L.shim_error.restype = ctypes.c_char_p
            raise RuntimeError(f"{what}: {self.L.shim_error().decode()}")

    def new_params(self, pos, sc, rot, sh, op, deg):
        n, k = pos.shape[0], sh.shape[1]
        arrs = [np.ascontiguousarray(a, np.float32) for a in (pos, sc, rot, sh, op)]
        return self.params_new(n, deg, k, *[_p(a) for a in arrs])

    def get_params(self, h, k):
        n = self.params_count(h)
        out = [np.zeros((n, 3), np.float32), np.zeros((n, 3), np.float32), np.zeros((n, 4), np.float32),
               np.zeros((n, k, 3), np.float32), np.zeros(n, np.float32)]
        if n:
            self.params_get(h, *[_p(a) for a in out])
        return out

    def reset(self, dc, n):
        st = getattr(self.L, self.pf + "dc_reset")(dc, n)
        self._check(st if self.pf == "shim_" else 0, "reset")

    def accumulate(self, dc, d_means2D, radii=None):
        d = np.ascontiguousarray(d_means2D, np.float32)
        r = None if radii is None else np.ascontiguousarray(radii, np.int32)
        n = d.size // 2
        st = getattr(self.L, self.pf + "dc_accumulate")(dc, _p(d), _p(r) if r is not None else None, n)
        self._check(st if self.pf == "shim_" else 0, "accumulate")

    def densify(self, dc, params, extent, **cfg):
        ic, fc = dcfg_arrays(**cfg)
        if self.pf == "ref_":
            f = self.L.ref_dc_densify
            f.restype = _i32
            f.argtypes = [_vp, _vp, _vp, _vp, _f32]
            return int(f(dc, params, _p(ic), _p(fc), extent))
        f = self.L.shim_dc_densify
        f.argtypes = [_vp, _vp, _vp, _vp, _f32, _vp]
        n = _i32()
        self._check(f(dc, params, _p(ic), _p(fc), extent, ctypes.byref(n)), "densify")
        return n.value

    def reset_opacity(self, params):
        getattr(self.L, self.pf + "dc_reset_opacity")(params)

    def new_trainer(self):
        if self.pf == "ref_":
            f = self.L.ref_trainer_new
            f.restype = _vp
            f.argtypes = [ctypes.c_size_t]
            return f(256 << 20)
        f = self.L.shim_trainer_new
        f.restype = _vp
        return f()

    def step_densify(self, tr, params, cam, cfg, gt, tc, extent, iteration, **dcfg):
        from oracle.oracle import cam_struct, cfg_struct
        ic, fc = dcfg_arrays(**dcfg)
        tcf, tci = tc
        gt = np.ascontiguousarray(gt, np.float32)
        n = _i32()
        if self.pf == "ref_":
            c, f = cam_struct(cam), cfg_struct(cfg)
            fn = self.L.ref_trainer_step_densify
            fn.restype = _f32
            fn.argtypes = [_vp] * 9 + [_f32, _i32, _vp]
            loss = fn(tr, params, ctypes.byref(c), ctypes.byref(f), _p(gt), _p(tcf), _p(tci), _p(ic), _p(fc), extent,
                      iteration, ctypes.byref(n))
            return float(loss), n.value
        c, f = cam.to_c(), cfg.to_c()
        fn = self.L.shim_trainer_step_densify
        fn.argtypes = [_vp] * 9 + [_f32, _i32, _vp, _vp]
        loss = _f32()
        self._check(fn(tr, params, ctypes.byref(c), ctypes.byref(f), _p(gt), _p(tcf), _p(tci), _p(ic), _p(fc), extent,
                       iteration, ctypes.byref(loss), ctypes.byref(n)), "step_with_densify")
        return float(loss.value), n.value


def reference_side():
    from oracle.oracle import REF_LIB
    return Side(ctypes.CDLL(REF_LIB), "ref_")


def hip_side():
    return Side(ctypes.CDLL(SHIM), "shim_")


def inv_sigmoid(x):
    return np.float32(np.log(np.float32(x) / (np.float32(1.0) - np.float32(x))))


def make_test_params(n, sh_degree=0):
    """make_test_params of the reference's tests/test_density_controller.cpp:10-28."""
    k = (sh_degree + 1) ** 2
    pos = np.zeros((n, 3), np.float32)
    pos[:, 0] = np.arange(n, dtype=np.float32) * np.float32(0.5)
    pos[:, 2] = 5.0
    sc = np.full((n, 3), np.log(np.float32(0.1)), np.float32)
    rot = np.zeros((n, 4), np.float32)
    rot[:, 0] = 1.0
    op = np.full(n, inv_sigmoid(0.5), np.float32)
    sh = np.ones((n, k, 3), np.float32)
    return pos, sc, rot, sh, op
