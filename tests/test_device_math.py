"""The library's restatements of glibc's expf and logf (exp_ref, log_ref in
aaags-claude_amd/csrc/gs_device.h), which the device uses wherever the
reference's CPU path calls them (activation, the rasterizers' exact alpha near
the 1/255 cut, the eval_3D cull threshold).  The same source is compiled for
the host (tests/cpp/mathcheck.cpp) and compared bit for bit with this
machine's glibc: every float of the ranges the path reaches, and a stride
through all 2^32 bit patterns."""
import ctypes
import os
import struct

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "cpp", "_build", "libgs_mathcheck.so")


def _bits(x):
    return struct.unpack("<I", struct.pack("<f", x))[0]


@pytest.fixture(scope="module")
def mc():
    if not os.path.exists(LIB):
        pytest.skip("tests/cpp/_build/libgs_mathcheck.so not built (make -C tests/cpp)")
    L = ctypes.CDLL(LIB)
    for f in (L.check_log_ref, L.check_exp_ref):
        f.restype = ctypes.c_longlong
        f.argtypes = [ctypes.c_uint32, ctypes.c_uint64, ctypes.c_uint32]
    return L


def test_log_ref_equals_glibc_logf(mc):
    # eval_3D: logf(opacity / (1/255)) with the argument in [1, 255]
    assert mc.check_log_ref(_bits(0.99), _bits(256.0) + 1, 1) == 0
    # all bit patterns (negatives, zeros, subnormals, infinities, NaNs included), stride 97
    assert mc.check_log_ref(0, 1 << 32, 97) == 0


def test_exp_ref_equals_glibc_expf(mc):
    # the rasterizers' alpha: exp of a non-positive power; activation: exp(log-scale), exp(-logit)
    assert mc.check_exp_ref(_bits(-12.0), _bits(-16.0) + 1, 1) == 0  # [-16, -12] (negative floats grow in bits)
    assert mc.check_exp_ref(_bits(0.0), _bits(4.0) + 1, 3) == 0
    assert mc.check_exp_ref(0, 1 << 32, 97) == 0
