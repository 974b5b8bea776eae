"""FrameAllocator, the host bump arena of the C++ interface
(aaags-claude_amd/cpp/gs3d_types.cpp), against the contract the reference's
tests/test_types.cpp:10-41 checks for include/types.h:8-31: 64-byte aligned
blocks, std::runtime_error on exhaustion (no wrap-around on huge requests), the
whole arena reusable after reset().  Host code only: runs without a GPU
through the test shim (tests/cpp/stage_shim.cpp)."""
import ctypes
import os

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SHIM = os.path.join(HERE, "cpp", "_build", "libgs_stage_shim.so")


def test_frame_allocator_contract():
    if not os.path.exists(SHIM):
        pytest.skip("tests/cpp/_build/libgs_stage_shim.so not built (make -C tests/cpp)")
    lib = ctypes.CDLL(SHIM)
    assert lib.shim_frame_allocator_contract() == 0
