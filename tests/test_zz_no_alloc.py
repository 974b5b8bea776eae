"""gs_scan_tiles_noalloc (the scan through the library's fixed per-device
scratch, for callers that cannot allocate) against gs_scan_tiles on the same
inputs through the C ABI: offsets and pair count bit-exact, on two calls in a
row, the second smaller (the scratch keeps the first call's contents); the
empty input; and a clear error beyond the fixed maximum."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

_vp = ctypes.c_void_p
GS_NOALLOC_SCAN_MAX = 16777216


def _p(t):
    return _vp(t.data_ptr())


def test_scan_tiles_noalloc_matches_scan_tiles(hip):
    from gs_hip import GsHip
    print("[case] gs_scan_tiles_noalloc vs gs_scan_tiles", flush=True)
    fresh = GsHip(0)  # a context on which nothing has allocated a workspace yet
    lib = fresh.lib
    lib.gs_scan_tiles_noalloc.argtypes = [_vp, ctypes.c_int, _vp, _vp, _vp]
    rng = np.random.default_rng(7)
    for n in (3_000_017, 70_001):
        t_np = rng.integers(0, 48, n, dtype=np.int32)
        t_np[rng.integers(0, n, n // 3)] = 0
        tiles = torch.tensor(t_np, device=fresh.dev)
        off = torch.full((n,), -1, dtype=torch.int32, device=fresh.dev)
        tot = torch.full((1,), -1, dtype=torch.int32, device=fresh.dev)
        fresh._check(lib.gs_scan_tiles_noalloc(fresh.ctx, n, _p(tiles), _p(off), _p(tot)), "gs_scan_tiles_noalloc")
        ref_off = torch.empty(n, dtype=torch.int32, device=fresh.dev)
        ref_tot = torch.zeros(1, dtype=torch.int32, device=fresh.dev)
        total_host = ctypes.c_longlong()
        fresh._check(lib.gs_scan_tiles(fresh.ctx, n, _p(tiles), _p(ref_off), _p(ref_tot), ctypes.byref(total_host)),
                     "gs_scan_tiles")
        torch.cuda.synchronize()
        assert torch.equal(off, ref_off)
        assert int(tot.item()) == int(ref_tot.item()) == total_host.value == int(t_np.astype(np.int64).sum())
        np.testing.assert_array_equal(off.cpu().numpy().view(np.uint32).astype(np.int64),
                                      np.cumsum(t_np.astype(np.int64)))
    # empty input: the pair count is zeroed
    tot.fill_(-1)
    fresh._check(lib.gs_scan_tiles_noalloc(fresh.ctx, 0, None, None, _p(tot)), "gs_scan_tiles_noalloc(0)")
    torch.cuda.synchronize()
    assert int(tot.item()) == 0
    # beyond the fixed scratch: refused before any launch
    st = lib.gs_scan_tiles_noalloc(fresh.ctx, GS_NOALLOC_SCAN_MAX + 1, _p(tiles), _p(off), _p(tot))
    assert st == 1
    assert b"exceed its fixed scratch" in lib.gs_last_error()
