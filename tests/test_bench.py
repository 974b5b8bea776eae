"""bench.py's own logic, on the CPU: it refuses more than one GPU, and the
byte models behind its roofline and bandwidth figures give the per-unit
numbers the analysis states (1M Gaussians, SH degree 3, 1920x1080: 284 B per
Gaussian for the preprocess, 1456 B for the fused backward + Adam, 152 B per
pair for the reference-compatible sort of 45-bit keys)."""
import importlib.util
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_byte_models_at_the_headline_workload():
    b = _bench()
    N, npix, tiles = 1_000_000, 1920 * 1080, 120 * 68
    K = 47_215_016
    # 59 parameters per Gaussian at SH degree 3 (3 + 3 + 4 + 16 x 3 + 1)
    assert b.algorithmic_bytes("preprocess", N, K, npix, 16, tiles) == N * (59 * 4 + 48) == 284_000_000
    assert b.algorithmic_bytes("backward", N, K, npix, 16, tiles) == N * 1456
    assert b.algorithmic_bytes("adam", N, K, npix, 16, tiles) == N * 59 * 28
    # ceil(log2 8160) = 13 tile bits + 32 depth bits = 45 -> six 8-bit passes
    assert b.algorithmic_bytes("sort", N, K, npix, 16, tiles) == K * 152
    assert b.algorithmic_bytes("raster_bwd", N, K, npix, 16, tiles) == K * 40 + npix * 20 + N * 72
    assert b.algorithmic_bytes("l1", N, K, npix, 16, tiles) == npix * 36
    # SH degree 0 (one basis): 3 + 3 + 4 + 3 + 1 = 14 parameters
    assert b.algorithmic_bytes("preprocess", N, K, npix, 1, tiles) == N * (14 * 4 + 48)
    assert b.sort_bytes(47_194_816, 45, 8) == 47_194_816 * 152
    assert b.sort_bytes(946_668, 32, 8) == 946_668 * (24 * 4 + 8)
    assert b.sort_bytes(1000, 35, 8) == 1000 * (24 * 5 + 8)


def test_committed_traffic_summary_is_readable():
    b = _bench()
    st = b.load_traffic("1m_1080p")
    for k in ("preprocess", "binning", "raster_fwd", "l1", "raster_bwd", "backward"):
        assert st[k]["traffic_bytes"] == st[k]["fetch_bytes"] + st[k]["write_bytes"] > 0
    # the preprocess moves its algorithmic bytes to within 1%
    assert abs(st["preprocess"]["traffic_bytes"] - 284_000_000) < 0.01 * 284_000_000
    for batch in ("64", "128", "256"):
        assert st["raster_bwd"]["valu_wave_insts_by_batch"][batch] > 0
    assert b.load_traffic("no_such_workload") == {}


def test_reference_opencl_status_names_what_is_missing(monkeypatch):
    b = _bench()
    monkeypatch.delenv("REFERENCE_PATH", raising=False)
    s = b.reference_opencl_status()
    assert s.startswith("reference OpenCL on PoCL: unavailable") and "reference tree" in s
