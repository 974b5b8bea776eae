"""3DGS PLY input (loadPly, ply_loader.cpp:32-189): the reference's loader,
compiled from source, against the layout's closed form on the CPU; the
device loader (gs_ply_load) against the reference's, bit for bit."""
import numpy as np
import pytest

from helpers import write_ply
from oracle.oracle import RefLib

CASES = [(300, 45, True), (200, 24, False), (150, 9, False), (100, 0, True)]


def _expected(raw, names, rest):
    """Copies and the SH reorder exactly; activations within 1 ulp (numpy's exp)."""
    col = {nm: raw[:, k] for k, nm in enumerate(names)}
    deg = 3 if rest >= 45 else 2 if rest >= 24 else 1 if rest >= 9 else 0
    R = {3: 15, 2: 8, 1: 3, 0: 0}[deg]
    K = (deg + 1) ** 2
    n = raw.shape[0]
    sh = np.zeros((n, K, 3), np.float32)
    for ch in range(3):
        sh[:, 0, ch] = col[f"f_dc_{ch}"]
        for k in range(R):
            sh[:, 1 + k, ch] = col[f"f_rest_{ch * R + k}"]
    q = np.stack([col[f"rot_{k}"] for k in range(4)], 1)
    ln = np.sqrt((q[:, 0] * q[:, 0] + q[:, 1] * q[:, 1]) + q[:, 2] * q[:, 2] + q[:, 3] * q[:, 3])
    inv = np.where(ln > 0, np.float32(1.0) / np.where(ln > 0, ln, 1), 1).astype(np.float32)
    rot = np.where(ln[:, None] > 0, q * inv[:, None], q).astype(np.float32)
    return dict(pos=np.stack([col["x"], col["y"], col["z"]], 1), sh=sh, rot=rot,
                scale=np.exp(np.stack([col[f"scale_{k}"] for k in range(3)], 1).astype(np.float64)),
                op=1.0 / (1.0 + np.exp(-col["opacity"].astype(np.float64))), max_coeffs=K,
                filter3d=col.get("filter_3D"))


@pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("n,rest,filt", CASES)
def test_reference_loader_layout(tmp_path, n, rest, filt):
    path = str(tmp_path / "m.ply")
    raw, names = write_ply(path, n, rest, seed=n + rest, filter3d=filt, zero_rot_at=7)
    got = RefLib().load_ply(path)
    exp = _expected(raw, names, rest)
    assert got["n"] == n and got["max_coeffs"] == exp["max_coeffs"]
    np.testing.assert_array_equal(got["pos"], exp["pos"])
    np.testing.assert_array_equal(got["sh"], exp["sh"])
    np.testing.assert_array_equal(got["rot"], exp["rot"])
    np.testing.assert_allclose(got["scale"], exp["scale"], rtol=3e-7)
    np.testing.assert_allclose(got["op"], exp["op"], rtol=3e-7)
    if filt:
        np.testing.assert_array_equal(got["filter3d"], exp["filter3d"])
    else:
        assert got["filter3d"] is None


@pytest.mark.gpu
@pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("n,rest,filt", CASES + [(70_000, 45, False)])
def test_gpu_ply_load_matches_reference(hip, tmp_path, n, rest, filt):
    print("[case] ply", n, rest, filt, flush=True)
    path = str(tmp_path / "m.ply")
    write_ply(path, n, rest, seed=n + rest + 1, filter3d=filt, zero_rot_at=3)
    ref = RefLib().load_ply(path)
    got = hip.load_ply(path)
    hip.sync()
    assert got["n"] == n and got["max_coeffs"] == ref["max_coeffs"]
    for k in ("pos", "scale", "rot", "sh", "op") + (("filter3d",) if filt else ()):
        np.testing.assert_array_equal(got[k].cpu().numpy().view(np.uint32), ref[k].view(np.uint32), err_msg=k)
    assert (got["filter3d"] is None) == (not filt)


def test_ply_header_errors(tmp_path):
    """Header rules of ply_loader.cpp:39-96 through gs_ply_info (no GPU call)."""
    import ctypes
    from gs_hip import load_library
    lib = load_library()
    p = tmp_path / "bad.ply"
    p.write_bytes(b"not a ply\n")
    a = [ctypes.c_int() for _ in range(4)]
    assert lib.gs_ply_info(str(p).encode(), *[ctypes.byref(x) for x in a]) != 0
    p.write_bytes(b"ply\nformat binary_little_endian 1.0\nelement vertex 2\nproperty float x\nend_header\n")
    assert lib.gs_ply_info(str(p).encode(), *[ctypes.byref(x) for x in a]) != 0
    write_ply(str(p), 5, 24, seed=1)
    assert lib.gs_ply_info(str(p).encode(), *[ctypes.byref(x) for x in a]) == 0
    assert [x.value for x in a[:3]] == [5, 2, 9]


@pytest.mark.gpu
@pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built")
@pytest.mark.parametrize("n,deg", [(500, 3), (300, 1), (200, 0)])
def test_gpu_ply_save_roundtrip(hip, oracle, tmp_path, n, deg):
    """gs_ply_save (savePly, train_main.cpp:82-146) read back by the
    reference's loadPly equals the reference activation of the raw parameters
    (the loader applies the same exp, sigmoid and normalisation)."""
    from gs_scenes import gaussians_normal
    print("[case] ply save", n, deg, flush=True)
    scene = gaussians_normal(n, 60 + deg, deg)
    path = str(tmp_path / "s.ply")
    hip.save_ply(path, hip.upload_scene(scene))
    got = RefLib().load_ply(path)
    act = oracle.activate(scene)
    assert got["n"] == n and got["max_coeffs"] == scene.max_coeffs
    for k in ("pos", "scale", "rot", "sh", "op"):
        np.testing.assert_array_equal(got[k].reshape(act[k].shape).view(np.uint32), act[k].view(np.uint32), err_msg=k)


def test_ply_load_refuses_mismatched_header(tmp_path):
    """gs_ply_load checks the file's header against the sizes the caller's
    arrays were allocated for, before any device work (no GPU needed: with the
    right sizes the call gets as far as the missing context)."""
    import ctypes
    import os
    lib_path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "aaags-claude_amd", "lib",
                            "libgs_hip.so")
    if not os.path.exists(lib_path):
        pytest.skip("libgs_hip.so not built")
    lib = ctypes.CDLL(lib_path)
    lib.gs_last_error.restype = ctypes.c_char_p
    path = str(tmp_path / "m.ply")
    write_ply(path, 50, 24, seed=5, filter3d=False)
    n, deg, mc, hf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    assert lib.gs_ply_info(path.encode(), ctypes.byref(n), ctypes.byref(deg), ctypes.byref(mc), ctypes.byref(hf)) == 0
    assert (n.value, mc.value, hf.value) == (50, 9, 0)
    fake = [ctypes.c_void_p(4096 * (k + 1)) for k in range(6)]  # never dereferenced: refused first

    def load(nn, kk, filt):
        return lib.gs_ply_load(None, path.encode(), nn, kk, *fake[:5], fake[5] if filt else None)

    assert load(51, 9, False) == 1 and b"holds 50 Gaussians" in lib.gs_last_error()
    assert load(50, 16, False) == 1 and b"9 SH coefficients" in lib.gs_last_error()
    assert load(50, 9, True) == 1 and b"no filter_3D" in lib.gs_last_error()
    assert load(50, 9, False) == 1 and b"null context" in lib.gs_last_error()
