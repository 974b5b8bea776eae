"""The camera file and image formats around the training path (host code of
libgs_hip.so, no GPU needed):
- gs_cameras_json on the reference's own cameras.json (76 views) equals the
  reference's loadCamerasJson (train_main.cpp:262-377, compiled from source)
  bit for bit, and its matrices satisfy the closed form: a rigid world-to-camera
  transform taking the camera centre to the origin, and viewproj = proj * view;
- gs_write_ppm / gs_read_ppm equal the reference's writePPM / readPPM byte for
  byte and bit for bit, out-of-range and NaN values included."""
import ctypes
import os

import numpy as np
import pytest

from oracle.oracle import REF_LIB, RefLib

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB = os.path.join(ROOT, "aaags-claude_amd", "lib", "libgs_hip.so")
CAMS = os.path.join(HERE, "golden", "cameras", "cameras.json")
_vp = ctypes.c_void_p


def _p(a):
    return a.ctypes.data_as(_vp)


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.skip("libgs_hip.so not built")
    L = ctypes.CDLL(LIB)
    L.gs_last_error.restype = ctypes.c_char_p
    L.gs_read_ppm.argtypes = [ctypes.c_char_p, _vp, ctypes.c_size_t, _vp, _vp]
    return L


class GsCam(ctypes.Structure):
    _fields_ = [("view", ctypes.c_float * 16), ("viewproj", ctypes.c_float * 16), ("cam_pos", ctypes.c_float * 3),
                ("tanx", ctypes.c_float), ("tany", ctypes.c_float), ("width", ctypes.c_int), ("height", ctypes.c_int)]


def load_cameras(lib, path):
    n = ctypes.c_int()
    assert lib.gs_cameras_json(path.encode(), None, 0, ctypes.byref(n), None, 0) == 0, lib.gs_last_error()
    cams = (GsCam * n.value)()
    names = ctypes.create_string_buffer(64 * n.value)
    assert lib.gs_cameras_json(path.encode(), cams, n.value, ctypes.byref(n), names, 64) == 0
    out = np.zeros((n.value, 39), np.float32)
    for k, c in enumerate(cams):
        out[k, :16] = c.view
        out[k, 16:32] = c.viewproj
        out[k, 32:35] = c.cam_pos
        out[k, 35:39] = (c.tanx, c.tany, c.width, c.height)
    img = [names.raw[64 * k:64 * (k + 1)].split(b"\0")[0].decode() for k in range(n.value)]
    return out, img


@pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (no reference tree)")
def test_cameras_json_equals_reference_loader(lib):
    got, names = load_cameras(lib, CAMS)
    assert got.shape[0] == 76 and names[0] == "78899858295079"
    R = ctypes.CDLL(REF_LIB)
    n = ctypes.c_int()
    assert R.ref_cameras_json(CAMS.encode(), None, 0, ctypes.byref(n)) == 0
    ref = np.zeros((n.value, 39), np.float32)
    assert R.ref_cameras_json(CAMS.encode(), _p(ref), n.value, ctypes.byref(n)) == 0
    np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_cameras_json_closed_form(lib):
    got, _ = load_cameras(lib, CAMS)
    for row in got:
        V = row[:16].astype(np.float64).reshape(4, 4).T  # column-major -> matrix
        VP = row[16:32].astype(np.float64).reshape(4, 4).T
        c = row[32:35].astype(np.float64)
        Rm = V[:3, :3]
        np.testing.assert_allclose(Rm @ Rm.T, np.eye(3), atol=1e-5)
        np.testing.assert_allclose(V @ np.append(c, 1.0), [0, 0, 0, 1], atol=1e-5)
        tx, ty = float(row[35]), float(row[36])
        zn, zf = 0.01, 100.0
        P = np.zeros((4, 4))
        P[0, 0], P[1, 1] = 1 / tx, 1 / ty
        P[2, 2], P[3, 2], P[2, 3] = zf / (zf - zn), 1.0, -(zf * zn) / (zf - zn)
        np.testing.assert_allclose(VP, P @ V, rtol=1e-5, atol=1e-6)


def test_cameras_json_rejects_bad_files(lib, tmp_path):
    bad = tmp_path / "bad.json"
    bad.write_text('[{"id": 0, "width": 4, ')
    n = ctypes.c_int()
    assert lib.gs_cameras_json(str(bad).encode(), None, 0, ctypes.byref(n), None, 0) == 1
    none = tmp_path / "none.json"
    none.write_text('[{"id": 0, "fx": 0, "fy": 1, "position": [0,0,0], "rotation": [[1,0,0],[0,1,0],[0,0,1]]}]')
    assert lib.gs_cameras_json(str(none).encode(), None, 0, ctypes.byref(n), None, 0) == 0 and n.value == 0
    assert lib.gs_cameras_json(str(tmp_path / "missing.json").encode(), None, 0, ctypes.byref(n), None, 0) == 1


@pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built (no reference tree)")
def test_ppm_matches_reference(lib, tmp_path):
    R = ctypes.CDLL(REF_LIB)
    R.ref_read_ppm.argtypes = [ctypes.c_char_p, _vp, ctypes.c_size_t, _vp, _vp]
    rng = np.random.default_rng(3)
    h, w = 17, 23
    img = rng.uniform(-0.2, 1.2, (h, w, 3)).astype(np.float32)
    img[0, 0] = [np.nan, np.inf, -np.inf]
    ours, theirs = str(tmp_path / "a.ppm"), str(tmp_path / "b.ppm")
    assert lib.gs_write_ppm(ours.encode(), _p(img), w, h) == 0
    assert R.ref_write_ppm(theirs.encode(), _p(img), w, h) == 0
    assert open(ours, "rb").read() == open(theirs, "rb").read()
    W, H = ctypes.c_int(), ctypes.c_int()
    a = np.zeros(h * w * 3, np.float32)
    b = np.zeros(h * w * 3, np.float32)
    assert lib.gs_read_ppm(ours.encode(), _p(a), a.size, ctypes.byref(W), ctypes.byref(H)) == 0
    assert (W.value, H.value) == (w, h)
    assert R.ref_read_ppm(ours.encode(), _p(b), b.size, ctypes.byref(W), ctypes.byref(H)) == 0
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    # a comment line before the sizes
    raw = open(ours, "rb").read()
    cm = str(tmp_path / "c.ppm")
    open(cm, "wb").write(raw.replace(b"P6\n", b"P6\n# made here\n", 1))
    assert lib.gs_read_ppm(cm.encode(), _p(a), a.size, ctypes.byref(W), ctypes.byref(H)) == 0
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
