"""The training run around the step (train_main.cpp:31-63, 613-673) on the
device:
- gs_ply_load_raw (loadPly then initRawFromGaussianData) equals the reference's
  two functions, compiled from source, bit for bit;
- gs_trainer_train_views draws the reference's view sequence (its
  std::mt19937 / uniform_int_distribution, seeded 42), its first step's loss
  equals the oracle's for that view, the loss on a revisited view falls, and
  its checkpoints "<out>.iter<N>.ply" and final file hold the trainer's raw
  parameters bit for bit."""
import ctypes
import os

import numpy as np
import pytest
import torch

from gs_scenes import RenderCfg, make_camera, random_small_scene, sphere_eye
from helpers import write_ply
from oracle.oracle import REF_LIB, RefLib

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not RefLib.available(), reason="oracle/_ref not built")]
_vp = ctypes.c_void_p


def _ref():
    R = ctypes.CDLL(REF_LIB)
    R.ref_ply_raw.argtypes = [ctypes.c_char_p, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    R.ref_view_sequence.argtypes = [ctypes.c_int, ctypes.c_int, _vp]
    return R


def ref_ply_raw(path):
    R = _ref()
    n, k = ctypes.c_int(), ctypes.c_int()
    assert R.ref_ply_raw(path.encode(), ctypes.byref(n), ctypes.byref(k), None, None, None, None, None) == 0
    n, k = n.value, k.value
    out = dict(pos=np.zeros((n, 3), np.float32), scale=np.zeros((n, 3), np.float32), rot=np.zeros((n, 4), np.float32),
               sh=np.zeros((n, k, 3), np.float32), op=np.zeros(n, np.float32))
    a = [out[x].ctypes.data_as(_vp) for x in ("pos", "scale", "rot", "sh", "op")]
    assert R.ref_ply_raw(path.encode(), ctypes.byref(ctypes.c_int()), ctypes.byref(ctypes.c_int()), *a) == 0
    return out


def read_raw_ply(path):
    """The float table of a binary little-endian PLY, by property name."""
    data = open(path, "rb").read()
    end = data.index(b"end_header\n") + len(b"end_header\n")
    names, n = [], 0
    for line in data[:end].decode().splitlines():
        if line.startswith("element vertex"):
            n = int(line.split()[2])
        elif line.startswith("property float"):
            names.append(line.split()[2])
    tab = np.frombuffer(data[end:], np.float32).reshape(n, len(names))
    return {nm: tab[:, k] for k, nm in enumerate(names)}


@pytest.mark.parametrize("n,rest,filt", [(300, 45, True), (150, 9, False), (120, 0, False), (70_000, 45, False)])
def test_gpu_ply_load_raw_matches_reference(hip, tmp_path, n, rest, filt):
    print("[case] ply raw", n, rest, filt, flush=True)
    path = str(tmp_path / "m.ply")
    write_ply(path, n, rest, seed=n + rest + 5, filter3d=filt, zero_rot_at=4)
    got = hip.load_ply_raw(path)
    torch.cuda.synchronize()
    ref = ref_ply_raw(path)
    assert got["n"] == n
    for k in ("pos", "scale", "rot", "sh", "op"):
        np.testing.assert_array_equal(got[k].cpu().numpy().view(np.uint32), ref[k].view(np.uint32), err_msg=k)


def test_train_views_sequence_loss_and_checkpoints(hip, oracle, tmp_path):
    from gs_hip import GsTrainer, train_cfg
    print("[case] train views", flush=True)
    scene, _, _ = random_small_scene(1500, 3, 231)
    rng = np.random.Generator(np.random.Philox(232))
    cams = [make_camera(sphere_eye(rng, 4.0), w, h, 60.0) for (w, h) in [(96, 64), (80, 60), (112, 72), (64, 64)]]
    # targets: the scene itself with its colours shifted, so that the loss
    # has somewhere to go on every view
    tscene = scene.copy()
    tscene.sh[:, 0, :] += np.float32(0.6)
    cfg = RenderCfg(sh_degree=3, training=True)
    gts = [torch.tensor(oracle.forward(tscene, c, cfg, backward_cache=False)["img"], device=hip.dev) for c in cams]
    tr = GsTrainer(hip, scene)
    out = str(tmp_path / "out.ply")
    iters = 30
    losses, views = tr.train_views(cams, cfg, gts, train_cfg(lr_feature=0.01), iters, seed=42, save_every=10, output=out)
    torch.cuda.synchronize()

    seq = np.zeros(iters, np.int32)
    _ref().ref_view_sequence(len(cams), iters, seq.ctypes.data_as(_vp))
    np.testing.assert_array_equal(views, seq)

    first = oracle.forward_backward(scene, cams[views[0]], cfg, gts[views[0]].cpu().numpy())["loss"]
    assert abs(losses[0] - first) <= 1e-6, (losses[0], first)
    for v in range(len(cams)):
        at = np.nonzero(views == v)[0]
        if len(at) >= 2:
            assert losses[at[-1]] < losses[at[0]], (v, losses[at])

    names = [os.path.basename(p) for p in sorted(tmp_path.iterdir())]
    for it in (10, 20, 30):
        assert f"out.ply.iter{it}.ply" in names, names
    assert open(out, "rb").read() == open(out + ".iter30.ply", "rb").read()
    pos, ls, rot, sh, op = tr.params()
    tab = read_raw_ply(out)
    R = (scene.max_coeffs - 1)
    np.testing.assert_array_equal(np.stack([tab["x"], tab["y"], tab["z"]], 1), pos)
    np.testing.assert_array_equal(np.stack([tab[f"scale_{k}"] for k in range(3)], 1), ls)
    np.testing.assert_array_equal(np.stack([tab[f"rot_{k}"] for k in range(4)], 1), rot)
    np.testing.assert_array_equal(tab["opacity"], op)
    for ch in range(3):
        np.testing.assert_array_equal(tab[f"f_dc_{ch}"], sh[:, 0, ch])
        for k in range(R):
            np.testing.assert_array_equal(tab[f"f_rest_{ch * R + k}"], sh[:, 1 + k, ch])
    # the checkpoint at 10 differs from the final parameters (training moved on)
    assert not np.array_equal(read_raw_ply(out + ".iter10.ply")["x"], tab["x"])


def test_train_views_rejects_bad_arguments(hip, tmp_path):
    from gs_hip import GsError, GsTrainer, train_cfg
    scene, cam, gt = random_small_scene(50, 0, 233)
    tr = GsTrainer(hip, scene)
    g = torch.tensor(gt, device=hip.dev)
    with pytest.raises(GsError):
        tr.train_views([cam], RenderCfg(sh_degree=0, training=True), [g], train_cfg(), 3, save_every=1, output=None)
