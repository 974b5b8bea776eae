"""Benchmark of the 3DGS render-and-train hot path on one MI355X.

Workload: 1M Gaussians (SH degree 3; means U[-1,1]^3, log-scales N(-4,0.5),
opacity logits N(0,1)), one 1920x1080 pinhole camera (fov 60 deg, distance 3),
L1 against a uniform random target, Adam (lr 1e-3, betas 0.9/0.999).  A step
is preprocess -> bin -> sort -> raster -> L1 -> raster backward -> preprocess
backward -> Adam, run by the native driver (gs_trainer_step) with inputs
resident in HBM.

Prints one JSON line: train steps/s (value), forward Mpix/s, a per-stage
breakdown (HIP events around each stage of the native step), the roofline of
the dominant kernel (its duration measured with HIP events inside the timed
steps), the HBM throughput of preprocess and of the two sorts, the batched and
large batched steps, and the reference CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "aaags-claude_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# a wave64 FP32 VALU instruction occupies a SIMD for 4 cycles (16 lanes a
# cycle; gs_device_peaks measures it): 256 CUs x 4 SIMDs x 2.4 GHz / 4
VALU_WAVE_INSTS_PER_S = 256 * 4 * 2.4e9 / 4

# FP32 operations per pixel-Gaussian evaluation of the reference's raster
# loops (cpu/rasterizer_cpu.cpp:141-176, cpu/rasterizer_backward_cpu.cpp:34-134),
# counted operation by operation:
# - every evaluated entry: dx and dy (2), the power -0.5 (a dx^2 + c dy^2) - b dx dy
#   (6 multiplies, 2 adds, 1 scaling: 9), the power > 0 test, exp, opacity * exp,
#   the 0.99 clamp and the 1/255 test (5): 16;
# - forward, per contributor: 1 - alpha, T (1 - alpha), the 1e-4 test and three
#   colour terms rgb * alpha * T accumulated (9): 12;
# - backward, per contributor: 1 - alpha and T_i (2), d_rgb (7), d_alpha (12),
#   accum_rec (9), d_power and d_opacity (3) and its sum (1), d_means2D (10),
#   d_conics (12): 56.
FLOPS_PER_EVAL, FLOPS_FWD_CONTRIB, FLOPS_BWD_CONTRIB = 16, 12, 56
# LDS bytes an evaluation reads: mean 8, conic 12, opacity 4, colour 12
LDS_BYTES_PER_EVAL = 36


def load_traffic(workload):
    """Per-stage HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (calibrated
    with a 1 GiB device copy) and VALU wave-instruction counts, from the committed summary for this
    workload; {} when none is committed."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{workload}.json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("stages", {})


def algorithmic_bytes(stage, N, K, npix, max_coeffs, num_tiles, table_entries=0):
    """Bytes each stage must move at minimum (per launch), from the per-unit
    figures of the hot-path analysis: preprocess 236 B/G read (SH3) + 48 B/G
    written; sort K*(24*P8 + 8), P8 = ceil((ceil(log2 tiles) + 32) / 8);
    raster forward 40 B/pair + 20 B/px; raster backward 40 B/pair + 20 B/px +
    72 B/G; preprocess backward ~580 B/G; Adam 28 B per parameter element."""
    import math
    params_per_g = 3 + 3 + 4 + max_coeffs * 3 + 1
    p8 = math.ceil((math.ceil(math.log2(max(num_tiles, 2))) + 32) / 8)
    return {
        "preprocess": N * (4 * params_per_g + 48),
        "scan": N * 8,
        "duplicate": N * 20 + K * 12,
        "sort": K * (24 * p8 + 8),
        # direct-placement binning: depth keys (16 B/G) + 4 depth-sort passes (32 B/G each), two walks
        # over the depth-ordered Gaussians (16 B/G each), the (tile, chunk) table written, scanned
        # (read + write) and read (16 B/entry), each pair's index written once (4 B/pair)
        "binning": N * (16 + 4 * 32 + 2 * 16) + table_entries * 16 + K * 4,
        "raster_fwd": K * 40 + npix * 20,
        "l1": npix * 36,
        "raster_bwd": K * 40 + npix * 20 + N * 72,
        "preprocess_bwd": N * 580,
        "adam": N * params_per_g * 28,
        # fused preprocess backward + Adam: parameters, m and v read and written (6 x 4 B per
        # parameter element), screen-space gradients (36 B/G) and radii (4 B/G) read
        "backward": N * (params_per_g * 24 + 40),
    }.get(stage)


_FLUSH = []


def cold_ms(f, reps=5):
    """Device time (ms) of one f() with cold caches: before each run a 512 MiB
    buffer is written (twice the 256 MB Infinity Cache, far beyond the L2s),
    so that f reads its inputs from HBM as the stages of a step do; HIP events
    on torch's current stream (the library's stream) around the single run;
    the median over reps, after one untimed run."""
    import torch
    if not _FLUSH:
        _FLUSH.append(torch.empty(1 << 27, dtype=torch.float32, device="cuda"))
    f()
    ts = []
    for _ in range(reps):
        _FLUSH[0].fill_(1.0)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        f()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b))
    return sorted(ts)[len(ts) // 2]


def median_ms(f, reps=5, groups=3):
    """Device time (ms) of one f(): reps back-to-back runs between one pair of
    HIP events on torch's current stream (the library's stream), so launch
    ramps overlap as in a pipeline; the median over `groups` such groups,
    after one untimed run."""
    import torch
    f()
    ts = []
    for _ in range(groups):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            f()
        b.record()
        torch.cuda.synchronize()
        ts.append(a.elapsed_time(b) / reps)
    return sorted(ts)[len(ts) // 2]


def sort_bytes(count, key_bits, radix_bits):
    """An LSD radix sort of (key, 4-byte value) pairs moves at least one key read
    for the digit histograms and, per pass, each 12-byte pair read and written."""
    import math
    return count * (24 * math.ceil(key_bits / radix_bits) + 8)


def bench_sorts(hip, p, cam, cfg, radix_bits, reps=5):
    """HBM throughput of the two sorts at the headline workload, timed with HIP
    events on the library's stream, caches flushed before each run (cold_ms):
    - the reference-compatible pair sort (gs_sort_pairs): the K (tile | depth)
      64-bit keys and pair indices that the reference's duplicate step makes;
    - the depth sort of the direct-placement binning (gs_radix_sort_kv on the
      N visible Gaussians' 32-bit depth keys)."""
    import math
    import torch

    timed = lambda f: cold_ms(f, reps)
    pre = hip.preprocess(p, cam, cfg)
    offsets, K = hip.scan_tiles(pre["tiles"])
    keys, vals = hip.duplicate(pre, offsets, K, cam.width, cam.height)
    tile_bits = max(1, math.ceil(math.log2(max(cam.num_tiles, 2))))
    ms_pairs = timed(lambda: hip.sort_pairs(keys, vals, cam.num_tiles))
    b_pairs = sort_bytes(K, tile_bits + 32, radix_bits)
    del keys, vals, offsets
    vis = pre["radii"] > 0
    dkeys = pre["depths"][vis].contiguous().view(torch.int32).to(torch.int64) & 0xFFFFFFFF
    didx = torch.nonzero(vis).flatten().to(torch.int32)
    nv = int(dkeys.numel())
    ms_depth = timed(lambda: hip.radix_sort(dkeys, didx, key_bits=32, radix_bits=radix_bits))
    b_depth = sort_bytes(nv, 32, radix_bits)
    out = {}
    for name, ms, byts, n, kb in (("pair_sort", ms_pairs, b_pairs, K, tile_bits + 32),
                                  ("depth_sort", ms_depth, b_depth, nv, 32)):
        gbs = byts / (ms * 1e-3) / 1e9
        out[name] = {"count": int(n), "key_bits": kb, "radix_bits": radix_bits, "ms": round(ms, 4),
                     "algorithmic_bytes": int(byts), "gb_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def bench_batched(hip, cfg, args):
    """16 views of uneven resolution over the same 1M-Gaussian scene: the plain
    single-view path (one gs_trainer_step per view, pair count read back per
    step) against the batched, graph-captured step (gs_batch_step, all views
    in one launch per stage, Adam fused), both in views trained per second."""
    import torch
    from gs_hip import GsBatch, GsTrainer, train_cfg
    from gs_scenes import make_workload

    scene, cams, gts = make_workload(args.batched_workload)
    V = len(cams)
    pix = sum(c.width * c.height for c in cams)
    tc = train_cfg(lr=1e-3)
    gts_t = [torch.tensor(g, device=hip.dev) for g in gts]
    # plain single-view path over the same views
    tr = GsTrainer(hip, scene)
    for v in range(V):
        tr.step(cams[v], cfg, gts_t[v], tc, v)
    torch.cuda.synchronize()
    rounds = max(1, args.steps // 10)
    t0 = time.perf_counter()
    for r in range(rounds):
        for v in range(V):
            tr.step(cams[v], cfg, gts_t[v], tc, V + r * V + v)
    torch.cuda.synchronize()
    plain = rounds * V / (time.perf_counter() - t0)
    del tr
    # batched, graph-captured step
    tr = GsTrainer(hip, scene)
    b = GsBatch(tr, cams, cfg, gts, persistent=True, fused_adam=True, use_graph=True)
    tuning = None if args.no_autotune else b.autotune(reps=2)
    for i in range(2):
        b.step(tc, i, sync=True)
    steps = max(2, args.steps // 2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        b.step(tc, 2 + i)
    res, losses = b.result()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    batched = steps * V / dt
    # batched forward only
    b.render()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        b.render()
    torch.cuda.synchronize()
    fwd_mpix = pix * steps / (time.perf_counter() - t0) / 1e6
    tiles, _, cap = b.info()
    return {"workload": args.batched_workload, "views": V, "pixels": pix, "pairs": int(res.pairs),
            "views_per_s": round(batched, 2), "ms_per_batch_step": round(1e3 * dt / steps, 3),
            "plain_views_per_s": round(plain, 2), "batched_vs_single": round(batched / plain, 3),
            "fwd_mpix_s": round(fwd_mpix, 1), "loss": round(float(res.loss), 6), "tuning": tuning}


def bench_large(hip, cfg, args):
    """The large batched step: 6M Gaussians in 32 clusters, 8 views at 1080p,
    ~10^8 pairs, graph-captured with Adam fused.  Reports views/s, the pair
    count and the device memory the trainer and batch hold (hipMemGetInfo
    before and after)."""
    import torch
    from gs_hip import GsBatch, GsTrainer, train_cfg
    from gs_scenes import make_workload

    scene, cams, gts = make_workload(args.large_workload)
    torch.cuda.synchronize()
    free0, total = torch.cuda.mem_get_info()
    tr = GsTrainer(hip, scene)
    b = GsBatch(tr, cams, cfg, gts, persistent=True, fused_adam=True, use_graph=True)
    tc = train_cfg(lr=1e-3)
    for i in range(2):
        b.step(tc, i, sync=True)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    steps = 3
    t0 = time.perf_counter()
    for i in range(steps):
        b.step(tc, 2 + i)
    res, _ = b.result()
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    out = {"workload": args.large_workload, "gaussians": scene.n, "views": len(cams), "pairs": int(res.pairs),
           "views_per_s": round(steps * len(cams) / dt, 2), "ms_per_batch_step": round(1e3 * dt / steps, 3),
           "device_gb_held": round((free0 - free1) / 1e9, 2), "device_gb_total": round(total / 1e9, 1)}
    del b, tr
    torch.cuda.empty_cache()
    # the HBM-bound stages of this workload, each timed alone with cold caches:
    # preprocess of one view over all Gaussians, and the batch's depth sort (the
    # V x N (view << 32 | depth bits) keys with their indices, view bits + 32
    # key bits)
    p = hip.upload_scene(scene)
    pre1 = hip.alloc_pre(scene.n)
    ms_pre = cold_ms(lambda: hip.preprocess(p, cams[0], cfg, pre=pre1))
    b_pre = algorithmic_bytes("preprocess", scene.n, 0, 0, scene.max_coeffs, 1)
    V = len(cams)
    view_bits = max(1, (V - 1).bit_length())
    keys = torch.cat([(hip.preprocess(p, c, cfg)["depths"].view(torch.int32).to(torch.int64) & 0xFFFFFFFF) | (v << 32)
                      for v, c in enumerate(cams)])
    vals = torch.arange(keys.numel(), device=hip.dev, dtype=torch.int32)
    ms_sort = cold_ms(lambda: hip.radix_sort(keys, vals, key_bits=32 + view_bits, radix_bits=8))
    b_sort = sort_bytes(keys.numel(), 32 + view_bits, 8)
    out["hbm_targets"] = {
        "preprocess": {"ms": round(ms_pre, 4), "algorithmic_bytes": int(b_pre),
                       "gb_s": round(b_pre / ms_pre / 1e6, 1), "frac": round(b_pre / ms_pre / 1e6 / HBM_PEAK_GBS, 4)},
        "depth_sort": {"count": int(keys.numel()), "key_bits": 32 + view_bits, "radix_bits": 8, "ms": round(ms_sort, 4),
                       "algorithmic_bytes": int(b_sort), "gb_s": round(b_sort / ms_sort / 1e6, 1),
                       "frac": round(b_sort / ms_sort / 1e6 / HBM_PEAK_GBS, 4)}}
    del p, pre1, keys, vals
    torch.cuda.empty_cache()
    return out


def bench_eval3d(hip, scene, cam, cfg, reps=5):
    """The eval_3D forward (AAA-Gaussians 3D evaluation with the 16-entry
    per-pixel k-buffer) on the headline scene and view, timed on the host
    around synchronized runs (the pair count is read back, as the reference's
    binner does): `ms` with the fast binning (gs_bin_tiles_3d: direct placement
    + per-tile sort), `ms_stage_path` with the reference's stage chain (scan,
    duplicate, 64-bit pair sort); and each stage's device time alone."""
    import torch
    p = hip.upload_scene(scene)
    out = hip.render_3d(p, cam, cfg, raw=True)  # warm-up: allocations, code load
    hip.sync()
    K = int(out["K"])
    pre, pre3d = out["pre"], out["pre3d"]
    bg = tuple(cfg.bg_color)
    img, T, nc = out["img"], out["T_final"], out["n_contrib"]

    def host_ms(f):
        f()
        hip.sync()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            f()
            hip.sync()
            ts.append(time.perf_counter() - t0)
        return 1e3 * sorted(ts)[len(ts) // 2]

    ms_stage = host_ms(lambda: hip.render_3d(p, cam, cfg, raw=True))
    vo = torch.empty(K + 4, dtype=torch.int32, device=hip.dev)
    ranges = torch.empty(cam.num_tiles * 2, dtype=torch.int32, device=hip.dev)

    def fast():
        hip.preprocess_3d(p, cam, cfg, True, None, pre, pre3d)
        vs, _, rg, _ = hip.bin_tiles_3d(pre, pre3d, cam, keys=False, out=(vo, None, ranges))
        hip.raster_fwd_3d(pre, pre3d, vs, rg, cam, bg, img, T, nc)

    ms = host_ms(fast)
    st = {}
    st["preprocess"] = median_ms(lambda: hip.preprocess_3d(p, cam, cfg, True, None, pre, pre3d))
    st["bin_fast"] = median_ms(lambda: hip.bin_tiles_3d(pre, pre3d, cam, keys=False, out=(vo, None, ranges)))
    offsets, _ = hip.scan_tiles(pre["tiles"])
    st["scan"] = median_ms(lambda: hip.scan_tiles(pre["tiles"]))
    st["duplicate"] = median_ms(lambda: hip.duplicate_3d(pre, pre3d, offsets, K, cam))
    st["sort"] = median_ms(lambda: hip.sort_pairs(out["keys"], out["values"], cam.num_tiles))
    st["raster"] = median_ms(lambda: hip.raster_fwd_3d(pre, pre3d, out["values_sorted"], out["ranges"], cam, bg, img,
                                                       T, nc), reps=3)
    npix = cam.width * cam.height
    res = {"workload": f"{scene.n} Gaussians SH{scene.sh_degree}, {cam.width}x{cam.height}, eval_3D",
           "pairs": K, "ms": round(ms, 3), "fwd_mpix_s": round(npix / (ms * 1e-3) / 1e6, 1),
           "ms_stage_path": round(ms_stage, 3),
           "stages_ms": {k: round(v, 4) for k, v in st.items()},
           "mean_blended_per_pixel": round(float(nc.float().mean().item()), 2)}
    del out, pre, pre3d, p, vo
    torch.cuda.empty_cache()
    return res


def raster_work(hip, tr, scene, cam, cfg):
    """The reference loops' pixel-Gaussian evaluations on the trainer's current
    parameters and the headline view (gs_raster_work on a forward with the
    complete pair list)."""
    from gs_scenes import RawScene
    import torch
    cur = RawScene(*tr.params(), scene.sh_degree)
    p = hip.upload_scene(cur)
    f = hip.forward_fast(p, cam, cfg)
    w = hip.raster_work(f["pre"], f["values_sorted"], f["ranges"], cam.width, cam.height, f["px"])
    del f, p
    torch.cuda.empty_cache()
    return w


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or the platform string."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or platform.machine()


def cpu_baseline(scene, cam, cfg, gt, threads=1):
    """The reference's own Trainer::step (compiled from source into
    oracle/_ref) on the same workload: one full step on one core; with
    threads > 1, also `threads` independent steps at once, one per core (ctypes
    releases the interpreter lock during the call; the reference keeps no
    global state), as the host's all-core training throughput.  Falls back to
    the oracle (C restatement, forward + backward only) if that library is
    absent."""
    from concurrent.futures import ThreadPoolExecutor
    from oracle.oracle import Oracle, RefLib
    tc = dict(lr_pos_init=1e-3, lr_pos_final=1e-3, lr_feature=1e-3, lr_opacity=1e-3, lr_scaling=1e-3,
              lr_rotation=1e-3, max_steps=30000, beta1=0.9, beta2=0.999, eps=1e-8)
    if RefLib.available():
        ref = RefLib()
        one = lambda _=None: ref.train_steps(scene, cam, cfg, gt, tc, 0, 1)
        kind = "reference"
        what = "full Trainer::step (reference CPU classes, Adam) at 1M Gaussians, 1920x1080"
    else:
        orc = Oracle()
        one = lambda _=None: orc.forward_backward(scene, cam, cfg, gt)
        kind = "port"
        what = "forward+backward of the C oracle at 1M Gaussians, 1920x1080 (no optimizer)"
    t0 = time.perf_counter()
    one()
    dt = time.perf_counter() - t0
    out = {"value": round(1.0 / dt, 5), "unit": "train steps/s", "cores": 1, "kind": kind, "sample": "one " + what,
           "seconds": round(dt, 2), "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}
    if threads > 1:
        t0 = time.perf_counter()
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(one, range(threads)))
        dtn = time.perf_counter() - t0
        out["all_cores"] = {"value": round(threads / dtn, 5), "unit": "train steps/s", "cores": threads,
                            "sample": f"{threads} independent {what}, one per core, at once",
                            "seconds": round(dtn, 2)}
    out["reference_opencl"] = reference_opencl_status()
    return out


def reference_opencl_status():
    """Whether the reference's OpenCL kernels could be timed on a CPU OpenCL
    device here: that needs the reference tree (its .cl files, named by
    REFERENCE_PATH, never copied into this repository) and a CPU OpenCL
    platform such as PoCL.  Reports which is missing; nothing is loaded."""
    import glob
    ref = os.environ.get("REFERENCE_PATH", "")
    has_ref = bool(ref) and os.path.isdir(os.path.join(ref, "harmonyos_3dgs", "src", "gpu", "kernels"))
    icds = [os.path.basename(f) for f in glob.glob("/etc/OpenCL/vendors/*.icd")]
    has_cpu_cl = any("pocl" in f.lower() for f in icds)
    if has_ref and has_cpu_cl:
        return "available (not run by this benchmark)"
    missing = []
    if not has_ref:
        missing.append("no reference tree (REFERENCE_PATH unset or without the .cl kernels)")
    if not has_cpu_cl:
        missing.append("no CPU OpenCL platform (ICDs: " + (", ".join(icds) or "none") + ")")
    return "reference OpenCL on PoCL: unavailable: " + "; ".join(missing)


def traffic_probe(args):
    """The headline step for hardware-counter runs (rocprofv3 --pmc ...): the
    native trainer with Adam fused, autotuned as in the timed benchmark, then
    two training steps; plus device copies of a known size, whose counter
    values per byte calibrate the stage kernels' values into HBM bytes per
    launch.  Prints the tuning, so that the counters can be matched to the
    kernel instances (LDS batch depths) the benchmark runs."""
    import torch
    from gs_hip import GsHip, GsTrainer, train_cfg
    from gs_scenes import RenderCfg, make_workload

    scene, cams, gts = make_workload(args.workload)
    cam = cams[0]
    cfg = RenderCfg(sh_degree=3, training=True)
    hip = GsHip(0)
    gt = torch.tensor(gts[0], device=hip.dev)
    tr = GsTrainer(hip, scene)
    tr.set_options(fused_adam=True)
    # the autotuner runs every rasterizer instance (LDS batch depth), so the
    # counters cover the instance the timed benchmark picks
    tuning = tr.autotune(cam, cfg, gt, reps=5)
    if args.probe_tuning:
        # fixed knobs for the measured steps: counter passes slow every
        # dispatch down, so the autotuner may pick differently from pass to pass
        hip.apply_tuning({k: int(v) for k, v in (kv.split("=") for kv in args.probe_tuning.split(","))})
        tuning = hip.tuning()
    tc = train_cfg(lr=1e-3)
    for i in range(2):
        r = tr.step(cam, cfg, gt, tc, i)
    torch.cuda.synchronize()
    copy_bytes = 1 << 30
    src = torch.ones(copy_bytes // 4, device=hip.dev)
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"traffic_probe": args.workload, "pairs": int(r.pairs), "tuning": tuning,
                      "copy_bytes": copy_bytes, "copies": 3}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="1m_1080p")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--breakdown-reps", type=int, default=3)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batched-workload", default="1m_16views")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default knobs")
    ap.add_argument("--no-large", action="store_true", help="skip the 6M-Gaussian 8-view batched leg")
    ap.add_argument("--no-fused-adam", action="store_true", help="single-view step with the separate Adam kernel")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="only run each stage kernel and a 1 GiB device copy (for rocprofv3 --pmc runs)")
    ap.add_argument("--probe-tuning", default="",
                    help="with --traffic-probe: knobs for the measured steps, as key=value,key=value (every knob of "
                         "gs_tuning: radix_bits, raster_batch, raster_bwd_batch, bin_group, bin_tiles_per_lane, "
                         "persistent_per_cu, split_groups)")
    ap.add_argument("--large-workload", default="6m_8x1080p")
    ap.add_argument("--no-eval3d", action="store_true", help="skip the eval_3D (k-buffer) forward leg")
    ap.add_argument("--eval3d-only", action="store_true",
                    help="only the eval_3D forward leg on the workload's first view (for profiling)")
    ap.add_argument("--sorts-only", action="store_true",
                    help="only the two sorts of hbm_targets on the workload's first view (for profiling)")
    ap.add_argument("--peaks-only", action="store_true",
                    help="only the measured device ceilings (gs_device_peaks), as one JSON line")
    args = ap.parse_args()

    if args.traffic_probe:
        traffic_probe(args)
        return
    if args.peaks_only:
        from gs_hip import GsHip
        print(json.dumps({"device_peaks": GsHip(0).device_peaks()}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU path is design-only",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from gs_hip import GsHip, GsTrainer, train_cfg
    from gs_scenes import RenderCfg, make_workload

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    t0 = time.perf_counter()
    scene, cams, gts = make_workload(args.workload)
    cam, gt_np = cams[0], gts[0]
    log(f"workload {args.workload}: N={scene.n} {cam.width}x{cam.height} generated in {time.perf_counter() - t0:.1f}s")
    cfg = RenderCfg(sh_degree=3, training=True)
    tc = train_cfg(lr=1e-3)
    hip = GsHip(0)
    if args.eval3d_only:
        print(json.dumps({"eval3d": bench_eval3d(hip, scene, cam, RenderCfg(sh_degree=3, training=False))}),
              flush=True)
        return
    if args.sorts_only:
        print(json.dumps({"hbm_targets": bench_sorts(hip, hip.upload_scene(scene), cam, cfg, 8)}), flush=True)
        return
    gt = torch.tensor(gt_np, device=hip.dev)
    tr = GsTrainer(hip, scene)
    tr.set_options(fused_adam=not args.no_fused_adam)  # the step at this size specifies a fused Adam

    tuning = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        tuning = tr.autotune(cam, cfg, gt, reps=5)
        log(f"autotune ({time.perf_counter() - t0:.1f}s): {tuning}")
    for i in range(args.warmup):
        r = tr.step(cam, cfg, gt, tc, i)
    torch.cuda.synchronize()
    log(f"warmup done: loss {r.loss:.5f} pairs {r.pairs}")
    it = args.warmup
    # per-stage device time of the native step: a HIP event pair around every
    # stage on the library's stream, over a few untimed steps
    tr.set_timing(None)
    for i in range(args.breakdown_reps):
        tr.step(cam, cfg, gt, tc, it)
        it += 1
    stages, _ = tr.stage_times()
    log("stages (ms): " + ", ".join(f"{k} {v:.3f}" for k, v in stages.items()))
    # the dominant kernel: the longest stage that is a single kernel launch
    single = ["preprocess", "raster_fwd", "raster_bwd"] + ([] if args.no_fused_adam else ["backward"])
    dom = max(single, key=lambda k: stages[k])
    # timed region; the dominant kernel alone is bracketed by events in it
    tr.set_timing([dom])
    t0 = time.perf_counter()
    for i in range(args.steps):
        r = tr.step(cam, cfg, gt, tc, it)
        it += 1
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    dom_times, dom_steps = tr.stage_times()
    tr.set_timing([])
    steps_s = args.steps / dt
    K = r.pairs
    log(f"train: {steps_s:.2f} steps/s ({1e3 * dt / args.steps:.3f} ms/step), last loss {r.loss:.5f}; "
        f"{dom} {dom_times[dom]:.4f} ms over {dom_steps} timed steps")

    for i in range(2):
        tr.render(cam, cfg)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        tr.render(cam, cfg)
    torch.cuda.synchronize()
    dtf = time.perf_counter() - t0
    mpix_s = cam.width * cam.height * args.steps / dtf / 1e6
    log(f"forward: {mpix_s:.1f} Mpix/s ({1e3 * dtf / args.steps:.3f} ms/frame)")

    npix = cam.width * cam.height
    kernel_names = {"preprocess": "preprocess_fwd_kernel", "raster_fwd": "raster_fwd_kernel",
                    "raster_bwd": "raster_bwd_kernel", "backward": "preprocess_bwd_multi_kernel (Adam fused)"}
    dom_ms = dom_times[dom]
    byts = algorithmic_bytes(dom, scene.n, K, npix, scene.max_coeffs, cam.num_tiles)
    hbm_gbs = byts / (dom_ms * 1e-3) / 1e9
    st_meas = load_traffic(args.workload).get(dom, {})
    # VALU wave-instructions of the rasterizer instance (LDS batch depth) this run uses
    batch = (tuning or {}).get("raster_batch" if dom == "raster_fwd" else "raster_bwd_batch")
    by_batch = st_meas.get("valu_wave_insts_by_batch", {})
    insts = by_batch.get(str(batch)) or (sorted(by_batch.values())[len(by_batch) // 2] if by_batch else None)
    peaks = hip.device_peaks()
    log(f"measured peaks: {peaks}")
    if dom in ("raster_fwd", "raster_bwd"):
        # algorithmic work of this frame, counted on the device (outside the timed region)
        work = raster_work(hip, tr, scene, cam, cfg)
        log(f"raster work: {work}")
        if dom == "raster_fwd":
            evals = work["fwd_evals"]
            flops = FLOPS_PER_EVAL * evals + FLOPS_FWD_CONTRIB * work["contributors"]
        else:
            evals = work["bwd_evals"]
            flops = FLOPS_PER_EVAL * evals + FLOPS_BWD_CONTRIB * work["contributors"]
        sec = dom_ms * 1e-3
        tflops = flops / sec / 1e12
        lds_gbs = evals * LDS_BYTES_PER_EVAL / sec / 1e9
        roofline = {"kernel": kernel_names[dom], "bound": "valu", "achieved": round(tflops, 2),
                    "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 4),
                    "traffic": st_meas.get("traffic_bytes"), "ms": round(dom_ms, 4),
                    "algorithmic_flops": int(flops), "evaluations": int(evals),
                    "contributors": int(work["contributors"]),
                    "flops_per_evaluation": FLOPS_PER_EVAL,
                    "flops_per_contributor": FLOPS_FWD_CONTRIB if dom == "raster_fwd" else FLOPS_BWD_CONTRIB,
                    # the same work against the FP32 rate these kernels can reach: one
                    # unfused, unpacked FP32 operation per lane and instruction at the
                    # measured issue rate (no FMA contraction, no packed FP32)
                    "measured_fp32_unfused_tflops": round(peaks["valu_wave_insts_per_s"] * 64 / 1e12, 2),
                    "lds": {"achieved_gbs": round(lds_gbs, 1), "peak_gbs_measured": peaks["lds_read_gbs"],
                            "frac": round(lds_gbs / peaks["lds_read_gbs"], 4) if peaks["lds_read_gbs"] else None,
                            "bytes_per_evaluation": LDS_BYTES_PER_EVAL},
                    "hbm": {"achieved_gbs": round(hbm_gbs, 1), "peak_gbs": HBM_PEAK_GBS,
                            "frac": round(hbm_gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": int(byts)}}
        if insts:
            # secondary: VALU issue occupancy (instructions executed, from the committed
            # rocprofv3 SQ_INSTS_VALU summary, per second against 4 cycles per
            # wave64 instruction); a measure of how busy the SIMDs are, not of work done
            ach = insts / sec / 1e9
            roofline["valu_issue"] = {"achieved": round(ach, 1), "peak": round(VALU_WAVE_INSTS_PER_S / 1e9, 1),
                                      "unit": "G VALU wave-instructions/s",
                                      "frac": round(ach / (VALU_WAVE_INSTS_PER_S / 1e9), 4),
                                      "valu_wave_insts": insts}
    else:
        roofline = {"kernel": kernel_names[dom], "bound": "hbm", "achieved": round(hbm_gbs, 1), "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": round(hbm_gbs / HBM_PEAK_GBS, 4), "traffic": st_meas.get("traffic_bytes"),
                    "algorithmic_bytes": int(byts), "ms": round(dom_ms, 4)}

    # the bandwidth targets: preprocess (from the stage timing) and the two sorts
    p = hip.upload_scene(scene)
    rb = (tuning or {}).get("radix_bits") or 8
    hbm_targets = bench_sorts(hip, p, cam, cfg, rb)
    # preprocess on its own, its 284 MB read cold as inside the step (repeated
    # back to back, the 236 MB of parameters would come from the Infinity
    # Cache)
    pre = hip.alloc_pre(scene.n)
    import ctypes
    g_c, c_c, f_c, p_c = hip._gauss(p, True), cam.to_c(), cfg.to_c(), hip._pre_struct(pre)
    pre_args = (hip.ctx, ctypes.byref(g_c), ctypes.byref(c_c), ctypes.byref(f_c), ctypes.byref(p_c))
    hip._check(hip.lib.gs_preprocess_fwd(*pre_args), "gs_preprocess_fwd")
    ms_pre = cold_ms(lambda: hip.lib.gs_preprocess_fwd(*pre_args), reps=7)
    pb = algorithmic_bytes("preprocess", scene.n, K, npix, scene.max_coeffs, cam.num_tiles)
    pg = pb / (ms_pre * 1e-3) / 1e9
    hbm_targets["preprocess"] = {"ms": round(ms_pre, 4), "algorithmic_bytes": int(pb),
                                 "gb_s": round(pg, 1), "frac": round(pg / HBM_PEAK_GBS, 4)}
    del pre
    del p
    log(f"hbm targets: {hbm_targets}")

    eval3d = None
    if not args.no_eval3d:
        eval3d = bench_eval3d(hip, scene, cam, RenderCfg(sh_degree=3, training=False))
        log(f"eval3d: {eval3d}")

    batched = None
    if not args.no_batched:
        batched = bench_batched(hip, cfg, args)
        # The batch takes one Adam update per batch; the plain path one per view.
        # Like for like: charge the batch the per-Gaussian backward + Adam sweeps
        # it skips, at the single-view step's measured time for that stage
        # (same 1M Gaussians).
        batched["adam_updates_per_batch"] = 1
        batched["plain_adam_updates"] = batched["views"]
        bwd_ms = stages.get("backward")
        if bwd_ms:
            ms_eq = batched["ms_per_batch_step"] + (batched["views"] - 1) * bwd_ms
            batched["batched_vs_single_same_updates"] = round(
                batched["views"] / (ms_eq * 1e-3) / batched["plain_views_per_s"], 3)
        log(f"batched: {batched}")

    large = None
    if not args.no_large:
        large = bench_large(hip, cfg, args)
        log(f"large: {large}")

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(scene, cam, cfg, gt_np, threads=min(16, os.cpu_count() or 1))
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": "train steps/s + fwd Mpix/s, 1M Gaussians 1080p; batched vs single-view",
        "value": round(steps_s, 3),
        "unit": "train steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * dt / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded Philox scene and target image)",
        "config": {"workload": args.workload, "gaussians": scene.n, "sh_degree": scene.sh_degree,
                   "width": cam.width, "height": cam.height, "pairs": int(K),
                   "optimizer": "adam lr 1e-3" + ("" if args.no_fused_adam else " (fused into the backward)")},
        "fwd_mpix_s": round(mpix_s, 2),
        "batched": batched,
        "large": large,
        "eval3d": eval3d,
        "stages_ms": {k: round(v, 4) for k, v in stages.items()},
        "tuning": tuning,
        "roofline": roofline,
        "device_peaks": peaks,
        "hbm_targets": hbm_targets,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
