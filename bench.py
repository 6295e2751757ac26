#!/usr/bin/env python3
"""End-to-end benchmark of the CBMV cost-volume path on one MI355X.

Workload: one seeded slanted-plane stereo pair, 1500 x 1000, D = 256
(subpixel ground truth, right-image gain/bias, noise sigma 3), through the
whole path on the GPU: four matchers, per-pixel statistics, 20 features per
voxel, the classifier, and the LR-check post-processing chain with 8-path SGM
and a 5 x 5 median. A step is one pair end to end, from u8 images resident in
device memory to the f32 disparity map in device memory.

Three classifier paths are measured: the gated MLP (20-128-64-1, GELU,
bf16 MFMA; pixels whose best match probability is below 0.5 fall back to the
forest), the same network with ReLU hidden units (a faster variant trained
on the same samples) and the plain 50-tree random forest. ``value`` is the
GELU MLP path's W*H*D/s when a trained MLP is present, else the forest
path's. Both paths' throughput and bad-2.0 are in
the JSON line, with per-stage times, the dominant kernel's roofline and the
CPU baseline (the oracle's C++/OpenMP restatement of the reference on a
bounded sample of the same workload, on this host's cores).

python bench.py [--gpus 1] [--steps K] [--warmup W]
                [--config 1500x1000x256|2960x1988x512|batch64|694x554x128]

--config 2960x1988x512 runs the same line on the 2960 x 1988, D = 512 full-resolution
stress pair (3.0e9 voxels; every volume of the path resident in HBM).
--config batch64 runs a batch of 64 seeded pairs of uneven size (seeds 1000-1063,
W ~ U{320..1500}, H ~ U{240..1000}, D in {64, 128, 192, 256}) as per-pair
hipGraphs on --streams streams with the tile width autotuned per (W, D)
bucket: aggregate W*H*D/s and per-pair p50/p99 latency (a step = the batch),
with the classifier kernel's roofline over every eighth pair and the CPU
baseline on a strip of the largest of those.
--config 694x554x128 runs the reference's own configuration (694 x 554, D = 128
fronto-parallel pair, seed 0; forest path, `_mem` chain with 4-path SGM) and
times the CPU path on the whole pair beside it (median of 3 after a warm-up).
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

MODELS = os.path.join(ROOT, "cbmv_amd", "models")
RF_MODEL = os.path.join(MODELS, "rf_fronto50.npz")
MLP_MODEL = os.path.join(MODELS, "mlp_slanted.npz")            # GELU, as specified
MLP_RELU_MODEL = os.path.join(MODELS, "mlp_slanted_relu.npz")  # ReLU variant
HBM_PEAK = 8.0e12          # B/s, MI355X HBM3E
BF16_PEAK = 2.5e15         # FLOP/s, dense
MLP_FLOP_PER_VOXEL = 2 * (20 * 128 + 128 * 64 + 64)   # 21,632


def log(*a):
    print(*a, file=sys.stderr, flush=True)


class Events:
    def __init__(self):
        from cbmv_amd import _abi
        self.abi = _abi
        self.a = _abi.check_ptr(_abi.lib.cbmv_event_create(), "event")
        self.b = _abi.check_ptr(_abi.lib.cbmv_event_create(), "event")

    def start(self, stream=None):
        self.abi.check(self.abi.lib.cbmv_event_record(self.a, stream))

    def stop(self, stream=None):
        self.abi.check(self.abi.lib.cbmv_event_record(self.b, stream))

    def ms(self):
        v = ctypes.c_float()
        self.abi.check(self.abi.lib.cbmv_event_elapsed_ms(self.a, self.b, ctypes.byref(v)))
        return v.value


def run_path(name, pair, rf, mlp_model, steps, warmup, params):
    from cbmv_amd import _abi, metrics
    from cbmv_amd.pipeline import DeviceBuffer, Pipeline

    H, W = pair.left.shape
    D = pair.ndisp
    cls = {"rf": 0, "mlp": 1, "mlp_gated": 2, "mlp_gated_relu": 2}[name]
    pl = Pipeline(H, W, D, forest=rf, mlp=mlp_model, classifier=cls, **params)
    dL, dR = DeviceBuffer.from_array(pair.left), DeviceBuffer.from_array(pair.right)
    dO = DeviceBuffer(H * W * 4)
    for _ in range(warmup):
        pl.run(dL.ptr, dR.ptr, dO.ptr)
    _abi.check(_abi.lib.cbmv_device_synchronize())
    ev = Events()
    t0 = time.perf_counter()
    ev.start()
    for _ in range(steps):
        pl.run(dL.ptr, dR.ptr, dO.ptr)
    ev.stop()
    ms_ev = ev.ms()
    _abi.check(_abi.lib.cbmv_device_synchronize())
    wall = (time.perf_counter() - t0) * 1000.0
    ms = max(ms_ev, wall) / steps
    disp = dO.to_array((H, W), np.float32)
    bad = metrics.bad_tau(disp, pair.gt, 2.0, pair.occluded)
    # per-stage times (one more pass, each stage bracketed by events)
    stages = {}
    for st in ("matchers", "stats", "classifier", "post"):
        e = Events()
        e.start()
        pl.run_stage(st, dL.ptr, dR.ptr, dO.ptr)
        e.stop()
        stages[st] = round(e.ms(), 3)
    voxels = H * W * D
    res = {"ms_per_step": round(ms, 3), "voxels_per_s": voxels / (ms / 1000.0),
           "bad2_all": round(bad["all"], 3), "bad2_nonocc": round(bad["nonocc"], 3),
           "stage_ms": stages}
    if cls == 2:
        n_gated = int(pl.buffer("gate")[H * W])
        res["gated_pixel_fraction"] = round(n_gated / (H * W), 5)
    del pl
    return res, disp


def kernel_ms(head, pair, rf, mlp_model, params, reps=5):
    """Average duration of the classifier's main kernel, from HIP events on the
    library stream around ``reps`` launches of the classifier stage of an
    ungated pipeline (there the stage is that one kernel: k_mlp_classify or
    k_features_rf)."""
    from cbmv_amd import _abi
    from cbmv_amd.pipeline import DeviceBuffer, Pipeline

    H, W = pair.left.shape
    cls = 0 if head == "rf" else 1
    pl = Pipeline(H, W, pair.ndisp, forest=rf, mlp=mlp_model, classifier=cls, **params)
    dL, dR = DeviceBuffer.from_array(pair.left), DeviceBuffer.from_array(pair.right)
    pl.run_stage("matchers", dL.ptr, dR.ptr)
    pl.run_stage("stats")
    pl.run_stage("classifier")
    _abi.check(_abi.lib.cbmv_device_synchronize())
    ev = Events()
    ev.start()
    for _ in range(reps):
        pl.run_stage("classifier")
    ev.stop()
    ms = ev.ms() / reps
    del pl
    return ms


# HBM bytes per launch of the dominant kernel at 1500x1000x256, from two
# rocprofv3 --pmc runs (FETCH_SIZE, then WRITE_SIZE) of
# "bench.py --steps 1 --warmup 0 --paths mlp_gated,mlp_gated_relu --no-cpu-baseline --copy-mib 1024",
# each counter scaled by its reading for the 1 GiB k_copy<double> of the same
# run (FETCH_SIZE read 524,316 KiB for 1,048,576 KiB fetched: x2.0; WRITE_SIZE
# read 1,048,576 KiB: x1.0). The profile summaries under profiles/ hold them.
TRAFFIC = {
    "k_mlp_classify": (5055198.2 * 2.0 + 1506234.4) * 1024,
}


def cpu_baseline(pair, rf_dict, rows=16, params=None):
    """The oracle (C++/OpenMP restatement of the reference) on a strip of the pair."""
    import oracle

    Hh, Ww = pair.left.shape
    y0 = Hh // 2
    L = np.ascontiguousarray(pair.left[y0:y0 + rows])
    R = np.ascontiguousarray(pair.right[y0:y0 + rows])
    D = pair.ndisp
    t0 = time.perf_counter()
    vols = oracle.get_costs(L, R, D)
    feats = oracle.test_features(vols)
    cost = oracle.forest_predict(feats, rf_dict).reshape(rows, L.shape[1], D)
    del feats
    oracle.post_interp(oracle.normalise(L), oracle.normalise(R), cost, npaths=8,
                       median_w=params.get("median_w", 5))
    dt = time.perf_counter() - t0
    vox = rows * L.shape[1] * D
    return {"value": vox / dt, "unit": "voxels/s", "cores": oracle.num_threads(), "kind": "port",
            "cpu": _cpu_model(), "host_cpus": os.cpu_count(),
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "sample": f"{rows}-row strip (rows {y0}-{y0 + rows - 1}) of the {Ww}x{Hh} D={D} pair, "
                      f"whole path on the CPU: four matchers, 20 features, the 50-tree forest, "
                      f"LR-check post chain with 8-path SGM; {dt:.1f} s"}


def reference_path_c0(pair, rf_dict):
    """The reference's whole path on the CPU (the oracle's C++/OpenMP
    restatement): matchers, batched features and the forest's float64 cost
    (test.py:597-626, main.py:70), then the `_mem` post chain (test.py:153-338)."""
    import oracle

    vols = oracle.get_costs(pair.left, pair.right, pair.ndisp)
    H, W, D = vols[0].shape
    r_s = oracle.batch_rows(W)
    cost = np.empty((H, W, D))
    for b in range(0, H, r_s):
        feats = oracle.features_lr(*(v[b:b + r_s] for v in vols))
        cost[b:b + r_s] = oracle.forest_predict(feats, rf_dict).reshape(-1, W, D)
        del feats
    del vols
    disp, _ = oracle.post_mem(oracle.normalise(pair.left), oracle.normalise(pair.right), cost)
    return disp


def run_c0(a):
    """The reference's configuration: one 694 x 554, D = 128 fronto-parallel
    pair (seed 0) through the forest path with the `_mem` chain (4-path SGM,
    3 x 3 median): the GPU's W*H*D/s, and the CPU path's (one warm-up, then
    the median of --cpu-runs runs of the whole pair) with the agreement of
    the two disparity maps."""
    from cbmv_amd import forest, metrics, synth

    pair = synth.fronto_pair(0)
    H, W = pair.left.shape
    D = pair.ndisp
    rf_dict = forest.load(RF_MODEL)
    rf = forest.Forest(rf_dict)
    params = {}  # the reference's Testing defaults
    log("running rf path ...")
    res, disp = run_path("rf", pair, rf, None, a.steps, a.warmup, params)
    log(f"  rf: {res}")
    kms = kernel_ms("rf", pair, rf, None, params)
    V = H * W * D
    ach = 23.0 * V / (kms / 1000.0)
    roof = {"kernel": "k_features_rf (features + 50-tree traversal)", "bound": "hbm",
            "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": ach / HBM_PEAK, "traffic": None, "kernel_ms": round(kms, 3)}
    cpu = None
    agree = None
    if not a.no_cpu_baseline:
        import oracle
        times = []
        for i in range(1 + a.cpu_runs):
            t0 = time.perf_counter()
            cpu_disp = reference_path_c0(pair, rf_dict)
            dt = time.perf_counter() - t0
            log(f"  CPU path run {i}{' (warm-up)' if i == 0 else ''}: {dt:.1f} s")
            if i:
                times.append(dt)
        med = float(np.median(times))
        agree = metrics.agreement(disp, cpu_disp, tol=0.05)
        cpu = {"value": V / med, "unit": "voxels/s", "cores": oracle.num_threads(), "kind": "port",
               "cpu": _cpu_model(), "host_cpus": os.cpu_count(),
               "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
               "sample": f"the whole {W}x{H} D={D} pair, median of {len(times)} runs after one "
                         f"warm-up ({', '.join(f'{t:.1f}' for t in times)} s): four matchers, "
                         "batched features, the 50-tree forest (float64 cost), _mem post chain"}
    line = {
        "metric": "end-to-end W*H*D/s + bad-2.0 %, MLP vs RF path on 1 MI355X",
        "value": res["voxels_per_s"],
        "unit": "voxels/s (W*H*D/s)",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/i32 matchers, f64 features, f32-compare forest, f32 post",
        "data": "synthetic (seeded fronto-parallel pair, seed 0)",
        "config": {"workload": f"fronto-parallel pair {W}x{H}, D={D}, the reference's _mem chain "
                               "(4-path SGM, 3x3 median)", "W": W, "H": H, "D": D, "path": "rf",
                   "forest": "50 trees, max depth 20, min_samples_leaf 700, fitted on four "
                             "694x554 D=128 fronto-parallel pairs (seeds 1-4)"},
        "paths": {"rf": res},
        "agreement_with_cpu_path": agree,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _render_batch_pair(seed):
    from cbmv_amd import synth
    return synth.batch_pair(seed)


def run_batch(a, rf, rf_dict, models, params):
    """The batched workload: 64 pairs through per-pair graphs on a few streams."""
    import multiprocessing as mp

    from cbmv_amd import metrics
    from cbmv_amd.batch import Batch

    seeds = list(range(1000, 1064))
    t0 = time.time()
    with mp.get_context("spawn").Pool(16) as pool:
        pairs = pool.map(_render_batch_pair, seeds, chunksize=1)
    log(f"generated {len(pairs)} pairs in {time.time() - t0:.1f}s")
    sizes = [(p.left.shape[0], p.left.shape[1], p.ndisp) for p in pairs]
    vox = sum(h * w * d for h, w, d in sizes)
    head = "mlp_gated" if models.get("mlp_gated") is not None else "rf"
    cls = 2 if head == "mlp_gated" else 0
    out = {}
    for ns in (a.streams, 1):
        b = Batch(sizes, forest=rf, mlp=models.get(head), n_streams=ns, classifier=cls, **params)
        for i, p in enumerate(pairs):
            b.set_pair(i, p.left, p.right)
        tiles = b.tune(reps=2)
        b.capture()
        for _ in range(a.warmup):
            b.run()
        runs = [b.run() for _ in range(a.steps)]
        ex = np.concatenate([r["exec_ms"] for r in runs])
        done = np.concatenate([r["done_ms"] for r in runs])
        tot = float(np.mean([r["total_ms"] for r in runs]))
        bad = [metrics.bad_tau(b.disparity(i), p.gt, 2.0, p.occluded) for i, p in enumerate(pairs)]
        out[ns] = {"streams": ns, "ms_per_batch": round(tot, 3), "voxels_per_s": vox / (tot / 1000.0),
                   "pair_exec_ms": {"p50": round(float(np.percentile(ex, 50)), 3),
                                    "p99": round(float(np.percentile(ex, 99)), 3)},
                   "pair_done_ms": {"p50": round(float(np.percentile(done, 50)), 3),
                                    "p99": round(float(np.percentile(done, 99)), 3)},
                   "bad2_all_mean": round(float(np.mean([x["all"] for x in bad])), 3),
                   "bad2_nonocc_mean": round(float(np.mean([x["nonocc"] for x in bad])), 3),
                   "tiles": {str(t): int((tiles == t).sum()) for t in (4, 8, 16)}}
        log(f"  {ns} stream(s): {out[ns]}")
        del b
    r = out[a.streams]
    # roofline of the classifier kernel over a sample of the batch (every
    # eighth pair: all four D values), each pair's kernel timed with events
    sample = list(range(0, len(pairs), 8))
    kms = [kernel_ms(head, pairs[i], rf, models.get(head), params) for i in sample]
    svox = [sizes[i][0] * sizes[i][1] * sizes[i][2] for i in sample]
    t = sum(kms) / 1000.0
    if head.startswith("mlp"):
        ach = MLP_FLOP_PER_VOXEL * sum(svox) / t
        roof = {"kernel": f"k_mlp_classify<{models[head]['activation']}>", "bound": "mfma",
                "achieved": ach / 1e12, "peak": BF16_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": ach / BF16_PEAK, "traffic": None}
    else:
        ach = 23.0 * sum(svox) / t
        roof = {"kernel": "k_features_rf", "bound": "hbm", "achieved": ach / 1e9,
                "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK, "traffic": None}
    roof["sample"] = (f"{len(sample)} of the {len(pairs)} pairs (seeds "
                      f"{', '.join(str(seeds[i]) for i in sample)}), "
                      f"{sum(kms):.2f} ms of kernel time for {sum(svox):.3g} voxels")
    cpu = None
    if not a.no_cpu_baseline:
        i = max(sample, key=lambda k: sizes[k][2] * sizes[k][1])
        log("timing the CPU baseline ...")
        cpu = cpu_baseline(pairs[i], rf_dict, rows=a.cpu_rows // 4, params=params)
    line = {
        "metric": "end-to-end W*H*D/s + bad-2.0 %, MLP vs RF path on 1 MI355X",
        "value": r["voxels_per_s"],
        "unit": "voxels/s (W*H*D/s)",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_batch"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/i32 matchers, f64 features, bf16/f16 MFMA MLP (f16 activations) or f32-compare forest, f32 post",
        "data": "synthetic (64 seeded slanted-plane pairs, seeds 1000-1063)",
        "config": {"workload": "batch of 64 pairs, W~U{320..1500}, H~U{240..1000}, "
                               "D in {64,128,192,256}, per-pair hipGraphs, LR-check chain, "
                               "8-path SGM, median 5x5", "pairs": len(pairs), "voxels": vox,
                   "path": head},
        "batch": r,
        "single_stream": out[1],
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--paths", default="auto", help="auto | rf | mlp_gated | mlp | comma list")
    ap.add_argument("--config", default="1500x1000x256",
                    choices=["1500x1000x256", "batch64", "2960x1988x512", "694x554x128"],
                    help="1500x1000x256: the slanted pair of the headline line; batch64: a "
                         "batch of 64 pairs of uneven size; 2960x1988x512: the full-resolution "
                         "stress pair; 694x554x128: the reference's configuration (forest "
                         "path, _mem chain) with the CPU path timed beside it")
    ap.add_argument("--streams", type=int, default=2, help="batch64: HIP streams of the batch")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-rows", type=int, default=128)
    ap.add_argument("--cpu-runs", type=int, default=3,
                    help="694x554x128: timed runs of the CPU path (after one warm-up)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first copy this many MiB device-to-device with the library's copy "
                         "kernel (k_copy), the reference of known size for HBM counters")
    a = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if a.gpus != 1 or world != 1:
        sys.exit("bench.py measures one GPU: every box here has one MI355X and this path runs "
                 "one pair per GPU (multi-GPU sharding is described, not run). Use --gpus 1.")

    from cbmv_amd import forest, mlp, synth

    if a.copy_mib:
        from cbmv_amd.compat import post
        n = a.copy_mib * (1 << 20) // 8
        src = post.CUDA_double_array(np.ones(n))
        dst = post.CUDA_double_array(np.zeros(n))
        post.CUDA_copy_double(src, dst)     # reads and writes n*8 bytes
        log(f"copied {a.copy_mib} MiB device-to-device (k_copy<double>)")
        del src, dst
    if a.config == "694x554x128":
        run_c0(a)
        return
    if a.config == "batch64":
        rf_dict = forest.load(RF_MODEL)
        models = {"mlp_gated": mlp.load(MLP_MODEL) if os.path.exists(MLP_MODEL) else None}
        if a.paths not in ("auto", "mlp_gated"):
            models = {}
        run_batch(a, forest.Forest(rf_dict), rf_dict, models,
                  dict(lr_check=1, sgm_paths=8, median_w=5))
        return
    t0 = time.time()
    pair = synth.slanted_pair(100) if a.config == "1500x1000x256" else synth.stress_pair(3000)
    H, W = pair.left.shape
    D = pair.ndisp
    log(f"generated the {W}x{H} D={D} pair in {time.time() - t0:.1f}s")
    rf_dict = forest.load(RF_MODEL)
    rf = forest.Forest(rf_dict)
    models = {"mlp_gated": mlp.load(MLP_MODEL) if os.path.exists(MLP_MODEL) else None,
              "mlp_gated_relu": mlp.load(MLP_RELU_MODEL) if os.path.exists(MLP_RELU_MODEL) else None}
    models["mlp"] = models["mlp_gated"]
    params = dict(lr_check=1, sgm_paths=8, median_w=5)
    if a.paths == "auto":
        paths = [p for p in ("mlp_gated", "mlp_gated_relu") if models[p] is not None] + ["rf"]
    else:
        paths = a.paths.split(",")
    results = {}
    for p in paths:
        log(f"running {p} path ...")
        results[p], _ = run_path(p, pair, rf, models.get(p), a.steps, a.warmup, params)
        log(f"  {p}: {results[p]}")
    head = paths[0]
    r = results[head]
    V = H * W * D
    # roofline of the dominant kernel of the headline path, timed on its own
    kms = kernel_ms(head, pair, rf, models.get(head), params)
    t = kms / 1000.0
    if head.startswith("mlp"):
        ach = MLP_FLOP_PER_VOXEL * V / t
        roof = {"kernel": f"k_mlp_classify<{models[head]['activation']}> (features + 3-layer MLP: bf16 and f16 MFMA, f16 activations)",
                "bound": "mfma",
                "achieved": ach / 1e12, "peak": BF16_PEAK / 1e12, "unit": "TFLOP/s",
                "frac": ach / BF16_PEAK,
                "traffic": TRAFFIC["k_mlp_classify"] if a.config == "1500x1000x256" else None,
                "kernel_ms": round(kms, 3)}
    else:
        # algorithmic bytes of the feature + forest kernel: 19 B of matcher
        # volumes read + 4 B of cost written per voxel
        ach = 23.0 * V / t
        roof = {"kernel": "k_features_rf (features + 50-tree traversal)", "bound": "hbm",
                "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": ach / HBM_PEAK, "traffic": None, "kernel_ms": round(kms, 3)}
    cpu = None
    if not a.no_cpu_baseline and a.config == "1500x1000x256":
        log("timing the CPU baseline ...")
        cpu = cpu_baseline(pair, rf_dict, rows=a.cpu_rows, params=params)
    line = {
        "metric": "end-to-end W*H*D/s + bad-2.0 %, MLP vs RF path on 1 MI355X",
        "value": r["voxels_per_s"],
        "unit": "voxels/s (W*H*D/s)",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/i32 matchers, f64 features, bf16/f16 MFMA MLP (f16 activations) or f32-compare forest, f32 post",
        "data": f"synthetic (seeded slanted-plane pair, seed {pair.seed})",
        "config": {"workload": f"slanted-plane pair {W}x{H}, D={D}, LR-check chain, 8-path SGM, "
                               "median 5x5", "W": W, "H": H, "D": D, "path": head,
                   "forest": "50 trees, max depth 20, min_samples_leaf 700, fitted on four "
                             "694x554 D=128 fronto-parallel pairs (seeds 1-4)"},
        "paths": results,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    if "rf" in results and head != "rf":
        line["mlp_over_rf"] = results["rf"]["ms_per_step"] / r["ms_per_step"]
        line["bad2_mlp_minus_rf"] = round(r["bad2_all"] - results["rf"]["bad2_all"], 3)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
