"""Train the matching MLP on pairs the pipeline generates itself.

``python -m cbmv_amd.train_mlp --out model.npz`` renders slanted-plane pairs
(1500 x 1000, D = 256; seeds 2000.. by default, disjoint from the test seed
100; optionally also full-resolution 2960 x 1988, D = 512 pairs, seeds
4000.., disjoint from the stress test seed 3000, and pairs of the
uneven-batch generator, seeds 5000.., disjoint from the batch seeds
1000-1063) on the host CPUs in parallel, runs the pipeline's matcher and statistics
stages on the GPU for each, samples voxels 1:1 (positive iff |d - gt| <= 1,
non-occluded pixels) and trains the network with the GPU forward/backward
kernels and Adam. The result is saved with ``cbmv_amd.mlp.save``. The
network uses GELU hidden units; ``--out-relu`` also trains the ReLU variant
on the same samples.
"""
from __future__ import annotations

import argparse
import multiprocessing as mp
import time

import numpy as np

from cbmv_amd import synth


def _render(kind, seed, H, W, D):
    if kind == "stress":
        return synth.stress_pair(seed)
    if kind == "batch":
        return synth.batch_pair(seed)
    return synth.slanted_pair(seed, H=H, W=W, ndisp=D)


def render_pairs(jobs, H=1000, W=1500, D=256, workers=16):
    """Pairs of (kind, seed) jobs: "slanted" (H x W, D), "stress" (the
    2960 x 1988, D = 512 generator with up to 40 planes) or "batch" (the
    uneven-batch generator's size for that seed)."""
    ctx = mp.get_context("spawn")
    with ctx.Pool(max(1, min(workers, len(jobs)))) as pool:
        for p in pool.starmap(_render, [(k, s, H, W, D) for k, s in jobs], chunksize=1):
            yield p


def main(argv=None):
    from cbmv_amd import mlp

    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--out", required=True, help="GELU model (.npz)")
    ap.add_argument("--out-relu", default=None, help="also train the ReLU variant into this file")
    ap.add_argument("--first-seed", type=int, default=2000)
    ap.add_argument("--pairs", type=int, default=200)
    ap.add_argument("--per-pair", type=int, default=20000)
    ap.add_argument("--epochs", type=int, default=6)
    ap.add_argument("--batch", type=int, default=4096)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--height", type=int, default=1000)
    ap.add_argument("--width", type=int, default=1500)
    ap.add_argument("--ndisp", type=int, default=256)
    ap.add_argument("--stress-pairs", type=int, default=0,
                    help="also this many 2960x1988, D=512 pairs (seeds from --stress-first-seed)")
    ap.add_argument("--stress-first-seed", type=int, default=4000)
    ap.add_argument("--batch-pairs", type=int, default=0,
                    help="also this many pairs of the uneven-batch generator (seeds from "
                         "--batch-first-seed)")
    ap.add_argument("--batch-first-seed", type=int, default=5000)
    ap.add_argument("--occluded-too", action="store_true",
                    help="sample pixels the right view does not see as well")
    a = ap.parse_args(argv)
    jobs = [("slanted", s) for s in range(a.first_seed, a.first_seed + a.pairs)]
    jobs += [("stress", s) for s in range(a.stress_first_seed, a.stress_first_seed + a.stress_pairs)]
    jobs += [("batch", s) for s in range(a.batch_first_seed, a.batch_first_seed + a.batch_pairs)]
    t0 = time.time()
    Xs, ys = [], []
    batch = []
    for k, p in enumerate(render_pairs(jobs, a.height, a.width, a.ndisp, a.workers)):
        batch.append(p)
        if len(batch) == 8 or k == len(jobs) - 1:
            X, y = mlp.training_set(batch, a.per_pair, seed=k, occluded_too=a.occluded_too)
            Xs.append(X)
            ys.append(y)
            batch = []
            print(f"{k + 1}/{len(jobs)} pairs, {sum(x.shape[0] for x in Xs)} samples, "
                  f"{time.time() - t0:.0f}s", flush=True)
    X, y = np.concatenate(Xs), np.concatenate(ys)
    outs = [(a.out, "gelu")] + ([(a.out_relu, "relu")] if a.out_relu else [])
    for path, act in outs:
        model = mlp.train(X, y, epochs=a.epochs, batch=a.batch, lr=a.lr, activation=act)
        mlp.save(path, model)
        print(f"saved {path} ({act}) after {time.time() - t0:.0f}s", flush=True)


if __name__ == "__main__":
    main()
