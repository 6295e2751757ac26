"""Python-3 replay of the reference's test entry point (main.py:49-70) on the
device pipeline, for one rectified pair:

    python -m cbmv_amd.main --l im0.png --r im1.png --d 128 --model rf.npz \\
        --test_set scene --prob_save_path ./results/ --disp_save_path disp.pfm \\
        [--saveCostVolume] [--loadCostVolume] [--interpolate] [--mlp mlp.npz]

Images are read as 8-bit greyscale (scipy.misc.imread(..., mode='L') in
test.py:506-507, PIL's 'L' conversion) or from .npy; the rectification step
is the identity (inputs are rectified). The cost-volume cache follows
main.py:55-67: the file is ``prob_save_path + test_set + ".prob.npy"``
holding predict_proba's float64 [H*W*D, 2] (column 0 = P(no match), the
cost); with --loadCostVolume an existing file replaces the matchers,
features and classifier, with --saveCostVolume a computed one is written.
The cache is read with numpy.load(allow_pickle=False). The disparity map is
written as PFM (pylibs/pfmutil.py conventions). --interpolate selects the
LR-check chain (__postprocessing_mem_interp) instead of __postprocessing_mem,
as do_post's `interpolate` argument does (test.py:497-502).
"""
from __future__ import annotations

import argparse
import os

import numpy as np

from cbmv_amd import _abi
from cbmv_amd.compat import pfmutil
from cbmv_amd.compat._util import ptr


def read_grey(path):
    if path.endswith(".npy"):
        img = np.load(path, allow_pickle=False)
    else:
        from PIL import Image
        with Image.open(path) as im:
            img = np.asarray(im.convert("L"))
    img = np.asarray(img)
    if img.ndim != 2:
        raise ValueError(f"{path}: expected a single-channel image")
    return np.ascontiguousarray(img, dtype=np.uint8)


def cost_file(prob_save_path, test_set):
    """main.py:55: prob_save_path + test_set + '.prob.npy' (plain concatenation)."""
    return prob_save_path + test_set + ".prob.npy"


def proba_from_cost(cost):
    """float64 [H*W*D, 2] from the pipeline's cost volume P(no match)."""
    c = np.asarray(cost, np.float64).reshape(-1)
    return np.stack([c, 1.0 - c], axis=1)


def load_proba(path, n):
    """The cached predict_proba [n, 2] (numpy.load without pickle: a file
    holding Python objects is refused)."""
    proba = np.load(path, allow_pickle=False)
    if proba.shape != (n, 2):
        raise ValueError(f"{path}: expected [{n}, 2], got {proba.shape}")
    return proba


def run(left, right, ndisp, forest=None, mlp=None, prob_file=None, load_cost=False,
        save_cost=False, interpolate=False, **params):
    """Disparity map of one pair; returns (disparity, proba or None, loaded)."""
    from cbmv_amd.pipeline import DeviceBuffer, Pipeline
    H, W = left.shape
    classifier = params.pop("classifier", 2 if mlp is not None else 0)
    pl = Pipeline(H, W, ndisp, forest=forest, mlp=mlp, classifier=classifier,
                  lr_check=1 if interpolate else 0, **params)
    dL, dR = DeviceBuffer.from_array(left), DeviceBuffer.from_array(right)
    dO = DeviceBuffer(H * W * 4)
    loaded = bool(load_cost and prob_file and os.path.exists(prob_file))
    proba = None
    if loaded:
        proba = load_proba(prob_file, H * W * ndisp)
        # eval_prob (test.py:143-149): column 0 as the cost, into the
        # pipeline's cost buffer, then the post-processing stage alone
        cost = np.ascontiguousarray(proba[:, 0].reshape(H, W, ndisp), dtype=np.float32)
        import ctypes
        p = ctypes.c_void_p()
        n = ctypes.c_int64()
        _abi.check(_abi.lib.cbmv_pipeline_buffer(pl._h, 6, ctypes.byref(p), ctypes.byref(n)), "buffer")
        if n.value < cost.nbytes:
            raise ValueError("pipeline cost buffer too small")
        _abi.check(_abi.lib.cbmv_memcpy_h2d(p, ptr(cost), cost.nbytes), "h2d")
        pl.run_stage("post", dL.ptr, dR.ptr, dO.ptr)
    else:
        pl.run(dL.ptr, dR.ptr, dO.ptr)
        if save_cost and prob_file:
            proba = proba_from_cost(pl.buffer("cost"))
            os.makedirs(os.path.dirname(prob_file) or ".", exist_ok=True)
            np.save(prob_file, proba)
    _abi.check(_abi.lib.cbmv_device_synchronize(), "synchronize")
    return dO.to_array((H, W), np.float32), proba, loaded


def main(argv=None):
    ap = argparse.ArgumentParser(description="CBMV on the device pipeline (main.py:49-70).")
    ap.add_argument("--l", dest="lim", required=True)
    ap.add_argument("--r", dest="rim", required=True)
    ap.add_argument("--d", dest="d", type=int, required=True)
    ap.add_argument("--model", dest="model", required=True, help="forest .npz (cbmv_amd.forest)")
    ap.add_argument("--mlp", dest="mlp", default="", help="optional MLP .npz (gated to the forest)")
    ap.add_argument("--test_set", dest="test_set", default="")
    ap.add_argument("--prob_save_path", dest="prob_save_path", default="./results")
    ap.add_argument("--disp_save_path", dest="disp_save_path", default="./results/disp.pfm")
    ap.add_argument("--saveCostVolume", dest="saveCostVolume", action="store_true")
    ap.add_argument("--loadCostVolume", dest="loadCostVolume", action="store_true")
    ap.add_argument("--interpolate", dest="interpolate", action="store_true")
    a = ap.parse_args(argv)
    from cbmv_amd import forest as forest_mod
    from cbmv_amd import mlp as mlp_mod
    left, right = read_grey(a.lim), read_grey(a.rim)
    rf = forest_mod.Forest(forest_mod.load(a.model))
    model = mlp_mod.load(a.mlp) if a.mlp else None
    pf = cost_file(a.prob_save_path, a.test_set)
    disp, _, loaded = run(left, right, a.d, forest=rf, mlp=model, prob_file=pf,
                          load_cost=a.loadCostVolume, save_cost=a.saveCostVolume,
                          interpolate=a.interpolate)
    print(f"{'Loaded cost volume from' if loaded else 'Computed cost volume'} {pf}")
    os.makedirs(os.path.dirname(a.disp_save_path) or ".", exist_ok=True)
    pfmutil.save(a.disp_save_path, disp.astype(np.float32))
    print(f"Saved {a.disp_save_path}")


if __name__ == "__main__":
    main()
