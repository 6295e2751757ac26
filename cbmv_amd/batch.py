"""Batches of stereo pairs of different sizes (cbmv_batch in include/cbmv_abi.h).

``Batch(sizes, forest=..., mlp=..., n_streams=2, **params)`` spreads the
pairs (H, W, D) over ``n_streams`` HIP streams, each with one pipeline sized
for its largest pair. ``set_pair`` copies a pair's images into the batch's
device buffers, ``tune`` picks the fused aggregation's tile width per
(W bucket, D), ``capture`` records each pair's path as a hipGraph and
``run`` replays them all, returning per-pair graph times, completion times
and the batch time in milliseconds.
"""
from __future__ import annotations

import ctypes

import numpy as np

from cbmv_amd import _abi
from cbmv_amd.compat._util import ptr
from cbmv_amd.pipeline import make_params


class Batch:
    def __init__(self, sizes, forest=None, mlp=None, n_streams=2, **params):
        sizes = [(int(h), int(w), int(d)) for h, w, d in sizes]
        if not sizes:
            raise ValueError("empty batch")
        self.sizes = sizes
        self.n = len(sizes)
        H = np.array([s[0] for s in sizes], np.int32)
        W = np.array([s[1] for s in sizes], np.int32)
        D = np.array([s[2] for s in sizes], np.int32)
        self.params = make_params(**params)
        self._h = _abi.check_ptr(
            _abi.lib.cbmv_batch_create(self.n, ptr(H), ptr(W), ptr(D), ctypes.byref(self.params),
                                       int(n_streams)), "batch_create")
        self._bufs = []
        for i in range(self.n):
            l, r, o = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
            _abi.check(_abi.lib.cbmv_batch_pair_buffers(self._h, i, ctypes.byref(l), ctypes.byref(r),
                                                        ctypes.byref(o)), "batch_pair_buffers")
            self._bufs.append((l.value, r.value, o.value))
        if forest is not None:
            _abi.check(_abi.lib.cbmv_batch_set_forest(self._h, forest.handle), "batch_set_forest")
            self._forest = forest
        if mlp is not None:
            from cbmv_amd.mlp import activation_code
            arrs = [np.ascontiguousarray(mlp[k], dtype=np.float32)
                    for k in ("mean", "inv_std", "w1", "b1", "w2", "b2", "w3", "b3")]
            _abi.check(_abi.lib.cbmv_batch_set_mlp(self._h, *[ptr(a) for a in arrs],
                                                   activation_code(mlp.get("activation", "gelu"))),
                       "batch_set_mlp")

    def set_pair(self, i, left, right):
        h, w, _ = self.sizes[i]
        L = np.ascontiguousarray(left, dtype=np.uint8)
        R = np.ascontiguousarray(right, dtype=np.uint8)
        if L.shape != (h, w) or R.shape != (h, w):
            raise ValueError(f"pair {i} must be {h}x{w}")
        _abi.check(_abi.lib.cbmv_memcpy_h2d(self._bufs[i][0], ptr(L), L.nbytes), "h2d")
        _abi.check(_abi.lib.cbmv_memcpy_h2d(self._bufs[i][1], ptr(R), R.nbytes), "h2d")

    def tune(self, reps=2):
        tiles = np.zeros(self.n, np.int32)
        _abi.check(_abi.lib.cbmv_batch_tune(self._h, int(reps), ptr(tiles)), "batch_tune")
        return tiles

    def capture(self):
        _abi.check(_abi.lib.cbmv_batch_capture(self._h), "batch_capture")

    def run(self):
        ex = np.zeros(self.n, np.float32)
        done = np.zeros(self.n, np.float32)
        tot = ctypes.c_float()
        _abi.check(_abi.lib.cbmv_batch_run(self._h, ptr(ex), ptr(done), ctypes.byref(tot)), "batch_run")
        return {"exec_ms": ex, "done_ms": done, "total_ms": tot.value}

    def disparity(self, i):
        h, w, _ = self.sizes[i]
        out = np.empty((h, w), np.float32)
        _abi.check(_abi.lib.cbmv_memcpy_d2h(ptr(out), self._bufs[i][2], out.nbytes), "d2h")
        return out

    def stream_of(self, i):
        return int(_abi.lib.cbmv_batch_stream_of(self._h, int(i)))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _abi.lib.cbmv_batch_destroy(h)
            self._h = None
