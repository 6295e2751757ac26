"""The fused device pipeline (cbmv_pipeline in include/cbmv_abi.h).

``Pipeline(H, W, D, forest=..., **params)`` allocates every device buffer
for one pair size; ``run`` takes device pointers (u8 images, f32 disparity
output) and a HIP stream, ``run_host`` is a convenience that copies numpy
images in and the disparity map out. Parameters are the reference's
``Testing`` defaults (test.py:44-47) unless overridden, plus:

* ``sgm_paths`` 4 (reference) or 8 (adds the diagonals),
* ``lr_check`` 0 = ``__postprocessing_mem`` chain, 1 = ``_mem_interp`` chain,
* ``classifier`` 0 = random forest, 1 = MLP, 2 = MLP gated to the forest.
"""
from __future__ import annotations

import ctypes

import numpy as np

from cbmv_amd import _abi
from cbmv_amd.compat._util import ptr

STAGES = {"matchers": 0, "stats": 1, "classifier": 2, "post": 3}

# index, dtype, shape(H, W, D)
BUFFERS = {
    "census": (0, np.uint8, lambda H, W, D: (H, D, W)),
    "ncc": (1, np.float64, lambda H, W, D: (H, D, W)),
    "sob": (2, np.uint16, lambda H, W, D: (H, D, W)),
    "sad": (3, np.float64, lambda H, W, D: (H, D, W)),
    "stats_left": (4, np.float64, lambda H, W, D: (8, H, W)),
    "stats_right": (5, np.float64, lambda H, W, D: (8, H, W)),
    "cost": (6, np.float32, lambda H, W, D: (H, W, D)),
    "wta_left": (7, np.float32, lambda H, W, D: (H, W)),
    "wta_right": (8, np.float32, lambda H, W, D: (H, W)),
    "outlier": (9, np.float32, lambda H, W, D: (H, W)),
    "pmax": (10, np.float32, lambda H, W, D: (H, W)),
    "gate": (11, np.int32, lambda H, W, D: (H * W + 1,)),
    "cost_right": (12, np.float32, lambda H, W, D: (H, W, D)),
    "agg_left": (13, np.float32, lambda H, W, D: (H, W, D)),
    "img_left": (14, np.float32, lambda H, W, D: (H, W)),
    "img_right": (15, np.float32, lambda H, W, D: (H, W)),
}


class DeviceBuffer:
    """A raw device allocation (cbmv_device_alloc)."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        self.ptr = _abi.check_ptr(_abi.lib.cbmv_device_alloc(self.nbytes), "device_alloc")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _abi.check(_abi.lib.cbmv_memcpy_h2d(b.ptr, ptr(a), a.nbytes), "h2d")
        return b

    def to_array(self, shape, dtype):
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("buffer smaller than the requested array")
        _abi.check(_abi.lib.cbmv_memcpy_d2h(ptr(out), self.ptr, out.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            _abi.lib.cbmv_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        self.free()


def make_params(**kw):
    p = _abi.default_params()
    names = {f[0] for f in _abi.Params._fields_}
    for k, v in kw.items():
        if k not in names:
            raise TypeError(f"unknown pipeline parameter {k!r}")
        setattr(p, k, v)
    return p


class Pipeline:
    def __init__(self, H, W, D, forest=None, mlp=None, **params):
        self.H, self.W, self.D = int(H), int(W), int(D)
        self.params = make_params(**params)
        self._h = _abi.check_ptr(
            _abi.lib.cbmv_pipeline_create(self.H, self.W, self.D, ctypes.byref(self.params)),
            "pipeline_create")
        if forest is not None:
            self.set_forest(forest)
        if mlp is not None:
            self.set_mlp(mlp)

    def set_forest(self, forest):
        _abi.check(_abi.lib.cbmv_pipeline_set_forest(self._h, forest.handle), "set_forest")
        self._forest = forest

    def set_mlp(self, mlp):
        """mlp: dict with mean, inv_std, w1, b1, w2, b2, w3, b3 (float32) and
        optionally activation ("gelu", the default, or "relu")."""
        from cbmv_amd.mlp import activation_code
        arrs = [np.ascontiguousarray(mlp[k], dtype=np.float32)
                for k in ("mean", "inv_std", "w1", "b1", "w2", "b2", "w3", "b3")]
        act = activation_code(mlp.get("activation", "gelu"))
        _abi.check(_abi.lib.cbmv_pipeline_set_mlp(self._h, *[ptr(a) for a in arrs], act), "set_mlp")

    def reshape(self, H, W, D):
        """Point the pipeline at a pair of size H x W x D, each at most the
        size it was created with (no reallocation)."""
        _abi.check(_abi.lib.cbmv_pipeline_reshape(self._h, int(H), int(W), int(D)), "reshape")
        self.H, self.W, self.D = int(H), int(W), int(D)

    def set_sgm_pairs(self, on):
        """One SGM sweep per path (False) or paired sweeps (True, the default
        of an 8-path LR-check pipeline with D <= 256)."""
        _abi.check(_abi.lib.cbmv_pipeline_set_sgm_pairs(self._h, int(bool(on))), "set_sgm_pairs")
        return self

    def run(self, left_ptr, right_ptr, disp_ptr, stream=None):
        _abi.check(_abi.lib.cbmv_pipeline_run(self._h, left_ptr, right_ptr, disp_ptr, stream), "pipeline_run")

    def run_stage(self, stage, left_ptr=None, right_ptr=None, disp_ptr=None, stream=None):
        st = STAGES[stage] if isinstance(stage, str) else int(stage)
        _abi.check(_abi.lib.cbmv_pipeline_run_stage(self._h, st, left_ptr, right_ptr, disp_ptr, stream),
                   "pipeline_run_stage")

    def run_host(self, left, right):
        """numpy u8 images in, float32 disparity out (synchronous)."""
        L = np.ascontiguousarray(left, dtype=np.uint8)
        R = np.ascontiguousarray(right, dtype=np.uint8)
        if L.shape != (self.H, self.W) or R.shape != (self.H, self.W):
            raise ValueError(f"images must be {self.H}x{self.W}")
        dL, dR = DeviceBuffer.from_array(L), DeviceBuffer.from_array(R)
        dO = DeviceBuffer(self.H * self.W * 4)
        self.run(dL.ptr, dR.ptr, dO.ptr)
        self._last_inputs = (dL, dR)
        return dO.to_array((self.H, self.W), np.float32)

    def buffer(self, name, rows=None):
        """Host copy of an internal buffer (BUFFERS); rows=(y0, y1) copies only
        that range of the leading (row) axis of an [H, ...] buffer."""
        idx, dtype, shape = BUFFERS[name]
        p = ctypes.c_void_p()
        n = ctypes.c_int64()
        _abi.check(_abi.lib.cbmv_pipeline_buffer(self._h, idx, ctypes.byref(p), ctypes.byref(n)), "buffer")
        shp = shape(self.H, self.W, self.D)
        if np.prod(shp, dtype=np.int64) * np.dtype(dtype).itemsize > n.value:
            raise ValueError(f"buffer {name} not allocated for this configuration")
        src = p.value
        if rows is not None:
            y0, y1 = int(rows[0]), int(rows[1])
            if shp[0] != self.H or not 0 <= y0 < y1 <= self.H:
                raise ValueError(f"rows {rows} do not select rows of buffer {name}")
            row_bytes = int(np.prod(shp[1:], dtype=np.int64)) * np.dtype(dtype).itemsize
            src += y0 * row_bytes
            shp = (y1 - y0,) + tuple(shp[1:])
        out = np.empty(shp, dtype)
        _abi.check(_abi.lib.cbmv_memcpy_d2h(ptr(out), ctypes.c_void_p(src), out.nbytes), "d2h")
        return out

    def sgm_timeouts(self):
        """Waits between workgroups of the 8-path SGM sweeps that gave up
        (synchronises the device; nonzero means invalid disparities)."""
        n = ctypes.c_int64()
        _abi.check(_abi.lib.cbmv_pipeline_sgm_timeouts(self._h, ctypes.byref(n)), "sgm_timeouts")
        return int(n.value)

    def row_features(self, y):
        """float32 [D, W, 20] features of row y (after the stats stage)."""
        buf = DeviceBuffer(self.D * self.W * 20 * 4)
        _abi.check(_abi.lib.cbmv_pipeline_row_features(self._h, int(y), buf.ptr, None), "row_features")
        return buf.to_array((self.D, self.W, 20), np.float32)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _abi.lib.cbmv_pipeline_destroy(h)
            self._h = None
