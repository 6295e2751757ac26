"""ctypes binding of libcbmv_hip.so (declarations in include/cbmv_abi.h).

The library is the only compute path: importing this module raises if it is
missing, and every call raises ``CbmvError`` when the C ABI reports a failure.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcbmv_hip.so")


class CbmvError(RuntimeError):
    pass


class CbmvValueError(CbmvError, ValueError):
    pass


if not os.path.exists(LIB_PATH):
    raise ImportError(
        f"{LIB_PATH} is missing: build it with `python -m cbmv_amd.build` "
        "(there is no CPU fallback)")

lib = ctypes.CDLL(LIB_PATH)

P = ctypes.c_void_p
I32 = ctypes.c_int32
I64 = ctypes.c_int64
U64 = ctypes.c_uint64
F32 = ctypes.c_float
F64 = ctypes.c_double
INT = ctypes.c_int


class Params(ctypes.Structure):
    _fields_ = [
        ("census_w", I32), ("ncc_w", I32), ("sad_w", I32), ("sob_w", I32),
        ("pi1", F32), ("pi2", F32), ("sgm_q1", F32), ("sgm_q2", F32), ("alpha1", F32),
        ("tau_so", F32),
        ("sgm_iters", I32), ("L1", I32),
        ("tau1", F32),
        ("cbca_i1", I32), ("cbca_i2", I32), ("median_iters", I32), ("median_w", I32),
        ("blur_sigma", F32), ("blur_t", F32),
        ("sgm_paths", I32), ("lr_check", I32), ("test_batch", I32), ("classifier", I32),
        ("gate_threshold", F32), ("cbca_tile", I32),
    ]


class CartParams(ctypes.Structure):
    _fields_ = [
        ("n_trees", I32), ("max_features", I32), ("min_samples_leaf", I32), ("max_depth", I32),
        ("bootstrap", I32), ("n_threads", I32), ("seed", U64), ("splitter_state", ctypes.c_uint32),
    ]


_SIGS = {
    "cbmv_last_error": ([], ctypes.c_char_p),
    "cbmv_abi_version": ([], INT),
    "cbmv_device_alloc": ([I64], P),
    "cbmv_device_free": ([P], INT),
    "cbmv_memcpy_h2d": ([P, P, I64], INT),
    "cbmv_memcpy_d2h": ([P, P, I64], INT),
    "cbmv_device_synchronize": ([], INT),
    "cbmv_event_create": ([], P),
    "cbmv_event_record": ([P, P], INT),
    "cbmv_event_elapsed_ms": ([P, P, ctypes.POINTER(F32)], INT),
    "cbmv_event_destroy": ([P], INT),
    "cbmv_census": ([P, P, I64, I64, I32, I32, P], INT),
    "cbmv_census_noalloc": ([P, P, I64, I64, I32, I32, P, P], INT),
    "cbmv_ncc_nister": ([P, P, I64, I64, I32, I32, P], INT),
    "cbmv_sadsob": ([P, P, I64, I64, I32, I32, P], INT),
    "cbmv_zsad": ([P, P, I64, I64, I32, I32, P], INT),
    "cbmv_sobel": ([P, I64, I64, P], INT),
    "cbmv_initthreads": ([], INT),
    "cbmv_swap_axes": ([P, I64, I64, I64, P], INT),
    "cbmv_swap_axes_back": ([P, I64, I64, I64, P], INT),
    "cbmv_get_cost": ([P, I64, I64, I64, P], INT),
    "cbmv_get_right_cost": ([P, I64, I64, I64, P], INT),
    "cbmv_get_left_cost": ([P, I64, I64, I64, P], INT),
    "cbmv_generate_d_indices": ([P, I64, I32, I32, U64, P], INT),
    "cbmv_get_samples": ([P, I64, I64, P, P], INT),
    "cbmv_extract_ratio": ([P, I64, I64, F64, P], INT),
    "cbmv_extract_ratio_samples": ([P, I64, I64, P, F64, P], INT),
    "cbmv_extract_likelihood": ([P, I64, I64, F64, P], INT),
    "cbmv_extract_likelihood_samples": ([P, I64, I64, P, F64, P], INT),
    "cbmv_generate_labels": ([I64, P], INT),
    "cbmv_init_device": ([], INT),
    "cbmv_array_create": ([I32, I32, P, P], P),
    "cbmv_array_dtype": ([P], INT),
    "cbmv_array_ndims": ([P], INT),
    "cbmv_array_shape": ([P, P], INT),
    "cbmv_array_get": ([P, P], INT),
    "cbmv_array_set_to_zero": ([P], INT),
    "cbmv_array_free": ([P], INT),
    "cbmv_array_destroy": ([P], None),
    "cbmv_post_sgm": ([P, P, P, P, P, F32, F32, F32, F32, F32, F32, I32], INT),
    "cbmv_post_cross": ([P, P, I32, F32], INT),
    "cbmv_post_cbca": ([P, P, P, P, I32], INT),
    "cbmv_post_subpixel": ([P, P, P], INT),
    "cbmv_post_median2d": ([P, P, I32], INT),
    "cbmv_post_mean2d": ([P, P, P, I32, I32, F32], INT),
    "cbmv_post_swap_axis": ([P, P], INT),
    "cbmv_post_swap_axis_back": ([P, P], INT),
    "cbmv_post_copy": ([P, P], INT),
    "cbmv_post_divide": ([P, F64], INT),
    "cbmv_post_outlier_detection": ([P, P, P, I32], INT),
    "cbmv_post_interpolate_mismatch": ([P, P, P], INT),
    "cbmv_post_interpolate_occlusion": ([P, P, P], INT),
    "cbmv_forest_create": ([I32, P, I64, P, P, P, P, P, I32], P),
    "cbmv_forest_n_trees": ([P], INT),
    "cbmv_forest_predict_proba": ([P, P, I64, P], INT),
    "cbmv_forest_destroy": ([P], None),
    "cbmv_cart_params_default": ([ctypes.POINTER(CartParams)], None),
    "cbmv_cart_fit": ([P, P, I64, I32, P, ctypes.POINTER(CartParams)], P),
    "cbmv_cart_n_nodes": ([P], I64),
    "cbmv_cart_n_trees": ([P], I32),
    "cbmv_cart_export": ([P] * 7, INT),
    "cbmv_cart_destroy": ([P], None),
    "cbmv_params_default": ([ctypes.POINTER(Params)], None),
    "cbmv_pipeline_create": ([I64, I64, I32, ctypes.POINTER(Params)], P),
    "cbmv_pipeline_set_forest": ([P, P], INT),
    "cbmv_pipeline_set_mlp": ([P, P, P, P, P, P, P, P, P, I32], INT),
    "cbmv_pipeline_reshape": ([P, I64, I64, I32], INT),
    "cbmv_batch_create": ([I32, P, P, P, P, I32], P),
    "cbmv_batch_set_forest": ([P, P], INT),
    "cbmv_batch_set_mlp": ([P, P, P, P, P, P, P, P, P, I32], INT),
    "cbmv_batch_pair_buffers": ([P, I32, P, P, P], INT),
    "cbmv_batch_tune": ([P, I32, P], INT),
    "cbmv_batch_capture": ([P], INT),
    "cbmv_batch_run": ([P, P, P, P], INT),
    "cbmv_batch_stream_of": ([P, I32], I32),
    "cbmv_batch_destroy": ([P], None),
    "cbmv_pipeline_set_cbca_tile": ([P, I32], INT),
    "cbmv_pipeline_set_sgm_pairs": ([P, I32], INT),
    "cbmv_pipeline_run": ([P, P, P, P, P], INT),
    "cbmv_pipeline_run_stage": ([P, I32, P, P, P, P], INT),
    "cbmv_pipeline_buffer": ([P, I32, ctypes.POINTER(P), ctypes.POINTER(I64)], INT),
    "cbmv_pipeline_sgm_timeouts": ([P, P], INT),
    "cbmv_pipeline_row_features": ([P, I64, P, P], INT),
    "cbmv_pipeline_sample_features": ([P, P, I64, P, P], INT),
    "cbmv_pipeline_destroy": ([P], None),
    "cbmv_mlp_trainer_create": ([P] * 8 + [I32], P),
    "cbmv_mlp_trainer_step": ([P, P, P, I64, F32, P, P], INT),
    "cbmv_mlp_trainer_predict": ([P, P, I64, P, P], INT),
    "cbmv_mlp_trainer_gradients": ([P, P, P, I64, P, P], INT),
    "cbmv_mlp_trainer_weights": ([P] * 7, INT),
    "cbmv_mlp_trainer_destroy": ([P], None),
    "cbmv_mfma_probe": ([P, P, P], INT),
    "cbmv_act_probe": ([P, P, I64, I32], INT),
}

for _name, (_args, _res) in _SIGS.items():
    _f = getattr(lib, _name)
    _f.argtypes = _args
    _f.restype = _res

EXPORTED = tuple(_SIGS)


def last_error() -> str:
    msg = lib.cbmv_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str = "") -> int:
    """Raise on a negative status (EINVAL -> ValueError subclass)."""
    if rc < 0:
        msg = f"{what}: {last_error()}" if what else last_error()
        if rc == -1:
            raise CbmvValueError(msg)
        raise CbmvError(msg)
    return rc


def check_ptr(ptr, what: str = ""):
    if not ptr:
        raise CbmvError(f"{what}: {last_error()}")
    return ptr


def default_params() -> Params:
    p = Params()
    lib.cbmv_params_default(ctypes.byref(p))
    return p
