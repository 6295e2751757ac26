"""Drop-in for the reference's ``post`` module (cpp/post/post.cu:905-933).

``CUDA_double_array`` / ``CUDA_float_array`` own device copies of numpy
arrays; their logical shape is updated by ``swap_axis``, ``swap_axis_back``
and ``CUDA_copy_double`` as in the reference (post.cu:612-614,663-665,702-704).
``free_array`` is idempotent. Every function runs HIP kernels.
"""
from __future__ import annotations

import ctypes

import numpy as np

from cbmv_amd import _abi
from cbmv_amd.compat._util import arr, ptr


class _DeviceArray:
    _dtype = None
    _code = None

    def __init__(self, a):
        a = arr(a, self._dtype, None, type(self).__name__)
        if not 1 <= a.ndim <= 4:
            raise ValueError(f"{type(self).__name__}: 1 to 4 dimensions supported")
        shape = (ctypes.c_int64 * a.ndim)(*a.shape)
        self._h = _abi.check_ptr(_abi.lib.cbmv_array_create(self._code, a.ndim, shape, ptr(a)),
                                 type(self).__name__)

    @property
    def shape(self):
        n = _abi.lib.cbmv_array_ndims(self._h)
        s = (ctypes.c_int64 * n)()
        _abi.check(_abi.lib.cbmv_array_shape(self._h, s))
        return tuple(s)

    def set_to_zero(self):
        _abi.check(_abi.lib.cbmv_array_set_to_zero(self._h), "set_to_zero")

    def get_array(self):
        out = np.empty(self.shape, self._dtype)
        _abi.check(_abi.lib.cbmv_array_get(self._h, ptr(out)), "get_array")
        return out

    def free_array(self):
        _abi.check(_abi.lib.cbmv_array_free(self._h), "free_array")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            _abi.lib.cbmv_array_destroy(h)
            self._h = None


class CUDA_double_array(_DeviceArray):
    _dtype = np.float64
    _code = 1


class CUDA_float_array(_DeviceArray):
    _dtype = np.float32
    _code = 0


def InitCUDA():
    ok = bool(_abi.lib.cbmv_init_device())
    if ok:
        print("HIP device initialized.")
    return ok


def sgm(fl, fr, din, dout, dtmp, pi1, pi2, tau_so, alpha1, sgm_q1, sgm_q2, direction):
    _abi.check(_abi.lib.cbmv_post_sgm(fl._h, fr._h, din._h, dout._h, dtmp._h, pi1, pi2, tau_so,
                                      alpha1, sgm_q1, sgm_q2, int(direction)), "sgm")


def cross(fin, fout, L1, tau1):
    _abi.check(_abi.lib.cbmv_post_cross(fin._h, fout._h, int(L1), float(tau1)), "cross")


def cbca(fleftc, frightc, din, dout, direction):
    _abi.check(_abi.lib.cbmv_post_cbca(fleftc._h, frightc._h, din._h, dout._h, int(direction)), "cbca")


def subpixel_enchancement(fd, dcost, fout):
    _abi.check(_abi.lib.cbmv_post_subpixel(fd._h, dcost._h, fout._h), "subpixel_enchancement")


def median2d(fin, fout, ksize):
    _abi.check(_abi.lib.cbmv_post_median2d(fin._h, fout._h, int(ksize)), "median2d")


def mean2d(fin, fout, kernel, alpha2):
    k = arr(kernel, np.float32, 2, "mean2d kernel")
    _abi.check(_abi.lib.cbmv_post_mean2d(fin._h, fout._h, ptr(k), k.shape[0], k.shape[1], float(alpha2)),
               "mean2d")


def swap_axis(din, dout):
    _abi.check(_abi.lib.cbmv_post_swap_axis(din._h, dout._h), "swap_axis")


def swap_axis_back(din, dout):
    _abi.check(_abi.lib.cbmv_post_swap_axis_back(din._h, dout._h), "swap_axis_back")


def CUDA_copy_double(din, dout):
    _abi.check(_abi.lib.cbmv_post_copy(din._h, dout._h), "CUDA_copy_double")


def CUDA_copy_float(fin, fout):
    _abi.check(_abi.lib.cbmv_post_copy(fin._h, fout._h), "CUDA_copy_float")


def CUDA_divide_double(din, val):
    _abi.check(_abi.lib.cbmv_post_divide(din._h, float(val)), "CUDA_divide_double")


def outlier_detection(fd0, fd1, fout, disp_max):
    _abi.check(_abi.lib.cbmv_post_outlier_detection(fd0._h, fd1._h, fout._h, int(disp_max)),
               "outlier_detection")


def interpolate_mismatch(fd0, foutlier, fout):
    _abi.check(_abi.lib.cbmv_post_interpolate_mismatch(fd0._h, foutlier._h, fout._h), "interpolate_mismatch")


def interpolate_occlusion(fd0, foutlier, fout):
    _abi.check(_abi.lib.cbmv_post_interpolate_occlusion(fd0._h, foutlier._h, fout._h), "interpolate_occlusion")
