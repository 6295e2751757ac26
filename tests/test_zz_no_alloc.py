"""cbmv_census_noalloc (census on device buffers with its descriptors in a
fixed device array, no allocation) against the census drop-in
(cbmv_census, matchers.cpp:209-350) bit for bit, on two calls in a row,
the second with a smaller input; and its size limit enforced before any
device work."""
import numpy as np
import pytest

D1, D2 = 48, 20


def _noalloc(L, R, D, w):
    from cbmv_amd import _abi
    from cbmv_amd.pipeline import DeviceBuffer
    H, W = L.shape
    dL, dR = DeviceBuffer.from_array(L), DeviceBuffer.from_array(R)
    dO = DeviceBuffer(H * W * D * 8)
    _abi.check(_abi.lib.cbmv_census_noalloc(dL.ptr, dR.ptr, H, W, D, w, dO.ptr, None), "census_noalloc")
    _abi.check(_abi.lib.cbmv_device_synchronize(), "synchronize")
    return dO.to_array((H, W, D), np.float64)


def _pair(seed, H, W, shift):
    rng = np.random.default_rng(seed)
    L = rng.integers(0, 256, (H, W)).astype(np.uint8)
    R = np.clip(np.roll(L.astype(int), -shift, axis=1) + rng.integers(-8, 9, (H, W)), 0, 255)
    return L, R.astype(np.uint8)


@pytest.mark.gpu
def test_noalloc_census_equals_census_on_consecutive_calls():
    from cbmv_amd.compat import libmatchers
    for seed, (H, W, D, w) in enumerate([(1000, 1500, D1, 11), (130, 257, D2, 11), (61, 90, D2, 7)]):
        L, R = _pair(seed, H, W, 5)
        got = _noalloc(L, R, D, w)
        ref = libmatchers.census(L, R, D, w)
        np.testing.assert_array_equal(got, ref)
        assert (got < 2147483647.0).any()


def test_noalloc_census_rejects_sizes_beyond_its_array():
    """Checked before any device call (no GPU needed): 1600 x 1000 pixels at
    two descriptor words each exceed the 3 * 2^20 words per image."""
    from cbmv_amd import _abi
    fake = 256  # never dereferenced: the size check comes first
    rc = _abi.lib.cbmv_census_noalloc(fake, fake, 1000, 1600, 64, 11, fake, None)
    assert rc != 0
    assert b"more than the fixed" in _abi.lib.cbmv_last_error()
