"""The order in which numpy reduces a contiguous float32 array, which the
image normalisation (test.py:489-492: np.mean, np.std) depends on and the
pipeline's normalisation kernels reproduce:

* the array is summed in chunks of numpy's ufunc buffer (8192 elements),
  chunk sums added in order from 0;
* each chunk by pairwise summation: halves rounded down to a multiple of 8
  until at most 128 elements remain; a block of 8..128 with 8 interleaved
  accumulators combined ((r0+r1)+(r2+r3))+((r4+r5)+(r6+r7)) then its tail
  in order; fewer than 8 in order from 0;
* mean and variance divide in double and round to float32; std is the
  float32 square root of the variance of the float32 squares.

The second check restates the kernel's parallel form of one chunk (each of
64 lanes descends 6 levels by the bits of its index, sums the node it
reaches, and the lanes combine up the tree) and holds it to numpy too.
"""
import numpy as np
import pytest

f32 = np.float32


def _block(a, lo, n):
    if n < 8:
        r = f32(0)
        for i in range(n):
            r = f32(r + a[lo + i])
        return r
    r = [f32(a[lo + j]) for j in range(8)]
    i = 8
    while i < n - n % 8:
        for j in range(8):
            r[j] = f32(r[j] + a[lo + i + j])
        i += 8
    res = f32(f32(f32(r[0] + r[1]) + f32(r[2] + r[3])) + f32(f32(r[4] + r[5]) + f32(r[6] + r[7])))
    for k in range(i, n):
        res = f32(res + a[lo + k])
    return res


def _half(n):
    return n // 2 - (n // 2) % 8


def _pairwise(a, lo, n):
    if n <= 128:
        return _block(a, lo, n)
    h = _half(n)
    return f32(_pairwise(a, lo, h) + _pairwise(a, lo + h, n - h))


def _lanes(a, lo, m):
    """The kernel's form: lane descent by index bits, then a tree combine."""
    v = [f32(0)] * 64
    split = [0] * 64
    for lane in range(64):
        s, n, alive, sp = lo, m, True, 0
        for k in range(6):
            bit = (lane >> (5 - k)) & 1
            if n > 128:
                sp |= 1 << k
                h = _half(n)
                s, n = (s + h, n - h) if bit else (s, h)
            elif bit:
                alive = False
        if alive:
            if n > 128:
                h = _half(n)
                assert n - h <= 128
                v[lane] = f32(_block(a, s, h) + _block(a, s + h, n - h))
            else:
                v[lane] = _block(a, s, n)
        split[lane] = sp
    for k in range(5, -1, -1):
        st = 1 << (5 - k)
        v = [f32(v[l] + v[l + st]) if (l & (2 * st - 1)) == 0 and (split[l] >> k) & 1 else v[l]
             for l in range(64)]
    return v[0]


def _sum(a, chunk_sum):
    s = f32(0)
    for c in range(0, a.size, 8192):
        s = f32(s + chunk_sum(a, c, min(8192, a.size - c)))
    return s


def _mean_std(img, chunk_sum):
    a = img.astype(f32).ravel()
    mean = f32(np.float64(_sum(a, chunk_sum)) / a.size)
    x = (a - mean).astype(f32)
    x = (x * x).astype(f32)
    var = f32(np.float64(_sum(x, chunk_sum)) / a.size)
    return mean, f32(np.sqrt(np.float64(var)))


SIZES = [(1, 5), (2, 9), (16, 16), (16, 23), (91, 90), (92, 90), (64, 128), (77, 211)]


@pytest.mark.parametrize("h,w", SIZES)
@pytest.mark.parametrize("form", ["recursive", "lanes"])
def test_reduction_order_equals_numpy(h, w, form):
    rng = np.random.default_rng(h * 1000 + w)
    chunk_sum = _pairwise if form == "recursive" else _lanes
    for _ in range(3):
        img = rng.integers(0, 256, (h, w)).astype(np.uint8)
        a = img.astype(f32)
        mean, std = _mean_std(img, chunk_sum)
        assert mean == np.mean(a) and std == np.std(a)
        norm = ((a - mean) / std).astype(f32)
        np.testing.assert_array_equal(norm, (a - np.mean(a)) / np.std(a))


def test_one_pass_would_differ():
    """The chunking matters: summing a large image as one pairwise tree (or
    row by row) gives a different float32 mean than numpy for this image."""
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (554, 694)).astype(np.uint8)
    a = img.astype(f32).ravel()
    whole = f32(np.float64(f32(_pairwise(a, 0, a.size))) / a.size)
    assert _mean_std(img, _pairwise)[0] == np.mean(img.astype(f32))
    rows = f32(0)
    for r in range(554):
        rows = f32(rows + _pairwise(a, r * 694, 694))
    assert whole != np.mean(img.astype(f32)) or f32(np.float64(rows) / a.size) != np.mean(img.astype(f32))
