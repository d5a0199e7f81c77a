"""The FP32 reciprocal division of the tile kernel's index arithmetic:
n / d == int((float(n) + 0.5f) * float(1/d)) for 0 <= n < 2^22, every
divisor up to 4096 (numpy float32 rounds each operation as the device does)."""
import numpy as np


def _fdiv(n, d):
    rd = np.float32(1.0) / np.float32(d)
    return ((n.astype(np.float32) + np.float32(0.5)) * rd).astype(np.int64)


def test_fdiv_exact_near_every_multiple():
    for d in range(1, 4097):
        q = np.arange(0, (1 << 22) // d + 1, max(1, ((1 << 22) // d) // 2000), dtype=np.int64)
        n = np.concatenate([q * d - 1, q * d, q * d + 1, q * d + d - 1])
        n = n[(n >= 0) & (n < (1 << 22))]
        assert np.array_equal(_fdiv(n, d), n // d), d


def test_fdiv_exact_small_range_exhaustive():
    n = np.arange(0, 1 << 16, dtype=np.int64)
    for d in range(1, 300):
        assert np.array_equal(_fdiv(n, d), n // d), d


def _fast_div_params(d):
    """FastDiv of uasd_kernels.hpp: l = ceil(log2 d), m = floor(2^(31+l)/d) + 1."""
    l = 0
    while (1 << l) < d:
        l += 1
    return (1 << (31 + l)) // d + 1, 31 + l


def test_fast_index_division_is_exact_below_2_31():
    """q / d as (q * m) >> s (Granlund-Montgomery, N = 31) for every index a
    kernel divides by the ensemble count: m fits in 32 bits for every d up
    to 2^31, and the quotient is exact for q < 2^31 -- exhaustively for
    q < 2^16, d < 300, and at q = k d - 1, k d, k d + 1 across the range for
    every d up to 4096 and for large and power-of-two divisors."""
    for d in list(range(1, 4097)) + [2 ** 20 + 7, 2 ** 30, 2 ** 31 - 1, 2 ** 31, 65535, 65536]:
        m, s = _fast_div_params(d)
        assert 0 < m < 2 ** 32
        k = np.unique(np.concatenate([np.arange(0, 2 ** 31 // d + 1, max(1, (2 ** 31 // d) // 4000)),
                                      [2 ** 31 // d]])).astype(np.uint64)
        q = np.concatenate([k * d - 1, k * d, k * d + 1]).astype(np.int64)
        q = q[(q >= 0) & (q < 2 ** 31)].astype(np.uint64)
        got = (q * np.uint64(m)) >> np.uint64(s)
        assert np.array_equal(got, q // np.uint64(d)), d
    q = np.arange(2 ** 16, dtype=np.uint64)
    for d in range(1, 300):
        m, s = _fast_div_params(d)
        assert np.array_equal((q * np.uint64(m)) >> np.uint64(s), q // np.uint64(d)), d
