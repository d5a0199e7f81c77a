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
