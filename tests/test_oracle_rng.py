"""The oracle's restatements of the LLG thermal deviates of the HIP kernels:
the FP32 engine's single-precision Box-Muller (its polynomial ln, sin and
cos against numpy's double precision functions) and the FP64 engine's
double-precision Box-Muller on 52-bit uniforms; both against N(0,1)."""
import math

import numpy as np


def test_bm_ln_accuracy(O):
    L = O.lib()
    rng = np.random.default_rng(3)
    u = np.concatenate([rng.random(20000), np.exp(-rng.random(5000) * 23.0),
                        np.float32([2.0 ** -33, 2.0 ** -32, 0.5, 1.0, 1.41421354, 0.70710677])])
    u = u.astype(np.float32)
    got = np.array([L.or_bm_ln(float(x)) for x in u], np.float64)
    want = np.log(u.astype(np.float64))
    err = np.abs(got - want)
    assert np.all(err <= 3e-7 * np.maximum(np.abs(want), 1e-30) + 1e-12), err.max()
    assert got[-3] == 0.0 or abs(got[-3]) < 1e-30     # ln(1) = 0 exactly


def test_bm_sincos_accuracy(O):
    L = O.lib()
    k = np.arange(0, 1 << 23, 997, dtype=np.int64)
    v = ((2 * k + 1).astype(np.float32) * np.float32(2.0 ** -22)).astype(np.float32)
    out = np.zeros(2, np.float32)
    s, c = np.zeros(len(v)), np.zeros(len(v))
    for i, x in enumerate(v):
        L.or_bm_sincos(float(x), out)
        s[i], c[i] = out
    th = np.pi / 2 * v.astype(np.float64)
    assert np.max(np.abs(s - np.sin(th))) < 3e-7
    assert np.max(np.abs(c - np.cos(th))) < 3e-7
    assert np.max(np.abs(s * s + c * c - 1)) < 1e-6


def test_thermal3_distribution(O):
    n = 60000
    z = np.array([O.thermal3(0xC0FFEE, 0x11, 7, i, 0) for i in range(n)], np.float64).ravel()
    se = 1.0 / math.sqrt(z.size)
    assert abs(z.mean()) < 5 * se
    assert abs(z.var() - 1.0) < 5 * math.sqrt(2.0) * se
    assert abs(np.mean(z ** 4) - 3.0) < 5 * math.sqrt(96.0) * se
    # tail fractions against the normal CDF
    for t in (1.0, 2.0, 3.0):
        p = math.erfc(t / math.sqrt(2.0))
        assert abs(np.mean(np.abs(z) > t) - p) < 5 * math.sqrt(p * (1 - p) / z.size)
    # components are uncorrelated
    a = z.reshape(n, 3)
    cc = np.corrcoef(a.T)
    assert np.max(np.abs(cc - np.eye(3))) < 5 / math.sqrt(n)


def _normal_checks(z):
    se = 1.0 / math.sqrt(z.size)
    assert abs(z.mean()) < 5 * se
    assert abs(z.var() - 1.0) < 5 * math.sqrt(2.0) * se
    assert abs(np.mean(z ** 4) - 3.0) < 5 * math.sqrt(96.0) * se
    for t in (1.0, 2.0, 3.0):
        p = math.erfc(t / math.sqrt(2.0))
        assert abs(np.mean(np.abs(z) > t) - p) < 5 * math.sqrt(p * (1 - p) / z.size)


def test_thermal3_f64_distribution_and_resolution(O):
    """FP64 deviates: N(0,1) moments and tails, uncorrelated components,
    and values that single precision cannot hold (the FP32 deviates are all
    float values), with distinct low-order bits."""
    n = 60000
    z = np.array([O.thermal3_f64(0xC0FFEE, 7, i, 0) for i in range(n)]).ravel()
    _normal_checks(z)
    cc = np.corrcoef(z.reshape(n, 3).T)
    assert np.max(np.abs(cc - np.eye(3))) < 5 / math.sqrt(n)
    assert np.mean(z.astype(np.float32).astype(np.float64) != z) > 0.999
    low = (z.view(np.uint64) & np.uint64(0xFFFF)).astype(np.int64)
    assert len(np.unique(low)) > 0.6 * 65536
    # the two Philox blocks are independent streams: no component repeats
    assert len(np.unique(z)) == z.size
