"""Pins the oracle's anisotropy fields by the mathematics.

The field routine (hamiltonianactions.f90:842-919) and the MC energy
(montecarlo_common.f90:534-569) are separate restatements in the oracle.  A
field is minus the gradient of its energy, so along any direction t tangent to
the moment, dE/dh = -m B.t.  That holds in the reference for the cubic term
(K1 and K2) and the uniaxial K1 term; the reference's own uniaxial K2 field
(-2t(K1 + 2K2(1-t^2)) e) and its type-7 cubic energy sign do not follow from
its MC energy, so those are pinned by composition instead: the type-7 field is
the uniaxial field plus sb times the cubic field (hamiltonianactions.f90:232-238).
A swapped component, a dropped term or a wrong sign in either restatement
fails these checks.
"""
import ctypes

import numpy as np

MUB = 9.274009994e-24


def _ham(O, taniso, eaniso, kaniso, sb):
    N = len(taniso)
    nl = np.zeros((N, 1), dtype=np.int32)
    ns = np.zeros(N, dtype=np.int32)
    nc = np.zeros((N, 1))
    return O.OracleHam(N, nl, ns, nc, taniso=np.asarray(taniso, dtype=np.int32),
                       eaniso=np.ascontiguousarray(eaniso), kaniso=np.ascontiguousarray(kaniso),
                       sb=np.ascontiguousarray(sb))


def _setup(rng, N):
    e = rng.normal(size=(N, 3))
    e /= np.linalg.norm(e, axis=1, keepdims=True)
    ea = rng.normal(size=(N, 3))
    ea /= np.linalg.norm(ea, axis=1, keepdims=True)
    m = rng.uniform(0.5, 2.5, size=N)
    return e, ea, m


def _grad_check(O, ham, e, m, i, rng, h=1e-5):
    N = e.shape[0]
    emom = np.ascontiguousarray(e[None])
    mm = np.ascontiguousarray(m[None])
    emomM = np.ascontiguousarray(emom * mm[..., None])
    beff, _ = ham.field(emomM)
    B = beff[0, i]
    L = O.lib()
    zero = np.zeros(3)
    for _ in range(3):
        t = rng.normal(size=3)
        t -= t.dot(e[i]) * e[i]
        t /= np.linalg.norm(t)
        d = []
        for s in (+1.0, -1.0):
            n = e[i] + s * h * t
            n /= np.linalg.norm(n)
            d.append(L.or_mc_energy_change(ctypes.byref(ham.h), N, emomM, emom, mm, i, 0,
                                           np.ascontiguousarray(n), zero))
        dEdh = (d[0] - d[1]) / (2 * h) / MUB
        want = -m[i] * B.dot(t)
        assert abs(dEdh - want) <= 1e-6 * max(abs(want), np.abs(B).max() * m[i] * 1e-3), (i, dEdh, want)


def test_cubic_field_is_gradient_of_mc_energy(O):
    rng = np.random.default_rng(60)
    N = 6
    e, ea, m = _setup(rng, N)
    kan = np.stack([rng.uniform(-0.02, 0.02, N), rng.uniform(-0.02, 0.02, N)], axis=1)
    ham = _ham(O, [2] * N, ea, kan, np.zeros(N))
    for i in range(N):
        _grad_check(O, ham, e, m, i, rng)


def test_uniaxial_k1_field_is_gradient_of_mc_energy(O):
    rng = np.random.default_rng(61)
    N = 6
    e, ea, m = _setup(rng, N)
    kan = np.stack([rng.uniform(-0.02, 0.02, N), np.zeros(N)], axis=1)
    ham = _ham(O, [1] * N, ea, kan, np.zeros(N))
    for i in range(N):
        _grad_check(O, ham, e, m, i, rng)


def test_type7_field_is_uniaxial_plus_sb_cubic(O):
    rng = np.random.default_rng(62)
    N = 6
    e, ea, m = _setup(rng, N)
    kan = np.stack([rng.uniform(-0.02, 0.02, N), rng.uniform(-0.02, 0.02, N)], axis=1)
    sb = rng.uniform(0.1, 2.0, N)
    emomM = np.ascontiguousarray((e * m[:, None])[None])
    b7, _ = _ham(O, [7] * N, ea, kan, sb).field(emomM)
    b1, _ = _ham(O, [1] * N, ea, kan, sb).field(emomM)
    b2, _ = _ham(O, [2] * N, ea, kan, sb).field(emomM)
    assert np.allclose(b7, b1 + sb[None, :, None] * b2, rtol=1e-14, atol=1e-14 * np.abs(b7).max())


def test_uniaxial_k2_field_closed_form(O):
    """K1 = 0: B = -4 K2 t (1 - t^2) e with t = m.e (the reference's field
    convention, hamiltonianactions.f90:858-876), checked at t = 0 (no field),
    t = +-|m| (no field) and a generic angle."""
    ea = np.array([[0.0, 0.0, 1.0]] * 3)
    m = np.array([1.0, 1.0, 1.0])
    e = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 1.0], [0.6, 0.0, 0.8]])
    K2 = 0.01
    kan = np.array([[0.0, K2]] * 3)
    b, _ = _ham(O, [1] * 3, ea, kan, np.zeros(3)).field(np.ascontiguousarray((e * m[:, None])[None]))
    assert np.allclose(b[0, 0], 0.0, atol=1e-18)
    assert np.allclose(b[0, 1], 0.0, atol=1e-18)
    assert np.allclose(b[0, 2], [0.0, 0.0, -4 * K2 * 0.8 * (1 - 0.64)], rtol=1e-14)
