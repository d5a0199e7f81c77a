"""Pins the oracle's semi-implicit midpoint solvers (SDEalgh 6, sibt/sibf;
SDEalgh 7, sispht/sisphf in sphmidpoint.f90) by closed forms.  The reference
has no regression case for them.

A free spin in a constant field B at T = 0 and lam = 0: both halves solve
e - e1 = w x u with w = dtg B (dtg = delta_t*gama*Landeg), so a step is an
exact rotation about B.
- SIB, u = (e1 + e)/2: the Cayley map, angle 2 atan(|w|/2) for any start.
- Spherical, u = (e1 + e)/|e1 + e|: for a start perpendicular to B,
  |e1 + e| = 2 cos(phi/2) and the angle is 2 asin(|w|/2).
Both turn e counter-clockwise about B (de/dt = gama B x e) and keep |e| and
e.B.  A dropped factor 1/2, a swapped cross product or a wrong time scale
fails these.  With damping the spin relaxes towards B.
"""
import ctypes

import numpy as np
import pytest

import refcase as RC



def _free_spins(O, N=3):
    nl = np.zeros((N, 1), dtype=np.int32)
    pairs = dict(nlist=nl, nlistsize=np.zeros(N, dtype=np.int32), ncoup=np.zeros((N, 1)))
    class _Lat:
        natom = N
    return RC.OracleSD(O, _Lat, pairs, np.full(N, 2.0))


def _angle(e, ref):
    """Angle about z from the projection of ref to that of e."""
    return np.arctan2(ref[..., 0] * e[..., 1] - ref[..., 1] * e[..., 0],
                      ref[..., 0] * e[..., 0] + ref[..., 1] * e[..., 1])


@pytest.mark.parametrize("alg", [6, 7])
def test_free_precession_closed_form(O, alg):
    sim = _free_spins(O)
    Bz, dt, n = 30.0, 1e-15, 200
    e0 = np.array([[[1.0, 0.0, 0.0], [0.0, -1.0, 0.0], [np.sqrt(0.5), 0.0, -np.sqrt(0.5)]]])
    sim.set_state(e0)
    ext = np.zeros_like(e0)
    ext[..., 2] = Bz
    for _ in range(n):
        sim.step(alg, dt, 0.0, 0.0, ext=ext)
    gama = ctypes.c_double.in_dll(O.lib(), "OR_GAMA").value   # Parameters/constants.f90
    w = dt * gama * Bz
    phi = 2 * np.arctan(w / 2) if alg == 6 else 2 * np.arcsin(w / 2)
    e = sim.e[0]
    assert np.allclose(np.linalg.norm(e, axis=1), 1.0, atol=1e-11)
    assert np.allclose(e[:, 2], e0[0, :, 2], atol=1e-11)          # e.B kept
    # the in-plane angle advanced by n*phi (perpendicular starts for alg 7)
    rows = [0, 1, 2] if alg == 6 else [0, 1]
    got = np.unwrap(_angle(e[rows], e0[0, rows]))
    want = np.remainder(n * phi + np.pi, 2 * np.pi) - np.pi
    assert np.allclose(np.remainder(got - want + np.pi, 2 * np.pi) - np.pi, 0.0, atol=1e-9)
    # distinct from the other solver's angle by far more than the tolerance
    other = 2 * np.arcsin(w / 2) if alg == 6 else 2 * np.arctan(w / 2)
    assert abs(n * (phi - other)) > 1e-6


@pytest.mark.parametrize("alg", [6, 7])
def test_damping_relaxes_towards_the_field(O, alg):
    sim = _free_spins(O, 1)
    e0 = np.array([[[1.0, 0.0, 0.0]]])
    sim.set_state(e0)
    ext = np.zeros_like(e0)
    ext[..., 2] = 30.0
    ez = []
    for _ in range(400):
        sim.step(alg, 1e-15, 0.5, 0.0, ext=ext)
        ez.append(sim.e[0, 0, 2])
    assert np.all(np.diff(ez) > 0) and ez[-1] > 0.5


@pytest.mark.parametrize("alg", [1, 6, 7])
def test_free_spins_reach_boltzmann_equilibrium(O, alg):
    """Thermal noise of SIB and the spherical midpoint (the rannum-scaled
    fluctuations in a1 and s1): free spins of 2 muB in 20 T along z at 15 K
    relax to the Boltzmann distribution, <cos theta> = L(x) = coth x - 1/x
    with x = m B / (kB T) = 1.79 (L = 0.50).  A noise amplitude off by sqrt(2)
    would give L(x/2) = 0.28 or L(2x) = 0.72.  The midpoint solver, whose
    noise the bccFe golden run pins, is the control."""
    rng = np.random.default_rng(alg)
    N, m, B, T = 4000, 2.0, 20.0, 15.0
    sim = _free_spins(O, N)          # moments of 2 muB
    e0 = rng.normal(size=(1, N, 3))
    e0 /= np.linalg.norm(e0, axis=2, keepdims=True)
    sim.set_state(e0)
    ext = np.zeros_like(e0)
    ext[..., 2] = B
    samples = []
    for it in range(3000):
        sim.step(alg, 1e-15, 1.0, T, normals=rng.normal(size=(1, N, 3)), ext=ext)
        if it >= 1000 and it % 20 == 0:
            samples.append(sim.e[0, :, 2].mean())
    mub = ctypes.c_double.in_dll(O.lib(), "OR_MUB").value
    kb = ctypes.c_double.in_dll(O.lib(), "OR_K_BOLT").value
    x = m * mub * B / (kb * T)
    want = 1.0 / np.tanh(x) - 1.0 / x
    got = float(np.mean(samples))
    print(f"alg {alg}: <cos theta> {got:.4f}, L(x) {want:.4f}")
    assert abs(got - want) < 0.02, (got, want)


def test_heun_noise_gives_the_reference_effective_temperature(O):
    """Heun (heun_proper.f90): the drift carries dt/(1+lam^2), the noise
    sqrt(2 D dt) dW with D = lam/(1+lam^2) kB T/(muB m) does not, so the
    reference's scheme samples an effective temperature (1+lam^2) T.  Free
    spins of 2 muB in 20 T at 15 K with lam = 0.3 then give
    <cos theta> = L(x/1.09) = 0.469, not the Boltzmann 0.499; a noise
    amplitude off by sqrt(2) would give 0.27 or 0.70."""
    rng = np.random.default_rng(4)
    N, m, B, T, lam = 4000, 2.0, 20.0, 15.0, 0.3
    sim = _free_spins(O, N)
    e0 = rng.normal(size=(1, N, 3))
    e0 /= np.linalg.norm(e0, axis=2, keepdims=True)
    sim.set_state(e0)
    ext = np.zeros_like(e0)
    ext[..., 2] = B
    samples = []
    for it in range(8000):
        sim.step(4, 1e-15, lam, T, normals=rng.normal(size=(1, N, 3)), ext=ext)
        if it >= 3000 and it % 20 == 0:
            samples.append(sim.e[0, :, 2].mean())
    mub = ctypes.c_double.in_dll(O.lib(), "OR_MUB").value
    kb = ctypes.c_double.in_dll(O.lib(), "OR_K_BOLT").value
    x = m * mub * B / (kb * T * (1.0 + lam * lam))
    want = 1.0 / np.tanh(x) - 1.0 / x
    got = float(np.mean(samples))
    assert abs(got - want) < 0.01, (got, want)
