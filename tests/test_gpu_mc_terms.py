"""Device Monte Carlo against the oracle on a Hamiltonian with every term of
calculate_energy on the accelerated path (montecarlo_common.f90:431-865):
exchange, DM (with the reference's emom/emomM mix at lines 612 and 615,
visible because |m| = 2.2), uniaxial, cubic and type-7 anisotropy on
different sites, and a Zeeman field; heat bath through the full local field
(effective_field_single, hamiltonianactions.f90:254-392).

The coloured sweeps draw the same Philox counters on both sides.  FP64: the
spins agree to rounding.  FP32: the engine draws its trial moves in single
precision (restated by the oracle's or_trial_move_f32) and sums energies in
single precision; the oracle keeps double energies, so the spins agree to
single-precision rounding, except where a Metropolis decision lies within
rounding of its threshold -- at most a few sites, counted.
"""
import ctypes

import numpy as np
import pytest

import uppasd_amd as U
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu

FC = CF.MRY_PER_MUB


def _system(L, moment):
    sysd = CF.triangular_dmi(L, J=1.0, D=0.2, K=0.0, B=0.0, moment=moment)
    N = sysd["lattice"].natom
    rng = np.random.default_rng(2024)
    taniso = rng.choice([1, 2, 7], size=N).astype(np.int32)
    ea = rng.normal(size=(N, 3))
    ea /= np.linalg.norm(ea, axis=1, keepdims=True)
    K1 = rng.uniform(-0.3, 0.3, N)
    K2 = rng.uniform(-0.1, 0.1, N)
    # setup_anisotropies' scaling (hamiltonianinit.f90:900-909)
    cub = taniso == 2
    kan = np.stack([np.where(cub, FC * K1 / moment ** 4, FC * K1 / moment ** 2),
                    np.where(cub, FC * K2 / moment ** 6, FC * K2 / moment ** 4)], axis=1)
    sb = rng.uniform(0.2, 1.5, N)
    sysd.update(taniso=taniso, eaniso=np.ascontiguousarray(ea), kaniso=np.ascontiguousarray(kan),
                sb=sb, ext=np.array([0.4, -0.3, 1.1]))
    return sysd


def _engine(sysd, M, precision):
    N = sysd["lattice"].natom
    eng = U.Engine(N, M, precision)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_dm(sysd["dmlist"], sysd["dmlistsize"], sysd["dm_vect"])
    eng.set_anisotropy(sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"])
    eng.set_external_field(sysd["ext"])
    return eng


def _oracle_ham(O, sysd):
    return O.OracleHam(sysd["lattice"].natom, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"],
                       dmlist=sysd["dmlist"], dmlistsize=sysd["dmlistsize"],
                       dm_vect=sysd["dm_vect"], taniso=sysd["taniso"], eaniso=sysd["eaniso"],
                       kaniso=sysd["kaniso"], sb=sysd["sb"])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("mode", ["M", "H"])
def test_coloured_sweeps_full_hamiltonian_match_oracle(O, mode, precision):
    moment = 2.2
    sysd = _system(8, moment)
    N, M = sysd["lattice"].natom, 4
    temps = np.array([60.0, 250.0, 600.0, 1200.0])
    rng = np.random.default_rng(7)
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    mm = np.full((M, N), moment)
    eng = _engine(sysd, M, precision)
    eng.set_moments(emom, mm)
    ncol, col = eng.coloring()
    order = np.argsort(col, kind="stable").astype(np.int32)
    ptr = np.zeros(ncol + 1, dtype=np.int32)
    for k in range(ncol):
        ptr[k + 1] = ptr[k] + np.count_nonzero(col == k)
    seed, nsweeps = 55555, 3
    eng.mc_sweeps(mode, temps, nsweeps, seed)
    e_gpu, _ = eng.get_moments()
    ham = _oracle_ham(O, sysd)
    e = emom.copy()
    eM = np.ascontiguousarray(e * mm[..., None])
    f32 = 1 if precision == "fp32" else 0
    for s in range(nsweeps):
        O.lib().or_mc_sweep_colored_ex(ctypes.byref(ham.h), N, M, temps, mode.encode(), seed, s,
                                       ncol, ptr, np.ascontiguousarray(order), eM, e, mm,
                                       np.ascontiguousarray(sysd["ext"]), f32)
    d = np.abs(e_gpu - e).max(axis=2)
    moved = np.abs(e - emom).max(axis=2) > 1e-6
    assert moved.mean() > 0.3          # the sweeps did accept moves
    if precision == "fp64":
        assert d.max() <= 1e-12, d.max()
    else:
        off = np.count_nonzero(d > 2e-5)
        assert off <= 2, (off, np.sort(d.ravel())[-5:])
