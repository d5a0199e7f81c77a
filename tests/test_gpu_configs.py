"""Parity at the benchmark configurations' real sizes (FP64 device path vs the
FP64 oracle on the same inputs).

  C1  bcc 20^3 x 2 (16,000 spins), J1 1.34 / J2 0.77 mRy, Depondt alpha 0.1,
      dt 1e-16, 300 K on an injected N(0,1) stream: 40 steps, <= 1e-12
  C2  bcc 64^3 x 2 (524,288 spins), five shells (58 neighbours), midpoint:
      T = 0, 6 steps, and 300 K on an injected stream, 3 steps, <= 1e-12;
      and as benchmarked (tile kernel, basis-major order, 300 K on the FP64
      Philox stream), 3 steps, <= 1e-12
  C3  triangular 512^2 (262,144 spins), J + DMI + Kz + Bz, Depondt alpha 0.05:
      T = 0, 8 steps, <= 1e-12
  C4  random binary bcc alloy replica (16^3 cells, species-pair couplings and
      moments): field and energy, relative 1e-13
  C5  kagome antiferromagnet with out-of-plane DMI (48^2 cells): field and
      energy, relative 1e-13
At these sizes the oracle finishes in seconds; the long runs of the
benchmark are covered by the size-independent checks of the bench (|m|
drift, energy) and by the golden runs.
"""
import numpy as np
import pytest

import refcase as RC
import uppasd_amd as U
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu


def _engine(sysd, M=1):
    N = sysd["lattice"].natom
    eng = U.Engine(N, M)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    if "dmlist" in sysd:
        eng.set_dm(sysd["dmlist"], sysd["dmlistsize"], sysd["dm_vect"])
    if "taniso" in sysd:
        eng.set_anisotropy(sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"])
    if "ext" in sysd:
        eng.set_external_field(sysd["ext"])
    return eng


def _oracle_sim(O, sysd, M=1):
    lat = sysd["lattice"]
    latO = O.LatticeSpec(lat.cell, lat.bas, lat.atype, lat.n, lat.bc)
    aniso = None
    if "taniso" in sysd:
        aniso = (sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"])
    dm = None
    if "dmlist" in sysd:
        dm = dict(nlist=sysd["dmlist"], nlistsize=sysd["dmlistsize"], ncoup=sysd["dm_vect"])
    return RC.OracleSD(O, latO, sysd, sysd["mmom"], M=M, aniso=aniso, dm=dm)


def _run_pair(O, sysd, alg, lam, temp, steps, rng, injected):
    N = sysd["lattice"].natom
    e0 = np.ascontiguousarray(CF.random_unit(N, rng, 1))
    eng = _engine(sysd)
    eng.set_moments(e0, sysd["mmom"][None])
    eng.set_llg(lam, temp, 1e-16)
    sim = _oracle_sim(O, sysd)
    sim.set_state(e0)
    ext = None
    if "ext" in sysd:
        ext = np.ascontiguousarray(np.broadcast_to(sysd["ext"], (1, N, 3)))
    for _ in range(steps):
        normals = rng.normal(size=(1, N, 3)) if injected else None
        if injected:
            eng.set_noise(U.UASD_NOISE_INJECTED, 0, normals)
        else:
            eng.set_noise(U.UASD_NOISE_NONE)
        eng.sd_steps(alg, 1)
        sim.step(alg, 1e-16, lam, temp, normals=None if normals is None else normals.copy(),
                 ext=ext)
    e, _ = eng.get_moments()
    return np.max(np.abs(e - sim.e))


def test_c1_depondt_injected_noise(O):
    sysd = CF.bcc_shells(20, [1.34, 0.77])
    assert sysd["lattice"].natom == 16000
    assert _run_pair(O, sysd, 5, 0.1, 300.0, 40, np.random.default_rng(70), True) <= 1e-12


def test_c2_full_size_midpoint(O):
    sysd = CF.bcc_shells(64, CF.bcc5_couplings())
    assert sysd["lattice"].natom == 524288 and int(sysd["nlistsize"].max()) == 58
    assert _run_pair(O, sysd, 1, 0.1, 0.0, 6, np.random.default_rng(71), False) <= 1e-12
    assert _run_pair(O, sysd, 1, 0.1, 300.0, 3, np.random.default_rng(72), True) <= 1e-12


def test_c2_as_benchmarked_tile_kernel_philox(O):
    """The headline run as the benchmark makes it: bcc 64^3 x 2, basis-major
    site order, lattice declared, the LDS-tiled kernel with (16,4,4)-cell
    tiles and the resident loop, midpoint at 300 K on the FP64 Philox stream.
    Three steps equal the FP64 oracle's on the same counter stream
    (or_sd_run, noise 2) within 1e-12."""
    import ctypes
    sysd = CF.bcc_shells(64, CF.bcc5_couplings())
    lat = sysd["lattice"]
    N = lat.natom
    e0 = np.ascontiguousarray(CF.random_unit(N, np.random.default_rng(77), 1))
    seed, steps = 0xC2C2C2C2, 3
    eng = U.Engine(N, 1, "fp64", site_order=lat.NA)
    eng.set_lattice_dims(lat.NA, *lat.n)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_moments(e0, sysd["mmom"][None])
    eng.set_llg(0.1, 300.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, seed)
    eng.set_tuning(eng.TUNE_LLG, 4, 256 | (4 << 18) | (1 << 22))   # tile (16,4,4), resident
    eng.sd_steps(1, steps)
    assert eng.llg_tile() == 4
    e, _ = eng.get_moments()
    ham = O.OracleHam(N, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    mm = np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (1, N)), dtype=np.float64)
    eo = e0.copy()
    eMo = np.ascontiguousarray(eo * mm[..., None])
    O.lib().or_sd_run(ctypes.byref(ham.h), N, 1, 1, steps, 0, 1e-16, np.full(N, 0.1),
                      np.full(N, 300.0), np.ones(N), mm, eo, eMo, seed, 2, None)
    assert np.max(np.abs(e - eo)) <= 1e-12


def test_c3_full_size_depondt(O):
    sysd = CF.triangular_dmi(512)
    assert sysd["lattice"].natom == 262144
    assert _run_pair(O, sysd, 5, 0.05, 0.0, 8, np.random.default_rng(73), False) <= 1e-12


def _field_parity(O, sysd, M, rng):
    N = sysd["lattice"].natom
    emom = CF.random_unit(N, rng, M)
    emomM = np.ascontiguousarray(emom * sysd["mmom"][None, :, None])
    beff, en = _engine(sysd, M).effective_field(emomM)
    kw = {}
    if "dmlist" in sysd:
        kw.update(dmlist=sysd["dmlist"], dmlistsize=sysd["dmlistsize"], dm_vect=sysd["dm_vect"])
    bref, eref = O.OracleHam(N, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"], **kw).field(emomM)
    assert np.max(np.abs(beff - bref)) <= 1e-13 * np.max(np.abs(bref))
    assert abs(en - eref) <= 1e-13 * abs(eref)


def test_c4_alloy_replica_field(O):
    rng = np.random.default_rng(74)
    L = 16
    occ = np.where(rng.random(2 * L ** 3) < 0.5, 1, 2)
    sysd = CF.bcc_alloy(L, occ)
    assert set(np.unique(sysd["mmom"])) == {2.6, 1.7}
    _field_parity(O, sysd, 4, rng)


def test_c5_kagome_field(O):
    rng = np.random.default_rng(75)
    sysd = CF.kagome_afm(48, J=-1.0, Dz=0.08)
    _field_parity(O, sysd, 2, rng)
