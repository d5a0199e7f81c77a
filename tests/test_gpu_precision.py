"""The reduced-precision form against the FP64 oracle (the north star's
precision form: FP32 field gather and spin rotation, FP64 reductions, the
error stated per run).

* Triangular J + DMI + Kz + Bz monolayer (the workload of the precision
  benchmark, 96^2 sites here), Depondt, alpha 0.05, T = 0 from a random
  start, 400 steps: energy per spin within 1e-5 relative of the oracle's at
  every checkpoint, the FP32 field of the oracle's state within 1e-6 of
  max |B|, |m_i| norm drift below 1e-6.
* Kagome antiferromagnet with out-of-plane DMI, four ensembles of uneven
  size (6^2 .. 13^2 cells) joined into one lattice, each at its own
  temperature (log-spaced 1 .. 300 K), FP32 Depondt with hipGraph-replayed
  steps and the Philox thermal field, alpha 0.2, dt 5e-17, 200 steps from a
  perturbed 120-degree state; the oracle integrates the same deviates in
  FP64.  Per ensemble: energy per spin and sublattice magnetisation within
  1e-5 relative, and the device's time-averaged energy (accumulators) within
  1e-5 of the oracle's average over the same samples.
"""

import numpy as np
import pytest

import refcase as RC
import uppasd_amd as U
from uppasd_amd import batch as Bt
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu

MUB, MRY = 9.274009994e-24, 2.179872325e-21


def _energy_per_spin(ham, eM, ext):
    """calc_energy's total per spin (mRy) of the oracle state, per ensemble."""
    t = ham.energy_terms(np.ascontiguousarray(eM), ext)
    return t.sum(axis=1) * MUB / MRY


def test_fp32_depondt_triangular_against_fp64_oracle(O):
    sysd = CF.triangular_dmi(96)
    N = sysd["lattice"].natom
    rng = np.random.default_rng(3)
    e0 = np.ascontiguousarray(CF.random_unit(N, rng, 1))
    eng = U.Engine(N, 1, "fp32")
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_dm(sysd["dmlist"], sysd["dmlistsize"], sysd["dm_vect"])
    eng.set_anisotropy(sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"])
    eng.set_external_field(sysd["ext"])
    eng.set_moments(e0, sysd["mmom"][None])
    eng.set_llg(0.05, 0.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_NONE)
    sim = RC.OracleSD(O, sysd["lattice"], sysd, sysd["mmom"],
                      aniso=(sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"]),
                      dm=dict(nlist=sysd["dmlist"], nlistsize=sysd["dmlistsize"],
                              ncoup=sysd["dm_vect"]))
    sim.set_state(e0)
    ext = np.ascontiguousarray(np.broadcast_to(sysd["ext"], (1, N, 3)))
    done, errs = 0, {}
    for c in (10, 100, 400):
        eng.sd_steps(5, c - done)
        for _ in range(c - done):
            sim.step(5, 1e-16, 0.05, 0.0, ext=ext)
        done = c
        ref = _energy_per_spin(sim.ham, sim.eM, ext)[0]
        got = eng.energy_terms()[0].sum()
        errs[c] = abs(got - ref) / abs(ref)
    print("FP32 triangular energy rel err", errs, flush=True)
    assert max(errs.values()) <= 1e-5, errs
    e32, eM32 = eng.get_moments()
    drift = np.max(np.abs(np.linalg.norm(e32[0], axis=1) - 1.0))
    assert drift <= 1e-6, drift
    b64, _ = sim.ham.field(sim.eM, ext)
    b32, _ = eng.effective_field(np.ascontiguousarray(sim.eM))
    ferr = np.max(np.abs(b32 - b64)) / np.max(np.abs(b64))
    assert ferr <= 1e-6, ferr


def _kagome_batch():
    Ls = [6, 8, 11, 13]
    Ts = np.logspace(0.0, np.log10(300.0), len(Ls))
    reps = []
    for L, T in zip(Ls, Ts):
        d = CF.kagome_afm(L, J=-1.0, Dz=0.08)
        d["temp"] = T
        reps.append(d)
    return reps, Bt.join(reps, per_site=("temp",))


def test_fp32_batched_kagome_against_fp64_oracle(O):
    reps, j = _kagome_batch()
    N = j.natom
    rng = np.random.default_rng(9)
    ang = 2 * np.pi / 3 * (np.arange(N) % 3)             # basis atom of every site
    e0 = np.stack([np.cos(ang), np.sin(ang), np.zeros(N)], axis=1) + 0.05 * rng.normal(size=(N, 3))
    e0 = np.ascontiguousarray((e0 / np.linalg.norm(e0, axis=1, keepdims=True))[None])
    lam, dt, seed, steps, K = 0.2, 5e-17, 5, 200, 10
    eng = Bt.engine_for(j, 1, "fp32")
    eng.set_moments(e0, j.mmom[None])
    eng.set_llg(np.full(N, lam), j.extra["temp"], dt)
    eng.set_noise(U.UASD_NOISE_PHILOX, seed)
    eng.set_graph_steps(K)
    hams = [O.OracleHam(r["lattice"].natom, r["nlist"], r["nlistsize"], r["ncoup"],
                        dmlist=r["dmlist"], dmlistsize=r["dmlistsize"], dm_vect=r["dm_vect"])
            for r in reps]
    sim = _joined_oracle(O, j)
    sim.set_state(e0)
    th = O.lib().or_thermal3
    normals = np.zeros((1, N, 3))
    f = np.zeros(3, np.float32)
    e_samples = []
    for s in range(steps):
        for i in range(N):
            th(seed, 0x11, s, i, 0, f)
            normals[0, i] = f
        sim.step(5, dt, lam, j.extra["temp"], normals=normals)
        if (s + 1) % K == 0:
            e_samples.append([_energy_per_spin(hams[r], sim.eM[:, a:b], None)[0]
                              for r, (a, b) in enumerate(zip(j.offsets[:-1], j.offsets[1:]))])
    for _ in range(steps // K):
        eng.sd_steps(5, K)
        eng.accumulate(with_energy=True)
    ene = eng.energy_terms()              # (nseg, M, 4) per spin, mRy
    acc = eng.accumulators()
    _, eM32 = eng.get_moments()
    e_ref = np.array(e_samples)
    errs = []
    for r, (a, b) in enumerate(zip(j.offsets[:-1], j.offsets[1:])):
        got = ene[r, 0].sum()
        want = e_ref[-1, r]
        errs.append(abs(got - want) / abs(want))
        for k in range(3):                # sublattice magnetisation, FP64 sums
            m32 = eM32[0, a + k:b:3].sum(axis=0) / ((b - a) // 3)
            m64 = sim.eM[0, a + k:b:3].sum(axis=0) / ((b - a) // 3)
            assert np.linalg.norm(m32 - m64) <= 1e-5 * np.linalg.norm(m64), (r, k)
        # running means weighted 1..n (calc_and_print_cumulant) of the samples
        w = np.arange(1, len(e_ref) + 1, dtype=np.float64)
        avg = np.sum(w * e_ref[:, r]) / np.sum(w)
        assert abs(acc["e"][r, 0] - avg) <= 1e-5 * abs(avg), (r, acc["e"][r, 0], avg)
    print("FP32 kagome batch energy rel err per ensemble", errs, flush=True)
    assert max(errs) <= 1e-5, errs


def _joined_oracle(O, j):
    class _Lat:
        natom = j.natom
    return RC.OracleSD(O, _Lat, dict(nlist=j.nlist, nlistsize=j.nlistsize, ncoup=j.ncoup),
                       j.mmom, dm=dict(nlist=j.extra["dmlist"], nlistsize=j.extra["dmlistsize"],
                                       ncoup=j.extra["dm_vect"]))
