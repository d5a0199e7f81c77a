"""The Fortran binding on the device: a Fortran driver (tests/fortran/driver.f90)
runs sd_mphase's loop body and mc_evolve through the drop-in wrappers of
module uasd_hip, which keep the reference routines' argument lists.

  * Solvers chain (regulartests.yaml), T = 0, 8000 steps of midpoint and
    Depondt from the oracle's MT19937 start: the reference's golden <m>
    within 1e-8
  * bcc five-shell lattice at 300 K on an injected noise stream, 40 steps:
    spins within 1e-12 of the oracle's sd_mphase restatement, field-routine
    energies within 1e-12 relative
  * MC: Fortran mc_evolve sweeps equal the engine's own sweeps bit for bit
    (same seed, same colouring, same Philox stream)
"""
import os
import subprocess

import numpy as np
import pytest

import refcase as RC
import uppasd_amd as U
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(HERE, "fortran", "uasd_driver")


def _run_driver(tmp_path, kind, nlist, nlistsize, ncoup, mmom, emom, alg, nsteps, dt, lam,
                temp, ext=(0.0, 0.0, 0.0), normals=None, lattice=(0, 0, 0, 0)):
    """lattice = (NA, N1, N2, N3): NA > 1 selects the basis-major site order,
    N1 > 0 passes the lattice dimensions to the library."""
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (tests/fortran/build_driver.py)")
    M, N, _ = emom.shape
    W = nlist.shape[1]
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fin, "wb") as f:
        np.array([kind, N, M, W, alg, nsteps, 0 if normals is None else 1, *lattice],
                 np.int32).tofile(f)
        np.array([dt, lam, temp], np.float64).tofile(f)
        np.ascontiguousarray(nlist, np.int32).tofile(f)
        np.ascontiguousarray(nlistsize, np.int32).tofile(f)
        np.ascontiguousarray(ncoup.reshape(N, W), np.float64).tofile(f)
        np.ascontiguousarray(mmom, np.float64).tofile(f)
        np.ascontiguousarray(emom, np.float64).tofile(f)
        np.asarray(ext, np.float64).tofile(f)
        if normals is not None:
            np.ascontiguousarray(normals, np.float64).tofile(f)
    print(f"fortran driver kind={kind} N={N} M={M} alg={alg} steps={nsteps}", flush=True)
    r = subprocess.run([DRIVER, str(fin), str(fout)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    out = np.fromfile(fout, np.float64)
    n3 = M * N * 3
    return (out[:n3].reshape(M, N, 3), out[n3:2 * n3].reshape(M, N, 3), out[2 * n3:])


@pytest.mark.parametrize("solver,name", [(1, "Midpoint"), (4, "Heun"), (5, "Depondt")])
def test_fortran_solvers_golden(O, tmp_path, solver, name):
    c = RC.read_case("Solvers", "inpsd.dat.base")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    sh = U.Shells.from_lines(lat, c["jlines"])
    nl, ns, nc = U.build_pair_list(lat, sh, c["sym"], c["ammom"].reshape(-1, 1))
    latO = O.LatticeSpec(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    sim = RC.OracleSD(O, latO, dict(nlist=nl, nlistsize=ns, ncoup=nc), np.ones(lat.natom))
    sim.random_init(O.mt(1), latO)
    g = [x for x in RC.golden("Solvers") if x["type"] == name][0]
    nsteps = g["select"]["#_Iter."]
    e, eM, _ = _run_driver(tmp_path, 1, nl, ns, nc, np.ones((1, lat.natom)), sim.e, solver,
                           nsteps, 1.0e-15, 1.0, 0.0)
    m = eM[0].sum(axis=0) / lat.natom
    assert np.allclose([*m, np.linalg.norm(m)], g["expected"], atol=1e-8)


@pytest.mark.parametrize("ordered", [False, True])
@pytest.mark.parametrize("alg", [1, 4, 5, 6, 7])
def test_fortran_injected_noise_matches_oracle(O, tmp_path, alg, ordered):
    rng = np.random.default_rng(21)
    sysd = CF.bcc_shells(3, CF.bcc5_couplings())
    lat = sysd["lattice"]
    M, N, nsteps = 2, lat.natom, 40
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    normals = rng.normal(size=(nsteps, M, N, 3))
    mm = np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (M, N)))
    lattice = (2, *lat.n) if ordered else (0, 0, 0, 0)
    e, eM, en = _run_driver(tmp_path, 1, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"], mm,
                            emom, alg, nsteps, 1e-16, 0.1, 300.0, normals=normals,
                            lattice=lattice)
    latO = O.LatticeSpec(lat.cell, lat.bas, lat.atype, lat.n, lat.bc)
    sim = RC.OracleSD(O, latO, sysd, sysd["mmom"], M=M)
    sim.set_state(emom)
    en_ref = []
    for s in range(nsteps):
        en_ref.append(sim.ham.field(sim.eM)[1])
        sim.step(alg, 1e-16, 0.1, 300.0, normals=normals[s].copy())
    assert np.max(np.abs(e - sim.e)) <= 1e-12
    assert np.max(np.abs(eM - sim.eM)) <= 1e-12 * np.max(sysd["mmom"])
    assert np.max(np.abs(en - np.array(en_ref)) / np.abs(en_ref)) <= 1e-12


@pytest.mark.parametrize("mode,alg", [("M", 1), ("H", 2)])
def test_fortran_mc_evolve_matches_engine(tmp_path, mode, alg):
    rng = np.random.default_rng(22)
    sysd = CF.bcc_shells(4, CF.bcc5_couplings())
    N, M = sysd["lattice"].natom, 2
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    mm = np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (M, N)))
    ext = (0.0, 0.0, 2.5)
    e, eM, _ = _run_driver(tmp_path, 2, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"], mm, emom,
                           alg, 6, 0.0, 0.0, 700.0, ext=ext)
    eng = U.Engine(N, M)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_external_field(np.array(ext))
    eng.set_moments(emom, mm)
    eng.mc_sweeps(mode, [700.0] * M, 6, 4242)
    e_ref, eM_ref = eng.get_moments()
    assert np.array_equal(e, e_ref)
    assert np.array_equal(eM, eM_ref)
