"""The oracle against the reference's own golden values (tests/regulartests.yaml).

These runs are stream-identical to the reference: MT19937 (mtprng quirk
included), ziggurat normals, magninit's random start, the neighbour map and
both LLG solvers.  A wrong sign, index or operand order in any of them moves
the printed digits.  Tolerance: the yaml prints 9 significant digits; its
'similar' comparison allows |delta| <= 1e-8.
"""
import ctypes
import os

import numpy as np
import pytest

import refcase as RC

MUB, KB, MRY = 9.274009994e-24, 1.38064852e-23, 2.179872325e-21


def _lattice(O, c):
    return O.LatticeSpec(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])


def _pairs(O, c, lat):
    return O.build_pairs(lat, c["jlines"], c["sym"], ammom=c["ammom"].reshape(-1, 1))


def _sites(lat, per_basis):
    return np.tile(np.asarray(per_basis, dtype=np.float64), lat.natom // lat.NA)


def _expected(case, typ):
    return [g for g in RC.golden(case) if g["type"] == typ]


@pytest.mark.parametrize("solver,name", [(1, "Midpoint"), (4, "Heun"), (5, "Depondt")])
def test_solvers_chain_golden(O, solver, name):
    c = RC.read_case("Solvers", "inpsd.dat.base")
    lat = _lattice(O, c)
    sim = RC.OracleSD(O, lat, _pairs(O, c, lat), _sites(lat, c["ammom"]))
    st = O.mt(1)                       # tseed default 1
    sim.random_init(st, lat)
    z = O.zig(1)
    g = _expected("Solvers", name)[0]
    nstep = g["select"]["#_Iter."]
    noise = np.zeros_like(sim.e)
    for _ in range(nstep):
        O.lib().or_zig_fill(ctypes.byref(z), noise.reshape(-1), noise.size)
        sim.step(solver, 1.0e-15, 1.0, 0.0, normals=noise)
    m = sim.mavg()[0]
    got = [m[0], m[1], m[2], np.linalg.norm(m)]
    assert np.allclose(got, g["expected"], rtol=0, atol=1e-8), (got, g["expected"])


def test_bccfe_sd_golden(O):
    """bcc Fe, 432 spins, midpoint at 500 K: 2000 initial-phase steps
    (dt = 1.0d-16), then 3000 measurement steps with the default timestep
    (the single-precision literal 1.0e-16).  Checks <m> at iterations 0 and
    2700 and cumulant row 41 (<M>, <M^2>, <M^4>, U4, chi)."""
    import ctypes
    c = RC.read_case("bccFe", "inpsd.dat.base")
    lat = _lattice(O, c)
    sim = RC.OracleSD(O, lat, _pairs(O, c, lat), _sites(lat, c["ammom"]))
    e0 = np.tile(c["edir"], (lat.natom // lat.NA, 1))
    sim.set_state(e0[None])
    z = O.zig(5)
    noise = np.zeros_like(sim.e)

    def step(T, dt, lam):
        O.lib().or_zig_fill(ctypes.byref(z), noise.reshape(-1), noise.size)
        sim.step(1, dt, lam, T, normals=noise)

    for _ in range(2000):
        step(500.0, 1.0e-16, 0.5)
    exp = {g["type"]: g for g in RC.golden("bccFe") if g["name"] == "bcc Fe (SD)"}
    m = sim.mavg()[0]
    assert np.allclose([*m, np.linalg.norm(m)], exp["Initial phase"]["expected"], atol=1e-8)
    dt = float(np.float32(1.0e-16))
    w = tot = mm = mm2 = mm4 = ee = ee2 = 0.0
    ncum, cumrow, m2700 = 0, None, None
    N = lat.natom
    en = None
    for mstep in range(1, 3001):
        if mstep - 1 == 2700:
            m2700 = sim.mavg()[0].copy()
        if mstep % 50 == 0:   # calc_and_print_cumulant, prn_averages.f90:919-1095
            me = np.linalg.norm(sim.eM[0].sum(axis=0)) / N
            w += 1.0
            mm = (mm * tot + me * w) / (tot + w)
            mm2 = (mm2 * tot + me ** 2 * w) / (tot + w)
            mm4 = (mm4 * tot + (me ** 2) ** 2 * w) / (tot + w)
            ee = (ee * tot + en * w) / (tot + w)
            ee2 = (ee2 * tot + en ** 2 * w) / (tot + w)
            tot += w
            ncum += 1
            if ncum == 41:
                U = 1 - (mm4 / 3 / mm2 ** 2)
                chi = (mm2 - mm ** 2) * MUB ** 2 * N / KB ** 2 / 500.0
                cv = (ee2 - ee ** 2) * MRY ** 2 * N / KB ** 2 / 500.0 ** 2
                cumrow = [mm, mm2, mm4, U, chi, cv]
        # calc_energy runs after measure when mod(mstep-1, avrg_step) == 0
        # (sd_driver.f90:554-575), so a cumulant sample uses the energy of 49
        # steps before it; energy per atom in mRy (energy.f90:346-372)
        if (mstep - 1) % 50 == 0:
            en = sim.ham.field(sim.eM)[1] / N
        step(500.0, dt, 0.5)
    assert np.allclose([*m2700, np.linalg.norm(m2700)], exp["Magnetization"]["expected"], atol=1e-8)
    want = exp["Cumulants"]["expected"][:6]
    assert np.allclose(cumrow, want, rtol=1e-8, atol=1e-12), (cumrow, want)


def test_heisstripe_anisotropy_golden(O):
    """10x1x100 stripe, vacuum in x, uniaxial K1 = -0.02 mRy along y, T = 0
    midpoint from a random start: <m> at iteration 1190 pins the anisotropy
    field and the vacuum-boundary neighbour lists."""
    import ctypes
    c = RC.read_case("HeisStripe")
    lat = _lattice(O, c)
    pairs = _pairs(O, c, lat)
    an = O.setup_anisotropies(lat, c["ktype"], c["kvals"], c["ammom"])
    sim = RC.OracleSD(O, lat, pairs, _sites(lat, c["ammom"]), aniso=an)
    st = O.mt(1)
    sim.random_init(st, lat)
    z = O.zig(1)
    noise = np.zeros_like(sim.e)
    g = [x for x in RC.golden("HeisStripe") if x["type"] == "Magnetization"][0]
    for _ in range(g["select"]["#_Iter."]):
        O.lib().or_zig_fill(ctypes.byref(z), noise.reshape(-1), noise.size)
        sim.step(1, 1.0e-16, 0.10, 0.0, normals=noise)
    m = sim.mavg()[0]
    assert np.allclose([*m, np.linalg.norm(m)], g["expected"], atol=1e-8), (m, g["expected"])


def test_kagome_dmi_golden(O):
    """Kagome lattice 12x12 (PBC in-plane), J = 1 mRy and a DM vector on
    every bond (dmfile, sym 0), moments from the momfile directions, T = 0
    midpoint with damping 0.5: <m> at iteration 13000 and atom 2's moment at
    iteration 2400 (trajectory file) pin the DMI field term and the DM
    neighbour map."""
    c = RC.read_case("kagome")
    lat = _lattice(O, c)
    pairs = _pairs(O, c, lat)
    dm = O.build_pairs(lat, c["dmlines"], 0, hdim=3, use_nntype=False,
                       ammom=c["ammom"].reshape(-1, 1))
    sim = RC.OracleSD(O, lat, pairs, _sites(lat, c["ammom"]), dm=dm)
    sim.set_state(np.tile(c["edir"], (lat.natom // lat.NA, 1))[None])
    gm = _expected("kagome", "Magnetization")[0]
    gt = _expected("kagome", "Trajectory")[0]
    traj = None
    for it in range(gm["select"]["#_Iter."]):
        if it == gt["select"]["#_Iter."]:
            traj = sim.e[0, 1].copy()
        sim.step(1, 1.0e-16, 0.5, 0.0)
    m = sim.mavg()[0]
    assert np.allclose([*m, np.linalg.norm(m)], gm["expected"], atol=1e-8), (m, gm["expected"])
    assert np.allclose([2.0, *traj, 2.0], gt["expected"], atol=1e-8), (traj, gt["expected"])


def test_kagome_golden_rejects_flipped_dm(O):
    """The kagome golden values are sensitive to the DM term: the same run
    with every DM vector negated (a plausible sign mistake) misses them."""
    c = RC.read_case("kagome")
    lat = _lattice(O, c)
    pairs = _pairs(O, c, lat)
    dm = O.build_pairs(lat, c["dmlines"], 0, hdim=3, use_nntype=False,
                       ammom=c["ammom"].reshape(-1, 1))
    dm["ncoup"] = np.ascontiguousarray(-dm["ncoup"])
    sim = RC.OracleSD(O, lat, pairs, _sites(lat, c["ammom"]), dm=dm)
    sim.set_state(np.tile(c["edir"], (lat.natom // lat.NA, 1))[None])
    gm = _expected("kagome", "Magnetization")[0]
    for _ in range(gm["select"]["#_Iter."]):
        sim.step(1, 1.0e-16, 0.5, 0.0)
    m = sim.mavg()[0]
    assert not np.allclose([*m, np.linalg.norm(m)], gm["expected"], atol=1e-4)


def test_heischain_from_restart_file_golden(O):
    """Heisenberg chain (100 spins, J = 2.5 mRy, vacuum in x and y): the start
    comes from the reference's restart file `startmom` through
    uppasd_amd.restart.read_restart (header skipped, moments normalised as
    read_mag_conf_std does).  20000 initial-phase midpoint steps at 0.1 K with
    damping 4 on the ziggurat stream (default tseed 1), then the measurement
    phase at T = 0 without damping: <M> at iteration 1000 and the total
    energy per atom at iteration 1400 against the regression values."""
    import ctypes
    from uppasd_amd import restart as RS
    c = RC.read_case("HeisChain", "inpsd.dat")
    lat = _lattice(O, c)
    N = lat.natom
    rstep, mmom, emom = RS.read_restart(os.path.join(RC.CASES, "HeisChain", "startmom"), N, 1)
    assert rstep == 0 and np.all(mmom == 1.0)
    sim = RC.OracleSD(O, lat, _pairs(O, c, lat), mmom[0])
    sim.set_state(emom)
    z = O.zig(1)
    noise = np.zeros_like(sim.e)
    for _ in range(20000):
        O.lib().or_zig_fill(ctypes.byref(z), noise.reshape(-1), noise.size)
        sim.step(1, 1e-16, 4.0, 0.1, normals=noise)
    zero = np.zeros_like(sim.e)
    exp = {g["type"]: g["expected"] for g in RC.golden("HeisChain")}
    for _ in range(1000):
        sim.step(1, 1e-16, 0.0, 0.0, normals=zero)
    m = sim.mavg()[0]
    assert np.allclose([*m, np.linalg.norm(m)], exp["Magnetization"], rtol=0, atol=1e-8)
    for _ in range(400):
        sim.step(1, 1e-16, 0.0, 0.0, normals=zero)
    assert abs(sim.ham.field(sim.eM)[1] / N - exp["Total energy"][0]) <= 1e-8


def test_heischain_af_golden(O):
    """Antiferromagnetic two-sublattice chain (100 cells, |m| = 1 and 2,
    vacuum in x and y): random start on the MT19937 stream (default tseed 1),
    midpoint at T = 0 with damping 0.5.  <M> at iteration 1000, the total
    energy per atom at 1500 (calc_energy's terms, here exchange only) and the
    per-type projected averages (do_proj_avrg Y: the sum of emomM over a
    type's sites per cell, prn_averages.f90:704-716) at 5000."""
    from uppasd_amd import output as OUT
    c = RC.read_case("HeisChainAF", "inpsd.dat")
    lat = _lattice(O, c)
    sim = RC.OracleSD(O, lat, _pairs(O, c, lat), _sites(lat, c["ammom"]))
    sim.random_init(O.mt(1), lat)
    exp = {g["type"]: g["expected"] for g in RC.golden("HeisChainAF")}
    ncell = lat.natom // lat.NA
    for it in range(1, 5001):
        sim.step(1, 1e-16, 0.5, 0.0)
        if it == 1000:
            m = sim.mavg()[0]
            assert np.allclose([*m, np.linalg.norm(m)], exp["Magnetization"], rtol=0,
                               atol=1e-8)
        if it == 1500:
            t = OUT.TotEnergy.per_ensemble(sim.ham.energy_terms(sim.eM) * OUT.MUB / OUT.MRY)
            assert abs(t[0, 0] - exp["Total energy"][0]) <= 1e-8
    for k in (1, 2):
        v = sim.eM[0].reshape(ncell, lat.NA, 3)[:, k - 1].sum(axis=0) / ncell
        assert np.allclose([np.linalg.norm(v), *v], exp[f"Sublattice {k}"], rtol=0, atol=1e-8), k
