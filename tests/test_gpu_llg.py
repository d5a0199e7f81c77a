"""GPU parity of the field and LLG kernels against the FP64 oracle.

Tolerances (FP64 path, -ffp-contract=off on both sides):
  * field, single LLG half-steps: relative 1e-13 (the sums run in the same
    order; only libm ulps of sin/cos/log/sqrt differ between host and device)
  * T = 0 trajectories and injected-noise trajectories: max |e_gpu - e_cpu|
    <= 1e-12 after the stated number of steps
  * reference golden values (Solvers, bccFe): |delta| <= 1e-8, as the
    reference's own 'similar' comparison
"""
import ctypes
import os

import numpy as np
import pytest

import refcase as RC
import uppasd_amd as U
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu


def oracle_system(O, sysd, M=1):
    lat = sysd["lattice"]
    latO = O.LatticeSpec(lat.cell, lat.bas, lat.atype, lat.n, lat.bc)
    return latO


def engine_for(sysd, M=1, precision="fp64"):
    lat = sysd["lattice"]
    eng = U.Engine(lat.natom, M, precision)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    if "dmlist" in sysd:
        eng.set_dm(sysd["dmlist"], sysd["dmlistsize"], sysd["dm_vect"])
    if "taniso" in sysd:
        eng.set_anisotropy(sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"])
    if "ext" in sysd:
        eng.set_external_field(sysd["ext"])
    return eng


def oracle_ham(O, sysd):
    kw = {}
    if "dmlist" in sysd:
        kw.update(dmlist=sysd["dmlist"], dmlistsize=sysd["dmlistsize"], dm_vect=sysd["dm_vect"])
    if "taniso" in sysd:
        kw.update(taniso=sysd["taniso"], eaniso=sysd["eaniso"], kaniso=sysd["kaniso"], sb=sysd["sb"])
    return O.OracleHam(sysd["lattice"].natom, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"], **kw)


def ext_array(sysd, M):
    N = sysd["lattice"].natom
    if "ext" not in sysd:
        return None
    return np.ascontiguousarray(np.broadcast_to(sysd["ext"], (M, N, 3)))


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("kind", ["bcc5", "tri_dmi"])
def test_effective_field_matches_oracle(O, kind):
    rng = np.random.default_rng(11)
    sysd = CF.bcc_shells(6, CF.bcc5_couplings()) if kind == "bcc5" else CF.triangular_dmi(16)
    M = 3
    N = sysd["lattice"].natom
    emom = CF.random_unit(N, rng, M)
    emomM = emom * sysd["mmom"][None, :, None]
    eng = engine_for(sysd, M)
    beff, en, b1, b2 = eng.effective_field(np.ascontiguousarray(emomM), parts=True)
    ham = oracle_ham(O, sysd)
    bref, eref, b1ref, b2ref = ham.field_parts(np.ascontiguousarray(emomM), ext_array(sysd, M))
    assert rel(beff, bref) <= 1e-13
    assert rel(b1, b1ref) <= 1e-13
    assert np.max(np.abs(b2 - b2ref)) <= 1e-13 * max(np.max(np.abs(bref)), 1.0)
    assert abs(en - eref) <= 1e-12 * abs(eref)


def test_anisotropy_types_field(O):
    """Uniaxial (1), cubic (2) and combined (7) anisotropy terms."""
    rng = np.random.default_rng(3)
    sysd = CF.bcc_shells(4, [1.0, 0.3])
    N = sysd["lattice"].natom
    t = rng.choice([1, 2, 7], size=N).astype(np.int32)
    e = rng.normal(size=(N, 3))
    e /= np.linalg.norm(e, axis=1)[:, None]
    sysd.update(taniso=t, eaniso=e, kaniso=rng.normal(size=(N, 2)) * 30.0, sb=rng.random(N))
    emomM = CF.random_unit(N, rng, 2) * 2.2
    eng = engine_for(sysd, 2)
    beff, en = eng.effective_field(np.ascontiguousarray(emomM))
    bref, eref = oracle_ham(O, sysd).field(np.ascontiguousarray(emomM))
    assert rel(beff, bref) <= 1e-13
    assert abs(en - eref) <= 1e-12 * abs(eref)


def _oracle_step(O, sim, alg, dt, lam, T, normals, ext):
    sim.step(alg, dt, lam, T, normals=normals, ext=ext)


@pytest.mark.parametrize("alg", [1, 4, 5, 6, 7])
def test_dropin_step_matches_oracle(O, alg):
    """field -> evolve_first -> field -> evolve_second -> moment_update through
    the host-array entry points, with an injected noise stream at 300 K."""
    rng = np.random.default_rng(5)
    sysd = CF.triangular_dmi(12)
    M, N = 2, sysd["lattice"].natom
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    mm = np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (M, N)))
    eng = engine_for(sysd, M)
    eng.set_moments(emom, mm)
    lam, T, dt = 0.05, 300.0, 1e-16
    eng.set_llg(lam, T, dt)
    sim = RC.OracleSD(O, oracle_system(O, sysd), dict(nlist=sysd["nlist"],
                      nlistsize=sysd["nlistsize"], ncoup=sysd["ncoup"]), sysd["mmom"], M=M,
                      aniso=(sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"]),
                      dm=dict(nlist=sysd["dmlist"], nlistsize=sysd["dmlistsize"],
                              ncoup=sysd["dm_vect"]))
    sim.set_state(emom)
    ext = ext_array(sysd, M)
    e, eM = emom.copy(), emom * mm[..., None]
    e2, b2 = emom.copy(), np.zeros_like(e)   # after moment_update emom2 equals emom
    for it in range(5):
        normals = rng.normal(size=(M, N, 3))
        eng.set_noise(U.UASD_NOISE_INJECTED, 0, normals)
        beff, _ = eng.effective_field(eM)
        eng.evolve_first(alg, beff, e, e2, eM, b2)
        beff, _ = eng.effective_field(eM)
        eng.evolve_second(alg, beff, b2, e, e2)
        U.moment_update(mm, e, e2, eM)
        sim.step(alg, dt, lam, T, normals=normals.copy(), ext=ext)
        assert np.max(np.abs(e - sim.e)) <= 1e-13, it


@pytest.mark.parametrize("alg", [1, 4, 5, 6, 7])
def test_resident_T0_trajectory(O, alg):
    """Deterministic T = 0 trajectory, fused kernels vs oracle, 500 steps."""
    rng = np.random.default_rng(7)
    sysd = CF.bcc_shells(4, CF.bcc5_couplings())
    N = sysd["lattice"].natom
    emom = np.ascontiguousarray(CF.random_unit(N, rng, 1))
    eng = engine_for(sysd)
    eng.set_moments(emom, sysd["mmom"][None])
    eng.set_llg(0.1, 0.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_NONE)
    sim = RC.OracleSD(O, oracle_system(O, sysd), sysd, sysd["mmom"])
    sim.set_state(emom)
    for _ in range(500):
        sim.step(alg, 1e-16, 0.1, 0.0)
    eng.sd_steps(alg, 500)
    e, eM = eng.get_moments()
    assert np.max(np.abs(e - sim.e)) <= 1e-12
    assert np.max(np.abs(eM - sim.eM)) <= 1e-12 * 2.2


@pytest.mark.parametrize("alg", [1, 4, 5, 6, 7])
def test_resident_injected_noise_trajectory(O, alg):
    """Stochastic LLG at 300 K on an identical injected noise stream."""
    rng = np.random.default_rng(9)
    sysd = CF.bcc_shells(4, CF.bcc5_couplings())
    M, N = 2, sysd["lattice"].natom
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    eng = engine_for(sysd, M)
    eng.set_moments(emom, np.broadcast_to(sysd["mmom"], (M, N)))
    eng.set_llg(0.1, 300.0, 1e-16)
    sim = RC.OracleSD(O, oracle_system(O, sysd), sysd, sysd["mmom"], M=M)
    sim.set_state(emom)
    for _ in range(100):
        normals = rng.normal(size=(M, N, 3))
        eng.set_noise(U.UASD_NOISE_INJECTED, 0, normals)
        eng.sd_steps(alg, 1)
        sim.step(alg, 1e-16, 0.1, 300.0, normals=normals.copy())
    e, _ = eng.get_moments()
    assert np.max(np.abs(e - sim.e)) <= 1e-12


@pytest.mark.parametrize("alg", [4, 5, 7])
def test_philox_noise_stream_matches_oracle(O, alg):
    """The FP64 engine's in-kernel Philox deviates (double-precision
    Box-Muller on 52-bit uniforms) equal the oracle's restatement of the same
    counter stream to a few ulps, so the trajectories agree to 1e-12."""
    rng = np.random.default_rng(13)
    sysd = CF.bcc_shells(3, [1.34, 0.77])
    M, N = 2, sysd["lattice"].natom
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    seed = 0x5EED1234ABCD
    eng = engine_for(sysd, M)
    eng.set_moments(emom, np.broadcast_to(sysd["mmom"], (M, N)))
    eng.set_llg(0.1, 300.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, seed)
    sim = RC.OracleSD(O, oracle_system(O, sysd), sysd, sysd["mmom"], M=M)
    sim.set_state(emom)
    nsteps = 20
    for step in range(nsteps):
        normals = np.zeros((M, N, 3))
        for k in range(M):
            for i in range(N):
                normals[k, i] = O.thermal3_f64(seed, step, i, k)
        sim.step(alg, 1e-16, 0.1, 300.0, normals=normals)
    eng.sd_steps(alg, nsteps)
    e, _ = eng.get_moments()
    assert np.max(np.abs(e - sim.e)) <= 1e-12


@pytest.mark.parametrize("solver,name", [(1, "Midpoint"), (4, "Heun"), (5, "Depondt")])
def test_solvers_golden_on_gpu(O, solver, name):
    """The reference's Solvers chain (regulartests.yaml) on the device: the
    random start comes from the oracle's MT19937 stream, T = 0."""
    c = RC.read_case("Solvers", "inpsd.dat.base")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    sh = U.Shells.from_lines(lat, c["jlines"])
    nl, ns, nc = U.build_pair_list(lat, sh, c["sym"], c["ammom"].reshape(-1, 1))
    latO = O.LatticeSpec(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    sim = RC.OracleSD(O, latO, dict(nlist=nl, nlistsize=ns, ncoup=nc), np.ones(lat.natom))
    sim.random_init(O.mt(1), latO)
    eng = U.Engine(lat.natom, 1)
    eng.set_exchange(nl, ns, nc)
    eng.set_moments(sim.e, np.ones((1, lat.natom)))
    eng.set_llg(1.0, 0.0, 1.0e-15)
    eng.set_noise(U.UASD_NOISE_NONE)
    g = [x for x in RC.golden("Solvers") if x["type"] == name][0]
    eng.sd_steps(solver, g["select"]["#_Iter."])
    _, eM = eng.get_moments()
    m = eM[0].sum(axis=0) / lat.natom
    assert np.allclose([*m, np.linalg.norm(m)], g["expected"], atol=1e-8)


def test_bccfe_sd_golden_on_gpu(O):
    """bcc Fe at 500 K with the reference's ziggurat stream injected each
    step: <m> after the initial phase and at measurement iteration 2700."""
    c = RC.read_case("bccFe", "inpsd.dat.base")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    sh = U.Shells.from_lines(lat, c["jlines"])
    nl, ns, nc = U.build_pair_list(lat, sh, c["sym"], c["ammom"].reshape(-1, 1))
    N = lat.natom
    eng = U.Engine(N, 1)
    eng.set_exchange(nl, ns, nc)
    e0 = np.tile(c["edir"], (N // lat.NA, 1))[None]
    eng.set_moments(np.ascontiguousarray(e0), np.full((1, N), 2.23))
    z = O.zig(5)
    noise = np.zeros((1, N, 3))
    exp = {g["type"]: g for g in RC.golden("bccFe") if g["name"] == "bcc Fe (SD)"}

    def run(nsteps, dt, stop=None):
        eng.set_llg(0.5, 500.0, dt)
        for s in range(nsteps):
            if stop is not None and s == stop:
                return
            O.lib().or_zig_fill(ctypes.byref(z), noise.reshape(-1), noise.size)
            eng.set_noise(U.UASD_NOISE_INJECTED, 0, noise)
            eng.sd_steps(1, 1)

    run(2000, 1.0e-16)
    _, eM = eng.get_moments()
    m = eM[0].sum(axis=0) / N
    assert np.allclose([*m, np.linalg.norm(m)], exp["Initial phase"]["expected"], atol=1e-8)
    run(3000, float(np.float32(1.0e-16)), stop=2700)
    _, eM = eng.get_moments()
    m = eM[0].sum(axis=0) / N
    assert np.allclose([*m, np.linalg.norm(m)], exp["Magnetization"]["expected"], atol=1e-8)


def test_measure_reductions(O):
    rng = np.random.default_rng(21)
    sysd = CF.triangular_dmi(10)
    M, N = 4, sysd["lattice"].natom
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    eng = engine_for(sysd, M)
    eng.set_moments(emom, np.broadcast_to(sysd["mmom"], (M, N)))
    msum, en = eng.measure(with_energy=True)
    eM = emom * sysd["mmom"][None, :, None]
    assert np.allclose(msum, eM.sum(axis=1), rtol=1e-13, atol=1e-12)
    ham = oracle_ham(O, sysd)
    ext = ext_array(sysd, M)
    for k in range(M):
        _, ek = ham.field(np.ascontiguousarray(eM[k:k + 1]), None if ext is None else ext[k:k + 1])
        assert abs(en[k] - ek) <= 1e-12 * abs(ek)


def test_kagome_dmi_golden_on_gpu(O):
    """The reference's kagome case (DMI on every bond, T = 0 midpoint) on the
    device: <m> at iteration 13000 and atom 2 at iteration 2400."""
    c = RC.read_case("kagome")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    latO = O.LatticeSpec(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    pairs = O.build_pairs(latO, c["jlines"], c["sym"], ammom=c["ammom"].reshape(-1, 1))
    dm = O.build_pairs(latO, c["dmlines"], 0, hdim=3, use_nntype=False,
                       ammom=c["ammom"].reshape(-1, 1))
    N = lat.natom
    eng = U.Engine(N, 1)
    eng.set_exchange(pairs["nlist"], pairs["nlistsize"], pairs["ncoup"])
    eng.set_dm(dm["nlist"], dm["nlistsize"], dm["ncoup"])
    e0 = np.ascontiguousarray(np.tile(c["edir"], (N // lat.NA, 1))[None])
    eng.set_moments(e0, np.tile(c["ammom"], N // lat.NA)[None])
    eng.set_llg(0.5, 0.0, 1.0e-16)
    eng.set_noise(U.UASD_NOISE_NONE)
    gm = [g for g in RC.golden("kagome") if g["type"] == "Magnetization"][0]
    gt = [g for g in RC.golden("kagome") if g["type"] == "Trajectory"][0]
    eng.sd_steps(1, gt["select"]["#_Iter."])
    e, _ = eng.get_moments()
    assert np.allclose([2.0, *e[0, 1], 2.0], gt["expected"], atol=1e-8)
    eng.sd_steps(1, gm["select"]["#_Iter."] - gt["select"]["#_Iter."])
    _, eM = eng.get_moments()
    m = eM[0].sum(axis=0) / N
    assert np.allclose([*m, np.linalg.norm(m)], gm["expected"], atol=1e-8)


def test_index_width_selection_and_int32_path(O):
    """Periodic lattices use int16 index differences; a list whose differences
    do not fit in int16 (random partners over 70,000 sites) keeps int32
    indices, and both give the oracle's field."""
    sysd = CF.bcc_shells(6, CF.bcc5_couplings())
    eng = engine_for(sysd)
    assert eng.index_bytes() == (2, 4)
    eng = engine_for(CF.triangular_dmi(16))
    assert eng.index_bytes() == (2, 2)
    rng = np.random.default_rng(12)
    N, W = 70000, 8
    nlist = rng.integers(1, N + 1, size=(N, W)).astype(np.int32)
    nsize = rng.integers(3, W + 1, size=N).astype(np.int32)
    ncoup = rng.normal(size=(N, W)) * 1e-3
    eng = U.Engine(N, 1)
    eng.set_exchange(nlist, nsize, ncoup)
    assert eng.index_bytes()[0] == 4
    emom = CF.random_unit(N, rng, 1)
    emomM = np.ascontiguousarray(emom * 1.5)
    beff, en = eng.effective_field(emomM)
    bref, eref = O.OracleHam(N, nlist, nsize, ncoup).field(emomM)
    assert rel(beff, bref) <= 1e-13
    assert abs(en - eref) <= 1e-12 * abs(eref)


def test_heischain_from_restart_file_on_gpu(O):
    """The reference's Heisenberg-chain case on the device: the start read from
    its restart file (uppasd_amd.restart), 20000 midpoint steps at 0.1 K with
    damping 4 on the injected ziggurat stream (default tseed 1), then T = 0
    without damping: <M> at iteration 1000 and the energy per atom at
    iteration 1400 against the regression values."""
    from uppasd_amd import restart as RS
    c = RC.read_case("HeisChain", "inpsd.dat")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    sh = U.Shells.from_lines(lat, c["jlines"])
    nl, ns, nc = U.build_pair_list(lat, sh, c["sym"], c["ammom"].reshape(-1, 1))
    N = lat.natom
    _, mmom, emom = RS.read_restart(os.path.join(RC.CASES, "HeisChain", "startmom"), N, 1)
    eng = U.Engine(N, 1)
    eng.set_exchange(nl, ns, nc)
    eng.set_moments(emom, mmom)
    eng.set_llg(4.0, 0.1, 1e-16)
    z = O.zig(1)
    noise = np.zeros((1, N, 3))
    for _ in range(20000):
        O.lib().or_zig_fill(ctypes.byref(z), noise.reshape(-1), noise.size)
        eng.set_noise(U.UASD_NOISE_INJECTED, 0, noise)
        eng.sd_steps(1, 1)
    eng.set_llg(0.0, 0.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_NONE)
    exp = {g["type"]: g["expected"] for g in RC.golden("HeisChain")}
    eng.sd_steps(1, 1000)
    ms, _ = eng.measure()
    m = ms[0] / N
    assert np.allclose([*m, np.linalg.norm(m)], exp["Magnetization"], rtol=0, atol=1e-8)
    eng.sd_steps(1, 400)
    _, en = eng.measure(with_energy=True)
    assert abs(en[0] / N - exp["Total energy"][0]) <= 1e-8


def test_heischain_af_golden_on_gpu(O):
    """The antiferromagnetic two-sublattice chain on the device: the random
    start from the oracle's MT19937 stream (default tseed 1), midpoint at
    T = 0 with damping 0.5; <M> at iteration 1000, calc_energy's total per
    atom at 1500 (uasd_energy_terms) and the per-type averages at 5000
    against the regression values."""
    c = RC.read_case("HeisChainAF", "inpsd.dat")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    nl, ns, nc = U.build_pair_list(lat, U.Shells.from_lines(lat, c["jlines"]), c["sym"],
                                   c["ammom"].reshape(-1, 1))
    latO = O.LatticeSpec(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    N, ncell = lat.natom, lat.natom // lat.NA
    mm = np.tile(c["ammom"], ncell)
    sim = RC.OracleSD(O, latO, dict(nlist=nl, nlistsize=ns, ncoup=nc), mm)
    sim.random_init(O.mt(1), latO)
    print(f"AF chain N={N}", flush=True)
    eng = U.Engine(N, 1)
    eng.set_exchange(nl, ns, nc)
    eng.set_moments(sim.e, mm[None])
    eng.set_llg(0.5, 0.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_NONE)
    exp = {g["type"]: g["expected"] for g in RC.golden("HeisChainAF")}
    eng.sd_steps(1, 1000)
    ms, _ = eng.measure()
    m = ms[0] / N
    assert np.allclose([*m, np.linalg.norm(m)], exp["Magnetization"], rtol=0, atol=1e-8)
    eng.sd_steps(1, 500)
    xc, ani, dm, ze = eng.energy_terms()[0]
    assert abs((xc + dm + ze + ani) - exp["Total energy"][0]) <= 1e-8
    eng.sd_steps(1, 3500)
    _, eM = eng.get_moments()
    for k in (1, 2):
        v = eM[0].reshape(ncell, lat.NA, 3)[:, k - 1].sum(axis=0) / ncell
        assert np.allclose([np.linalg.norm(v), *v], exp[f"Sublattice {k}"], rtol=0, atol=1e-8), k
