"""The C oracle against the reference's own compiled routines.

oracle/_ref/libuppasd_ref.so holds the reference's effective_field_full,
depondt_evolve_first/second, smodeulermpt/modeulermpf, rannum, the MT19937 /
ziggurat generators and moment_update, compiled from its sources
(oracle/Makefile, target ref) and driven in the order of sd_mphase
(oracle/ref_driver.f90).  These tests pin the oracle's restatement of those
routines to rounding: fields with every term of the accelerated Hamiltonian,
T = 0 trajectories of both solvers, and stochastic trajectories on the
reference's own ziggurat stream (one thread, where the reference's draws are
reproducible).  They skip where the library is not built (no reference
sources or amdflang)."""
import ctypes

import numpy as np
import pytest

from uppasd_amd import configs as CF


@pytest.fixture(scope="module")
def R(O):
    if O.ref_lib() is None and O.build_ref() is None:
        pytest.skip("the reference's routines are not built (no sources or amdflang here)")
    L = O.ref_lib()
    L.ref_set_threads(1)
    return L


def _triangular(rng, L=12):
    """J + DMI, anisotropy of types 1, 2 and 7 with random axes, constants and
    sb, uneven moments."""
    s = CF.triangular_dmi(L)
    N = s["lattice"].natom
    e = rng.normal(size=(N, 3))
    s.update(taniso=rng.choice([1, 2, 7], size=N).astype(np.int32),
             eaniso=e / np.linalg.norm(e, axis=1)[:, None],
             kaniso=rng.normal(size=(N, 2)) * 5.0, sb=rng.uniform(0.2, 1.5, size=N),
             mmom=rng.uniform(0.8, 2.4, size=N))
    return s


def _pair(O, s, M):
    N = s["lattice"].natom
    kw = {}
    if "dmlist" in s:
        kw.update(dmlist=s["dmlist"], dmlistsize=s["dmlistsize"], dm_vect=s["dm_vect"])
    if "taniso" in s:
        kw.update(taniso=s["taniso"], eaniso=s["eaniso"], kaniso=s["kaniso"], sb=s["sb"])
    ref = O.RefSD(N, M, s["nlist"], s["nlistsize"], s["ncoup"], **kw)
    ham = O.OracleHam(N, s["nlist"], s["nlistsize"], s["ncoup"], **kw)
    return ref, ham


def _state(rng, s, M):
    N = s["lattice"].natom
    e = np.ascontiguousarray(CF.random_unit(N, rng, M))
    mm = np.ascontiguousarray(np.broadcast_to(s["mmom"], (M, N)), dtype=np.float64)
    return e, mm, np.ascontiguousarray(e * mm[..., None])


@pytest.mark.parametrize("system", ["triangular_dmi_aniso_zeeman", "bcc_58nn"])
def test_field_equals_the_reference(O, R, system):
    rng = np.random.default_rng(71)
    M = 2
    if system == "bcc_58nn":
        s = CF.bcc_shells(5, CF.bcc5_couplings(), moment=2.2)
        ext = None
    else:
        s = _triangular(rng)
        ext = np.ascontiguousarray(rng.normal(size=(M, s["lattice"].natom, 3)) * 3.0)
    ref, ham = _pair(O, s, M)
    ref.set_external_field(ext)
    _, mm, eM = _state(rng, s, M)
    b_ref, en_ref = ref.field(eM, mm)
    b_or, en_or = ham.field(eM, ext)
    scale = np.max(np.abs(b_ref))
    # a few ulps: the exchange sums round alike up to the fused multiply-adds
    assert np.max(np.abs(b_or - b_ref)) <= 8 * 2.2e-16 * scale, np.max(np.abs(b_or - b_ref)) / scale
    # the reference sums the energy over every ensemble into one number
    assert abs(np.sum(en_or) - en_ref) <= 1e-13 * abs(en_ref)


@pytest.mark.parametrize("alg", [1, 5])
@pytest.mark.parametrize("system", ["triangular_dmi_aniso_zeeman", "bcc_58nn"])
def test_zero_temperature_trajectory_equals_the_reference(O, R, alg, system):
    """300 steps at T = 0 (midpoint 1, Depondt 5), two ensembles: every
    moment within 1e-12 of the reference's."""
    rng = np.random.default_rng(72 + alg)
    M = 2
    if system == "bcc_58nn":
        s = CF.bcc_shells(5, CF.bcc5_couplings(), moment=2.2)
        ext = None
    else:
        s = _triangular(rng)
        ext = np.ascontiguousarray(rng.normal(size=(M, s["lattice"].natom, 3)) * 3.0)
    N = s["lattice"].natom
    ref, ham = _pair(O, s, M)
    ref.set_external_field(ext)
    ref.seed(3)
    e, mm, eM = _state(rng, s, M)
    lam = np.full(N, 0.1)
    T = np.zeros(N)
    er, eMr, mr = e.copy(), eM.copy(), mm.copy()
    ref.run(alg, 300, 1e-16, lam, T, mr, er, eMr)
    eo, eMo = e.copy(), eM.copy()
    O.lib().or_sd_run(ctypes.byref(ham.h), N, M, alg, 300, 0, 1e-16, lam, T, np.ones(N), mm, eo,
                      eMo, 0, 0, None if ext is None else ext.ctypes.data)
    d = np.max(np.abs(eo - er))
    print(f"{system} SDEalgh {alg}: max |e_oracle - e_reference| = {d:.2e}")
    assert d <= 1e-12


def test_stochastic_midpoint_on_the_reference_ziggurat_stream(O, R):
    """Midpoint at 300 K: the reference draws its thermal field with rannum
    from its ziggurat stream (seed 5, one thread); the oracle, fed its own
    restatement of that stream (or_zig_fill) through the same scaling, keeps
    every moment within 1e-12 over 200 steps."""
    rng = np.random.default_rng(74)
    s = CF.bcc_shells(5, CF.bcc5_couplings(), moment=2.2)
    N, M = s["lattice"].natom, 1
    ref, ham = _pair(O, s, M)
    ref.seed(5)
    e, mm, eM = _state(rng, s, M)
    lam = np.full(N, 0.1)
    T = np.full(N, 300.0)
    er, eMr, mr = e.copy(), eM.copy(), mm.copy()
    ref.run(1, 200, 1e-16, lam, T, mr, er, eMr)
    z = O.zig(5)
    noise = np.zeros((M, N, 3))
    import refcase as RC

    class _Lat:
        natom = N
    sim = RC.OracleSD(O, _Lat, s, s["mmom"])
    sim.set_state(e)
    for _ in range(200):
        O.lib().or_zig_fill(ctypes.byref(z), noise.reshape(-1), noise.size)
        sim.step(1, 1e-16, 0.1, 300.0, normals=noise)
    d = np.max(np.abs(sim.e - er))
    print(f"stochastic midpoint: max |e_oracle - e_reference| = {d:.2e}")
    assert d <= 1e-12


def test_stochastic_depondt_on_the_reference_ziggurat_stream(O, R):
    """Depondt at 300 K: depondt_evolve_first draws its N(0,1) deviates with
    rng_gaussianP from the reference's ziggurat stream (seed 5, one thread)
    and scales them itself; the oracle, fed its restatement of the stream,
    keeps every moment within 1e-12 over 200 steps."""
    rng = np.random.default_rng(75)
    s = CF.bcc_shells(5, CF.bcc5_couplings(), moment=2.2)
    N, M = s["lattice"].natom, 1
    ref, ham = _pair(O, s, M)
    ref.seed(5)
    e, mm, eM = _state(rng, s, M)
    er, eMr, mr = e.copy(), eM.copy(), mm.copy()
    ref.run(5, 200, 1e-16, np.full(N, 0.1), np.full(N, 300.0), mr, er, eMr)
    import refcase as RC

    class _Lat:
        natom = N
    sim = RC.OracleSD(O, _Lat, s, s["mmom"])
    sim.set_state(e)
    z = O.zig(5)
    noise = np.zeros((M, N, 3))
    for _ in range(200):
        O.lib().or_zig_fill(ctypes.byref(z), noise.reshape(-1), noise.size)
        sim.step(5, 1e-16, 0.1, 300.0, normals=noise)
    d = np.max(np.abs(sim.e - er))
    print(f"stochastic Depondt: max |e_oracle - e_reference| = {d:.2e}")
    assert d <= 1e-12


def _nm_cases():
    import refcase as RC
    import uppasd_amd as U
    from uppasd_amd.lattice import BCC_SHELLS
    bcc5 = [(1, 1, 1, 1, *BCC_SHELLS[k], j) for k, j in enumerate(CF.bcc5_couplings())]
    out = []
    for n, bc in (((6, 6, 6), "PPP"), ((4, 5, 3), "PPP"), ((4, 4, 4), "P0P"), ((3, 1, 7), "00P")):
        out.append((f"bcc5_{n[0]}{n[1]}{n[2]}_{bc}", U.bcc(n, tuple(bc)), bcc5, 1, 1, True))
    t = CF.triangular_dmi(8)
    out.append(("triangular_J_sym0", t["lattice"], t["jlines"], 0, 1, True))
    out.append(("triangular_DM_sym0", t["lattice"], t["dmlines"], 0, 3, False))
    out.append(("triangular_hexagonal_sym3", t["lattice"], [(1, 1, 1, 1, 1.0, 0.0, 0.0, 1.0)], 3,
                1, True))
    k = CF.kagome_afm(5)
    out.append(("kagome_sym0", k["lattice"], k["jlines"], 0, 1, True))
    c = RC.read_case("Regression", "inpsd.dat")
    out.append(("regression_maptype2", U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"]),
                c["jlines"], c["sym"], 1, True))
    return out


@pytest.mark.parametrize("case", _nm_cases(), ids=lambda c: c[0])
def test_neighbour_map_equals_the_reference_setup_nm(O, R, case):
    """setup_nm (neighbourmap.f90:32-359) of the reference, compiled from its
    source, and the oracle's restatement give the same neighbour map, entry
    for entry (atom, shell, equivalent site, in the reference's symmetry
    order): the order the neighbour lists and so the FP64 exchange sums
    follow.  The product's builder equals the oracle's lists bit for bit
    (test_builder.py)."""
    _, lat, lines, sym, hdim, use_nntype = case
    latO = O.LatticeSpec(lat.cell, lat.bas, lat.atype, lat.n, lat.bc)
    a = O.shells_and_map(latO, lines, sym, hdim=hdim, use_nntype=use_nntype)
    b = O.ref_setup_nm(latO, a, sym, use_nntype)
    assert (a["max_no_neigh"], a["max_no_equiv"]) == (b["max_no_neigh"], b["max_no_equiv"])
    assert a["nm"].shape == b["nm"].shape and np.array_equal(a["nm"], b["nm"])
    assert np.array_equal(a["nmdim"], b["nmdim"])
