"""GPU Monte Carlo against the oracle.

* Deterministic: the coloured Metropolis and heat-bath sweeps use the same
  Philox counters on both sides, so after a few sweeps the spins agree to
  rounding (the oracle applies the reference's per-site move, energy and
  acceptance rules in the same colour order).
* Statistical: bcc Fe at 500 K from the reference's tests/bccFe M and H
  cases; the cumulant averages <M>, <M^2>, <M^4>, U4 of the coloured GPU
  sweeps must match the reference's golden values within its 'sloppy'
  comparison (2 %).
"""
import ctypes

import numpy as np
import pytest

import refcase as RC
import uppasd_amd as U
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu


def _oracle_ham(O, sysd):
    return O.OracleHam(sysd["lattice"].natom, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])


def _colour_lists(ncol, col):
    order = np.argsort(col, kind="stable").astype(np.int32)
    ptr = np.zeros(ncol + 1, dtype=np.int32)
    for k in range(ncol):
        ptr[k + 1] = ptr[k] + np.count_nonzero(col == k)
    return ptr, np.ascontiguousarray(order)


@pytest.mark.parametrize("mode", ["M", "H"])
def test_coloured_sweeps_match_oracle(O, mode):
    rng = np.random.default_rng(17)
    L = 4
    occ = rng.integers(1, 3, size=2 * L ** 3).astype(np.int32)
    sysd = CF.bcc_alloy(L, occ)
    N, M = sysd["lattice"].natom, 4
    temps = np.array([200.0, 600.0, 1000.0, 1600.0])
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    mm = np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (M, N)))
    eng = U.Engine(N, M)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_moments(emom, mm)
    ncol, col = eng.coloring()
    assert ncol == 2
    ptr, sites = _colour_lists(ncol, col)
    seed, nsweeps = 987654321, 6
    eng.mc_sweeps(mode, temps, nsweeps, seed)
    e_gpu, _ = eng.get_moments()
    ham = _oracle_ham(O, sysd)
    e = emom.copy()
    eM = np.ascontiguousarray(e * mm[..., None])
    ext = np.zeros(3)
    for s in range(nsweeps):
        O.lib().or_mc_sweep_colored(ctypes.byref(ham.h), N, M, temps, mode.encode(), seed, s,
                                    ncol, ptr, sites, eM, e, mm, ext)
    assert np.max(np.abs(e_gpu - e)) <= 1e-12


def test_flip_moves_reverse_the_spin(O):
    """Regression: the spin-flip trial (newmom = -emom) must reverse the spin.
    All spins along +z in a strong field along -z: every site that draws the
    flip move (3*u0 >= 2) lowers its Zeeman energy and ends exactly at -z."""
    sysd = CF.bcc_shells(4, [1.0])
    N = sysd["lattice"].natom
    eng = U.Engine(N, 1)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_external_field(np.array([0.0, 0.0, -20000.0]))  # Zeeman gain >> exchange cost
    e0 = np.zeros((1, N, 3))
    e0[..., 2] = 1.0
    eng.set_moments(e0, sysd["mmom"][None])
    seed = 31337
    eng.mc_sweeps("M", [10.0], 1, seed)
    e, _ = eng.get_moments()
    flips = 0
    for i in range(N):
        u = np.zeros(4)
        O.lib().or_uniform4(seed, 0x21, 0, i, 0, u)
        if int(np.floor(3 * u[0])) == 2:
            flips += 1
            assert np.array_equal(np.abs(e[0, i]), [0.0, 0.0, 1.0]) and e[0, i, 2] == -1.0, i
    assert flips > N // 5


def test_colouring_is_proper():
    sysd = CF.bcc_shells(5, CF.bcc5_couplings())
    eng = U.Engine(sysd["lattice"].natom, 1)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    ncol, col = eng.coloring()
    nl, ns = sysd["nlist"], sysd["nlistsize"]
    for i in range(nl.shape[0]):
        nb = nl[i, :ns[i]] - 1
        assert not np.any(col[nb] == col[i])
    assert ncol >= 2


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("mode,label", [("M", "bcc Fe (MC)"), ("H", "bcc Fe (HB)")])
def test_bccfe_mc_statistics_vs_reference(mode, label, precision):
    """bccFe at 500 K through the device sweeps (both precisions; the FP32
    engine draws its trial moves in single precision): the cumulant row of
    the reference's regression run within 2 %."""
    c = RC.read_case("bccFe", "inpsd.dat.base")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    sh = U.Shells.from_lines(lat, c["jlines"])
    nl, ns, nc = U.build_pair_list(lat, sh, c["sym"], c["ammom"].reshape(-1, 1))
    N, M = lat.natom, 16
    eng = U.Engine(N, M, precision)
    eng.set_exchange(nl, ns, nc)
    e0 = np.ascontiguousarray(np.broadcast_to(np.tile(c["edir"], (N // lat.NA, 1)), (M, N, 3)))
    eng.set_moments(e0, np.full((M, N), 2.23))
    temps = np.full(M, 500.0)
    eng.mc_sweeps(mode, temps, 2000, 4242)          # ip_mcanneal 2000 sweeps
    w = tot = 0.0
    acc = np.zeros((M, 3))
    for s in range(1, 2051):                       # cumulant every 50 sweeps (row 41)
        if s % 50 == 0:
            ms, _ = eng.measure()
            me = np.linalg.norm(ms, axis=1) / N
            w += 1.0
            x = np.stack([me, me ** 2, me ** 4], axis=1)
            acc = (acc * tot + x * w) / (tot + w)
            tot += w
        eng.mc_sweeps(mode, temps, 1, 4242)
    mean = acc.mean(axis=0)
    U4 = 1 - mean[2] / 3 / mean[1] ** 2
    g = [x for x in RC.golden("bccFe") if x["name"] == label and x["type"] == "Cumulants"][0]
    want = g["expected"][:4]
    got = [mean[0], mean[1], mean[2], U4]
    assert np.allclose(got, want, rtol=0.02), (got, want)


def test_bccfe_mc_files_through_writer(tmp_path):
    """The device MC run written by output.MeasurementPhase: cumulant row 41
    of cumulants.<simid>.out against the reference's bccFe (MC) golden
    values, 2 % (a statistical comparison: the sweep order and random
    stream differ from the reference's)."""
    import sys, os
    sys.path.insert(0, os.path.dirname(__file__))
    from test_output import read_selected
    from uppasd_amd import output as OUT
    c = RC.read_case("bccFe", "inpsd.dat.base")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    sh = U.Shells.from_lines(lat, c["jlines"])
    nl, ns, nc = U.build_pair_list(lat, sh, c["sym"], c["ammom"].reshape(-1, 1))
    N, M = lat.natom, 16
    eng = U.Engine(N, M)
    eng.set_exchange(nl, ns, nc)
    e0 = np.ascontiguousarray(np.broadcast_to(np.tile(c["edir"], (N // lat.NA, 1)), (M, N, 3)))
    eng.set_moments(e0, np.full((M, N), 2.23))
    temps = np.full(M, 500.0)
    eng.mc_sweeps("M", temps, 2000, 4242)
    ph = OUT.MeasurementPhase(eng, "bccFeM", 500.0, avrg_step=50, cumu_step=50,
                              outdir=str(tmp_path), cumu_buff=10)
    ph.run(2100, lambda n: eng.mc_sweeps("M", temps, n, 4242))
    g = [x for x in RC.golden("bccFe") if x["name"] == "bcc Fe (MC)" and x["type"] == "Cumulants"][0]
    got = read_selected(tmp_path / "cumulants.bccFeM.out", g["headers"], g["select"],
                        g["extract"][:4])
    assert np.allclose(got, g["expected"][:4], rtol=0.02), (got, g["expected"])
