"""Term-resolved energies on the device (uasd_energy_terms, the terms of
calc_energy, energy.f90:60-366) against the oracle's restatement, per
ensemble and per segment, and the reference's measurement regression case
(tests/Regression, simid megaTest) run on the device through the
reference-format writers, compared with the reference's printed values."""
import numpy as np
import pytest

import refcase as RC
import uppasd_amd as U
from uppasd_amd import batch as BT
from uppasd_amd import configs as CF
from uppasd_amd import output as OUT

pytestmark = pytest.mark.gpu
FC = OUT.MUB / OUT.MRY


def _system(rng, L):
    """Triangular lattice with exchange, DM, anisotropy of types 1, 2 and 7
    (random axes, constants and sb) and moments of uneven size."""
    s = CF.triangular_dmi(L)
    N = s["lattice"].natom
    e = rng.normal(size=(N, 3))
    s.update(taniso=rng.choice([1, 2, 7], size=N).astype(np.int32),
             eaniso=e / np.linalg.norm(e, axis=1)[:, None],
             kaniso=rng.normal(size=(N, 2)) * 5.0, sb=rng.uniform(0.2, 1.5, size=N),
             mmom=rng.uniform(0.8, 2.4, size=N))
    return s


def _oracle(O, s):
    return O.OracleHam(s["lattice"].natom if "lattice" in s else s["nlist"].shape[0],
                       s["nlist"], s["nlistsize"], s["ncoup"], dmlist=s["dmlist"],
                       dmlistsize=s["dmlistsize"], dm_vect=s["dm_vect"], taniso=s["taniso"],
                       eaniso=s["eaniso"], kaniso=s["kaniso"], sb=s["sb"])


def _moments(rng, M, N, mmom, precision):
    emom = CF.random_unit(N, rng, M)
    mm = np.broadcast_to(mmom, (M, N))
    if precision == "fp32":     # the device holds e and |m| in FP32 and forms e*|m| in FP32
        eM = (emom.astype(np.float32) * mm.astype(np.float32)[..., None]).astype(np.float64)
    else:
        eM = emom * mm[..., None]
    return emom, np.ascontiguousarray(eM)


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5)])
@pytest.mark.parametrize("field", ["uniform", "per_site"])
def test_energy_terms_match_oracle(O, precision, tol, field):
    rng = np.random.default_rng(21)
    s = _system(rng, 16)
    N, M = s["lattice"].natom, 3
    ext = (np.broadcast_to(rng.normal(size=3) * 2.0, (M, N, 3)) if field == "uniform"
           else rng.normal(size=(M, N, 3)) * 2.0)
    ext = np.ascontiguousarray(ext)
    print(f"energy terms {precision} field={field} N={N} M={M}", flush=True)
    eng = U.Engine(N, M, precision)
    eng.set_exchange(s["nlist"], s["nlistsize"], s["ncoup"])
    eng.set_dm(s["dmlist"], s["dmlistsize"], s["dm_vect"])
    eng.set_anisotropy(s["taniso"], s["eaniso"], s["kaniso"], s["sb"])
    eng.set_external_field(ext[0, 0] if field == "uniform" else ext)
    emom, eM = _moments(rng, M, N, s["mmom"], precision)
    eng.set_moments(emom, s["mmom"])
    got = eng.energy_terms()
    want = _oracle(O, s).energy_terms(eM, ext) * FC
    assert got.shape == (M, 4)
    for t in range(4):
        scale = np.max(np.abs(want[:, t]))
        assert scale > 0.0
        assert np.max(np.abs(got[:, t] - want[:, t])) <= tol * scale, (t, got[:, t], want[:, t])


def test_energy_terms_per_segment(O):
    """Two replicas of different size in one context: each segment's terms
    are its own system's, divided by its own number of sites."""
    rng = np.random.default_rng(22)
    systems = [_system(rng, 8), _system(rng, 12)]
    j = BT.join(systems)
    M = 2
    print(f"energy terms per segment offsets={list(j.offsets)} M={M}", flush=True)
    eng = BT.engine_for(j, M)
    emom = CF.random_unit(j.natom, rng, M)
    eng.set_moments(emom, j.mmom)
    got = eng.energy_terms()
    assert got.shape == (2, M, 4)
    for r, s in enumerate(systems):
        a, b = j.offsets[r], j.offsets[r + 1]
        eM = np.ascontiguousarray(emom[:, a:b] * s["mmom"][None, :, None])
        want = _oracle(O, s).energy_terms(eM) * FC
        assert np.allclose(got[r], want, rtol=1e-12, atol=1e-12 * np.max(np.abs(want)))


def test_regression_case_files_on_gpu(tmp_path):
    """The measurement regression case on the device: exchange read in
    maptype 2 by the product's reader, the Zeeman field uploaded, 11000
    midpoint steps at T = 0 through MeasurementPhase; averages at 11000,
    totenergy at 10900 and cumulant row 211 against the reference's values
    (regressionResaro.yaml)."""
    from test_output import read_selected
    c = RC.read_case("Regression", "inpsd.dat")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    raw = [(int(r[0]), int(r[1]), 1, 1, *[float(x) for x in r[2:6]])
           for r in RC._rows(RC.os.path.join(RC.CASES, "Regression", "jfile"))]
    nl, ns, nc = U.build_pair_list(lat, U.Shells.from_lines(lat, raw, maptype=2), c["sym"],
                                   c["ammom"].reshape(-1, 1))
    N = lat.natom
    print(f"regression case N={N} steps=11000", flush=True)
    eng = U.Engine(N, 1)
    eng.set_exchange(nl, ns, nc)
    eng.set_external_field(c["hfield"])
    eng.set_moments(np.tile(c["edir"], (N // lat.NA, 1))[None], np.tile(c["ammom"], N // lat.NA))
    eng.set_llg(0.3, 0.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_NONE)
    ph = OUT.MeasurementPhase(eng, "megaTest", 0.0, avrg_step=100, cumu_step=50,
                              outdir=str(tmp_path), cumu_buff=10)
    ph.run(11000, lambda n: eng.sd_steps(1, n))
    checked = 0
    for g in RC.golden("Regression"):
        got = read_selected(tmp_path / g["outfile"], g["headers"], g["select"], g["extract"])
        assert np.allclose(got, g["expected"], rtol=1e-8, atol=1e-8), (g["type"], got,
                                                                       g["expected"])
        checked += 1
    assert checked == 3


def test_feco_case_files_on_gpu(tmp_path):
    """FeCo B2 on the device (maptype-2 exchange read by the product's
    reader, 255 neighbours, T = 0 midpoint, damping 0.001, 5000 steps)
    through MeasurementPhase: the averages and cumulant rows against the
    reference's values with its own comparison functions, the cumulant row
    also to 1e-8."""
    from test_output import check_feco_files
    c = RC.read_case("FeCo", "inpsd.dat")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    raw = [(int(r[0]), int(r[1]), 1, 1, *[float(x) for x in r[2:6]])
           for r in RC._rows(RC.os.path.join(RC.CASES, "FeCo", "jASD"))]
    nl, ns, nc = U.build_pair_list(lat, U.Shells.from_lines(lat, raw, maptype=2), c["sym"],
                                   c["ammom"].reshape(-1, 1))
    N = lat.natom
    print(f"FeCo B2 N={N} steps=5000", flush=True)
    eng = U.Engine(N, 1)
    eng.set_exchange(nl, ns, nc)
    eng.set_moments(np.tile(c["edir"], (N // lat.NA, 1))[None], np.tile(c["ammom"], N // lat.NA))
    eng.set_llg(0.001, 0.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_NONE)
    ph = OUT.MeasurementPhase(eng, "test", 0.0, avrg_step=100, cumu_step=20,
                              outdir=str(tmp_path), cumu_buff=10)
    ph.run(5000, lambda n: eng.sd_steps(1, n))
    check_feco_files(str(tmp_path))


def test_heischain_af_files_on_gpu(O, tmp_path):
    """The antiferromagnetic chain on the device through MeasurementPhase
    with the projected-averages writer: averages, projavgs (both types) and
    totenergy against the reference's values to 1e-8."""
    from test_output import check_af_files
    c = RC.read_case("HeisChainAF", "inpsd.dat")
    lat = U.Lattice(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    nl, ns, nc = U.build_pair_list(lat, U.Shells.from_lines(lat, c["jlines"]), c["sym"],
                                   c["ammom"].reshape(-1, 1))
    latO = O.LatticeSpec(c["cell"], c["bas"], c["atype"], c["n"], c["bc"])
    N, ncell = lat.natom, lat.natom // lat.NA
    mm = np.tile(c["ammom"], ncell)
    sim = RC.OracleSD(O, latO, dict(nlist=nl, nlistsize=ns, ncoup=nc), mm)
    sim.random_init(O.mt(1), latO)          # the reference's random start
    print(f"AF chain files N={N}", flush=True)
    eng = U.Engine(N, 1)
    eng.set_exchange(nl, ns, nc)
    eng.set_moments(sim.e, mm[None])
    eng.set_llg(0.5, 0.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_NONE)
    proj = OUT.ProjAverages("AF_WireT", c["atype"], ncell, "Y", str(tmp_path))
    ph = OUT.MeasurementPhase(eng, "AF_WireT", 0.0, avrg_step=100, cumu_step=50,
                              outdir=str(tmp_path), proj=proj)
    ph.run(5000, lambda n: eng.sd_steps(1, n))
    check_af_files(str(tmp_path))
