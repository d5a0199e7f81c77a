"""Reference case directories run on the device through the product's input
layer (uppasd_amd.inputs.load_case -> Case.engine): one call reads
inpsd.dat, posfile, momfile, jfile / dmfile / kfile, builds the lists and
anisotropy and uploads everything; the runs reproduce the golden values of
the reference's regression suite (tests/regulartests.yaml).  Random starts
(initmag 1) come from the oracle's MT19937 magninit, as the reference draws
them; initmag 3 starts from the momfile directions.
"""
import os

import numpy as np
import pytest

import refcase as RC
from uppasd_amd import inputs as IN

pytestmark = pytest.mark.gpu


def _mt_start(O, c):
    lat = c.lattice
    latO = O.LatticeSpec(lat.cell, lat.bas, lat.atype, lat.n, lat.bc)
    h = c.hamiltonian()
    mm, _ = c.site_moments()
    sim = RC.OracleSD(O, latO, h, mm)
    sim.random_init(O.mt(1), latO)
    return sim.e.copy()


def _mavg(eng, N):
    ms, _ = eng.measure()
    m = ms[0] / N
    return [*m, np.linalg.norm(m)]


def test_heisstripe_case_on_gpu(O):
    """Uniaxial anisotropy from the case's kfile (K1 = -0.02 mRy along y),
    vacuum in x, T = 0 midpoint from the MT19937 start: <m> at iteration 1190."""
    c = IN.load_case(os.path.join(RC.CASES, "HeisStripe"))
    assert c.params["initmag"] == 1
    print(f"HeisStripe N={c.natom}", flush=True)
    eng = c.engine(emom=_mt_start(O, c))
    g = [x for x in RC.golden("HeisStripe") if x["type"] == "Magnetization"][0]
    eng.sd_steps(c.params["sdealgh"], g["select"]["#_Iter."])
    got = _mavg(eng, c.natom)
    assert np.allclose(got, g["expected"], rtol=0, atol=1e-8), (got, g["expected"])


@pytest.mark.parametrize("solver,name", [(1, "Midpoint"), (4, "Heun"), (5, "Depondt")])
def test_solvers_case_on_gpu(O, solver, name):
    c = IN.load_case(os.path.join(RC.CASES, "Solvers"), "inpsd.dat.base", subst={"SOLVER": solver})
    assert c.params["sdealgh"] == solver
    eng = c.engine(emom=_mt_start(O, c))
    g = [x for x in RC.golden("Solvers") if x["type"] == name][0]
    eng.sd_steps(solver, g["select"]["#_Iter."])
    assert np.allclose(_mavg(eng, c.natom), g["expected"], rtol=0, atol=1e-8)


def test_kagome_case_on_gpu():
    """DM vectors from the case's dmfile (sym 0 map), initmag 3: <m> at 13000."""
    c = IN.load_case(os.path.join(RC.CASES, "kagome"))
    assert c.params["initmag"] == 3
    eng = c.engine()
    g = [x for x in RC.golden("kagome") if x["type"] == "Magnetization"][0]
    eng.sd_steps(1, g["select"]["#_Iter."])
    assert np.allclose(_mavg(eng, c.natom), g["expected"], rtol=0, atol=1e-8)


def test_feco_case_directory_on_gpu(tmp_path):
    """maptype-2 exchange (255 neighbours) and two species from the case's
    files, through the measurement writer: the reference's rows with its own
    comparison functions."""
    from test_output import check_feco_files
    from uppasd_amd import output as OUT
    c = IN.load_case(os.path.join(RC.CASES, "FeCo"))
    assert c.params["maptype"] == 2
    eng = c.engine()
    ph = OUT.MeasurementPhase(eng, "test", c.params["temp"], avrg_step=100, cumu_step=20,
                              outdir=str(tmp_path), cumu_buff=10)
    ph.run(c.params["nstep"], lambda n: eng.sd_steps(c.params["sdealgh"], n))
    check_feco_files(str(tmp_path))
