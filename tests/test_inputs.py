"""The product's input layer (uppasd_amd.inputs) against the test-side
readers and the oracle's builders, which were written independently.

* Every committed reference case directory loads with one call; its
  exchange (and DM) lists and couplings equal the oracle's read_exchange +
  setup_nm + setup_neighbour_hamiltonian bit for bit, from the test-side
  reader's parse of the same files.
* setup_anisotropies equals the oracle's restatement bit for bit (types 1, 2,
  7, unnormalised axes, moments 0.7 ... 2.3), and its scaling makes the MC
  anisotropy energy (a separate oracle routine) the input's K1, K2 expression
  in mRy for any |m| (uniaxial, cubic).
* read_inpsd: list-directed values across lines, comments, case-insensitive
  keywords, defaults (the single-precision default timestep), template
  placeholders, refused options outside the accelerated path.
"""
import ctypes
import os

import numpy as np
import pytest

import refcase as RC
from uppasd_amd import inputs as IN

CASES = [("Solvers", "inpsd.dat.base"), ("HeisStripe", "inpsd.dat"), ("kagome", "inpsd.dat"),
         ("FeCo", "inpsd.dat"), ("HeisChain", "inpsd.dat"), ("HeisChainAF", "inpsd.dat"),
         ("bccFe", "inpsd.dat.base"), ("Regression", "inpsd.dat")]

MUB, MRY = 9.274009994e-24, 2.179872325e-21


@pytest.mark.parametrize("case,base", CASES)
def test_case_lists_equal_oracle_builder(O, case, base):
    c = IN.load_case(os.path.join(RC.CASES, case), base, subst={"SOLVER": 1})
    h = c.hamiltonian()
    r = RC.read_case(case, base)
    assert np.array_equal(c.lattice.bas, r["bas"]) and np.array_equal(c.lattice.atype, r["atype"])
    assert tuple(c.lattice.n) == r["n"] and np.array_equal(c.lattice.cell, r["cell"])
    mm, em = c.site_moments()
    na = c.lattice.NA
    assert np.array_equal(mm, np.tile(r["ammom"], c.natom // na))
    assert np.allclose(em, np.tile(r["edir"], (c.natom // na, 1)), rtol=0, atol=1e-15)
    latO = O.LatticeSpec(r["cell"], r["bas"], r["atype"], r["n"], r["bc"])
    pairs = O.build_pairs(latO, r["jlines"], r["sym"], ammom=r["ammom"].reshape(-1, 1))
    W = pairs["nlist"].shape[1]
    assert np.array_equal(h["nlistsize"], pairs["nlistsize"])
    for i in range(c.natom):
        n = h["nlistsize"][i]
        assert np.array_equal(h["nlist"][i, :n], pairs["nlist"][i, :n]), i
        assert np.array_equal(h["ncoup"][i, :n], pairs["ncoup"][i, :n]), i
    if "dmlines" in r:
        dm = O.build_pairs(latO, r["dmlines"], 0, hdim=3, use_nntype=False,
                           ammom=r["ammom"].reshape(-1, 1))
        assert np.array_equal(h["dmlistsize"], dm["nlistsize"])
        for i in range(c.natom):
            n = h["dmlistsize"][i]
            assert np.array_equal(h["dmlist"][i, :n], dm["nlist"][i, :n])
            assert np.array_equal(h["dm_vect"][i, :n], dm["ncoup"][i, :n])
    else:
        assert "dmlist" not in h
    if "ktype" in r:
        an = O.setup_anisotropies(latO, r["ktype"], r["kvals"], r["ammom"])
        for a, b in zip((h["taniso"], h["eaniso"], h["kaniso"], h["sb"]), an):
            assert np.array_equal(a, b)
    assert W >= 1


@pytest.mark.parametrize("m", [0.7, 1.0, 2.3])
def test_setup_anisotropies_equals_oracle_and_gives_mry(O, m):
    rng = np.random.default_rng(11)
    na = 3
    types = np.array([1, 2, 7], dtype=np.int32)
    vals = np.column_stack([rng.uniform(-0.4, 0.4, na), rng.uniform(-0.2, 0.2, na),
                            rng.normal(size=(na, 3)) * 3.0, rng.uniform(0.1, 2.0, na)])
    ammom = np.array([m, 1.3 * m, 0.8 * m])
    lat = O.LatticeSpec(np.eye(3), rng.uniform(size=(na, 3)), [1, 1, 1], (2, 1, 1))
    got = IN.setup_anisotropies(lat.natom, na, types.reshape(na, 1), vals.reshape(na, 1, 6),
                                ammom.reshape(na, 1))
    want = O.setup_anisotropies(lat, types, vals, ammom)
    for a, b in zip(got, want):
        assert np.array_equal(a, b)
    taniso, eaniso, kaniso, sb = got
    for i in (0, 1):                       # uniaxial, cubic: energy in mRy, any |m|
        mi = ammom[i]
        ham = O.OracleHam(1, np.zeros((1, 1), np.int32), np.zeros(1, np.int32), np.zeros((1, 1)),
                          taniso=taniso[i:i + 1], eaniso=eaniso[i:i + 1], kaniso=kaniso[i:i + 1],
                          sb=sb[i:i + 1])
        a = eaniso[i]
        e0 = np.cross(a, [1.0, 0.0, 0.0]) if i == 0 else np.array([1.0, 0.0, 0.0])
        e0 /= np.linalg.norm(e0)
        for _ in range(3):
            ed = rng.normal(size=3)
            ed /= np.linalg.norm(ed)
            emom = np.ascontiguousarray(e0[None, None])
            mm = np.full((1, 1), mi)
            de = O.lib().or_mc_energy_change(ctypes.byref(ham.h), 1,
                                             np.ascontiguousarray(emom * mi), emom, mm, 0, 0,
                                             np.ascontiguousarray(ed), np.zeros(3)) / MRY
            K1, K2 = vals[i, 0], vals[i, 1]
            if i == 0:
                t, t0 = ed.dot(a), e0.dot(a)
                want_e = K1 * (t * t - t0 * t0) + K2 * (t ** 4 - t0 ** 4)
            else:
                c4 = ed[0] ** 2 * ed[1] ** 2 + ed[1] ** 2 * ed[2] ** 2 + ed[2] ** 2 * ed[0] ** 2
                want_e = -K1 * c4 - K2 * (ed[0] * ed[1] * ed[2]) ** 2
            assert de == pytest.approx(want_e, rel=1e-11, abs=1e-15), (i, m)


def test_read_inpsd_forms(tmp_path):
    p = tmp_path / "inpsd.dat"
    p.write_text("""simid  test
# a comment line
NCELL  4 5
       6
Cell   1.0 0.0 0.0
       0.0 2.0 0.0   0 0 3.0
bc P 0 P      trailing words are ignored
SYM 1
posfile ./posfile
Temp 12.5
timestep 2.5d-16
hfield 0.0 0.1
       0.2
damping  LAMBDA
""")
    v = IN.read_inpsd(str(p), subst={"LAMBDA": "0.25"})
    assert v["ncell"] == [4, 5, 6]
    assert np.array_equal(v["cell"], [[1, 0, 0], [0, 2, 0], [0, 0, 3]])
    assert v["bc"] == ["P", "0", "P"] and v["sym"] == 1 and v["posfile"] == "./posfile"
    assert v["temp"] == 12.5 and v["timestep"] == 2.5e-16 and v["damping"] == 0.25
    assert v["hfield"] == [0.0, 0.1, 0.2]
    assert v["mode"] == "S" and v["sdealgh"] == 1 and v["mensemble"] == 1
    assert v["timestep"] != 1.0e-16
    q = tmp_path / "inpsd2.dat"
    q.write_text("cell 1 0 0 0 1 0 0 0 1\n")
    assert IN.read_inpsd(str(q))["timestep"] == float(np.float32(1.0e-16))
    q.write_text("cell 1 0 0 0 1 0 0 0 1\naunits Y\n")
    with pytest.raises(IN.InputError):
        IN.read_inpsd(str(q))


def test_random_alloy_case_takes_species(tmp_path):
    """do_ralloy 1: alloy posfile and jfile columns, per-species moments and
    couplings; the species of every site come from the caller."""
    (tmp_path / "inpsd.dat").write_text("""ncell 3 3 3
BC P P P
cell 1 0 0
     0 1 0
     0 0 1
sym 1
do_ralloy 1
posfile ./posfile
momfile ./momfile
exchange ./jfile
anisotropy ./kfile
""")
    (tmp_path / "posfile").write_text("1 1 1 0.5 0.0 0.0 0.0\n1 1 2 0.5 0.0 0.0 0.0\n"
                                      "2 1 1 0.5 0.5 0.5 0.5\n2 1 2 0.5 0.5 0.5 0.5\n")
    (tmp_path / "momfile").write_text("1 1 2.6 0 0 1\n1 2 1.7 0 0 1\n2 1 2.6 0 0 1\n2 2 1.7 0 0 1\n")
    (tmp_path / "jfile").write_text("1 1 1 1 0.5 0.5 0.5 1.2\n1 1 1 2 0.5 0.5 0.5 2.0\n"
                                    "1 1 2 1 0.5 0.5 0.5 2.0\n1 1 2 2 0.5 0.5 0.5 0.4\n")
    (tmp_path / "kfile").write_text("1 1 1 0.1 0.0 0 0 1 0\n1 2 2 0.2 0.0 1 0 0 0\n"
                                    "2 1 1 0.1 0.0 0 0 1 0\n2 2 2 0.2 0.0 1 0 0 0\n")
    with pytest.raises(IN.InputError):
        IN.load_case(str(tmp_path))
    rng = np.random.default_rng(3)
    achem = rng.integers(1, 3, size=54).astype(np.int32)
    c = IN.load_case(str(tmp_path), achem=achem)
    h = c.hamiltonian()
    mm, _ = c.site_moments()
    assert np.array_equal(mm, np.where(achem == 1, 2.6, 1.7))
    from uppasd_amd import configs as CF
    ref = CF.bcc_alloy(3, achem)
    assert np.array_equal(h["nlist"], ref["nlist"]) and np.array_equal(h["ncoup"], ref["ncoup"])
    assert np.array_equal(h["taniso"], np.where(achem == 1, 1, 2))
    fc = IN.MRY_OVER_MUB
    want_k1 = np.where(achem == 1, fc * 0.1 / 2.6 ** 2, fc * 0.2 / (1.7 ** 2) ** 2)
    assert np.allclose(h["kaniso"][:, 0], want_k1, rtol=1e-15)
