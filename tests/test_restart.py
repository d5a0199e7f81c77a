"""Restart and moment files in the reference's layout (uppasd_amd/restart.py):
the reference's own restart file of the HeisChain case is read, files written
here have the reference's header and edit descriptors and read back, and the
legacy layout (step line, no header, no normalisation) is read too."""
import os

import numpy as np
import pytest

import refcase as RC
from uppasd_amd import output as OUT
from uppasd_amd import restart as RS


def test_reads_reference_restart_file():
    path = os.path.join(RC.CASES, "HeisChain", "startmom")
    rstep, mmom, emom = RS.read_restart(path, 100, 1)
    assert rstep == 0 and mmom.shape == (1, 100) and np.all(mmom == 1.0)
    first = np.array([-2.97458276e-03, -4.15411622e-03, 9.99986948e-01])
    assert np.allclose(emom[0, 0], first / np.linalg.norm(first), rtol=0, atol=1e-15)
    assert np.allclose(np.linalg.norm(emom, axis=2), 1.0, rtol=0, atol=1e-15)
    with pytest.raises(ValueError):
        RS.read_restart(path, 101, 1)


def test_restart_layout_and_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    M, N = 2, 7
    e = rng.normal(size=(M, N, 3))
    e /= np.linalg.norm(e, axis=2, keepdims=True)
    m = rng.uniform(0.5, 3.0, size=(M, N))
    path = RS.write_restart("bcc_Fe", 1234, m, e, mode="S", outdir=str(tmp_path))
    assert os.path.basename(path) == "restart.bcc_Fe.out"
    lines = open(path).read().splitlines()
    assert lines[0] == "#" * 80 and lines[5] == "#" * 80
    assert lines[1] == "# File type: R" and lines[2] == "# Simulation type: S"
    assert lines[3] == "# Number of atoms:  " + OUT.fmt_i(N)
    assert lines[4] == "# Number of ensembles:  " + OUT.fmt_i(M)
    assert lines[6] == "  # iter     ens   iatom           |Mom|             M_x             M_y             M_z"
    assert len(lines) == 7 + M * N
    assert all(len(l) == 3 * 8 + 2 + 4 * 16 for l in lines[7:])      # (3i8, 2x, 4es16.8)
    assert lines[7 + N].split()[:3] == ["1234", "2", "1"]             # ensembles outer
    rstep, m2, e2 = RS.read_restart(path, N, M)
    assert rstep == 1234
    assert np.allclose(m2, m, rtol=1e-8, atol=0)
    assert np.allclose(e2, e, rtol=0, atol=2e-8)


def test_moment_file_appends_snapshots(tmp_path):
    e = np.zeros((1, 3, 3))
    e[..., 2] = 1.0
    m = np.ones((1, 3))
    for it in (0, 10, 20):
        path = RS.write_moments("chain", it, m, e, outdir=str(tmp_path))
    lines = open(path).read().splitlines()
    assert os.path.basename(path) == "moment.chain.out"
    assert lines[1] == "# File type: M" and len(lines) == 7 + 3 * 3
    assert [l.split()[0] for l in lines[7::3]] == ["0", "10", "20"]


def test_reads_legacy_layout(tmp_path):
    path = tmp_path / "restart.legacy.out"
    rows = ["     500"]
    for i, v in enumerate([(0.0, 0.0, 2.0), (0.6, 0.8, 0.0)]):
        rows.append(OUT.fmt_i(1) + OUT.fmt_i(i + 1) + "  " + OUT.fmt_es(2.2) +
                    "".join(OUT.fmt_es(x) for x in v))
    path.write_text("\n".join(rows) + "\n")
    rstep, m, e = RS.read_restart(str(path), 2, 1, legacy=True)
    assert rstep == 500 and np.all(m == 2.2)
    assert np.array_equal(e[0, 0], [0.0, 0.0, 2.0])                   # legacy: not normalised
