"""Parallel tempering host logic (uppasd_amd/tempering.py) against the
oracle's restatement of pt_driver.f90 (or_pt_exchange, or_pt_select), and
the driver's cadence on a recording stand-in for the engine (no GPU)."""
import ctypes

import numpy as np
import pytest

from uppasd_amd import tempering as PT


def _oracle_exchange(O, temps, repene, natom, u, temprescale=1.0):
    t = np.ascontiguousarray(temps, dtype=np.float64).copy()
    hit, trial = ctypes.c_double(0.0), ctypes.c_double(0.0)
    O.lib().or_pt_exchange(len(t), t, np.ascontiguousarray(repene, dtype=np.float64), natom,
                           np.ascontiguousarray(u, dtype=np.float64), temprescale,
                           ctypes.byref(hit), ctypes.byref(trial))
    return list(t), int(hit.value), int(trial.value)


@pytest.mark.parametrize("nrep", [2, 3, 4, 7, 8])
def test_exchange_equals_oracle(O, nrep):
    rng = np.random.default_rng(nrep)
    natom = 250
    accepted = rejected = 0
    for _ in range(200):
        temps = np.sort(rng.uniform(100.0, 1500.0, size=nrep))
        rng.shuffle(temps)
        # energies in mRy for the whole replica, spread so that the
        # acceptance probabilities cover (0, 1)
        repene = -natom * (5.0 + rng.normal(scale=0.05, size=nrep))
        u = rng.random(nrep)
        got = PT.exchange(temps, repene, natom, u)
        want = _oracle_exchange(O, temps, repene, natom, u)
        assert got[0] == want[0] and got[1:] == want[1:]
        assert sorted(got[0]) == sorted(temps)
        accepted += got[1]
        rejected += got[2] - got[1]
    assert accepted > 0 and rejected > 0


def test_total_energy_rule_scales_the_exponent():
    """per_atom=False uses the replica's total energy: the exponent is natom
    times the reference's, so a pair the reference swaps on a draw of 0.9
    is kept apart."""
    temps, repene, natom = [300.0, 310.0], [-1000.0, -990.0], 100
    assert PT.exchange(temps, repene, natom, [0.9, 0.9])[1] == 1
    assert PT.exchange(temps, repene, natom, [0.9, 0.9], per_atom=False)[1] == 0


def test_downhill_exchanges_always_accepted(O):
    """Moving the lower energy to the lower temperature has
    (E2 - E1)(beta2 - beta1) >= 0 and is accepted for any draw u <= 1."""
    temps = [200.0, 400.0]
    repene = [-1000.0, -1200.0]          # replica 2 is lower in energy, hotter
    t, h, n = PT.exchange(temps, repene, 200, [1.0, 1.0])
    assert t == [400.0, 200.0] and (h, n) == (1, 1)
    assert _oracle_exchange(O, temps, repene, 200, [1.0, 1.0]) == (t, h, n)


def test_select_replica_equals_oracle(O):
    rng = np.random.default_rng(3)
    for _ in range(100):
        e = rng.integers(-5, 5, size=6).astype(np.float64)
        want = O.lib().or_pt_select(6, np.ascontiguousarray(e))
        assert PT.select_replica(e) == want == int(np.argmin(e))


class _Recorder:
    """Stands in for the engine: records the sweeps and returns energies
    that depend on the temperature each ensemble was last swept at."""

    def __init__(self, N, M):
        self.N, self.M = N, M
        self.calls = []
        self.t = np.zeros(M)

    def mc_sweeps(self, mode, temps, n, seed):
        self.calls.append((mode, np.array(temps), n, seed))
        self.t = np.array(temps)

    def measure(self, with_energy=False):
        return np.zeros((self.M, 3)), -self.N * 10.0 / (1.0 + self.t / 100.0)


def test_driver_cadence():
    eng = _Recorder(64, 6)
    draws = []

    def uniform(n):
        draws.append(n)
        return np.zeros(n)               # accept every proposal

    pt = PT.ParallelTempering(eng, [100.0, 200.0, 300.0], mensemble=2, pt_step=7, seed=9,
                              uniform=uniform)
    best = pt.run(100)
    # sweeps in runs that end on multiples of pt_step, and the rest
    assert [c[2] for c in eng.calls] == [7] * 14 + [2]
    assert all(c[0] == "H" and c[3] == 9 for c in eng.calls)
    # replica-major ensembles: each replica's temperature on its mensemble ensembles
    assert np.array_equal(eng.calls[0][1], [100.0, 100.0, 200.0, 200.0, 300.0, 300.0])
    # attempts at 7, 14, ..., 84 (84/100 < 0.9) but not at 91 or 98
    assert [h[0] for h in pt.history] == [7 * k for k in range(1, 15)]
    assert len(draws) == 12 and all(n == 3 for n in draws)
    assert pt.n_trial == 12 * 2
    assert sorted(pt.temps) == [100.0, 200.0, 300.0]
    assert best == int(np.argmin(pt.repene))


def test_pt_restart_layout(tmp_path):
    rng = np.random.default_rng(1)
    e = rng.normal(size=(2, 5, 3))
    e /= np.linalg.norm(e, axis=2, keepdims=True)
    p = PT.write_pt_restart("ptcase", 40, [300.0, 612.5], np.full((2, 5), 2.2), e,
                            outdir=str(tmp_path))
    lines = open(p).read().splitlines()
    assert lines[0] == "#" * 80 and lines[5] == "#" * 80
    assert lines[3] == "# Number of atoms:         5"
    assert lines[4] == "# Number of ensembles:         2"
    assert lines[6].split() == ["#iterTemp", "iatom", "|Mom|", "M_x", "M_y", "M_z"]
    rows = lines[7:]
    assert len(rows) == 10 and all(len(r) == 8 + 8 + 8 + 2 + 4 * 16 for r in rows)
    t = rows[5].split()
    assert t[:3] == ["40", "612.500", "1"]
    assert np.allclose([float(x) for x in t[4:]], e[1, 0], atol=1e-8)
