"""hipGraph-captured stepping: replays give the same numbers as launching
the kernels one by one, for every replay, with eager work and host
synchronisations between replays, and after a change of parameters (which
must re-capture)."""
import numpy as np
import pytest

import uppasd_amd as U
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu


def _llg_engine(sysd, emom, precision="fp64"):
    N, M = sysd["lattice"].natom, emom.shape[0]
    eng = U.Engine(N, M, precision)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_moments(emom, np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (M, N))))
    eng.set_llg(0.1, 300.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, 2024)
    return eng


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("alg", [1, 4, 5])
def test_llg_graph_replays_equal_eager(precision, alg):
    rng = np.random.default_rng(40)
    sysd = CF.bcc_shells(5, CF.bcc5_couplings())
    emom = np.ascontiguousarray(CF.random_unit(sysd["lattice"].natom, rng, 2))
    eager = _llg_engine(sysd, emom, precision)
    graph = _llg_engine(sysd, emom, precision)
    graph.set_graph_steps(7)
    # 3 eager, 2 replays + 1 eager, 4 eager, 3 replays + 2 eager: every
    # replay checked against the eager run
    for n in (3, 15, 4, 23):
        eager.sd_steps(alg, n)
        graph.sd_steps(alg, n)
        a, _ = eager.get_moments()
        b, _ = graph.get_moments()
        assert np.array_equal(a, b), n
    # a parameter change re-captures
    eager.set_llg(0.2, 100.0, 1e-16)
    graph.set_llg(0.2, 100.0, 1e-16)
    eager.sd_steps(alg, 14)
    graph.sd_steps(alg, 14)
    assert np.array_equal(eager.get_moments()[0], graph.get_moments()[0])


@pytest.mark.parametrize("mode", ["M", "H"])
def test_mc_graph_replays_equal_eager(mode):
    rng = np.random.default_rng(41)
    sysd = CF.bcc_shells(6, CF.bcc5_couplings())
    N, M = sysd["lattice"].natom, 3
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    temps = [200.0, 700.0, 1300.0]
    engs = []
    for k in (0, 4):
        eng = U.Engine(N, M)
        eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
        eng.set_moments(emom, np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (M, N))))
        eng.set_graph_steps(k)
        engs.append(eng)
    for n in (2, 9, 8, 1):
        for eng in engs:
            eng.mc_sweeps(mode, temps, n, 31)
        assert np.array_equal(engs[0].get_moments()[0], engs[1].get_moments()[0]), n
    # accumulate between replays (eager reductions on the same stream)
    for eng in engs:
        for _ in range(3):
            eng.mc_sweeps(mode, temps, 4, 31)
            eng.accumulate(with_energy=True)
    a, b = engs[0].accumulators(), engs[1].accumulators()
    for k in ("m", "m2", "m4", "e", "e2"):
        assert np.array_equal(a[k], b[k]), k


def test_graph_steps_rejected_when_out_of_range():
    sysd = CF.bcc_shells(3, CF.bcc5_couplings())
    eng = U.Engine(sysd["lattice"].natom, 1)
    with pytest.raises(U.UasdError):
        eng.set_graph_steps(-1)
    with pytest.raises(U.UasdError):
        eng.set_graph_steps(5000)
