"""Heat-bath parallel tempering on the device (uppasd_amd/tempering.py):
replicas as ensembles of one context, swept in one launch per colour.

- With the reference's exchange rule (energies per atom, pt_driver.f90:
  216-217), the temperature sequence of a device run is replayed through the
  oracle's restatement of the exchange step (or_pt_exchange) with the
  recorded replica energies and draws: every attempt agrees.
- With the total-energy rule, replica exchange must leave the canonical
  distribution at every temperature unchanged (detailed balance), so the
  mean energy sampled at each temperature equals that of plain heat-bath
  runs at the fixed temperatures within their statistical errors; a wrong
  sign, a wrong energy unit or a mix-up of replicas and ensembles biases
  these means by many standard errors."""
import numpy as np
import pytest

import uppasd_amd as U
from uppasd_amd import configs as CF
from uppasd_amd import tempering as PT

pytestmark = pytest.mark.gpu


def _engine(sysd, M, e0):
    N = sysd["lattice"].natom
    eng = U.Engine(N, M)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_moments(e0, np.broadcast_to(sysd["mmom"], (M, N)))
    return eng


def _batch_mean(x, nb=20):
    x = np.asarray(x)
    b = x[: len(x) // nb * nb].reshape(nb, -1).mean(axis=1)
    return b.mean(), b.std(ddof=1) / np.sqrt(nb)


def test_reference_rule_replays_through_oracle(O):
    import ctypes
    sysd = CF.bcc_shells(4, CF.bcc5_couplings())
    N = sysd["lattice"].natom
    temps = np.linspace(700.0, 1300.0, 8)
    nrep, mens = len(temps), 2
    e0 = np.ascontiguousarray(CF.random_unit(N, np.random.default_rng(4), nrep * mens))
    print(f"parallel tempering, reference rule N={N} nrep={nrep} mensemble={mens}", flush=True)
    pt = PT.ParallelTempering(_engine(sysd, nrep * mens, e0), temps, mensemble=mens,
                              pt_step=10, seed=3)
    best = pt.run(1000)
    attempts = [h for h in pt.history if h[3] is not None]
    assert len(attempts) == 89 and len(pt.history) == 100
    for a, b in zip(pt.history[:-1], pt.history[1:]):
        t = np.ascontiguousarray(a[1], dtype=np.float64)
        if a[3] is not None:
            hit, trial = ctypes.c_double(0.0), ctypes.c_double(0.0)
            O.lib().or_pt_exchange(nrep, t, np.ascontiguousarray(a[2]), N,
                                   np.ascontiguousarray(a[3]), 1.0, ctypes.byref(hit),
                                   ctypes.byref(trial))
        assert list(t) == b[1], a[0]
    # per-atom energies make nearly every proposal succeed
    assert pt.hit_rate() > 0.5
    assert best == O.lib().or_pt_select(nrep, np.ascontiguousarray(pt.repene))


def test_total_energy_rule_keeps_each_temperature_canonical():
    sysd = CF.bcc_shells(4, CF.bcc5_couplings())
    N = sysd["lattice"].natom
    temps = np.linspace(750.0, 1030.0, 8)
    nrep, mens, step, nsweep, burn = len(temps), 2, 5, 6000, 1000
    M = nrep * mens
    e0 = np.ascontiguousarray(CF.random_unit(N, np.random.default_rng(5), M))
    print(f"parallel tempering, total-energy rule N={N} nrep={nrep} mensemble={mens} "
          f"sweeps={nsweep}", flush=True)
    pt = PT.ParallelTempering(_engine(sysd, M, e0), temps, mensemble=mens, pt_step=step,
                              seed=11, per_atom=False)
    best = pt.run(nsweep)
    assert sorted(pt.temps) == sorted(temps.tolist())
    print(f"hit rate {pt.hit_rate():.3f}", flush=True)
    assert 0.05 < pt.hit_rate() < 0.95
    assert best == int(np.argmin(pt.repene))
    # energies per site at each temperature, from the attempts after burn-in
    pt_samples = {t: [] for t in temps.tolist()}
    for s, tl, ene, _ in pt.history:
        if s > burn:
            for r, t in enumerate(tl):
                pt_samples[t].append(ene[r] / (N * mens))
    # plain heat-bath at fixed temperatures, same start and cadence
    eng = _engine(sysd, M, e0)
    plain = {t: [] for t in temps.tolist()}
    for s in range(step, nsweep + 1, step):
        eng.mc_sweeps("H", np.repeat(temps, mens), step, 12)
        if s > burn:
            _, en = eng.measure(with_energy=True)
            e = en.reshape(nrep, mens).sum(axis=1) / (N * mens)
            for r, t in enumerate(temps.tolist()):
                plain[t].append(e[r])
    means = []
    for t in temps.tolist():
        a, sa = _batch_mean(pt_samples[t])
        b, sb = _batch_mean(plain[t])
        means.append(b)
        print(f"T={t:7.1f}  PT {a:.5f} +- {sa:.5f}   plain {b:.5f} +- {sb:.5f}", flush=True)
        assert abs(a - b) <= 5.0 * np.hypot(sa, sb) + 1e-4, t
    assert np.all(np.diff(means) > 0)    # energy rises with temperature
