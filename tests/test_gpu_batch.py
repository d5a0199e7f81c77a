"""Batches: replicas of uneven size joined into one lattice.

  * LLG on the joined lattice equals each replica run alone, bit for bit, at
    T = 0 and on an injected noise stream (the same arithmetic per site; only
    the site numbering is shifted)
  * per-replica FP64 sums (measure_segments) equal each replica's own
    measure to 1e-13 relative (only the reduction tree differs)
  * the device cumulant accumulators equal the running weighted means of
    calc_and_print_cumulant (prn_averages.f90:956-1000) restated on the host
    from measure_segments samples, to 1e-12 relative
"""
import numpy as np
import pytest

import uppasd_amd as U
from uppasd_amd import batch as Bt
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu


def _replicas(Ls):
    return [CF.bcc_shells(L, CF.bcc5_couplings(), moment=2.2) for L in Ls]


@pytest.mark.parametrize("alg", [1, 5])
@pytest.mark.parametrize("noise", ["none", "injected"])
def test_joined_llg_equals_separate_runs(alg, noise):
    rng = np.random.default_rng(30)
    reps = _replicas([3, 5, 4])
    M = 2
    states = [np.ascontiguousarray(CF.random_unit(r["lattice"].natom, rng, M)) for r in reps]
    j = Bt.join(reps)
    eng = Bt.engine_for(j, M)
    eng.set_moments(np.ascontiguousarray(np.concatenate(states, axis=1)),
                    np.ascontiguousarray(np.broadcast_to(j.mmom, (M, j.natom))))
    eng.set_llg(0.1, 300.0 if noise == "injected" else 0.0, 1e-16)
    steps = 30
    normals = rng.normal(size=(steps, M, j.natom, 3))
    for s in range(steps):
        if noise == "injected":
            eng.set_noise(U.UASD_NOISE_INJECTED, 0, normals[s])
        else:
            eng.set_noise(U.UASD_NOISE_NONE)
        eng.sd_steps(alg, 1)
    e_joined, _ = eng.get_moments()
    for r, (rep, st) in enumerate(zip(reps, states)):
        a, b = j.offsets[r], j.offsets[r + 1]
        one = U.Engine(rep["lattice"].natom, M)
        one.set_exchange(rep["nlist"], rep["nlistsize"], rep["ncoup"])
        one.set_moments(st, np.ascontiguousarray(np.broadcast_to(rep["mmom"], (M, b - a))))
        one.set_llg(0.1, 300.0 if noise == "injected" else 0.0, 1e-16)
        for s in range(steps):
            if noise == "injected":
                one.set_noise(U.UASD_NOISE_INJECTED, 0, np.ascontiguousarray(normals[s][:, a:b]))
            else:
                one.set_noise(U.UASD_NOISE_NONE)
            one.sd_steps(alg, 1)
        e_one, _ = one.get_moments()
        assert np.array_equal(e_joined[:, a:b], e_one), r


def test_segment_sums_match_separate_measure():
    rng = np.random.default_rng(31)
    reps = _replicas([4, 3, 6, 5])
    M = 3
    j = Bt.join(reps)
    eng = Bt.engine_for(j, M)
    e = np.ascontiguousarray(CF.random_unit(j.natom, rng, M))
    eng.set_moments(e, np.ascontiguousarray(np.broadcast_to(j.mmom, (M, j.natom))))
    ms, en = eng.measure_segments(with_energy=True)
    assert ms.shape == (4, M, 3) and en.shape == (4, M)
    for r, rep in enumerate(reps):
        a, b = j.offsets[r], j.offsets[r + 1]
        one = U.Engine(b - a, M)
        one.set_exchange(rep["nlist"], rep["nlistsize"], rep["ncoup"])
        one.set_moments(np.ascontiguousarray(e[:, a:b]),
                        np.ascontiguousarray(np.broadcast_to(rep["mmom"], (M, b - a))))
        ms1, en1 = one.measure(with_energy=True)
        assert np.allclose(ms[r], ms1, rtol=1e-13, atol=1e-13 * np.abs(ms1).max())
        assert np.allclose(en[r], en1, rtol=1e-13)


# wide (M | 256) and strided reduction forms; 128 and 256 ensembles have more
# (ensemble, value) partials per workgroup than the workgroup has threads
@pytest.mark.parametrize("M", [4, 3, 128, 256])
def test_accumulators_match_host_recursion(M):
    rng = np.random.default_rng(32)
    reps = _replicas([3, 4])
    j = Bt.join(reps)
    eng = Bt.engine_for(j, M)
    e = np.ascontiguousarray(CF.random_unit(j.natom, rng, M))
    eng.set_moments(e, np.ascontiguousarray(np.broadcast_to(j.mmom, (M, j.natom))))
    temps = np.linspace(300.0, 1500.0, M)
    sizes = np.diff(j.offsets).astype(np.float64)
    tot = np.zeros((2, M))
    w = np.zeros((2, M))
    acc = np.zeros((5, 2, M))
    for _ in range(12):
        eng.mc_sweeps("M", temps, 3, 99)
        ms, en = eng.measure_segments(with_energy=True)
        eng.accumulate(with_energy=True)
        me = np.linalg.norm(ms, axis=2) / sizes[:, None]
        ee = en / sizes[:, None]
        x = np.stack([me, me ** 2, (me ** 2) ** 2, ee, ee ** 2])
        w = w + 1.0
        acc = (acc * tot + x * w) / (tot + w)
        tot = tot + w
    got = eng.accumulators()
    assert got["samples"] == 12
    assert np.array_equal(got["n"], w)
    assert np.allclose(got["W"], tot)
    for i, k in enumerate(("m", "m2", "m4", "e", "e2")):
        assert np.allclose(got[k], acc[i], rtol=1e-12, atol=0), k
    eng.reset_accumulators()
    assert eng.accumulators()["samples"] == 0
    assert np.all(eng.accumulators()["W"] == 0)


@pytest.mark.parametrize("M", [64, 128, 256, 5])
def test_segment_sums_many_ensembles_match_host(M):
    """measure_segments against numpy sums of the uploaded state, for ensemble
    counts whose (ensemble, value) partials outnumber a workgroup's threads."""
    rng = np.random.default_rng(33)
    reps = _replicas([3, 2, 4])
    j = Bt.join(reps)
    eng = Bt.engine_for(j, M)
    e = np.ascontiguousarray(CF.random_unit(j.natom, rng, M))
    mm = np.ascontiguousarray(np.broadcast_to(j.mmom, (M, j.natom)))
    eng.set_moments(e, mm)
    ms, _ = eng.measure_segments()
    eM = e * mm[..., None]
    for r in range(len(reps)):
        a, b = j.offsets[r], j.offsets[r + 1]
        want = eM[:, a:b].sum(axis=1)
        assert np.allclose(ms[r], want, rtol=0, atol=1e-12 * (b - a)), (M, r)


def test_segments_rejected_when_malformed():
    reps = _replicas([3])
    j = Bt.join(reps)
    eng = Bt.engine_for(j, 1)
    with pytest.raises(U.UasdError):
        eng.set_segments([0, 10, 5, j.natom])
    with pytest.raises(U.UasdError):
        eng.set_segments([1, j.natom])
