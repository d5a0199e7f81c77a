"""Launch shapes and the autotuner.

The gather batch only changes when loads are issued, not the order in which
terms are added, and the workgroup size only changes which lanes share a
workgroup, so every shape must give bit-identical results.  The autotuner
must leave the state and the step counters as it found them.
"""
import numpy as np
import pytest

import uppasd_amd as U
from uppasd_amd import configs as CF

pytestmark = pytest.mark.gpu

NT = 0x10000   # non-temporal neighbour-list loads
I32 = 0x20000  # int32 entries although packed int16 entries exist
PAIR = 0x1000000  # MC: two adjacent ensembles per thread
SHAPES = [(2, 64), (4, 128 | NT), (8, 256), (8, 256 | NT | I32), (16, 64 | I32), (16, 256 | NT)]
MC_SHAPES = SHAPES + [(2, 64 | PAIR), (4, 256 | PAIR | I32), (8, 128 | PAIR | NT)]


def _llg_run(sysd, emom, gather, block, precision, alg, steps=20):
    N = sysd["lattice"].natom
    M = emom.shape[0]
    eng = U.Engine(N, M, precision)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    if "dmlist" in sysd:
        eng.set_dm(sysd["dmlist"], sysd["dmlistsize"], sysd["dm_vect"])
    if "taniso" in sysd:
        eng.set_anisotropy(sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"])
    if "ext" in sysd:
        eng.set_external_field(sysd["ext"])
    eng.set_moments(emom, np.broadcast_to(sysd["mmom"], (M, N)))
    eng.set_llg(0.1, 300.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, 1234)
    eng.set_tuning(eng.TUNE_LLG, gather, block)
    eng.set_tuning(eng.TUNE_FIELD, gather, block)
    eng.sd_steps(alg, steps)
    e, _ = eng.get_moments()
    beff, en = eng.field(with_energy=True)
    return e, beff, en


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("alg", [1, 5])
def test_llg_results_independent_of_launch_shape(precision, alg):
    rng = np.random.default_rng(5)
    sysd = CF.triangular_dmi(24)  # exchange + DM + anisotropy + field
    N = sysd["lattice"].natom
    emom = np.ascontiguousarray(CF.random_unit(N, rng, 2))
    ref = _llg_run(sysd, emom, 4, 256, precision, alg)
    for g, b in SHAPES:
        got = _llg_run(sysd, emom, g, b, precision, alg)
        for x, y in zip(got, ref):
            assert np.array_equal(x, y), (g, b)


def test_mc_results_independent_of_launch_shape():
    rng = np.random.default_rng(6)
    sysd = CF.bcc_shells(6, CF.bcc5_couplings())
    N = sysd["lattice"].natom
    emom = np.ascontiguousarray(CF.random_unit(N, rng, 2))
    out = []
    for g, b in [(4, 256)] + MC_SHAPES:
        eng = U.Engine(N, 2, "fp64")
        eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
        eng.set_moments(emom, np.broadcast_to(sysd["mmom"], (2, N)))
        eng.set_tuning(eng.TUNE_MC, g, b)
        eng.mc_sweeps("M", [300.0, 900.0], 5, 77)
        eng.mc_sweeps("H", [300.0, 900.0], 5, 78)
        out.append(eng.get_moments()[0])
    for x in out[1:]:
        assert np.array_equal(x, out[0])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("M", [4, 3])
def test_mc_pair_of_ensembles_per_thread_is_bit_identical(precision, M):
    """Two adjacent ensembles per thread (shared neighbour entries and
    couplings) against one per thread, on a Hamiltonian with every MC term
    (exchange, DM with |m| != 1, anisotropy types 1, 2 and 7, Zeeman):
    Metropolis and heat bath give the same moments bit for bit.  With M odd
    the flag leaves one ensemble per thread."""
    rng = np.random.default_rng(64)
    sysd = CF.triangular_dmi(12, J=1.0, D=0.2, K=0.0, B=0.0, moment=2.2)
    N = sysd["lattice"].natom
    taniso = rng.choice([1, 2, 7], size=N).astype(np.int32)
    ea = rng.normal(size=(N, 3))
    ea /= np.linalg.norm(ea, axis=1, keepdims=True)
    kan = np.ascontiguousarray(rng.uniform(-0.02, 0.02, size=(N, 2)))
    sb = rng.uniform(0.2, 1.5, N)
    temps = np.linspace(100.0, 900.0, M)
    emom = np.ascontiguousarray(CF.random_unit(N, rng, M))
    out = []
    for g, b in [(4, 256), (4, 256 | PAIR), (2, 64 | PAIR | I32), (8, 128 | PAIR | NT)]:
        eng = U.Engine(N, M, precision)
        eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
        eng.set_dm(sysd["dmlist"], sysd["dmlistsize"], sysd["dm_vect"])
        eng.set_anisotropy(taniso, np.ascontiguousarray(ea), kan, sb)
        eng.set_external_field(np.array([0.4, -0.3, 1.1]))
        eng.set_moments(emom, np.full((M, N), 2.2))
        eng.set_tuning(eng.TUNE_MC, g, b)
        assert eng.tuning(eng.TUNE_MC) == (g, b)
        eng.mc_sweeps("M", temps, 4, 321)
        eng.mc_sweeps("H", temps, 3, 322)
        out.append(eng.get_moments()[0])
    assert np.abs(out[0] - emom).max() > 0.1      # the sweeps moved the spins
    for x in out[1:]:
        assert np.array_equal(x, out[0])


def test_bad_shape_rejected():
    sysd = CF.bcc_shells(3, CF.bcc5_couplings())
    eng = U.Engine(sysd["lattice"].natom, 1, "fp64")
    with pytest.raises(U.UasdError):
        eng.set_tuning(eng.TUNE_LLG, 3, 256)
    with pytest.raises(U.UasdError):
        eng.set_tuning(eng.TUNE_LLG, 4, 512)
    with pytest.raises(U.UasdError):
        eng.set_tuning(7, 4, 256)
    with pytest.raises(U.UasdError):
        eng.set_tuning(eng.TUNE_LLG, 4, 256 | PAIR)   # MC only
    with pytest.raises(U.UasdError):
        eng.set_tuning(eng.TUNE_MC, 4, 256 | (PAIR << 1))   # no such flag


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_autotune_llg_leaves_trajectory_unchanged(precision):
    rng = np.random.default_rng(8)
    sysd = CF.bcc_shells(8, CF.bcc5_couplings())
    N = sysd["lattice"].natom
    emom = np.ascontiguousarray(CF.random_unit(N, rng, 1))

    def run(tune):
        eng = U.Engine(N, 1, precision)
        eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
        eng.set_moments(emom, sysd["mmom"][None])
        eng.set_llg(0.1, 300.0, 1e-16)
        eng.set_noise(U.UASD_NOISE_PHILOX, 99)
        eng.sd_steps(1, 3)
        if tune:
            ms = eng.autotune_llg(1, steps=2)
            assert ms > 0.0
            g, b = eng.tuning(eng.TUNE_LLG)
            assert g in (2, 4, 8, 16) and b & 0xFFFF in (64, 128, 256) and b >> 16 in (0, 1, 2, 3)
            assert eng.tuning(eng.TUNE_FIELD) == (g, b)
        eng.sd_steps(1, 7)
        return eng.get_moments()[0]

    assert np.array_equal(run(True), run(False))


def test_autotune_mc_leaves_chain_unchanged():
    rng = np.random.default_rng(10)
    sysd = CF.bcc_shells(6, CF.bcc5_couplings())
    N = sysd["lattice"].natom
    emom = np.ascontiguousarray(CF.random_unit(N, rng, 2))

    def run(tune):
        eng = U.Engine(N, 2, "fp64")
        eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
        eng.set_moments(emom, np.broadcast_to(sysd["mmom"], (2, N)))
        eng.mc_sweeps("M", [500.0, 700.0], 2, 5)
        if tune:
            assert eng.autotune_mc("M", [500.0, 700.0], sweeps=1) > 0.0
        eng.mc_sweeps("M", [500.0, 700.0], 4, 5)
        return eng.get_moments()[0]

    assert np.array_equal(run(True), run(False))


def test_sampled_kernel_timing_counts_and_leaves_results():
    """HIP-event timing of one launch in n of each kind: the launch counts are
    the sampled ones, the durations positive, and the trajectory is the one
    of an untimed run."""
    rng = np.random.default_rng(61)
    sysd = CF.bcc_shells(6, CF.bcc5_couplings())
    emom = np.ascontiguousarray(CF.random_unit(sysd["lattice"].natom, rng, 1))
    ref, _, _ = _llg_run(sysd, emom, 4, 256, "fp64", 1, steps=24)
    N = sysd["lattice"].natom
    eng = U.Engine(N, 1)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_moments(emom, np.broadcast_to(sysd["mmom"], (1, N)))
    eng.set_llg(0.1, 300.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, 1234)
    eng.set_tuning(eng.TUNE_LLG, 4, 256)
    eng.set_kernel_timing(True, every=4)
    eng.sd_steps(1, 24)
    (n0, ms0), (n1, ms1) = eng.kernel_timing(0), eng.kernel_timing(1)
    eng.set_kernel_timing(False)
    assert (n0, n1) == (6, 6) and ms0 > 0.0 and ms1 > 0.0
    assert np.array_equal(eng.get_moments()[0], ref)
    with pytest.raises(U.UasdError):
        eng.set_kernel_timing(True, every=-2)


def test_call_timing_times_whole_calls_only():
    """Events around whole sd_steps calls (kind 6) and no launch events: one
    timed call per call, a positive duration, no per-launch samples, and the
    trajectory of an untimed run."""
    rng = np.random.default_rng(62)
    sysd = CF.bcc_shells(6, CF.bcc5_couplings())
    emom = np.ascontiguousarray(CF.random_unit(sysd["lattice"].natom, rng, 1))
    ref, _, _ = _llg_run(sysd, emom, 4, 256, "fp64", 1, steps=24)
    N = sysd["lattice"].natom
    eng = U.Engine(N, 1)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_moments(emom, np.broadcast_to(sysd["mmom"], (1, N)))
    eng.set_llg(0.1, 300.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, 1234)
    eng.set_tuning(eng.TUNE_LLG, 4, 256)
    eng.set_kernel_timing(True, calls_only=True)
    eng.sd_steps(1, 10)
    eng.sd_steps(1, 14)
    n6, ms6 = eng.kernel_timing(6)
    assert n6 == 2 and ms6 > 0.0
    assert eng.kernel_timing(0)[0] == 0 and eng.kernel_timing(1)[0] == 0
    eng.set_kernel_timing(False)
    assert np.array_equal(eng.get_moments()[0], ref)


def test_tune_cache_reuses_winners_and_persists(tmp_path):
    """With the cache on, a second context of the same workload takes the
    first one's winner without timing (the kept time comes back), the file
    holds one line per workload, and a fresh load of that file serves the
    next context; switching the cache off restores timing."""
    rng = np.random.default_rng(63)
    sysd = CF.bcc_shells(6, CF.bcc5_couplings())
    N = sysd["lattice"].natom
    emom = np.ascontiguousarray(CF.random_unit(N, rng, 1))
    path = tmp_path / "tune.txt"

    def ctx():
        eng = U.Engine(N, 1)
        eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
        eng.set_moments(emom, np.broadcast_to(sysd["mmom"], (1, N)))
        eng.set_llg(0.1, 300.0, 1e-16)
        eng.set_noise(U.UASD_NOISE_PHILOX, 1234)
        return eng

    try:
        U.set_tune_cache(True, str(path))
        a = ctx()
        ms_a = a.autotune_llg(1, steps=2)
        b = ctx()
        ms_b = b.autotune_llg(1, steps=2)
        assert ms_b == ms_a and b.tuning(b.TUNE_LLG) == a.tuning(a.TUNE_LLG)
        lines = path.read_text().splitlines()
        assert len(lines) == 1 and lines[0].split()[0].startswith("v2/gfx950/llg/a1/fp64/")
        U.set_tune_cache(False)
        U.set_tune_cache(True, str(path))      # reload from the file alone
        c = ctx()
        assert abs(c.autotune_llg(1, steps=2) - ms_a) <= 1e-5     # kept to 6 decimals
        assert c.tuning(c.TUNE_LLG) == a.tuning(a.TUNE_LLG)
        e_c = (c.sd_steps(1, 5), c.get_moments()[0])[1]
        U.set_tune_cache(False)
        d = ctx()
        d.autotune_llg(1, steps=2)            # timed again: a fresh measurement
        assert len(path.read_text().splitlines()) == 1
        d.sd_steps(1, 5)
        assert np.array_equal(d.get_moments()[0], e_c)
    finally:
        U.set_tune_cache(False)
