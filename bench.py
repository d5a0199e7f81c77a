#!/usr/bin/env python3
"""Benchmark of the spin-update hot path on one MI355X.

Workload (default): the plain FP64 drop-in configuration -- bcc ferromagnet,
64x64x64 conventional cells (524,288 spins), periodic, five exchange shells
(58 neighbours) with J_k = 1.34 exp(-1.2(k-1)) mRy alternating in sign from
shell 3, moment 2.2 muB, semi-implicit midpoint LLG (SDEalgh 1), alpha 0.1,
dt 1e-16 s, T = 300 K with the in-kernel Philox thermal field.  A step is one
full LLG step of every spin (predictor + corrector, both fused with the field
gather); the metric is spin-updates/s = spins x steps / wall time, inputs
resident in HBM.

Prints one JSON line (see the keys below).  The roofline object is for the
dominant kernel (the corrector), timed with HIP events on the stream it runs
on; the cpu_baseline object times the reference's own routines, compiled from
its sources, on every host core over a bounded sample of the same workload
(the FP64 C oracle stands in where that library is not built).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "spin-updates/s (LLG step, MC sweep); FP32+batched vs FP64 at stated error"
HBM_PEAK_GBS = 8000.0


SOLVER_NAMES = {1: "midpoint", 4: "Heun", 5: "Depondt", 6: "SIB", 7: "spherical midpoint"}


def kernel_bytes(W: int, s: int, ib: int = 4, cls: int = 4, noise: str = "philox",
                 b2: bool = False) -> dict:
    """Algorithmic HBM bytes per site of the fused LLG kernels as built.

    Spins are stored as quads {x, y, z, mmom} of s-byte reals; neighbour
    entries take ib bytes (4 for int32 indices, 2 for int16 differences, 0 for
    the LDS-tiled stencil kernels, which read no lists); cls bytes of site
    class per site (0 for the tiled kernels on a one-species lattice, whose
    classes come with the launch).
    predictor: gather e_j quads (4s), index stream ib*W, class, write p (4s)
    corrector: gather p_j quads (4s), own e (4s), index ib*W, class, write e (4s)
    noise: "philox" -- the predictor stores its deviates as a quad of reals
    for the corrector (4s written, 4s read); "injected" -- both halves read the
    caller's N(0,1) quads (4s each); "none" -- nothing.
    Each site's own spin is one of its neighbours' gathers, so the gathered
    quads count once per site (perfect reuse through L2 / MALL).  b2: the
    solver keeps a field quad between the halves (Depondt's bdup, Heun's
    b2eff), written by the predictor and read by the corrector (4s each).
    """
    q = 4 * s if b2 else 0
    nz = 0 if noise == "none" else 4 * s
    pred = 4 * s + ib * W + cls + 4 * s + nz + q
    corr = 4 * s + 4 * s + ib * W + cls + 4 * s + nz + q
    return {"predictor": pred, "corrector": corr, "step": pred + corr}


def model_bytes(W: int, s: int, injected: bool = False) -> dict:
    """Algorithmic bytes per spin-update of the hot path's two-pass LLG step,
    as the byte model of the workload counts them: pass 1 gathers the moments
    and writes the predicted moments and field (9 s + 4 W), pass 2 gathers the
    predicted moments, reads the moments and the predicted field and writes
    the new moments (12 s + 4 W); 4-byte neighbour indices, couplings held per
    basis site, noise from counters (an injected N(0,1) stream adds 3 s to each
    pass).  21 s + 8 W per step: 632 B in FP64 at W = 58."""
    extra = 3 * s if injected else 0
    p1 = 9 * s + 4 * W + extra
    p2 = 12 * s + 4 * W + extra
    return {"predictor": p1, "corrector": p2, "step": p1 + p2}


def timed_steps(eng, solver: int, steps: int, timing_every: int):
    """Two runs of `steps` steps on the context stream.  The first carries HIP
    events around the whole call only (no events between kernels): its wall
    time gives the value and its event time the step time.  The second
    records events around both halves of one step in `timing_every`, which
    gives each half's share of the step; the halves' durations are those
    shares of the first run's step time (events between kernels stretch the
    sampled durations, the shares are what they are used for)."""
    eng.set_kernel_timing(True, calls_only=True)
    eng.sync()
    t0 = time.perf_counter()
    eng.sd_steps(solver, steps)
    eng.sync()
    wall = time.perf_counter() - t0
    _, call_ms = eng.kernel_timing(6)
    eng.set_kernel_timing(True, timing_every)
    eng.sd_steps(solver, steps)
    eng.sync()
    n_pred, ms_pred = eng.kernel_timing(0)
    n_corr, ms_corr = eng.kernel_timing(1)
    eng.set_kernel_timing(False)
    sp, sc = ms_pred / max(n_pred, 1), ms_corr / max(n_corr, 1)
    step_s = call_ms * 1e-3 / steps
    return {"wall_s": wall, "step_s": step_s,
            "pred_s": step_s * sp / (sp + sc), "corr_s": step_s * sc / (sp + sc),
            "sampled_ms": [round(sp, 4), round(sc, 4)], "sampled_launches": [n_pred, n_corr]}


def build_c2(L: int):
    from uppasd_amd import configs as CF
    return CF.bcc_shells(L, CF.bcc5_couplings(), moment=2.2)


def _ref_rate(O, sysd, alg, temp, threads, seconds_target, lam=0.1, dt=1e-16):
    """Spin-updates/s of the reference's own routines (oracle/_ref, the
    sd_mphase loop body: field, evolve_first, field, evolve_second,
    moment_update, thermal noise from its ziggurat stream) on `threads` host
    threads, over about `seconds_target` seconds of steps."""
    L = O.ref_lib()
    L.ref_set_threads(threads)
    used = L.ref_threads()
    N = sysd["lattice"].natom
    ref = O.RefSD(N, 1, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    ref.seed(1)
    rng = np.random.default_rng(1)
    e = np.ascontiguousarray(CF_random_unit(N, rng))
    mm = np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (1, N)), dtype=np.float64)
    eM = np.ascontiguousarray(e * mm[..., None])
    t0 = time.perf_counter()
    ref.run(alg, 1, dt, lam, temp, mm, e, eM)
    one = time.perf_counter() - t0
    nsteps = max(1, int(seconds_target / max(one, 1e-6)))
    t0 = time.perf_counter()
    ref.run(alg, nsteps, dt, lam, temp, mm, e, eM)
    el = time.perf_counter() - t0
    return N * nsteps / el, used, nsteps, el


def CF_random_unit(N, rng):
    e = rng.normal(size=(1, N, 3))
    return e / np.linalg.norm(e, axis=2, keepdims=True)


def cpu_baseline(sysd, seconds_target: float = 12.0):
    """The reference's own CPU path on the host cores: its routines compiled
    from its sources (oracle/_ref/libuppasd_ref.so, OpenMP) on the headline
    workload (midpoint, 300 K, its ziggurat noise) with every core the
    process may use, plus the C1 configuration (bcc 20^3 x 2, two shells,
    Depondt, 300 K) on one thread and on all.  Where that library is not
    built, the C oracle (OpenMP, FP64, the FP64 engine's Philox deviates)
    stands in, kind "port"."""
    import ctypes
    import oracle as O
    n_cores = len(os.sched_getaffinity(0))
    N = sysd["lattice"].natom
    if O.ref_lib() is not None:
        v, used, nst, el = _ref_rate(O, sysd, 1, 300.0, n_cores, seconds_target)
        from uppasd_amd import configs as CF
        c1 = CF.bcc_shells(20, [1.34, 0.77], moment=2.2)     # J1 to 8 NN, J2 to 6 NNN
        c1_1, _, n1, e1 = _ref_rate(O, c1, 5, 300.0, 1, 3.0)
        c1_n, usedn, nn_, en_ = _ref_rate(O, c1, 5, 300.0, n_cores, 3.0)
        return {"value": v, "unit": "spin-updates/s", "cores": used, "kind": "reference",
                "host_cores_available": n_cores,
                "sample": f"the reference's field/midpoint/rannum/moment_update routines "
                          f"(compiled from its sources, -O2, OpenMP) on the same workload "
                          f"({N} spins, 58 neighbours, 300 K ziggurat noise): {nst} steps in "
                          f"{el:.1f} s on {used} threads",
                "c1_depondt_bcc20": {"spins": c1["lattice"].natom,
                                     "one_thread": c1_1, "all_threads": c1_n,
                                     "threads": usedn, "steps": [n1, nn_]}}
    L = O.lib()
    L.or_set_threads(n_cores)
    used = L.or_threads()
    ham = O.OracleHam(N, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    rng = np.random.default_rng(1)
    e = np.ascontiguousarray(CF_random_unit(N, rng))
    mm = np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (1, N)))
    eM = np.ascontiguousarray(e * mm[..., None])
    lam, T, lg = np.full(N, 0.1), np.full(N, 300.0), np.ones(N)
    t0 = time.perf_counter()
    L.or_sd_run(ctypes.byref(ham.h), N, 1, 1, 1, 0, 1e-16, lam, T, lg, mm, e, eM, 7, 2, None)
    one = time.perf_counter() - t0
    nsteps = max(1, int(seconds_target / max(one, 1e-6)))
    t0 = time.perf_counter()
    L.or_sd_run(ctypes.byref(ham.h), N, 1, 1, nsteps, 1, 1e-16, lam, T, lg, mm, e, eM, 7, 2, None)
    dt = time.perf_counter() - t0
    return {"value": N * nsteps / dt, "unit": "spin-updates/s", "cores": used, "kind": "port",
            "host_cores_available": n_cores,
            "sample": f"same workload ({N} spins, 58 neighbours, midpoint FP64, FP64 Philox "
                      f"noise at 300 K), {nsteps} steps in {dt:.1f} s on {used} host threads"}


def copy_calibration(mib: int = 512, reps: int = 5):
    """A plain device-to-device copy of known size (torch), so that rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters can be scaled to bytes."""
    import torch
    n = mib * 1024 * 1024 // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read_per_dispatch": n * 4,
                                           "bytes_written_per_dispatch": n * 4,
                                           "dispatches": reps}}), flush=True)


# ---------------------------------------------------------------- forms
# Each form runs a bounded sample of one configuration and reports the
# throughput of the accelerated form beside the plain FP64 path on the same
# workload, with the error of the reduced-precision arithmetic.

def _sync_time(fn):
    t0 = time.perf_counter()
    fn()
    return time.perf_counter() - t0


def form_c3(U, CF, steps=5000):
    """Precision: triangular 512x512 monolayer, J 1 mRy (6 NN), interfacial
    DMI 0.2 mRy, Kz 0.05 mRy, Bz 0.3 mRy/muB, Depondt, alpha 0.05, T = 0 from
    a random start.  FP32 field gather + rotation with FP64 reductions against
    the FP64 path (itself within 1e-12 of the FP64 oracle); errors stated per
    run: energy per spin at checkpoints, max field error on one state, max
    |m_i| norm drift."""
    sysd = CF.triangular_dmi(512)
    N = sysd["lattice"].natom
    rng = np.random.default_rng(3)
    e0 = np.ascontiguousarray(CF.random_unit(N, rng, 1))
    # control: FP64 from a start perturbed at the 1e-7 level, which measures
    # how far any rounding-level difference carries in this relaxation
    pert = e0 + 1e-7 * rng.normal(size=e0.shape)
    pert = np.ascontiguousarray(pert / np.linalg.norm(pert, axis=2, keepdims=True))
    engs = {}
    lat = sysd["lattice"]
    for prec in ("fp64", "fp32", "fp64_perturbed"):
        eng = U.Engine(N, 1, prec[:4])
        eng.set_lattice_dims(lat.NA, *lat.n)    # the LDS-tiled kernels are tuning candidates
        eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
        eng.set_dm(sysd["dmlist"], sysd["dmlistsize"], sysd["dm_vect"])
        eng.set_anisotropy(sysd["taniso"], sysd["eaniso"], sysd["kaniso"], sysd["sb"])
        eng.set_external_field(sysd["ext"])
        eng.set_moments(pert if prec == "fp64_perturbed" else e0, sysd["mmom"][None])
        eng.set_llg(0.05, 0.0, 1e-16)
        eng.set_noise(U.UASD_NOISE_NONE)
        eng.autotune_llg(5, steps=2)
        engs[prec] = eng
    checkpoints = [s for s in (10, 100, 1000, steps) if s <= steps]
    rec = {p: {} for p in engs}
    times = {p: 0.0 for p in engs}
    done = 0
    for c in checkpoints:
        for p, eng in engs.items():
            def run(eng=eng, n=c - done):
                eng.sd_steps(5, n)
                eng.sync()
            times[p] += _sync_time(run)
            ms, en = eng.measure(with_energy=True)
            rec[p][c] = (float(en[0]) / N, float(np.linalg.norm(ms[0])) / N)
        done = c
    e64, _ = engs["fp64"].get_moments()
    e32, eM32 = engs["fp32"].get_moments()
    drift = float(np.max(np.abs(np.linalg.norm(eM32[0], axis=1) / sysd["mmom"] - 1.0)))
    # field error of the FP32 gather on the FP64 final state
    emomM = np.ascontiguousarray(e64 * sysd["mmom"][None, :, None])
    b64, _ = engs["fp64"].effective_field(emomM)
    b32, _ = engs["fp32"].effective_field(emomM)
    ferr = float(np.max(np.abs(b32 - b64)) / np.max(np.abs(b64)))
    rel = {c: abs(rec["fp32"][c][0] - rec["fp64"][c][0]) / abs(rec["fp64"][c][0]) for c in checkpoints}
    relp = {c: abs(rec["fp64_perturbed"][c][0] - rec["fp64"][c][0]) / abs(rec["fp64"][c][0])
            for c in checkpoints}
    return {
        "workload": f"triangular 512^2 ({N} spins), J+DMI+Kz+Bz, Depondt alpha 0.05, T 0, {steps} steps",
        "fp64_value": N * steps / times["fp64"],
        "fp32_value": N * steps / times["fp32"],
        "fp32_over_fp64": times["fp64"] / times["fp32"],
        "energy_per_spin_mRy": {str(c): rec["fp64"][c][0] for c in checkpoints},
        "energy_rel_err": {str(c): rel[c] for c in checkpoints},
        "control_fp64_perturbed_1e-7_energy_rel_err": {str(c): relp[c] for c in checkpoints},
        "field_max_rel_err": ferr,
        "norm_drift_fp32": drift,
        "unit": "spin-updates/s",
    }


def _c4_replicas(CF, n_rep=64):
    Ls = np.round(np.linspace(8, 48, n_rep)).astype(int)
    reps = []
    for r, L in enumerate(Ls):
        rng = np.random.default_rng(4000 + r)
        occ = np.where(rng.random(2 * L ** 3) < 0.5, 1, 2)
        reps.append(CF.bcc_alloy(int(L), occ))
    return reps


def _batch_means(x, nblk=8):
    """Mean and standard error of a time series by batch means over nblk
    blocks (x: (T, ...), the time axis first)."""
    n = x.shape[0] // nblk * nblk
    b = x[:n].reshape(nblk, n // nblk, *x.shape[1:]).mean(axis=1)
    return b.mean(axis=0), b.std(axis=0, ddof=1) / np.sqrt(nblk)


def _binder_blocks(m, nblk=8):
    """U4 = 1 - <m^4>/(3<m^2>^2) of a series and its jackknife error over
    nblk blocks."""
    n = m.shape[0] // nblk * nblk
    b2 = (m[:n] ** 2).reshape(nblk, n // nblk, *m.shape[1:]).mean(axis=1)
    b4 = (m[:n] ** 4).reshape(nblk, n // nblk, *m.shape[1:]).mean(axis=1)
    u = 1.0 - b4.mean(axis=0) / (3.0 * b2.mean(axis=0) ** 2)
    jk = np.stack([1.0 - np.delete(b4, i, 0).mean(0) / (3.0 * np.delete(b2, i, 0).mean(0) ** 2)
                   for i in range(nblk)])
    err = np.sqrt((nblk - 1) / nblk * ((jk - jk.mean(0)) ** 2).sum(0))
    return u, err


def _z_stats(a, sa, b, sb):
    z = np.abs(a - b) / np.sqrt(sa ** 2 + sb ** 2 + 1e-300)
    return {"within_2sigma": float(np.mean(z <= 2.0)), "max_z": float(np.max(z)),
            "median_z": float(np.median(z))}


def form_c4(U, CF, sweeps=100, stat_sweeps=600, every=5, modes=("M", "H"), plain=True):
    """Batched MC: random binary bcc alloy, 64 replicas of 8^3..48^3 cells
    joined into one lattice, 32 temperatures 200..1600 K as ensembles,
    checkerboard Metropolis and heat bath; FP32 spins with FP64 observables.

    Throughput: `sweeps` sweeps in graphs of 10, one cumulant sample (energy
    included) per 10 sweeps, FP32 and FP64, against the plain FP64 engine run
    one replica at a time with all 32 temperatures per launch (graphs of 10,
    every 8th replica).  Observables: a further `stat_sweeps` sweeps from the
    ordered state, the first sixth discarded, |m| and energy per spin of every
    (replica, temperature) measured every `every` sweeps; <E>, <|m|> and U4
    with 2-sigma errors by 8 batch means (jackknife for U4), FP32 against
    FP64, per temperature for the largest replica and as z-scores over all
    2048 (replica, temperature) pairs."""
    from uppasd_amd import batch as Bt
    reps = _c4_replicas(CF)
    j = Bt.join(reps)
    M = 32
    temps = np.linspace(200.0, 1600.0, M)
    Ntot = j.natom
    nrep = len(reps)
    sizes = np.diff(j.offsets).astype(np.float64)
    out = {"workload": f"bcc alloy 64 replicas (8^3..48^3 cells, {Ntot} spins) x 32 T "
                       f"200..1600 K, Metropolis and heat bath",
           "unit": "spin-updates/s", "sweeps_timed": sweeps, "stat_sweeps": stat_sweeps}
    up = np.zeros((M, Ntot, 3))
    up[..., 2] = 1.0
    mm = np.ascontiguousarray(np.broadcast_to(j.mmom, (M, Ntot)))
    series = {}
    for mode in modes:
        for prec in ("fp32", "fp64"):
            eng = Bt.engine_for(j, M, prec)
            eng.set_moments(np.ascontiguousarray(up), mm)
            eng.autotune_mc(mode, temps, sweeps=1)
            eng.set_graph_steps(10)

            def run(eng=eng):
                for _ in range(sweeps // 10):
                    eng.mc_sweeps(mode, temps, 10, 7)
                    eng.accumulate(with_energy=True)
                eng.sync()
            t = _sync_time(run)
            out[f"{mode}_batched_{prec}_value"] = Ntot * M * sweeps / t
            if not stat_sweeps:
                del eng
                continue
            eng.set_moments(np.ascontiguousarray(up), mm)
            ms, es = [], []
            for k in range(stat_sweeps // every):
                eng.mc_sweeps(mode, temps, every, 11)
                if k >= stat_sweeps // every // 6:
                    msum, en = eng.measure_segments(with_energy=True)
                    ms.append(np.linalg.norm(msum, axis=2) / sizes[:, None])
                    es.append(en / sizes[:, None])
            series[(mode, prec)] = (np.array(ms), np.array(es))
            del eng
        out[f"{mode}_batched_fp32_over_fp64"] = (out[f"{mode}_batched_fp32_value"]
                                                 / out[f"{mode}_batched_fp64_value"])
        if not stat_sweeps:
            continue
        # FP32 against FP64 observables
        m32, e32 = series[(mode, "fp32")]
        m64, e64 = series[(mode, "fp64")]
        am32, sm32 = _batch_means(m32)
        am64, sm64 = _batch_means(m64)
        ae32, se32 = _batch_means(e32)
        ae64, se64 = _batch_means(e64)
        u32, su32 = _binder_blocks(m32)
        u64, su64 = _binder_blocks(m64)
        big = nrep - 1
        out[f"{mode}_observables"] = {
            "samples": int(m32.shape[0]),
            "z_all_pairs": {"E": _z_stats(ae32, se32, ae64, se64),
                            "abs_m": _z_stats(am32, sm32, am64, sm64),
                            "U4": _z_stats(u32, su32, u64, su64)},
            "largest_replica_per_T": {
                "T": [float(x) for x in temps],
                "E_fp32": [[float(a), 2 * float(b)] for a, b in zip(ae32[big], se32[big])],
                "E_fp64": [[float(a), 2 * float(b)] for a, b in zip(ae64[big], se64[big])],
                "abs_m_fp32": [[float(a), 2 * float(b)] for a, b in zip(am32[big], sm32[big])],
                "abs_m_fp64": [[float(a), 2 * float(b)] for a, b in zip(am64[big], sm64[big])],
                "U4_fp32": [[float(a), 2 * float(b)] for a, b in zip(u32[big], su32[big])],
                "U4_fp64": [[float(a), 2 * float(b)] for a, b in zip(u64[big], su64[big])]},
            "units": "E in mRy per spin, |m| in muB per spin; pairs [value, 2 sigma]"}
    # plain FP64: replica by replica (all temperatures per launch), graphs of 10
    for mode in modes if plain else ():
        t_plain, n_plain = 0.0, 0
        for r in range(0, nrep, 8):
            rep = reps[r]
            N = rep["lattice"].natom
            eng = U.Engine(N, M, "fp64")
            eng.set_exchange(rep["nlist"], rep["nlistsize"], rep["ncoup"])
            upr = np.zeros((M, N, 3))
            upr[..., 2] = 1.0
            eng.set_moments(upr, np.ascontiguousarray(np.broadcast_to(rep["mmom"], (M, N))))
            eng.set_graph_steps(10)
            eng.mc_sweeps(mode, temps, 10, 7)
            eng.sync()

            def run(eng=eng):
                for _ in range(3):
                    eng.mc_sweeps(mode, temps, 10, 7)
                    eng.accumulate(with_energy=True)
                eng.sync()
            t_plain += _sync_time(run)
            n_plain += N * M * 30
            del eng
        out[f"{mode}_plain_fp64_value"] = n_plain / t_plain
        out[f"{mode}_batched_fp32_over_plain_fp64"] = (out[f"{mode}_batched_fp32_value"]
                                                       / out[f"{mode}_plain_fp64_value"])
    return out


def form_c5(U, CF, steps=10000, graph=50, oracle_check=True):
    """Batched LLG: kagome antiferromagnet (J1 -1 mRy, Dz 0.08 mRy), 256
    ensembles of 24^2..96^2 cells at log-spaced 1..300 K joined into one
    lattice (each ensemble's sites numbered sublattice by sublattice, so its
    three sublattices are segments), Depondt with hipGraph-captured steps,
    alpha 0.2, dt 5e-17, `steps` steps; FP64 time-averaged energy and
    sublattice magnetisation from the device accumulators, one sample per
    graph of `graph` steps.  FP32 and FP64 batched; the plain FP64 path is
    the FP64 engine one ensemble at a time with the same graphs and samples
    (every 32nd ensemble, 1000 steps).  Error: the smallest ensemble (1 K)
    integrated by the FP64 oracle on the FP32 engine's own deviates, energy
    per spin at the end and time-averaged over the same samples."""
    from uppasd_amd import batch as Bt
    Ls = np.round(np.linspace(24, 96, 256)).astype(int)
    Ts = np.logspace(0.0, np.log10(300.0), 256)
    cache = {}
    reps = []
    for L, T in zip(Ls, Ts):
        if L not in cache:
            cache[L] = Bt.sublattice_major(CF.kagome_afm(int(L), J=-1.0, Dz=0.08), 3)
        d = dict(cache[L])
        d["temp"] = T
        reps.append(d)
    j = Bt.join(reps, per_site=("temp",))
    Ntot = j.natom
    seg = np.concatenate([[a + k * (b - a) // 3 for k in range(3)]
                          for a, b in zip(j.offsets[:-1], j.offsets[1:])] + [[Ntot]])
    rng = np.random.default_rng(55)
    e0 = np.ascontiguousarray(CF.random_unit(Ntot, rng, 1))
    out = {"workload": f"kagome AFM+DMI, 256 ensembles (24^2..96^2 cells, {Ntot} spins), "
                       f"T 1..300 K, Depondt alpha 0.2, dt 5e-17, {steps} steps, time-averaged "
                       f"energy and sublattice |m| per ensemble every {graph} steps",
           "unit": "spin-updates/s"}
    lam = np.full(Ntot, 0.2)
    avg = {}
    for prec in ("fp32", "fp64"):
        eng = Bt.engine_for(j, 1, prec)
        eng.set_segments(seg)
        eng.set_moments(e0, j.mmom[None])
        eng.set_llg(lam, j.extra["temp"], 5e-17)
        eng.set_noise(U.UASD_NOISE_PHILOX, 5)
        eng.autotune_llg(5, steps=2)
        eng.set_graph_steps(graph)

        def run(eng=eng):
            for _ in range(steps // graph):
                eng.sd_steps(5, graph)
                eng.accumulate(with_energy=True)
            eng.sync()
        t = _sync_time(run)
        out[f"batched_{prec}_value"] = Ntot * steps / t
        acc = eng.accumulators()
        e_ens = acc["e"][:, 0].reshape(256, 3).mean(axis=1)      # equal sublattices
        avg[prec] = (e_ens, acc["m"][:, 0].reshape(256, 3))
        if prec == "fp32":
            # ensemble 0 = segments 0..2 (its sublattices, equal in size)
            ene_final = float(eng.energy_terms()[0:3, 0].sum(axis=1).mean())
        del eng
    out["batched_fp32_over_fp64"] = out["batched_fp32_value"] / out["batched_fp64_value"]
    out["time_avg_energy_per_spin_mRy"] = {"T_1K": float(avg["fp32"][0][0]),
                                           "T_300K": float(avg["fp32"][0][-1])}
    out["time_avg_sublattice_m"] = {"T_1K": [float(x) for x in avg["fp32"][1][0]],
                                    "T_300K": [float(x) for x in avg["fp32"][1][-1]]}
    out["fp32_vs_fp64_time_avg_energy_rel_diff"] = {
        "median": float(np.median(np.abs(avg["fp32"][0] / avg["fp64"][0] - 1.0))),
        "max": float(np.max(np.abs(avg["fp32"][0] / avg["fp64"][0] - 1.0))),
        "note": "independent thermal streams (FP32 and FP64 deviates): a statistical difference"}
    if oracle_check:
        out["oracle_error_smallest_ensemble"] = _c5_oracle_error(reps[0], e0[:, :j.offsets[1]],
                                                                 steps, graph, ene_final,
                                                                 avg["fp32"][0][0])
    t_plain, n_plain = 0.0, 0
    for r in range(0, 256, 32):
        rep = reps[r]
        N = rep["lattice"].natom
        eng = U.Engine(N, 1, "fp64")
        eng.set_exchange(rep["nlist"], rep["nlistsize"], rep["ncoup"])
        eng.set_dm(rep["dmlist"], rep["dmlistsize"], rep["dm_vect"])
        eng.set_segments(np.array([0, N // 3, 2 * N // 3, N]))
        eng.set_moments(np.ascontiguousarray(CF.random_unit(N, rng, 1)), rep["mmom"][None])
        eng.set_llg(0.2, rep["temp"], 5e-17)
        eng.set_noise(U.UASD_NOISE_PHILOX, 5)
        eng.set_graph_steps(graph)
        eng.sd_steps(5, graph)
        eng.sync()

        def run(eng=eng):
            for _ in range(1000 // graph):
                eng.sd_steps(5, graph)
                eng.accumulate(with_energy=True)
            eng.sync()
        t_plain += _sync_time(run)
        n_plain += N * 1000
        del eng
    out["plain_fp64_value"] = n_plain / t_plain
    out["batched_fp32_over_plain_fp64"] = out["batched_fp32_value"] / out["plain_fp64_value"]
    return out


def _c5_oracle_error(rep, e0, steps, graph, e_final_dev, e_avg_dev):
    """The smallest C5 ensemble (first in the batch, so its device sites and
    Philox keys are 0..N-1) integrated by the FP64 oracle on the FP32
    engine's deviates: relative error of the FP32 device run's final and
    time-averaged energy per spin."""
    import ctypes
    import oracle as O
    N = rep["lattice"].natom
    ham = O.OracleHam(N, rep["nlist"], rep["nlistsize"], rep["ncoup"], dmlist=rep["dmlist"],
                      dmlistsize=rep["dmlistsize"], dm_vect=rep["dm_vect"])
    e = np.ascontiguousarray(e0.copy())
    mm = np.ascontiguousarray(rep["mmom"][None].astype(np.float64))
    eM = np.ascontiguousarray(e * mm[..., None])
    lam, T, lg = np.full(N, 0.2), np.full(N, float(rep["temp"])), np.ones(N)
    fc = 9.274009994e-24 / 2.179872325e-21
    samples = []
    for k in range(steps // graph):
        O.lib().or_sd_run(ctypes.byref(ham.h), N, 1, 5, graph, k * graph, 5e-17, lam, T, lg, mm,
                          e, eM, 5, 1, None)
        samples.append(ham.energy_terms(eM)[0].sum() * fc)
    w = np.arange(1, len(samples) + 1, dtype=np.float64)
    e_avg = float(np.sum(w * np.array(samples)) / np.sum(w))
    return {"N": N, "T_K": float(rep["temp"]), "steps": steps,
            "final_energy_rel_err": abs(e_final_dev - samples[-1]) / abs(samples[-1]),
            "time_avg_energy_rel_err": abs(e_avg_dev - e_avg) / abs(e_avg)}


def host_halves_variant(U, sysd, steps: int = 20, warmup: int = 3) -> dict:
    """The headline workload through the host-array drop-in halves, in the
    order a reference driver calls them every step (sd_driver.f90:517-836):
    effective_field -> evolve_first -> effective_field -> evolve_second on
    host arrays (each call uploads its inputs and downloads its outputs over
    PCIe), then moment_update on the host.  FP64, midpoint, 300 K, the
    in-kernel Philox stream.  A PCIe-inclusive rate beside the
    device-resident headline, never the metric."""
    N = sysd["lattice"].natom
    eng = U.Engine(N, 1, "fp64")
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    rng = np.random.default_rng(1)
    e = np.ascontiguousarray(CF_random_unit(N, rng))
    mm = np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (1, N)), dtype=np.float64)
    eng.set_moments(e, mm)
    eng.set_llg(0.1, 300.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, 0xC2C2C2C2)
    eM = np.ascontiguousarray(e * mm[..., None])
    e2 = e.copy()
    b2 = np.zeros_like(e)            # unused by the midpoint solver

    def step():
        beff, _ = eng.effective_field(eM)
        eng.evolve_first(1, beff, e, e2, eM)
        beff, _ = eng.effective_field(eM)
        eng.evolve_second(1, beff, b2, e, e2)
        U.moment_update(mm, e, e2, eM)

    for _ in range(warmup):
        step()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    dt = time.perf_counter() - t0
    drift = float(np.max(np.abs(np.linalg.norm(e, axis=2) - 1.0)))
    del eng
    return {"workload": f"bcc 64^3 x 2 ({N} spins), 58 neighbours, midpoint, noise philox, "
                        f"T 300 K, host arrays through the drop-in halves",
            "value": N * steps / dt, "ms_per_step": dt / steps * 1e3, "steps": steps,
            "max_norm_drift": drift,
            "note": "PCIe-inclusive: every call moves its arrays between host and device; "
                    "moment_update runs on the host; not the metric"}


def c2_variant(U, args, label: str, sysd=None, L: int = 64, temp: float = 300.0,
               noise: str = "philox", stencil: bool = True, steps: int = 400,
               warmup: int = 40, M: int = 1):
    """The plain FP64 workload under another noise source, size or kernel
    family, bounded to a few hundred steps: value, step time and roofline
    fractions by the workload's byte model and by the kernels' own bytes,
    with the counter traffic of the committed profile of the same run where
    there is one.  noise: "philox" (in-kernel), "injected" (one N(0,1) array
    uploaded once and read by both halves of every step; a drop-in caller
    would upload a fresh one per step over PCIe), "none" (T = 0).  M > 1: M
    ensembles of the lattice, each with its own start and noise."""
    if sysd is None:
        sysd = build_c2(L)
    lat = sysd["lattice"]
    N = lat.natom
    W = int(sysd["nlistsize"].max())
    eng = U.Engine(N, M, args.precision, site_order=lat.NA)
    if stencil:
        eng.set_lattice_dims(lat.NA, *lat.n)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    rng = np.random.default_rng(2)
    e0 = rng.normal(size=(M, N, 3))
    e0 /= np.linalg.norm(e0, axis=2, keepdims=True)
    eng.set_moments(np.ascontiguousarray(e0), np.broadcast_to(sysd["mmom"], (M, N)))
    eng.set_llg(0.1, temp if noise != "none" else 0.0, 1e-16)
    if noise == "philox":
        eng.set_noise(U.UASD_NOISE_PHILOX, 0xC2C2C2C2)
    elif noise == "injected":
        eng.set_noise(U.UASD_NOISE_INJECTED, 0, np.ascontiguousarray(rng.normal(size=(M, N, 3))))
    else:
        eng.set_noise(U.UASD_NOISE_NONE)
    eng.autotune_llg(args.solver, steps=3)
    eng.sd_steps(args.solver, warmup)
    eng.sync()
    tm = timed_steps(eng, args.solver, steps, args.timing_every)
    g, b = eng.tuning(eng.TUNE_LLG)
    tile, pipe = eng.llg_tile(), (b >> 22) & 1     # the kernel the step launches
    s = 8 if args.precision == "fp64" else 4
    ib = 0 if tile else eng.index_bytes()[0]
    kb = kernel_bytes(W, s, ib, cls=0 if tile else 4, noise=noise, b2=args.solver in (4, 5))
    sb = model_bytes(W, s, injected=noise == "injected")
    del eng
    variant = f"tile{tile}{'r' if pipe else ''}" if tile else f"idx{ib}"
    tr = profiled_traffic(label, variant)
    traffic = tr["predictor"] + tr["corrector"] if tr else None
    NM = N * M
    out = {"workload": f"bcc {lat.n[0]}^3 x 2 ({N} spins{f' x {M} ensembles' if M > 1 else ''}), "
                       f"{W} neighbours, {SOLVER_NAMES[args.solver]}, noise {noise}"
                       f"{'' if noise == 'none' else f', T {temp:g} K'}",
           "value": NM * steps / tm["wall_s"], "ms_per_step": tm["step_s"] * 1e3,
           "kernel_ms": [round(tm["pred_s"] * 1e3, 4), round(tm["corr_s"] * 1e3, 4)],
           "kernel": variant,
           "bytes_per_update_model": sb["step"], "own_bytes_per_site": kb["step"],
           "frac_model": sb["step"] * NM / tm["step_s"] / 1e9 / HBM_PEAK_GBS,
           "frac_own_bytes": kb["step"] * NM / tm["step_s"] / 1e9 / HBM_PEAK_GBS,
           "traffic_per_step": traffic, "steps": steps}
    if traffic:
        out["traffic_over_own_bytes"] = traffic / (kb["step"] * NM)
    return out


def profiled_traffic(label: str, variant: str) -> dict:
    """HBM bytes per launch of each LLG kernel half from the committed
    rocprofv3 summary `profiles/k_llg_<label>_<variant>.md` (its FETCH_SIZE /
    WRITE_SIZE table, already scaled by a plain copy of known size):
    {"predictor": B, "corrector": B}, or {} when there is no such summary."""
    path = os.path.join(ROOT, "profiles", f"k_llg_{label}_{variant}.md")
    if not os.path.exists(path):
        return {}
    in_table, out = False, {}
    for line in open(path):
        cells = [c.strip() for c in line.split("|")]
        if len(cells) > 3 and cells[1] == "kernel":
            in_table = cells[2].startswith("FETCH_SIZE")
            continue
        if in_table and len(cells) > 6 and cells[1].startswith("`uasd::k_llg"):
            half = int(cells[1].split(",")[2])      # k_llg[_tile]<R, ALG, HALF, ...>
            out["corrector" if half else "predictor"] = (float(cells[4]) + float(cells[5])) * 1e6
    return out if len(out) == 2 else {}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--tune-cache", default=os.path.join(ROOT, "uppasd_amd", "tune_gfx950.txt"),
                    help="autotuner cache file: winners kept there are reused without timing, "
                         "new ones appended; 'none' times every candidate")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain copy kernel of known size (for counter scaling)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--solver", type=int, choices=[1, 4, 5, 6, 7], default=1,
                    help="SDEalgh: 1 midpoint, 4 Heun, 5 Depondt, 6 SIB, 7 spherical midpoint")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--timing-every", type=int, default=16,
                    help="HIP events around one step in this many (events cost time of their own)")
    ap.add_argument("--gather", type=int, default=0,
                    help="fixed gather batch (2/4/8/16) instead of autotuning")
    ap.add_argument("--block", type=int, default=256,
                    help="workgroup size with --gather; add 65536 for non-temporal index loads")
    ap.add_argument("--no-list-compare", action="store_true",
                    help="skip timing the list-gather kernels beside the tiled ones")
    ap.add_argument("--no-stencil", action="store_true",
                    help="do not declare the lattice (list-gather kernels only)")
    ap.add_argument("--site-order", type=int, default=0,
                    help="device site order: 0 = basis-major (NA of the lattice), 1 = reference")
    ap.add_argument("--noise", choices=["philox", "injected", "none"], default="philox",
                    help="thermal noise of the main run: in-kernel Philox at 300 K, an injected "
                         "N(0,1) array at 300 K, or none (T = 0)")
    ap.add_argument("--variants", default="T0,injected,L128,M8,host",
                    help="comma list of runs of the plain workload beside the headline: T0 "
                         "(no noise), injected (caller's N(0,1) stream), L128 (bcc 128^3 x 2, "
                         "above the MALL), M8 (eight ensembles of bcc 32^3 x 2), host (host "
                         "arrays through the drop-in halves, PCIe-inclusive), or 'none'")
    ap.add_argument("--c4-sweeps", type=int, default=100,
                    help="timed sweeps of each batched MC run of the c4 form")
    ap.add_argument("--c4-stat-sweeps", type=int, default=600,
                    help="sweeps of the c4 form's FP32-vs-FP64 observables (0: none)")
    ap.add_argument("--c4-modes", default="M,H", help="MC modes of the c4 form")
    ap.add_argument("--no-c4-plain", action="store_true",
                    help="skip the c4 form's plain FP64 replica-by-replica runs")
    ap.add_argument("--forms", default="c3,c4,c5",
                    help="comma list of extra forms to run (c3,c4,c5) or 'none'")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(the engine is single-GPU; a multi-GPU run would shard the lattice by slabs).",
              file=sys.stderr)
        sys.exit(2)

    import uppasd_amd as U
    if args.tune_cache != "none":
        U.set_tune_cache(True, args.tune_cache)

    t_setup = time.perf_counter()
    sysd = build_c2(args.L)
    N = sysd["lattice"].natom
    W = int(sysd["nlistsize"].max())
    order = sysd["lattice"].NA if args.site_order == 0 else args.site_order
    eng = U.Engine(N, 1, args.precision, site_order=order)
    lat = sysd["lattice"]
    if not args.no_stencil:
        eng.set_lattice_dims(lat.NA, *lat.n)     # enables the LDS-tiled stencil kernels
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    rng = np.random.default_rng(2)
    e0 = rng.normal(size=(1, N, 3))
    e0 /= np.linalg.norm(e0, axis=2, keepdims=True)
    eng.set_moments(np.ascontiguousarray(e0), sysd["mmom"][None])
    eng.set_llg(0.1, 0.0 if args.noise == "none" else 300.0, 1e-16)
    if args.noise == "philox":
        eng.set_noise(U.UASD_NOISE_PHILOX, 0xC2C2C2C2)
    elif args.noise == "injected":
        eng.set_noise(U.UASD_NOISE_INJECTED, 0, np.ascontiguousarray(rng.normal(size=(1, N, 3))))
    else:
        eng.set_noise(U.UASD_NOISE_NONE)
    setup_s = time.perf_counter() - t_setup

    t_tune = time.perf_counter()
    if args.gather:
        eng.set_tuning(eng.TUNE_LLG, args.gather, args.block)
    else:
        eng.autotune_llg(args.solver, steps=3)
    tune_s = time.perf_counter() - t_tune
    gather, block = eng.tuning(eng.TUNE_LLG)
    nt, idx32, tile = (block >> 16) & 1, (block >> 17) & 1, eng.llg_tile()
    pipe = (block >> 22) & 1
    block &= 0xFFFF

    eng.sd_steps(args.solver, args.warmup)
    eng.sync()
    tm = timed_steps(eng, args.solver, args.steps, args.timing_every)
    elapsed = tm["wall_s"]
    t_pred, t_corr = tm["pred_s"], tm["corr_s"]

    s = 8 if args.precision == "fp64" else 4
    ib = 0 if tile else eng.index_bytes()[0]     # the tile kernel reads no lists
    kb = kernel_bytes(W, s, ib, cls=0 if tile else 4,   # one species: classes per basis
                      noise=args.noise, b2=args.solver in (4, 5))
    sb = model_bytes(W, s, injected=args.noise == "injected")
    # the dominant kernel is the corrector: the workload's algorithmic bytes
    # per update of its pass over every site, per its duration
    achieved = sb["corrector"] * N / t_corr / 1e9
    achieved_own = kb["corrector"] * N / t_corr / 1e9
    value = N * args.steps / elapsed
    traffic = None
    if args.precision == "fp64" and args.solver == 1 and order == 2:
        variant = f"tile{tile}{'r' if pipe else ''}" if tile else f"idx{ib}"
        noise_label = {"philox": "philox", "injected": "injected", "none": "T0"}[args.noise]
        traffic = profiled_traffic(f"fp64_midpoint_bcc{args.L}_58nn_{noise_label}",
                                   variant).get("corrector")
    _, eM = eng.get_moments()
    norm_drift = float(np.max(np.abs(np.linalg.norm(eM[0], axis=1) / sysd["mmom"] - 1.0)))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": tm["step_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"bcc {args.L}^3 x 2 ({N} spins), 5 exchange shells ({W} neighbours), "
                         f"{SOLVER_NAMES[args.solver]} LLG, alpha 0.1, dt 1e-16 s, "
                         + {"philox": "T 300 K Philox thermal field",
                            "injected": "T 300 K injected N(0,1) thermal field",
                            "none": "T 0"}[args.noise]
                         + f", {args.precision}, device-resident"),
            "spins": N, "neighbours": W, "solver": args.solver, "precision": args.precision,
            "setup_s": round(setup_s, 2),
            "launch": {"gather": gather, "block": block, "nt_index_loads": bool(nt),
                       "int32_index": bool(idx32),
                       "tile_shape": U.engine.TILE_SHAPES[tile - 1] if tile else None,
                       "resident_tile_loop": bool(pipe),
                       "site_order": order,
                       "autotuned": not args.gather, "tune_s": round(tune_s, 3),
                       "tune_cache": None if args.tune_cache == "none" else
                       os.path.relpath(args.tune_cache, ROOT)},
            "wall_ms_per_step": elapsed / args.steps * 1e3,
            "kernel_ms": {"predictor": round(t_pred * 1e3, 4), "corrector": round(t_corr * 1e3, 4),
                          "sampled": tm["sampled_ms"], "sampled_launches": tm["sampled_launches"],
                          "method": ("step time from HIP events around the timed call, split "
                                     "by the halves' shares of events around one step in "
                                     f"{args.timing_every} of a second run")},
            "bytes_per_update_model": sb,
            "bytes_per_site": kb,
            "index_bytes": ib,
            "max_norm_drift": norm_drift,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "bytes_per_update": sb["corrector"],
            "own_bytes_per_site": kb["corrector"],
            "own_bytes_achieved": achieved_own,
            "own_bytes_frac": achieved_own / HBM_PEAK_GBS,
            "step_frac": sb["step"] * N / tm["step_s"] / 1e9 / HBM_PEAK_GBS,
        },
    }
    del eng
    variants = [v for v in args.variants.split(",") if v and v != "none"]
    if tile and not args.no_list_compare:
        variants = ["list"] + variants
    if variants:
        out["variants"] = {}
    for v in variants:
        t0 = time.perf_counter()
        if v == "list":      # the same workload on the list-gather kernels
            res = c2_variant(U, args, "fp64_midpoint_bcc64_58nn_philox", sysd=sysd,
                             stencil=False)
        elif v == "T0":
            res = c2_variant(U, args, "fp64_midpoint_bcc64_58nn_T0", sysd=sysd, noise="none")
        elif v == "injected":
            res = c2_variant(U, args, "fp64_midpoint_bcc64_58nn_injected", sysd=sysd,
                             noise="injected")
        elif v == "L128":    # above the 256 MB MALL
            res = c2_variant(U, args, "fp64_midpoint_bcc128_58nn_philox", L=128, steps=200,
                             warmup=20)
        elif v == "M8":      # the same spin count as eight ensembles of bcc 32^3 x 2
            res = c2_variant(U, args, "fp64_midpoint_bcc32_58nn_M8_philox", L=32, M=8)
        elif v == "host":    # host arrays through the drop-in halves (PCIe-inclusive)
            res = host_halves_variant(U, sysd)
        else:
            raise SystemExit(f"unknown variant {v}")
        res["wall_s"] = round(time.perf_counter() - t0, 1)
        out["variants"][v] = res
        print(f"variant {v} done in {res['wall_s']} s", file=sys.stderr, flush=True)
    forms = [f for f in args.forms.split(",") if f and f != "none"]
    if forms:
        from uppasd_amd import configs as CF
        out["forms"] = {}
        for f in forms:
            t0 = time.perf_counter()
            if f == "c4":
                res = form_c4(U, CF, sweeps=args.c4_sweeps, stat_sweeps=args.c4_stat_sweeps,
                              modes=tuple(args.c4_modes.split(",")), plain=not args.no_c4_plain)
            else:
                res = {"c3": form_c3, "c5": form_c5}[f](U, CF)
            res["wall_s"] = round(time.perf_counter() - t0, 1)
            out["forms"][f] = res
            print(f"form {f} done in {res['wall_s']} s", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sysd)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
