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
on; the cpu_baseline object times the FP64 C oracle (one thread) on a bounded
sample of the same workload.
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


def kernel_bytes(W: int, s: int, ib: int = 4, cls: int = 4, nsave: int = 16,
                 b2: bool = False) -> dict:
    """Algorithmic HBM bytes per site of the fused LLG kernels.

    Spins are stored as quads {x, y, z, mmom} of s-byte reals; neighbour
    entries take ib bytes (4 for int32 indices, 2 for int16 differences, 0 for
    the LDS-tiled stencil kernels, which read no lists); cls bytes of site
    class per site (0 for the tiled kernels on a one-species lattice, whose
    classes come with the launch); nsave bytes of thermal deviates the
    predictor stores for its corrector (a float4; 0 without Philox noise).
    predictor: gather e_j quads (4s), index stream ib*W, class, write p (4s),
               write deviates
    corrector: gather p_j quads (4s), own e (4s), index ib*W, class, write e (4s),
               read deviates
    Each site's own spin is one of its neighbours' gathers, so the gathered
    quads count once per site (perfect reuse through L2 / MALL).  b2: the
    solver keeps a field quad between the halves (Depondt's bdup, Heun's
    b2eff), written by the predictor and read by the corrector (4s each).
    """
    q = 4 * s if b2 else 0
    pred = 4 * s + ib * W + cls + 4 * s + nsave + q
    corr = 4 * s + 4 * s + ib * W + cls + 4 * s + nsave + q
    return {"predictor": pred, "corrector": corr, "step": pred + corr}


def profile_path(order: int, variant: str) -> str:
    """Committed rocprofv3 summary of the FP64 midpoint workload for a device
    site order and kernel variant (idx4 / idx2 list gathers, tileK stencil with
    one tile per workgroup, tileKr with resident workgroups looping over tiles)."""
    return os.path.join(ROOT, "profiles",
                        f"k_llg_fp64_midpoint_bcc64_58nn_order{order}_{variant}.md")


def profiled_traffic(kernel: str, order: int, variant: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary of
    the same configuration (FETCH_SIZE + WRITE_SIZE scaled by a plain copy),
    or None when there is none for it."""
    path = profile_path(order, variant)
    if not os.path.exists(path):
        return None
    in_table = False          # rows of the FETCH_SIZE / WRITE_SIZE table only
    for line in open(path):
        cells = [c.strip() for c in line.split("|")]
        if len(cells) > 3 and cells[1] == "kernel":
            in_table = len(cells) > 3 and cells[2].startswith("FETCH_SIZE")
            continue
        if in_table and len(cells) > 6 and cells[1] == f"`{kernel}`":
            return (float(cells[4]) + float(cells[5])) * 1e6
    return None


def tile_plane(sysd, tile: int) -> int:
    """Component plane size the tile kernel launches with (its template
    argument): the smallest of 1160, 2312, 3080, 3848 holding NA * HV, for the
    bcc five-shell stencil (reach 2 cells each way)."""
    from uppasd_amd.engine import TILE_SHAPES
    tx, ty, tz = TILE_SHAPES[tile - 1]
    nahv = sysd["lattice"].NA * (tx + 4) * (ty + 4) * (tz + 4)
    return next(p for p in (1160, 2312, 3080, 3848) if nahv <= p)


def build_c2(L: int):
    from uppasd_amd import configs as CF
    return CF.bcc_shells(L, CF.bcc5_couplings(), moment=2.2)


def cpu_baseline(sysd, seconds_target: float = 12.0):
    """Oracle (plain C, FP64, 1 thread) on the same Hamiltonian."""
    import ctypes
    import oracle as O
    L = O.lib()
    N = sysd["lattice"].natom
    ham = O.OracleHam(N, sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    rng = np.random.default_rng(1)
    e = rng.normal(size=(1, N, 3))
    e /= np.linalg.norm(e, axis=2, keepdims=True)
    e = np.ascontiguousarray(e)
    mm = np.ascontiguousarray(np.broadcast_to(sysd["mmom"], (1, N)))
    eM = np.ascontiguousarray(e * mm[..., None])
    lam, T, lg = np.full(N, 0.1), np.full(N, 300.0), np.ones(N)
    t0 = time.perf_counter()
    L.or_sd_run(ctypes.byref(ham.h), N, 1, 1, 1, 0, 1e-16, lam, T, lg, mm, e, eM, 7, 1, None)
    one = time.perf_counter() - t0
    nsteps = max(1, int(seconds_target / max(one, 1e-6)))
    t0 = time.perf_counter()
    L.or_sd_run(ctypes.byref(ham.h), N, 1, 1, nsteps, 1, 1e-16, lam, T, lg, mm, e, eM, 7, 1, None)
    dt = time.perf_counter() - t0
    return {"value": N * nsteps / dt, "unit": "spin-updates/s", "cores": 1, "kind": "port",
            "sample": f"same workload ({N} spins, 58 neighbours, midpoint FP64, Philox noise "
                      f"at 300 K), {nsteps} steps in {dt:.1f} s on one host core"}


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
    for prec in ("fp64", "fp32", "fp64_perturbed"):
        eng = U.Engine(N, 1, prec[:4])
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


def form_c4(U, CF, sweeps=40):
    """Batched MC: random binary bcc alloy, 64 replicas of 8^3..48^3 cells
    joined into one lattice, 32 temperatures 200..1600 K as ensembles,
    checkerboard Metropolis; FP32 spins with FP64 observables (one cumulant
    sample per sweep).  The plain path is the FP64 engine run one replica at
    a time (all 32 temperatures per launch), on a bounded sample."""
    from uppasd_amd import batch as Bt
    reps = _c4_replicas(CF)
    j = Bt.join(reps)
    M = 32
    temps = np.linspace(200.0, 1600.0, M)
    Ntot = j.natom
    rng = np.random.default_rng(44)
    out = {"workload": f"bcc alloy 64 replicas (8^3..48^3 cells, {Ntot} spins) x 32 T, Metropolis",
           "unit": "spin-updates/s"}
    for prec in ("fp32", "fp64"):
        eng = Bt.engine_for(j, M, prec)
        e0 = np.ascontiguousarray(CF.random_unit(Ntot, rng, M))
        eng.set_moments(e0, np.ascontiguousarray(np.broadcast_to(j.mmom, (M, Ntot))))
        eng.autotune_mc("M", temps, sweeps=1)
        eng.set_graph_steps(10)
        eng.mc_sweeps("M", temps, 10, 7)
        eng.sync()

        def run(eng=eng):
            for _ in range(sweeps // 10):
                eng.mc_sweeps("M", temps, 10, 7)
                eng.accumulate(with_energy=True)
            eng.sync()
        t = _sync_time(run)
        out[f"batched_{prec}_value"] = Ntot * M * sweeps / t
        if prec == "fp32":
            acc = eng.accumulators()
            out["U4_first_replica_lowT_highT"] = [float(acc["U4"][0, 0]), float(acc["U4"][0, -1])]
        del eng
    # plain FP64: replica by replica, eager launches, bounded to 8 replicas
    t_plain, n_plain = 0.0, 0
    for r in range(0, len(reps), 8):
        rep = reps[r]
        N = rep["lattice"].natom
        eng = U.Engine(N, M, "fp64")
        eng.set_exchange(rep["nlist"], rep["nlistsize"], rep["ncoup"])
        eng.set_moments(np.ascontiguousarray(CF.random_unit(N, rng, M)),
                        np.ascontiguousarray(np.broadcast_to(rep["mmom"], (M, N))))
        eng.mc_sweeps("M", temps, 2, 7)
        eng.sync()

        def run(eng=eng):
            eng.mc_sweeps("M", temps, 10, 7)
            eng.sync()
        t_plain += _sync_time(run)
        n_plain += N * M * 10
    out["plain_fp64_value"] = n_plain / t_plain
    out["batched_fp32_over_plain_fp64"] = out["batched_fp32_value"] / out["plain_fp64_value"]
    return out


def form_c5(U, CF, steps=2000):
    """Batched LLG: kagome antiferromagnet (J1 -1 mRy, Dz 0.08 mRy), 256
    ensembles of 24^2..96^2 cells at log-spaced 1..300 K joined into one
    lattice, FP32 Depondt with hipGraph-captured steps, alpha 0.2, dt 5e-17.
    Plain path: the FP64 engine one ensemble at a time with eager launches
    (bounded sample)."""
    from uppasd_amd import batch as Bt
    Ls = np.round(np.linspace(24, 96, 256)).astype(int)
    Ts = np.logspace(0.0, np.log10(300.0), 256)
    cache = {}
    reps = []
    for L, T in zip(Ls, Ts):
        if L not in cache:
            cache[L] = CF.kagome_afm(int(L), J=-1.0, Dz=0.08)
        d = dict(cache[L])
        d["temp"] = T
        reps.append(d)
    j = Bt.join(reps, per_site=("temp",))
    Ntot = j.natom
    rng = np.random.default_rng(55)
    e0 = np.ascontiguousarray(CF.random_unit(Ntot, rng, 1))
    out = {"workload": f"kagome AFM+DMI, 256 ensembles (24^2..96^2 cells, {Ntot} spins), "
                       f"T 1..300 K, Depondt alpha 0.2, dt 5e-17, {steps} steps",
           "unit": "spin-updates/s"}
    lam = np.full(Ntot, 0.2)
    for prec, graph in (("fp32", 50), ("fp64", 50)):
        eng = Bt.engine_for(j, 1, prec)
        eng.set_moments(e0, j.mmom[None])
        eng.set_llg(lam, j.extra["temp"], 5e-17)
        eng.set_noise(U.UASD_NOISE_PHILOX, 5)
        eng.autotune_llg(5, steps=2)
        eng.set_graph_steps(graph)
        eng.sd_steps(5, graph)
        eng.sync()

        def run(eng=eng):
            for _ in range(steps // graph):
                eng.sd_steps(5, graph)
                eng.accumulate(with_energy=False)
            eng.sync()
        t = _sync_time(run)
        out[f"batched_{prec}_value"] = Ntot * steps / t
        del eng
    t_plain, n_plain = 0.0, 0
    for r in range(0, 256, 32):
        rep = reps[r]
        N = rep["lattice"].natom
        eng = U.Engine(N, 1, "fp64")
        eng.set_exchange(rep["nlist"], rep["nlistsize"], rep["ncoup"])
        eng.set_dm(rep["dmlist"], rep["dmlistsize"], rep["dm_vect"])
        eng.set_moments(np.ascontiguousarray(CF.random_unit(N, rng, 1)), rep["mmom"][None])
        eng.set_llg(0.2, rep["temp"], 5e-17)
        eng.set_noise(U.UASD_NOISE_PHILOX, 5)
        eng.sd_steps(5, 5)
        eng.sync()

        def run(eng=eng):
            eng.sd_steps(5, 100)
            eng.sync()
        t_plain += _sync_time(run)
        n_plain += N * 100
    out["plain_fp64_value"] = n_plain / t_plain
    out["batched_fp32_over_plain_fp64"] = out["batched_fp32_value"] / out["plain_fp64_value"]
    return out


def list_kernel_compare(U, sysd, args, order, e0, steps=300):
    """The same workload on the list-gather kernels (lattice not declared, so
    no stencil), autotuned: its corrector time and roofline beside the tiled
    kernel's.  Bounded to a few hundred steps."""
    N = sysd["lattice"].natom
    W = int(sysd["nlistsize"].max())
    eng = U.Engine(N, 1, args.precision, site_order=order)
    eng.set_exchange(sysd["nlist"], sysd["nlistsize"], sysd["ncoup"])
    eng.set_moments(np.ascontiguousarray(e0), sysd["mmom"][None])
    eng.set_llg(0.1, 300.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, 0xC2C2C2C2)
    eng.autotune_llg(args.solver, steps=3)
    eng.sd_steps(args.solver, 20)
    eng.sync()
    eng.set_kernel_timing(True, args.timing_every)
    t0 = time.perf_counter()
    eng.sd_steps(args.solver, steps)
    eng.sync()
    elapsed = time.perf_counter() - t0
    n, ms = eng.kernel_timing(1)
    t = ms / max(n, 1) * 1e-3
    s = 8 if args.precision == "fp64" else 4
    ib = eng.index_bytes()[0]
    kb = kernel_bytes(W, s, ib, b2=args.solver in (4, 5))
    ach = kb["corrector"] * N / t / 1e9
    g, b = eng.tuning(eng.TUNE_LLG)
    return {"value": N * steps / elapsed, "corrector_ms": round(t * 1e3, 4),
            "bytes_per_site": kb, "index_bytes": ib, "gather": g, "block": b & 0xFFFF,
            "roofline": {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": ach / HBM_PEAK_GBS}}


def main():
    ap = argparse.ArgumentParser()
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
    eng.set_llg(0.1, 300.0, 1e-16)
    eng.set_noise(U.UASD_NOISE_PHILOX, 0xC2C2C2C2)
    setup_s = time.perf_counter() - t_setup

    t_tune = time.perf_counter()
    if args.gather:
        eng.set_tuning(eng.TUNE_LLG, args.gather, args.block)
    else:
        eng.autotune_llg(args.solver, steps=3)
    tune_s = time.perf_counter() - t_tune
    gather, block = eng.tuning(eng.TUNE_LLG)
    nt, idx32, tile = (block >> 16) & 1, (block >> 17) & 1, (block >> 18) & 0xF
    pipe = (block >> 22) & 1
    block &= 0xFFFF

    eng.sd_steps(args.solver, args.warmup)
    eng.sync()
    eng.set_kernel_timing(True, args.timing_every)
    eng.sync()
    t0 = time.perf_counter()
    eng.sd_steps(args.solver, args.steps)
    eng.sync()
    elapsed = time.perf_counter() - t0
    n_corr, ms_corr = eng.kernel_timing(1)
    n_pred, ms_pred = eng.kernel_timing(0)
    eng.set_kernel_timing(False)

    s = 8 if args.precision == "fp64" else 4
    ib = 0 if tile else eng.index_bytes()[0]     # the tile kernel reads no lists
    kb = kernel_bytes(W, s, ib, cls=0 if tile else 4,   # one species: classes per basis
                      b2=args.solver in (4, 5))
    t_corr = ms_corr / max(n_corr, 1) * 1e-3
    t_pred = ms_pred / max(n_pred, 1) * 1e-3
    achieved = kb["corrector"] * N / t_corr / 1e9
    value = N * args.steps / elapsed
    traffic = None
    if args.precision == "fp64" and args.L == 64 and args.solver == 1 and not nt:
        if tile:
            plane = tile_plane(sysd, tile)
            # exchange-only Hamiltonian: the kernel without DM/anisotropy/Zeeman terms
            traffic = profiled_traffic(f"uasd::k_llg_tile<double, 1, 1, {plane}, false>", order,
                                       f"tile{tile}{'r' if pipe else ''}")
        else:
            traffic = profiled_traffic(f"uasd::k_llg<double, 1, 1, true, {gather}>", order,
                                       f"idx{ib}")
    _, eM = eng.get_moments()
    norm_drift = float(np.max(np.abs(np.linalg.norm(eM[0], axis=1) / sysd["mmom"] - 1.0)))

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "spin-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {
            "workload": (f"bcc {args.L}^3 x 2 ({N} spins), 5 exchange shells ({W} neighbours), "
                         f"{SOLVER_NAMES[args.solver]} LLG, alpha 0.1, "
                         f"dt 1e-16 s, T 300 K Philox thermal field, {args.precision}, "
                         "device-resident"),
            "spins": N, "neighbours": W, "solver": args.solver, "precision": args.precision,
            "setup_s": round(setup_s, 2),
            "launch": {"gather": gather, "block": block, "nt_index_loads": bool(nt),
                       "int32_index": bool(idx32),
                       "tile_shape": U.engine.TILE_SHAPES[tile - 1] if tile else None,
                       "resident_tile_loop": bool(pipe),
                       "site_order": order,
                       "autotuned": not args.gather, "tune_s": round(tune_s, 3)},
            "kernel_ms": {"predictor": round(t_pred * 1e3, 4), "corrector": round(t_corr * 1e3, 4),
                          "timed_launches": [n_pred, n_corr],
                          "method": f"HIP events around one step in {args.timing_every}"},
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
        },
    }
    del eng
    if tile and not args.no_list_compare:
        out["list_kernel"] = list_kernel_compare(U, sysd, args, order, e0)
    forms = [f for f in args.forms.split(",") if f and f != "none"]
    if forms:
        from uppasd_amd import configs as CF
        out["forms"] = {}
        for f in forms:
            t0 = time.perf_counter()
            res = {"c3": form_c3, "c4": form_c4, "c5": form_c5}[f](U, CF)
            res["wall_s"] = round(time.perf_counter() - t0, 1)
            out["forms"][f] = res
            print(f"form {f} done in {res['wall_s']} s", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sysd)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
