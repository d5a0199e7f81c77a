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


def kernel_bytes(W: int, s: int) -> dict:
    """Algorithmic HBM bytes per site of the fused midpoint kernels.

    Spins are stored as quads {x, y, z, mmom} of s-byte reals.
    predictor: gather e_j quads (4s), index stream 4W, class id 4, write p (4s)
    corrector: gather p_j quads (4s), own e (4s), index 4W, class id 4, write e (4s)
    Each site's own spin is one of its neighbours' gathers, so the gathered
    quads count once per site (perfect reuse through L2 / MALL).
    """
    pred = 4 * s + 4 * W + 4 + 4 * s
    corr = 4 * s + 4 * s + 4 * W + 4 + 4 * s
    return {"predictor": pred, "corrector": corr, "step": pred + corr}


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain copy kernel of known size (for counter scaling)")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--L", type=int, default=64)
    ap.add_argument("--precision", choices=["fp64", "fp32"], default="fp64")
    ap.add_argument("--solver", type=int, choices=[1, 5], default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--gather", type=int, default=0,
                    help="fixed gather batch (2/4/8/16) instead of autotuning")
    ap.add_argument("--block", type=int, default=256)
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported in this build "
              "(the engine is single-GPU; a multi-GPU run would shard the lattice by slabs).",
              file=sys.stderr)
        sys.exit(2)

    import uppasd_amd as U

    t_setup = time.perf_counter()
    sysd = build_c2(args.L)
    N = sysd["lattice"].natom
    W = int(sysd["nlistsize"].max())
    eng = U.Engine(N, 1, args.precision)
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

    eng.sd_steps(args.solver, args.warmup)
    eng.sync()
    eng.set_kernel_timing(True)
    eng.sync()
    t0 = time.perf_counter()
    eng.sd_steps(args.solver, args.steps)
    eng.sync()
    elapsed = time.perf_counter() - t0
    n_corr, ms_corr = eng.kernel_timing(1)
    n_pred, ms_pred = eng.kernel_timing(0)
    eng.set_kernel_timing(False)

    s = 8 if args.precision == "fp64" else 4
    kb = kernel_bytes(W, s)
    t_corr = ms_corr / max(n_corr, 1) * 1e-3
    t_pred = ms_pred / max(n_pred, 1) * 1e-3
    achieved = kb["corrector"] * N / t_corr / 1e9
    value = N * args.steps / elapsed
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
                         f"{'midpoint' if args.solver == 1 else 'Depondt'} LLG, alpha 0.1, "
                         f"dt 1e-16 s, T 300 K Philox thermal field, {args.precision}, "
                         "device-resident"),
            "spins": N, "neighbours": W, "solver": args.solver, "precision": args.precision,
            "setup_s": round(setup_s, 2),
            "launch": {"gather": gather, "block": block,
                       "autotuned": not args.gather, "tune_s": round(tune_s, 3)},
            "kernel_ms": {"predictor": round(t_pred * 1e3, 4), "corrector": round(t_corr * 1e3, 4)},
            "bytes_per_site": kb,
            "max_norm_drift": norm_drift,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sysd)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
