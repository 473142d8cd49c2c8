#!/usr/bin/env python3
"""Benchmark of the MI355X NDDO SCF engine on its headline workload.

Metric (north star): PM3 SCF solves/s with p50/p99 latency, mixed precision vs
the FP64 GPU path, at |dE| <= 1e-5 eV.

Workload of one step: the batched PM3 mix -- 4096 neutral closed-shell
molecules, even index an all-trans alkane C_nH_{2n+2} with n ~ U{1..40}, odd
index a water cluster (H2O)_m with m ~ U{2..64}, seeded geometries with
N(0, 0.02 A) jitter -- solved to density RMS < 1e-8 (core-Hamiltonian guess,
DIIS k = 5 / start 0.1 / end 0, damping 0.7 above RMS 0.02) in ONE batch: one launch per SCF
stage for all molecules.  Inputs are resident on the device before the timed
region (each step re-uploads the geometry -- a few MB -- and rebuilds all
integrals on the GPU: that setup is part of every solve).

    python bench.py [--steps K] [--warmup W] [--count N] [--gpus 1]

Prints ONE JSON line.  `value` is the solves/s of the fastest precision mode
whose energies agree with the FP64 GPU path within 1e-5 eV.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "PM3 SCF solves/s + p50/p99 latency, mixed vs FP64 GPU at |ΔE|≤1e-5 eV"
EV = 1.0 / 0.03674903
FP64_PEAK_TFLOPS = 78.6       # MI355X dense FP64 (matrix and vector)
# MolDS SCF knobs for the mix: DIIS as test/ch3f_pm3.in; damping 0.7 while RMS > 0.02
# keeps the randomly oriented water clusters out of occupation-swapping cycles
SCF = dict(rms_density=1e-8, diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0,
           damping_weight=0.7, damping_thresh=0.02, max_iter=100)
CPU_SCF = dict(rms_density=1e-8, diis_k=5, diis_start=0.1, diis_end=0.0, damping_weight=0.7, damping_thresh=0.02,
               max_iter=100)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_jobs(count, precision):
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    mols = list(W.batch_mix(count))
    return mols, [Job("pm3", Z, W.to_bohr(x), precision=precision, **SCF) for _, _, Z, x in mols]


def gpu_leg(engine, jobs, steps, warmup):
    for _ in range(warmup):
        engine.scf_batch(jobs)
    engine.profile(True)
    engine.stage_times(reset=True)
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        res = engine.scf_batch(jobs)
    wall = time.perf_counter() - t0
    st = engine.stage_times(reset=True)
    engine.profile(False)
    return res, wall, st


def cpu_leg(mols, seconds):
    """The oracle (C restatement + numpy/LAPACK SCF driver) on the same molecules,
    in order, until `seconds` of CPU work: a bounded sample of the workload."""
    from oracle import scf as O
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] + [1])
    except Exception:
        threads = 1
    t0 = time.perf_counter()
    done = conv = 0
    for kind, size, Z, x in mols:
        try:
            O.scf("pm3", Z, np.asarray(x) * O.ANG2AU, **CPU_SCF)
            conv += 1
        except O.OracleError:
            pass
        done += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return dict(value=done / dt, unit="solves/s", cores=threads, kind="port",
                sample=f"first {done} molecules of the same PM3 mix ({conv} converged), solved one at a time "
                       f"(oracle: C Fock/integrals + numpy LAPACK dsyevd, {dt:.1f} s)")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--modes", default="fp64,mixed")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build targets one GPU; --gpus {args.gpus} is not supported "
              f"(the batched workload shards by molecule across GPUs, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    from molds_amd.nddo import Engine
    engine = Engine(0)
    t = time.perf_counter()
    mols, jobs64 = make_jobs(args.count, "fp64")
    log(f"built {len(mols)} molecules in {time.perf_counter() - t:.1f} s")
    nao = np.array([sum(1 if z == 1 else 4 for z in m[2]) for m in mols])

    legs = {}
    for mode in args.modes.split(","):
        jobs = jobs64 if mode == "fp64" else make_jobs(args.count, mode)[1]
        res, wall, st = gpu_leg(engine, jobs, args.steps, args.warmup)
        ok = np.array([r["status"] == 0 for r in res])
        iters = np.array([r["iterations"] for r in res])
        legs[mode] = dict(res=res, wall=wall, stages=st, ok=ok, iters=iters,
                          solves_per_s=len(jobs) * args.steps / wall, ms_per_step=1e3 * wall / args.steps)
        log(f"{mode}: {legs[mode]['solves_per_s']:.1f} solves/s, {legs[mode]['ms_per_step']:.1f} ms/step, "
            f"converged {ok.sum()}/{len(ok)}, mean iters {iters.mean():.1f}, stages {st}")

    base = legs["fp64"]
    e64 = np.array([r["energy"] for r in base["res"]])
    best_mode, best = "fp64", base
    max_de = {}
    for mode, lg in legs.items():
        if mode == "fp64":
            continue
        e = np.array([r["energy"] for r in lg["res"]])
        both = base["ok"] & lg["ok"]
        max_de[mode] = float(np.max(np.abs(e[both] - e64[both])) * EV) if both.any() else float("nan")
        if max_de[mode] <= 1e-5 and lg["solves_per_s"] > best["solves_per_s"]:
            best_mode, best = mode, lg

    # roofline of the dominant kernel (the eigensolver): algorithmic FP64 flops
    # (10/3) n^3 per eigensolve (one per SCF iteration per molecule) over the
    # measured eigen-stage time on the engine's stream
    eig_ms, eig_launches = best["stages"]["eigen"]
    flops = args.steps * float(np.sum(best["iters"] * (10.0 / 3.0) * nao.astype(float) ** 3))
    achieved = flops / (eig_ms * 1e-3) / 1e12 if eig_ms > 0 else 0.0
    lat = [best["wall"] / args.steps] * len(best["res"])

    out = {
        "metric": METRIC,
        "value": round(best["solves_per_s"], 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(best["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if best_mode == "fp64" else "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"batched PM3 SCF, {args.count} molecules: 50% all-trans alkanes C_nH_2n+2 "
                               f"n~U[1,40], 50% water clusters (H2O)_m m~U[2,64], seeded jitter, "
                               f"density RMS < 1e-8, one launch per SCF stage",
                   "molecules": args.count, "mode": best_mode},
        "fp64_solves_per_s": round(base["solves_per_s"], 3),
        "modes": {m: {"solves_per_s": round(lg["solves_per_s"], 3), "converged": int(lg["ok"].sum()),
                      "mean_iterations": round(float(lg["iters"].mean()), 2),
                      "max_abs_dE_vs_fp64_eV": max_de.get(m, 0.0),
                      "stage_ms_per_step": {k: round(v[0] / args.steps, 3) for k, v in lg["stages"].items()}}
                  for m, lg in legs.items()},
        "latency_s": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                      "note": "batched mode: every molecule retires with its batch"},
        "roofline": {"bound": "mfma", "kernel": "k_jacobi (FP64 eigensolve)", "achieved": round(achieved, 4),
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": round(achieved / FP64_PEAK_TFLOPS, 6),
                     "traffic": None},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_leg(mols, args.cpu_seconds)
    engine.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
