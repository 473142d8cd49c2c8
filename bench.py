#!/usr/bin/env python3
"""Benchmark of the MI355X NDDO SCF engine on its headline workload.

Metric: PM3 SCF solves/s with p50/p99 latency, mixed precision vs the FP64
GPU path, at |dE| <= 1e-5 eV.

Step (the timed unit): the batched PM3 mix -- 4096 neutral closed-shell
molecules, even index an all-trans alkane C_nH_{2n+2} with n ~ U{1..40}, odd
index a water cluster (H2O)_m with m ~ U{2..64}, seeded geometries with
N(0, 0.02 A) jitter -- solved to density RMS < 1e-8 (core-Hamiltonian guess,
DIIS k = 5 / start 0.1 / end 0, damping 0.7 above RMS 0.02) in ONE batch: one
launch per SCF stage for all molecules.  The geometry (a few MB) is uploaded
at the start of each step and every integral is rebuilt on the GPU inside the
timed region.

Beside it, in the same run:
  * served legs: Poisson arrivals at 5..2000 requests/s (2 s of arrivals,
    at least 50 and at most 400 requests per leg, sizes log-uniform on 5..600
    atoms over the alkane / water generators, retired at density RMS < 1e-7)
    through the continuous-batching request loop, mixed and FP64, with
    p50/p99 latency (submit -> retire), overall and by request size; the 5/s
    leg runs below saturation;
  * the requests of the mixed 500/s served leg as one nddo_scf_batch, the
    batched rate of the same work;
  * (H2O)_256 AM1 (1536 AOs), one molecule, FP64 and mixed precision;
    (H2O)_1000 PM3 (6000 AOs), seconds per SCF iteration;
  * analytic ground-state forces of C_20H_42 and (H2O)_64 against the
    central-difference stencil of 12 N displaced SCFs;
  * the plain FP64 GPU path: the same 4096-molecule mix solved one molecule at
    a time (nddo_scf_once), on a bounded prefix;
  * the CPU oracle on a bounded prefix of the same mix (cpu_baseline).

    python bench.py [--steps K] [--warmup W] [--count N] [--gpus 1]

Prints ONE JSON line.  `value` is the batched solves/s of the fastest
precision mode whose energies agree with the FP64 GPU path within 1e-5 eV.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

METRIC = "PM3 SCF solves/s + p50/p99 latency, mixed vs FP64 GPU at |ΔE|≤1e-5 eV"
EV = 1.0 / 0.03674903
FP64_PEAK_TFLOPS = 78.6       # MI355X dense FP64 (matrix and vector)
FP32_PEAK_TFLOPS = 157.3      # MI355X dense FP32 MFMA
HBM_PEAK_TBPS = 8.0           # MI355X HBM3E
# HBM bytes per launch of the dominant kernel on the 4096-molecule pass of a
# precision mode, from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised
# in these files
TRAFFIC_PROFILE = os.path.join(HERE, "profiles", "hbm_traffic_tridiag_blk_batched_pm3_4096_{mode}.md")


def measured_traffic(mode):
    """The `traffic_bytes_per_launch:` line of the committed counter summary
    for `mode`, or None when it is absent."""
    try:
        with open(TRAFFIC_PROFILE.format(mode=mode)) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    return int(float(line.split(":", 1)[1]))
    except OSError:
        pass
    return None
# MolDS SCF knobs for the mix: DIIS as test/ch3f_pm3.in; damping 0.7 while RMS > 0.02
# keeps the randomly oriented water clusters out of occupation-swapping cycles
SCF = dict(rms_density=1e-8, diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0,
           damping_weight=0.7, damping_thresh=0.02, max_iter=100)
SERVED_SCF = dict(SCF, rms_density=1e-7)
CPU_SCF = dict(rms_density=1e-8, diis_k=5, diis_start=0.1, diis_end=0.0, damping_weight=0.7, damping_thresh=0.02,
               max_iter=100)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def nao_of(Z):
    return int(sum(1 if z == 1 else 4 for z in Z))


def make_jobs(mols, precision, opts=SCF):
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    return [Job("pm3", Z, W.to_bohr(x), precision=precision, **opts) for _, _, Z, x in mols]


def batched_leg(engine, jobs, steps, warmup):
    """Warmup passes, then `steps` timed passes on the product path (hipGraph
    replay of every iteration after a batch's first, no profiling events),
    then one more pass with per-stage hipEvents (eager launches, the same
    kernels) for the stage times and the roofline, outside the timed region."""
    for _ in range(warmup):
        engine.scf_batch(jobs)
    engine.host_times(reset=True)
    g0 = engine.graph_stats()
    t0 = time.perf_counter()
    res = None
    for _ in range(steps):
        res = engine.scf_batch(jobs)
    wall = time.perf_counter() - t0
    g1 = engine.graph_stats()
    log(f"host phases: {engine.host_times(reset=True)}; graph replays in the timed passes: "
        f"{g1['replays'] - g0['replays']}")
    engine.profile(True)
    engine.stage_times(reset=True)
    t1 = time.perf_counter()
    prof_res = engine.scf_batch(jobs)
    prof_wall = time.perf_counter() - t1
    st = engine.stage_times(reset=True)
    engine.profile(False)
    same = all(a["iterations"] == b["iterations"] and a["energy"] == b["energy"] for a, b in zip(res, prof_res))
    return res, wall, st, dict(graph_replays=g1["replays"] - g0["replays"], profiled_pass_s=round(prof_wall, 4),
                               profiled_pass_bitwise_equal=bool(same))


SERVED_LO, SERVED_HI = 5, 600   # request sizes (atoms), log-uniform


def served_leg(engine, rate, count, precision, seed=3, hi=SERVED_HI):
    """Poisson arrivals through nddo_submit; returns throughput and latency percentiles."""
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    reqs = [(t, Job("pm3", Z, W.to_bohr(x), precision=precision, **SERVED_SCF))
            for t, _, _, Z, x in W.poisson_requests(rate, count, seed=seed, lo=SERVED_LO, hi=hi)]
    ids = []
    t0 = time.perf_counter()
    for t, job in reqs:
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        ids.append((engine.submit(job), time.perf_counter() - t0, job))
    from molds_amd.nddo import NddoError
    lat, done_at, ok, fp64_it, it, errors, nao = [], [], 0, [], [], [], []
    for jid, ts, job in ids:
        engine.wait(jid, 300.0)
        try:
            r = engine.get_result(jid, job)
        except NddoError as ex:   # counted and reported, not retried
            errors.append(str(ex))
            engine.release(jid)
            continue
        engine.release(jid)
        if r["status"] == 0:
            ok += 1
        lat.append(r["latency_s"])
        nao.append(len(r["eps"]))
        done_at.append(ts + r["latency_s"])
        it.append(r["iterations"])
        fp64_it.append(r["fp64_iterations"])
    if errors:
        log(f"served {precision} {rate:g}/s: {len(errors)} failed requests, first: {errors[0]}")
    span = max(done_at)
    lat, nao = np.array(lat), np.array(nao)
    by_size = {}
    for name, n0, n1 in (("<=137", 0, 137), ("138-512", 138, 512), (">512", 513, 1 << 30)):
        m = (nao >= n0) & (nao <= n1)
        if m.any():
            by_size[name] = dict(requests=int(m.sum()), p50_s=round(float(np.percentile(lat[m], 50)), 4))
    return dict(precision=precision, rate_per_s=rate, requests=count, converged=ok, failed=len(errors),
                first_error=errors[0] if errors else None,
                solves_per_s=round(count / span, 3), span_s=round(span, 3),
                p50_s=round(float(np.percentile(lat, 50)), 5), p99_s=round(float(np.percentile(lat, 99)), 5),
                mean_iterations=round(float(np.mean(it)), 2), mean_fp64_iterations=round(float(np.mean(fp64_it)), 2),
                p50_by_AOs=by_size,
                sizes=f"log-uniform {SERVED_LO}..{hi} atoms, alkane / water clusters",
                rms_density=SERVED_SCF["rms_density"])


def served_as_batch_leg(engine, rate, count, precision, seed=3):
    """The requests of a served leg solved as ONE nddo_scf_batch (all present
    at once, one launch per stage): the batched rate of the same work, beside
    which the served rate is read."""
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    jobs = [Job("pm3", Z, W.to_bohr(x), precision=precision, **SERVED_SCF)
            for _, _, _, Z, x in W.poisson_requests(rate, count, seed=seed, lo=SERVED_LO, hi=SERVED_HI)]
    t0 = time.perf_counter()
    res = engine.scf_batch(jobs)
    dt = time.perf_counter() - t0
    return dict(precision=precision, requests=count, solves_per_s=round(count / dt, 3), seconds=round(dt, 3),
                converged=int(sum(r["status"] == 0 for r in res)),
                note=f"the {count} requests of the served {precision} {rate:g}/s leg as one batch")


def plain_served_leg(engine, seconds, seed=3):
    """The plain FP64 GPU path on the served request mix: the requests of the
    served legs, one nddo_scf_once each, in arrival order, for `seconds`."""
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    t0 = time.perf_counter()
    done = 0
    for _, _, _, Z, x in W.poisson_requests(500.0, 100000, seed=seed, lo=SERVED_LO, hi=SERVED_HI):
        engine.scf_batch([Job("pm3", Z, W.to_bohr(x), precision="fp64", **SERVED_SCF)])
        done += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return dict(solves_per_s=round(done / dt, 3),
                sample=f"first {done} requests of the served mix ({SERVED_LO}..{SERVED_HI} atoms), one "
                       f"nddo_scf_once each, FP64 ({dt:.1f} s)")


def forces_leg(engine):
    """Ground-state forces of C_20H_42 (62 atoms, 122 AOs) and (H2O)_64 (192
    atoms, 384 AOs), PM3: seconds per force evaluation (SCF to density RMS
    1e-10 from the core-Hamiltonian guess, then the analytic forces), and the
    same by the central-difference stencil of 12 N displaced SCFs in one
    batch (warm-started), with the largest difference between the two."""
    from molds_amd import workloads as W
    from molds_amd.forces import scf_forces
    out = {}
    for name, (Z, x) in (("C20H42", W.alkane(20, seed=2)), ("(H2O)_64", W.water_cluster(64, seed=2, jitter=0.02))):
        xyz = W.to_bohr(x)
        opts = dict(diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0, damping_weight=0.7,
                    damping_thresh=0.02, max_iter=200)
        scf_forces(engine, "pm3", Z, xyz, **opts)   # warm-up
        t0 = time.perf_counter()
        a = scf_forces(engine, "pm3", Z, xyz, **opts)
        ta = time.perf_counter() - t0
        t0 = time.perf_counter()
        s = scf_forces(engine, "pm3", Z, xyz, method="stencil", **opts)
        ts = time.perf_counter() - t0
        out[name] = dict(atoms=len(Z), analytic_s=round(ta, 4), stencil_s=round(ts, 3),
                         stencil_scfs=12 * len(Z), speedup=round(ts / ta, 1),
                         max_abs_diff_hartree_per_bohr=float(np.max(np.abs(a["forces"] - s["forces"]))))
    return out


def large_leg(engine):
    """(H2O)_256, AM1, 1536 AOs, one molecule: FP64 and mixed precision, seconds
    per solve and the mixed-vs-FP64 differences."""
    from molds_amd import workloads as W
    Z, x = W.water_cluster(256, seed=1, jitter=0.0)
    xyz = W.to_bohr(x)
    opts = dict(rms_density=1e-8, diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0,
                damping_weight=0.0, max_iter=200)
    out = {"workload": "(H2O)_256 AM1, 1536 AOs, cube at 1 g/cm3, density RMS < 1e-8"}
    res = {}
    for prec in ("fp64", "mixed"):
        engine.scf("am1", Z, xyz, precision=prec, want_C=False, want_P=False, **opts)   # warm-up
        t0 = time.perf_counter()
        r = engine.scf("am1", Z, xyz, precision=prec, want_C=False, **opts)
        res[prec] = r
        out[prec] = dict(seconds=round(time.perf_counter() - t0, 4), status=r["status"], iterations=r["iterations"],
                         fp64_iterations=r["fp64_iterations"], energy_hartree=r["energy"])
    a, b = res["fp64"], res["mixed"]
    out["mixed_vs_fp64"] = dict(abs_dE_eV=float(abs(a["energy"] - b["energy"]) * EV),
                                max_abs_deps_eV=float(np.max(np.abs(a["eps"] - b["eps"])) * EV),
                                density_rms=float(np.sqrt(np.mean((a["P"] - b["P"]) ** 2))))
    return out


def water1000_leg(engine, iterations=6):
    """(H2O)_1000 PM3, 6000 AOs, one molecule (the largest single-molecule
    workload): seconds per SCF iteration in FP64 and mixed precision, stage
    times of the FP64 run, and the density-RMS trajectory of the first four
    iterations against the oracle's (tests/golden/large_vectors.json; this
    randomly packed cluster does not converge under MolDS's DIIS / damping, in
    the oracle either, so energies at convergence do not exist)."""
    from molds_amd import workloads as W
    Z, x = W.water_cluster(1000, seed=4, jitter=0.0)
    xyz = W.to_bohr(x)
    opts = dict(rms_density=1e-8, diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0,
                damping_weight=0.0, max_iter=iterations)
    try:
        with open(os.path.join(HERE, "tests", "golden", "large_vectors.json")) as f:
            ref = np.array(json.load(f)["water1000_pm3_first4"]["traj"])
    except (OSError, KeyError, ValueError):
        ref = None
    out = {"workload": f"(H2O)_1000 PM3, 6000 AOs, first {iterations} SCF iterations (stored integrals, device-wide "
                       f"Householder + divide-and-conquer eigensolver)"}
    for prec in ("fp64", "mixed"):
        if prec == "fp64":
            engine.profile(True)
            engine.stage_times(reset=True)
        t0 = time.perf_counter()
        r = engine.scf("pm3", Z, xyz, precision=prec, want_C=False, want_P=False, want_traj=True, **opts)
        dt = time.perf_counter() - t0
        leg = dict(seconds=round(dt, 3), iterations=r["iterations"], seconds_per_iteration=round(dt / max(r["iterations"], 1), 4),
                   fp64_iterations=r["fp64_iterations"])
        if prec == "fp64":
            st = engine.stage_times(reset=True)
            engine.profile(False)
            leg["stage_ms"] = {k: round(v[0], 2) for k, v in st.items() if v[0] > 0}
        if ref is not None:
            traj = np.array([[t[1], t[2]] for t in r["traj"][:len(ref)]])
            leg["max_rel_dev_first4_vs_oracle"] = float(np.max(np.abs(traj[:, 0] - ref[:, 0]) / np.abs(ref[:, 0])))
        out[prec] = leg
    return out


def headline_latency(served, rate):
    """p50/p99 (submit -> retire) of the mixed-precision served leg at `rate`."""
    leg = min(served["mixed"], key=lambda g: abs(g["rate_per_s"] - rate))
    return {"p50": leg["p50_s"], "p99": leg["p99_s"], "rate_per_s": leg["rate_per_s"],
            "source": "served mixed-precision leg, Poisson arrivals (submit -> retire)"}


def plain_leg(engine, jobs, seconds):
    """The plain FP64 GPU path: one molecule per nddo_scf_once call, in order."""
    t0 = time.perf_counter()
    done = 0
    for j in jobs:
        engine.scf_batch([j])
        done += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return dict(solves_per_s=round(done / dt, 3), sample=f"first {done} molecules of the mix, one nddo_scf_once "
                                                         f"each, FP64 ({dt:.1f} s)")


def _cpu_one(args):
    """One oracle solve on one thread (process-pool worker)."""
    Z, x = args
    from threadpoolctl import threadpool_limits
    from oracle import scf as O
    with threadpool_limits(1):
        try:
            O.scf("pm3", Z, np.asarray(x) * O.ANG2AU, **CPU_SCF)
            return 1
        except O.OracleError:
            return 0


def _host_cpu():
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(nproc=os.cpu_count(), model=model)


def cpu_leg(mols, seconds, workers):
    """The oracle (C restatement of MolDS's integrals / Fock / core repulsion,
    OpenMP over atom rows and pairs as MolDS, + numpy LAPACK dsyevd SCF
    driver) on the same molecules, in order, on the host: (a) one thread, one
    molecule after another; (b) one molecule after another with the OpenMP
    Fock and LAPACK threaded over every core, as MolDS runs (the reported
    `value`); (c) `workers` processes, one molecule per process, one thread
    each.  Each runs until `seconds` of wall time: a bounded sample of the
    workload.  Runs before the engine touches the GPU, so the pool's processes
    start from a GPU-free parent."""
    import multiprocessing as mp
    from threadpoolctl import threadpool_limits
    from oracle import scf as O
    O.build() if not os.path.exists(os.path.join(HERE, "oracle", "_build", "liboracle.so")) else None
    out = dict(host=_host_cpu(), kind="port")
    t0 = time.perf_counter()
    done = conv = 0
    with threadpool_limits(1):
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
    out["one_thread"] = dict(value=round(done / dt, 3), unit="solves/s", cores=1,
                             sample=f"first {done} molecules of the mix ({conv} converged), one thread, {dt:.1f} s")
    # MolDS's own way: molecules one after another, LAPACK threaded over all cores
    nthr = max(1, os.cpu_count() or 1)
    t0 = time.perf_counter()
    done = conv = 0
    with threadpool_limits(nthr):
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
    seq = dict(value=round(done / dt, 3), unit="solves/s", cores=nthr,
               sample=f"first {done} molecules of the mix ({conv} converged), in order, OpenMP Fock and LAPACK on "
                      f"{nthr} threads (MolDS's way: one molecule at a time, every core on it), {dt:.1f} s")
    ctx = mp.get_context("fork")
    t0 = time.perf_counter()
    done = conv = 0
    with ctx.Pool(workers) as pool:
        for ok in pool.imap(_cpu_one, [(Z, x) for _, _, Z, x in mols], chunksize=1):
            done += 1
            conv += ok
            if time.perf_counter() - t0 > seconds:
                break
        pool.terminate()
    dt = time.perf_counter() - t0
    out["molecule_parallel"] = dict(value=round(done / dt, 3), unit="solves/s", cores=workers,
                                    sample=f"first {done} molecules of the mix ({conv} converged), one molecule per "
                                           f"process on {workers} processes of one thread each, {dt:.1f} s")
    out.update(seq)
    out["sample"] += (" (oracle: C Fock / integrals + numpy LAPACK dsyevd); one_thread: the same code on one core; "
                      "molecule_parallel: one molecule per single-threaded process")
    return out


def unique_integrals(Z):
    """I = 100 C(N_hv, 2) + 10 N_hv N_H + C(N_H, 2): the packed two-centre integrals."""
    nh = int(sum(1 for z in Z if z == 1))
    nx = len(Z) - nh
    return 100 * nx * (nx - 1) // 2 + 10 * nx * nh + nh * (nh - 1) // 2


def stage_rooflines(lg, nao, nocc, nints, lds_max_n, steps_prof=1):
    """Each SCF stage of the profiled pass against the roofline that bounds it,
    with the algorithmic work per molecule-iteration of the scope table:
    Fock B = 8 I + 16 n^2 bytes (FP32 iterations of a mixed job: 4 I + 8 n^2;
    one more FP64 build per solve for the energy), eigensolve (10/3) n^3 FP64
    flops (tridiagonalization (4/3) n^3), density 2 n^2 n_occ flops."""
    st = lg["stages"]
    it = lg["iters"]
    it64 = lg["it64"]
    it32 = it - it64
    glb = nao > lds_max_n
    out = {}
    ms = st["fock"][0] / steps_prof
    b = float(np.sum(it32 * (4.0 * nints + 8.0 * nao ** 2) + (it64 + 1) * (8.0 * nints + 16.0 * nao ** 2)))
    out["fock"] = dict(bound="hbm", ms=round(ms, 3), algorithmic_GB=round(b / 1e9, 3),
                       achieved_GBps=round(b / (ms * 1e-3) / 1e9, 1), frac=round(b / (ms * 1e-3) / 8e12, 4))
    ms = st["eigen"][0] / steps_prof
    f = float(np.sum(it * (10.0 / 3.0) * nao ** 3))
    out["eigensolve"] = dict(bound="fp64", ms=round(ms, 3), algorithmic_TFLOP=round(f / 1e12, 4),
                             achieved_TFLOPs=round(f / (ms * 1e-3) / 1e12, 3),
                             frac=round(f / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4),
                             note="both eigensolver classes, which run concurrently on two streams")
    ms = st["tridiag_glb"][0] / steps_prof
    f = float(np.sum(it[glb] * (4.0 / 3.0) * nao[glb] ** 3))
    out["tridiag_device_memory"] = dict(bound="fp64", ms=round(ms, 3), algorithmic_TFLOP=round(f / 1e12, 4),
                                        achieved_TFLOPs=round(f / (ms * 1e-3) / 1e12, 3) if ms > 0 else None,
                                        frac=round(f / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4) if ms > 0 else None,
                                        molecules=int(glb.sum()))
    ms = st["tridiag_lds"][0] / steps_prof
    f = float(np.sum(it[~glb] * (10.0 / 3.0) * nao[~glb] ** 3))
    out["eigensolve_lds_class"] = dict(bound="fp64", ms=round(ms, 3), algorithmic_TFLOP=round(f / 1e12, 4),
                                       achieved_TFLOPs=round(f / (ms * 1e-3) / 1e12, 3) if ms > 0 else None,
                                       frac=round(f / (ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS, 4) if ms > 0 else None,
                                       molecules=int((~glb).sum()),
                                       note="its whole eigensolve on the second stream, beside the other class")
    ms = st["density"][0] / steps_prof
    f = float(np.sum(it * 2.0 * nao ** 2 * nocc))
    f32 = float(np.sum(it32 * 2.0 * nao ** 2 * nocc))
    peak = FP64_PEAK_TFLOPS if f32 == 0 else FP32_PEAK_TFLOPS
    out["density"] = dict(bound="mfma", ms=round(ms, 3), algorithmic_TFLOP=round(f / 1e12, 4),
                          achieved_TFLOPs=round(f / (ms * 1e-3) / 1e12, 3),
                          frac=round(f / (ms * 1e-3) / 1e12 / peak, 4), peak_TFLOPs=peak,
                          note="includes the DIIS ring write and error dot products of the same tiles")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--modes", default="fp64,mixed")
    ap.add_argument("--served-rates", default="5,50,500,2000", help="mixed-precision served legs (requests/s)")
    ap.add_argument("--served-rates-fp64", default="500", help="FP64 served legs (requests/s)")
    ap.add_argument("--served-seconds", type=float, default=2.0, help="seconds of arrivals per served leg")
    ap.add_argument("--served-max", type=int, default=400, help="at most this many requests per served leg")
    ap.add_argument("--no-large", action="store_true", help="skip the (H2O)_256 single-molecule leg")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (reference for the HBM byte "
                         "counters FETCH_SIZE / WRITE_SIZE) and exit")
    ap.add_argument("--latency-rate", type=float, default=500.0, help="served rate whose p50/p99 is the headline")
    ap.add_argument("--plain-seconds", type=float, default=8.0)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--cpu-workers", type=int, default=16, help="host processes of the CPU baseline (at most nproc)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-batched", action="store_true",
                    help="skip the batched legs (profiling runs of the other legs; the line then has no value)")
    ap.add_argument("--served-only", action="store_true",
                    help="only the mixed-precision served legs of --served-rates (profiling runs; the line has no value)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0").fill_(1.0)
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": a.numel() * 8, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported (the batched and served "
              f"workloads shard by molecule across GPUs, one engine per GPU)", file=sys.stderr)
        sys.exit(2)

    from molds_amd import workloads as W
    t = time.perf_counter()
    mols = list(W.batch_mix(args.count))
    nao = np.array([nao_of(m[2]) for m in mols], dtype=float)
    nocc = np.array([sum(1 if z == 1 else z - 2 for z in m[2]) // 2 for m in mols], dtype=float)
    nints = np.array([unique_integrals(m[2]) for m in mols], dtype=float)
    log(f"built {len(mols)} molecules in {time.perf_counter() - t:.1f} s")
    cpu = None
    if not args.no_cpu:   # before the engine touches the GPU (the pool forks this process)
        cpu = cpu_leg(mols, args.cpu_seconds, max(1, min(args.cpu_workers, os.cpu_count() or 1)))
        log(f"cpu baseline: {cpu}")

    from molds_amd.nddo import Engine
    engine = Engine(0)
    if args.served_only:   # profiling run of the served legs only
        legs = []
        for rate in [float(r) for r in args.served_rates.split(",") if r]:
            count = min(args.served_max, max(50, int(round(rate * args.served_seconds))))
            legs.append(served_leg(engine, rate, count, "mixed"))
            log(f"served mixed {rate:g}/s: {legs[-1]}")
        engine.close()
        print(json.dumps({"metric": METRIC, "value": None, "unit": "solves/s", "n_gpus": 1,
                          "note": "served legs only", "served_mixed": legs}), flush=True)
        return
    if args.no_batched:   # profiling run of the single-molecule legs only
        out = {"metric": METRIC, "value": None, "unit": "solves/s", "n_gpus": 1, "note": "single-molecule legs only",
               "large_single": large_leg(engine), "water1000_pm3": water1000_leg(engine),
               "forces_pm3": forces_leg(engine)}
        engine.close()
        print(json.dumps(out), flush=True)
        return
    legs = {}
    for mode in args.modes.split(","):
        jobs = make_jobs(mols, mode)
        res, wall, st, extra = batched_leg(engine, jobs, args.steps, args.warmup)
        ok = np.array([r["status"] == 0 for r in res])
        iters = np.array([r["iterations"] for r in res], dtype=float)
        it64 = np.array([r["fp64_iterations"] for r in res], dtype=float)
        legs[mode] = dict(res=res, wall=wall, stages=st, ok=ok, iters=iters, it64=it64, extra=extra,
                          solves_per_s=len(jobs) * args.steps / wall, ms_per_step=1e3 * wall / args.steps)
        log(f"{mode}: {legs[mode]['solves_per_s']:.1f} solves/s, {legs[mode]['ms_per_step']:.1f} ms/step, "
            f"converged {ok.sum()}/{len(ok)}, mean iters {iters.mean():.1f}, {extra}, stages {st}")
        del jobs
    mem = engine.memory()
    engine.trim()   # the served legs and the (H2O)_256 leg get the batch workspace back

    # a mode counts only if its energies agree with the FP64 leg within 1e-5 eV;
    # without an FP64 leg (profiling runs of one mode) the first mode stands
    if "fp64" in legs:
        base_mode = "fp64"
    else:
        base_mode = next(iter(legs))
        log(f"no fp64 leg: {base_mode} is reported without the 1e-5 eV check")
    base = legs[base_mode]
    e64 = np.array([r["energy"] for r in base["res"]])
    best_mode, best = base_mode, base
    max_de = {base_mode: 0.0 if base_mode == "fp64" else None}
    for mode, lg in legs.items():
        if mode == base_mode:
            continue
        e = np.array([r["energy"] for r in lg["res"]])
        both = base["ok"] & lg["ok"]
        max_de[mode] = float(np.max(np.abs(e[both] - e64[both])) * EV) if both.any() else float("nan")
        if max_de[mode] <= 1e-5 and lg["solves_per_s"] > best["solves_per_s"]:
            best_mode, best = mode, lg

    # roofline of the dominant kernel of the step: the tridiagonalization of
    # the matrices that do not fit LDS (k_tridiag_blk, blocked), timed with
    # hipEvents on the engine's stream around its launches in the profiled
    # pass (the same kernels and inputs as the timed passes); algorithmic work
    # (4/3) n^3 FP64 flops per eigensolve, one eigensolve per SCF iteration per
    # molecule. The same launches seen as HBM traffic: the upper triangle of
    # the trailing block read once per Householder step (4 n^3 / 3 bytes) and
    # read and written once per panel of 8 steps (n^3 / 3), FP64 storage in
    # both precision modes: 5 n^3 / 3 bytes, 0.8 flop per byte.
    info = engine.info()
    glb = nao > info["lds_tridiag_max_n"]
    ms_tri, launches_tri = best["stages"]["tridiag_glb"]
    flops = float(np.sum(best["iters"][glb] * (4.0 / 3.0) * nao[glb] ** 3))
    achieved = flops / (ms_tri * 1e-3) / 1e12 if ms_tri > 0 else 0.0
    alg_bytes = float(np.sum(best["iters"][glb] * (5.0 / 3.0) * nao[glb] ** 3))
    hbm_tbs = alg_bytes / (ms_tri * 1e-3) / 1e12 if ms_tri > 0 else 0.0
    traffic = measured_traffic(best_mode) if args.count == 4096 else None
    stages = {m: stage_rooflines(lg, nao, nocc, nints, info["lds_tridiag_max_n"]) for m, lg in legs.items()}

    served = {}
    if not args.no_served:
        for prec, rates in (("mixed", args.served_rates), ("fp64", args.served_rates_fp64)):
            served[prec] = []
            for rate in [float(r) for r in rates.split(",") if r]:
                count = min(args.served_max, max(50, int(round(rate * args.served_seconds))))
                leg = served_leg(engine, rate, count, prec)
                served[prec].append(leg)
                log(f"served {prec} {rate:g}/s: {leg}")
        if args.latency_rate in [float(r) for r in args.served_rates.split(",") if r]:
            count = min(args.served_max, max(50, int(round(args.latency_rate * args.served_seconds))))
            served["mixed_as_one_batch"] = served_as_batch_leg(engine, args.latency_rate, count, "mixed")
            log(f"served mix as one batch: {served['mixed_as_one_batch']}")
        engine.trim()
    large = water1000 = forces = None
    if not args.no_large:
        large = large_leg(engine)
        log(f"large: {large}")
        water1000 = water1000_leg(engine)
        log(f"(H2O)_1000: {water1000}")
        forces = forces_leg(engine)
        log(f"forces: {forces}")
        engine.trim()
    plain = plain_served = None
    if not args.no_plain:
        plain = plain_leg(engine, make_jobs(mols, "fp64"), args.plain_seconds)
        log(f"plain fp64: {plain}")
        if served:
            plain_served = plain_served_leg(engine, args.plain_seconds * 2)
            log(f"plain fp64 on the served mix: {plain_served}")

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
                               f"n~U[1,40], 50% water clusters (H2O)_m m~U[2,64], seeded N(0,0.02 A) jitter, "
                               f"density RMS < 1e-8, one launch per SCF stage",
                   "molecules": args.count, "mode": best_mode},
        "fp64_solves_per_s": round(legs["fp64"]["solves_per_s"], 3) if "fp64" in legs else None,
        "modes": {m: {"solves_per_s": round(lg["solves_per_s"], 3), "ms_per_step": round(lg["ms_per_step"], 3),
                      "converged": int(lg["ok"].sum()),
                      "mean_iterations": round(float(lg["iters"].mean()), 2),
                      "mean_fp64_iterations": round(float(lg["it64"].mean()), 2),
                      "max_abs_dE_vs_fp64_eV": max_de.get(m, 0.0),
                      "timed_path": f"hipGraph replay, {lg['extra']['graph_replays']} replays in the timed passes",
                      "profiled_pass": lg["extra"],
                      "stage_ms_profiled_pass": {k: round(v[0], 3) for k, v in lg["stages"].items()},
                      "stage_rooflines": stages[m]}
                  for m, lg in legs.items()},
        "batch_workspace_GiB": round(mem["workspace"] / 2**30, 3),
        "latency_s": (headline_latency(served, args.latency_rate) if served else None),
        "served": served or None,
        "large_single": large,
        "water1000_pm3": water1000,
        "forces_pm3": forces,
        "plain_fp64_gpu": plain,
        "plain_fp64_gpu_served_mix": plain_served,
        "speedup_vs_plain_fp64": ({"batched": round(best["solves_per_s"] / plain["solves_per_s"], 2)}
                                  if plain else None),
        # (4/3) n^3 flops over (5/3) n^3 bytes, 0.8 flop per byte, far below
        # the FP64 ridge point (78.6 TF/s / 8 TB/s ~ 10): HBM bounds it
        "roofline": {"bound": "hbm", "kernel": "k_tridiag_blk (blocked Householder tridiagonalization, device memory)",
                     "achieved": round(hbm_tbs * 1e3, 2), "peak": HBM_PEAK_TBPS * 1e3, "unit": "GB/s",
                     "frac": round(hbm_tbs / HBM_PEAK_TBPS, 4), "traffic": traffic,
                     "algorithmic_bytes_per_launch": round(alg_bytes / max(launches_tri, 1)),
                     "avg_launch_ms": round(ms_tri / max(launches_tri, 1), 4), "launches": launches_tri,
                     "timed_in": "the profiled pass after the timed passes (hipEvents on the kernel's stream)",
                     "fp64_view": {"achieved_TFLOPs": round(achieved, 5), "peak_TFLOPs": FP64_PEAK_TFLOPS,
                                   "frac": round(achieved / FP64_PEAK_TFLOPS, 6)},
                     "lds_tridiag_max_n": info["lds_tridiag_max_n"]},
    }
    if plain_served and served:
        peak = max(leg["solves_per_s"] for leg in served["mixed"])
        out["speedup_vs_plain_fp64"]["served_mixed_peak_same_mix"] = round(peak / plain_served["solves_per_s"], 2)
    if cpu is not None:
        out["cpu_baseline"] = cpu
    engine.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
