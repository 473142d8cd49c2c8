"""Writes tests/golden/large_vectors.json: oracle (C restatement + numpy LAPACK)
results for the two large single-molecule workloads, so that the GPU tests can
check them without running the oracle on the GPU host:

  * (H2O)_256, AM1, 1536 AOs (water_cluster(256, seed=1)): the converged SCF --
    total and core energy, iteration count, per-iteration density RMS and DIIS
    error, all orbital energies, and density invariants (trace, Frobenius norm,
    the diagonal);
  * (H2O)_1000, PM3, 6000 AOs (water_cluster(1000, seed=4)): the first four
    iterations' density RMS and DIIS error (this geometry does not converge
    under MolDS's DIIS / damping: an orbital swap recurs).

Run from the repository root:  python tests/golden/make_large_vectors.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from molds_amd import workloads as W  # noqa: E402
from oracle import scf as O  # noqa: E402

SCF = dict(rms_density=1e-8, diis_k=5, diis_start=0.1, diis_end=0.0, damping_weight=0.0)


def main():
    out = {}
    Z, x = W.water_cluster(256, seed=1, jitter=0.0)
    log = []
    t0 = time.perf_counter()
    r = O.scf("am1", Z, W.to_bohr(x), max_iter=200, log=lambda *a: log.append([float(a[1]), float(a[2])]), **SCF)
    P = np.asarray(r["P"])
    out["water256_am1"] = dict(theory="am1", m=256, seed=1, energy=float(r["energy"]), e_core=float(r["e_core"]),
                               iterations=int(r["iterations"]), traj=log, eps=[float(v) for v in r["eps"]],
                               trace_P=float(np.trace(P)), frob_P=float(np.linalg.norm(P)),
                               diag_P=[float(v) for v in np.diag(P)])
    print(f"(H2O)_256: {r['iterations']} iterations, {time.perf_counter() - t0:.0f} s", flush=True)
    Z, x = W.water_cluster(1000, seed=4, jitter=0.0)
    log = []
    t0 = time.perf_counter()
    try:
        O.scf("pm3", Z, W.to_bohr(x), max_iter=4, log=lambda *a: log.append([float(a[1]), float(a[2])]), **SCF)
    except O.OracleError:
        pass
    out["water1000_pm3_first4"] = dict(theory="pm3", m=1000, seed=4, traj=log)
    print(f"(H2O)_1000: {len(log)} iterations, {time.perf_counter() - t0:.0f} s", flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "large_vectors.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
