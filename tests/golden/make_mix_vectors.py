"""Writes tests/golden/mix_vectors.json: the oracle's converged PM3 SCF for a
size-stratified sample of the first 256 molecules of the batched workload
(molds_amd.workloads.batch_mix: even index an all-trans alkane C_nH_2n+2,
n ~ U{1..40}; odd index a water cluster (H2O)_m, m ~ U{2..64}), with the
workload's SCF options (density RMS < 1e-8, DIIS k = 5 / start 0.1 / end 0,
damping 0.7 above RMS 0.02).  Per molecule: total and core energy, iteration
count, the per-iteration density RMS and DIIS error, every orbital energy and
the density's diagonal.  The GPU tests compare the batched FP64 path with it.

Run from the repository root:  python tests/golden/make_mix_vectors.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from molds_amd import workloads as W  # noqa: E402
from oracle import scf as O  # noqa: E402

COUNT = 256
SAMPLE = 24
OPTS = dict(rms_density=1e-8, diis_k=5, diis_start=0.1, diis_end=0.0, damping_weight=0.7, damping_thresh=0.02,
            max_iter=100)


def sample_indices(mols):
    nao = np.array([int(sum(1 if z == 1 else 4 for z in m[2])) for m in mols])
    order = np.argsort(nao, kind="stable")
    pick = order[np.linspace(0, len(order) - 1, SAMPLE).round().astype(int)]
    return sorted(set(int(i) for i in pick))


def main():
    mols = list(W.batch_mix(COUNT))
    out = dict(count=COUNT, opts=OPTS, cases=[])
    for i in sample_indices(mols):
        kind, size, Z, x = mols[i]
        log = []
        r = O.scf("pm3", Z, W.to_bohr(x), log=lambda *a: log.append([float(a[1]), float(a[2])]), **OPTS)
        P = np.asarray(r["P"])
        out["cases"].append(dict(index=i, kind=kind, size=int(size), nao=int(r["nao"]), energy=float(r["energy"]),
                                 e_core=float(r["e_core"]), iterations=int(r["iterations"]), traj=log,
                                 eps=[float(v) for v in r["eps"]], diag_P=[float(v) for v in np.diag(P)]))
        print(f"molecule {i}: {kind} {size}, {r['nao']} AOs, {r['iterations']} iterations", flush=True)
    with open(os.path.join(ROOT, "tests", "golden", "mix_vectors.json"), "w") as f:
        json.dump(out, f)


if __name__ == "__main__":
    main()
