"""Device memory of the synchronous batch path.

nddo_scf_batch carves every batch out of one engine-owned workspace that grows
to the largest batch and is reused: repeated batches leave the device's free
memory where the first one left it.  A batch larger than the engine's budget
runs as smaller batches with the same per-job results (a job's SCF does not
depend on the batch it shares), and a job that does not fit on its own ends
ADMISSION_DENIED instead of a HIP allocation error.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EV = 1.0 / 0.03674903
SCF = dict(rms_density=1e-8, diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0,
           damping_weight=0.7, damping_thresh=0.02, max_iter=100)


def _mix(count, precision="fp64", start=0):
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    mols = list(W.batch_mix(start + count))[start:]
    return [Job("pm3", Z, W.to_bohr(x), precision=precision, **SCF) for _, _, Z, x in mols]


def test_repeated_batches_keep_free_memory(engine):
    print("repeated batches: free memory", flush=True)
    jobs = {m: _mix(384, m) for m in ("fp64", "mixed")}
    engine.scf_batch(jobs["mixed"])          # the workspace takes the larger (mixed) size
    base = engine.memory()
    assert base["workspace"] > 0
    for k in range(6):
        mode = ("fp64", "mixed")[k % 2]
        res = engine.scf_batch(jobs[mode])
        assert all(r["status"] == 0 for r in res)
        m = engine.memory()
        print(f"  call {k} {mode}: free {m['free'] / 2**30:.3f} GiB, workspace {m['workspace'] / 2**30:.3f} GiB",
              flush=True)
        assert m["workspace"] == base["workspace"]
        assert abs(m["free"] - base["free"]) <= 64 << 20


def test_budget_splits_batch_and_denies_oversized_job():
    print("batch split under a small budget", flush=True)
    from molds_amd import workloads as W
    from molds_amd.nddo import Engine, Job
    jobs = _mix(48, start=100)
    with Engine(0, flags=Engine.NO_AUTOTUNE) as full:
        ref = full.scf_batch(jobs, want_P=True)
    # 48 molecules need several hundred MB together; 96 MB holds a few at a time
    with Engine(0, hbm_budget_bytes=96 << 20, flags=Engine.NO_AUTOTUNE) as small:
        got = small.scf_batch(jobs, want_P=True)
        assert 0 < small.memory()["workspace"] <= 96 << 20
        Z, x = W.water_cluster(64, seed=3, jitter=0.02)    # 384 AOs: about 25 MB alone, and over 1 MB
        with Engine(0, hbm_budget_bytes=1 << 20, flags=Engine.NO_AUTOTUNE) as tiny:
            r = tiny.scf_batch([Job("pm3", Z, W.to_bohr(x), **SCF)])[0]
            assert r["status_name"] == "ADMISSION_DENIED"
    for a, b in zip(ref, got):
        assert a["status"] == 0 and b["status"] == 0
        assert a["iterations"] == b["iterations"]
        assert abs(a["energy"] - b["energy"]) * EV <= 1e-9
        assert np.max(np.abs(a["P"] - b["P"])) <= 1e-10


def test_trim_returns_workspace(engine):
    engine.scf_batch(_mix(64))
    assert engine.memory()["workspace"] > 0
    engine.trim()
    assert engine.memory()["workspace"] == 0
    r = engine.scf_batch(_mix(8))
    assert all(x["status"] == 0 for x in r)
