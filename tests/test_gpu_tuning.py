"""The autotuner's launch shapes (pair-Fock workgroup size, back-transformation
row block) are timed at engine creation; every candidate must compute the same
numbers, so a batch run under each forced setting is compared bit for bit."""
import numpy as np
import pytest

from conftest import case_xyz_bohr, golden_opts, load_goldens

pytestmark = pytest.mark.gpu


def test_tuning_is_a_candidate(engine):
    t = engine.tuning()
    print("tuning", t, flush=True)
    assert t["fock_pairs_per_workgroup"] in (1, 2, 4, 8)
    assert set(t["backtransform_rows"].values()) <= {32, 64}


def test_launch_shapes_do_not_change_results(engine):
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    print("launch shapes vs results", flush=True)
    jobs = [Job(c["theory"], c["Z"], case_xyz_bohr(c), **golden_opts(c)) for c in load_goldens()[:4]]
    for m, seed in ((20, 51), (40, 52), (60, 53), (80, 54)):   # n = 120, 240, 360, 480: every row-block class
        Z, x = W.water_cluster(m, seed=seed, jitter=0.02)
        jobs.append(Job("pm3", Z, W.to_bohr(x), rms_density=1e-8))
    keep = engine.tuning()
    try:
        ref = None
        for wpb, rows in ((4, (64, 64, 64, 64)), (1, (32, 32, 32, 32)), (8, (64, 32, 64, 32)), (2, (32, 64, 32, 64))):
            engine.set_tuning(wpb, rows)
            res = engine.scf_batch(jobs, want_P=True)
            assert all(r["status"] == 0 for r in res)
            if ref is None:
                ref = res
                continue
            for a, b in zip(ref, res):
                assert a["iterations"] == b["iterations"]
                assert a["energy"] == b["energy"]
                assert np.array_equal(a["eps"], b["eps"])
                assert np.array_equal(a["P"], b["P"])
    finally:
        engine.set_tuning(keep["fock_pairs_per_workgroup"], [keep["backtransform_rows"][k] for k in (128, 256, 384, 512)])
