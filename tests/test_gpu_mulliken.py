"""Mulliken charges of the final density (the "Mulliken charge(SCF)" lines,
Cndo2.cpp:733, 1187-1199, 1733-1747) through the C ABI: against MolDS's printed
goldens (7 digits) and against the oracle's charges on the same molecule, from
the batch path (with and without P copied back) and the served path."""
import numpy as np
import pytest

from conftest import case_xyz_bohr, golden_opts, load_goldens, oracle_opts

pytestmark = pytest.mark.gpu

REL = 5.01e-7   # half a unit in the 7th printed digit (%e)


def _unique_cases():
    seen, out = set(), []
    for c in load_goldens():
        if not c["expect"]["mulliken"]:
            continue
        key = (c["theory"], tuple(c["Z"]), tuple(map(tuple, c["xyz_angstrom"])), tuple(sorted(c["scf"].items())))
        if key not in seen:
            seen.add(key)
            out.append(c)
    return out


def test_mulliken_matches_goldens(engine):
    from molds_amd.nddo import Job
    cases = _unique_cases()
    jobs = [Job(c["theory"], c["Z"], case_xyz_bohr(c), **golden_opts(c)) for c in cases]
    with_p = engine.scf_batch(jobs, want_P=True, want_mulliken=True)
    diag_only = engine.scf_batch(jobs, want_mulliken=True)
    for c, r, d in zip(cases, with_p, diag_only):
        g = np.array(c["expect"]["mulliken"])
        assert r["status"] == 0, c["name"]
        assert np.all(np.abs(r["mulliken"] - g) <= REL * np.abs(g) + 1e-9), (c["name"], r["mulliken"], g)
        # the diagonal-only copy gives the same numbers as the full density
        assert np.array_equal(r["mulliken"], d["mulliken"]), c["name"]
        assert abs(r["mulliken"].sum()) < 1e-10


@pytest.mark.parametrize("theory,tag", [("pm3", "decane"), ("am1", "water8")])
def test_mulliken_matches_oracle(engine, theory, tag):
    from molds_amd import workloads as W
    from oracle import scf as O
    if tag == "decane":
        Z, x = W.decane()
    else:
        Z, x = W.water_cluster(8, seed=5, jitter=0.02)
    Z = np.asarray(Z)
    xyz = W.to_bohr(x)
    opts = dict(rms_density=1e-8, diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0,
                damping_weight=0.0, max_iter=200)
    r = engine.scf(theory, Z, xyz, **opts)
    o = O.scf(theory, Z, xyz, rms_density=1e-8, diis_k=5, diis_start=0.1, diis_end=0.0, damping_weight=0.0,
              max_iter=200)
    assert r["status"] == 0
    assert np.max(np.abs(r["mulliken"] - O.mulliken(Z, o["P"]))) <= 1e-8


def test_served_mulliken(engine):
    from molds_amd.nddo import Job
    c = [c for c in load_goldens() if c["name"] == "ch3f_pm3"][0]
    job = Job(c["theory"], c["Z"], case_xyz_bohr(c), **golden_opts(c))
    jid = engine.submit(job)
    assert engine.wait(jid, 60.0) == 0
    r = engine.get_result(jid, job, want_mulliken=True)
    engine.release(jid)
    g = np.array(c["expect"]["mulliken"])
    assert r["status"] == 0
    assert np.all(np.abs(r["mulliken"] - g) <= REL * np.abs(g) + 1e-9)
