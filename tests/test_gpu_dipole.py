"""Ground-state dipole moments on the GPU (k_dipole: STO-6G position integrals,
centre of core masses, as MolDS prints them, Cndo2.cpp:748-753, 1871-2131).

Against MolDS's printed total / electronic / core dipoles for every golden
case (7 digits), against the oracle on decane (PM3), (H2O)_8 (AM1) and (H2O)_4
(MNDO) to 1e-9 a.u., and the no-alloc entry point against the batch path.
"""
import numpy as np
import pytest

from conftest import case_xyz_bohr, golden_opts, load_goldens

pytestmark = pytest.mark.gpu

REL = 5.01e-7


def _close(v, e):
    v, e = np.asarray(v), np.asarray(e)
    return np.all(np.abs(v - e) <= REL * np.abs(e) + 5.01e-7 * max(np.max(np.abs(e)), 1e-3) + 1e-12)


def test_dipole_goldens_batched(engine):
    from molds_amd.nddo import Job
    cases = load_goldens()
    jobs = [Job(c["theory"], c["Z"], case_xyz_bohr(c), **golden_opts(c)) for c in cases]
    res = engine.scf_batch(jobs, want_dipole=True)
    for c, r in zip(cases, res):
        assert r["status"] == 0, c["name"]
        for k in ("dipole_total", "dipole_electronic", "dipole_core"):
            assert _close(r[k], c["expect"][k]), (c["name"], k, r[k], c["expect"][k])


@pytest.mark.parametrize("theory,tag", [("pm3", "decane"), ("am1", "water8"), ("mndo", "water4")])
def test_dipole_vs_oracle(engine, theory, tag):
    from molds_amd import workloads as W
    from oracle import scf as O
    if tag == "decane":
        Z, x = W.decane()
    else:
        Z, x = W.water_cluster(int(tag[5:]), seed=5, jitter=0.02)
    xyz = W.to_bohr(x)
    r = engine.scf(theory, Z, xyz, want_dipole=True, rms_density=1e-8, diis_num_error_vect=5,
                   diis_start_error=0.1, diis_end_error=0.0, damping_weight=0.0, max_iter=200)
    o = O.scf(theory, Z, xyz, dipole=True, rms_density=1e-8, diis_k=5, diis_start=0.1, diis_end=0.0,
              damping_weight=0.0, max_iter=200)
    assert r["status"] == 0
    for k in ("dipole_total", "dipole_electronic", "dipole_core"):
        assert np.max(np.abs(r[k] - o[k])) <= 1e-9, (k, r[k], o[k])


def test_dipole_noalloc_equals_batch(engine):
    from molds_amd import workloads as W
    from molds_amd.nddo import Job, Result, _dp
    import ctypes
    Z, x = W.alkane(6, seed=9)
    job = Job("pm3", Z, W.to_bohr(x), rms_density=1e-8)
    ref = engine.scf_batch([job], want_dipole=True)[0]
    r = Result()
    d = np.full(9, np.nan)
    r.dipole = d.ctypes.data_as(_dp)
    assert engine.L.nddo_scf_once_noalloc(engine.h, ctypes.byref(job.desc), ctypes.byref(r)) == 0
    assert np.allclose(d, np.concatenate([ref["dipole_total"], ref["dipole_electronic"], ref["dipole_core"]]),
                       rtol=0, atol=1e-12)
