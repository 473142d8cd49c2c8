"""nddo_scf_once_noalloc: one SCF on the library's fixed device scratch (no
device allocation), checked against nddo_scf_once -- the same kernels on
buffers of their own -- on two calls in a row, the second with a smaller
molecule, so whatever the first call left in the scratch cannot leak into the
second.  Tolerances are the FP64 parity bar (1e-6 eV, density RMS 1e-8); the
paths launch the same kernels in the same order, so they agree far closer."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EV = 1.0 / 0.03674903
SCF = dict(rms_density=1e-8, diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0,
           damping_weight=0.7, damping_thresh=0.02, max_iter=100)


def _job(kind, size, seed, precision="fp64"):
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    Z, x = W.alkane(size, seed=seed) if kind == "alkane" else W.water_cluster(size, seed=seed, jitter=0.02)
    return Job("pm3", Z, W.to_bohr(x), precision=precision, **SCF)


def _check(a, b):
    assert a["status"] == 0 and b["status"] == 0, (a["status_name"], b["status_name"])
    assert a["iterations"] == b["iterations"]
    assert abs(a["energy"] - b["energy"]) * EV <= 1e-6
    assert np.sqrt(np.mean((a["P"] - b["P"]) ** 2)) <= 1e-8
    assert np.max(np.abs(a["eps"] - b["eps"])) <= 1e-9
    assert np.max(np.abs(np.abs(a["C"]) - np.abs(b["C"]))) <= 1e-8
    assert np.allclose(a["mulliken"], b["mulliken"], atol=1e-9)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_scf_once_twice(engine, precision):
    from molds_amd.nddo import Job  # noqa: F401
    # first a 384-AO water cluster (device-memory eigensolver class), then a
    # 62-AO alkane (LDS class): the second call reuses a dirtier, larger scratch
    for kind, size, seed in (("water", 64, 501), ("alkane", 10, 502)):
        print(f"noalloc {precision} {kind} {size}", flush=True)
        job = _job(kind, size, seed, precision)
        got = engine.scf_once_noalloc(job)
        ref = engine.scf_batch([job], want_C=True, want_P=True, want_mulliken=True)[0]
        _check(got, ref)


def test_noalloc_rejects_oversized(engine):
    from molds_amd.nddo import NddoError, load_library
    limit = load_library().nddo_noalloc_scratch_bytes()
    assert limit == 64 << 20
    job = _job("water", 180, 503)          # 1080 AOs: far more than 64 MiB of state
    with pytest.raises(NddoError) as ex:
        engine.scf_once_noalloc(job)
    assert "ADMISSION_DENIED" in str(ex.value) and str(limit) in str(ex.value)
    # the scratch still serves a job that fits
    job = _job("alkane", 4, 504)
    _check(engine.scf_once_noalloc(job), engine.scf_batch([job], want_C=True, want_P=True, want_mulliken=True)[0])
