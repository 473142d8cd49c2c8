"""Mixed-precision SCF (FP32 two-electron contraction + split-FP32 MFMA density,
FP64 polish iterations) against the FP64 GPU path and the oracle.

Tolerances: total energy, heat of formation and orbital energies within
1e-5 eV of the FP64 path and of the oracle; density RMS within 1e-7 of the
FP64 density (both paths stop at a density RMS below 1e-8).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EV = 1.0 / 0.03674903
TOL_MIXED_EV = 1e-5
SCF = dict(rms_density=1e-8, diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0,
           damping_weight=0.0, max_iter=200)


def _geom(tag):
    from molds_amd import workloads as W
    if tag == "decane":
        Z, x = W.decane()
    elif tag.startswith("water"):
        Z, x = W.water_cluster(int(tag[5:]), seed=5, jitter=0.02)
    else:
        Z, x = W.alkane(int(tag[6:]), seed=7)
    return Z, W.to_bohr(x)


@pytest.mark.parametrize("theory,tag", [("pm3", "decane"), ("am1", "water8"), ("mndo", "water4"),
                                        ("pm3", "water24"), ("pm3", "alkane20")])
def test_mixed_matches_fp64_and_oracle(engine, theory, tag):
    print(f"mixed {theory} {tag}", flush=True)
    from oracle import scf as O
    Z, xyz = _geom(tag)
    r64 = engine.scf(theory, Z, xyz, precision="fp64", **SCF)
    rmx = engine.scf(theory, Z, xyz, precision="mixed", **SCF)
    o = O.scf(theory, Z, xyz, rms_density=1e-8, diis_k=5, diis_start=0.1, diis_end=0.0, damping_weight=0.0,
              max_iter=200)
    assert r64["status"] == 0 and rmx["status"] == 0
    assert abs(rmx["energy"] - r64["energy"]) * EV <= TOL_MIXED_EV
    assert abs(rmx["energy"] - o["energy"]) * EV <= TOL_MIXED_EV
    assert np.max(np.abs(rmx["eps"] - r64["eps"])) * EV <= TOL_MIXED_EV
    assert np.sqrt(np.mean((rmx["P"] - r64["P"]) ** 2)) <= 1e-7
    # heat of formation for every theory (AM1 / PM3 with E_isol from the one-centre parameters)
    assert abs(rmx["hof"] - r64["hof"]) * EV <= TOL_MIXED_EV
    assert abs(r64["hof"] - o["hof"]) * EV <= 1e-6
    # some iterations ran in FP32, and the FP64 polish finished the solve
    assert 1 <= rmx["fp64_iterations"] < rmx["iterations"]
    assert r64["fp64_iterations"] == r64["iterations"]


def test_mixed_alone_equals_batch(engine):
    """A mixed-precision job gives the same result alone and inside an uneven
    batch of FP64 and mixed jobs."""
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    mols = [W.alkane(5, seed=21), W.water_cluster(9, seed=22, jitter=0.02), W.alkane(12, seed=23),
            W.water_cluster(4, seed=24, jitter=0.02)]
    jobs = [Job("pm3", Z, W.to_bohr(x), precision="mixed" if i % 2 == 0 else "fp64", rms_density=1e-8)
            for i, (Z, x) in enumerate(mols)]
    batch = engine.scf_batch(jobs)
    for j, m in zip(jobs, batch):
        alone = engine.scf_batch([j])[0]
        assert alone["status"] == 0 and m["status"] == 0
        assert abs(alone["energy"] - m["energy"]) * EV <= 1e-9
        assert alone["iterations"] == m["iterations"]
        assert alone["fp64_iterations"] == m["fp64_iterations"]
