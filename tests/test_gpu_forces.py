"""Analytic SCF forces on the GPU (nddo_result.forces: the pair energies'
gradient at the converged density, evaluated in dual numbers) against
independent derivatives of the SCF energy: central differences of the
oracle's converged energies and of the engine's own (12 N displaced SCFs),
on MNDO, AM1 and PM3 molecules with H, C, N, O and F, a molecule whose bonds
lie along the space z axis (where the Euler-angle frame of the integrals is
singular), batched and mixed-precision jobs; a free molecule has no net force
or torque.  Forces of a density converged to RMS 1e-10 agree with the
oracle's differences to 4e-10 Hartree/bohr or better (forces up to 0.1); a
dropped term, a sign or a wrong weight misses by orders of magnitude."""
import numpy as np
import pytest

from conftest import case_xyz_bohr, load_goldens

pytestmark = pytest.mark.gpu

OPTS = dict(diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0, damping_weight=0.0,
            rms_density=1e-10, max_iter=200)


def _oracle_fd(theory, Z, xyz):
    from molds_amd.forces import stencil_forces
    from oracle import scf as O
    oo = dict(rms_density=1e-10, diis_k=5, diis_start=0.1, diis_end=0.0, damping_weight=0.0, max_iter=200)
    P = O.scf(theory, list(Z), xyz, **oo)["P"]
    return stencil_forces(lambda gs: [O.scf(theory, list(Z), g, requires_guess=False, P0=P, **oo)["energy"]
                                      for g in gs], xyz, 1e-3)


def _molecules():
    from molds_amd import workloads as W
    out = []
    for c in load_goldens():
        if c["name"] in ("c2h6_pm3", "ch3f_am1", "c2h6_mndo", "ch3f_mndo", "h2o_pm3_freq_num_proj_s1"):
            out.append((c["name"], c["theory"], np.array(c["Z"]), case_xyz_bohr(c).reshape(-1, 3)))
    ang = 1.0 / 0.5291772
    # nitrogen: methylamine-like CH3-NH2, slightly off its symmetric geometry
    mna = np.array([[0.0, 0.0, 0.0], [1.47, 0.02, -0.01], [-0.37, 1.02, 0.03], [-0.36, -0.52, 0.89],
                    [-0.38, -0.50, -0.88], [1.80, -0.93, 0.05], [1.79, 0.49, 0.83]]) * ang
    for th in ("mndo", "am1", "pm3"):
        out.append((f"ch3nh2_{th}", th, np.array([6, 7, 1, 1, 1, 1, 1]), mna))
    # bonds along the space z axis: H-C-C-H and F-H on z
    lin = np.array([[0.0, 0.0, -1.66], [0.0, 0.0, -0.6], [0.0, 0.0, 0.6], [0.0, 0.0, 1.66]]) * ang
    out.append(("c2h2_on_z", "pm3", np.array([1, 6, 6, 1]), lin))
    hf = np.array([[0.0, 0.0, 0.0], [0.0, 0.0, 0.93]]) * ang
    out.append(("hf_on_z", "am1", np.array([9, 1]), hf))
    Z, x = W.water_cluster(3, seed=5, jitter=0.02)
    out.append(("h2o3_mndo", "mndo", np.array(Z), W.to_bohr(x)))
    return out


MOLS = _molecules()


@pytest.mark.parametrize("mol", MOLS, ids=lambda m: m[0])
def test_analytic_forces_match_oracle_differences(engine, mol):
    name, theory, Z, xyz = mol
    print(f"analytic forces {name}", flush=True)
    r = engine.scf(theory, Z, xyz, want_forces=True, want_C=False, **OPTS)
    assert r["status"] == 0
    F = r["forces"]
    Fo = _oracle_fd(theory, Z, xyz)
    print(f"max |F - F_fd(oracle)| = {np.max(np.abs(F - Fo)):.2e}, max |F| = {np.max(np.abs(Fo)):.2e}", flush=True)
    assert np.max(np.abs(F - Fo)) <= 2e-9, np.max(np.abs(F - Fo))
    # free molecule: no net force (exact pair antisymmetry) and no net torque
    assert np.max(np.abs(F.sum(axis=0))) <= 1e-12 * max(1.0, np.max(np.abs(F))) * len(Z)
    assert np.max(np.abs(np.cross(xyz, F).sum(axis=0))) <= 1e-8


@pytest.mark.parametrize("mol", MOLS[:2] + MOLS[5:6] + MOLS[-3:], ids=lambda m: m[0])
def test_analytic_forces_match_gpu_stencil(engine, mol):
    from molds_amd.forces import scf_forces
    name, theory, Z, xyz = mol
    print(f"analytic vs stencil {name}", flush=True)
    a = scf_forces(engine, theory, Z, xyz, method="analytic", **OPTS)["forces"]
    s = scf_forces(engine, theory, Z, xyz, method="stencil", **OPTS)["forces"]
    assert np.max(np.abs(a - s)) <= 2e-8 * max(1.0, np.max(np.abs(s)))


def test_batched_and_mixed_forces(engine):
    """Forces of jobs batched with others equal the jobs alone (the batch's
    launch shapes change the last bits of the density, not more); a
    mixed-precision job's forces (FP64 polish to 1e-10) agree with FP64 to
    1e-7."""
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    print("batched forces", flush=True)
    jobs = []
    for k in range(6):
        Z, x = W.alkane(2 + k, seed=40 + k) if k % 2 else W.water_cluster(2 + k, seed=40 + k, jitter=0.02)
        jobs.append(Job("pm3", Z, W.to_bohr(x), **OPTS))
    res = engine.scf_batch(jobs, want_forces=True)
    for j, r in zip(jobs, res):
        alone = engine.scf_batch([j], want_forces=True)[0]
        assert r["status"] == 0 and np.max(np.abs(r["forces"] - alone["forces"])) <= 1e-10
    plain = engine.scf_batch(jobs[:3])
    assert all("forces" not in r for r in plain)
    for j, r in zip(jobs[:3], res[:3]):
        m = engine.scf_batch([Job("pm3", j.Z, j.xyz, precision="mixed", **OPTS)], want_forces=True)[0]
        assert m["status"] == 0
        assert np.max(np.abs(m["forces"] - r["forces"])) <= 1e-7


def test_noalloc_forces_equal_batch(engine):
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    Z, x = W.alkane(4, seed=9)
    j = Job("am1", Z, W.to_bohr(x), **OPTS)
    ref = engine.scf_batch([j], want_forces=True)[0]["forces"]
    r = engine.scf_once_noalloc(j, want_forces=True)
    assert r["status"] == 0 and np.array_equal(r["forces"], ref)
