"""SCF forces by central differences of batched SCF energies, and one MD step
(molds_amd/forces.py), against MolDS's ground-state MD goldens: the atom
momenta and coordinates printed after MD step 1 of test/c2h6_{am1,mndo,pm3}_force
(velocity Verlet from rest, md/MD.cpp:87-133; fixtures written by
tests/golden/make_md_goldens.py).

MolDS converges those SCFs to density RMS 1e-6 only, and its analytic forces
(Mndo::CalcForce) carry an error linear in that residual: the printed momenta
agree with forces of a tightly converged SCF to about 1e-5 of their largest
component. The bar is 2e-5 of the largest component for momenta; a dropped
term, a sign, a wrong mass or time unit misses it by orders of magnitude.
The displacement after the step is held to the same bar."""
import json
import os

import numpy as np
import pytest

from conftest import ROOT

GOLD = os.path.join(ROOT, "tests", "golden", "molds_md_goldens.json")
ANG2AU = 1.0 / 0.5291772


def _cases():
    with open(GOLD) as f:
        return json.load(f)["cases"]


def _opts_oracle():
    return dict(rms_density=1e-10, diis_k=5, diis_start=0.1, diis_end=0.0, damping_weight=0.0, max_iter=200)


def _oracle_forces_of(theory, Z):
    from molds_amd import forces as FM
    from oracle import scf as O

    def forces_of(xyz):
        r = O.scf(theory, Z, xyz, **_opts_oracle())
        F = FM.stencil_forces(
            lambda gs: [O.scf(theory, Z, g, requires_guess=False, P0=r["P"], **_opts_oracle())["energy"] for g in gs],
            xyz, 1e-3)
        return r["energy"], F
    return forces_of


def _check_step(c, x0, x1, p1):
    g = np.array(c["step1_momenta"])
    assert np.max(np.abs(p1 - g)) <= 2e-5 * np.max(np.abs(g)), np.max(np.abs(p1 - g))
    # the displacement dt^2 F / 2m to the same bar, plus the coordinates' print rounding
    gx = np.array(c["step1_xyz_bohr"])
    dg = gx - x0
    assert np.all(np.abs((x1 - x0) - dg) <= 2e-5 * np.max(np.abs(dg)) + 5.01e-7 * np.abs(gx))


def test_stencil_on_a_closed_form():
    """The stencil's sign, coefficients and displacement pattern on an energy
    with known forces: E = sum_a k_a |x_a - c|^2 / 2 + sum_a<b 1 / |x_a - x_b|."""
    from molds_amd.forces import stencil_forces
    rng = np.random.default_rng(0)
    x = rng.normal(size=(4, 3)) * 2.0
    k = np.array([0.3, 1.1, 0.7, 2.0])
    c = np.array([0.1, -0.2, 0.3])

    def E(g):
        e = 0.5 * np.sum(k[:, None] * (g - c) ** 2)
        for a in range(len(g)):
            for b in range(a):
                e += 1.0 / np.linalg.norm(g[a] - g[b])
        return e
    F = stencil_forces(lambda gs: [E(g) for g in gs], x, 1e-3)
    ref = -k[:, None] * (x - c)
    for a in range(len(x)):
        for b in range(len(x)):
            if a != b:
                d = x[a] - x[b]
                ref[a] += d / np.linalg.norm(d) ** 3
    assert np.max(np.abs(F - ref)) < 1e-9


@pytest.mark.parametrize("case", _cases(), ids=lambda c: c["name"])
def test_oracle_md_step_matches_molds(case):
    from molds_amd import forces as FM
    Z = case["Z"]
    x0 = np.array(case["xyz_angstrom"]) * ANG2AU
    forces_of = _oracle_forces_of(case["theory"], Z)
    E0, F0 = forces_of(x0)
    # a free molecule: no net force or torque
    assert np.max(np.abs(F0.sum(axis=0))) < 1e-9
    assert np.max(np.abs(np.cross(x0, F0).sum(axis=0))) < 1e-8
    x1, p1, F1, E1 = FM.md_step(forces_of, Z, x0, np.zeros_like(x0), F0, case["dt_fs"] * FM.FS2AU)
    _check_step(case, x0, x1, p1)
    assert abs(E1 - case["step1_energy"]) <= 5.01e-7 * abs(case["step1_energy"])


@pytest.mark.gpu
@pytest.mark.parametrize("case", _cases(), ids=lambda c: c["name"])
def test_gpu_forces_and_md_step(engine, case):
    from molds_amd import forces as FM
    print(f"forces {case['name']}", flush=True)
    Z = np.array(case["Z"])
    th = case["theory"]
    x0 = np.array(case["xyz_angstrom"]) * ANG2AU
    opts = dict(diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0, damping_weight=0.0)

    def forces_of(xyz):
        r = FM.scf_forces(engine, th, Z, xyz, **opts)
        return r["energy"], r["forces"]
    E0, F0 = forces_of(x0)
    E0o, F0o = _oracle_forces_of(th, list(Z))(x0)
    assert np.max(np.abs(F0 - F0o)) <= 1e-8
    x1, p1, F1, E1 = FM.md_step(forces_of, Z, x0, np.zeros_like(x0), F0, case["dt_fs"] * FM.FS2AU)
    _check_step(case, x0, x1, p1)
