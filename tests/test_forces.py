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


def _pair_energy_sum(theory, Z, xyz, P):
    """Sum over pairs A < B of the two-centre energy the GPU force kernel
    differentiates (resonance 2 P beta_bar S, electron-core, Coulomb P_AA P_BB
    (mu nu|lam sig), exchange -1/2 P_AB P_AB (mu nu|lam sig), core repulsion),
    from the oracle's space-frame integrals and overlap."""
    import ctypes
    from oracle import scf as O
    m = O.Molecule(theory, Z, xyz)
    g = m.integrals()
    S = m.overlap()
    el = {}
    for z in set(Z):
        v = np.zeros(33)
        O.lib().orc_elem_dump(int(z), O.THEORIES[theory], v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        el[z] = v
    nao = [1 if z == 1 else 4 for z in Z]
    off = np.concatenate(([0], np.cumsum(nao)))
    beta = np.concatenate([[el[z][5]] if z == 1 else [el[z][5]] + [el[z][6]] * 3 for z in Z])
    E = m.ecore
    for A in range(len(Z)):
        for B in range(A + 1, len(Z)):
            a, b = slice(off[A], off[A + 1]), slice(off[B], off[B + 1])
            na, nb = nao[A], nao[B]
            I = g[A, B][:na, :na, :nb, :nb]
            E += 2.0 * np.sum(P[a, b] * 0.5 * (beta[a][:, None] + beta[b][None, :]) * S[a, b])
            E += np.einsum("mn,ls,mnls->", P[a, a], P[b, b], I)
            E -= 0.5 * np.einsum("ml,ns,mnls->", P[a, b], P[a, b], I)
            E -= el[Z[B]][0] * np.sum(P[a, a] * I[:, :, 0, 0]) + el[Z[A]][0] * np.sum(P[b, b] * I[0, 0, :, :])
    return E


@pytest.mark.parametrize("theory", ["mndo", "am1", "pm3"])
def test_pair_energy_decomposition_and_frozen_density_gradient(theory):
    """The analytic forces differentiate the pair energies at fixed density.
    Pinned here on the oracle: (1) the energy at fixed P minus the pair sum is
    the same at two geometries (the remainder is one-centre only), so the pair
    terms and weights are complete; (2) the gradient of the energy at fixed P
    equals the gradient of the SCF energy (stationarity), both by central
    differences."""
    from molds_amd.forces import stencil_forces
    from oracle import scf as O
    Z = [6, 7, 1, 1, 1, 1, 1] if theory != "mndo" else [8, 1, 1]
    ang = 1.0 / 0.5291772
    x0 = (np.array([[0.0, 0.0, 0.0], [1.47, 0.02, -0.01], [-0.37, 1.02, 0.03], [-0.36, -0.52, 0.89],
                    [-0.38, -0.50, -0.88], [1.80, -0.93, 0.05], [1.79, 0.49, 0.83]]) if theory != "mndo" else
          np.array([[0.0, 0.0, 0.0], [0.96, 0.05, 0.0], [-0.25, 0.93, 0.02]])) * ang
    P = O.scf(theory, Z, x0, **_opts_oracle())["P"]

    def frozen(x):
        m = O.Molecule(theory, Z, x)
        return m.elec_energy(P) + m.ecore
    x1 = x0 + np.random.default_rng(3).normal(scale=0.05, size=x0.shape)
    d0 = frozen(x0) - _pair_energy_sum(theory, Z, x0, P)
    d1 = frozen(x1) - _pair_energy_sum(theory, Z, x1, P)
    assert abs(d0 - d1) <= 1e-10 * abs(d0), (d0, d1)
    Ff = stencil_forces(lambda gs: [frozen(x) for x in gs], x0, 1e-3)
    Fs = _oracle_forces_of(theory, Z)(x0)[1]
    assert np.max(np.abs(Ff - Fs)) <= 1e-8, np.max(np.abs(Ff - Fs))
