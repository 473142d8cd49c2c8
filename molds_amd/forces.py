"""SCF forces and MolDS's velocity-Verlet MD step on them.

MolDS computes the ground-state force analytically (Mndo::CalcForce,
Mndo.cpp:2711-2986, with derivative integrals).  The engine does the same on
the GPU (nddo_result.forces: the gradient of every pair's two-centre energy at
the converged density, evaluated in dual numbers), one SCF per force.  The
central-difference form is kept as an independent check: every atom
coordinate is displaced by -2h, -h, +h, +2h and all 12 N displaced SCFs run as
one batch, warm-started from the converged density of the undisplaced
geometry (DoSCF(requiresGuess=false)); the force is the fourth-order stencil
  F = -[E(-2h) - 8 E(-h) + 8 E(+h) - E(+2h)] / (12 h).
The SCF energy is stationary in the density, so a density converged to RMS
1e-10 leaves an energy error far below the stencil's rounding (about 1e-14 /
h Hartree/bohr).

* stencil_forces(energies_of, xyz, h)   the stencil over any batched energy
                                        function (list of geometries -> energies)
* scf_forces(engine, ...)               GPU forces: analytic (default) or the
                                        stencil of 12 N displaced SCFs
* md_step(...)                          one step of MD::DoMD (md/MD.cpp:99-133):
                                        p += dt/2 F; x += dt p / m_core; SCF and
                                        forces at x; p += dt/2 F
"""
from __future__ import annotations

import numpy as np

FS2AU = 1.0 / 2.418884326505e-2              # Parameters.cpp:52
GMOL2AU = 1.0e5 / (6.0221415 * 9.1095)       # Parameters.cpp:53
ATOMIC_MASS = {1: 1.00794, 6: 12.0107, 7: 14.00674, 8: 15.9994, 9: 18.9984032}   # atoms/*atom.cpp


def core_masses(Z):
    """Atomic mass minus the valence electrons' mass, a.u. (Atom::GetCoreMass, Atom.h:34)."""
    return np.array([ATOMIC_MASS[int(z)] * GMOL2AU - (1.0 if z == 1 else float(z - 2)) for z in Z])


def stencil_forces(energies_of, xyz, h=1e-3):
    """Forces (N x 3, Hartree/bohr) from energies_of(list of N x 3 geometries)."""
    xyz = np.asarray(xyz, dtype=np.float64)
    geoms = []
    for a in range(xyz.shape[0]):
        for k in range(3):
            for s in (-2.0, -1.0, 1.0, 2.0):
                g = xyz.copy()
                g[a, k] += s * h
                geoms.append(g)
    E = np.asarray(energies_of(geoms), dtype=np.float64).reshape(xyz.shape[0], 3, 4)
    return -(E[..., 0] - 8.0 * E[..., 1] + 8.0 * E[..., 2] - E[..., 3]) / (12.0 * h)


def scf_forces(engine, theory, Z, xyz_bohr, P0=None, h=1e-3, precision="fp64", method="analytic", **opts):
    """SCF at xyz_bohr (warm from P0 when given) to density RMS <= 1e-10, and
    its forces: analytic from the converged density (method "analytic"), or
    from 12 N displaced SCFs in one batch (method "stencil").
    Returns dict(energy, forces, P, result)."""
    from molds_amd.nddo import Job
    o = dict(opts)
    o["rms_density"] = min(o.get("rms_density", 1e-8), 1e-10)
    o.setdefault("max_iter", 200)
    center = engine.scf(theory, Z, xyz_bohr, precision=precision, requires_guess=P0 is None, P0=P0,
                        want_C=False, want_P=True, want_traj=False, want_forces=method == "analytic", **o)
    if center["status"] != 0:
        raise RuntimeError(f"SCF at the undisplaced geometry ended with {center['status_name']}")
    Pc = center["P"]
    if method == "analytic":
        return dict(energy=center["energy"], forces=center["forces"], P=Pc, result=center)
    if method != "stencil":
        raise ValueError(f"unknown force method {method!r}")

    def energies(geoms):
        res = engine.scf_batch([Job(theory, Z, g, precision, False, Pc, **o) for g in geoms])
        bad = [r["status_name"] for r in res if r["status"] != 0]
        if bad:
            raise RuntimeError(f"{len(bad)} displaced SCFs did not converge ({bad[0]})")
        return [r["energy"] for r in res]

    F = stencil_forces(energies, xyz_bohr, h)
    return dict(energy=center["energy"], forces=F, P=Pc, result=center)


def md_step(forces_of, Z, xyz, p, F, dt_au):
    """One velocity-Verlet step as MD::DoMD: returns (xyz, p, F, energy) after it.
    forces_of(xyz) -> (energy, forces)."""
    m = core_masses(Z)[:, None]
    p = p + 0.5 * dt_au * F
    xyz = xyz + dt_au * p / m
    E, F = forces_of(xyz)
    p = p + 0.5 * dt_au * F
    return xyz, p, F, E
