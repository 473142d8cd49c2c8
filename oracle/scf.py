"""CPU oracle for the NDDO SCF path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  The product
(``molds_amd``) never imports it.

The per-geometry pieces (overlap, two-centre integrals, core repulsion, Fock
build, electronic energy) are the C restatement in ``nddo_oracle.c``; this file
restates the SCF driver of MolDS around them:

* ``Cndo2::DoSCF``                    src/cndo/Cndo2.cpp:585-729
* ``Lapack::Dsyevd`` (+ phase rule)   src/wrappers/Lapack.cpp:102-161
  (numpy.linalg.eigh is LAPACK ?syevd, the same routine MolDS calls)
* ``CalcOrbitalElectronPopulation``   src/cndo/Cndo2.cpp:1702-1731
* ``SatisfyConvergenceCriterion``     src/cndo/Cndo2.cpp:1472-1517
* ``DoDIIS`` / ``DoDamp``             src/cndo/Cndo2.cpp:920-1061
* ``CalcSCFProperties``               src/cndo/Cndo2.cpp:732-754, 1324-1416
* ``Mndo::CalcHeatsFormation``        src/mndo/Mndo.cpp:395-404
* dipole moments (STO-6G)             src/cndo/Cndo2.cpp:748-753, 1871-2131
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "_build", "liboracle.so")
THEORIES = {"mndo": 0, "am1": 1, "pm3": 2}
EV2AU = 0.03674903
ANG2AU = 1.0 / 0.5291772
KCAL2AU = 0.00159360175

# SCF defaults, src/base/Parameters.cpp:132-144
DEFAULT_OPTS = dict(max_iter=100, rms_density=1.0e-8, damping_thresh=1.0, damping_weight=0.8,
                    diis_k=5, diis_start=1.0e-2, diis_end=1.0e-8, dipole=False)

_lib = None


def build():
    """Compile the C oracle into oracle/_build (gcc only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
        # the C oracle's OpenMP threads and numpy's BLAS threads alternate
        # (Fock, then eigensolve): threads that spin while idle would take the
        # cores from the other pool, so they sleep instead (read when the
        # OpenMP runtime loads, with the library below)
        os.environ.setdefault("OMP_WAIT_POLICY", "PASSIVE")
        L = ctypes.CDLL(LIBPATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ip, dp, ip]
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        for f in ("orc_nao", "orc_nelec"):
            getattr(L, f).restype = ctypes.c_int
            getattr(L, f).argtypes = [ctypes.c_void_p]
        for f in ("orc_ecore", "orc_hof_offset"):
            getattr(L, f).restype = ctypes.c_double
            getattr(L, f).argtypes = [ctypes.c_void_p]
        L.orc_eisol_ev.restype = ctypes.c_double
        L.orc_eisol_ev.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_get_overlap.argtypes = [ctypes.c_void_p, dp]
        L.orc_get_integrals.argtypes = [ctypes.c_void_p, dp]
        L.orc_fock.argtypes = [ctypes.c_void_p, dp, ctypes.c_int, dp]
        L.orc_elec_energy.restype = ctypes.c_double
        L.orc_elec_energy.argtypes = [ctypes.c_void_p, dp]
        L.orc_param.restype = ctypes.c_int
        L.orc_param.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        L.orc_multipole_probe.restype = ctypes.c_double
        L.orc_multipole_probe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          ctypes.c_int, ctypes.c_double]
        L.orc_local_integral_probe.restype = ctypes.c_double
        L.orc_local_integral_probe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int] + \
            [ctypes.c_int] * 4 + [ctypes.c_double]
        L.orc_overlap_local.restype = ctypes.c_double
        L.orc_overlap_local.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                        ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double]
        L.orc_aux_A.restype = ctypes.c_double
        L.orc_aux_A.argtypes = [ctypes.c_int, ctypes.c_double]
        L.orc_aux_B.restype = ctypes.c_double
        L.orc_aux_B.argtypes = [ctypes.c_int, ctypes.c_double]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleError(RuntimeError):
    pass


class Molecule:
    """Per-geometry state: S, integrals, E_core (what DoSCF builds before its loop)."""

    def __init__(self, theory, Z, xyz_bohr):
        L = lib()
        self.theory = theory
        self.Z = np.ascontiguousarray(Z, dtype=np.int32)
        self.xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        err = ctypes.c_int(0)
        h = L.orc_create(THEORIES[theory], len(self.Z), self.Z.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                         _dp(self.xyz), ctypes.byref(err))
        if not h:
            raise OracleError({1: "unsupported element", 2: "odd number of valence electrons"}[err.value])
        self.h = h
        self.nao = L.orc_nao(h)
        self.nelec = L.orc_nelec(h)
        self.ecore = L.orc_ecore(h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def overlap(self):
        S = np.zeros((self.nao, self.nao))
        lib().orc_get_overlap(self.h, _dp(S))
        return S

    def integrals(self):
        N = len(self.Z)
        g = np.zeros((N, N, 4, 4, 4, 4))
        lib().orc_get_integrals(self.h, _dp(g))
        return g

    def fock(self, P, guess=False):
        P = np.ascontiguousarray(P, dtype=np.float64)
        F = np.zeros((self.nao, self.nao))
        lib().orc_fock(self.h, _dp(P), 1 if guess else 0, _dp(F))
        return F

    def elec_energy(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        return lib().orc_elec_energy(self.h, _dp(P))

    def hof_offset(self):
        return lib().orc_hof_offset(self.h)


# STO-6G expansions of unit-exponent Slater orbitals (Stewart, J. Chem. Phys.
# 52, 431 (1970); the table MolDS keeps in src/base/GTOExpansionSTO.cpp):
# (exponents, coefficients of normalized primitive Gaussians) for 1s, 2s, 2p.
STO6G = {
    "1s": ((23.10303149, 4.235915534, 1.185056519, 0.4070988982, 0.1580884151, 0.06510953954),
           (0.00916359628, 0.04936149294, 0.1685383049, 0.3705627997, 0.4164915298, 0.1303340841)),
    "2s": ((27.68496241, 5.077140627, 1.42678605, 0.2040335729, 0.09260298399, 0.04416183978),
           (-0.004151277819, -0.02067024148, -0.05150303337, 0.3346271174, 0.5621061301, 0.1712994697)),
    "2p": ((5.868285913, 1.530329631, 0.5475665231, 0.2288932733, 0.1046655969, 0.04948220127),
           (0.007924233646, 0.05144104825, 0.189840006, 0.4049863191, 0.4012362861, 0.1051855189)),
}
# standard atomic weights (g/mol) as the atom classes set them (Hatom.cpp:52, Catom.cpp:49, ...)
ATOMIC_MASS = {1: 1.00794, 6: 12.0107, 7: 14.00674, 8: 15.9994, 9: 18.9984032}
GMOL2AU = 1.0e5 / (6.0221415 * 9.1095)   # Parameters.cpp:52
DEBYE2AU = 0.393430191                   # Parameters.cpp:57


def _ov1d(a, b, A, B):
    """1-D overlaps of (x-A)^i e^{-a(x-A)^2} and (x-B)^j e^{-b(x-B)^2} for
    i <= 2, j <= 1 (Obara-Saika recurrence); a, b arrays of primitive pairs."""
    p = a + b
    P = (a * A + b * B) / p
    e00 = np.sqrt(np.pi / p) * np.exp(-a * b / p * (A - B) ** 2)
    xa, xb, h = P - A, P - B, 0.5 / p
    e10 = xa * e00
    e01 = xb * e00
    e11 = xb * e10 + h * e00
    e20 = xa * e10 + h * e00
    e21 = xb * e20 + 2.0 * h * e10
    return {(0, 0): e00, (1, 0): e10, (0, 1): e01, (1, 1): e11, (2, 0): e20, (2, 1): e21}


def cartesian_matrix(theory, Z, xyz_bohr):
    """<mu| r_c |nu> (c = x, y, z; absolute coordinates) between the valence
    Slater AOs expanded in STO-6G (Cndo2::CalcCartesianMatrixByGTOExpansion,
    Cndo2.cpp:1946-2131): AO order s, py, pz, px; normalized primitives;
    exponents zeta^2 * alpha_i with the theory's Slater exponents."""
    Z = list(Z)
    xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
    prm = np.zeros(14)
    aos = []   # (atom, shell key, cartesian direction or -1, zeta)
    for a, z in enumerate(Z):
        lib().orc_param(int(z), THEORIES[theory], _dp(prm))
        zs, zp = prm[6], prm[7]
        if z == 1:
            aos.append((a, "1s", -1, zs))
        else:
            aos += [(a, "2s", -1, zs), (a, "2p", 1, zp), (a, "2p", 2, zp), (a, "2p", 0, zp)]
    n = len(aos)
    X = np.zeros((3, n, n))
    for mu, (A, ka, da, za) in enumerate(aos):
        ea, ca = STO6G[ka]
        ea = za * za * np.asarray(ea)
        na = (2 * ea / np.pi) ** 0.75 * (np.sqrt(4 * ea) if da >= 0 else 1.0) * np.asarray(ca)
        for nu, (B, kb, db, zb) in enumerate(aos):
            eb, cb = STO6G[kb]
            eb = zb * zb * np.asarray(eb)
            nb = (2 * eb / np.pi) ** 0.75 * (np.sqrt(4 * eb) if db >= 0 else 1.0) * np.asarray(cb)
            a2, b2 = np.meshgrid(ea, eb, indexing="ij")
            w = np.outer(na, nb)
            E = [_ov1d(a2, b2, xyz[A, c], xyz[B, c]) for c in range(3)]
            la = [1 if c == da else 0 for c in range(3)]
            lb = [1 if c == db else 0 for c in range(3)]
            ov = [E[c][(la[c], lb[c])] for c in range(3)]
            for c in range(3):
                raised = E[c][(la[c] + 1, lb[c])] * ov[(c + 1) % 3] * ov[(c + 2) % 3]
                X[c, mu, nu] = np.sum(w * (raised + xyz[A, c] * ov[0] * ov[1] * ov[2]))
    return X


def dipole(theory, Z, xyz_bohr, P, S):
    """Ground-state dipole moments in a.u. as MolDS computes them
    (Cndo2.cpp:748-753, 1871-1934): about the centre of core masses
    (Molecule::CalcXyzCOC, core mass = atomic mass - valence electrons),
    electronic = -sum P_mu,nu <mu|r|nu> + centre * sum P_mu,nu S_mu,nu,
    core = sum_A Z_A (R_A - centre); returns (total, electronic, core)."""
    xyz = np.asarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
    core = np.array([1.0 if z == 1 else float(z - 2) for z in Z])
    mass = np.array([ATOMIC_MASS[int(z)] * GMOL2AU for z in Z]) - core
    centre = (mass[:, None] * xyz).sum(axis=0) / mass.sum()
    X = cartesian_matrix(theory, Z, xyz)
    el = -np.einsum("ij,cij->c", P, X) + centre * np.sum(P * S)
    cd = (core[:, None] * (xyz - centre)).sum(axis=0)
    return el + cd, el, cd


def dsyevd_rows(F):
    """Lapack::Dsyevd: eigenvalues ascending, row i = eigenvector i, and each
    row's sign flipped so that its element sum is >= 0 (Lapack.cpp:117-150)."""
    eps, V = np.linalg.eigh(F)
    C = np.ascontiguousarray(V.T)
    neg = C.sum(axis=1) < 0
    C[neg] *= -1.0
    return eps, C


def mulliken(Z, P):
    """Mulliken charges of a density: core charge minus the atom's sum of P_mu,mu
    (Cndo2::CalcAtomicElectronPopulation, Cndo2.cpp:1733-1747; printed as
    core charge - population, Cndo2.cpp:1187-1199). Core charges are the valence
    electron counts (H 1, C 4, N 5, O 6, F 7); AOs per atom: H 1, heavy 4."""
    q = np.zeros(len(Z))
    d = np.diag(P)
    mu = 0
    for a, z in enumerate(Z):
        na = 1 if z == 1 else 4
        q[a] = (1.0 if z == 1 else float(z - 2)) - d[mu:mu + na].sum()
        mu += na
    return q


def scf(theory, Z, xyz_bohr, requires_guess=True, P0=None, log=None, **opts):
    """Restatement of Cndo2::DoSCF for MNDO/AM1/PM3 closed-shell RHF.

    Returns a dict with the same quantities MolDS prints: MO energies, the
    electronic energy including core repulsion, core repulsion, heat of
    formation (MNDO only), plus P, C (rows = MOs) and the iteration trajectory
    [(iter, rms, diis_error, diis_on, damping_on)].
    Raises OracleError('not converged') like Cndo2.cpp:700-704."""
    o = dict(DEFAULT_OPTS)
    o.update(opts)
    t0 = time.perf_counter()
    mol = Molecule(theory, Z, xyz_bohr)
    n = mol.nao
    nocc = mol.nelec // 2
    k = int(o["diis_k"])
    P = np.zeros((n, n)) if P0 is None else np.array(P0, dtype=np.float64)
    storedP = np.zeros((max(k, 1), n, n))
    storedE = np.zeros((max(k, 1), n, n))
    B = np.zeros((k + 1, k + 1))
    diis_error = 0.0
    diis_on = False
    damp_on = False
    traj = []
    res = None
    for it in range(int(o["max_iter"])):
        Pold = P.copy()
        guess = (it == 0 and requires_guess)
        F = mol.fock(P, guess)
        eps, C = dsyevd_rows(F)
        Cocc = C[:nocc]
        P = 2.0 * (Cocc.T @ Cocc)
        P = 0.5 * (P + P.T)  # dsyrk fills one triangle and mirrors it
        rms = float(np.sqrt(np.sum((Pold - P) ** 2) / (n * n)))
        traj.append((it, rms, diis_error, diis_on, damp_on))
        if log is not None:
            log(it, rms, diis_error, diis_on, damp_on)
        if rms < o["rms_density"]:
            e_el = mol.elec_energy(P)
            res = dict(eps=eps, C=C, P=P, F_last=F, e_el=e_el, e_core=mol.ecore,
                       energy=e_el + mol.ecore, iterations=it + 1, rms=rms, traj=traj,
                       nao=n, nocc=nocc)
            if o.get("dipole"):
                res["dipole_total"], res["dipole_electronic"], res["dipole_core"] = dipole(
                    theory, Z, xyz_bohr, P, mol.overlap())
            res["hof"] = res["energy"] + mol.hof_offset()
            break
        if not guess:
            if k > 0:
                # DoDIIS, Cndo2.cpp:920-1028
                storedP[:-1] = storedP[1:].copy()
                storedE[:-1] = storedE[1:].copy()
                storedP[k - 1] = P
                storedE[k - 1] = P - Pold
                B[: k - 1, : k - 1] = B[1:k, 1:k].copy()
                B[k - 1, :k] = storedE[:k].reshape(k, -1) @ storedE[k - 1].ravel()
                B[:k, k - 1] = B[k - 1, :k]
                B[:k, k] = -1.0
                B[k, :k] = -1.0
                B[k, k] = 0.0
                rhs = np.zeros(k + 1)
                rhs[k] = -1.0
                diis_error = float(np.max(np.abs(storedE[k - 1])))
                diis_on = False
                if k <= it and o["diis_end"] < diis_error < o["diis_start"]:
                    diis_on = True
                    try:
                        c = np.linalg.solve(B, rhs)
                    except np.linalg.LinAlgError:
                        diis_on = False
                        c = None
                    if c is not None:
                        P = np.tensordot(c[:k], storedP[:k], axes=1)
            # DoDamp, Cndo2.cpp:1030-1061
            damp_on = False
            w = o["damping_weight"]
            if 0.0 < w and o["damping_thresh"] < rms:
                damp_on = True
                P = P * (1.0 - w) + w * Pold
        if it == int(o["max_iter"]) - 1:
            raise OracleError("not converged")
    res["seconds"] = time.perf_counter() - t0
    return res
