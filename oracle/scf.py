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
                    diis_k=5, diis_start=1.0e-2, diis_end=1.0e-8)

_lib = None


def build():
    """Compile the C oracle into oracle/_build (gcc only)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            build()
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
