"""The two tridiagonalizations of the 138-512 AO class (matrices in device
memory): blocked (panels of 8 reflectors, rank-16 FP64 MFMA update per panel,
the default) and per-reflector (one rank-2 update per step).

Both replace the Householder reduction inside Lapack::Dsyevd
(src/wrappers/Lapack.cpp:102-161). Checked against LAPACK (numpy eigvalsh) on
the eigenvalues (1e-13 of ||A||), the residual and orthonormality (1e-12), the
phase rule, and through whole SCFs against the oracle (FP64 energy within
1e-6 eV, mixed within 1e-5 eV) and against each other.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EV = 1.0 / 0.03674903


@pytest.fixture
def tri(engine):
    yield engine
    engine.set_tridiag(True)


def _check_eigh(engine, A):
    n = A.shape[0]
    eps, C, _ = engine.dbg_eigh(A)
    w = np.linalg.eigvalsh(A)
    nrm = np.linalg.norm(A)
    assert np.max(np.abs(eps - w)) <= 1e-13 * nrm
    assert np.all(np.diff(eps) >= 0)
    assert np.max(np.abs(C @ A - eps[:, None] * C)) <= 1e-12 * nrm
    assert np.max(np.abs(C @ C.T - np.eye(n))) <= 1e-12
    assert np.all(C.sum(axis=1) >= 0)
    return eps


@pytest.mark.parametrize("blocked", [True, False])
@pytest.mark.parametrize("n", [138, 139, 145, 200, 255, 256, 263, 300, 384, 401, 450, 511, 512])
def test_eigh_random(tri, n, blocked):
    print(f"tridiag blocked={blocked} n={n}", flush=True)
    tri.set_tridiag(blocked)
    rng = np.random.default_rng(1000 + n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    A[0, :] = A[:, 0] = 0.0
    A[0, 0] = A[1, 1]          # an exact degeneracy and a decoupled row
    _check_eigh(tri, A)


@pytest.mark.parametrize("blocked", [True, False])
def test_eigh_structured(tri, blocked):
    """Zero reflector tails (tau = 0) inside and at the ends of panels: a block
    diagonal matrix with 1x1, 7x7 and 9x9 blocks, a diagonal one, and a graded
    one (entries over 12 orders of magnitude)."""
    tri.set_tridiag(blocked)
    rng = np.random.default_rng(7)
    n = 203
    A = np.zeros((n, n))
    i = 0
    sizes = [1, 7, 9, 8, 16, 1, 33, 64, 40, 24]
    for s in sizes:
        B = rng.standard_normal((s, s))
        A[i:i + s, i:i + s] = B + B.T
        i += s
    assert i == n
    _check_eigh(tri, A)
    _check_eigh(tri, np.diag(rng.standard_normal(160)))
    g = 10.0 ** np.linspace(0, -12, 220)
    B = rng.standard_normal((220, 220))
    _check_eigh(tri, (B + B.T) * np.outer(g, g) + np.diag(g))


def test_blocked_matches_unblocked(tri):
    """Both forms give the same eigenvalues to 1e-13 ||A|| on a Fock-like matrix."""
    rng = np.random.default_rng(11)
    n = 384
    B = rng.standard_normal((n, n)) * 0.05
    A = np.diag(np.linspace(-1.5, 1.0, n)) + B + B.T
    tri.set_tridiag(True)
    e1 = _check_eigh(tri, A)
    tri.set_tridiag(False)
    e0 = _check_eigh(tri, A)
    assert np.max(np.abs(e1 - e0)) <= 1e-13 * np.linalg.norm(A)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_scf_blocked_vs_unblocked_and_oracle(tri, precision):
    """Water cluster (H2O)_40 (240 AOs) and C_30H_62 (182 AOs) in one PM3 batch,
    both tridiagonalizations, against each other and the oracle."""
    print(f"tridiag scf {precision}", flush=True)
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    from oracle import scf as O
    mols = [W.water_cluster(40, seed=31, jitter=0.02), W.alkane(30, seed=32)]
    opts = dict(rms_density=1e-8, diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0,
                damping_weight=0.0, max_iter=200)
    jobs = [Job("pm3", Z, W.to_bohr(x), precision=precision, **opts) for Z, x in mols]
    res = {}
    for blocked in (True, False):
        tri.set_tridiag(blocked)
        res[blocked] = tri.scf_batch(jobs)
    tri.set_tridiag(True)
    tol = 1e-6 if precision == "fp64" else 1e-5
    for j, (Z, x) in enumerate(mols):
        a, b = res[True][j], res[False][j]
        assert a["status"] == 0 and b["status"] == 0
        assert abs(a["energy"] - b["energy"]) * EV <= 1e-8
        o = O.scf("pm3", Z, W.to_bohr(x), rms_density=1e-8, diis_k=5, diis_start=0.1, diis_end=0.0,
                  damping_weight=0.0, max_iter=200)
        assert abs(a["energy"] - o["energy"]) * EV <= tol
        assert np.max(np.abs(a["eps"] - o["eps"])) * EV <= tol
        if precision == "fp64":
            assert a["iterations"] == o["iterations"]
