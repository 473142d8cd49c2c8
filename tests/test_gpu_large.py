"""Molecules above 512 AOs: the one-workgroup Householder + divide-and-conquer
solver (up to 1200 AOs), the device-wide eigensolver (above 1200, or above 512
with NDDO_ENGINE_DEVICE_WIDE_EIGEN): Householder + divide and conquer over the
whole grid (kernels_xlh.hip), or block Jacobi with NDDO_ENGINE_BLOCK_JACOBI
(kernels_xl.hip), and the chunked DIIS / energy reductions of the device-wide
class.

Tolerances are those of the smaller sizes: eigenvalues within 1e-13 ||A||_F
of LAPACK, residual and orthonormality 1e-12; FP64 SCF within 1e-6 eV total
energy and 1e-8 density RMS of the oracle; mixed precision within 1e-5 eV of
FP64.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EV = 1.0 / 0.03674903
SCF = dict(rms_density=1e-8, diis_num_error_vect=5, diis_start_error=0.1, diis_end_error=0.0,
           damping_weight=0.0, max_iter=200)


def _check_eigh(e, A, label):
    n = A.shape[0]
    t0 = time.perf_counter()
    eps, C, sweeps = e.dbg_eigh(A)
    print(f"  {label} n={n}: {sweeps} sweeps, {time.perf_counter() - t0:.3f} s with transfers", flush=True)
    w = np.linalg.eigvalsh(A)
    nrm = np.linalg.norm(A)
    assert sweeps > 0
    assert np.max(np.abs(eps - w)) <= 1e-13 * nrm
    assert np.all(np.diff(eps) >= 0)
    assert np.max(np.abs(C @ A - eps[:, None] * C)) <= 1e-12 * nrm
    assert np.max(np.abs(C @ C.T - np.eye(n))) <= 1e-12
    assert np.all(C.sum(axis=1) >= 0)


def _random_sym(n):
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    A = A + A.T
    A[0, :] = A[:, 0] = 0.0
    A[0, 0] = A[1, 1]          # an exact degeneracy and a decoupled row
    return A


# one workgroup: 513 (past the 512-AO kernels), 777 / 1000 (back-transform
# variants, trailing block wider than the workgroup past 1025), 1200 (the
# largest); device-wide: 1536 (the (H2O)_256 size)
@pytest.mark.parametrize("n", [513, 777, 1000, 1100, 1200, 1536])
def test_eigh_large(engine, n):
    print(f"eigh large n={n}", flush=True)
    _check_eigh(engine, _random_sym(n), "default")


# the device-wide solver at sizes the one-workgroup path also covers: ragged
# last block, an odd block count (a padding block in the round-robin)
@pytest.mark.parametrize("n", [513, 600, 1000])
def test_eigh_device_wide(engine_dw, n):
    print(f"eigh device-wide n={n}", flush=True)
    _check_eigh(engine_dw, _random_sym(n), "device-wide")


# the device-wide Householder + divide-and-conquer path well past the
# one-workgroup limit: odd sizes, several merge levels with unpaired blocks
@pytest.mark.parametrize("n", [2049, 3001, 4100])
def test_eigh_device_wide_householder_sizes(engine, n):
    print(f"eigh device-wide Householder n={n}", flush=True)
    _check_eigh(engine, _random_sym(n), "householder")


@pytest.fixture
def dw_tri(engine_dw):
    yield engine_dw
    engine_dw.set_xl_tridiag(None)


# both device-wide tridiagonalizations, forced at sizes where the engine picks
# the per-reflector one: blocked (panels of 8, the default from 5600 AOs) and
# per reflector; sizes with a ragged last panel and last MFMA tile
@pytest.mark.parametrize("blocked", [True, False])
@pytest.mark.parametrize("n", [513, 530, 777, 2049])
def test_eigh_device_wide_tridiagonalizations(dw_tri, n, blocked):
    print(f"eigh device-wide blocked={blocked} n={n}", flush=True)
    dw_tri.set_xl_tridiag(blocked)
    _check_eigh(dw_tri, _random_sym(n), f"device-wide blocked={blocked}")


@pytest.mark.parametrize("blocked", [True, False])
def test_eigh_device_wide_structured(dw_tri, blocked):
    """Zero reflector tails (tau = 0) inside and at the ends of panels: a block
    diagonal matrix with 1x1, 7x7, 9x9 and 16x16 blocks, a diagonal one, and a
    graded one (entries over 12 orders of magnitude)."""
    print(f"eigh device-wide structured blocked={blocked}", flush=True)
    dw_tri.set_xl_tridiag(blocked)
    rng = np.random.default_rng(17)
    sizes = [1, 7, 9, 8, 16, 1, 33, 64, 40, 24, 200, 130]
    n = sum(sizes)
    A = np.zeros((n, n))
    i = 0
    for s in sizes:
        B = rng.standard_normal((s, s))
        A[i:i + s, i:i + s] = B + B.T
        i += s
    _check_eigh(dw_tri, A, "block diagonal")
    _check_eigh(dw_tri, np.diag(rng.standard_normal(600)), "diagonal")
    g = 10.0 ** np.linspace(0, -12, 560)
    B = rng.standard_normal((560, 560))
    _check_eigh(dw_tri, (B + B.T) * np.outer(g, g) + np.diag(g), "graded")


def test_scf_device_wide_blocked_equals_unblocked(dw_tri):
    """(H2O)_90, 540 AOs, on the device-wide solver with either
    tridiagonalization: the same SCF (energy within 1e-9 eV, same iteration
    count)."""
    print("scf device-wide blocked vs unblocked", flush=True)
    Z, xyz = _water(90, 11)
    dw_tri.set_xl_tridiag(True)
    a = dw_tri.scf("pm3", Z, xyz, precision="fp64", **SCF)
    dw_tri.set_xl_tridiag(False)
    b = dw_tri.scf("pm3", Z, xyz, precision="fp64", **SCF)
    assert a["status"] == 0 and b["status"] == 0
    assert abs(a["energy"] - b["energy"]) * EV <= 1e-9
    assert a["iterations"] == b["iterations"]
    assert np.max(np.abs(a["eps"] - b["eps"])) * EV <= 1e-9


def test_device_wide_blocked_side_by_side(dw_tri):
    """The blocked tridiagonalization with several molecules of different sizes
    in one launch (a workgroup group and a workspace each, different iteration
    counts), forced below its size threshold: each equals its alone solve."""
    print("device-wide blocked, molecules side by side", flush=True)
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    dw_tri.set_xl_tridiag(True)
    Zc, xc = W.alkane(90, seed=33)
    big = [Job("pm3", *_water(95, 12), rms_density=1e-8), Job("pm3", Zc, W.to_bohr(xc), rms_density=1e-8),
           Job("am1", *_water(88, 13), rms_density=1e-8)]
    mix = dw_tri.scf_batch(big)
    assert all(r["status"] == 0 for r in mix)
    for j, r in zip(big, mix):
        alone = dw_tri.scf_batch([j])[0]
        assert alone["status"] == 0
        assert abs(alone["energy"] - r["energy"]) * EV <= 1e-9
        assert alone["iterations"] == r["iterations"]


# the block-Jacobi alternative (NDDO_ENGINE_BLOCK_JACOBI) stays available
@pytest.mark.parametrize("n", [600, 1536])
def test_eigh_block_jacobi(engine_bj, n):
    print(f"eigh block Jacobi n={n}", flush=True)
    _check_eigh(engine_bj, _random_sym(n), "block-jacobi")


def test_scf_block_jacobi_equals_householder(engine_dw, engine_bj):
    """(H2O)_90, 540 AOs, on both device-wide eigensolvers: the same SCF (energy
    within 1e-9 eV, same iteration count)."""
    Z, xyz = _water(90, 11)
    a = engine_dw.scf("pm3", Z, xyz, precision="fp64", **SCF)
    b = engine_bj.scf("pm3", Z, xyz, precision="fp64", **SCF)
    assert a["status"] == 0 and b["status"] == 0
    assert abs(a["energy"] - b["energy"]) * EV <= 1e-9
    assert a["iterations"] == b["iterations"]


@pytest.mark.parametrize("which", ["default", "device-wide"])
def test_eigh_large_clustered(engine, engine_dw, which):
    """Tight eigenvalue clusters (like the orbital bands of a water cluster)."""
    n = 700
    rng = np.random.default_rng(7)
    Q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    lam = np.repeat(rng.standard_normal(70), 10) + 1e-10 * rng.standard_normal(n)
    A = (Q * lam) @ Q.T
    A = 0.5 * (A + A.T)
    _check_eigh(engine if which == "default" else engine_dw, A, which)


def _water(m, seed):
    from molds_amd import workloads as W
    Z, x = W.water_cluster(m, seed=seed, jitter=0.02)
    return Z, W.to_bohr(x)


@pytest.mark.parametrize("which", ["default", "device-wide"])
def test_scf_large_fp64_and_mixed_vs_oracle(engine, engine_dw, which):
    """(H2O)_90, 540 AOs, PM3: FP64 against the oracle, mixed against both."""
    print(f"scf large (H2O)_90 {which}", flush=True)
    from oracle import scf as O
    e = engine if which == "default" else engine_dw
    Z, xyz = _water(90, 11)
    t0 = time.perf_counter()
    r64 = e.scf("pm3", Z, xyz, precision="fp64", **SCF)
    t1 = time.perf_counter()
    rmx = e.scf("pm3", Z, xyz, precision="mixed", **SCF)
    t2 = time.perf_counter()
    print(f"  fp64 {t1 - t0:.2f} s ({r64['iterations']} it), mixed {t2 - t1:.2f} s ({rmx['iterations']} it)", flush=True)
    o = O.scf("pm3", Z, xyz, rms_density=1e-8, diis_k=5, diis_start=0.1, diis_end=0.0, damping_weight=0.0,
              max_iter=200)
    assert r64["status"] == 0 and rmx["status"] == 0
    assert abs(r64["energy"] - o["energy"]) * EV <= 1e-6
    assert np.sqrt(np.mean((r64["P"] - o["P"]) ** 2)) <= 1e-8
    assert np.max(np.abs(r64["eps"] - o["eps"])) * EV <= 1e-6
    assert abs(rmx["energy"] - r64["energy"]) * EV <= 1e-5
    assert np.max(np.abs(rmx["eps"] - r64["eps"])) * EV <= 1e-5
    assert 1 <= rmx["fp64_iterations"] < rmx["iterations"]


@pytest.mark.parametrize("which", ["default", "device-wide"])
def test_large_alone_equals_batch(engine, engine_dw, which):
    """Large molecules give the same result alone and batched together with
    small ones (with the device-wide solver: several in one persistent launch,
    of different sizes and iteration counts)."""
    from molds_amd import workloads as W
    from molds_amd.nddo import Job
    e = engine if which == "default" else engine_dw
    Zc, xc = W.alkane(90, seed=33)
    big = [Job("pm3", *_water(95, 12), rms_density=1e-8), Job("pm3", Zc, W.to_bohr(xc), rms_density=1e-8),
           Job("am1", *_water(88, 13), rms_density=1e-8)]
    small = [Job("pm3", *(lambda t: (t[0], W.to_bohr(t[1])))(W.alkane(6, seed=31)), rms_density=1e-8),
             Job("am1", *_water(5, 32), rms_density=1e-8)]
    mix = e.scf_batch([small[0], big[0], big[1], small[1], big[2]])
    assert all(r["status"] == 0 for r in mix)
    for j, r in zip(small + big, (mix[0], mix[3], mix[1], mix[2], mix[4])):
        alone = e.scf_batch([j])[0]
        assert alone["status"] == 0
        assert abs(alone["energy"] - r["energy"]) * EV <= 1e-9
        assert alone["iterations"] == r["iterations"]


def _large_vectors():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "large_vectors.json")
    with open(path) as f:
        return json.load(f)


def test_water256_matches_oracle_vectors(engine):
    """(H2O)_256 AM1, 1536 AOs (the single-molecule FP64-vs-mixed workload)
    against the oracle's converged SCF (tests/golden/make_large_vectors.py):
    FP64 within 1e-6 eV total energy, the same iteration count and RMS
    trajectory, orbital energies within 1e-6 eV, density diagonal within 1e-8;
    mixed precision within 1e-5 eV."""
    from molds_amd import workloads as W
    print("(H2O)_256 vs oracle vectors", flush=True)
    ref = _large_vectors()["water256_am1"]
    Z, x = W.water_cluster(256, seed=1, jitter=0.0)
    xyz = W.to_bohr(x)
    r = engine.scf("am1", Z, xyz, precision="fp64", **SCF)
    rm = engine.scf("am1", Z, xyz, precision="mixed", want_C=False, want_P=False, **SCF)
    assert r["status"] == 0 and rm["status"] == 0
    assert abs(r["energy"] - ref["energy"]) * EV <= 1e-6
    assert abs(r["e_core"] - ref["e_core"]) * EV <= 1e-9
    assert r["iterations"] == ref["iterations"]
    traj = np.array([[t[1], t[2]] for t in r["traj"]])
    assert np.allclose(traj, np.array(ref["traj"]), rtol=1e-6, atol=1e-10)
    assert np.max(np.abs(r["eps"] - np.array(ref["eps"]))) * EV <= 1e-6
    P = r["P"]
    assert np.max(np.abs(np.diag(P) - np.array(ref["diag_P"]))) <= 1e-8
    assert abs(np.trace(P) - ref["trace_P"]) <= 1e-8
    assert abs(np.linalg.norm(P) - ref["frob_P"]) <= 1e-8
    assert abs(rm["energy"] - ref["energy"]) * EV <= 1e-5


def test_water1000_first_iterations_match_oracle(engine):
    """(H2O)_1000 PM3, 6000 AOs (device-wide eigensolver, stored integrals):
    the first four iterations' density RMS and DIIS error against the oracle's
    (this geometry does not converge under MolDS's DIIS / damping, in the
    oracle either), plus tr P = n_elec and P^2 = 2P on the last density."""
    from molds_amd import workloads as W
    print("(H2O)_1000 first iterations vs oracle vectors", flush=True)
    ref = _large_vectors()["water1000_pm3_first4"]
    Z, x = W.water_cluster(1000, seed=4, jitter=0.0)
    r = engine.scf("pm3", Z, W.to_bohr(x), precision="fp64", want_C=False, **dict(SCF, max_iter=4))
    assert r["status"] == -1 and r["iterations"] == 4
    traj = np.array([[t[1], t[2]] for t in r["traj"]])
    assert np.allclose(traj, np.array(ref["traj"]), rtol=1e-9, atol=1e-12)
    P = r["P"]
    assert abs(np.trace(P) - 8000.0) <= 1e-7
    Ps = P[:400]
    assert np.max(np.abs(Ps @ P - 2.0 * Ps)) <= 1e-9
