import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden", "molds_goldens.json")
ANG2AU = 1.0 / 0.5291772


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built libnddo.so")


def load_goldens():
    with open(GOLDEN) as f:
        return json.load(f)["cases"]


def golden_opts(case):
    s = case["scf"]
    return dict(max_iter=int(s["max_iter"]), rms_density=s["rms_density"], damping_thresh=s["damping_thresh"],
                damping_weight=s["damping_weight"], diis_num_error_vect=int(s["diis_num_error_vect"]),
                diis_start_error=s["diis_start_error"], diis_end_error=s["diis_end_error"])


def oracle_opts(case):
    o = golden_opts(case)
    return dict(max_iter=o["max_iter"], rms_density=o["rms_density"], damping_thresh=o["damping_thresh"],
                damping_weight=o["damping_weight"], diis_k=o["diis_num_error_vect"],
                diis_start=o["diis_start_error"], diis_end=o["diis_end_error"])


def case_xyz_bohr(case):
    return np.array(case["xyz_angstrom"]) * ANG2AU


@pytest.fixture(scope="session")
def engine():
    from molds_amd.nddo import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.fixture(scope="session")
def engine_dw():
    """An engine that solves every molecule above 512 AOs with the device-wide
    block-Jacobi eigensolver (NDDO_ENGINE_DEVICE_WIDE_EIGEN)."""
    from molds_amd.nddo import Engine
    e = Engine(0, flags=Engine.DEVICE_WIDE_EIGEN)
    yield e
    e.close()


@pytest.fixture(scope="session")
def engine_bj():
    """Large molecules (above 512 AOs) on the block-Jacobi eigensolver instead
    of the default device-wide Householder + divide and conquer."""
    from molds_amd.nddo import Engine
    e = Engine(0, flags=Engine.DEVICE_WIDE_EIGEN | Engine.BLOCK_JACOBI | Engine.NO_AUTOTUNE)
    yield e
    e.close()
