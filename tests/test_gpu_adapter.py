"""The C++ ElectronicStructure-shaped adapter (include/nddo_es.hpp) run the way
MolDS runs Cndo2 -- SetMolecule, DoSCF, getters, DoSCF(false), GetForce(0),
exceptions -- on MolDS golden cases; energies within the goldens' 7 printed
digits, forces against the oracle's central differences."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, case_xyz_bohr, golden_opts, load_goldens, oracle_opts

BIN = os.path.join(ROOT, "tests", "cpp", "_build", "test_es_adapter")
THEORY = {"mndo": 0, "am1": 1, "pm3": 2}


def _run_case(tmp_path, c):
    o = golden_opts(c)
    xyz = case_xyz_bohr(c).reshape(-1, 3)
    lines = [f"{THEORY[c['theory']]} {int(o['max_iter'])} {float(o['rms_density'])!r} {float(o['damping_thresh'])!r} "
             f"{float(o['damping_weight'])!r} {int(o['diis_num_error_vect'])} {float(o['diis_start_error'])!r} "
             f"{float(o['diis_end_error'])!r}",
             str(len(c["Z"]))]
    lines += [f"{int(z)} {float(x)!r} {float(y)!r} {float(w)!r}" for z, (x, y, w) in zip(c["Z"], xyz)]
    f = tmp_path / f"{c['name']}.txt"
    f.write_text("\n".join(lines) + "\n")
    env = dict(os.environ, LD_LIBRARY_PATH=os.path.join(ROOT, "molds_amd", "lib"))
    r = subprocess.run([BIN, str(f)], capture_output=True, text=True, timeout=60, env=env)
    return r


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["c2h6_pm3", "ch3f_am1", "c2h6_mndo"])
def test_cpp_adapter_on_goldens(tmp_path, name):
    print(f"adapter {name}", flush=True)
    assert os.path.exists(BIN), "build first (python -m molds_amd.build)"
    c = [c for c in load_goldens() if c["name"] == name][0]
    r = _run_case(tmp_path, c)
    assert r.returncode == 0, r.stdout + r.stderr
    vals = dict(zip(*[iter(r.stdout.split("\n")[0].split())] * 2))
    warm = dict(zip(*[iter([l for l in r.stdout.split("\n") if l.startswith("warm_")][0].split())] * 2))
    e = c["expect"]
    assert abs(float(vals["energy"]) - e["energy"]) <= 5.01e-7 * abs(e["energy"])
    assert abs(float(vals["core"]) - e["ecore"]) <= 5.01e-7 * abs(e["ecore"])
    assert int(warm["warm_iterations"]) <= 3
    dline = [l for l in r.stdout.split("\n") if l.startswith("dipole ")][0]
    dip = np.array([float(v) for v in dline.split()[1:]])
    g = np.array(e["dipole_total"])
    assert np.all(np.abs(dip - g) <= 5.01e-7 * np.abs(g) + 5.01e-7 * max(np.max(np.abs(g)), 1e-3) + 1e-12)
    # GetForce(0): the same central differences on the oracle's SCF energies
    from oracle import scf as O
    from molds_amd.forces import stencil_forces
    line = [l for l in r.stdout.split("\n") if l.startswith("forces ")][0]
    F = np.array([float(v) for v in line.split()[1:]]).reshape(-1, 3)
    oo = dict(oracle_opts(c), rms_density=1e-10, max_iter=200)
    x0 = case_xyz_bohr(c).reshape(-1, 3)
    P = O.scf(c["theory"], c["Z"], x0, **oo)["P"]
    Fo = stencil_forces(lambda gs: [O.scf(c["theory"], c["Z"], g, requires_guess=False, P0=P, **oo)["energy"]
                                    for g in gs], x0, 1e-3)
    assert np.max(np.abs(F - Fo)) <= 1e-8
    assert r.stdout.strip().endswith("OK")


def test_case_file_parses(tmp_path):
    """Host-only: the case file the GPU test writes holds plain numbers."""
    c = [c for c in load_goldens() if c["name"] == "c2h6_pm3"][0]
    o = golden_opts(c)
    xyz = case_xyz_bohr(c).reshape(-1, 3)
    line = f"{int(c['Z'][0])} {float(xyz[0][0])!r} {float(xyz[0][1])!r} {float(xyz[0][2])!r}"
    assert all(float(t) == float(t) for t in line.split()), line
    assert "np." not in line and "np." not in repr(float(o["rms_density"]))


def test_adapter_header_compiles():
    """Host-only: the adapter header is standalone C++17 over nddo.h."""
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I" + os.path.join(ROOT, "include"),
                        "-x", "c++", "-"], input='#include "nddo_es.hpp"\nint main() { return 0; }\n',
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
