"""MolDS input files and the printed SCF results block (molds_amd/molds_io.py).

The inputs under tests/golden/inputs are four of MolDS's own test inputs; the
expected numbers are the printed results of the matching MolDS outputs
(tests/golden/molds_goldens.json), compared to their 7 printed digits.
Host-only tests parse the inputs and print the oracle's results; the GPU test
runs the inputs end to end through the engine."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, case_xyz_bohr, golden_opts, load_goldens

INPUTS = os.path.join(ROOT, "tests", "golden", "inputs")
NAMES = ["c2h6_pm3", "ch3f_am1", "c2h6_mndo", "h2o_pm3_freq_num_proj_s1"]
REL = 5.01e-7


def _case(name):
    return [c for c in load_goldens() if c["name"] == name][0]


def _text(name):
    with open(os.path.join(INPUTS, name + ".in")) as f:
        return f.read()


def _read_block(text):
    """The numbers of a printed SCF block, read with the line patterns of MolDS's output."""
    out = dict(mo=[], mulliken=[], iters=[])
    for line in text.splitlines():
        s = line.strip()
        m = re.match(r"SCF iter (\d+)\s+(\S+)\s+(\S+)", s)
        if m:
            out["iters"].append((int(m.group(1)), float(m.group(2)), float(m.group(3)), "on" in s.split()[5:6]))
        m = re.match(r"Energy of MO:\s+(\d+)\s+(\S+)\s+(\S+)\s+(\S+)", s)
        if m:
            out["mo"].append(float(m.group(3)))
        m = re.match(r"Electronic energy\(SCF\):\s+(\S+)", s)
        if m:
            out["energy"] = float(m.group(1))
        m = re.match(r"Core repulsion energy:\s+(\S+)", s)
        if m:
            out["ecore"] = float(m.group(1))
        m = re.match(r"Mulliken charge\(SCF\):\s+\d+\s+\d+\s+\S+\s+(\S+)\s+(\S+)", s)
        if m:
            out["mulliken"].append(float(m.group(2)))
        m = re.match(r"Heats of formation:\s+(\S+)\s+(\S+)", s)
        if m:
            out["hof"] = float(m.group(1))
        m = re.match(r"(Total|Electronic) Dipole moment\(SCF\):\s+(\S+)\s+(\S+)\s+(\S+)", s)
        if m:
            out["dipole_" + m.group(1).lower()] = [float(m.group(k)) for k in (2, 3, 4)]
        m = re.match(r"Core Dipole moment:\s+(\S+)\s+(\S+)\s+(\S+)", s)
        if m:
            out["dipole_core"] = [float(m.group(k)) for k in (1, 2, 3)]
    return out


def _check_against_golden(block, c):
    e = c["expect"]
    assert abs(block["energy"] - e["energy"]) <= REL * abs(e["energy"])
    assert abs(block["ecore"] - e["ecore"]) <= REL * abs(e["ecore"])
    mo = np.array(e["mo"])
    assert len(block["mo"]) == len(mo)
    assert np.all(np.abs(np.array(block["mo"]) - mo) <= REL * np.abs(mo) + 1e-12)
    g = np.array(e["mulliken"])
    assert np.all(np.abs(np.array(block["mulliken"]) - g) <= REL * np.abs(g) + 1e-9)
    if c["theory"] == "mndo":
        assert abs(block["hof"] - e["hof"]) <= REL * abs(e["hof"])
    if c["scf"]["rms_density"] >= 1e-8:
        assert len(block["iters"]) == len(e["iters"])
    for k in ("dipole_total", "dipole_electronic", "dipole_core"):
        g = np.array(e[k])
        v = np.array(block[k])
        assert np.all(np.abs(v - g) <= REL * np.abs(g) + 5.01e-7 * max(np.max(np.abs(g)), 1e-3) + 1e-12), k


@pytest.mark.parametrize("name", NAMES)
def test_parse_input_matches_golden_case(name):
    from molds_amd.molds_io import parse_input
    inp = parse_input(_text(name))
    c = _case(name)
    assert inp["theory"] == c["theory"]
    assert list(inp["Z"]) == list(c["Z"])
    assert np.allclose(inp["xyz_bohr"], case_xyz_bohr(c), rtol=0, atol=1e-12)
    o = golden_opts(c)
    for k, v in inp["opts"].items():
        assert v == o[k], k


def test_parse_input_rules():
    from molds_amd.molds_io import InputError, parse_input
    txt = "# comment\n  // also a comment\nTHEORY\n\tPM3\nTHEORY_END\nSCF\n max_iter 7\n vdw yes\n sum_charges 0 1\n" \
          " rms_density 1e-9\nSCF_END\nGEOMETRY\nO 0 0 0\nH\t0.96\t0 0\nh 0 0.96 0\nGEOMETRY_END\n"
    inp = parse_input(txt)
    assert inp["theory"] == "pm3" and inp["opts"] == {"max_iter": 7, "rms_density": 1e-9}
    assert list(inp["Z"]) == [8, 1, 1]
    assert abs(inp["xyz_bohr"][1, 0] - 0.96 / 0.5291772) < 1e-14
    with pytest.raises(InputError):
        parse_input(txt.replace("PM3", "zindo/s"))
    with pytest.raises(InputError):
        parse_input(txt.replace("O 0 0 0", "S 0 0 0"))
    with pytest.raises(InputError):
        parse_input("")


@pytest.mark.parametrize("name", NAMES)
def test_printed_block_of_oracle_results(name):
    """The printer's lines, read back with MolDS's output patterns, carry the
    oracle's results to the goldens' 7 digits."""
    from molds_amd.molds_io import format_scf_results, parse_input
    from oracle import scf as O
    inp = parse_input(_text(name))
    o = golden_opts(_case(name))
    r = O.scf(inp["theory"], inp["Z"], inp["xyz_bohr"], max_iter=o["max_iter"], rms_density=o["rms_density"],
              damping_thresh=o["damping_thresh"], damping_weight=o["damping_weight"], diis_k=o["diis_num_error_vect"],
              diis_start=o["diis_start_error"], diis_end=o["diis_end_error"], dipole=True)
    res = dict(r, status=0, status_name="OK", e_core=r["e_core"], mulliken=O.mulliken(inp["Z"], r["P"]),
               hof=r.get("hof", float("nan")))
    text = format_scf_results(inp["theory"], inp["Z"], res)
    assert text.startswith("**********  START: ") and text.rstrip().endswith("-SCF  **********")
    _check_against_golden(_read_block(text), _case(name))


@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_run_input_on_gpu(engine, name):
    from molds_amd.molds_io import run
    print(f"molds input {name}", flush=True)
    inp, res, text = run(_text(name), engine=engine)
    assert res["status"] == 0
    _check_against_golden(_read_block(text), _case(name))
