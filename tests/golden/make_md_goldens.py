"""Extract MolDS's ground-state MD goldens into tests/golden/molds_md_goldens.json.

Reads the reference test inputs/outputs test/c2h6_{am1,mndo,pm3}_force.{in,dat}
as text and keeps only data: the geometry (Angstrom), the SCF knobs, the MD
time step (fs), and the printed atom momenta and coordinates (a.u.) and the
electronic energy after MD step 1 (velocity Verlet from rest, md/MD.cpp:87-133).  Run where the
reference tree is available:  python tests/golden/make_md_goldens.py /root/reference/test
"""
import json
import os
import re
import sys

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from make_goldens import parse_in  # noqa: E402


def parse_md(path):
    txt = open(path).read()
    a = txt.find("START: MD step 1")
    b = txt.find("DONE: MD step 1", a)
    blk = txt[a:b]
    # each line holds p / (angstrom2AU gMolin2AU / fs2AU), then p in a.u.
    # (Molecule::OutputMomenta, Molecule.cpp:553-573): keep the a.u. triple
    mom = [[float(v) for v in m.groups()[3:]] for m in
           re.finditer(r"Atom momenta:\s+\d+\s+\S+" + r"\s+(\S+)" * 6, blk)]
    xyz = [[float(v) for v in m.groups()[:3]] for m in
           re.finditer(r"Atom coordinates:\s+\d+\s+\S+" + r"\s+(\S+)" * 6, blk)]
    e = re.search(r"Electronic energy\(SCF\):\s+(\S+)", blk)
    return mom, xyz, float(e.group(1))


def main(test_dir):
    cases = []
    for name in ("c2h6_am1_force", "c2h6_mndo_force", "c2h6_pm3_force"):
        text = open(os.path.join(test_dir, name + ".in")).read()
        inp = parse_in(os.path.join(test_dir, name + ".in"))
        dt = float(re.search(r"(?im)^\s*dt\s+(\S+)", text).group(1))
        mom, xyz1, e1 = parse_md(os.path.join(test_dir, name + ".dat"))
        Z = [{"h": 1, "c": 6, "n": 7, "o": 8, "f": 9}[g[0]] for g in inp["geometry"]]
        cases.append(dict(name=name, theory=inp["theory"], scf=inp["scf"], Z=Z,
                          xyz_angstrom=[g[1:] for g in inp["geometry"]], dt_fs=dt,
                          step1_momenta=mom, step1_xyz_bohr=xyz1, step1_energy=e1))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "molds_md_goldens.json")
    with open(out, "w") as f:
        json.dump({"source": "MolDS test/*_force.{in,dat}, MD step 1", "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test")
