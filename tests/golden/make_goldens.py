"""Extract MolDS's own NDDO golden fixtures into tests/golden/molds_goldens.json.

Reads the reference test inputs (test/<case>.in) and golden outputs
(test/<case>.dat) as text and keeps only data: the SCF knobs, the geometry,
and the printed results of the ground-state SCF (iteration lines, MO energies,
electronic energy incl. core repulsion, core repulsion, MNDO heat of formation,
Mulliken charges, total / electronic / core dipole moments in a.u.).  Run where the reference tree is available (the machines
that run the GPU tests have none):  python tests/golden/make_goldens.py /root/reference/test
"""
import glob
import json
import os
import re
import sys

ELEM = {"h": 1, "c": 6, "n": 7, "o": 8, "f": 9}


def parse_in(path):
    toks = []
    for line in open(path):
        line = line.split("//")[0]
        toks += line.split()
    toks = [t.lower() for t in toks]
    out = {"scf": {}, "geometry": []}
    i = 0
    while i < len(toks):
        t = toks[i]
        if t == "theory":
            out["theory"] = toks[i + 1]
            i += 2
            continue
        if t == "scf":
            i += 1
            while toks[i] != "scf_end":
                try:
                    out["scf"][toks[i]] = float(toks[i + 1])
                    i += 2
                except ValueError:  # flag-style directive without a value
                    i += 1
            continue
        if t == "geometry":
            i += 1
            while toks[i] != "geometry_end":
                out["geometry"].append([toks[i], float(toks[i + 1]), float(toks[i + 2]), float(toks[i + 3])])
                i += 4
            continue
        i += 1
    return out


def parse_dat(path, theory):
    txt = open(path).read()
    tag = theory.upper() + "-SCF"
    start = txt.find("START: " + tag)
    end = txt.find("DONE: " + tag, start)
    if start < 0 or end < 0:
        return None
    blk = txt[start:end]
    res = {"iters": [], "mo": [], "mulliken": []}
    for line in blk.splitlines():
        s = line.strip()
        if s.startswith("SCF iter "):
            # "SCF iter %d\t%e\t%e\t%s\t\t%s" (Cndo2.cpp:1505-1510)
            f = line.lstrip("\t").rstrip("\n").split("\t")
            f += [""] * (6 - len(f))
            res["iters"].append([int(f[0].split()[2]), float(f[1]), float(f[2]),
                                 f[3].strip() != "", f[5].strip() != ""])
            continue
        m = re.match(r"Energy of MO:\s+(\d+)\s+(occ|unocc)\s+(\S+)\s+(\S+)", s)
        if m:
            res["mo"].append(float(m.group(3)))
            continue
        m = re.match(r"Electronic energy\(SCF\):\s+(\S+)\s+(\S+)", s)
        if m:
            res["energy"] = float(m.group(1))
            continue
        m = re.match(r"Core repulsion energy:\s+(\S+)\s+(\S+)", s)
        if m:
            res["ecore"] = float(m.group(1))
            continue
        m = re.match(r"Heats of formation:\s+(\S+)\s+(\S+)", s)
        if m:
            res["hof"] = float(m.group(1))
            continue
        m = re.match(r"(Total|Electronic) Dipole moment\(SCF\):\s+(\S+)\s+(\S+)\s+(\S+)", s)
        if m:
            res["dipole_" + m.group(1).lower()] = [float(m.group(k)) for k in (2, 3, 4)]
            continue
        m = re.match(r"Core Dipole moment:\s+(\S+)\s+(\S+)\s+(\S+)", s)
        if m:
            res["dipole_core"] = [float(m.group(k)) for k in (1, 2, 3)]
            continue
        m = re.match(r"Mulliken charge\(SCF\):\s+\d+\s+\d+\s+\S+\s+(\S+)\s+(\S+)", s)
        if m:
            res["mulliken"].append(float(m.group(2)))
            continue
    # heat of formation is printed after the DONE line of the MNDO block
    m = re.search(r"Heats of formation:\s+(\S+)\s+(\S+)", txt[start:])
    if m and theory == "mndo":
        res["hof"] = float(m.group(1))
    if "converged" not in blk and "met convergence criterion" not in blk:
        return None
    return res


def main(refdir):
    cases = []
    for inp in sorted(glob.glob(os.path.join(refdir, "*.in"))):
        name = os.path.basename(inp)[:-3]
        dat = inp[:-3] + ".dat"
        if not os.path.exists(dat):
            continue
        if "vdw" in name or "UEP" in name or "MC" in name:
            continue
        try:
            d = parse_in(inp)
        except (ValueError, IndexError):  # inputs outside the SCF directive subset
            continue
        th = d.get("theory")
        if th not in ("mndo", "am1", "pm3"):
            continue
        if any(a[0] not in ELEM for a in d["geometry"]):
            continue
        # keep only plain ground-state SCF outputs (no vdW/EPC variants)
        if "vdw" in name or "UEP" in name or "MC" in name:
            continue
        r = parse_dat(dat, th)
        if r is None or "energy" not in r:
            continue
        cases.append({"name": name, "theory": th, "scf": d["scf"],
                      "Z": [ELEM[a[0]] for a in d["geometry"]],
                      "xyz_angstrom": [a[1:] for a in d["geometry"]], "expect": r})
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "molds_goldens.json")
    with open(out, "w") as f:
        json.dump({"source": "MolDS test/*.in + test/*.dat (ground-state SCF block)", "cases": cases}, f, indent=1)
    print(f"wrote {len(cases)} cases to {out}")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/test")
