"""MolDS input files and the ground-state SCF results block.

* parse_input(text)        the THEORY, SCF and GEOMETRY directives of a MolDS
                           input (InputParser.cpp: tokens split on spaces and
                           tabs and lower-cased, AddInputTermsFromString
                           :515-534; lines whose first characters after trimming
                           are "#" or "//" skipped, IsCommentOut :2414-2440; SCF
                           knobs ParseConditionsSCF :646-734; geometry in
                           Angstrom converted with angstrom2AU = 1/0.5291772,
                           ParseMolecularConfiguration :542-582, Parameters.cpp:45-48).
                           Directives of other MolDS features (CIS, MD, MC,
                           optimisation, memory, ...) are ignored: this engine
                           runs the ground-state SCF only.
* format_scf_results(...)  the lines MolDS prints for that SCF: iteration
                           table (Cndo2.cpp:1505-1510), convergence line, MO
                           energies (:1063-1081), electronic and core-repulsion
                           energies (:1083-1107), the STO-6G dipole moments
                           (:1126-1184), Mulliken charges (:1187-1199) and, for
                           MNDO, the heat of formation (Mndo.cpp:412-420; MolDS
                           prints none for AM1/PM3, Am1.cpp:261-263, although
                           the engine returns it).
* main()                   python -m molds_amd.molds_io input.in: solve on GPU 0
                           and print the block.
"""
from __future__ import annotations

import sys

import numpy as np

ANGSTROM2AU = 1.0 / 0.5291772     # Parameters.cpp:48
EV2AU = 0.03674903                # Parameters.cpp:45
KCALMOL2AU = 0.00159360175        # Parameters.cpp:47
DEBYE2AU = 0.393430191            # Parameters.cpp:57

THEORIES = {"mndo": "MNDO", "am1": "AM1", "pm3": "PM3"}
OTHER_THEORIES = ("cndo/2", "indo", "zindo/s", "am1-d", "pm3-d", "pm3/pddg")
ELEMENTS = {"h": 1, "c": 6, "n": 7, "o": 8, "f": 9}
SYMBOL = {1: "H", 6: "C", 7: "N", 8: "O", 9: "F"}

# SCF block keyword -> engine option (defaults as Parameters.cpp:132-144)
SCF_KEYS = {"max_iter": ("max_iter", int), "rms_density": ("rms_density", float),
            "damping_thresh": ("damping_thresh", float), "damping_weight": ("damping_weight", float),
            "diis_num_error_vect": ("diis_num_error_vect", int), "diis_start_error": ("diis_start_error", float),
            "diis_end_error": ("diis_end_error", float)}
# SCF keywords with one argument that this engine does not use
SCF_IGNORED_1 = ("vdw", "vdw_s6", "vdw_d", "mpi", "scalapack")


class InputError(ValueError):
    pass


def _terms(text):
    out = []
    for line in text.splitlines():
        s = line.strip()
        if s.startswith("#") or s.startswith("//"):
            continue
        out += [t.lower() for t in s.replace("\t", " ").split(" ") if t]
    return out


def parse_input(text):
    """Returns dict(theory, opts, Z, xyz_bohr) for the SCF the input asks for."""
    t = _terms(text)
    if not t:
        raise InputError("input is empty")
    theory, opts, Z, xyz = None, {}, [], []
    i = 0
    while i < len(t):
        w = t[i]
        if w == "theory":
            i += 1
            while i < len(t) and t[i] != "theory_end":
                if t[i] in THEORIES:
                    theory = t[i]
                elif t[i] in OTHER_THEORIES:
                    raise InputError(f"theory {t[i]} is not an NDDO theory this engine runs (mndo, am1, pm3)")
                i += 1
        elif w == "scf":
            i += 1
            while i < len(t) and t[i] != "scf_end":
                k = t[i]
                if k in SCF_KEYS:
                    name, conv = SCF_KEYS[k]
                    try:
                        opts[name] = conv(float(t[i + 1])) if conv is int else conv(t[i + 1])
                    except (IndexError, ValueError):
                        raise InputError(f"SCF {k}: missing or bad value") from None
                    i += 2
                elif k == "sum_charges":
                    i += 3          # two atom indices
                elif k in SCF_IGNORED_1:
                    i += 2
                else:
                    i += 1
        elif w == "geometry":
            i += 1
            while i < len(t) and t[i] != "geometry_end":
                if t[i] not in ELEMENTS:
                    raise InputError(f"element {t[i]} is not supported (H, C, N, O, F)")
                try:
                    x, y, z = (float(v) * ANGSTROM2AU for v in t[i + 1:i + 4])
                except ValueError:
                    raise InputError(f"bad coordinates for atom {len(Z)}") from None
                Z.append(ELEMENTS[t[i]])
                xyz.append((x, y, z))
                i += 4
        i += 1
    if theory is None:
        raise InputError("no THEORY block naming mndo, am1 or pm3")
    if not Z:
        raise InputError("no GEOMETRY block")
    return dict(theory=theory, opts=opts, Z=np.array(Z, dtype=np.int32), xyz_bohr=np.array(xyz, dtype=np.float64))


def _e(v):
    return "%e" % v


def format_scf_results(theory, Z, res):
    """The SCF block for a result of Engine.scf (with traj and mulliken)."""
    T = THEORIES[theory]
    L = [f"**********  START: {T}-SCF  **********",
         "\t\t\t|  RMS density  | DIIS error | DIIS on/off | damping on/off |"]
    for it, rms, derr, don, dmp in res.get("traj") or []:
        L.append(f"\tSCF iter {it}\t{_e(rms)}\t{_e(derr)}\t{'on' if don else ''}\t\t{'on' if dmp else ''}")
    if res["status"] != 0:
        L.append(f"\t\t{T}-SCF ended with status {res['status_name']}")
        L.append(f"**********  DONE: {T}-SCF  **********")
        return "\n".join(L) + "\n"
    L += ["", "", "", f"\t\t{T}-SCF met convergence criterion(^^b", "", "", ""]
    L.append("\t\t\t| i-th | occ/unocc |  e[a.u.]  |  e[eV]  | ")
    for i, e in enumerate(res["eps"]):
        L.append(f"\tEnergy of MO:\t{i}\t{'occ' if i < res['nocc'] else 'unocc'}\t{_e(e)}\t{_e(e / EV2AU)}")
    L += ["", "\t\t\t\t|   [a.u.]   |   [eV]   |",
          f"\tElectronic energy(SCF):\t{_e(res['energy'])}\t{_e(res['energy'] / EV2AU)}",
          "\tNote that this electronic energy includes core-repulsions.", "",
          "\t\t\t\t|   [a.u.]   |   [eV]   |",
          f"\tCore repulsion energy:\t{_e(res['e_core'])}\t{_e(res['e_core'] / EV2AU)}", ""]
    if res.get("dipole_total") is not None:
        # Cndo2::OutputSCFDipole (Cndo2.cpp:1126-1184): x, y, z, magnitude in a.u. and debye
        for title, key, label in (("   ", "dipole_total", "Total Dipole moment(SCF):"),
                                  ("  ", "dipole_electronic", "Electronic Dipole moment(SCF):"),
                                  ("  ", "dipole_core", "Core Dipole moment:\t")):
            v = [float(x) for x in res[key]]
            mag = float(np.sqrt(sum(x * x for x in v)))
            L.append(f"\t\t\t\t\t|{title}x[a.u.]{title}|{title}y[a.u.]{title}|{title}z[a.u.]{title}|  "
                     f"magnitude[a.u.]  |\t\t|  x[debye]  |  y[debye]  |  z[debye]  |  magnitude[debye]  |")
            L.append(f"\t{label}\t" + "\t".join(_e(x) for x in v + [mag]) + "\t\t" +
                     "\t".join(_e(x / DEBYE2AU) for x in v + [mag]))
            L.append("")
    q = res.get("mulliken")
    if q is not None:
        L.append("\t\t\t\t| k-th eigenstate | i-th atom | atom type | core charge[a.u.] | Mulliken charge[a.u.]| ")
        for a, z in enumerate(Z):
            core = 1.0 if z == 1 else float(z - 2)
            L.append(f"\tMulliken charge(SCF):\t0\t{a}\t{SYMBOL[int(z)]}\t{_e(core)}\t{_e(q[a])}")
        L.append("")
    if theory == "mndo":
        L += ["\t\t\t\t|  [a.u.]  |  [Kcal/mol]  | ",
              f"\tHeats of formation:\t{_e(res['hof'])}\t{_e(res['hof'] / KCALMOL2AU)}", ""]
    L.append(f"**********  DONE: {T}-SCF  **********")
    return "\n".join(L) + "\n"


def run(text, engine=None, precision="fp64"):
    """Solves the SCF an input asks for; returns (parsed input, result, printed block)."""
    from molds_amd.nddo import Engine
    inp = parse_input(text)
    eng = engine if engine is not None else Engine(0)
    res = eng.scf(inp["theory"], inp["Z"], inp["xyz_bohr"], precision=precision, want_dipole=True, **inp["opts"])
    return inp, res, format_scf_results(inp["theory"], inp["Z"], res)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    if len(argv) != 1:
        print("usage: python -m molds_amd.molds_io input.in", file=sys.stderr)
        return 2
    with open(argv[0]) as f:
        _, res, text = run(f.read())
    sys.stdout.write(text)
    return 0 if res["status"] == 0 else 1


if __name__ == "__main__":
    sys.exit(main())
