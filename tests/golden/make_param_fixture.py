"""Writes tests/golden/params_fixture.json: the MNDO / AM1 / PM3 parameters of
H, C, N, O and F as MolDS's atom classes set them, read as text from the
reference's src/base/atoms/{H,C,N,O,F}atom.cpp (SetAtomicParameters).

For every assignment `this-><theory><Field>[<k>] = <expr>;` of the fields the
NDDO SCF uses, the fixture records the numeric literal and the unit the
expression converts it from, as the expression spells it:
  "eV"      literal * GetEV2AU()                (U, beta, G, H, K, E_isol)
  "kcal"    literal * GetKcalMolin2AU()         (MNDO atomic heat of formation)
  "1/A"     literal / GetAngstrom2AU()          (alpha)
  "1/A2"    literal / pow(GetAngstrom2AU(), 2)  (L)
  "A"       literal * GetAngstrom2AU()          (M)
  "half/"   0.5 / literal                       (rho, stored as 0.5 / (1/2 rho))
  "1"       literal                             (zeta, D in bohr, zeros)
An assignment from another field (`this->am1Gss = this->mndoGss;`) takes that
field's entry.  Only the data is kept; the test applies MolDS's conversion
constants (Parameters.cpp:45-48) itself.

Run in the build container (it reads /root/reference):
    python tests/golden/make_param_fixture.py
"""
import json
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
REF = "/root/reference/src/base/atoms"
ELEMENTS = {1: "Hatom.cpp", 6: "Catom.cpp", 7: "Natom.cpp", 8: "Oatom.cpp", 9: "Fatom.cpp"}
FIELDS = ["CoreintegralS", "CoreintegralP", "OrbitalExponentS", "OrbitalExponentP", "BondingParameterS",
          "BondingParameterP", "Alpha", "Gss", "Gpp", "Gsp", "Gpp2", "Hsp", "DerivedParameterD",
          "DerivedParameterRho", "ParameterK", "ParameterL", "ParameterM", "ElecEnergyAtom", "HeatsFormAtom"]
NUM = r"([-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?)"
ASSIGN = re.compile(r"^\s*this->(mndo|am1|pm3)(" + "|".join(FIELDS) + r")(?:\[(\d)\])?\s*=\s*(.*?);", re.M)


def parse_expr(expr):
    e = expr.replace(" ", "")
    m = re.fullmatch(r"this->(mndo|am1|pm3)(\w+?)(?:\[(\d)\])?", e)
    if m:
        return ("alias", m.group(1), m.group(2), int(m.group(3)) if m.group(3) else None)
    m = re.fullmatch(r"0\.5/" + NUM, e)
    if m:
        return (float(m.group(1)), "half/")
    m = re.fullmatch(NUM + r"(.*)", e)
    if not m:
        raise ValueError(f"unrecognised expression: {expr}")
    val, rest = float(m.group(1)), m.group(2)
    unit = {"": "1", "*Parameters::GetInstance()->GetEV2AU()": "eV",
            "*Parameters::GetInstance()->GetKcalMolin2AU()": "kcal",
            "/Parameters::GetInstance()->GetAngstrom2AU()": "1/A",
            "/pow(Parameters::GetInstance()->GetAngstrom2AU(),2.0)": "1/A2",
            "*Parameters::GetInstance()->GetAngstrom2AU()": "A"}.get(rest)
    if unit is None:
        raise ValueError(f"unrecognised conversion: {expr}")
    return (val, unit)


def main():
    out = {"source": "src/base/atoms/{H,C,N,O,F}atom.cpp, SetAtomicParameters", "elements": {}}
    for Z, fname in ELEMENTS.items():
        text = open(os.path.join(REF, fname)).read()
        core = float(re.search(r"this->coreCharge\s*=\s*" + NUM + ";", text).group(1))
        raw = {}
        for th, field, idx, expr in ASSIGN.findall(text):
            key = (th, field, int(idx) if idx else None)
            raw[key] = parse_expr(expr)
        for key, v in list(raw.items()):
            while v[0] == "alias":
                v = raw[(v[1], v[2], v[3])]
            raw[key] = v
        el = {"core": core}
        for (th, field, idx), (val, unit) in sorted(raw.items(), key=lambda kv: str(kv[0])):
            name = field if idx is None else f"{field}[{idx}]"
            el.setdefault(th, {})[name] = [val, unit]
        out["elements"][str(Z)] = el
    with open(os.path.join(ROOT, "tests", "golden", "params_fixture.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", sum(len(v.get(t, {})) for v in out["elements"].values() for t in ("mndo", "am1", "pm3")), "values")


if __name__ == "__main__":
    main()
