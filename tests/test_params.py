"""The MNDO / AM1 / PM3 parameter tables against MolDS's own values.

tests/golden/params_fixture.json holds the literals of MolDS's atom classes
(src/base/atoms/{H,C,N,O,F}atom.cpp, read as text by make_param_fixture.py)
with the unit each is converted from.  Both hand-typed tables -- the engine's
(molds_amd/csrc/params.cpp, through nddo_element_params) and the oracle's
(oracle/nddo_oracle.c, through orc_elem_dump) -- must reproduce every value
after MolDS's conversions (Parameters.cpp:45-48), to the last bit or one
rounding of the conversion.  Host-only: no GPU call.
"""
import ctypes
import json
import os

import numpy as np
import pytest

from conftest import ROOT

EV2AU = 0.03674903
KCAL2AU = 0.00159360175
ANG2AU = 1.0 / 0.5291772
THEORIES = ("mndo", "am1", "pm3")
LAYOUT = (["core", "CoreintegralS", "CoreintegralP", "OrbitalExponentS", "OrbitalExponentP", "BondingParameterS",
           "BondingParameterP", "Alpha", "Gss", "Gpp", "Gsp", "Gpp2", "Hsp", "Hpp"]
          + [f"DerivedParameterD[{k}]" for k in range(3)] + [f"DerivedParameterRho[{k}]" for k in range(3)]
          + [f"Parameter{x}[{k}]" for x in "KLM" for k in range(4)] + ["hof_offset"])


def _fixture():
    with open(os.path.join(ROOT, "tests", "golden", "params_fixture.json")) as f:
        return json.load(f)["elements"]


def _convert(val, unit):
    return {"eV": lambda v: v * EV2AU, "kcal": lambda v: v * KCAL2AU, "1/A": lambda v: v / ANG2AU,
            "1/A2": lambda v: v / ANG2AU ** 2, "A": lambda v: v * ANG2AU,
            "half/": lambda v: 0.5 / v if v != 0.0 else 0.0, "1": lambda v: v}[unit](val)


def expected(el, theory):
    t = el[theory]
    out = {"core": el["core"]}
    for name in LAYOUT[1:-1]:
        if name == "Hpp":
            continue
        if name in t:
            out[name] = _convert(*t[name])
        else:   # MNDO has no Gaussian core terms
            assert name.startswith("Parameter") and theory == "mndo", (theory, name)
            out[name] = 0.0
    out["Hpp"] = 0.5 * (out["Gpp"] - out["Gpp2"])   # Atom.cpp:1051-1070
    m = el["mndo"]
    out["hof_offset"] = -_convert(*m["ElecEnergyAtom"]) + _convert(*m["HeatsFormAtom"])
    return np.array([out[k] for k in LAYOUT])


def _tables():
    from molds_amd.nddo import LIBPATH
    from oracle import scf as O
    eng = ctypes.CDLL(LIBPATH)
    orc = O.lib()
    return eng.nddo_element_params, orc.orc_elem_dump


@pytest.mark.parametrize("Z", [1, 6, 7, 8, 9])
@pytest.mark.parametrize("theory", THEORIES)
def test_tables_match_molds_literals(Z, theory):
    el = _fixture()[str(Z)]
    want = expected(el, theory)
    for fn in _tables():
        got = np.zeros(33)
        assert fn(Z, THEORIES.index(theory), got.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
        bad = [(LAYOUT[i], got[i], want[i]) for i in range(33)
               if abs(got[i] - want[i]) > 2.3e-16 * abs(want[i])]
        assert not bad, (fn.__name__, Z, theory, bad)


def test_fixture_covers_the_scf_fields():
    el = _fixture()
    for Z in ("1", "6", "7", "8", "9"):
        for th in THEORIES:
            names = set(el[Z][th])
            for f in ("CoreintegralS", "OrbitalExponentS", "BondingParameterS", "Alpha", "Gss",
                      "DerivedParameterRho[0]"):
                assert f in names, (Z, th, f)
            if th != "mndo":
                assert "ParameterK[0]" in names and "ParameterM[1]" in names


# Published isolated-atom energies (eV): AM1 [DZHS_1985] and PM3 [S_1989]
# (the EISOL values of those parameter sets).
EISOL_PUBLISHED = {
    "am1": {1: -11.396427, 6: -120.815794, 7: -202.407743, 8: -316.099520, 9: -482.290583},
    "pm3": {1: -13.073321, 6: -111.229917, 7: -157.6137755, 9: -437.517169},
}


@pytest.mark.parametrize("theory", ["am1", "pm3"])
def test_eisol_from_one_centre_parameters_matches_published(theory):
    """The AM1 / PM3 heat of formation subtracts E_isol derived from the
    one-centre parameters; the derivation reproduces the published atomic
    energies of both parameter sets, in the engine and in the oracle."""
    from molds_amd.nddo import LIBPATH
    from oracle import scf as O
    eng = ctypes.CDLL(LIBPATH)
    th = THEORIES.index(theory)
    for Z, ref in EISOL_PUBLISHED[theory].items():
        v = ctypes.c_double()
        assert eng.nddo_element_eisol(Z, th, ctypes.byref(v)) == 0
        assert abs(v.value - ref) <= 2e-6, (theory, Z, v.value, ref)
        assert abs(O.lib().orc_eisol_ev(Z, th) - v.value) <= 1e-12 * abs(ref)


def test_eisol_derivation_against_molds_mndo_table():
    """The same derivation against MolDS's tabulated MNDO atomic energies
    (mndoElecEnergyAtom): exact for H, C and F; the tabulated N and O values
    differ by 0.015 and 0.0024 eV (MolDS keeps MNDO's table, which the MNDO
    heat of formation uses unchanged)."""
    from molds_amd.nddo import LIBPATH
    eng = ctypes.CDLL(LIBPATH)
    el = _fixture()
    for Z, tol in ((1, 1e-9), (6, 1e-9), (9, 1e-9), (7, 0.016), (8, 0.0025)):
        v = ctypes.c_double()
        assert eng.nddo_element_eisol(Z, 0, ctypes.byref(v)) == 0
        tab = el[str(Z)]["mndo"]["ElecEnergyAtom"][0]
        assert abs(v.value - tab) <= tol, (Z, v.value, tab)


def test_sto6g_tables_match_molds():
    """The STO-6G expansion of the dipole integrals (the oracle's STO6G table and
    the engine's kernels_props.hip constants) against MolDS's own table
    (src/base/GTOExpansionSTO.cpp, extracted by make_sto6g_fixture.py)."""
    import re
    from oracle import scf as O
    with open(os.path.join(ROOT, "tests", "golden", "sto6g_fixture.json")) as f:
        ref = json.load(f)["expansions"]
    src = open(os.path.join(ROOT, "molds_amd", "csrc", "kernels_props.hip")).read()
    blocks = [list(map(float, re.findall(r"[-+]?\d+\.\d+(?:[eE][-+]?\d+)?", b)))
              for b in re.findall(r"c_sto_(?:exp|coef)\[3\]\[6\] = \{(.*?)\};", src, re.S)]
    eng_exp, eng_coef = blocks[0], blocks[1]
    for i, key in enumerate(("1s", "2s", "2p")):
        exps, coefs = ref[key]["exponents"], ref[key]["coefficients"]
        assert list(O.STO6G[key][0]) == exps and list(O.STO6G[key][1]) == coefs, key
        assert eng_exp[6 * i:6 * i + 6] == exps and eng_coef[6 * i:6 * i + 6] == coefs, key
