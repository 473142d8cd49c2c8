"""Writes tests/golden/sto6g_fixture.json: the STO-6G expansion of MolDS
(src/base/GTOExpansionSTO.cpp, read as text): for the 1s (K shell) and the 2s,
2p (L shell) Slater orbitals of exponent 1, the six Gaussian exponents and
contraction coefficients (Stewart's least-squares fits, J. Chem. Phys. 52, 431
(1970)).  The dipole integrals scale the exponents by zeta^2.

Run in the build container (it reads /root/reference):
    python tests/golden/make_sto6g_fixture.py
"""
import json
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
SRC = "/root/reference/src/base/GTOExpansionSTO.cpp"
NUM = r"([-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?)"
PAT = re.compile(r"(exponents|coefficients)\[STO6G\]\[(kShell|lShell)\]\[(sAzimuthal|pAzimuthal)\]\[(\d)\]\s*=\s*" + NUM)


def main():
    text = open(SRC).read()
    out = {}
    for kind, shell, az, i, val in PAT.findall(text):
        key = {"kShell": "1", "lShell": "2"}[shell] + {"sAzimuthal": "s", "pAzimuthal": "p"}[az]
        out.setdefault(key, {"exponents": [None] * 6, "coefficients": [None] * 6})[kind][int(i)] = float(val)
    for key in ("1s", "2s", "2p"):
        assert None not in out[key]["exponents"] and None not in out[key]["coefficients"], key
    out = {k: out[k] for k in ("1s", "2s", "2p")}
    with open(os.path.join(ROOT, "tests", "golden", "sto6g_fixture.json"), "w") as f:
        json.dump({"source": "src/base/GTOExpansionSTO.cpp (STO6G, K and L shells)", "expansions": out}, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
