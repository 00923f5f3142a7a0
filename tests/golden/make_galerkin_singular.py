"""Writes galerkin_singular.json: the five singular face-pair integrals [ST, EAc, EAo, VAc, VAo] of the
Galerkin N operator as the reference computes them (its DIRECTFN C++, Np_1D = 20 as in
assembly_nop_const.m:17, combined as singular_*_const.m do), for a few (ko, dx).  Needs oracle/_ref
(oracle/build_ref.sh, this container only).  Run from the repository root:
    python tests/golden/make_galerkin_singular.py
"""
import json
import math
import os
import sys

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__)))))
from oracle import galerkin as OG  # noqa: E402

CASES = [(2 * math.pi, 1.0 / 35.0), (2 * math.pi, 0.1), (2 * math.pi, 0.05), (1e-9, 1.0)]


def main():
    if not OG.ref_available():
        raise SystemExit("oracle/_ref not built (oracle/build_ref.sh)")
    out = []
    for ko, dx in CASES:
        rec = {"ko": ko, "dx": dx, "n": 20, "names": ["ST", "EAc", "EAo", "VAc", "VAo"]}
        for n in (20, 30, 64):
            vals = OG.reference_singular(ko, dx, n)
            rec[f"directfn_{n}"] = [[v.real, v.imag] for v in vals]
        out.append(rec)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "galerkin_singular.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
