"""Writes spher_u_reference.json: the Cartesian -> real solid harmonic
coefficients U(l) (rows m = -l..l, columns the Cartesian components in the
Auto2e order) for l = 2..5 as the reference's generated code states them in
auto1e_SpherTransf_U_<l> (src/integrals/Auto2e/src/auto2e_SpherTransf.f90,
lines 3938-4130).  The numbers are read from the text of that file (the
statements S(m) = c1 * C(i1) + c2 * C(i2) + ...); nothing is run.

    python tests/golden/make_spher_u.py [path/to/auto2e_SpherTransf.f90]
"""
import json
import os
import re
import sys

SRC = "/root/reference/src/integrals/Auto2e/src/auto2e_SpherTransf.f90"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "spher_u_reference.json")


def parse(text, l):
    m = re.search(r"subroutine auto1e_SpherTransf_U_%d\(S, C\)(.*?)end subroutine" % l, text, re.S)
    body = m.group(1).replace("&\n", " ")
    nc = (l + 1) * (l + 2) // 2
    U = [[0.0] * nc for _ in range(2 * l + 1)]
    for row, rhs in re.findall(r"S\((\d+)\)\s*=\s*([^\n]+)", body):
        terms = re.findall(r"([+-]?)\s*([0-9.]+(?:[eE][+-]?\d+)?)_F64\s*\*\s*C\((\d+)\)", rhs)
        assert terms, rhs
        for sign, val, col in terms:
            U[int(row) - 1][int(col) - 1] += (-1.0 if sign == "-" else 1.0) * float(val)
    return U


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else SRC
    text = open(src).read()
    out = {"source": "src/integrals/Auto2e/src/auto2e_SpherTransf.f90, auto1e_SpherTransf_U_<l>",
           "layout": "U[l][m + l][cartesian index], Cartesian order lx = l..0, ly = l - lx..0",
           "U": {str(l): parse(text, l) for l in range(2, 6)}}
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print(OUT)


if __name__ == "__main__":
    main()
