"""Writes tests/golden/math_golden.json: closed-form values that pin the
oracle's special functions and tables independently of the oracle.

Sources: scipy.special (Legendre P_n, spherical Bessel j_n / y_n), numpy
Gauss-Legendre, and the known-answer values of the reference's
test/tMath.cpp (the 15 M2L distances at test/tMath.cpp:152-154 and the
Hankel arguments 5 x {distances}, orders 10..20 at test/tMath.cpp:175-177).
Nothing from the reference is executed.

Run:  python tests/golden/make_golden.py
"""
import json
import os

import numpy as np
from scipy.special import eval_legendre, spherical_jn, spherical_yn

HERE = os.path.dirname(os.path.abspath(__file__))

# test/tMath.cpp:152-154 (rounded as printed there)
TMATH_DISTANCES = [2, 2.23607, 2.44949, 2.82843, 3, 3.16228, 3.31662, 3.4641,
                   3.60555, 3.74166, 4.12311, 4.24264, 4.3589, 4.69042, 5.19615]


def main():
    out = {}
    xs = [round(-1.0 + 0.1 * i, 10) for i in range(21)]
    out["legendre"] = {"x": xs, "n": list(range(0, 21)),
                       "P": [[float(eval_legendre(n, x)) for n in range(0, 21)] for x in xs]}
    d = np.array(TMATH_DISTANCES)
    args = (5.0 * d).tolist()
    H = [[[float(spherical_jn(n, a)), float(spherical_yn(n, a))] for n in range(10, 21)] for a in args]
    out["hankel"] = {"x": args, "n": list(range(10, 21)), "h": H}
    # small orders: h_n(x) over a wide range of arguments
    xs2 = [0.5, 1.0, 2.5, 6.283185307179586, 12.566370614359172, 31.4, 80.0]
    out["hankel_low"] = {"x": xs2, "n": list(range(0, 9)),
                         "h": [[[float(spherical_jn(n, a)), float(spherical_yn(n, a))] for n in range(0, 9)]
                               for a in xs2]}
    gl = {}
    for l in [1, 2, 3, 4, 5, 9, 14, 21, 34, 58]:
        x, w = np.polynomial.legendre.leggauss(l)
        gl[str(l)] = {"x": x.tolist(), "w": w.tolist()}
    out["gauss_legendre"] = gl
    out["tmath_distances"] = TMATH_DISTANCES
    exact = sorted({float(np.sqrt(a * a + b * b + c * c)) for a in range(4) for b in range(4) for c in range(4)
                    if max(a, b, c) > 1})
    out["inode_distances"] = exact
    dirs = set()
    for z in range(-3, 4):
        for y in range(-3, 4):
            for x in range(-3, 4):
                if max(abs(x), abs(y), abs(z)) > 1:
                    g = np.gcd.reduce([abs(x), abs(y), abs(z)])
                    dirs.add((x // g, y // g, z // g))
    out["inode_direction_count"] = len(dirs)
    out["inode_offset_count"] = sum(1 for z in range(-3, 4) for y in range(-3, 4) for x in range(-3, 4)
                                    if max(abs(x), abs(y), abs(z)) > 1)
    with open(os.path.join(HERE, "math_golden.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", os.path.join(HERE, "math_golden.json"))


if __name__ == "__main__":
    main()
