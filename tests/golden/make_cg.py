"""Writes cg_2j8.json: Clebsch-Gordan coefficients <j1 m1; j2 m2 | j m> for
all 2j1, 2j2, 2j <= 8 from sympy (an independent library routine), indexed
like the SNAP tables: (2j1, 2j2, 2j, k1, k2) with m1 = k1 - j1/2,
m2 = k2 - j2/2, m = m1 + m2.  Run: python tests/golden/make_cg.py
"""
import json
import os

from sympy import Rational
from sympy.physics.quantum.cg import CG

TWOJMAX = 8
out = []
for j1 in range(TWOJMAX + 1):
    for j2 in range(TWOJMAX + 1):
        for j in range(abs(j1 - j2), min(TWOJMAX, j1 + j2) + 1, 2):
            for k1 in range(j1 + 1):
                for k2 in range(j2 + 1):
                    aa2, bb2 = 2 * k1 - j1, 2 * k2 - j2
                    m = (aa2 + bb2 + j) // 2
                    if m < 0 or m > j:
                        continue
                    c = CG(Rational(j1, 2), Rational(aa2, 2), Rational(j2, 2), Rational(bb2, 2),
                           Rational(j, 2), Rational(aa2 + bb2, 2)).doit()
                    out.append([j1, j2, j, k1, k2, float(c)])
path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cg_2j8.json")
with open(path, "w") as f:
    json.dump({"twojmax": TWOJMAX, "source": "sympy.physics.quantum.cg.CG", "cg": out}, f)
print(len(out), "coefficients ->", path)
