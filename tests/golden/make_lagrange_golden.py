"""Write lagrange_o25.npz: the order-25 Lagrange tables in exact rational arithmetic.

A(eps) = prod_{i=1}^{12} (i + eps)(i + 1 - eps) / (12! 13!) at eps = the float64 values
k * (1/1000), k = 0..1000 (the grid the response uses), and the barycentric factors
E_j = (-1)^j 12!/(12-j)! * 13!/(13+j)!, j = 1..12, each rounded once to float64.
Pure mathematics: no reference code is run.
"""
import os
from fractions import Fraction
from math import factorial

import numpy as np

H = 13


def main():
    deps = 1.0 / 1000
    eps = np.arange(1001) * deps
    den = factorial(H - 1) * factorial(H)
    A = []
    for e in eps:
        ef = Fraction(float(e))
        p = Fraction(1)
        for i in range(1, H):
            p *= (i + ef) * (i + 1 - ef)
        A.append(float(p / den))
    E = [float(Fraction((-1) ** j * factorial(H - 1) * factorial(H),
                        factorial(H - 1 - j) * factorial(H + j))) for j in range(1, H)]
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lagrange_o25.npz")
    np.savez(out, A_exact=np.array(A), E_exact=np.array(E))
    print(out)


if __name__ == "__main__":
    main()
