"""The folded stencil recurrence of the float32-output forms (flr_device.hpp:stencil_fold_c_lo),
restated in numpy and checked against the stencil's defining sum in exact rational arithmetic.

The fast stencil's sum S = sum_{j=1}^{h-1} E[j-1] (x[d+1+j]/(j+e) + x[d-j]/(j+1-e))
(LISAResponse.cu:149-207; E from response.py:300-310) is evaluated as a running fraction T/P;
the folded form divides each pair's Q_j = (j+e)(j+1-e) by E[j-1] so that the pair's weights are
(j+1-e) P and (j+e) P. A wrong index, sign or pairing in the recurrence moves S far from the
exact sum; the folded form must stay within FP64 rounding of it, like the unfolded one, and its
running product within the range the kernel relies on (|P| < 1e47 at order 25).
"""
from fractions import Fraction

import numpy as np

from fastlisaresponse_amd.tables import lagrange_tables

H = 13


def _exact(x, e, E):
    ef = Fraction(e)
    s = Fraction(0)
    for j in range(1, H):
        s += Fraction(E[j - 1]) * (Fraction(x[H + j]) / (j + ef) + Fraction(x[H - 1 - j]) / (j + 1 - ef))
    return s


def _unfolded(x, e, E):
    D = e * (1.0 - e)
    T = 0.0
    P = 1.0
    for j in range(1, H):
        Q = float(j * (j + 1)) + D
        EP = E[j - 1] * P
        wu = (j + 1) * EP - e * EP
        wd = j * EP + e * EP
        n = x[H + j] * wu + x[H - 1 - j] * wd
        T = n if j == 1 else T * Q + n
        P = P * Q
    return T / P


def _folded(x, e, E):
    D = e * (1.0 - e)
    Ei = 1.0 / E[:H - 1]
    Ci = np.array([j * (j + 1) for j in range(1, H)], dtype=np.float64) * Ei
    T = 0.0
    P = 1.0
    Pmax = 1.0
    for j in range(1, H):
        Q = D * Ei[j - 1] + Ci[j - 1]
        eP = e * P
        wu = (j + 1) * P - eP
        wd = j * P + eP
        n = x[H + j] * wu + x[H - 1 - j] * wd
        T = n if j == 1 else T * Q + n
        P = P * Q
        Pmax = max(Pmax, abs(P))
    return T / P, Pmax


def test_folded_recurrence_matches_the_exact_sum():
    _, _, E = lagrange_tables(25)
    rng = np.random.default_rng(7)
    worst_fold = worst_plain = 0.0
    for trial in range(300):
        x = rng.standard_normal(2 * H) * 10.0 ** rng.uniform(-22, 0)
        e = float(rng.uniform(0.0, 1.0)) if trial else 0.0
        exact = _exact(x, e, E)
        scale = float(np.sum(np.abs(x)))
        got, pmax = _folded(x, e, E)
        plain = _unfolded(x, e, E)
        worst_fold = max(worst_fold, abs(float(Fraction(got) - exact)) / scale)
        worst_plain = max(worst_plain, abs(float(Fraction(plain) - exact)) / scale)
        assert pmax < 1e47
    # both within a few FP64 rounding steps of the exact sum (on the scale of the taps)
    assert worst_fold < 1e-15, worst_fold
    assert worst_plain < 1e-15, worst_plain


def test_a_planted_index_error_is_caught():
    """The same recurrence with the lower tap of each pair shifted by one is far off."""
    _, _, E = lagrange_tables(25)
    rng = np.random.default_rng(8)
    x = rng.standard_normal(2 * H)
    e = 0.37
    exact = float(_exact(x, e, E))
    shifted = x.copy()
    shifted[:H - 1] = x[1:H]  # x[d-j] read as x[d-j+1]
    bad, _ = _folded(shifted, e, E)
    good, _ = _folded(x, e, E)
    assert abs(good - exact) < 1e-14 * np.sum(np.abs(x))
    assert abs(bad - exact) > 1e-3 * abs(exact)
