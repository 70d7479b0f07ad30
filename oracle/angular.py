"""Oracle: angular-momentum algebra for P123 (test infrastructure only).

Restates the GSL calls the reference makes:
  gsl_sf_coupling_3j/6j/9j   (CPP/auxiliary.cpp:127-166, CPP/make_permutation_matrix.cpp:409-439)
  ClebschGordan              (CPP/auxiliary.cpp:127-135)
  gsl_sf_legendre_array_e(GSL_SF_LEGENDRE_SPHARM, ..., csphase=-1)
                             (CPP/make_permutation_matrix.cpp:544-589)
  Atilde                     (CPP/auxiliary.cpp:863-884)
Wigner symbols are exact: Racah's single-sum formulas in integer arithmetic,
the square root taken last.  Arguments are doubled (two_j), as in GSL.

Pins (tests/test_oracle_pins.py): 3j/6j/9j against sympy.physics.wigner;
Plm against scipy.special.sph_harm at phi = 0 (Condon-Shortley phase included).
tests/test_oracle_recoupling.py: Atilde against explicit overlaps of coupled states (LS-jj
recouplings by Clebsch-Gordan sums in the product basis, and the spin and isospin overlaps of
the cyclic permutation in the 8-dim space of three spin-1/2 particles), for channels with L, l
up to 4; the S-wave block against the textbook [[1/4, -3/4], [-3/4, 1/4]].
"""
import math
from fractions import Fraction
from functools import lru_cache

import numpy as np


def _f(n):
    return math.factorial(n)


def _tri_ok(a, b, c):
    """two_j arguments form a triangle with integer perimeter."""
    return (a + b + c) % 2 == 0 and abs(a - b) <= c <= a + b


def _delta2(a, b, c):
    """Delta(abc)^2 as a Fraction (two_j arguments)."""
    return Fraction(_f((a + b - c) // 2) * _f((a - b + c) // 2) * _f((-a + b + c) // 2), _f((a + b + c) // 2 + 1))


def _signed_sqrt(fr):
    """sign(fr) * sqrt(|fr|) as float from an exact rational."""
    if fr == 0:
        return 0.0
    s = 1.0 if fr > 0 else -1.0
    a = abs(fr)
    return s * math.sqrt(a.numerator) / math.sqrt(a.denominator)


@lru_cache(maxsize=None)
def wigner3j(tj1, tj2, tj3, tm1, tm2, tm3):
    """(j1 j2 j3; m1 m2 m3) with doubled arguments (Racah formula)."""
    if tm1 + tm2 + tm3 != 0:
        return 0.0
    if not _tri_ok(tj1, tj2, tj3):
        return 0.0
    for tj, tm in ((tj1, tm1), (tj2, tm2), (tj3, tm3)):
        if abs(tm) > tj or (tj + tm) % 2:
            return 0.0
    j1p, j1m = (tj1 + tm1) // 2, (tj1 - tm1) // 2
    j2p, j2m = (tj2 + tm2) // 2, (tj2 - tm2) // 2
    j3p, j3m = (tj3 + tm3) // 2, (tj3 - tm3) // 2
    pref = _delta2(tj1, tj2, tj3) * _f(j1p) * _f(j1m) * _f(j2p) * _f(j2m) * _f(j3p) * _f(j3m)
    a1 = (tj3 - tj2 + tm1) // 2
    a2 = (tj3 - tj1 - tm2) // 2
    b1 = (tj1 + tj2 - tj3) // 2
    b2 = j1m
    b3 = j2p
    kmin = max(0, -a1, -a2)
    kmax = min(b1, b2, b3)
    s = 0
    for k in range(kmin, kmax + 1):
        den = _f(k) * _f(b1 - k) * _f(b2 - k) * _f(b3 - k) * _f(a1 + k) * _f(a2 + k)
        s += Fraction((-1) ** k, den)
    phase = (-1) ** ((tj1 - tj2 - tm3) // 2)
    val2 = pref * s * s
    return phase * (1.0 if s >= 0 else -1.0) * _signed_sqrt(val2) if s != 0 else 0.0


def clebsch_gordan(tj1, tj2, tj3, tm1, tm2, tm3):
    """<j1 m1 j2 m2 | j3 m3> = (-1)^{j1-j2+m3} sqrt(2j3+1) (j1 j2 j3; m1 m2 -m3) (CPP/auxiliary.cpp:127-135)."""
    if abs(tm1) > tj1 or abs(tm2) > tj2:
        return 0.0
    w = wigner3j(tj1, tj2, tj3, tm1, tm2, -tm3)
    if w == 0.0:
        return 0.0
    return (-1) ** ((tj1 - tj2 + tm3) // 2) * math.sqrt(tj3 + 1) * w


@lru_cache(maxsize=None)
def wigner6j_exact(a, b, c, d, e, f):
    """{j1 j2 j3; j4 j5 j6} (doubled) as (sign, Fraction of the square), Racah formula."""
    for t in ((a, b, c), (a, e, f), (d, b, f), (d, e, c)):
        if not _tri_ok(*t):
            return Fraction(0)
    pref = _delta2(a, b, c) * _delta2(a, e, f) * _delta2(d, b, f) * _delta2(d, e, c)
    t1, t2, t3, t4 = (a + b + c) // 2, (a + e + f) // 2, (d + b + f) // 2, (d + e + c) // 2
    u1, u2, u3 = (a + b + d + e) // 2, (a + c + d + f) // 2, (b + c + e + f) // 2
    s = 0
    for k in range(max(t1, t2, t3, t4), min(u1, u2, u3) + 1):
        s += Fraction((-1) ** k * _f(k + 1),
                      _f(k - t1) * _f(k - t2) * _f(k - t3) * _f(k - t4) * _f(u1 - k) * _f(u2 - k) * _f(u3 - k))
    # value = s * sqrt(pref); return signed square
    return (1 if s >= 0 else -1) * s * s * pref


def wigner6j(a, b, c, d, e, f):
    return _signed_sqrt(wigner6j_exact(a, b, c, d, e, f))


@lru_cache(maxsize=None)
def wigner9j(a, b, c, d, e, f, g, h, i):
    """{a b c; d e f; g h i} (doubled) = sum_x (-1)^{2x} (2x+1) {a b c; f i x}{d e f; b x h}{g h i; x a d}."""
    xmin = max(abs(a - i), abs(h - d), abs(b - f))
    xmax = min(a + i, h + d, b + f)
    tot = 0.0
    # accumulate exactly where possible: each 6j is sign*sqrt(rational); multiply floats at the end
    for x in range(xmin, xmax + 1, 2):
        w1 = wigner6j(a, b, c, f, i, x)
        if w1 == 0.0:
            continue
        w2 = wigner6j(d, e, f, b, x, h)
        if w2 == 0.0:
            continue
        w3 = wigner6j(g, h, i, x, a, d)
        tot += (-1) ** x * (x + 1) * w1 * w2 * w3
    return tot


def spharm_plm(lmax, x):
    """Array Y_l^m(theta, 0) for 0 <= m <= l <= lmax at x = cos(theta), normalised as
    sqrt((2l+1)/4pi) sqrt((l-m)!/(l+m)!) P_l^m(x) with the (-1)^m Condon-Shortley phase,
    indexed l(l+1)/2 + m (GSL_SF_LEGENDRE_SPHARM layout).  x may be an ndarray."""
    x = np.asarray(x, dtype=float)
    out = np.zeros(((lmax + 1) * (lmax + 2) // 2,) + x.shape)
    s = np.sqrt(np.maximum(0.0, 1.0 - x * x))
    # normalised recurrences (stable): start with Y_m^m
    ymm = np.full(x.shape, 1.0 / math.sqrt(4 * math.pi))
    for m in range(lmax + 1):
        if m > 0:
            ymm = -math.sqrt((2 * m + 1) / (2.0 * m)) * s * ymm
        out[m * (m + 1) // 2 + m] = ymm
        if m + 1 <= lmax:
            y1 = math.sqrt(2 * m + 3) * x * ymm
            out[(m + 1) * (m + 2) // 2 + m] = y1
            ylm2, ylm1 = ymm, y1
            for l in range(m + 2, lmax + 1):
                a = math.sqrt((4.0 * l * l - 1) / (l * l - m * m))
                b = math.sqrt(((l - 1.0) ** 2 - m * m) / (4.0 * (l - 1) ** 2 - 1))
                yl = a * (x * ylm1 - b * ylm2)
                out[l * (l + 1) // 2 + m] = yl
                ylm2, ylm1 = ylm1, yl
    return out


def plm_signed(table, l, m):
    """Y_l^m at phi = 0 for signed m from a spharm_plm table (CPP/auxiliary.cpp:1002-1035)."""
    v = table[l * (l + 1) // 2 + abs(m)]
    if m < 0 and (-m) % 2 == 1:
        v = -v
    return v


def atilde(pw, a, ap, Ltot, two_J, two_T):
    """CPP/auxiliary.cpp:863-884 (alpha = a, alpha' = ap)."""
    L, S, J, l, tj, T = (int(pw.L[a]), int(pw.S[a]), int(pw.J[a]), int(pw.l[a]), int(pw.two_j[a]), int(pw.T[a]))
    Lp, Sp, Jp, lp, tjp, Tp = (int(pw.L[ap]), int(pw.S[ap]), int(pw.J[ap]), int(pw.l[ap]), int(pw.two_j[ap]), int(pw.T[ap]))
    ret = 0.0
    for two_S in (1, 3):
        ret += (math.sqrt((2 * J + 1) * (tj + 1) * (2 * S + 1) * (2 * T + 1)
                          * (2 * Jp + 1) * (tjp + 1) * (2 * Sp + 1) * (2 * Tp + 1))
                * (-1) ** (Sp + Tp)
                * (two_S + 1)
                * wigner9j(2 * L, 2 * S, 2 * J, 2 * l, 1, tj, 2 * Ltot, two_S, two_J)
                * wigner9j(2 * Lp, 2 * Sp, 2 * Jp, 2 * lp, 1, tjp, 2 * Ltot, two_S, two_J)
                * wigner6j(1, 1, 2 * S, 1, two_S, 2 * Sp)
                * wigner6j(1, 1, 2 * T, 1, two_T, 2 * Tp))
    return ret
