"""Oracle: permutation matrix P123 in the wave-packet basis (test infrastructure only).

Restates:
  CPP/make_permutation_matrix.cpp:160-253   overlap mask and per-(q',p') phi meshes
  CPP/make_permutation_matrix.cpp:324-943   angular tables and the element loop (rows alpha' q' p',
                                            columns alpha q p; exact zeros dropped, line 824)
  CPP/auxiliary.cpp:453-462                 pi1/pi2 Jacobi momenta
  CPP/auxiliary.cpp:660-746                 WP-bin quadrature of one element
  CPP/auxiliary.cpp:950-1067                G-tilde
  CPP/General_functions/matrix_routines.cpp:222-244,287-326   column-major sort and CSC pointers
  CPP/make_permutation_matrix.cpp:1297-1332 x mesh = gauss(Nx) on [-1, 1]

The element loop is vectorised over the (phi, x) quadrature grid; the sum over
grid points is a numpy reduction rather than the reference's sequential loop,
so values agree to rounding, not bit for bit.

Pins (tests/test_oracle_pins.py): the S-wave special case G-tilde = A0/2 (the
closed form of CPP/auxiliary.cpp:991-1063 at L=l=L'=l'=0); for constant
G-tilde the element equals a brute-force Monte-Carlo/nested quadrature of the
bin-overlap volume; the mask is compared with a dense point test.
tests/test_oracle_recoupling.py: G-tilde for L > 0 against the rotation-invariant contraction
of two bipolar harmonics (complex spherical harmonics in randomly rotated frames) at the Jacobi
momenta pi1 = -p/2 - 3q/4, pi2 = p - q/2; the whole P123 against the symmetry P132 = P123^T =
P123 on antisymmetric states (the asymmetry falls >= 3x per doubling of Nphi = Nx).
"""
import math

import numpy as np

from . import angular as ang
from .states import gauss, gauss_ab


def pi1_tilde(p, q, x):
    return np.sqrt(0.25 * p * p + 9.0 / 16.0 * q * q + 0.75 * p * q * x)


def pi2_tilde(p, q, x):
    return np.sqrt(p * p + 0.25 * q * q - p * q * x)


def phi_mesh(pe, qe, Nphi):
    """Per bra bin (q', p'): GL on [atan(p'_l/q'_u), atan(p'_u/q'_l)] (lines 176-183)."""
    Np, Nq = len(pe) - 1, len(qe) - 1
    phi = np.zeros((Nq, Np, Nphi))
    wphi = np.zeros((Nq, Np, Nphi))
    for qp in range(Nq):
        for pp in range(Np):
            lo = math.atan(pe[pp] / qe[qp + 1])
            hi = math.atan(pe[pp + 1] / qe[qp]) if qe[qp] > 0 else math.pi / 2
            phi[qp, pp], wphi[qp, pp] = gauss_ab(Nphi, lo, hi)
    return phi, wphi


def overlap_mask(pe, qe, x, phi):
    """Boolean mask[q', p', q, p] (CPP/make_permutation_matrix.cpp:160-253)."""
    Np, Nq = len(pe) - 1, len(qe) - 1
    mask = np.zeros((Nq, Np, Nq, Np), dtype=bool)
    for qp in range(Nq):
        for pp in range(Np):
            lo = math.atan(pe[pp] / qe[qp + 1])
            hi = math.atan(pe[pp + 1] / qe[qp]) if qe[qp] > 0 else math.pi / 2
            if not lo < hi:
                continue
            s = np.sin(phi[qp, pp])[:, None]
            c = np.cos(phi[qp, pp])[:, None]
            with np.errstate(divide="ignore"):
                kmin = np.maximum(pe[pp] / s, qe[qp] / c)
                kmax = np.minimum(pe[pp + 1] / s, qe[qp + 1] / c)
            z1 = pi1_tilde(s, c, x[None, :])
            z2 = pi2_tilde(s, c, x[None, :])
            for q in range(Nq):
                for p in range(Np):
                    hl = np.maximum(pe[p] / z1, qe[q] / z2)
                    hu = np.minimum(pe[p + 1] / z1, qe[q + 1] / z2)
                    ok = (kmin < kmax) & (hl < hu) & (np.maximum(kmin, hl) < np.minimum(kmax, hu))
                    mask[qp, pp, q, p] = bool(ok.any())
    return mask


def gtilde(pw, ap, a, two_J, Atab, Lmax, s, c, x, tabs):
    """G-tilde_{alpha' alpha}(phi, x) (CPP/auxiliary.cpp:950-1067).

    ap: bra state (row), a: ket state (column), as called from
    CPP/make_permutation_matrix.cpp:771-788.  ``tabs`` = (Plm1(x), Plm2(cos th1), Plm3(cos th2)).
    """
    L12, l3 = int(pw.L[ap]), int(pw.l[ap])
    L12p, l3p = int(pw.L[a]), int(pw.l[a])
    P1, P2, P3 = tabs
    ret = np.zeros(np.broadcast(s, x).shape)
    Llo = max(abs(L12 - l3), abs(L12p - l3p))
    Lhi = min((two_J + 5) // 2, min(L12 + l3, L12p + l3p))
    for Lt in range(Llo, Lhi + 1):
        fac1 = 8.0 * math.pi * math.pi * Atab[ap, a, Lt]
        if fac1 == 0:
            continue
        for M in range(-min(l3, Lt), min(l3, Lt) + 1):
            fac2 = ang.clebsch_gordan(2 * L12, 2 * l3, 2 * Lt, 0, 2 * M, 2 * M) * (-1) ** (M % 2) \
                * ang.plm_signed(P1, l3, M)
            if np.all(fac2 == 0):
                continue
            for M12 in range(-L12p, L12p + 1):
                if abs(M - M12) > l3p:
                    continue
                cg = ang.clebsch_gordan(2 * L12p, 2 * l3p, 2 * Lt, 2 * M12, 2 * M - 2 * M12, 2 * M)
                if cg == 0:
                    continue
                ret = ret + fac1 * fac2 * cg * ang.plm_signed(P2, L12p, M12) * ang.plm_signed(P3, l3p, M - M12)
    return ret * math.sqrt((2.0 * L12 + 1) / (4 * math.pi))


def atilde_table(pw, two_J):
    """Atilde_store[alpha, alpha', Ltot] with Ltot <= Lmax (lines 444-465); zero unless T3 equal."""
    Na = pw.Nalpha
    Lmax = int(pw.L.max() + pw.l.max())
    A = np.zeros((Na, Na, Lmax + 1))
    for a in range(Na):
        for ap in range(Na):
            if pw.two_T3[a] == pw.two_T3[ap]:
                for Lt in range(Lmax + 1):
                    A[a, ap, Lt] = ang.atilde(pw, a, ap, Lt, two_J, int(pw.two_T3[a]))
    return A, Lmax


def p123_rows(pw, pe, qe, Nphi, Nx, row_bins=None):
    """P123 elements for the bra bins ``row_bins`` [(q', p'), ...] (all by default).

    Returns a dict {(row, col): value} with exact zeros dropped, plus the mask.
    """
    Np, Nq = len(pe) - 1, len(qe) - 1
    Na = pw.Nalpha
    two_J = int(pw.two_J3[0])
    x, wx = gauss(Nx)
    phi, wphi = phi_mesh(pe, qe, Nphi)
    mask = overlap_mask(pe, qe, x, phi)
    Atab, Lmax = atilde_table(pw, two_J)
    maxL, maxl = int(pw.L.max()), int(pw.l.max())
    P1 = ang.spharm_plm(maxl, x)[:, None, :]          # (n, 1, Nx) broadcast over phi
    out = {}
    if row_bins is None:
        row_bins = [(qp, pp) for qp in range(Nq) for pp in range(Np)]
    for qp, pp in row_bins:
        s = np.sin(phi[qp, pp])[:, None]
        c = np.cos(phi[qp, pp])[:, None]
        w2 = wphi[qp, pp][:, None] * wx[None, :]
        z1 = pi1_tilde(s, c, x[None, :])
        z2 = pi2_tilde(s, c, x[None, :])
        ct1 = np.clip(-(0.5 * s + 0.75 * c * x[None, :]) / z1, -1, 1)
        ct2 = np.clip((s - 0.5 * c * x[None, :]) / z2, -1, 1)
        tabs = (P1, ang.spharm_plm(maxL, ct1), ang.spharm_plm(maxl, ct2))
        cells = [(q, p) for q in range(Nq) for p in range(Np) if mask[qp, pp, q, p]]
        if not cells:
            continue
        # geometric weight per cell (lines 700-735 of CPP/auxiliary.cpp)
        W = np.zeros((len(cells),) + s.shape[:1] + x.shape)
        for k, (q, p) in enumerate(cells):
            kmin = np.maximum.reduce([pe[pp] / s + 0 * x, qe[qp] / c + 0 * x, pe[p] / z1, qe[q] / z2])
            kmax = np.minimum.reduce([pe[pp + 1] / s + 0 * x, qe[qp + 1] / c + 0 * x, pe[p + 1] / z1, qe[q + 1] / z2])
            g = w2 * c * s * (kmax * kmax - kmin * kmin) / (2 * z1 * z2)
            g[kmin > kmax] = 0.0
            norm = math.sqrt(qe[q + 1] - qe[q]) * math.sqrt(pe[p + 1] - pe[p]) \
                * math.sqrt(qe[qp + 1] - qe[qp]) * math.sqrt(pe[pp + 1] - pe[pp])
            W[k] = g / norm
        for ap in range(Na):
            for a in range(Na):
                G = gtilde(pw, ap, a, two_J, Atab, Lmax, s, c, x[None, :], tabs)
                vals = np.tensordot(W, G, axes=([1, 2], [0, 1]))
                row = ap * Nq * Np + qp * Np + pp
                for k, (q, p) in enumerate(cells):
                    if vals[k] != 0.0:
                        out[(row, a * Nq * Np + q * Np + p)] = float(vals[k])
    return out, mask


def to_csc(entries, dim):
    """COO dict -> CSC (val f64, row i32, colptr u64), rows ascending in each column
    (CPP/General_functions/matrix_routines.cpp:222-244,287-326; CPP/main.cpp:226-244)."""
    keys = sorted(entries.keys(), key=lambda rc: (rc[1], rc[0]))
    val = np.array([entries[k] for k in keys], dtype=np.float64)
    row = np.array([k[0] for k in keys], dtype=np.int32)
    col = np.array([k[1] for k in keys], dtype=np.int64)
    ptr = np.zeros(dim + 1, dtype=np.uint64)
    np.add.at(ptr, col + 1, 1)
    ptr = np.cumsum(ptr).astype(np.uint64)
    return val, row, ptr
