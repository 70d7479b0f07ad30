"""Oracle: diagonal channel resolvent in the SWP basis (test infrastructure only).

Restates CPP/make_resolvent.cpp:4-211 (Kukulin 2007 eqs. 32, 34-35):
  bound-continuum  R = ln|(E_q^l+Eb-E)/(E_q^u+Eb-E)|/dE_q - i pi [H(E_q^u+Eb-E)-H(E_q^l+Eb-E)]/dE_q
  continuum-continuum Q (log and Heaviside combination, lines 39-80)
with mu1 = Mn (Mn+Mp+Eb)/(2Mn+Mp+Eb), Eb = the p-bin's lower edge when negative, else 0.
Index order (alpha, q, p), as in the permutation matrix (line 201).

Pins (tests/test_oracle_pins.py): the closed forms are checked against a
brute-force bin average of 1/(E - e_p - e_q + i0) done by quadrature of the
principal value plus -i pi delta.
"""
import math

import numpy as np

from . import states as st


def heaviside(v):
    return 0.0 if v < 0 else 1.0


def resolvent_term_scale(E, Eb, q_u, q_l, e_u, e_l, bound):
    """Sum of the magnitudes of the terms the closed forms add up, divided by their
    denominators: the rounding of each term (a few ulp of its size) bounds how far two
    correct evaluations of the same formula may differ.  The continuum formula is a
    second difference of v ln|v| and loses up to ~10 digits for narrow bins far off shell."""
    m1 = st.Mn * (st.Mn + st.Mp + Eb) / (st.Mn + st.Mn + st.Mp + Eb)
    Eql = q_l * q_l / (2 * m1)
    Equ = q_u * q_u / (2 * m1)
    Dq = Equ - Eql
    if bound:
        return (abs(math.log(abs(Eql + Eb - E))) + abs(math.log(abs(Equ + Eb - E))) + math.pi) / Dq
    Dp = e_u - e_l
    Ep, Eq = 0.5 * (e_l + e_u), 0.5 * (Eql + Equ)
    DM, DP, D = 0.5 * (Dp - Dq), 0.5 * (Dp + Dq), Ep + Eq - E
    s = sum(abs(v) * (abs(math.log(abs(v))) + 1.0) for v in (D + DM, D - DM, D + DP, D - DP) if v != 0.0)
    return s / (Dp * Dq)


def resolvent_bound_continuum(E, Eb, q_u, q_l):
    """CPP/make_resolvent.cpp:14-36."""
    m1 = st.Mn * (st.Mn + st.Mp + Eb) / (st.Mn + st.Mn + st.Mp + Eb)
    Eql = q_l * q_l / (2 * m1)
    Equ = q_u * q_u / (2 * m1)
    Dq = Equ - Eql
    re = math.log(abs((Eql + Eb - E) / (Equ + Eb - E))) / Dq
    im = (heaviside(Equ + Eb - E) - heaviside(Eql + Eb - E)) * (-math.pi) / Dq
    return complex(re, im)


def _xlogx(v):
    return v * math.log(abs(v)) if v != 0.0 else 0.0


def resolvent_continuum_continuum(E, Eb, q_u, q_l, e_u, e_l):
    """CPP/make_resolvent.cpp:39-80."""
    m1 = st.Mn * (st.Mn + st.Mp + Eb) / (st.Mn + st.Mn + st.Mp + Eb)
    Eql = q_l * q_l / (2 * m1)
    Equ = q_u * q_u / (2 * m1)
    Ep = 0.5 * (e_l + e_u)
    Eq = 0.5 * (Eql + Equ)
    Dq = Equ - Eql
    Dp = e_u - e_l
    DM = 0.5 * (Dp - Dq)
    DP = 0.5 * (Dp + Dq)
    D = Ep + Eq - E
    # the reference evaluates v*log|v| literally; v == 0 gives nan there (0*-inf)
    re = ((D + DM) * math.log(abs(D + DM)) + (D - DM) * math.log(abs(D - DM))
          - (D + DP) * math.log(abs(D + DP)) - (D - DP) * math.log(abs(D - DP))) / (Dp * Dq)
    im = ((D + DM) * heaviside(D + DM) + (D - DM) * heaviside(D - DM)
          - (D + DP) * heaviside(D + DP) - (D - DP) * heaviside(D - DP)) * math.pi / (Dp * Dq)
    return complex(re, im)


def swp_edge_table(pw_chn, swp, tensor_force, isospin_breaking_1S0):
    """Per-alpha SWP energy edges (Np+1) chosen as in CPP/make_resolvent.cpp:118-156."""
    Np = swp.Np
    out = np.zeros((pw_chn.Nalpha, Np + 1))
    for a in range(pw_chn.Nalpha):
        L, S, J, T, t3 = (int(pw_chn.L[a]), int(pw_chn.S[a]), int(pw_chn.J[a]),
                          int(pw_chn.T[a]), int(pw_chn.two_T3[a]))
        s1S0 = (S == 0 and J == 0 and L == 0)
        via_L = tensor_force and L != J and J != 0
        via_T = s1S0 and isospin_breaking_1S0
        if via_L and via_T:
            raise ValueError("isospin breaking in a coupled channel")
        if via_L or via_T:
            idx = st.unique_2N_idx(L, S, J, T, True, tensor_force, isospin_breaking_1S0)
            upper = (L >= J) if via_L else (t3 != 1)
            off = Np + 1 if upper else 0
            out[a] = swp.e_coup[idx, off:off + Np + 1]
        else:
            idx = st.unique_2N_idx(L, S, J, T, False, tensor_force, isospin_breaking_1S0)
            out[a] = swp.e_unco[idx]
    return out


def resolvent_array(E, pw_chn, swp, tensor_force, isospin_breaking_1S0, with_scale=False):
    """G[(alpha, q, p)] at total energy E (CPP/make_resolvent.cpp:82-211); with_scale also
    returns resolvent_term_scale of every element."""
    Np, Nq = swp.Np, swp.Nq
    edges = swp_edge_table(pw_chn, swp, tensor_force, isospin_breaking_1S0)
    qe = swp.q_edges
    G = np.zeros(pw_chn.Nalpha * Nq * Np, dtype=complex)
    scale = np.zeros(len(G))
    for a in range(pw_chn.Nalpha):
        for p in range(Np):
            el, eu = edges[a, p], edges[a, p + 1]
            bound = el < 0
            Eb = el if bound else 0.0
            for q in range(Nq):
                if bound:
                    g = resolvent_bound_continuum(E, Eb, qe[q + 1], qe[q])
                else:
                    g = resolvent_continuum_continuum(E, Eb, qe[q + 1], qe[q], eu, el)
                G[a * Nq * Np + q * Np + p] = g
                if with_scale:
                    scale[a * Nq * Np + q * Np + p] = resolvent_term_scale(E, Eb, qe[q + 1], qe[q], eu, el, bound)
    return (G, scale) if with_scale else G
