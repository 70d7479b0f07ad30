"""Pins of the oracle's angular algebra, 2N-basis embedding, isospin mixing and kinematics
(CPU only), each by a route independent of the restated formula:

  Atilde        (CPP/auxiliary.cpp:863-884)       explicit overlaps of coupled states, built
                                                 from Clebsch-Gordan sums in the product basis;
                                                 the S-wave block is the textbook recoupling
  G~ for L > 0  (CPP/auxiliary.cpp:950-1067)      the rotation-invariant contraction of two
                                                 bipolar harmonics, evaluated with complex
                                                 spherical harmonics in randomly rotated frames
                                                 at the Jacobi momenta of the cyclic permutation
  P123          (CPP/make_permutation_matrix.cpp) P132 = P123^T and P132 = P123 on states
                                                 antisymmetric in (23): the asymmetry falls with
                                                 the quadrature order
  C_op, VC_op   (CPP/solve_faddeev.cpp:1588-1927) the embedded SWP basis is orthogonal,
                                                 diagonalises the embedded H0 + V and carries
                                                 the deuteron mostly in its S-wave rows
  nn/np mixing  (CPP/make_potential_matrix.cpp:363-387)  isospin Clebsch-Gordan coefficients
  kinematics    (CPP/General_functions/kinetic_conversion.cpp:5-53)  Mandelstam s and a
                                                 Lorentz boost; the constants against PDG values
"""
import math
from functools import lru_cache

import numpy as np
import pytest
import scipy.special as ss
from sympy import S as Sym
from sympy.physics.wigner import clebsch_gordan

from oracle import angular as ang
from oracle import p123 as pm
from oracle import potential as pot
from oracle import solve as so
from oracle import states as st
from oracle import swp as sw


@lru_cache(maxsize=None)
def cg(two_j1, two_m1, two_j2, two_m2, two_j, two_m):
    """<j1 m1 j2 m2 | j m> (doubled arguments), sympy exact, as float."""
    if two_m1 + two_m2 != two_m or abs(two_m1) > two_j1 or abs(two_m2) > two_j2 or abs(two_m) > two_j:
        return 0.0
    h = lambda t: Sym(t) / 2
    return float(clebsch_gordan(h(two_j1), h(two_j2), h(two_j), h(two_m1), h(two_m2), h(two_m)))


def _ms(two_j):
    return range(-two_j, two_j + 1, 2)


# ------------------------------------------------------------------ Atilde
def ls_jj(L, S, J, l, tj, Lt, tSt, tJ):
    """<[(L S)J, (l 1/2)j] Jtot | [(L l)Lt, (S 1/2)St] Jtot>, explicit sum over the magnetic
    quantum numbers of the product basis |L mL>|S mS>|l ml>|1/2 ms> at M = Jtot."""
    tM = tJ
    tot = 0.0
    for mL in _ms(2 * L):
        for mS in _ms(2 * S):
            for ml in _ms(2 * l):
                for ms in _ms(1):
                    a = cg(2 * L, mL, 2 * S, mS, 2 * J, mL + mS) * cg(2 * l, ml, 1, ms, tj, ml + ms) \
                        * cg(2 * J, mL + mS, tj, ml + ms, tJ, tM)
                    if a == 0.0:
                        continue
                    b = cg(2 * L, mL, 2 * l, ml, 2 * Lt, mL + ml) * cg(2 * S, mS, 1, ms, tSt, mS + ms) \
                        * cg(2 * Lt, mL + ml, tSt, mS + ms, tJ, tM)
                    tot += a * b
    return tot


def three_half_state(order, two_pair, two_tot, two_M):
    """|[(s_a s_b)S, s_c] Stot M> of three spin-1/2 particles (or isospins) in the 8-dim
    product basis |m1 m2 m3>, for particles order = (a, b, c) (1-based)."""
    v = np.zeros(8)
    a, b, c = order
    for ma in (-1, 1):
        for mb in (-1, 1):
            for mc in (-1, 1):
                coef = cg(1, ma, 1, mb, two_pair, ma + mb) * cg(two_pair, ma + mb, 1, mc, two_tot, two_M)
                if coef == 0.0:
                    continue
                m = {a: ma, b: mb, c: mc}
                k = sum(((m[i] + 1) // 2) << (3 - i) for i in (1, 2, 3))
                v[k] += coef
    return v


def recoupling(two_S, two_Sp, two_tot):
    """<(23)S, 1; Stot | (31)S', 2; Stot>: the overlap the cyclic permutation P123 gives between
    pair-(23) states (particle 1 the spectator) and pair-(31) states."""
    return float(three_half_state((2, 3, 1), two_S, two_tot, two_tot) @
                 three_half_state((3, 1, 2), two_Sp, two_tot, two_tot))


def atilde_explicit(pw, a, ap, Lt, two_J, two_T):
    """(2 Lt + 1) Atilde[alpha, alpha', Lt] from explicit overlaps: the LS-jj recouplings of the two
    states into a common [(L l)Lt, (S 1/2)St] J basis, times the spin and isospin recouplings of
    the cyclic permutation, summed over St = 1/2, 3/2; divided by 2 Lt + 1."""
    L, S, J, l, tj, T = (int(pw.L[a]), int(pw.S[a]), int(pw.J[a]), int(pw.l[a]), int(pw.two_j[a]), int(pw.T[a]))
    Lp, Sp, Jp, lp, tjp, Tp = (int(pw.L[ap]), int(pw.S[ap]), int(pw.J[ap]), int(pw.l[ap]), int(pw.two_j[ap]),
                               int(pw.T[ap]))
    tot = 0.0
    for tSt in (1, 3):
        c1 = ls_jj(L, S, J, l, tj, Lt, tSt, two_J)
        c2 = ls_jj(Lp, Sp, Jp, lp, tjp, Lt, tSt, two_J)
        if c1 == 0.0 or c2 == 0.0:
            continue
        tot += c1 * c2 * recoupling(2 * S, 2 * Sp, tSt) * recoupling(2 * T, 2 * Tp, two_T)
    return tot / (2 * Lt + 1)


def test_atilde_swave_block_is_the_textbook_recoupling():
    """S waves (L = l = 0): the pair states (S, T) = (0, 1) and (1, 0) recouple under the cyclic
    permutation with [[1/4, -3/4], [-3/4, 1/4]] (the spin and isospin recouplings of three
    spin-1/2 particles, each [[-1/2, +-sqrt(3)/2], ...], multiplied)."""
    pw = st.construct_pw_states(1, 1, False, swave_only=True).channel(0)
    A, _ = pm.atilde_table(pw, 1)
    assert [(int(pw.S[k]), int(pw.T[k])) for k in range(2)] == [(0, 1), (1, 0)]
    assert np.allclose(A[:, :, 0], [[0.25, -0.75], [-0.75, 0.25]], rtol=0, atol=1e-14)
    assert abs(recoupling(0, 0, 1) + 0.5) < 1e-14 and abs(abs(recoupling(0, 2, 1)) - math.sqrt(3) / 2) < 1e-14


@pytest.mark.parametrize("J2max,two_J", [(2, 1), (2, 3), (3, 5)])
def test_atilde_equals_explicit_recoupling_overlaps(J2max, two_J):
    """Every Atilde[alpha, alpha', Lt] of a channel with L, l up to 4 equals the explicit overlap
    construction; a transposed table fails (Atilde[a', a] = (-1)^(L - L') Atilde[a, a'] for
    antisymmetric pair states, and pairs with odd L - L' occur)."""
    pwf = st.construct_pw_states(J2max, two_J)
    c = [k for k in range(pwf.N_chn) if int(pwf.channel(k).two_J3[0]) == two_J][0]
    pw = pwf.channel(c)
    A, Lmax = pm.atilde_table(pw, two_J)
    odd = 0
    for a in range(pw.Nalpha):
        for ap in range(pw.Nalpha):
            if pw.two_T3[a] != pw.two_T3[ap]:
                continue
            odd += (int(pw.L[a]) - int(pw.L[ap])) % 2
            for Lt in range(Lmax + 1):
                ref = atilde_explicit(pw, a, ap, Lt, two_J, int(pw.two_T3[a]))
                assert abs(A[a, ap, Lt] - ref) <= 1e-12, (a, ap, Lt, A[a, ap, Lt], ref)
    assert odd > 0


# ------------------------------------------------------------------ G~ for L > 0
def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    return q if np.linalg.det(q) > 0 else -q


def _angles(v):
    v = v / np.linalg.norm(v)
    return math.acos(max(-1.0, min(1.0, v[2]))), math.atan2(v[1], v[0])


def _bipolar(L, l, Lp, lp, Lt, va, vb, vc, vd):
    """sum_M [Y_L(va) x Y_l(vb)]^{Lt M *} [Y_L'(vc) x Y_l'(vd)]^{Lt M} with complex spherical
    harmonics (scipy, Condon-Shortley phase); rotation invariant."""
    def Y(n, m, v):
        th, ph = _angles(v)
        return complex(ss.sph_harm_y(n, m, th, ph))
    tot = 0.0 + 0.0j
    for M in range(-Lt, Lt + 1):
        b1 = sum(cg(2 * L, 2 * m, 2 * l, 2 * (M - m), 2 * Lt, 2 * M) * Y(L, m, va) * Y(l, M - m, vb)
                 for m in range(-L, L + 1) if abs(M - m) <= l)
        b2 = sum(cg(2 * Lp, 2 * m, 2 * lp, 2 * (M - m), 2 * Lt, 2 * M) * Y(Lp, m, vc) * Y(lp, M - m, vd)
                 for m in range(-Lp, Lp + 1) if abs(M - m) <= lp)
        tot += np.conj(b1) * b2
    return tot


def test_gtilde_is_the_bipolar_contraction_at_the_jacobi_momenta():
    """G~_{a' a}(p, q, x) = 8 pi^2 sum_Lt Atilde[a', a, Lt] B_Lt with B the rotation-invariant
    contraction of [Y_L(p) Y_l(q)]^Lt (bra a') and [Y_L'(pi1) Y_l'(pi2)]^Lt (ket a) at
    p = p z, q = q (-sqrt(1 - x^2), 0, x), pi1 = -p/2 - 3q/4, pi2 = p - q/2 (the Jacobi momenta
    of the cyclic permutation), evaluated in random rotated frames.  Catches a lost (-1)^M
    phase, a bra/ket swap, a wrong Clebsch-Gordan order or the wrong Jacobi relation."""
    rng = np.random.default_rng(5)
    pw = st.construct_pw_states(2, 1).channel(0)
    two_J = 1
    A, Lmax = pm.atilde_table(pw, two_J)
    maxL, maxl = int(pw.L.max()), int(pw.l.max())
    pts = [(0.7, 0.4, -0.3), (0.35, 0.9, 0.55), (1.1, 0.2, 0.8)]
    pairs = [(ap, a) for ap in range(pw.Nalpha) for a in range(pw.Nalpha)
             if pw.two_T3[ap] == pw.two_T3[a] and (pw.L[ap] + pw.l[ap]) > 0 and (pw.L[a] + pw.l[a]) > 0]
    pairs = pairs[::max(1, len(pairs) // 12)]
    assert len(pairs) >= 8
    for p, q, x in pts:
        s_, c_ = np.array([[p]]), np.array([[q]])
        xa = np.array([[x]])
        z1 = pm.pi1_tilde(s_, c_, xa)
        z2 = pm.pi2_tilde(s_, c_, xa)
        ct1 = np.clip(-(0.5 * s_ + 0.75 * c_ * xa) / z1, -1, 1)
        ct2 = np.clip((s_ - 0.5 * c_ * xa) / z2, -1, 1)
        tabs = (ang.spharm_plm(maxl, xa), ang.spharm_plm(maxL, ct1), ang.spharm_plm(maxl, ct2))
        pv = np.array([0.0, 0.0, p])
        qv = q * np.array([-math.sqrt(1 - x * x), 0.0, x])
        pi1 = -0.5 * pv - 0.75 * qv
        pi2 = pv - 0.5 * qv
        R = _rotation(rng)
        for ap, a in pairs:
            g = float(pm.gtilde(pw, ap, a, two_J, A, Lmax, s_, c_, xa, tabs)[0, 0])
            L, l, Lp, lp = int(pw.L[ap]), int(pw.l[ap]), int(pw.L[a]), int(pw.l[a])
            ref = 0.0
            for Lt in range(Lmax + 1):
                if A[ap, a, Lt] != 0.0:
                    ref += 8 * math.pi ** 2 * A[ap, a, Lt] * _bipolar(L, l, Lp, lp, Lt, R @ pv, R @ qv, R @ pi1,
                                                                       R @ pi2).real
            assert abs(g - ref) <= 1e-12 * max(1.0, abs(ref)), (ap, a, p, q, x, g, ref)


# ------------------------------------------------------------------ P123 symmetry
def test_p123_is_symmetric_on_antisymmetric_states_in_the_quadrature_limit():
    """P132 = P123^T, and on states antisymmetric in (23) P132 = P23 P123 P23 = P123, so P123 is
    symmetric up to quadrature error: ||P - P^T|| / ||P|| falls at least 3x per doubling of
    Nphi = Nx (16, 32, 64), in the whole matrix and in its L + l > 0 block.  Dropping the
    (-1)^M phase of G~ leaves an asymmetry of order one."""
    N = 4
    fwp = st.make_fwp(N, N, st.chebyshev_scale_for_top(N, 8 * st.hbarc))
    pw = st.construct_pw_states(1, 1).channel(0)
    dim = pw.Nalpha * N * N
    lpos = np.repeat((pw.L + pw.l) > 0, N * N)
    blk = np.ix_(lpos, lpos)
    full, high = [], []
    for n in (16, 32, 64):
        ent, _ = pm.p123_rows(pw, fwp.p_edges, fwp.q_edges, n, n)
        P = so.p_matrix(*pm.to_csc(ent, dim), dim).toarray()
        D = P - P.T
        full.append(np.linalg.norm(D) / np.linalg.norm(P))
        high.append(np.linalg.norm(D[blk]) / np.linalg.norm(P[blk]))
    for seq in (full, high):
        assert seq[0] < 0.05
        assert seq[0] / seq[1] >= 3.0 and seq[1] / seq[2] >= 3.0, seq


# ------------------------------------------------------------------ embedding of the SWP basis
def _synthetic_small(N=6, J2max=3):
    fwp = st.make_fwp(N, N, st.chebyshev_scale_for_top(N, 8 * st.hbarc))
    pwf = st.construct_pw_states(J2max, 1)
    theta, beta = pot.synthetic_theta()
    Vu, Vc = sw.potential_matrices(fwp, pwf, pot.Synthetic(theta, beta), True, True, J2max)
    swp = sw.make_swp_states(fwp, pwf, Vu, Vc, True, True)
    return fwp, pwf, Vu, Vc, swp


def test_embedded_swp_basis_is_orthogonal_and_diagonalises_H():
    """C_op and VC_op on the (alpha, p) space of the 1/2+ channel (tensor force, isospin-breaking
    1S0): C_op^T C_op = 1; C_op^T (H0_op + V_op) C_op is diagonal with the SWP eigenvalues (H0
    per bin (p_l^2 + p_l p_u + p_u^2) / (6 mu23), CPP/make_swp_states.cpp:38); VC_op = V_op C_op;
    and the deuteron column (the first 3S1 alpha, p = 0) holds more than 90% of its norm in the
    L = 0 rows, with E_bound as its diagonal element.  A swapped half of a coupled block puts
    another eigenvector in that column."""
    fwp, pwf, Vu, Vc, swp = _synthetic_small()
    pw = pwf.channel(0)
    Na, Np = pw.Nalpha, fwp.Np
    C, VC = so.c_and_vc(pw, swp, Vu, Vc, True, True)
    Vop = so.embed_2N(pw, swp, Vu, Vc, True, True)
    Cm, VCm, Vm = (M.reshape(Na * Np, Na * Np) for M in (C, VC, Vop))
    assert np.abs(Cm.T @ Cm - np.eye(Na * Np)).max() < 1e-12
    pe = fwp.p_edges
    h0 = (pe[:-1] ** 2 + pe[:-1] * pe[1:] + pe[1:] ** 2) / (6 * st.mu23)
    H = np.diag(np.tile(h0, Na)) + Vm
    D = Cm.T @ H @ Cm
    off = D - np.diag(np.diag(D))
    assert np.abs(off).max() < 1e-9 * np.abs(D).max()
    assert np.abs(VCm - Vm @ Cm).max() < 1e-12 * np.abs(VCm).max()
    # the state (alpha_d, p = 0) of the deuteron alpha (L = 0, S = 1, J = 1, T = 0) is the bound
    # state (CPP/solve_faddeev.cpp:229-233): its diagonal element is E_bound
    deut = [a for a in range(Na) if (pw.L[a], pw.S[a], pw.J[a], pw.T[a]) == (0, 1, 1, 0)][0]
    assert abs(D[deut * Np, deut * Np] - swp.E_bound) < 1e-9 * abs(swp.E_bound)
    col = C[:, :, deut, 0]
    wS = sum(np.sum(col[a] ** 2) for a in range(Na) if pw.L[a] == 0)
    assert abs(np.sum(col ** 2) - 1.0) < 1e-12 and wS > 0.9, wS


# ------------------------------------------------------------------ isospin mixing
class _TwoLevel:
    """A potential that is a constant in every slot, vnn for Tz = -1 and vnp for Tz = 0."""

    def __init__(self, vnn, vnp):
        self.vnn, self.vnp = vnn, vnp

    def V(self, p_in, p_out, coupled, S, J, T, Tz):
        return [self.vnn if Tz == -1 else self.vnp] * 6


def test_isospin_mixing_coefficients_from_clebsch_gordan():
    """The 3N isospin states |(t23 = 1) t1; T3N, M = -1/2> hold the nn pair (m23 = -1) and the np
    pair (m23 = 0) with weights |<1 m23 1/2 M-m23 | T3N M>|^2: 2/3 nn + 1/3 np for T3N = 1/2,
    1/3 nn + 2/3 np for T3N = 3/2, and +-sqrt(2)/3 (nn - np) between them (the product of the
    two coefficients).  The sign between the T3N = 1/2 and 3/2 states is a convention: the
    reference's + is the coupling order spectator first, (1/2 x 1); pair first gives -.  P123
    does not connect different T3N (Atilde vanishes unless the T3N are equal), so the sign
    cancels in every on-shell U (test_isospin_mixing_sign_is_a_convention).  The oracle's V_WP
    of a constant potential carries exactly these weights times the bin integrals."""
    N = 3
    fwp = st.make_fwp(N, N, st.chebyshev_scale_for_top(N, 8 * st.hbarc))
    pwf = st.construct_pw_states(1, 1)
    vnn, vnp = 3.0, 1.0
    Vu, Vc = sw.potential_matrices(fwp, pwf, _TwoLevel(vnn, vnp), True, True, 1)
    Vu1, Vc1 = sw.potential_matrices(fwp, pwf, _TwoLevel(1.0, 1.0), True, True, 1)
    def w2(tT):
        return {m: cg(2, 2 * m, 1, -1 - 2 * m, tT, -1) ** 2 for m in (-1, 0)}
    def w(tT):   # spectator first: <1/2 M-m23, 1 m23 | T M>
        return {m: cg(1, -1 - 2 * m, 2, 2 * m, tT, -1) for m in (-1, 0)}
    half, three = w2(1), w2(3)
    assert abs(half[-1] - 2 / 3) < 1e-15 and abs(three[-1] - 1 / 3) < 1e-15
    mix = w(1)[-1] * w(3)[-1] * vnn + w(1)[0] * w(3)[0] * vnp
    assert abs(mix - math.sqrt(2) / 3 * (vnn - vnp)) < 1e-14
    # the isospin-breaking 1S0 block (index of the coupled 1S0 matrix)
    idx = st.unique_2N_idx(0, 0, 0, 1, True, True, True)
    B, B1 = Vc[idx], Vc1[idx]
    Np = N
    assert np.allclose(B[:Np, :Np], (half[-1] * vnn + half[0] * vnp) * B1[:Np, :Np], rtol=1e-14)
    assert np.allclose(B[Np:, Np:], (three[-1] * vnn + three[0] * vnp) * B1[Np:, Np:], rtol=1e-14)
    assert np.allclose(B[:Np, Np:], mix * B1[:Np, :Np], rtol=1e-14)
    assert np.allclose(B[Np:, :Np], mix * B1[:Np, :Np], rtol=1e-14)
    # an isoscalar T = 1 channel (3P0) gets the T3N = 1/2 weights
    k = st.unique_2N_idx(1, 1, 0, 1, False, True, True)
    assert np.allclose(Vu[k], (half[-1] * vnn + half[0] * vnp) * Vu1[k], rtol=1e-14)


# ------------------------------------------------------------------ kinematics
def test_kinematics_against_mandelstam_s_and_a_boost():
    """q_com(T_lab) for a nucleon of mass Mp on a deuteron at rest (Md = Mn + Mp + Eb): from the
    invariant s = (E_lab + Md)^2 - p_lab^2, p* = p_lab Md / sqrt(s), and by boosting the
    projectile's four-momentum into the centre-of-mass frame; E_com = 3 q^2 / (4 MN) is the
    kinetic energy q^2 / (2 mu) of the reduced mass mu = 2 MN / 3.  The masses and hbar c are
    the PDG values to the digits the reference keeps."""
    assert abs(st.Mp - 938.27208816) < 1e-3 and abs(st.Mn - 939.56542052) < 1e-3
    assert abs(st.hbarc - 197.3269804) < 1e-3
    Eb = -2.2241
    Md = st.Mn + st.Mp + Eb
    for T in (0.5, 1.0, 13.0, 57.86, 200.0):
        E = T + st.Mp
        plab = math.sqrt(E * E - st.Mp ** 2)
        s = (E + Md) ** 2 - plab ** 2
        pstar = plab * Md / math.sqrt(s)
        assert abs(st.lab_energy_to_com_momentum(T, Eb) - pstar) <= 1e-12 * pstar
        beta = plab / (E + Md)
        gam = 1.0 / math.sqrt(1 - beta * beta)
        pz = gam * (plab - beta * E)   # the projectile in the c.m. frame
        assert abs(abs(pz) - pstar) <= 1e-9 * pstar
        assert abs(st.com_momentum_to_lab_energy(pstar, Eb) - T) <= 1e-9 * T
        assert abs(st.com_q_momentum_to_com_energy(pstar) - pstar ** 2 / (2 * (2 * st.MN / 3))) <= 1e-12 * T


def test_isospin_mixing_sign_is_a_convention():
    """Flipping the sign of the T3N = 1/2 <-> 3/2 block of the isospin-breaking 1S0 potential
    (the other coupling order's convention) leaves the on-shell U of the oracle's dense solve
    unchanged to rounding, on the 1/2+ channel with tensor force at Np = Nq = 6."""
    from oracle import resolvent as rs
    N = 6
    fwp = st.make_fwp(N, N, st.chebyshev_scale_for_top(N, 8 * st.hbarc))
    pwf = st.construct_pw_states(1, 1)
    theta, beta = pot.synthetic_theta()
    Vu, Vc = sw.potential_matrices(fwp, pwf, pot.Synthetic(theta, beta), True, True, 1)
    pw = pwf.channel(0)
    dim = pw.Nalpha * N * N
    ent, _ = pm.p123_rows(pw, fwp.p_edges, fwp.q_edges, 16, 16)
    P = so.p_matrix(*pm.to_csc(ent, dim), dim)
    idx = st.unique_2N_idx(0, 0, 0, 1, True, True, True)
    Us = []
    for sign in (1.0, -1.0):
        Vc2 = Vc.copy()
        Vc2[idx, :N, N:] *= sign
        Vc2[idx, N:, :N] *= sign
        swp = sw.make_swp_states(fwp, pwf, Vu, Vc2, True, True)
        os_ = sw.find_on_shell_bins(fwp.q_edges, swp.E_bound, [10.0])
        deut = sw.find_deuteron_channels(pw)
        G = np.stack([rs.resolvent_array(e, pw, swp, True, True) for e in os_.E_com + swp.E_bound])
        A = so.KernelA(pw, swp, Vu, Vc2, P, True, True).dense()
        Us.append(so.dense_solve(A, G, deut, os_.q_idx, N, N))
    assert np.abs(Us[0] - Us[1]).max() <= 1e-12 * np.abs(Us[0]).max()
