"""Pins of the oracle against library routines, closed forms and brute force.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign, a transposed operand, a wrong Jacobian) fails it.
"""
import math

import mpmath
import numpy as np
import pytest
import scipy.integrate as si
import scipy.special as ss
from sympy.physics.wigner import wigner_3j, wigner_6j, wigner_9j

from oracle import angular as ang
from oracle import p123 as pm
from oracle import potential as pot
from oracle import resolvent as rs
from oracle import solve as so
from oracle import states as st
from oracle import swp as sw


# ------------------------------------------------------------------ states
@pytest.mark.parametrize("n", [1, 2, 5, 8, 16, 48])
def test_gauss_legendre_matches_numpy(n):
    x, w = st.gauss(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    assert np.allclose(x, xr, atol=1e-15, rtol=0)
    assert np.allclose(w, wr, atol=1e-14, rtol=0)


def test_chebyshev_edges_closed_form():
    N, s = 30, st.chebyshev_scale_for_top(30, 8 * st.hbarc)
    e = st.chebyshev_edges(N, s)
    assert e[0] == 0.0 and abs(e[-1] - 8 * st.hbarc) < 1e-9
    i = np.arange(1, N + 1)
    assert np.allclose(e[1:], s * np.tan((2 * i - 1) * np.pi / (4 * N)), rtol=1e-15)


def test_kinematics_round_trip():
    Eb = -2.2245
    for T in (1.0, 13.0, 200.0):
        q = st.lab_energy_to_com_momentum(T, Eb)
        assert abs(st.com_momentum_to_lab_energy(q, Eb) - T) < 1e-9 * T
        assert abs(st.com_energy_to_com_q_momentum(st.com_q_momentum_to_com_energy(q)) - q) < 1e-12 * q


def test_pw_channel_sizes():
    # N_alpha per J^Pi for J2max = 3 without / with isospin-breaking 1S0 (T3 = 3/2 adds one 1S0 state)
    pw0 = st.construct_pw_states(3, 17, False)
    pw1 = st.construct_pw_states(3, 17, True)
    sizes0 = [pw0.channel(c).Nalpha for c in range(pw0.N_chn)]
    sizes1 = [pw1.channel(c).Nalpha for c in range(pw1.N_chn)]
    assert sizes0[:2] == [26, 26] and max(sizes0) == 62 and len(sizes0) == 18
    assert sizes1 == [s + 1 for s in sizes0]
    # every state obeys the triangle, Pauli and parity rules
    for a in range(pw1.Nalpha):
        L, S, J, T, l, tj, tJ, tT, P = (int(pw1.L[a]), int(pw1.S[a]), int(pw1.J[a]), int(pw1.T[a]), int(pw1.l[a]),
                                        int(pw1.two_j[a]), int(pw1.two_J3[a]), int(pw1.two_T3[a]), int(pw1.P3[a]))
        assert abs(J - S) <= L <= J + S and (L + S + T) % 2 == 1
        assert abs(2 * l - 1) <= tj <= 2 * l + 1 and abs(tJ - 2 * J) <= tj <= tJ + 2 * J
        assert P == (1 if (L + l) % 2 == 0 else -1)


def test_unique_2N_index_without_tensor_force_overflows_beyond_J1():
    # reference quirk (CPP/make_pw_symm_states.cpp:50): J*(4*(J>1)+2) skips indices for J >= 2,
    # so J2max >= 2 without a tensor force addresses past num_2N_unco_states
    nu, _ = st.num_2N_states(2, False, False)
    assert st.unique_2N_idx(2, 1, 2, 0, False, False, False) >= nu


@pytest.mark.parametrize("tf,ib,J2", [(True, True, 3), (True, False, 3), (False, False, 1)])
def test_unique_2N_index_is_a_bijection(tf, ib, J2):
    nu, nc = st.num_2N_states(J2, tf, ib)
    seen_u, seen_c = set(), set()
    for J in range(J2 + 1):
        for S in range(2):
            for L in range(abs(J - S), J + S + 1):
                for T in range(2):
                    if (L + S + T) % 2 == 0:
                        continue
                    coupled = (tf and L != J and J != 0) or (S == 0 and J == 0 and L == 0 and ib)
                    if coupled and L > J:
                        continue  # the partner of a coupled pair shares its index
                    u = st.unique_2N_idx(L, S, J, T, coupled, tf, ib)
                    (seen_c if coupled else seen_u).add(u)
    assert seen_u == set(range(nu)) and seen_c == set(range(nc))


# --------------------------------------------------------------- potential
@pytest.mark.parametrize("L", [1, 2, 3, 4])
@pytest.mark.parametrize("z", [1.0004, 1.05, 1.7, 12.0, 300.0])
def test_legendre_Q_matches_mpmath(L, z):
    ref = float(mpmath.legenq(L, 0, z, type=3).real)
    assert abs(pot.legendre_Q(L, z) - ref) <= 1e-13 * abs(ref)


@pytest.mark.parametrize("L", [0, 1, 2])
def test_yukawa_projection_against_radial_integral(L):
    """(2/pi) int r^2 j_L(p'r) e^{-mu r}/r j_L(p r) dr = Q_L(z)/(pi p p')."""
    p, pp, mu = 1.3, 0.7, 0.9
    f = lambda r: (2 / np.pi) * r * ss.spherical_jn(L, pp * r) * np.exp(-mu * r) * ss.spherical_jn(L, p * r)
    ref = si.quad(f, 0, 80, limit=400, epsabs=1e-14)[0]
    assert abs(pot.yukawa_pw(L, pp, p, mu) - ref) < 1e-10


def test_mt_is_the_L0_yukawa_pair():
    mt = pot.MalflietTjon()
    p, pp = 150.0, 310.0
    y = lambda V, mu: V / st.hbarc * pot.yukawa_pw(0, pp, p, mu * st.hbarc)
    ref1 = y(1438.72, 3.11) - y(626.885, 1.55)
    assert abs(mt.V(p, pp, False, 1, 1, 0, 0)[2] - ref1) < 1e-14 * abs(ref1)
    ref0 = y(1438.72, 3.11) - y(513.968, 1.55)
    assert abs(mt.V(p, pp, False, 0, 0, 1, 0)[0] - ref0) < 1e-14 * abs(ref0)


# --------------------------------------------------------------------- SWP
def _mt_small(N=12):
    pw = st.construct_pw_states(1, 1, False, swave_only=True)
    fwp = st.make_fwp(N, N, st.chebyshev_scale_for_top(N, 8 * st.hbarc))
    Vu, Vc = sw.potential_matrices(fwp, pw, pot.MalflietTjon(), False, False, 1)
    return pw, fwp, Vu, Vc


def test_V_WP_element_against_dblquad():
    pw, fwp, Vu, _ = _mt_small()
    mt = pot.MalflietTjon()
    i, j = 3, 5
    pe = fwp.p_edges
    g = lambda po, pi: pi * po * mt.V(pi, po, False, 1, 1, 0, 0)[2]
    ref = si.dblquad(g, pe[j], pe[j + 1], pe[i], pe[i + 1], epsabs=1e-13, epsrel=1e-11)[0]
    ref /= math.sqrt(pe[i + 1] - pe[i]) * math.sqrt(pe[j + 1] - pe[j])
    idx = st.unique_2N_idx(0, 1, 1, 0, False, False, False)
    assert abs(Vu[idx, i, j] - ref) < 1e-8 * abs(ref)


def test_swp_eigen_invariants_and_variational_bound():
    pw, fwp, Vu, Vc = _mt_small()
    s = sw.make_swp_states(fwp, pw, Vu, Vc, False, False)
    idx = st.unique_2N_idx(0, 1, 1, 0, False, False, False)
    C = s.C_unco[idx]
    pe = fwp.p_edges
    H = Vu[idx] + np.diag((pe[1:] ** 2 + pe[1:] * pe[:-1] + pe[:-1] ** 2) / (6 * st.mu23))
    assert np.allclose(C.T @ C, np.eye(len(C)), atol=1e-13)
    D = C.T @ H @ C
    assert np.allclose(D - np.diag(np.diag(D)), 0, atol=1e-10 * np.abs(D).max())
    # Galerkin bound: WP energy above the converged MT-III value -2.2305 MeV (Nystrom, 80 points)
    assert -2.2305 < s.E_bound < 0
    pw, fwp2, Vu2, Vc2 = _mt_small(24)
    s2 = sw.make_swp_states(fwp2, pw, Vu2, Vc2, False, False)
    assert -2.2305 < s2.E_bound < s.E_bound
    # SWP edges: midpoints of consecutive eigenvalues; 3S1 bin 0 = [E_b, 0]
    ev = np.sort(np.diag(D))
    e = s.e_unco[idx]
    assert e[0] == s.E_bound and e[1] == 0.0
    assert np.allclose(e[2:-1], 0.5 * (ev[1:-1] + ev[2:]), rtol=1e-12)


# ---------------------------------------------------------------- resolvent
def _box_average(E, ep, dp, eq, dq, n=4000):
    """(1/(dp dq)) PV int int dx dy/(E - x - y) over the box, by the line-integral in y."""
    # inner integral over x analytic: int dx/(E-x-y) = -ln|E - x - y|
    y = eq - dq / 2 + (np.arange(n) + 0.5) * dq / n
    a, b = ep - dp / 2, ep + dp / 2
    inner = -np.log(np.abs(E - b - y)) + np.log(np.abs(E - a - y))
    return inner.mean() / dp


def test_resolvent_cc_against_box_average():
    # continuum-continuum bin with E away from and inside the box
    mu1 = st.Mn * (st.Mn + st.Mp) / (2 * st.Mn + st.Mp)
    ql, qu, el, eu = 40.0, 55.0, 3.0, 4.5
    Eql, Equ = ql ** 2 / (2 * mu1), qu ** 2 / (2 * mu1)
    for E in (0.5, 20.0, 0.5 * (el + eu) + 0.5 * (Eql + Equ) + 0.1):
        g = rs.resolvent_continuum_continuum(E, 0.0, qu, ql, eu, el)
        ref_re = _box_average(E, 0.5 * (el + eu), eu - el, 0.5 * (Eql + Equ), Equ - Eql)
        assert abs(g.real - ref_re) < 1e-5 * max(1.0, abs(ref_re))
        # imaginary part: -pi times the box-averaged delta(E - x - y) (length of the diagonal in the box)
        xs = np.linspace(el, eu, 200001)
        ys = E - xs
        inside = (ys >= Eql) & (ys <= Equ)
        ref_im = -math.pi * inside.mean() * (eu - el) / ((eu - el) * (Equ - Eql))
        assert abs(g.imag - ref_im) < 1e-4 * max(1e-3, abs(ref_im))


def test_resolvent_bc_closed_form():
    Eb, ql, qu = -2.2, 30.0, 41.0
    mu1 = st.mu1(Eb)
    Eql, Equ = ql ** 2 / (2 * mu1), qu ** 2 / (2 * mu1)
    E = Eb + 0.5 * (Eql + Equ)
    g = rs.resolvent_bound_continuum(E, Eb, qu, ql)
    x = Eql + (np.arange(2_000_000) + 0.5) * (Equ - Eql) / 2_000_000
    ref = np.mean(1.0 / (E - Eb - x))  # symmetric PV, midpoint rule
    assert abs(g.real - ref) < 1e-4 and abs(g.imag + math.pi / (Equ - Eql)) < 1e-15


# ----------------------------------------------------------------- angular
@pytest.mark.parametrize("args", [(2, 2, 4, 0, 2, -2), (4, 2, 2, 2, -2, 0), (3, 1, 2, 1, 1, -2), (6, 4, 4, -2, 2, 0),
                                  (5, 3, 4, 3, -1, -2)])
def test_wigner3j_sympy(args):
    ref = float(wigner_3j(*[sympy_half(a) for a in args]))
    assert abs(ang.wigner3j(*args) - ref) < 1e-15


def sympy_half(a):
    import sympy
    return sympy.Rational(a, 2)


@pytest.mark.parametrize("args", [(1, 1, 2, 1, 3, 2), (1, 1, 0, 1, 1, 2), (4, 2, 2, 4, 2, 6), (6, 5, 3, 4, 3, 5),
                                  (8, 6, 4, 6, 4, 6)])
def test_wigner6j_sympy(args):
    ref = float(wigner_6j(*[sympy_half(a) for a in args]))
    assert abs(ang.wigner6j(*args) - ref) < 1e-15


@pytest.mark.parametrize("args", [(0, 2, 2, 0, 1, 1, 0, 3, 1), (2, 2, 2, 4, 1, 5, 4, 3, 5), (4, 2, 6, 2, 1, 3, 4, 1, 5),
                                  (6, 0, 6, 8, 1, 7, 6, 1, 7), (2, 2, 4, 6, 1, 7, 6, 3, 9)])
def test_wigner9j_sympy(args):
    ref = float(wigner_9j(*[sympy_half(a) for a in args]))
    assert abs(ang.wigner9j(*args) - ref) < 1e-14


def test_spharm_plm_scipy():
    x = np.linspace(-0.99, 0.99, 7)
    tab = ang.spharm_plm(6, x)
    theta = np.arccos(x)
    for l in range(7):
        for m in range(-l, l + 1):
            ref = ss.sph_harm_y(l, m, theta, 0.0).real
            assert np.allclose(ang.plm_signed(tab, l, m), ref, atol=1e-14)


# -------------------------------------------------------------------- P123
def _swave_pw():
    return st.construct_pw_states(1, 1, False, swave_only=True).channel(0)


def test_gtilde_swave_closed_form():
    pw = _swave_pw()
    A, Lmax = pm.atilde_table(pw, 1)
    s = np.array([[0.3], [0.8]])
    c = np.sqrt(1 - s * s)
    x = np.array([[-0.5, 0.2, 0.9]])
    tabs = (ang.spharm_plm(0, x)[:, None, :] if False else ang.spharm_plm(0, x), ang.spharm_plm(0, 0 * s + 0 * x),
            ang.spharm_plm(0, 0 * s + 0 * x))
    for a in range(2):
        for ap in range(2):
            g = pm.gtilde(pw, ap, a, 1, A, Lmax, s, c, x, tabs)
            assert np.allclose(g, A[ap, a, 0] / 2, rtol=1e-14)


def test_p123_element_geometry_brute_force():
    """For G~ == 1 the element is (1/norm) int_bra p'q' dp'dq' int dx Theta(pi1 in p-bin) Theta(pi2 in q-bin)/(pi1 pi2);
    the polar quadrature of the oracle at high order must match a Cartesian brute force."""
    pe = np.array([0.0, 60.0, 130.0, 220.0])
    qe = np.array([0.0, 50.0, 120.0, 200.0])
    bra, ket = (1, 2), (1, 1)        # (q', p'), (q, p)
    Nphi = Nx = 160
    x, wx = st.gauss(Nx)
    phi, wphi = pm.phi_mesh(pe, qe, Nphi)
    qp, pp = bra
    q, p = ket
    s = np.sin(phi[qp, pp])[:, None]
    c = np.cos(phi[qp, pp])[:, None]
    z1, z2 = pm.pi1_tilde(s, c, x[None, :]), pm.pi2_tilde(s, c, x[None, :])
    kmin = np.maximum.reduce([pe[pp] / s + 0 * x, qe[qp] / c + 0 * x, pe[p] / z1, qe[q] / z2])
    kmax = np.minimum.reduce([pe[pp + 1] / s + 0 * x, qe[qp + 1] / c + 0 * x, pe[p + 1] / z1, qe[q + 1] / z2])
    g = wphi[qp, pp][:, None] * wx[None, :] * c * s * (kmax ** 2 - kmin ** 2) / (2 * z1 * z2)
    polar = g[kmin <= kmax].sum()
    n = 220
    P1 = pe[pp] + (np.arange(n) + 0.5) * (pe[pp + 1] - pe[pp]) / n
    Q1 = qe[qp] + (np.arange(n) + 0.5) * (qe[qp + 1] - qe[qp]) / n
    X = -1 + (np.arange(n) + 0.5) * 2 / n
    Pp, Qp, XX = np.meshgrid(P1, Q1, X, indexing="ij")
    a1 = pm.pi1_tilde(Pp, Qp, XX)
    a2 = pm.pi2_tilde(Pp, Qp, XX)
    ins = (a1 >= pe[p]) & (a1 <= pe[p + 1]) & (a2 >= qe[q]) & (a2 <= qe[q + 1])
    brute = np.sum(np.where(ins, Pp * Qp / (a1 * a2), 0.0)) * (pe[pp + 1] - pe[pp]) * (qe[qp + 1] - qe[qp]) * 2 / n ** 3
    assert polar > 0 and abs(polar - brute) < 3e-3 * brute


def test_overlap_mask_against_point_sampling():
    pe = np.array([0.0, 60.0, 130.0, 220.0])
    qe = np.array([0.0, 50.0, 120.0, 200.0])
    x, _ = st.gauss(24)
    phi, _ = pm.phi_mesh(pe, qe, 24)
    mask = pm.overlap_mask(pe, qe, x, phi)
    rng = np.random.default_rng(1)
    for qp in range(3):
        for pp in range(3):
            n = 60000
            P1 = rng.uniform(pe[pp], pe[pp + 1], n)
            Q1 = rng.uniform(qe[qp], qe[qp + 1], n)
            X = rng.uniform(-1, 1, n)
            a1, a2 = pm.pi1_tilde(P1, Q1, X), pm.pi2_tilde(P1, Q1, X)
            for q in range(3):
                for p in range(3):
                    hit = np.any((a1 > pe[p]) & (a1 < pe[p + 1]) & (a2 > qe[q]) & (a2 < qe[q + 1]))
                    if hit:  # a sampled overlap must be detected by the quadrature test
                        assert mask[qp, pp, q, p]


# ------------------------------------------------------------------- solve
def test_pade_geometric_series_exact():
    # [1/1] is exact for one geometric series, [2/2] for a sum of two (higher orders are degenerate 0/0)
    r1, r2, c1, c2 = 0.37 - 0.5j, -1.4 + 0.2j, 0.8, -0.3 + 0.1j
    a1 = np.array([r1 ** n for n in range(29)])
    assert abs(so.pade_approximant(a1, 1, 1, 1.0) - 1 / (1 - r1)) < 1e-13
    a2 = np.array([c1 * r1 ** n + c2 * r2 ** n for n in range(29)])
    assert abs(so.pade_approximant(a2, 2, 2, 1.0) - (c1 / (1 - r1) + c2 / (1 - r2))) < 1e-12


def test_pade_matches_scipy():
    from scipy.interpolate import pade
    a = np.array([1.0, 0.5, -0.3, 0.7, 0.2, -0.1, 0.05])
    p, q = pade(a[:5], 2)
    assert abs(so.pade_approximant(a, 2, 2, 1.0) - p(1.0) / q(1.0)) < 1e-12


def test_pade_select_divergent_series_recovers_resummed_value():
    # a_n = r^n with |r| > 1: Neumann sum diverges, [N/N] still gives 1/(1-r)
    r = 1.8 + 0.4j
    a = np.array([r ** n for n in range(29)])
    U, PA, best = so.pade_select(a)
    assert abs(U - 1 / (1 - r)) < 1e-9


def test_row_apply_dense_and_dense_solve_residual():
    rng = np.random.default_rng(3)
    pw = _swave_pw()
    N = 6
    fwp = st.make_fwp(N, N, st.chebyshev_scale_for_top(N, 8 * st.hbarc))
    pwf = st.construct_pw_states(1, 1, False, swave_only=True)
    Vu, Vc = sw.potential_matrices(fwp, pwf, pot.MalflietTjon(), False, False, 1)
    swp = sw.make_swp_states(fwp, pwf, Vu, Vc, False, False)
    ent, _ = pm.p123_rows(pw, fwp.p_edges, fwp.q_edges, 16, 16)
    dim = pw.Nalpha * N * N
    P = so.p_matrix(*pm.to_csc(ent, dim), dim)
    K = so.KernelA(pw, swp, Vu, Vc, P, False, False)
    A = K.dense()
    # elements of A = C^T (2P) VC from the per-alpha-pair blocks and P itself, by explicit sums
    # (the sparse assembly of the q-diagonal block operators is not used)
    Pd = P.toarray()
    Na = pw.Nalpha
    for (ar, qr, pr), (ac, qc, pc) in [((0, 1, 0), (1, 2, 3)), ((1, 4, 2), (0, 0, 5)), ((1, 3, 1), (1, 3, 1))]:
        ref = 0.0
        for ai in range(Na):
            for pi in range(N):
                cpart = K.Cb[ai, pi, ar, pr]
                if cpart == 0.0:
                    continue
                for aj in range(Na):
                    for pj in range(N):
                        ref += cpart * 2.0 * Pd[ai * N * N + qr * N + pi, aj * N * N + qc * N + pj] \
                            * K.VCb[aj, pj, ac, pc]
        assert abs(A[ar * N * N + qr * N + pr, ac * N * N + qc * N + pc] - ref) <= 1e-13 * np.abs(A).max()
    X = rng.normal(size=(2, dim)) + 1j * rng.normal(size=(2, dim))
    assert np.allclose(K.row_apply(X), X @ A, atol=1e-13 * np.abs(X @ A).max())
    assert np.allclose(K.col_apply(X.T), A @ X.T, atol=1e-13 * np.abs(A @ X.T).max())
    G = rng.normal(size=(1, dim)) * 0.01 + 1j * rng.normal(size=(1, dim)) * 0.01
    deut, q_idx = np.array([1]), np.array([2])
    U = so.dense_solve(A, G, deut, q_idx, N, N)
    r = int(deut[0]) * N * N + int(q_idx[0]) * N
    Xfull = np.linalg.solve(np.eye(dim) - A * G[0][None, :], A)
    assert abs(U[0, 0, 0] - Xfull[r, r]) < 1e-12 * abs(Xfull[r, r])
    assert np.allclose((np.eye(dim) - A * G[0][None, :]) @ Xfull, A, atol=1e-12)


# ---------------------------------------------------------------- breakup
def _expected_breakup_cells(ep, Eq, E):
    """Breakup cells from the geometry of each cell: the cell [ep_p, ep_p+1] x [Eq_q, Eq_q+1]
    brackets E when E lies between its smallest and largest corner sum; its four
    sub-rectangles (split at the midpoints) are tried in the order (+p, +q), (-p, +q),
    (-p, -q), (+p, -q), and the first whose corner sums bracket E names the 2x2 block of
    neighbours, listed p-major with q inside."""
    out = []
    Na, Np1 = ep.shape
    for a in range(Na):
        for q in range(len(Eq) - 1):
            for p in range(Np1 - 1):
                if ep[a, p] < 0 or ep[a, p + 1] < 0:
                    continue
                corners = [x + y for x in ep[a, p:p + 2] for y in Eq[q:q + 2]]
                if not (min(corners) <= E <= max(corners)):
                    continue
                out.append((a, q, p))
                pm_, qm_ = 0.5 * (ep[a, p + 1] + ep[a, p]), 0.5 * (Eq[q + 1] + Eq[q])
                for dp, dq in ((1, 1), (-1, 1), (-1, -1), (1, -1)):
                    pr = (pm_, ep[a, p + 1]) if dp > 0 else (ep[a, p], pm_)
                    qr = (qm_, Eq[q + 1]) if dq > 0 else (Eq[q], qm_)
                    sums = [x + y for x in pr for y in qr]
                    if min(sums) <= E <= max(sums):
                        block = sorted({(p + i, q + k) for i in (0, dp) for k in (0, dq)} - {(p, q)})
                        out += [(a, qq, pp) for pp, qq in block]
                        break
    return out


def _mt_breakup_case(N=10, T_lab=(0.5, 13.0, 40.0)):
    pwf = st.construct_pw_states(1, 1, False, swave_only=True)
    pw = pwf.channel(0)
    fwp = st.make_fwp(N, N, st.chebyshev_scale_for_top(N, 8 * st.hbarc))
    Vu, Vc = sw.potential_matrices(fwp, pwf, pot.MalflietTjon(), False, False, 1)
    swp = sw.make_swp_states(fwp, pwf, Vu, Vc, False, False)
    os_ = sw.find_on_shell_bins(fwp.q_edges, swp.E_bound, list(T_lab))
    return pw, fwp, Vu, Vc, swp, os_


def test_breakup_cells_against_cell_geometry():
    pw, fwp, _, _, swp, os_ = _mt_breakup_case()
    ptr, cells = sw.find_breakup_bins(pw, swp, os_.E_com, False, False)
    ep = np.stack([sw._p_edges_of_alpha(pw, swp, a, False, False) for a in range(pw.Nalpha)])
    Eq = 0.5 * fwp.q_edges ** 2 / sw.mu1_free()
    assert ptr[0] == 0 and len(ptr) == len(os_.E_com) + 1
    for j, E in enumerate(os_.E_com):
        got = [tuple(c) for c in cells[ptr[j]:ptr[j + 1]]]
        if E + swp.E_bound < 0:
            assert got == []
        else:
            assert got and got == _expected_breakup_cells(ep, Eq, E)
    assert ptr[-1] > 0 and (os_.E_com + swp.E_bound < 0).any()   # both branches exercised


def test_breakup_dense_values_against_explicit_inverse_and_pade():
    """U_BU from (1 - A G_j) X = A[:, cells] equals A (1 - G_j A)^-1 by explicit inversion at
    the same entries (the other order of the factors), and the Pade resummation of the breakup
    Neumann terms reaches it."""
    from oracle import p123 as pm_
    pw, fwp, Vu, Vc, swp, os_ = _mt_breakup_case()
    N = fwp.Np
    deut = sw.find_deuteron_channels(pw)
    ptr, cells = sw.find_breakup_bins(pw, swp, os_.E_com, False, False)
    ent, _ = pm_.p123_rows(pw, fwp.p_edges, fwp.q_edges, 16, 16)
    dim = pw.Nalpha * N * N
    K = so.KernelA(pw, swp, Vu, Vc, so.p_matrix(*pm_.to_csc(ent, dim), dim), False, False)
    A = K.dense()
    G = np.stack([rs.resolvent_array(E + swp.E_bound, pw, swp, False, False) for E in os_.E_com])
    Ubu = so.dense_solve_breakup(A, G, deut, os_.q_idx, N, N, ptr, cells)
    assert len(cells) > 0
    for j in range(len(os_.E_com)):
        Ufull = A @ np.linalg.inv(np.eye(dim) - G[j][:, None] * A)
        for dr, a in enumerate(deut):
            r = int(a) * N * N + int(os_.q_idx[j]) * N
            for b in range(ptr[j], ptr[j + 1]):
                c = so.breakup_flat(cells[b], N, N)
                if 0 <= c < dim:
                    assert abs(Ubu[dr, b] - Ufull[r, c]) <= 1e-11 * np.abs(Ufull[r]).max()
                else:
                    assert np.isnan(Ubu[dr, b])
    ok = ~np.isnan(Ubu)
    a_bu = so.neumann_terms_breakup(K, G, deut, os_.q_idx, N, N, ptr, cells, 29)
    Up = np.array([[so.pade_select(a_bu[d, b])[0] if ok[d, b] else np.nan for b in range(len(cells))]
                   for d in range(len(deut))])
    assert np.abs(Up[ok] - Ubu[ok]).max() <= 1e-6 * np.abs(Ubu[ok]).max()


def test_midpoint_branch_is_one_point_quadrature():
    """The mid-point branch of V_WP (CPP/make_potential_matrix.cpp:271-329: p_in p_out d_r d_c V at the
    bin midpoints) equals the quadrature branch (330-396, pinned against scipy dblquad) with one
    Gauss-Legendre point per bin, whose node is the bin midpoint and whose weight the bin width; on
    the synthetic potential with tensor force and isospin breaking (coupled and T-mixing blocks)."""
    N = 6
    pw = st.construct_pw_states(3, 1, True)
    s = 100.0
    fwp1 = st.make_fwp(N, N, s, 1.0, Np_per_WP=1)
    fwp8 = st.make_fwp(N, N, s, 1.0, Np_per_WP=8)
    assert np.allclose(fwp1.p_pts, 0.5 * (fwp1.p_edges[1:] + fwp1.p_edges[:-1]), rtol=1e-15)
    assert np.allclose(fwp1.p_wts, np.diff(fwp1.p_edges), rtol=1e-14)
    theta, beta = pot.synthetic_theta(7)
    orc = pot.Synthetic(theta, beta)
    Vu_q, Vc_q = sw.potential_matrices(fwp1, pw, orc, True, True, 3)
    Vu_m, Vc_m = sw.potential_matrices(fwp8, pw, orc, True, True, 3, midpoint_approx=True)
    for a, b in ((Vu_m, Vu_q), (Vc_m, Vc_q)):
        used = ~np.isnan(b)
        assert np.array_equal(used, ~np.isnan(a)) and used.any()
        assert np.abs(a[used] - b[used]).max() <= 1e-13 * np.abs(b[used]).max()
    # and it is not the 8-point quadrature (the branch is taken)
    Vu_8, _ = sw.potential_matrices(fwp8, pw, orc, True, True, 3)
    used = ~np.isnan(Vu_8)
    assert np.abs(Vu_m[used] - Vu_8[used]).max() > 1e-6 * np.abs(Vu_8[used]).max()
