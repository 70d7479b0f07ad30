"""Oracle: V in the WP basis, SWP states, on-shell bins (test infrastructure only).

Restates:
  CPP/make_potential_matrix.cpp:4-430   V_WP blocks (uncoupled Np x Np, coupled 2Np x 2Np),
                                        isoscalar nn/np mixing, isospin-breaking 1S0 block
  CPP/make_swp_states.cpp:28-183,221-471 H = H0 + V, eigenvectors C, SWP energy-bin edges
  CPP/run_organizer.cpp:4-131,342-378   on-shell q-bins and deuteron channels
  CPP/run_organizer.cpp:136-339         breakup on-shell cells (alpha, q, p) and their
                                        quadrant neighbours

The reference diagonalises with LAPACKE_dspevd; here numpy.linalg.eigh.  The
eigenvector sign is arbitrary in both, and every on-shell U element is
invariant under it (C enters as C ... C^T with the diagonal G in between).

Pins (tests/test_oracle_pins.py): C^T C = 1 and C^T H C diagonal; the nn/np mixing
weights against isospin Clebsch-Gordan coefficients, and the sign of the T3N-mixing block shown
to be a convention that cancels in U (tests/test_oracle_recoupling.py); MT-III deuteron
near -2.23 MeV at large N (CPP/Interactions/malfliet_tjon.cpp:11); V_WP of the
MT potential against scipy.integrate.dblquad over one bin pair.  Breakup cells: every
listed primary cell brackets the energy and every bracketing cell is listed, in loop
order; the neighbours follow from the geometry of the cell's four sub-rectangles (the
energy ranges of their corner sums), computed independently in the test.  The energy
convention itself (the kinetic E_com of the on-shell bin, compared with SWP p-bin energies
plus E_q = q^2 / (2 mu1_free)) is the reference's choice at run_organizer.cpp:146-220 and no
reference output pins it: "parity unpinned" for that convention.
"""
import math
from dataclasses import dataclass

import numpy as np

from . import states as st


def extract_potential_element(L, Lp, J, S, coupled, v):
    """CPP/make_potential_matrix.cpp:4-48."""
    if J < abs(S - L) or J > S + L or J < abs(S - Lp) or J > S + Lp:
        raise ValueError("unphysical state")
    sgn = -1.0
    if coupled:
        if L == Lp and L < J:
            return v[2]
        if L == Lp and L > J:
            return v[5]
        if L < Lp:
            return sgn * v[4]
        return sgn * v[3]
    if L == J + 1:
        return v[5]
    if L == J - 1:
        return v[2]
    if S == 0:
        return v[0]
    if S == 1 and L == J:
        return v[1]
    return 0.0


def potential_matrices(fwp, pw_full, pot, tensor_force, isospin_breaking_1S0, J_2N_max, midpoint_approx=False):
    """CPP/make_potential_matrix.cpp:51-430: the quadrature branch (330-396) or, with
    midpoint_approx, the mid-point branch (271-329): V_WP = p_in p_out f_in f_out d_r d_c V / (N_r N_c)
    at the bin midpoints p (CPP/make_wp_states.cpp:150-160), f = 1 for momentum wave packets.

    Returns (V_unco[nu, Np, Np], V_coup[nc, 2Np, 2Np]); blocks never touched stay NaN.
    """
    nu, nc = st.num_2N_states(J_2N_max, tensor_force, isospin_breaking_1S0)
    Np, npw = fwp.Np, fwp.Np_per_WP
    Vu = np.full((nu, Np, Np), np.nan)
    Vc = np.full((nc, 2 * Np, 2 * Np), np.nan)
    done_u, done_c = set(), set()
    pts, wts = fwp.p_pts, fwp.p_wts
    for ar in range(pw_full.Nalpha):
        Lr, Sr, Jr, Tr, t3r = (int(pw_full.L[ar]), int(pw_full.S[ar]), int(pw_full.J[ar]),
                               int(pw_full.T[ar]), int(pw_full.two_T3[ar]))
        for ac in range(pw_full.Nalpha):
            Lc, Sc, Jc, Tc, t3c = (int(pw_full.L[ac]), int(pw_full.S[ac]), int(pw_full.J[ac]),
                                   int(pw_full.T[ac]), int(pw_full.two_T3[ac]))
            checkL = (tensor_force and abs(Lr - Lc) <= 2) or Lr == Lc
            if not (Tr == Tc and Jr == Jc and Sr == Sc and checkL):
                continue
            s1S0 = (Sr == 0 and Jr == 0 and Lr == 0)
            # isoscalar type, lines 138-158 (operator precedence of line 147 kept)
            if s1S0 and t3r == 3 and t3c == 3:
                iso = 2
            elif (s1S0 and (t3r == 3 and t3c == 1)) or (t3r == 1 and t3c == 3):
                iso = 3
            elif Tr == 0 and t3r == t3c and t3r == 1:
                iso = 0
            elif Tr == 1 and t3r == t3c and t3r == 1:
                iso = 1
            else:
                raise ValueError("unknown isoscalar type")
            via_L, via_T = st.coupling_kind(Lr, Sr, Jr, Tr, Lc, tensor_force, isospin_breaking_1S0)
            coupled_matrix = via_L or via_T
            coupled_model = (Lr != Lc) or (Lr == Lc and Lr != Jr)
            idx = st.unique_2N_idx(Lr, Sr, Jr, Tr, coupled_matrix, tensor_force, isospin_breaking_1S0)
            if coupled_matrix:
                if idx in done_c:
                    continue
                done_c.add(idx)
            else:
                if idx in done_u:
                    continue
                done_u.add(idx)
            def isoscalar(p_in, p_out):
                """The six V slots of the isoscalar combination (lines 283-316 / 351-385)."""
                if not via_T and iso == 0:
                    return np.array(pot.V(p_in, p_out, coupled_model, Sr, Jr, Tr, 0))
                if not via_T and iso == 1:
                    vnn = np.array(pot.V(p_in, p_out, coupled_model, Sr, Jr, Tr, -1))
                    vnp = np.array(pot.V(p_in, p_out, coupled_model, Sr, Jr, Tr, 0))
                    return (2. / 3) * vnn + (1. / 3) * vnp
                if via_T:
                    vnn = pot.V(p_in, p_out, coupled_model, Sr, Jr, Tr, -1)
                    vnp = pot.V(p_in, p_out, coupled_model, Sr, Jr, Tr, 0)
                    enn = extract_potential_element(Lr, Lc, Jr, Sr, False, vnn)
                    enp = extract_potential_element(Lr, Lc, Jr, Sr, False, vnp)
                    vis = np.zeros(6)
                    vis[2] = (2. / 3) * enn + (1. / 3) * enp
                    vis[3] = (math.sqrt(2) / 3.) * (enn - enp)
                    vis[4] = (math.sqrt(2) / 3.) * (enn - enp)
                    vis[5] = (1. / 3) * enn + (2. / 3) * enp
                    return vis
                raise ValueError("unknown isoscalar potential")

            for br in range(Np):
                Nr = fwp.norm_p[br]
                for bc in range(Np):
                    Nc = fwp.norm_p[bc]
                    acc = np.zeros(6)
                    if midpoint_approx:
                        p_in = 0.5 * (fwp.p_edges[bc] + fwp.p_edges[bc + 1])
                        p_out = 0.5 * (fwp.p_edges[br] + fwp.p_edges[br + 1])
                        d_c = fwp.p_edges[bc + 1] - fwp.p_edges[bc]
                        d_r = fwp.p_edges[br + 1] - fwp.p_edges[br]
                        acc += p_in * p_out * d_r * d_c * isoscalar(p_in, p_out) / (Nr * Nc)
                    else:
                        for ir in range(npw):
                            pr = pts[br * npw + ir]
                            fr = wts[br * npw + ir] * pr
                            for ic in range(npw):
                                pc = pts[bc * npw + ic]
                                fc = wts[bc * npw + ic] * pc
                                acc += fr * fc * isoscalar(pc, pr) / (Nr * Nc)
                    if coupled_matrix:
                        if via_L:
                            Vc[idx, br, bc] = extract_potential_element(Jr - 1, Jr - 1, Jr, Sr, True, acc)
                            Vc[idx, br, bc + Np] = extract_potential_element(Jr - 1, Jr + 1, Jr, Sr, True, acc)
                            Vc[idx, br + Np, bc] = extract_potential_element(Jr + 1, Jr - 1, Jr, Sr, True, acc)
                            Vc[idx, br + Np, bc + Np] = extract_potential_element(Jr + 1, Jr + 1, Jr, Sr, True, acc)
                        else:
                            Vc[idx, br, bc] = acc[2]
                            Vc[idx, br, bc + Np] = acc[3]
                            Vc[idx, br + Np, bc] = acc[4]
                            Vc[idx, br + Np, bc + Np] = acc[5]
                    else:
                        Vu[idx, br, bc] = extract_potential_element(Lr, Lc, Jr, Sr, False, acc)
    return Vu, Vc


@dataclass
class SWP:
    Np: int
    Nq: int
    nu: int
    nc: int
    E_bound: float
    e_unco: np.ndarray     # (nu, Np+1)
    e_coup: np.ndarray     # (nc, 2(Np+1))
    C_unco: np.ndarray     # (nu, Np, Np)    C[basis, eigenstate]
    C_coup: np.ndarray     # (nc, 2Np, 2Np)
    q_edges: np.ndarray


def _swp_edges(ev, Np, coupled, chn_3S1):
    """CPP/make_swp_states.cpp:146-183."""
    e = np.zeros(2 * (Np + 1) if coupled else Np + 1)
    for i in range(Np - 1):
        e[i + 1] = 0.5 * (ev[i + 1] + ev[i])
        if coupled:
            e[i + 1 + Np + 1] = 0.5 * (ev[i + 1 + Np] + ev[i + Np])
    e[Np] = ev[Np - 1] + 0.5 * (e[Np - 1] - e[Np - 2])
    if coupled:
        e[2 * Np + 1] = ev[2 * Np - 1] + 0.5 * (e[2 * Np] - e[2 * Np - 1])
    if chn_3S1:
        e[0] = ev[0]
        e[1] = 0.0
    return e


def make_swp_states(fwp, pw_full, Vu, Vc, tensor_force, isospin_breaking_1S0):
    """CPP/make_swp_states.cpp:221-471; raises unless exactly one 3S1 bound state."""
    Np = fwp.Np
    pe = fwp.p_edges
    H0 = (pe[1:] ** 2 + pe[1:] * pe[:-1] + pe[:-1] ** 2) / (6 * st.mu23)
    nu, nc = Vu.shape[0], Vc.shape[0]
    Cu = np.full_like(Vu, np.nan)
    Cc = np.full_like(Vc, np.nan)
    eu = np.zeros((nu, Np + 1))
    ec = np.zeros((nc, 2 * (Np + 1)))
    done_u, done_c = set(), set()
    E_bound = 0.0
    for ar in range(pw_full.Nalpha):
        Lr, Sr, Jr, Tr = (int(pw_full.L[ar]), int(pw_full.S[ar]), int(pw_full.J[ar]), int(pw_full.T[ar]))
        for ac in range(pw_full.Nalpha):
            Lc, Sc, Jc, Tc = (int(pw_full.L[ac]), int(pw_full.S[ac]), int(pw_full.J[ac]), int(pw_full.T[ac]))
            if not (Tr == Tc and Jr == Jc and Sr == Sc and abs(Lr - Lc) <= 2):
                continue
            via_L, via_T = st.coupling_kind(Lr, Sr, Jr, Tr, Lc, tensor_force, isospin_breaking_1S0)
            coupled = via_L or via_T
            idx = st.unique_2N_idx(Lr, Sr, Jr, Tr, coupled, tensor_force, isospin_breaking_1S0)
            if coupled:
                if idx in done_c:
                    continue
                done_c.add(idx)
                V = Vc[idx]
                h0 = np.concatenate([H0, H0])
            else:
                if idx in done_u:
                    continue
                done_u.add(idx)
                V = Vu[idx]
                h0 = H0
            # H from the upper triangle of V (construct_full_hamiltonian, lines 45-81)
            Vs = np.triu(V) + np.triu(V, 1).T
            H = Vs + np.diag(h0)
            ev, C = np.linalg.eigh(H)
            chn_3S1 = (Lr == 0 and Sr == 1 and Jr == 1 and Tr == 0)
            nb = int(np.sum(ev < 0))
            if nb == 1:
                E_bound = float(ev[ev < 0][0])
            if (nb == 0 and not chn_3S1) or (nb == 1 and chn_3S1):
                pass
            elif nb != 0:
                raise ValueError("unphysical bound states in NN Hamiltonian (L=%d S=%d J=%d T=%d)" % (Lr, Sr, Jr, Tr))
            else:
                raise ValueError("no bound state in the 3S1 Hamiltonian")
            if coupled:
                # reorder_coupled_eigenspectrum, lines 85-112
                order = np.concatenate([np.arange(0, 2 * Np, 2), np.arange(1, 2 * Np, 2)])
                ev = ev[order]
                C = C[:, order]
                Cc[idx] = C
                ec[idx] = _swp_edges(ev, Np, True, chn_3S1)
            else:
                Cu[idx] = C
                eu[idx] = _swp_edges(ev, Np, False, chn_3S1)
    return SWP(Np, fwp.Nq, nu, nc, E_bound, eu, ec, Cu, Cc, fwp.q_edges.copy())


@dataclass
class OnShell:
    T_lab: np.ndarray
    E_com: np.ndarray
    q_com: np.ndarray
    q_idx: np.ndarray


def find_on_shell_bins(q_edges, E_bound, T_lab_input):
    """CPP/run_organizer.cpp:4-131 (elastic part)."""
    Nq = len(q_edges) - 1
    m1 = st.mu1(E_bound)
    Tmid = np.zeros(Nq)
    for k in range(Nq):
        Eql = 0.5 * q_edges[k] ** 2 / m1
        Equ = 0.5 * q_edges[k + 1] ** 2 / m1
        Ec = 0.5 * (Equ + Eql)
        Tmid[k] = st.com_momentum_to_lab_energy(st.com_energy_to_com_q_momentum(Ec), E_bound)
    mid = [False] * (Nq - 1)
    for k in range(Nq - 1):
        for T in T_lab_input:
            if Tmid[k] <= T <= Tmid[k + 1]:
                mid[k] = True
    onbin = [False] * Nq
    for k in range(Nq - 1):
        if mid[k]:
            onbin[k] = onbin[k + 1] = True
    bins = [k for k in range(Nq) if onbin[k]]
    T_lab = np.array([Tmid[k] for k in bins])
    q_com = np.array([st.lab_energy_to_com_momentum(T, E_bound) for T in T_lab])
    E_com = np.array([st.com_q_momentum_to_com_energy(q) for q in q_com])
    q_idx = []
    for q in q_com:
        found = -1
        for k in range(Nq):
            if q_edges[k] < q < q_edges[k + 1]:
                found = k
                break
        if found < 0:
            raise ValueError("on-shell T_lab outside the WP state space")
        q_idx.append(found)
    return OnShell(T_lab, E_com, q_com, np.array(q_idx, dtype=np.int64))


def find_deuteron_channels(pw_chn):
    """CPP/run_organizer.cpp:342-378 for one 3N channel."""
    return np.array([a for a in range(pw_chn.Nalpha)
                     if pw_chn.L[a] == st.deuteron_L and pw_chn.S[a] == st.deuteron_S
                     and pw_chn.J[a] == st.deuteron_J and pw_chn.T[a] == st.deuteron_T], dtype=np.int64)


def mu1_free():
    """Spectator reduced mass against a free pair (CPP/General_functions/kinetic_conversion.cpp:16-19)."""
    M12 = st.Mp + st.Mn
    return st.Mn * M12 / (st.Mn + M12)


def _p_edges_of_alpha(pw_chn, swp, a, tensor_force, isospin_breaking_1S0):
    """The SWP p-bin energy edges of alpha (CPP/run_organizer.cpp:157-203)."""
    L, S, J, T, t3 = (int(pw_chn.L[a]), int(pw_chn.S[a]), int(pw_chn.J[a]), int(pw_chn.T[a]),
                      int(pw_chn.two_T3[a]))
    s1S0 = S == 0 and J == 0 and L == 0
    via_L = bool(tensor_force) and L != J and J != 0
    via_T = s1S0 and bool(isospin_breaking_1S0)
    if via_L and via_T:
        raise ValueError("isospin breaking in a coupled channel")
    Np = swp.Np
    if via_L or via_T:
        idx = st.unique_2N_idx(L, S, J, T, True, tensor_force, isospin_breaking_1S0)
        if via_L:
            off = 0 if L < J else Np + 1
        else:
            off = 0 if t3 == 1 else Np + 1
        return swp.e_coup[idx, off:off + Np + 1]
    idx = st.unique_2N_idx(L, S, J, T, False, tensor_force, isospin_breaking_1S0)
    return swp.e_unco[idx]


def find_breakup_bins(pw_chn, swp, E_com, tensor_force, isospin_breaking_1S0):
    """CPP/run_organizer.cpp:136-339.  Returns (q_com_BU_idx [nE + 1], cells [n, 3] of
    (alpha, q, p)).  Energies with E_com + E_bound < 0 get no cells; an energy above that
    with no bracketing cell raises."""
    Nq, Np = swp.Nq, swp.Np
    qe = swp.q_edges
    m = mu1_free()
    cells, ptr = [], [0]
    for E in E_com:
        if E + swp.E_bound < 0:
            ptr.append(len(cells))
            continue
        found = False
        for a in range(pw_chn.Nalpha):
            ep = _p_edges_of_alpha(pw_chn, swp, a, tensor_force, isospin_breaking_1S0)
            for q in range(Nq):
                Eql = 0.5 * (qe[q] * qe[q]) / m
                Equ = 0.5 * (qe[q + 1] * qe[q + 1]) / m
                for p in range(Np):
                    Epl, Epu = ep[p], ep[p + 1]
                    lo, hi = Epl + Eql, Epu + Equ
                    if not (lo <= E <= hi and Epl >= 0 and Epu >= 0):
                        continue
                    found = True
                    cells.append((a, q, p))
                    Ep, Eq = 0.5 * (Epu + Epl), 0.5 * (Equ + Eql)
                    if Eq + Ep <= E <= hi:                      # quadrant 1
                        prange, qrange = (p, p + 1), (q, q + 1)
                    elif Eq + Epl <= E <= Ep + Equ:             # quadrant 2
                        prange, qrange = (p - 1, p), (q, q + 1)
                    elif lo <= E <= Eq + Ep:                    # quadrant 3
                        prange, qrange = (p - 1, p), (q - 1, q)
                    elif Eql + Ep <= E <= Epu + Eq:             # quadrant 4
                        prange, qrange = (p, p + 1), (q - 1, q)
                    else:
                        continue
                    for pi in range(prange[0], prange[1] + 1):
                        for qj in range(qrange[0], qrange[1] + 1):
                            if pi != p or qj != q:
                                cells.append((a, qj, pi))
        if not found:
            raise ValueError("on-shell breakup energy outside the WP state space")
        ptr.append(len(cells))
    return np.array(ptr, dtype=np.int64), np.array(cells, dtype=np.int64).reshape(-1, 3)
