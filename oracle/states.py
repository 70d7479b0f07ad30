"""Oracle: state spaces, constants, kinematics, Gauss-Legendre (test infrastructure only).

Restates:
  CPP/constants.h:19-35                     masses, hbar*c, deuteron quantum numbers
  CPP/General_functions/kinetic_conversion.cpp:5-53   kinematics
  CPP/General_functions/gauss_legendre.cpp:36-95      Gauss-Legendre nodes/weights
  CPP/make_wp_states.cpp:3-44,77-211        Chebyshev bins, per-bin GL meshes, normalisation
  CPP/make_pw_symm_states.cpp:4-283         partial-wave enumeration and unique 2N index

Pins (tests/test_oracle_pins.py): kinematics against the Mandelstam-s closed form and a
Lorentz boost, the masses and hbar c against PDG values (tests/test_oracle_recoupling.py);
GL against numpy.polynomial.legendre.leggauss,
Chebyshev edges against the closed form, kinematics round trips, and the channel
counts of the enumeration against a brute-force enumeration of angular-momentum
couplings.
"""
import math
from dataclasses import dataclass, field

import numpy as np

# CPP/constants.h:19-35
Mp = 938.272
Mn = 939.565
MN = 0.5 * (Mp + Mn)
mu23 = Mp * Mn / (Mn + Mp)
hbarc = 197.327
deuteron_L, deuteron_S, deuteron_J, deuteron_T = 0, 1, 1, 0


# ---------------------------------------------------------------- kinematics
# CPP/General_functions/kinetic_conversion.cpp:5-53
def deuteron_mass(Eb):
    return Mn + Mp + Eb


def mu1(Eb):
    Md = deuteron_mass(Eb)
    return Mn * Md / (Mn + Md)


def com_momentum_to_lab_energy(q_com, Eb):
    Md = deuteron_mass(Eb)
    b = 2 * (Mp - q_com * q_com / Md)
    c = -q_com * q_com * (Mp + Md) * (Mp + Md) / (Md * Md)
    return 0.5 * (-b + math.sqrt(b * b - 4 * c))


def lab_energy_to_com_momentum(T_lab, Eb):
    Md = deuteron_mass(Eb)
    return math.sqrt(Md * Md * T_lab * (T_lab + 2 * Mp) / ((Mp + Md) * (Mp + Md) + 2 * T_lab * Md))


def com_energy_to_com_q_momentum(E_com):
    return math.sqrt(4 * E_com * MN / 3)


def com_q_momentum_to_com_energy(q_com):
    return 3 * q_com * q_com / (4 * MN)


# ---------------------------------------------------------- Gauss-Legendre
def gauss(N):
    """Newton iteration on P_N, nodes ascending (CPP/General_functions/gauss_legendre.cpp:36-67)."""
    x = np.zeros(N)
    w = np.zeros(N)
    m = (N + 1) // 2
    for i in range(1, m + 1):
        t = math.cos(math.pi * (i - 0.25) / (N + 0.5))
        t1 = 1.0
        pp = 0.0
        while abs(t - t1) >= 3e-16:
            p1, p2 = 1.0, 0.0
            for j in range(1, N + 1):
                p3 = p2
                p2 = p1
                p1 = ((2 * j - 1) * t * p2 - (j - 1) * p3) / j
            pp = N * (t * p1 - p2) / (t * t - 1)
            t1 = t
            t = t1 - p1 / pp
        x[i - 1] = -t
        x[N - i] = t
        w[i - 1] = 2.0 / ((1 - t * t) * pp * pp)
        w[N - i] = w[i - 1]
    return x, w


def gauss_ab(N, a, b):
    """GL mesh mapped to [a, b] (CPP/General_functions/gauss_legendre.cpp:90-95)."""
    x, w = gauss(N)
    return 0.5 * (b - a) * x + 0.5 * (b + a), 0.5 * (b - a) * w


# --------------------------------------------------------------- FWP space
def chebyshev_edges(N, s, t=1.0):
    """b_0 = 0, b_i = s tan((2i-1) pi / 4N)^t (CPP/make_wp_states.cpp:26-44)."""
    e = np.zeros(N + 1)
    for i in range(1, N + 1):
        e[i] = s * math.tan((2 * i - 1) * math.pi / (4 * N)) ** t
    return e


def chebyshev_scale_for_top(N, top_MeV):
    """Scale s such that the top edge sits at top_MeV (t = 1)."""
    return top_MeV / math.tan((2 * N - 1) * math.pi / (4 * N))


@dataclass
class FWP:
    Np: int
    Nq: int
    p_edges: np.ndarray
    q_edges: np.ndarray
    Np_per_WP: int
    p_pts: np.ndarray = None     # (Np*Np_per_WP,)
    p_wts: np.ndarray = None
    norm_p: np.ndarray = None
    norm_q: np.ndarray = None


def make_fwp(Np, Nq, s, t=1.0, Np_per_WP=8):
    """CPP/make_wp_states.cpp:116-211 (momentum WPs, quadrature meshes, sqrt(width) norms)."""
    pe = chebyshev_edges(Np, s, t)
    qe = chebyshev_edges(Nq, s, t)
    pts = np.zeros(Np * Np_per_WP)
    wts = np.zeros(Np * Np_per_WP)
    for b in range(Np):
        x, w = gauss_ab(Np_per_WP, pe[b], pe[b + 1])
        pts[b * Np_per_WP:(b + 1) * Np_per_WP] = x
        wts[b * Np_per_WP:(b + 1) * Np_per_WP] = w
    return FWP(Np, Nq, pe, qe, Np_per_WP, pts, wts,
               np.sqrt(np.diff(pe)), np.sqrt(np.diff(qe)))


# ---------------------------------------------------------------- PW space
@dataclass
class PW:
    L: np.ndarray
    S: np.ndarray
    J: np.ndarray
    T: np.ndarray
    l: np.ndarray
    two_j: np.ndarray
    two_J3: np.ndarray
    two_T3: np.ndarray
    P3: np.ndarray
    chn_idx: list = field(default_factory=list)   # start of each 3N channel + total

    @property
    def Nalpha(self):
        return len(self.L)

    @property
    def N_chn(self):
        return len(self.chn_idx) - 1

    def channel(self, c):
        a, b = self.chn_idx[c], self.chn_idx[c + 1]
        return PW(self.L[a:b], self.S[a:b], self.J[a:b], self.T[a:b], self.l[a:b],
                  self.two_j[a:b], self.two_J3[a:b], self.two_T3[a:b], self.P3[a:b],
                  [0, b - a])


def construct_pw_states(J_2N_max, two_J_3N_max, isospin_breaking_1S0=True, swave_only=False):
    """Loop order and filters of CPP/make_pw_symm_states.cpp:122-228.

    ``swave_only`` reproduces the commented-out S-wave filter (lines 176-179)
    that the Malfliet-Tjon configuration uses.
    """
    two_T3_max = 3 if isospin_breaking_1S0 else 1
    rows = []
    chn = []
    n_prev = 0
    for two_J3 in range(1, two_J_3N_max + 1, 2):
        for Prem in range(2):
            n_in = 0
            for two_T3 in range(1, two_T3_max + 1, 2):
                for J in range(0, J_2N_max + 1):
                    for S in range(2):
                        for L in range(abs(J - S), J + S + 1):
                            Tmin = abs(two_T3 - 1) // 2
                            Tmax = min((two_T3 + 1) // 2, 1)
                            for T in range(Tmin, Tmax + 1):
                                if (L + S + T) % 2 == 0:
                                    continue
                                for two_j in range(abs(two_J3 - 2 * J), two_J3 + 2 * J + 1, 2):
                                    for l in range(abs(two_j - 1) // 2, (two_j + 1) // 2 + 1):
                                        if (L + l) % 2 != Prem:
                                            continue
                                        if two_T3 == 3 and not (S == 0 and L == 0 and J == 0):
                                            continue
                                        if swave_only and (L != 0 or l != 0):
                                            continue
                                        rows.append((L, S, J, T, l, two_j, two_J3, two_T3, 1 - 2 * Prem))
                                        n_in += 1
            if n_in:
                chn.append(n_prev)
            n_prev += n_in
    chn.append(len(rows))
    a = np.array(rows, dtype=np.int64).reshape(-1, 9)
    return PW(*(a[:, k].copy() for k in range(9)), chn_idx=chn)


def unique_2N_idx(L, S, J, T, coupled, tensor_force, isospin_breaking_1S0):
    """CPP/make_pw_symm_states.cpp:4-60."""
    state_1S0 = (S == 0 and J == 0 and L == 0 and T == 1)
    if coupled and not tensor_force:
        # the reference aborts here (line 6-8), also for isospin-breaking 1S0
        raise ValueError("coupled state without tensor force")
    if tensor_force:
        if coupled:
            u = J - 1
            if (not state_1S0) and isospin_breaking_1S0:
                u += 1
            if state_1S0 and isospin_breaking_1S0:
                u = 0
            return u
        u = 2 * J + S
        if isospin_breaking_1S0:
            u -= 1
        return u
    u = J * (4 * (J > 1) + 2) + (S == 1) + (S == 1 and J != 0) * (J - L + 1)
    if isospin_breaking_1S0:
        u -= 1
    return u


def num_2N_states(J_2N_max, tensor_force, isospin_breaking_1S0):
    """CPP/main.cpp:273-289."""
    if tensor_force:
        nu = 2 * (J_2N_max + 1)
        nc = J_2N_max
        if isospin_breaking_1S0:
            nu -= 1
            nc += 1
    else:
        nu = 4 * J_2N_max + 2
        nc = 0
        if isospin_breaking_1S0:
            nu -= 1
            nc = 1
    return nu, nc


def coupling_kind(Lr, Sr, Jr, Tr, Lc, tensor_force, isospin_breaking_1S0):
    """(coupled_via_L, coupled_via_T3) as in CPP/solve_faddeev.cpp:1703-1713."""
    state_1S0 = (Sr == 0 and Jr == 0 and Lr == 0)
    via_L = tensor_force and (Lr != Lc or (Lr == Lc and Lr != Jr and Jr != 0))
    via_T = state_1S0 and isospin_breaking_1S0
    if via_L and via_T:
        raise ValueError("isospin breaking in a coupled channel")
    return via_L, via_T
