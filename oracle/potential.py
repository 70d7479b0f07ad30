"""Oracle: NN potential models (test infrastructure only).

Restates the interaction plugin convention of
  CPP/Interactions/potential_model.h:11-22   V(i, j, p_in, p_out, coupled, S, J, T, Tz, V[6])
  CPP/make_potential_matrix.cpp:4-48          meaning of the six slots
and the two models the workloads use:
  * Malfliet-Tjon I/III (CPP/Interactions/malfliet_tjon.cpp:27-66);
  * the seeded synthetic non-local potential (new, no reference counterpart):
        V = f(p') f(p) [ sum_i Y_i/hbarc * Q_L(z_i)/(pi p p') delta_{L'L}
                         + S-wave, P-wave and S-D contacts
                         + three separable rank-1 terms ]
    with f(p) = exp(-(p/Lambda)^6), Lambda = 2.5 fm^-1, z_i = (p^2+p'^2+mu_i^2)/(2pp').
    It is linear in its 12 parameters theta (see THETA_NAMES).

Momenta are in MeV, V in MeV^-2, with the measure p^2 dp of the reference
(MT: V_0/(pi p p') Q_0(z) == V_0/(2 pi p p') ln(...)).

Pins (tests/test_oracle_pins.py): Q_L against mpmath.legenq; the MT
S-wave formula against the Yukawa partial-wave projection at L = 0.
"""
import math

import numpy as np

from .states import hbarc

# six-slot convention of CPP/make_potential_matrix.cpp:4-48
# [0] S=0 uncoupled, [1] S=1 L=J, [2] L=L'=J-1, [3] (+-), [4] (-+), [5] L=L'=J+1


def legendre_Q(L, z):
    """Q_L(z), z > 1, by Gautschi's backward continued fraction for Q_l/Q_{l-1}.

    Q_0 = 1/2 ln((z+1)/(z-1)); r_l = Q_l/Q_{l-1} = l / ((2l+1) z - (l+1) r_{l+1}).
    """
    z = float(z)
    q0 = 0.5 * math.log((z + 1.0) / (z - 1.0))
    if L == 0:
        return q0
    # convergence factor per step ~ 1/(z + sqrt(z^2-1))
    rho = z + math.sqrt(z * z - 1.0)
    nstart = L + 20 + int(40.0 / max(math.log(rho), 1e-3))
    nstart = min(nstart, L + 20000)
    r = 0.0
    ratios = {}
    for l in range(nstart, 0, -1):
        r = l / ((2 * l + 1) * z - (l + 1) * r)
        if l <= L:
            ratios[l] = r
    q = q0
    for l in range(1, L + 1):
        q *= ratios[l]
    return q


def yukawa_pw(L, p_out, p_in, mu):
    """Partial-wave projected Yukawa e^{-mu r}/r per unit strength (MeV^-2 per MeV fm / hbarc).

    (2/pi) int r^2 dr j_L(p'r) e^{-mu r}/r j_L(pr) = Q_L(z)/(pi p p').
    """
    z = (p_out * p_out + p_in * p_in + mu * mu) / (2.0 * p_out * p_in)
    if L == 0:
        # same closed form as the MT potential, no cancellation near z = 1
        q0 = 0.5 * math.log(((p_out + p_in) ** 2 + mu * mu) / ((p_out - p_in) ** 2 + mu * mu))
        return q0 / (math.pi * p_out * p_in)
    return legendre_Q(L, z) / (math.pi * p_out * p_in)


class MalflietTjon:
    """CPP/Interactions/malfliet_tjon.cpp:27-66 (J==0: MT-I in slot 0, J==1: MT-III in slot 2)."""
    name = "malfliet_tjon"

    def V(self, p_in, p_out, coupled, S, J, T, Tz):
        v = [0.0] * 6
        VA_I, VA_III, VR = 513.968, 626.885, 1438.720
        muA, muR = 1.550 * hbarc, 3.110 * hbarc
        if J in (0, 1):
            num_A = (p_out + p_in) ** 2 + muA * muA
            num_R = (p_out + p_in) ** 2 + muR * muR
            den_A = (p_out - p_in) ** 2 + muA * muA
            den_R = (p_out - p_in) ** 2 + muR * muR
            VA = VA_I if J == 0 else VA_III
            val = ((VR / hbarc) * math.log(num_R / den_R) - (VA / hbarc) * math.log(num_A / den_A)) \
                / (2 * math.pi * p_in * p_out)
            v[0 if J == 0 else 2] = val
        return v


THETA_NAMES = ["Y_0.7", "Y_2.0", "Y_4.0",
               "C_1S0", "C_3S1", "C_3P0", "C_1P1", "C_3P1", "C_SD",
               "lam_1S0", "lam_3S1", "lam_1P1"]

# Default contact / separable parameters of the synthetic potential.  The
# Yukawa strengths are drawn from numpy.random.default_rng(seed).normal()*100.
SYNTH_DEFAULT_SEED = 7
SYNTH_MU_FM = (0.7, 2.0, 4.0)
SYNTH_LAMBDA_FM = 2.5
SYNTH_CD_NN = 0.02          # C_1S0(nn) = (1 + 0.02) C_1S0(np): charge dependence


def synthetic_theta(seed=SYNTH_DEFAULT_SEED, contacts=None, separable=None):
    """Parameter vector theta (12) and the seeded separable ranges beta_k [fm^-1]."""
    rng = np.random.default_rng(seed)
    Y = rng.normal(size=3) * 100.0
    beta = rng.uniform(1.0, 3.0, size=3)
    lam = rng.normal(size=3) * 0.01
    if contacts is None:
        contacts = [-0.10, -0.12653, 0.20, 0.15, 0.20, 0.02]
    if separable is not None:
        lam = np.asarray(separable, dtype=float)
    theta = np.concatenate([Y, np.asarray(contacts, float), lam])
    return theta, beta


class Synthetic:
    """Seeded synthetic non-local potential, linear in theta (12 parameters).

    Units: Y_i in MeV fm; S-wave contacts in fm^2; P-wave and S-D contacts in fm^4;
    separable strengths in fm^2 with g(p) = (p/hbarc)^l exp(-(p/(beta hbarc))^2).
    """
    name = "synthetic"

    def __init__(self, theta, beta, mu_fm=SYNTH_MU_FM, Lambda_fm=SYNTH_LAMBDA_FM, cd_nn=SYNTH_CD_NN):
        self.theta = np.asarray(theta, dtype=float)
        self.beta = np.asarray(beta, dtype=float)
        self.mu = [m * hbarc for m in mu_fm]
        self.Lam = Lambda_fm * hbarc
        self.cd_nn = cd_nn

    def _reg(self, p):
        return math.exp(-(p / self.Lam) ** 6)

    def _g(self, l, beta, p):
        x = p / hbarc
        return x ** l * math.exp(-(x / beta) ** 2)

    def V(self, p_in, p_out, coupled, S, J, T, Tz):
        th = self.theta
        v = [0.0] * 6
        f = self._reg(p_in) * self._reg(p_out)
        hc2, hc4 = hbarc ** 2, hbarc ** 4

        def yuk(L):
            return sum(th[i] / hbarc * yukawa_pw(L, p_out, p_in, self.mu[i]) for i in range(3))

        # slot -> L (diagonal part)
        if S == 0:
            v[0] = yuk(J)
        else:
            v[1] = yuk(J)
            if J >= 1:
                v[2] = yuk(J - 1)
            v[5] = yuk(J + 1)
        # contacts
        if S == 0 and J == 0:
            c = th[3] * (1.0 + self.cd_nn if Tz == -1 else 1.0)
            v[0] += c / hc2
            v[0] += th[9] / hc2 * self._g(0, self.beta[0], p_in) * self._g(0, self.beta[0], p_out)
        if S == 1 and J == 1:
            v[2] += th[4] / hc2 + th[10] / hc2 * self._g(0, self.beta[1], p_in) * self._g(0, self.beta[1], p_out)
            v[4] += th[8] * p_in * p_in / hc4     # (-+): row L=0 (p_out), column L=2 (p_in)
            v[3] += th[8] * p_out * p_out / hc4   # (+-): row L=2 (p_out), column L=0 (p_in)
            v[1] += th[7] * p_in * p_out / hc4    # 3P1
        if S == 1 and J == 0:
            v[5] += th[5] * p_in * p_out / hc4    # 3P0
        if S == 0 and J == 1:
            v[0] += th[6] * p_in * p_out / hc4    # 1P1
            v[0] += th[11] / hc2 * self._g(1, self.beta[2], p_in) * self._g(1, self.beta[2], p_out)
        return [x * f for x in v]
