"""Parameters of the seeded synthetic NN potential (workload definition).

Three Yukawa strengths drawn as numpy.random.default_rng(seed).normal()*100 MeV fm
(masses 0.7, 2.0, 4.0 fm^-1), fixed S/P-wave and S-D contacts, and three separable
rank-1 terms g(p) = p^l exp(-(p/beta)^2) whose ranges beta ~ U[1, 3] fm^-1 and
strengths ~ N(0, 0.01) fm^2 come from the same generator.  The layout matches
tt_potential_update_parameters: 12 strengths theta followed by 3 ranges beta.
"""
import numpy as np

DEFAULT_SEED = 7
# C_1S0, C_3S1 [fm^2]; C_3P0, C_1P1, C_3P1, C_SD [fm^4]
DEFAULT_CONTACTS = (-0.10, -0.12653, 0.20, 0.15, 0.20, 0.02)
THETA_NAMES = ("Y_0.7", "Y_2.0", "Y_4.0", "C_1S0", "C_3S1", "C_3P0", "C_1P1", "C_3P1", "C_SD",
               "lam_1S0", "lam_3S1", "lam_1P1")


def synthetic_params(seed=DEFAULT_SEED, contacts=DEFAULT_CONTACTS, separable=None):
    rng = np.random.default_rng(seed)
    yuk = rng.normal(size=3) * 100.0
    beta = rng.uniform(1.0, 3.0, size=3)
    lam = rng.normal(size=3) * 0.01
    if separable is not None:
        lam = np.asarray(separable, dtype=float)
    return np.concatenate([yuk, np.asarray(contacts, float), lam, beta])
