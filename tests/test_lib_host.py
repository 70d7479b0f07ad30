"""CPU tests of libtictac.so: exported symbols, host-side builders against the oracle,
error statuses.  No GPU compute is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import tictac as tt
from helpers import ROOT, mt_config, oracle_pw_full, synthetic_config
from oracle import potential as opot
from oracle import states as ost
from oracle import swp as osw
from tictac import _lib as L


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "tictac.h")).read()
    declared = set(re.findall(r"\b(tt_[A-Za-z0-9_]+)\s*\(", hdr))
    assert len(declared) >= 25
    lib = C.CDLL(L.LIB_PATH)
    missing = [n for n in sorted(declared) if not hasattr(lib, n)]
    assert not missing, missing
    assert set(L.PROTOTYPES) == declared


def _oracle_problem(rp, pot_oracle):
    pw = ost.construct_pw_states(rp.J_2N_max, rp.two_J_3N_max, bool(rp.isospin_breaking_1S0),
                                 swave_only=bool(rp.swave_only))
    fwp = ost.make_fwp(rp.Np_WP, rp.Nq_WP, rp.chebyshev_s, rp.chebyshev_t, rp.Np_per_WP)
    Vu, Vc = osw.potential_matrices(fwp, pw, pot_oracle, bool(rp.tensor_force), bool(rp.isospin_breaking_1S0),
                                    rp.J_2N_max)
    swp = osw.make_swp_states(fwp, pw, Vu, Vc, bool(rp.tensor_force), bool(rp.isospin_breaking_1S0))
    return pw, fwp, Vu, Vc, swp


def _compare(prob, rp, pot_oracle, T_lab):
    pw, fwp, Vu, Vc, swp = _oracle_problem(rp, pot_oracle)
    d = prob.pw_arrays()
    for k, ok in (("L_2N", pw.L), ("S_2N", pw.S), ("J_2N", pw.J), ("T_2N", pw.T), ("L_1N", pw.l),
                  ("two_J_1N", pw.two_j), ("two_J_3N", pw.two_J3), ("two_T_3N", pw.two_T3), ("P_3N", pw.P3)):
        assert np.array_equal(d[k], ok), k
    assert list(d["chn_3N_idx"]) == list(pw.chn_idx)
    assert np.allclose(prob.p_edges, fwp.p_edges, rtol=1e-15)
    Np = rp.Np_WP
    Vlu = prob.V_unco[:prob.nu * Np * Np].reshape(prob.nu, Np, Np)
    Vlc = prob.V_coup[:prob.nc * 4 * Np * Np].reshape(prob.nc, 2 * Np, 2 * Np)
    for a, b in ((Vlu, Vu), (Vlc, Vc)):
        used = ~np.isnan(b)
        assert np.array_equal(used, ~np.isnan(a))
        if used.any():
            assert np.abs(a[used] - b[used]).max() <= 1e-12 * np.abs(b[used]).max()
    assert abs(prob.E_bound - swp.E_bound) < 1e-10 * abs(swp.E_bound)
    eu = prob.e_unco[:prob.nu * (Np + 1)].reshape(prob.nu, Np + 1)
    assert np.allclose(eu, swp.e_unco, rtol=1e-10, atol=1e-10)
    if prob.nc:
        ec = prob.e_coup[:prob.nc * 2 * (Np + 1)].reshape(prob.nc, 2 * (Np + 1))
        assert np.allclose(ec, swp.e_coup, rtol=1e-10, atol=1e-10)
    # eigenvectors agree up to the sign of each column
    Cl = prob.C_unco[:prob.nu * Np * Np].reshape(prob.nu, Np, Np)
    for c in range(prob.nu):
        if np.isnan(swp.C_unco[c]).any():
            continue
        s = np.sign(np.sum(Cl[c] * swp.C_unco[c], axis=0))
        assert np.allclose(Cl[c] * s[None, :], swp.C_unco[c], atol=1e-9)
    chan = prob.channel(0, T_lab)
    osh = osw.find_on_shell_bins(fwp.q_edges, swp.E_bound, T_lab)
    assert np.array_equal(chan.q_idx, osh.q_idx)
    assert np.allclose(chan.E_com, osh.E_com, rtol=1e-12)
    assert np.array_equal(chan.deut, osw.find_deuteron_channels(pw.channel(0)))


def test_host_builders_malfliet_tjon():
    rp, pot, T = mt_config(N=16)
    _compare(tt.Problem(rp, pot), rp, opot.MalflietTjon(), T)


def test_host_builders_synthetic_small():
    rp, pot, T = synthetic_config(N=6, T_lab=(60.0,))
    rp.Np_per_WP = 4
    prob = tt.Problem(rp, pot)
    theta, beta = opot.synthetic_theta(7)
    _compare(prob, rp, opot.Synthetic(theta, beta), T)


def test_synthetic_potential_values_match_oracle():
    prm = tt.synthetic_params(7)
    lib_pot = tt.Potential("synthetic", prm)
    theta, beta = opot.synthetic_theta(7)
    orc = opot.Synthetic(theta, beta)
    rng = np.random.default_rng(5)
    for _ in range(40):
        pi, po = rng.uniform(5, 900, 2)
        S, J = int(rng.integers(0, 2)), int(rng.integers(0, 4))
        for Tz in (-1, 0):
            a = lib_pot.V(pi, po, True, S, J, 0, Tz)
            b = np.array(orc.V(pi, po, True, S, J, 0, Tz))
            assert np.allclose(a, b, rtol=1e-12, atol=1e-300)


def test_error_statuses():
    rp = tt.run_params(two_J_3N_max=2)
    with pytest.raises(tt.TictacError) as e:
        tt.Problem(rp, tt.Potential("malfliet_tjon"))
    assert e.value.code == L.TT_ERR_ARG
    with pytest.raises(tt.TictacError) as e:
        tt.Potential("no_such_model")
    assert e.value.code == L.TT_ERR_ARG
    # seed 4 binds unphysical states in the synthetic potential
    rp, _, _ = synthetic_config(N=20)
    with pytest.raises(tt.TictacError) as e:
        tt.Problem(rp, tt.Potential("synthetic", tt.synthetic_params(4)))
    assert e.value.code == L.TT_ERR_PHYSICS
    # T_lab above the WP space
    rp, pot, _ = mt_config(N=16)
    prob = tt.Problem(rp, pot)
    ch = prob.channel(0, [1.0e5])
    assert ch.nE == 0
    # no tensor force with J2max >= 2: the reference's 2N index overruns its arrays; rejected here
    rp = tt.run_params(two_J_3N_max=1, J_2N_max=2, Np_WP=8, Nq_WP=8, tensor_force=0, isospin_breaking_1S0=0)
    with pytest.raises(tt.TictacError) as e:
        tt.Problem(rp, tt.Potential("malfliet_tjon"))
    assert e.value.code == L.TT_ERR_ARG


def test_synthetic_potential_is_linear_in_its_strengths():
    """V_WP(theta) = sum_i theta_i V_i with the basis of tt_potential_basis_matrices (the
    premise of tt_solve_gradient)."""
    rp, pot, _ = synthetic_config(N=8, J_2N_max=3, two_J_3N_max=1)
    prob = tt.Problem(rp, pot)
    Vb_u, Vb_c = prob.basis_matrices(12)
    th = tt.synthetic_params()[:12]
    n = rp.Np_WP
    Vu = prob.V_unco[:prob.nu * n * n]
    Vc = prob.V_coup[:prob.nc * 4 * n * n]
    assert np.abs(np.tensordot(th, Vb_u, 1) - Vu).max() <= 1e-14 * np.abs(Vu).max()
    assert np.abs(np.tensordot(th, Vb_c, 1) - Vc).max() <= 1e-14 * np.abs(Vc).max()
    assert np.array_equal(prob.pot.params, tt.synthetic_params())   # restored


_V_SCRIPT = """
import sys, numpy as np
sys.path[:0] = sys.argv[2:]
import tictac as tt
from helpers import synthetic_config
rp, pot, _ = synthetic_config(N=12, J_2N_max=3, two_J_3N_max=1)
prob = tt.Problem(rp, pot)
np.save(sys.argv[1], np.concatenate([prob.V_unco, prob.V_coup]))
"""


def test_potential_matrices_do_not_depend_on_the_thread_count(tmp_path):
    """V_WP rows are built on several host threads; each element is summed in one fixed order,
    so one thread and eight give the same bits."""
    import subprocess
    import sys
    out = {}
    for n in (1, 8):
        f = tmp_path / ("v%d.npy" % n)
        env = dict(os.environ, TT_HOST_THREADS=str(n))
        subprocess.run([sys.executable, "-c", _V_SCRIPT, str(f), os.path.join(ROOT, "tic-tac_amd"),
                        os.path.join(ROOT, "tests")], env=env, check=True, timeout=120)
        out[n] = np.load(f)
    assert out[1].size > 0 and np.array_equal(out[1], out[8])


def test_breakup_cells_match_oracle():
    """tt_find_breakup_bins against the oracle on the library's own SWP states (tensor force
    and isospin-breaking 1S0: coupled p-bin edges), energies below and above the breakup
    threshold."""
    from helpers import oracle_pw, oracle_swp
    rp, pot_, _ = synthetic_config(N=10, two_J_3N_max=3)
    prob = tt.Problem(rp, pot_)
    T = [0.5, 10.0, 60.0]
    for c in (0, 2):
        ch = prob.channel(c, T, breakup=True)
        ptr, cells = osw.find_breakup_bins(oracle_pw(ch), oracle_swp(prob), ch.E_com, True, True)
        assert ch.n_bu == len(cells) > 0
        assert np.array_equal(ch.bu_ptr, ptr)
        assert np.array_equal(ch.bu_cells, cells)
        assert ptr[1] == 0   # the lowest bin lies below the threshold


def test_midpoint_potential_matrices_match_oracle():
    """The library's mid-point V_WP (run option midpoint_approx, CPP/make_potential_matrix.cpp:271-329)
    against the oracle's restatement, on the synthetic potential with tensor force and isospin
    breaking; the builders are called directly (a mid-point V_WP of this potential need not bind the
    deuteron alone, which the SWP step checks)."""
    N = 8
    rp = tt.run_params(two_J_3N_max=1, J_2N_max=3, Np_WP=N, Nq_WP=N, chebyshev_s=tt.chebyshev_scale(N),
                       midpoint_approx=1)
    pot = tt.Potential("synthetic", tt.synthetic_params(7))
    lib = L.lib()
    pw, fwp = L.PwStates(), L.FwpStates()
    L.check(lib.tt_construct_pw_states(C.byref(rp), C.byref(pw)))
    L.check(lib.tt_make_fwp_statespace(C.byref(rp), C.byref(fwp)))
    nu, nc = C.c_int(), C.c_int()
    lib.tt_num_2N_states(C.byref(rp), C.byref(nu), C.byref(nc))
    Vu = np.full(nu.value * N * N, np.nan)
    Vc = np.full(nc.value * 4 * N * N, np.nan)
    L.check(lib.tt_calculate_potential_matrices(C.byref(rp), C.byref(fwp), C.byref(pw), pot._h, L.dptr(Vu),
                                                L.dptr(Vc)))
    lib.tt_pw_states_free(C.byref(pw))
    lib.tt_fwp_states_free(C.byref(fwp))
    pw_o = ost.construct_pw_states(3, 1, True)
    fwp_o = ost.make_fwp(N, N, rp.chebyshev_s, rp.chebyshev_t, rp.Np_per_WP)
    theta, beta = opot.synthetic_theta(7)
    Vuo, Vco = osw.potential_matrices(fwp_o, pw_o, opot.Synthetic(theta, beta), True, True, 3,
                                      midpoint_approx=True)
    for a, b in ((Vu.reshape(Vuo.shape), Vuo), (Vc.reshape(Vco.shape), Vco)):
        used = ~np.isnan(b)
        assert np.array_equal(used, ~np.isnan(a)) and used.any()
        assert np.abs(a[used] - b[used]).max() <= 1e-12 * np.abs(b[used]).max()
