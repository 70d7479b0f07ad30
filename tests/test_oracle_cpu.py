"""The oracle's C++/OpenMP restatement (oracle/cpu_solve.cpp, the CPU baseline) against the numpy
oracle on the same inputs: its GMRES U against the dense solve (1e-10), and its Neumann terms,
with every column of A regenerated per term as the reference does, against the numpy rows x A
(1e-12 of the largest term)."""
import numpy as np
import pytest

from oracle import cpu
from oracle import p123 as pm
from oracle import potential as pot
from oracle import resolvent as rs
from oracle import solve as so
from oracle import states as st
from oracle import swp as sw


def _case(kind):
    if kind == "mt":
        N, J2, tf, ib = 10, 1, False, False
        pwf = st.construct_pw_states(1, 1, False, swave_only=True)
        model = pot.MalflietTjon()
    else:
        N, J2, tf, ib = 6, 3, True, True
        pwf = st.construct_pw_states(3, 1)
        theta, beta = pot.synthetic_theta()
        model = pot.Synthetic(theta, beta)
    fwp = st.make_fwp(N, N, st.chebyshev_scale_for_top(N, 8 * st.hbarc))
    Vu, Vc = sw.potential_matrices(fwp, pwf, model, tf, ib, J2)
    swp = sw.make_swp_states(fwp, pwf, Vu, Vc, tf, ib)
    pw = pwf.channel(0)
    dim = pw.Nalpha * N * N
    ent, _ = pm.p123_rows(pw, fwp.p_edges, fwp.q_edges, 16, 16)
    v, r, p = pm.to_csc(ent, dim)
    os_ = sw.find_on_shell_bins(fwp.q_edges, swp.E_bound, [13.0, 40.0])
    deut = sw.find_deuteron_channels(pw)
    G = np.stack([rs.resolvent_array(e, pw, swp, tf, ib) for e in os_.E_com + swp.E_bound])
    K = so.KernelA(pw, swp, Vu, Vc, so.p_matrix(v, r, p, dim), tf, ib)
    prob = cpu.Problem(pw, swp, Vu, Vc, v, r, p, tf, ib)
    return N, K, prob, G, deut, os_.q_idx


@pytest.mark.parametrize("kind", ["mt", "synthetic"])
def test_cpu_restatement_matches_numpy_oracle(kind):
    N, K, prob, G, deut, q_idx = _case(kind)
    Ud = so.dense_solve(K.dense(), G, deut, q_idx, N, N)
    U, rc, its, res = prob.gmres_U(G, deut, q_idx, tol=1e-13)
    assert rc == 0 and (res <= 1e-13).all()
    assert np.abs(U - Ud).max() <= 1e-10 * np.abs(Ud).max()
    n = 9
    a, done, spt = prob.neumann_terms(G, deut, q_idx, n)
    assert done == n and spt > 0
    ao = so.neumann_terms(K, G, deut, q_idx, N, N, n)
    for t in range(n):
        assert np.abs(a[..., t] - ao[..., t]).max() <= 1e-12 * np.abs(ao[..., t]).max(), t
