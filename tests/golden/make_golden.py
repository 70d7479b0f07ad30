"""Writes tests/golden/mt_np16_elastic.npz: oracle outputs for the Malfliet-Tjon S-wave case
(J^Pi = 1/2+, N_alpha = 2, Np = Nq = 16, T_lab = 13 MeV), built from the oracle alone.

The reference has no tests or golden vectors for this path, and its build needs MKL, GSL
and HDF5, which are absent, so these fixtures are the oracle's own outputs.  Once the
oracle's pins (tests/test_oracle_pins.py) pass, they freeze its results: a later change to
the oracle shows up as a fixture mismatch.  Everything stored is independent of the
arbitrary signs of the SWP eigenvectors (nd = 1 here).

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import p123 as pm  # noqa: E402
from oracle import potential as po  # noqa: E402
from oracle import resolvent as rs  # noqa: E402
from oracle import solve as so  # noqa: E402
from oracle import states as st  # noqa: E402
from oracle import swp as sw  # noqa: E402

N, T_LAB, NPHI, NX = 16, 13.0, 48, 48
PATH = os.path.join(HERE, "mt_np16_elastic.npz")


def compute():
    pw = st.construct_pw_states(1, 1, False, swave_only=True)
    fwp = st.make_fwp(N, N, st.chebyshev_scale_for_top(N, 8 * st.hbarc))
    Vu, Vc = sw.potential_matrices(fwp, pw, po.MalflietTjon(), False, False, 1)
    swp = sw.make_swp_states(fwp, pw, Vu, Vc, False, False)
    os_ = sw.find_on_shell_bins(fwp.q_edges, swp.E_bound, [T_LAB])
    deut = sw.find_deuteron_channels(pw)
    E = os_.E_com + swp.E_bound
    G = np.stack([rs.resolvent_array(e, pw, swp, False, False) for e in E])
    ent, _ = pm.p123_rows(pw, fwp.p_edges, fwp.q_edges, NPHI, NX)
    dim = pw.Nalpha * N * N
    val, row, ptr = pm.to_csc(ent, dim)
    K = so.KernelA(pw, swp, Vu, Vc, so.p_matrix(val, row, ptr, dim), False, False)
    U = so.dense_solve(K.dense(), G, deut, os_.q_idx, N, N)
    a = so.neumann_terms(K, G, deut, os_.q_idx, N, N, 7)
    sel = np.arange(0, len(val), 7)
    return {
        "E_bound": np.array(swp.E_bound), "e_SWP_unco": swp.e_unco, "q_edges": fwp.q_edges,
        "p_edges": fwp.p_edges, "q_idx": os_.q_idx, "E_com": os_.E_com, "deut": deut,
        "G": G, "p123_colptr": ptr.astype(np.uint64), "p123_row": row.astype(np.int32),
        "p123_val_sample_idx": sel, "p123_val_sample": val[sel],
        "p123_abs_sum": np.array(np.abs(val).sum()), "U": U, "neumann": a,
    }


if __name__ == "__main__":
    np.savez_compressed(PATH, **compute())
    print("wrote", PATH, os.path.getsize(PATH), "bytes")
