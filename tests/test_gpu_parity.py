"""GPU parity of the gfx950 solve path against the oracle, through the C-ABI.

Tolerances (floating point; the reference itself is FP64):
  resolvent            64 ulp of the summed term magnitudes of the reference's closed form
                       (oracle.resolvent.resolvent_term_scale): the continuum formula is a second
                       difference of v ln|v| and is ill-conditioned for narrow bins far off shell
  P123 elements        structure identical; values 1e-13 relative to max|P|
  A application        1e-13 (FP64), 1e-5 (FP32 products)
  converged U          1e-10 relative (dense, GMRES); mixed vs FP64 1e-8 (the workload's target)
  Pade                 Neumann terms 1e-10; U within 1e-5 of the converged solve, the accuracy of
                       the reference's own stopping rule (1e-6 relative / 1e-7 absolute)
"""
import os

import numpy as np
import pytest
import scipy.sparse.linalg as sla

import tictac as tt
from helpers import flags, mt_config, oracle_pw, oracle_swp, oracle_V, synthetic_config
from oracle import p123 as pm
from oracle import resolvent as rs
from oracle import solve as so

pytestmark = pytest.mark.gpu


class Case:
    pass


def _setup(ctx, rp, pot, T, oracle_p123=None, channel=0):
    c = Case()
    c.prob = tt.Problem(rp, pot)
    c.ch = c.prob.channel(channel, T)
    c.P = c.ch.build_p123(ctx)
    c.G = c.ch.resolvent(ctx)
    c.och = oracle_pw(c.ch)
    c.oswp = oracle_swp(c.prob)
    c.Vu, c.Vc = oracle_V(c.prob)
    c.tf, c.ib = flags(c.prob)
    c.val, c.row, c.ptr = c.P.to_host()
    c.K = so.KernelA(c.och, c.oswp, c.Vu, c.Vc, so.p_matrix(c.val, c.row, c.ptr, c.ch.dim), c.tf, c.ib)
    return c


@pytest.fixture(scope="module")
def c1(ctx):
    rp, pot, T = mt_config(N=16)
    c = _setup(ctx, rp, pot, T)
    c.A = c.K.dense()
    c.Ud = so.dense_solve(c.A, c.G, c.ch.deut, c.ch.q_idx, rp.Nq_WP, rp.Np_WP)
    return c


@pytest.fixture(scope="module")
def c2(ctx):
    rp, pot, T = synthetic_config(N=30)
    return _setup(ctx, rp, pot, T)


# ------------------------------------------------------------------ Malfliet-Tjon, N = 16
def _check_resolvent(c):
    for j, E in enumerate(c.ch.energies()):
        Go, scale = rs.resolvent_array(E, c.och, c.oswp, c.tf, c.ib, with_scale=True)
        err = np.abs(c.G[j] - Go)
        assert (err <= 64 * np.finfo(float).eps * scale + 1e-15 * np.abs(Go)).all(), err.max()
        # well-conditioned elements (the bulk) agree to 1e-12 relative
        good = scale <= 1e3 * np.abs(Go)
        assert good.mean() > 0.2
        assert (err[good] <= 1e-12 * np.abs(Go[good]) + 1e-15).all()


def test_resolvent_c1(c1):
    _check_resolvent(c1)


def test_p123_c1_matches_oracle_structure_and_values(c1):
    rp = c1.prob.rp
    ent, mask = pm.p123_rows(c1.och, c1.prob.p_edges, c1.prob.q_edges, rp.Nphi, rp.Nx)
    vo, ro, po = pm.to_csc(ent, c1.ch.dim)
    assert c1.P.stats.mask_density == pytest.approx(mask.mean(), abs=0)
    assert np.array_equal(po, c1.ptr) and np.array_equal(ro, c1.row)
    assert np.abs(vo - c1.val).max() <= 1e-13 * np.abs(vo).max()


@pytest.mark.parametrize("n_ang", [47, 48, 64])
def test_p123_mask_forms_match_oracle(ctx, n_ang):
    """Nphi = Nx = 48 builds the overlap mask from per-bra-bin tables in LDS (38 KB); 64 would
    need 68 KB and takes the per-cell form.  The angular basis takes two points per lane at even
    Nx and one at odd Nx (47).  All give the oracle's mask, structure and values."""
    rp, pot, T = synthetic_config(N=8, J_2N_max=1, Nphi=n_ang, Nx=n_ang)
    prob = tt.Problem(rp, pot)
    ch = prob.channel(0, T)
    P = ch.build_p123(ctx)
    val, row, ptr = P.to_host()
    ent, mask = pm.p123_rows(oracle_pw(ch), prob.p_edges, prob.q_edges, rp.Nphi, rp.Nx)
    vo, ro, po = pm.to_csc(ent, ch.dim)
    assert P.stats.mask_density == pytest.approx(mask.mean(), abs=0)
    assert np.array_equal(po, ptr) and np.array_equal(ro, row)
    assert np.abs(vo - val).max() <= 1e-13 * np.abs(vo).max()


def test_apply_c1(ctx, c1):
    rng = np.random.default_rng(0)
    X = rng.normal(size=(3, c1.ch.dim)) + 1j * rng.normal(size=(3, c1.ch.dim))
    Gsel = np.stack([c1.G[0], c1.G[1], np.ones(c1.ch.dim)])
    Yo = (X * Gsel) @ c1.A
    Y = c1.ch.apply_rows(ctx, c1.P, X, G=c1.G, g_index=[0, 1, -1])
    assert np.abs(Y - Yo).max() <= 1e-13 * np.abs(Yo).max()
    Y32 = c1.ch.apply_rows(ctx, c1.P, X, G=c1.G, g_index=[0, 1, -1], fp32=True)
    assert np.abs(Y32 - Yo).max() <= 1e-5 * np.abs(Yo).max()
    Ycsc = c1.ch.apply_rows(ctx, c1.P, X, G=c1.G, g_index=[0, 1, -1], csc=True)
    assert np.abs(Ycsc - Yo).max() <= 1e-13 * np.abs(Yo).max()


def test_c1_against_golden_fixture(ctx, c1):
    """The library's SWP edges, P123, G and converged U at the Malfliet-Tjon N = 16 case equal
    the committed fixture (tests/golden/mt_np16_elastic.npz)."""
    from test_golden import load
    g = load()
    assert abs(c1.prob.E_bound - float(g["E_bound"])) <= 1e-12 * abs(float(g["E_bound"]))
    assert np.array_equal(np.asarray(c1.ch.q_idx), g["q_idx"])
    assert np.array_equal(c1.ptr, g["p123_colptr"]) and np.array_equal(c1.row, g["p123_row"])
    vs = c1.val[g["p123_val_sample_idx"]]
    assert np.abs(vs - g["p123_val_sample"]).max() <= 1e-13 * np.abs(g["p123_val_sample"]).max()
    assert abs(np.abs(c1.val).sum() - float(g["p123_abs_sum"])) <= 1e-12 * float(g["p123_abs_sum"])
    # resolvent end to end: the fixture's G comes from the oracle's own SWP edges (numpy eigh),
    # which differ from the library's (Jacobi) at ~1e-13; the continuum closed form amplifies
    # that by its conditioning, so the bound is 1e-9 of max|G| here (parity on identical
    # edges, to 64 ulp of the term sizes, is test_resolvent_c1)
    assert np.abs(c1.G - g["G"]).max() <= 1e-9 * np.abs(g["G"]).max()
    U, _ = c1.ch.solve(ctx, c1.P, c1.G, method=tt.SOLVE_GMRES)
    assert np.abs(U - g["U"]).max() <= 1e-10 * np.abs(g["U"]).max()


def test_column_operator_c2(ctx, c2):
    """The column form Y_k = A (G_k .* X_k) (products with P123 itself through the transposed
    cell blocks, used by the adjoint) against the oracle's A X."""
    rng = np.random.default_rng(3)
    X = rng.normal(size=(3, c2.ch.dim)) + 1j * rng.normal(size=(3, c2.ch.dim))
    g = [0, -1, 1]
    Gsel = np.stack([c2.G[k] if k >= 0 else np.ones(c2.ch.dim) for k in g])
    Yo = c2.K.col_apply((X * Gsel).T).T
    Y = c2.ch.apply_rows(ctx, c2.P, X, G=c2.G, g_index=g, column=True)
    assert np.abs(Y - Yo).max() <= 1e-13 * np.abs(Yo).max()
    Y32 = c2.ch.apply_rows(ctx, c2.P, X, G=c2.G, g_index=g, column=True, fp32=True)
    assert np.abs(Y32 - Yo).max() <= 1e-5 * np.abs(Yo).max()


def test_uploaded_reference_csc_gives_same_solution(ctx, c1):
    """A CSC supplied by the caller (the reference's format, here the oracle's P123) is
    converted to cell blocks on first use and solves to the same U."""
    rp = c1.prob.rp
    ent, _ = pm.p123_rows(c1.och, c1.prob.p_edges, c1.prob.q_edges, rp.Nphi, rp.Nx)
    vo, ro, po = pm.to_csc(ent, c1.ch.dim)
    Pu = tt.P123.from_host(ctx, c1.ch.dim, vo, ro, po)
    U, st = c1.ch.solve(ctx, Pu, c1.G, method=tt.SOLVE_GMRES)
    assert np.abs(U - c1.Ud).max() <= 1e-10 * np.abs(c1.Ud).max()


@pytest.mark.parametrize("method", [tt.SOLVE_DENSE, tt.SOLVE_GMRES, tt.SOLVE_MIXED])
def test_converged_solves_c1(ctx, c1, method):
    U, st = c1.ch.solve(ctx, c1.P, c1.G, method=method)
    assert np.abs(U - c1.Ud).max() <= 1e-10 * np.abs(c1.Ud).max()
    if method != tt.SOLVE_DENSE:
        assert st.residual <= 1e-12


def test_pade_c1(ctx, c1):
    U, st, a = c1.ch.solve(ctx, c1.P, c1.G, method=tt.SOLVE_PADE, neumann=True)
    ao = so.neumann_terms(c1.K, c1.G, c1.ch.deut, c1.ch.q_idx, c1.prob.rp.Nq_WP, c1.prob.rp.Np_WP, a.shape[-1])
    done = a != 0
    assert done[..., :3].all()
    for n in range(a.shape[-1]):
        m = done[..., n]
        if m.any():
            assert np.abs(a[..., n][m] - ao[..., n][m]).max() <= 1e-10 * np.abs(ao[..., n]).max()
    assert np.abs(U - c1.Ud).max() <= 1e-5 * np.abs(c1.Ud).max()
    # the library's resummation is the reference's Pade rules applied to its own terms: U is the
    # oracle's [k/k] approximant of the same terms (CPP/solve_faddeev.cpp:271-296) for a k the
    # selection rules (lines 1236-1298) can pick.  The high-order approximants are ratios of
    # near-singular Hankel determinants: the two implementations' rounding moves them by far more
    # than the terms' rounding (~1e-7 at [12/12]), and it decides near-ties of the rule (which
    # approximant differs least from its predecessor) and where the rule stops.  So: sens[k] = the
    # largest change of the oracle's [k/k] under random relative changes of 1e-13 in the terms;
    # the candidates are the orders k the computed terms reach whose difference d[k] to the
    # predecessor is within 20% plus the sensitivities of the smallest one; U must equal one of
    # them to 10x its sensitivity.
    rng = np.random.default_rng(2)
    for dr in range(a.shape[0]):
        for dc in range(a.shape[1]):
            for j in range(a.shape[2]):
                ae = a[dr, dc, j]
                K = min(14, (int(np.count_nonzero(ae)) - 1) // 2)   # [k/k] needs terms 0..2k
                PA = [so.pade_approximant(ae, k, k, 1.0) for k in range(K + 1)]
                sens = [max(abs(so.pade_approximant(ae * (1 + 1e-13 * rng.normal(size=ae.shape)), k, k, 1.0)
                                - PA[k]) for _ in range(8)) for k in range(K + 1)]
                d = [abs(PA[k] - (PA[k - 1] if k else 0.0)) for k in range(K + 1)]
                noise = [sens[k] + (sens[k - 1] if k else 0.0) for k in range(K + 1)]
                kmin = min(range(K + 1), key=lambda k: d[k])
                ok = [k for k in range(K + 1) if d[k] <= 1.2 * d[kmin] + 2 * (noise[k] + noise[kmin])]
                good = any(abs(U[dr, dc, j] - PA[k]) <= 1e-10 * abs(PA[k]) + 10 * sens[k] for k in ok)
                assert good, (dr, dc, j, kmin, ok)


# ---------------------------------------------------------- synthetic potential, N = 30, N_alpha = 27
def test_p123_c2_sampled_rows(c2):
    rp = c2.prob.rp
    bins = [(3, 4), (21, 22)]
    ent, _ = pm.p123_rows(c2.och, c2.prob.p_edges, c2.prob.q_edges, rp.Nphi, rp.Nx, row_bins=bins)
    Pr = so.p_matrix(c2.val, c2.row, c2.ptr, c2.ch.dim).tocsr()
    N = rp.Np_WP
    rows = [a * N * N + qp * N + pp for (qp, pp) in bins for a in range(c2.ch.Nalpha)]
    assert sum(Pr[r].nnz for r in rows) == len(ent)
    mx = max(abs(v) for v in ent.values())
    for (r, cidx), v in ent.items():
        assert abs(Pr[r, cidx] - v) <= 1e-13 * mx


def test_p123_c2_placeholder_structure(ctx, c2):
    """production_run = 0 (CPP/make_permutation_matrix.cpp:817-822): cos(row)cos(col) on every
    overlapping cell and alpha pair; the production matrix is its exact-zero-dropped subset."""
    Q = c2.ch.build_p123(ctx, production_run=False)
    v, r, p = Q.to_host()
    Na = c2.ch.Nalpha
    assert Q.nnz == Q.stats.n_cells * Na * Na
    cols = np.repeat(np.arange(c2.ch.dim), np.diff(p.astype(np.int64)))
    assert np.abs(v - np.cos(r.astype(float)) * np.cos(cols.astype(float))).max() < 1e-15
    # rows strictly ascending inside each column
    d = np.diff(r.astype(np.int64))
    starts = p[1:-1].astype(np.int64)
    ok = np.ones(len(d), bool)
    ok[starts[(starts > 0) & (starts < len(r))] - 1] = False
    assert (d[ok] > 0).all()
    prod_cols = np.repeat(np.arange(c2.ch.dim), np.diff(c2.ptr.astype(np.int64)))
    key_q = set((cols * c2.ch.dim + r).tolist())
    key_p = (prod_cols * c2.ch.dim + c2.row).tolist()
    assert all(k in key_q for k in key_p[::97])


def test_resolvent_c2(c2):
    _check_resolvent(c2)


def test_apply_c2(ctx, c2):
    rng = np.random.default_rng(1)
    X = rng.normal(size=(4, c2.ch.dim)) + 1j * rng.normal(size=(4, c2.ch.dim))
    g = [0, 1, 0, -1]
    Gsel = np.stack([c2.G[k] if k >= 0 else np.ones(c2.ch.dim) for k in g])
    Yo = c2.K.row_apply(X * Gsel)
    Y = c2.ch.apply_rows(ctx, c2.P, X, G=c2.G, g_index=g)
    assert np.abs(Y - Yo).max() <= 1e-13 * np.abs(Yo).max()
    Ycsc = c2.ch.apply_rows(ctx, c2.P, X, G=c2.G, g_index=g, csc=True)
    assert np.abs(Ycsc - Yo).max() <= 1e-13 * np.abs(Yo).max()
    Y32 = c2.ch.apply_rows(ctx, c2.P, X, G=c2.G, g_index=g, fp32=True)
    assert np.abs(Y32 - Yo).max() <= 1e-5 * np.abs(Yo).max()


@pytest.fixture(scope="module")
def c3(ctx):
    """A J^Pi = 3/2 channel with N_alpha = 47 at Np = Nq = 12: the 3-tile (MT = 3) MFMA path."""
    rp, pot, T = synthetic_config(N=12, two_J_3N_max=3, T_lab=(10.0, 50.0))
    prob = tt.Problem(rp, pot)
    big = max(range(prob.n_channels), key=lambda i: prob.channel(i, T).Nalpha)
    c = _setup(ctx, rp, pot, T, channel=big)
    assert c.ch.Nalpha > 32
    return c


def test_apply_large_channel(ctx, c3):
    """The cell-block product at N_alpha = 47 (3 M tiles) with 18 right-hand sides (48-wide N
    tiles) against the oracle's row application: 1e-13 in FP64, 1e-5 with FP32 products."""
    rng = np.random.default_rng(5)
    r = 18
    X = rng.normal(size=(r, c3.ch.dim)) + 1j * rng.normal(size=(r, c3.ch.dim))
    g = [k % c3.ch.nE for k in range(r)]
    Gsel = np.stack([c3.G[k] for k in g])
    Yo = c3.K.row_apply(X * Gsel)
    Y = c3.ch.apply_rows(ctx, c3.P, X, G=c3.G, g_index=g)
    assert np.abs(Y - Yo).max() <= 1e-13 * np.abs(Yo).max()
    Y32 = c3.ch.apply_rows(ctx, c3.P, X, G=c3.G, g_index=g, fp32=True)
    assert np.abs(Y32 - Yo).max() <= 1e-5 * np.abs(Yo).max()


def test_mixed_solve_large_channel(ctx, c3):
    """Iterative refinement (FP32 products and basis) on the N_alpha = 47 channel reaches the
    FP64 GMRES answer to 1e-10."""
    U64, _ = c3.ch.solve(ctx, c3.P, c3.G, method=tt.SOLVE_GMRES)
    Umx, st = c3.ch.solve(ctx, c3.P, c3.G, method=tt.SOLVE_MIXED)
    assert st.n_apply_fp32 > 0
    assert np.abs(Umx - U64).max() <= 1e-10 * np.abs(U64).max()


def test_solves_c2_against_oracle_gmres(ctx, c2):
    ch, N = c2.ch, c2.prob.rp.Np_WP
    Uo = np.zeros((ch.nd, ch.nd, ch.nE), complex)
    for dr in range(ch.nd):
        for j in range(ch.nE):
            e = np.zeros(ch.dim)
            e[ch.deut[dr] * N * N + ch.q_idx[j] * N] = 1.0
            b = c2.K.row_apply(e[None, :])[0].astype(complex)
            g = c2.G[j]
            op = sla.LinearOperator((ch.dim, ch.dim), dtype=complex,
                                    matvec=lambda v, g=g: v - c2.K.row_apply((v * g)[None, :])[0])
            y, info = sla.gmres(op, b, rtol=1e-13, restart=60, maxiter=50)
            assert info == 0
            for dc in range(ch.nd):
                Uo[dr, dc, j] = y[ch.deut[dc] * N * N + ch.q_idx[j] * N]
    U64, st64 = ch.solve(ctx, c2.P, c2.G, method=tt.SOLVE_GMRES)
    assert np.abs(U64 - Uo).max() <= 1e-10 * np.abs(Uo).max()
    Umx, stmx = ch.solve(ctx, c2.P, c2.G, method=tt.SOLVE_MIXED)
    assert stmx.n_apply_fp32 > 0
    assert np.abs(Umx - U64).max() <= 1e-8 * np.abs(U64).max()
    Upa, _ = ch.solve(ctx, c2.P, c2.G, method=tt.SOLVE_PADE)
    assert np.abs(Upa - U64).max() <= 1e-5 * np.abs(U64).max()


def test_dense_solve_c2_on_the_device(ctx, c2):
    """The dense path (solve_dense=true) at dim 24,300: L = 1 - A G_j formed and LU-factorised on
    the device (9.4 GB per energy, partial pivoting, GEMM updates on the MFMA) agrees with the
    converged GMRES solve, which test_solves_c2_against_oracle_gmres pins to the oracle: 1e-10."""
    U64, _ = c2.ch.solve(ctx, c2.P, c2.G, method=tt.SOLVE_GMRES)
    Ud, _ = c2.ch.solve(ctx, c2.P, c2.G, method=tt.SOLVE_DENSE)
    assert np.abs(Ud - U64).max() <= 1e-10 * np.abs(U64).max()


def test_dense_solve_large_channel_against_oracle_dense(ctx, c3):
    """The device LU at N_alpha = 47, dim 6,768 (106 panels), against the oracle's numpy dense
    solve of the same A: 1e-10."""
    Uo = so.dense_solve(c3.K.dense(), c3.G, c3.ch.deut, c3.ch.q_idx, c3.prob.rp.Nq_WP, c3.prob.rp.Np_WP)
    Ud, _ = c3.ch.solve(ctx, c3.P, c3.G, method=tt.SOLVE_DENSE)
    assert np.abs(Ud - Uo).max() <= 1e-10 * np.abs(Uo).max()


def test_graph_replays_match_eager_launches(ctx, c2):
    """Each Arnoldi step is captured once as a HIP graph and replayed by later solves of the
    channel (the context keeps the channel's plan).  Every replay, with eager work and
    synchronisation in between, gives the result of plain stream launches to the solve
    tolerance (the Gram-Schmidt sums add block partials by device atomics, whose order, and
    so the last bits, vary from run to run; the iteration counts are equal)."""
    ch = c2.ch
    for method in (tt.SOLVE_GMRES, tt.SOLVE_MIXED):
        U_eager, st_e = ch.solve(ctx, c2.P, c2.G, method=method, graphs=False)
        for rep in range(3):
            G = ch.resolvent(ctx)   # eager work between the replays
            U_g, st_g = ch.solve(ctx, c2.P, G, method=method, graphs=True)
            assert st_g.iterations == st_e.iterations
            assert np.abs(U_g - U_eager).max() <= 1e-12 * np.abs(U_eager).max(), (method, rep)
            assert st_g.n_spmm <= st_g.n_apply


def test_device_resolvent_matches_host_G(ctx, c2):
    """resolvent_on_device computes G of the on-shell energies inside the solve; the result
    equals the solve fed the host G of tt_calculate_resolvent (same kernel, same values)."""
    U_h, _ = c2.ch.solve(ctx, c2.P, c2.G, method=tt.SOLVE_GMRES)
    U_d, _ = c2.ch.solve(ctx, c2.P, None, method=tt.SOLVE_GMRES, device_resolvent=True)
    assert np.abs(U_d - U_h).max() <= 1e-14 * np.abs(U_h).max()


def test_plan_follows_potential_and_p123_changes(ctx):
    """A cached plan is reused only for the same P123 object and the same V, C and channel:
    changing the potential strengths on the same P123, or rebuilding P123, gives the
    oracle's result for the new inputs."""
    rp, pot, T = synthetic_config(N=12, J_2N_max=1)
    th = tt.synthetic_params()
    for scale, rebuild in ((1.0, False), (1.05, False), (1.05, True), (1.0, False)):
        p = th.copy()
        p[:12] *= scale
        prob = tt.Problem(rp, tt.Potential("synthetic", p))
        ch = prob.channel(0, T)
        if rebuild or 'P' not in locals():
            P = ch.build_p123(ctx)
        G = ch.resolvent(ctx)
        U, st = ch.solve(ctx, P, G, method=tt.SOLVE_GMRES)
        val, row, ptr = P.to_host()
        tf, ib = flags(prob)
        Vu, Vc = oracle_V(prob)
        K = so.KernelA(oracle_pw(ch), oracle_swp(prob), Vu, Vc, so.p_matrix(val, row, ptr, ch.dim), tf, ib)
        Uo = so.dense_solve(K.dense(), G, ch.deut, ch.q_idx, rp.Nq_WP, rp.Np_WP)
        assert np.abs(U - Uo).max() <= 1e-10 * np.abs(Uo).max(), (scale, rebuild)


def test_batch_over_channels_and_energies_matches_single_solves(ctx):
    """tt_solve_batch: four 3N channels (1/2+-, 3/2+-), two lab energies each (all their
    on-shell bins in one block system, G on the device) give the single-channel results, and
    the FP64 items agree with the oracle's dense solve."""
    rp, pot, _ = synthetic_config(N=10, J_2N_max=1, two_J_3N_max=3)
    T = [5.0, 50.0]
    prob = tt.Problem(rp, pot)
    chans = [prob.channel(c, T) for c in range(prob.n_channels)]
    chans = [ch for ch in chans if ch.nd > 0]
    assert len(chans) >= 3
    Ps = [ch.build_p123(ctx) for ch in chans]
    for method, tol in ((tt.SOLVE_GMRES, 1e-10), (tt.SOLVE_MIXED, 1e-8)):
        out = prob.solve_batch(ctx, chans, Ps, method=method)
        for (U, st, status), ch, P in zip(out, chans, Ps):
            assert status == 0 and st.residual <= 1e-12
            U1, _ = ch.solve(ctx, P, None, method=tt.SOLVE_GMRES, device_resolvent=True)
            assert np.abs(U - U1).max() <= tol * np.abs(U1).max()
    ch, P = chans[0], Ps[0]
    val, row, ptr = P.to_host()
    tf, ib = flags(prob)
    Vu, Vc = oracle_V(prob)
    K = so.KernelA(oracle_pw(ch), oracle_swp(prob), Vu, Vc, so.p_matrix(val, row, ptr, ch.dim), tf, ib)
    G = ch.resolvent(ctx)
    Uo = so.dense_solve(K.dense(), G, ch.deut, ch.q_idx, rp.Nq_WP, rp.Np_WP)
    U = prob.solve_batch(ctx, [ch], [P], method=tt.SOLVE_GMRES)[0][0]
    assert np.abs(U - Uo).max() <= 1e-10 * np.abs(Uo).max()


def test_batch_step_graphs_match_plain_launches(ctx):
    """Replayed step graphs in a lockstep batch (the default for 4 or more channels): three
    batches with graphs, with plain-launch batches, single solves and host synchronisations in
    between, each equal to the plain-launch batch to the solve tolerance (1e-10 of max|U|: the
    dot products sum by device atomics, so their rounding differs from run to run)."""
    rp, pot, _ = synthetic_config(N=10, J_2N_max=1, two_J_3N_max=3)
    T = [5.0, 50.0]
    prob = tt.Problem(rp, pot)
    chans = [ch for ch in (prob.channel(c, T) for c in range(prob.n_channels)) if ch.nd > 0]
    assert len(chans) >= 4
    Ps = [ch.build_p123(ctx) for ch in chans]
    for method in (tt.SOLVE_GMRES, tt.SOLVE_MIXED):
        ref = [U for U, _, _ in prob.solve_batch(ctx, chans, Ps, method=method, graphs=False)]
        for rep in range(3):
            out = prob.solve_batch(ctx, chans, Ps, method=method, graphs=True if rep else None)
            for (U, st, status), Ur in zip(out, ref):
                assert status == 0 and st.residual <= 1e-12
                assert np.abs(U - Ur).max() <= 1e-10 * np.abs(Ur).max(), (method, rep)
            chans[rep % len(chans)].solve(ctx, Ps[rep % len(chans)], None, method=method, device_resolvent=True)
            ctx.synchronize()
            prob.solve_batch(ctx, chans[:2], Ps[:2], method=method, graphs=False)


@pytest.fixture(scope="module")
def grad_case(ctx):
    rp, pot, _ = synthetic_config(N=10, J_2N_max=1, two_J_3N_max=1)
    c = Case()
    c.rp, c.T = rp, [5.0, 50.0]
    c.prob = tt.Problem(rp, pot)
    c.ch = c.prob.channel(0, c.T)
    c.P = c.ch.build_p123(ctx)
    c.basis = c.prob.basis_matrices(12)
    c.theta = tt.synthetic_params()[:12]
    c.U, c.dU, c.st = c.prob.gradient(ctx, 0, c.P, c.theta, c.T, basis=c.basis)
    return c


def test_gradient_forward_is_the_solve(ctx, grad_case):
    c = grad_case
    U1, _ = c.ch.solve(ctx, c.P, None, method=tt.SOLVE_GMRES, device_resolvent=True)
    assert np.abs(c.U - U1).max() <= 1e-10 * np.abs(U1).max()
    assert c.st.residual_forward <= 1e-12 and c.st.residual_adjoint <= 1e-12


def test_gradient_matches_central_differences(ctx, grad_case):
    """Adjoint dU/dtheta_i against central differences of the library's own U(theta),
    h = 1e-6 |theta_i|; the FD error (truncation ~h^2, GMRES tolerance / h) is ~1e-6."""
    c = grad_case
    for i in range(12):
        h = 1e-6 * abs(c.theta[i])
        tp, tm = c.theta.copy(), c.theta.copy()
        tp[i] += h
        tm[i] -= h
        Up = c.prob.gradient(ctx, 0, c.P, tp, c.T, basis=c.basis, grad=False)[0]
        Um = c.prob.gradient(ctx, 0, c.P, tm, c.T, basis=c.basis, grad=False)[0]
        fd = (Up - Um) / (2 * h)
        scale = max(np.abs(fd).max(), np.abs(c.dU[i]).max(), 1e-300)
        assert np.abs(c.dU[i] - fd).max() <= 1e-5 * scale, (i, np.abs(c.dU[i] - fd).max() / scale)


def test_gradient_against_dense_reference(ctx, grad_case):
    """Dense FP64 reference: central differences of the oracle's own chain (numpy SWP
    eigenvectors and edges, closed-form resolvent, dense A = C^T 2P VC, dense LU solve) on the
    same basis potential; the oracle has independent eigen-solvers and no Krylov tolerance."""
    from oracle import states as ost
    from oracle import swp as osw
    c = grad_case
    rp = c.rp
    Np = rp.Np_WP
    nu, nc = c.prob.nu, c.prob.nc
    fwp = ost.make_fwp(Np, rp.Nq_WP, rp.chebyshev_s)
    pw_full = helpers_pw_full(c.prob)
    pw = oracle_pw(c.ch)
    val, row, ptr = c.P.to_host()
    Pm = so.p_matrix(val, row, ptr, c.ch.dim)
    tf, ib = flags(c.prob)

    def U_oracle(theta):
        Vu = np.tensordot(theta, c.basis[0], 1).reshape(nu, Np, Np)
        Vc = np.tensordot(theta, c.basis[1], 1).reshape(nc, 2 * Np, 2 * Np)
        swp = osw.make_swp_states(fwp, pw_full, Vu, Vc, tf, ib)
        os_ = osw.find_on_shell_bins(fwp.q_edges, swp.E_bound, c.T)
        deut = osw.find_deuteron_channels(pw)
        G = np.stack([rs.resolvent_array(e, pw, swp, tf, ib) for e in os_.E_com + swp.E_bound])
        A = so.KernelA(pw, swp, Vu, Vc, Pm, tf, ib).dense()
        return so.dense_solve(A, G, deut, os_.q_idx, rp.Nq_WP, Np)

    U0 = U_oracle(c.theta)
    assert np.abs(U0 - c.U).max() <= 1e-9 * np.abs(U0).max()
    for i in (0, 3, 4, 9):   # a Yukawa, the 1S0 and 3S1 contacts, a separable
        h = 1e-5 * abs(c.theta[i])
        tp, tm = c.theta.copy(), c.theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (U_oracle(tp) - U_oracle(tm)) / (2 * h)
        assert np.abs(c.dU[i] - fd).max() <= 1e-5 * np.abs(fd).max(), i


def test_gradient_torch_autograd_gradcheck(ctx, grad_case):
    """torch.autograd.gradcheck of the autograd.Function over relative parameter changes
    (theta = theta0 (1 + s)), so one eps suits strengths of every size."""
    import torch
    from tictac.adjoint import ChannelGradient, u_matrix
    c = grad_case
    solver = ChannelGradient(c.prob, ctx, 0, c.P, c.T, basis=c.basis)
    th0 = torch.tensor(c.theta, dtype=torch.float64)
    s = torch.zeros(12, dtype=torch.float64, requires_grad=True)
    fn = lambda s_: u_matrix(th0 * (1 + s_), solver).reshape(-1)
    assert torch.autograd.gradcheck(fn, (s,), eps=1e-6, atol=1e-7, rtol=1e-4, fast_mode=False)


def helpers_pw_full(prob):
    from helpers import oracle_pw_full
    return oracle_pw_full(prob)


def test_error_paths_on_device(ctx, c1):
    with pytest.raises(tt.TictacError) as e:
        tt.P123.from_host(ctx, 4, np.ones(2), np.array([0, 9], np.int32), np.array([0, 1, 2, 2, 2], np.uint64))
    assert e.value.code == tt._lib.TT_ERR_ARG
    # P123 of another dimension
    bad = tt.P123.from_host(ctx, 8, np.ones(1), np.array([0], np.int32), np.array([0] + [1] * 8, np.uint64))
    with pytest.raises(tt.TictacError) as e:
        c1.ch.solve(ctx, bad, c1.G)
    assert e.value.code == tt._lib.TT_ERR_ARG


def test_parameter_walk_matches_per_set_solves(ctx):
    """tt_parameter_walk (the reference's PSI loop for several channels at once, each set one
    lockstep batch, the next set's host work overlapped) gives, per parameter set and channel,
    the U of a separate per-set Problem solved channel by channel (FP64 GMRES: 1e-10; mixed:
    1e-8).  A set without the single 3S1 bound state fails alone, with TT_ERR_PHYSICS on its
    items, and the walk goes on."""
    from helpers import oracle_chain_U
    rp, pot, T = synthetic_config(N=12, two_J_3N_max=3, T_lab=(10.0, 50.0))
    th = tt.synthetic_params()
    sets = []
    for scale in (1.0, 1.02, 0.3, 0.99):
        p = th.copy()
        p[:12] *= scale
        sets.append(p)
    sets = np.array(sets)
    prob = tt.Problem(rp, tt.Potential("synthetic", th))
    chans = [0, 2]
    Ps = [prob.channel(c, T).build_p123(ctx) for c in chans]
    for method, tol in ((tt.SOLVE_GMRES, 1e-10), (tt.SOLVE_MIXED, 1e-8)):
        U, status, code = prob.parameter_walk(ctx, chans, Ps, sets, T, method=method)
        assert code == tt._lib.TT_ERR_PHYSICS
        assert (status[2] == tt._lib.TT_ERR_PHYSICS).all() and U[2][0] is None
        for s in (0, 1, 3):
            assert (status[s] == tt._lib.TT_OK).all()
            ref = tt.Problem(rp, tt.Potential("synthetic", sets[s]))
            for k, c in enumerate(chans):
                Ur, _ = ref.channel(c, T).solve(ctx, Ps[k], None, method=tt.SOLVE_GMRES, device_resolvent=True)
                assert U[s][k].shape == Ur.shape
                assert np.abs(U[s][k] - Ur).max() <= tol * np.abs(Ur).max(), (method, s, c)
                if s == 1 and k == 0 and method == tt.SOLVE_GMRES:
                    # and against the oracle's own chain for that set (SWP states, bins, G, dense A)
                    Uo = oracle_chain_U(ref, ref.channel(c, T), Ps[k], T)
                    assert np.abs(U[s][k] - Uo).max() <= 1e-9 * np.abs(Uo).max()


# ------------------------------------------------------------------ breakup amplitudes
@pytest.fixture(scope="module")
def c1bu(ctx, c1):
    """C1 with breakup cells at energies below (1 MeV) and above the breakup threshold; the
    oracle's converged breakup values from the dense A (NaN where a neighbour cell's flat
    index leaves the state space)."""
    N = c1.prob.rp.Np_WP
    T = [1.0, 13.0, 60.0]
    ch = c1.prob.channel(0, T, breakup=True)
    G = ch.resolvent(ctx)
    Ud = so.dense_solve_breakup(c1.A, G, ch.deut, ch.q_idx, N, N, ch.bu_ptr, ch.bu_cells)
    return ch, G, Ud, T


@pytest.mark.parametrize("method", [tt.SOLVE_DENSE, tt.SOLVE_GMRES, tt.SOLVE_MIXED])
def test_breakup_amplitudes_converged_c1(ctx, c1, c1bu, method):
    ch, G, Ud, _ = c1bu
    print("breakup", method, ch.n_bu, flush=True)
    ch.solve(ctx, c1.P, G, method=method)
    ok = ~np.isnan(Ud)
    assert ok.sum() > 0 and np.array_equal(np.isnan(ch.U_BU), ~ok)
    assert np.abs(ch.U_BU[ok] - Ud[ok]).max() <= 1e-10 * np.abs(Ud[ok]).max()


def test_breakup_amplitudes_pade_c1(ctx, c1, c1bu):
    """Pade breakup values: Neumann terms at the breakup columns resummed by the same rules as
    the elastic ones; within 1e-5 of the converged values (as the elastic Pade test), and the
    elastic U equals that of the same solve without breakup cells."""
    ch, G, Ud, T = c1bu
    U, st = ch.solve(ctx, c1.P, G, method=tt.SOLVE_PADE)
    ok = ~np.isnan(Ud)
    assert np.array_equal(np.isnan(ch.U_BU), ~ok)
    assert np.abs(ch.U_BU[ok] - Ud[ok]).max() <= 1e-5 * np.abs(Ud[ok]).max()
    plain = c1.prob.channel(0, T)
    U0, _ = plain.solve(ctx, c1.P, G, method=tt.SOLVE_PADE)
    assert np.array_equal(U, U0)


@pytest.mark.parametrize("method", [tt.SOLVE_GMRES, tt.SOLVE_MIXED])
def test_breakup_amplitudes_in_batch(ctx, c1, c1bu, method):
    ch, G, Ud, _ = c1bu
    ok = ~np.isnan(Ud)
    ch.U_BU = None
    res = c1.prob.solve_batch(ctx, [ch], [c1.P], method=method, tol=1e-12)
    assert res[0][2] == tt._lib.TT_OK
    assert np.array_equal(np.isnan(ch.U_BU), ~ok)
    assert np.abs(ch.U_BU[ok] - Ud[ok]).max() <= 1e-9 * np.abs(Ud[ok]).max()


# ------------------------------------------------------------------ plan identity and failure paths
def _oracle_dense_U(prob, ch, P, G, Vu=None, Vc=None):
    from helpers import oracle_dense_U
    return oracle_dense_U(prob, ch, P, G, Vu, Vc)


def test_plan_follows_sign_flips_of_V(ctx):
    """The cached plan compares V and C word for word: flipping the sign of V (or of one
    symmetric pair of its elements) with the SWP basis unchanged gives the oracle's result
    for the new V, not the cached operator's."""
    rp, pot, T = synthetic_config(N=10, J_2N_max=1)
    prob = tt.Problem(rp, pot)
    ch = prob.channel(0, T)
    P = ch.build_p123(ctx)
    G = ch.resolvent(ctx)
    Np = rp.Np_WP
    V0u, V0c = prob.V_unco.copy(), prob.V_coup.copy()
    try:
        U0, _ = ch.solve(ctx, P, G, method=tt.SOLVE_GMRES)
        prob.V_unco[:] = -V0u
        prob.V_coup[:] = -V0c
        U1, _ = ch.solve(ctx, P, G, method=tt.SOLVE_GMRES)
        assert np.abs(U1 - U0).max() > 1e-3 * np.abs(U0).max()
        Uo = _oracle_dense_U(prob, ch, P, G)
        assert np.abs(U1 - Uo).max() <= 1e-10 * np.abs(Uo).max()
        prob.V_unco[:] = V0u
        prob.V_coup[:] = V0c
        i, j = 2, 5   # one symmetric pair of the first uncoupled block
        prob.V_unco[i * Np + j] *= -1.0
        prob.V_unco[j * Np + i] *= -1.0
        U2, _ = ch.solve(ctx, P, G, method=tt.SOLVE_GMRES)
        Uo = _oracle_dense_U(prob, ch, P, G)
        assert np.abs(U2 - Uo).max() <= 1e-10 * np.abs(Uo).max()
    finally:
        prob.V_unco[:] = V0u
        prob.V_coup[:] = V0c


@pytest.mark.parametrize("method", [tt.SOLVE_GMRES, tt.SOLVE_MIXED])
def test_nan_resolvent_is_reported_not_converged(ctx, c1, method):
    """A NaN in G makes the residual NaN, which no tolerance accepts: the solve returns
    TT_ERR_NOT_CONVERGED instead of a NaN (or zero) U with TT_OK."""
    G = c1.G.copy()
    G[0, 7] = np.nan
    with pytest.raises(tt.TictacError) as e:
        c1.ch.solve(ctx, c1.P, G, method=method)
    assert e.value.code == tt._lib.TT_ERR_NOT_CONVERGED
    U, st = c1.ch.solve(ctx, c1.P, c1.G, method=method)   # the plan still solves good input
    assert np.abs(U - c1.Ud).max() <= 1e-10 * np.abs(c1.Ud).max()


def test_channel_above_64_alphas_takes_the_csc_gather(ctx):
    """J_2N_max = 4 gives a 5/2+ channel with N_alpha = 83, beyond the cell-block product:
    the default solve takes the CSC gather and agrees with the oracle's dense solve; the
    gradient, which needs the cell blocks, reports TT_ERR_UNSUPPORTED."""
    rp, pot, T = synthetic_config(N=6, J_2N_max=4, two_J_3N_max=5, T_lab=(10.0,))
    prob = tt.Problem(rp, pot)
    c = max(range(prob.n_channels), key=lambda i: prob.channel(i, T).Nalpha)
    ch = prob.channel(c, T)
    assert ch.Nalpha > 64
    P = ch.build_p123(ctx)
    G = ch.resolvent(ctx)
    U, st = ch.solve(ctx, P, G, method=tt.SOLVE_GMRES)
    Uo = _oracle_dense_U(prob, ch, P, G)
    assert np.abs(U - Uo).max() <= 1e-10 * np.abs(Uo).max()
    Umx, _ = ch.solve(ctx, P, G, method=tt.SOLVE_MIXED)
    assert np.abs(Umx - Uo).max() <= 1e-8 * np.abs(Uo).max()
    with pytest.raises(tt.TictacError) as e:
        prob.gradient(ctx, c, P, tt.synthetic_params()[:12], T, basis=prob.basis_matrices(12))
    assert e.value.code == tt._lib.TT_ERR_UNSUPPORTED


@pytest.fixture
def grouped_launches():
    """Restores TT_GROUPED and TT_LANES (read by every batch call) after the test."""
    old = {k: os.environ.get(k) for k in ("TT_GROUPED", "TT_LANES")}
    yield
    for k, v in old.items():
        if v is None:
            os.environ.pop(k, None)
        else:
            os.environ[k] = v


def test_grouped_batch_matches_lane_batch(ctx, grouped_launches):
    """TT_GROUPED=1: each stage of the lockstep steps (the block ops, the cell product per
    kernel instance, the Gram-Schmidt kernels) and of the FP64 residual applications is one
    launch for all channels of a lane: all channels of the batch with TT_LANES=1, two or three
    per lane with three lanes; the steps captured as graphs and replayed (three batches in a
    row, with single solves and host synchronisations between), and as plain launches.  The channels 2J = 1..7 of both parities at Np = Nq = 8
    (N_alpha 27 to 63: the two-, three- and four-M-tile product instances) give the six-lane
    batch's U to the solve tolerance (1e-10 of max|U|: the dot products sum by device
    atomics), FP64 and mixed, twice in a row with single solves in between; the FP64 result of
    the first channel equals the oracle's dense solve."""
    rp, pot, _ = synthetic_config(N=8, two_J_3N_max=7)
    T = [5.0, 50.0]
    prob = tt.Problem(rp, pot)
    chans = [ch for ch in (prob.channel(c, T) for c in range(prob.n_channels)) if ch.nd > 0]
    assert {27, 47, 63} <= {ch.Nalpha for ch in chans}
    Ps = [ch.build_p123(ctx) for ch in chans]
    for method in (tt.SOLVE_GMRES, tt.SOLVE_MIXED):
        os.environ["TT_GROUPED"] = "0"
        os.environ.pop("TT_LANES", None)
        ref = [U for U, _, _ in prob.solve_batch(ctx, chans, Ps, method=method)]
        os.environ["TT_GROUPED"] = "1"
        for lanes in ("1", "3"):
            os.environ["TT_LANES"] = lanes
            for graphs in (None, False):   # captured grouped steps (the default here), plain
                for rep in range(3):
                    out = prob.solve_batch(ctx, chans, Ps, method=method, graphs=graphs)
                    for (U, st, status), Ur, ch in zip(out, ref, chans):
                        assert status == 0 and st.residual <= 1e-12, (method, lanes, graphs, rep, ch.Nalpha)
                        assert np.abs(U - Ur).max() <= 1e-10 * np.abs(Ur).max(), (method, lanes, graphs, rep)
                    chans[rep].solve(ctx, Ps[rep], None, method=method, device_resolvent=True)
                    ctx.synchronize()
    os.environ["TT_LANES"] = "1"
    U0 = prob.solve_batch(ctx, chans, Ps, method=tt.SOLVE_GMRES)[0][0]
    c = _setup(ctx, rp, pot, T, channel=0)
    assert c.ch.Nalpha == chans[0].Nalpha
    Uo = so.dense_solve(c.K.dense(), c.G, c.ch.deut, c.ch.q_idx, rp.Nq_WP, rp.Np_WP)
    assert np.abs(U0 - Uo).max() <= 1e-10 * np.abs(Uo).max()


def test_grouped_gradient_matches(ctx, grad_case, grouped_launches):
    """TT_GROUPED=1 with one lane: the gradient's row and column solves (two operators of
    opposite orientation in one lockstep group) as grouped launches give the same U and dU/dθ
    as the default lanes, to 1e-10 of their largest elements."""
    c = grad_case
    os.environ["TT_GROUPED"] = "1"
    os.environ["TT_LANES"] = "1"
    U, dU, _ = c.prob.gradient(ctx, 0, c.P, c.theta, c.T, basis=c.basis)
    assert np.abs(U - c.U).max() <= 1e-10 * np.abs(c.U).max()
    assert np.abs(dU - c.dU).max() <= 1e-10 * np.abs(c.dU).max()


@pytest.fixture
def split_half_product():
    """Forces the split-half FP16 product for the FP32 applications of a fresh context (the
    autotuner reads TT_CELL_H when it first meets a shape)."""
    old = os.environ.get("TT_CELL_H")
    os.environ["TT_CELL_H"] = "1"
    ctx2 = tt.Context(0)
    yield ctx2
    del ctx2
    if old is None:
        del os.environ["TT_CELL_H"]
    else:
        os.environ["TT_CELL_H"] = old


def test_split_half_product_against_oracle(ctx, c2, c3, split_half_product):
    """The FP32 product on the FP16 MFMA (every operand as a high and a low half) at the
    two-, three- and four-M-tile shapes, N_alpha 27 (C2), 47, and 63 (a 7/2+ channel at
    Np = Nq = 8): rows and columns against the oracle's applications to 1e-5 like the FP32 MFMA
    product, and the mixed solve against FP64 GMRES to 1e-10."""
    ctx2 = split_half_product
    rp, pot, T = synthetic_config(N=8, two_J_3N_max=7, T_lab=(10.0, 50.0))
    prob = tt.Problem(rp, pot)
    big = max(range(prob.n_channels), key=lambda i: prob.channel(i, T).Nalpha)
    c4 = _setup(ctx, rp, pot, T, channel=big)
    assert c4.ch.Nalpha > 48
    rng = np.random.default_rng(9)
    for c in (c2, c3, c4):
        r = 18
        X = rng.normal(size=(r, c.ch.dim)) + 1j * rng.normal(size=(r, c.ch.dim))
        g = [k % c.ch.nE for k in range(r)]
        Yo = c.K.row_apply(X * np.stack([c.G[k] for k in g]))
        Y32 = c.ch.apply_rows(ctx2, c.P, X, G=c.G, g_index=g, fp32=True)
        assert np.abs(Y32 - Yo).max() <= 1e-5 * np.abs(Yo).max(), c.ch.Nalpha
        U64, _ = c.ch.solve(ctx2, c.P, c.G, method=tt.SOLVE_GMRES)
        Umx, st = c.ch.solve(ctx2, c.P, c.G, method=tt.SOLVE_MIXED)
        assert st.n_apply_fp32 > 0
        assert np.abs(Umx - U64).max() <= 1e-10 * np.abs(U64).max(), c.ch.Nalpha
        # the column form (transposed cells, products with P123 itself) of an odd number of columns
        Xc = X[:3]
        Yco = c.K.col_apply((Xc * np.stack([c.G[k] for k in g[:3]])).T).T
        Yc32 = c.ch.apply_rows(ctx2, c.P, Xc, G=c.G, g_index=g[:3], column=True, fp32=True)
        assert np.abs(Yc32 - Yco).max() <= 1e-5 * np.abs(Yco).max(), c.ch.Nalpha
    assert any("fp16-split" in ln for ln in tt.cell_tuning())
    assert any("column fp32" in ln and "fp16-split" in ln for ln in tt.cell_tuning())
