"""GPU parity at the benchmark's own workloads (synthetic NN potential, seed 7), through the C-ABI.

  * all 18 3N channels 2J = 1..17 of both parities at Np = Nq = 30 with the 8 lab energies
    linspace(1, 200, 8) MeV in one batched solve, mixed precision and FP64; the N_alpha = 63
    channel (the four-tile cell product, k_cellmfma<*, 4, 3, *>) against the oracle's GMRES;
  * the four-tile product at small size against the oracle's row and column applications and
    a converged solve;
  * Np = Nq = 40 (dims 43,200 and 100,800);
  * the gradient of channels 2J = 1, 3 of both parities at Np = Nq = 30, T_lab 5, 50, 150 MeV.

Tolerances: FP64 solves 1e-10 relative to the largest element of the compared block (the
oracle's scipy GMRES at rtol 1e-13); mixed precision 1e-8 per element against FP64 (the
workload's accuracy target); products 1e-13 (FP64) and 1e-5 (FP32 products) of max|Y|;
gradients 1e-5 of max|dU| against central differences of solves at tol 1e-13.
"""
import numpy as np
import pytest

import tictac as tt
from helpers import oracle_gmres_row, oracle_kernel, synthetic_config

pytestmark = pytest.mark.gpu

BATCH_T = [float(t) for t in np.linspace(1.0, 200.0, 8)]


def _rp(N, two_J_max):
    return tt.run_params(two_J_3N_max=two_J_max, J_2N_max=3, Np_WP=N, Nq_WP=N, chebyshev_s=tt.chebyshev_scale(N))


def _jp(ch):
    return "%d/2%s" % (ch.two_J, "+" if ch.parity > 0 else "-")


@pytest.fixture(scope="module")
def batch30(ctx):
    prob = tt.Problem(_rp(30, 17), tt.Potential("synthetic", tt.synthetic_params(7)))
    chans = [prob.channel(c, BATCH_T) for c in range(prob.n_channels)]
    assert len(chans) == 18
    print("building P123 of the 18 channels at Np = Nq = 30", flush=True)
    Ps = [ch.build_p123(ctx) for ch in chans]
    return prob, chans, Ps


def test_batch_18_channels_np30_mixed_and_fp64(ctx, batch30):
    """The benchmark's batch: every item converges to 1e-12; mixed precision agrees with FP64
    to 1e-8 per element; the batch agrees with one-channel solves of the same channels."""
    prob, chans, Ps = batch30
    print("batched FP64 and mixed solves, 18 channels x 8 energies", flush=True)
    out64 = prob.solve_batch(ctx, chans, Ps, method=tt.SOLVE_GMRES)
    outmx = prob.solve_batch(ctx, chans, Ps, method=tt.SOLVE_MIXED)
    for (U64, s64, st64), (Umx, smx, stmx), ch in zip(out64, outmx, chans):
        assert st64 == 0 and stmx == 0
        assert s64.residual <= 1e-12 and smx.residual <= 1e-12
        assert smx.n_apply_fp32 > 0
        rel = np.abs(Umx - U64) / np.abs(U64)
        assert rel.max() <= 1e-8, (_jp(ch), rel.max())
    for i in (0, 6):   # 1/2+ (N_alpha 27) and 7/2+ (N_alpha 63)
        ch, P = chans[i], Ps[i]
        U1, _ = ch.solve(ctx, P, None, method=tt.SOLVE_GMRES, device_resolvent=True)
        assert np.abs(out64[i][0] - U1).max() <= 1e-10 * np.abs(U1).max(), _jp(ch)


def test_nalpha63_channel_np30_against_oracle_gmres(ctx, batch30):
    """The 7/2+ channel (N_alpha 63, dim 56,700, 8.1e7 nonzeros of P123): one on-shell row system
    (first deuteron channel, lowest on-shell bin) solved by the oracle's scipy GMRES over the
    oracle's A = C^T (2 P123) VC, against the batched FP64 and mixed results."""
    prob, chans, Ps = batch30
    i = next(k for k, ch in enumerate(chans) if ch.Nalpha == 63)
    ch, P = chans[i], Ps[i]
    N = prob.rp.Np_WP
    G = ch.resolvent(ctx)
    print("oracle GMRES on", _jp(ch), "dim", ch.dim, flush=True)
    K = oracle_kernel(prob, ch, P)
    Uo = oracle_gmres_row(K, ch, G, 0, 0, N, N)
    for method in (tt.SOLVE_GMRES, tt.SOLVE_MIXED):
        U = prob.solve_batch(ctx, [ch], [P], method=method)[0][0]
        assert np.abs(U[0, :, 0] - Uo).max() <= 1e-10 * np.abs(Uo).max(), method


@pytest.fixture(scope="module")
def mt4(ctx):
    """A J^Pi = 7/2+ channel with N_alpha = 63 at Np = Nq = 12: the four-tile (MT = 4) cell product."""
    prob = tt.Problem(_rp(12, 7), tt.Potential("synthetic", tt.synthetic_params(7)))
    c = next(k for k in range(prob.n_channels) if prob.channel(k, [10.0]).Nalpha == 63)
    ch = prob.channel(c, [10.0, 50.0])
    P = ch.build_p123(ctx)
    return prob, ch, P, ch.resolvent(ctx), oracle_kernel(prob, ch, P)


def test_four_tile_product_against_oracle(ctx, mt4):
    """Row and column applications at N_alpha = 63 with 18 right-hand sides (48-wide N tiles):
    1e-13 of max|Y| in FP64, 1e-5 with FP32 products."""
    prob, ch, P, G, K = mt4
    rng = np.random.default_rng(11)
    r = 18
    X = rng.normal(size=(r, ch.dim)) + 1j * rng.normal(size=(r, ch.dim))
    g = [k % ch.nE for k in range(r)]
    Gsel = np.stack([G[k] for k in g])
    Yo = K.row_apply(X * Gsel)
    for fp32, tol in ((False, 1e-13), (True, 1e-5)):
        Y = ch.apply_rows(ctx, P, X, G=G, g_index=g, fp32=fp32)
        assert np.abs(Y - Yo).max() <= tol * np.abs(Yo).max(), fp32
    Yc = K.col_apply((X * Gsel).T).T
    for fp32, tol in ((False, 1e-13), (True, 1e-5)):
        Y = ch.apply_rows(ctx, P, X, G=G, g_index=g, fp32=fp32, column=True)
        assert np.abs(Y - Yc).max() <= tol * np.abs(Yc).max(), ("column", fp32)


def test_four_tile_solves_against_oracle(ctx, mt4):
    prob, ch, P, G, K = mt4
    N = prob.rp.Np_WP
    Uo = np.zeros((ch.nd, ch.nd, ch.nE), complex)
    for dr in range(ch.nd):
        for j in range(ch.nE):
            Uo[dr, :, j] = oracle_gmres_row(K, ch, G, dr, j, N, N)
    U64, _ = ch.solve(ctx, P, G, method=tt.SOLVE_GMRES)
    assert np.abs(U64 - Uo).max() <= 1e-10 * np.abs(Uo).max()
    Umx, st = ch.solve(ctx, P, G, method=tt.SOLVE_MIXED)
    assert st.n_apply_fp32 > 0
    assert (np.abs(Umx - Uo) / np.abs(Uo)).max() <= 1e-8


def test_np40_mixed_precision_and_oracle(ctx):
    """Np = Nq = 40: the 7/2+ channel (N_alpha 63, dim 100,800) in mixed precision against its
    FP64 solve (1e-8 per element), and the 1/2+ channel (dim 43,200) against the oracle's
    GMRES (1e-10)."""
    prob = tt.Problem(_rp(40, 7), tt.Potential("synthetic", tt.synthetic_params(7)))
    N = 40
    for c in range(prob.n_channels):
        ch = prob.channel(c, [10.0])
        if ch.parity != 1 or ch.two_J not in (1, 7):
            continue
        print("Np = 40", _jp(ch), "dim", ch.dim, flush=True)
        P = ch.build_p123(ctx)
        U64, s64 = ch.solve(ctx, P, None, method=tt.SOLVE_GMRES, device_resolvent=True)
        Umx, smx = ch.solve(ctx, P, None, method=tt.SOLVE_MIXED, device_resolvent=True)
        assert s64.residual <= 1e-12 and smx.residual <= 1e-12
        assert (np.abs(Umx - U64) / np.abs(U64)).max() <= 1e-8, _jp(ch)
        if ch.two_J == 1:
            G = ch.resolvent(ctx)
            K = oracle_kernel(prob, ch, P)
            for j in range(ch.nE):
                Uo = oracle_gmres_row(K, ch, G, 0, j, N, N)
                assert np.abs(U64[0, :, j] - Uo).max() <= 1e-10 * np.abs(Uo).max()
        del P


@pytest.fixture(scope="module")
def grad30(ctx):
    T = [5.0, 50.0, 150.0]
    prob = tt.Problem(_rp(30, 3), tt.Potential("synthetic", tt.synthetic_params(7)))
    basis = prob.basis_matrices(12)
    theta = tt.synthetic_params()[:12]
    cases = []
    for c in range(prob.n_channels):
        ch = prob.channel(c, T)
        print("gradient at Np = 30,", _jp(ch), flush=True)
        P = ch.build_p123(ctx)
        U, dU, st = prob.gradient(ctx, c, P, theta, T, basis=basis, tol=1e-13)
        cases.append((c, ch, P, U, dU, st))
    return prob, T, basis, theta, cases


def test_gradient_np30_against_central_differences(ctx, grad30):
    """The four channels 2J = 1, 3 of both parities: the adjoint dU/dtheta of a Yukawa strength,
    the 3S1 contact and a separable strength against central differences (h = 1e-5 |theta_i|,
    solves to 1e-13), and the forward U against the plain solve."""
    prob, T, basis, theta, cases = grad30
    for c, ch, P, U, dU, st in cases:
        assert st.residual_forward <= 1e-13 and st.residual_adjoint <= 1e-13
        U1, _ = ch.solve(ctx, P, None, method=tt.SOLVE_GMRES, device_resolvent=True, tol=1e-13)
        assert np.abs(U - U1).max() <= 1e-10 * np.abs(U1).max()
        for i in (1, 4, 10):
            h = 1e-5 * abs(theta[i])
            tp, tm = theta.copy(), theta.copy()
            tp[i] += h
            tm[i] -= h
            Up = prob.gradient(ctx, c, P, tp, T, basis=basis, grad=False, tol=1e-13)[0]
            Um = prob.gradient(ctx, c, P, tm, T, basis=basis, grad=False, tol=1e-13)[0]
            fd = (Up - Um) / (2 * h)
            scale = max(np.abs(fd).max(), np.abs(dU[i]).max())
            assert np.abs(dU[i] - fd).max() <= 1e-5 * scale, (_jp(ch), i, np.abs(dU[i] - fd).max() / scale)


def test_gradient_np30_gradcheck(ctx, grad30):
    """torch.autograd.gradcheck of the autograd.Function on the 1/2+ channel at Np = Nq = 30."""
    import torch
    from tictac.adjoint import ChannelGradient, u_matrix
    prob, T, basis, theta, cases = grad30
    c, ch, P = cases[0][:3]
    solver = ChannelGradient(prob, ctx, c, P, T, basis=basis, tol=1e-13)
    th0 = torch.tensor(theta, dtype=torch.float64)
    s = torch.zeros(12, dtype=torch.float64, requires_grad=True)
    fn = lambda s_: u_matrix(th0 * (1 + s_), solver).reshape(-1)
    # eps 1e-5: the finite-difference noise (solve tolerance / eps) and truncation (~eps^2) stay
    # below the tolerances.  The largest entries of dU/ds are ~50 (the 3S1 contact, which sets the
    # deuteron binding); atol 1e-6 is 2e-8 of that.  The truncation of the central differences of
    # that strongly nonlinear dependence reaches 1.6e-4 relative at entries of 2.5e-2 (4e-6), so
    # rtol is 1e-3: this test checks the autograd wiring; the values are pinned by
    # test_gradient_np30_against_central_differences (1e-5 of the largest entry, per channel) and
    # by the dense oracle gradient at Np = 10 (test_gradient_against_dense_reference).
    assert torch.autograd.gradcheck(fn, (s,), eps=1e-5, atol=1e-6, rtol=1e-3, fast_mode=False)
