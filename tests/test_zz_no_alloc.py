"""tt_apply_rows_noalloc: the FP64 row application with every device buffer beyond the P123
object carved from the library's static scratch array (no allocation).  Its output equals that
of tt_apply_rows on two calls in a row, the second with a smaller input, over the cell blocks
and over the CSC; sizes above the scratch maximum and unsupported flags are errors."""
import numpy as np
import pytest

import tictac as tt
from helpers import synthetic_config

pytestmark = pytest.mark.gpu


def _case(ctx, N, two_J_max, T):
    rp, pot, _ = synthetic_config(N=N, two_J_3N_max=two_J_max, T_lab=T)
    prob = tt.Problem(rp, pot)
    c = max(range(prob.n_channels), key=lambda i: prob.channel(i, T).Nalpha)
    ch = prob.channel(c, T)
    P = ch.build_p123(ctx)
    return prob, ch, P, ch.resolvent(ctx)


def test_noalloc_apply_matches_apply_rows_on_two_calls(ctx):
    big = _case(ctx, 12, 3, (10.0, 50.0))     # N_alpha 47, dim 6768
    small = _case(ctx, 8, 1, (10.0,))         # N_alpha 27, dim 1728
    rng = np.random.default_rng(21)
    for (prob, ch, P, G), r in ((big, 7), (small, 3)):
        print("noalloc apply: N_alpha", ch.Nalpha, "dim", ch.dim, "r", r, flush=True)
        X = rng.normal(size=(r, ch.dim)) + 1j * rng.normal(size=(r, ch.dim))
        g = [k % ch.nE if k % 3 else -1 for k in range(r)]
        for csc in (False, True):
            Y0 = ch.apply_rows(ctx, P, X, G=G, g_index=g, csc=csc)
            Y1 = ch.apply_rows(ctx, P, X, G=G, g_index=g, csc=csc, noalloc=True)
            assert np.abs(Y1 - Y0).max() <= 1e-13 * np.abs(Y0).max(), (ch.Nalpha, r, csc)


def test_noalloc_apply_rejects_what_it_cannot_do(ctx):
    prob, ch, P, G = _case(ctx, 8, 1, (10.0,))
    X = np.ones((2, ch.dim), complex)
    with pytest.raises(tt.TictacError) as e:
        ch.apply_rows(ctx, P, X, G=G, g_index=[0, 0], fp32=True, noalloc=True)
    assert e.value.code == tt._lib.TT_ERR_UNSUPPORTED
    r = 1400   # 3 dim r complex vectors alone exceed the 128 MiB scratch at dim 1728
    Xb = np.ones((r, ch.dim), complex)
    with pytest.raises(tt.TictacError) as e:
        ch.apply_rows(ctx, P, Xb, G=G, g_index=[0] * r, noalloc=True)
    assert e.value.code == tt._lib.TT_ERR_ARG and "scratch" in str(e.value)
