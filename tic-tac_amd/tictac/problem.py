"""Host-side mirror of the reference driver (CPP/main.cpp:83-505) over the C-ABI.

``Problem`` builds the state spaces, potential matrices and SWP basis once
(CPP/main.cpp:144-153, 273-347); ``Channel`` holds one 3N channel J^Pi with
its on-shell indexing (CPP/main.cpp:177-198, 356-378) and calls the three
solve-path entry points: P123 build, resolvent, Faddeev solve.
"""
import ctypes as C

import numpy as np

from . import _lib as L
from ._lib import check, dptr, iptr

HBARC = 197.327


def chebyshev_scale(N, top_fm=8.0):
    """Scale s putting the top Chebyshev edge s*tan((2N-1)pi/4N) at top_fm fm^-1."""
    return top_fm * HBARC / np.tan((2 * N - 1) * np.pi / (4 * N))


class Context:
    """One HIP device and stream (tt_ctx)."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        check(L.lib().tt_ctx_create(device, C.byref(self._h)))

    @property
    def handle(self):
        return self._h

    def synchronize(self):
        check(L.lib().tt_ctx_synchronize(self._h))

    def close(self):
        if self._h:
            L.lib().tt_ctx_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_params(**kw):
    rp = L.RunParams()
    L.lib().tt_run_params_default(C.byref(rp))
    for k, v in kw.items():
        if not hasattr(rp, k):
            raise KeyError(k)
        setattr(rp, k, v)
    return rp


class Potential:
    def __init__(self, name, params=None):
        self._h = C.c_void_p()
        check(L.lib().tt_potential_create(name.encode(), C.byref(self._h)))
        self.name = name
        if params is not None:
            self.update(params)

    def update(self, params):
        a = np.ascontiguousarray(params, dtype=np.float64)
        check(L.lib().tt_potential_update_parameters(self._h, dptr(a), len(a)))
        self.params = a

    def V(self, p_in, p_out, coupled, S, J, T, Tz):
        out = np.zeros(6)
        check(L.lib().tt_potential_V(self._h, p_in, p_out, int(coupled), S, J, T, Tz, dptr(out)))
        return out

    def __del__(self):
        try:
            if self._h:
                L.lib().tt_potential_destroy(self._h)
        except Exception:
            pass


def _arr(ptr, n, dtype):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype).copy()


class Problem:
    """State spaces + 2N potential + SWP basis for one parameter set."""

    def __init__(self, rp, potential):
        lib = L.lib()
        self.rp = rp
        self.pot = potential
        self.pw = L.PwStates()
        check(lib.tt_construct_pw_states(C.byref(rp), C.byref(self.pw)))
        self.fwp = L.FwpStates()
        check(lib.tt_make_fwp_statespace(C.byref(rp), C.byref(self.fwp)))
        nu, nc = C.c_int(), C.c_int()
        lib.tt_num_2N_states(C.byref(rp), C.byref(nu), C.byref(nc))
        self.nu, self.nc = nu.value, nc.value
        Np = rp.Np_WP
        self.V_unco = np.full(max(self.nu * Np * Np, 1), np.nan)
        self.V_coup = np.full(max(self.nc * 4 * Np * Np, 1), np.nan)
        check(lib.tt_calculate_potential_matrices(C.byref(rp), C.byref(self.fwp), C.byref(self.pw), potential._h,
                                                  dptr(self.V_unco), dptr(self.V_coup)))
        self.e_unco = np.zeros(max(self.nu * (Np + 1), 1))
        self.e_coup = np.zeros(max(self.nc * 2 * (Np + 1), 1))
        self.C_unco = np.full(max(self.nu * Np * Np, 1), np.nan)
        self.C_coup = np.full(max(self.nc * 4 * Np * Np, 1), np.nan)
        self.swp = L.SwpStates()
        self.swp.e_SWP_unco = dptr(self.e_unco)
        self.swp.e_SWP_coup = dptr(self.e_coup)
        self.swp.C_SWP_unco = dptr(self.C_unco)
        self.swp.C_SWP_coup = dptr(self.C_coup)
        check(lib.tt_make_swp_states(C.byref(rp), C.byref(self.fwp), C.byref(self.pw), dptr(self.V_unco),
                                     dptr(self.V_coup), C.byref(self.swp)))
        self.E_bound = self.swp.E_bound
        self.p_edges = _arr(self.fwp.p_WP, Np + 1, np.float64)
        self.q_edges = _arr(self.fwp.q_WP, rp.Nq_WP + 1, np.float64)

    @property
    def n_channels(self):
        return self.pw.N_chn_3N

    def channel(self, c, T_lab, breakup=False):
        return Channel(self, c, T_lab, breakup)

    def solve_batch(self, ctx, channels, P123s, method=L.SOLVE_MIXED, tol=1e-12, max_iter=2000, restart=40,
                    graphs=None):
        """tt_solve_batch over channels (each with all its on-shell bins), G on the device.
        graphs: True / False force replayed step graphs on / off, None leaves the library's choice.
        Returns [(U[d_row, d_col, j], stats, status)] per channel; raises on the first failure
        after every channel was attempted."""
        o = L.SolveOpts()
        L.lib().tt_solve_opts_default(C.byref(o))
        o.method, o.tol, o.max_iter, o.restart = method, tol, max_iter, restart
        o.use_graphs = -1 if graphs is None else int(bool(graphs))
        n = len(channels)
        items = (L.BatchItem * max(n, 1))()
        Us, stats, Ubus = [], [], []
        for i, (ch, P) in enumerate(zip(channels, P123s)):
            U = np.zeros(ch.nd * ch.nd * ch.nE * 2)
            st = L.SolveStats()
            Us.append(U)
            stats.append(st)
            items[i].chn = C.pointer(ch.view)
            items[i].os = C.pointer(ch.os)
            items[i].P123 = C.pointer(P.csc)
            items[i].U = dptr(U)
            items[i].stats = C.pointer(st)
            Ubus.append(np.zeros(max(ch.nd * ch.n_bu, 1) * 2))
            if ch.n_bu:
                items[i].U_BU = dptr(Ubus[-1])
        check(L.lib().tt_solve_batch(ctx.handle, C.byref(self.rp), dptr(self.V_unco), dptr(self.V_coup),
                                     C.byref(self.swp), items, n, C.byref(o)))
        for ch, Ubu in zip(channels, Ubus):
            if ch.n_bu:
                ch.U_BU = Ubu.view(np.complex128)[:ch.nd * ch.n_bu].reshape(ch.nd, ch.n_bu).copy()
        return [(U.view(np.complex128).reshape(ch.nd, ch.nd, ch.nE), st, items[i].status)
                for i, (U, st, ch) in enumerate(zip(Us, stats, channels))]

    def parameter_walk(self, ctx, channels, P123s, param_sets, T_lab, method=L.SOLVE_MIXED, tol=1e-12,
                       max_iter=2000, restart=40):
        """tt_parameter_walk: the on-shell U of the listed channels (indices into the 3N channels,
        with their P123) for every parameter set (rows of param_sets), each set's channels
        solved as one lockstep batch.  Uses this problem's potential and gives it back its own
        parameters afterwards.  Returns (U, status, code): U[s][c] complex [nd, nd, nE] (None
        where the item failed), the status codes [n_sets, n_chn] and the call's status."""
        sets = np.ascontiguousarray(np.atleast_2d(param_sets), np.float64)
        n_sets, n_params = sets.shape
        chans = np.ascontiguousarray(channels, np.int32)
        n_chn = len(chans)
        T = np.ascontiguousarray(T_lab, np.float64)
        cap = 3 * 3 * 2 * (self.rp.Nq_WP + 1)
        U = np.zeros(2 * cap * n_sets * n_chn)
        n_out = np.zeros(n_sets * n_chn, np.int32)
        status = np.zeros(n_sets * n_chn, np.int32)
        Parr = (C.POINTER(L.CSC) * n_chn)(*[C.pointer(P.csc) for P in P123s])
        o = L.SolveOpts()
        L.lib().tt_solve_opts_default(C.byref(o))
        o.method, o.tol, o.max_iter, o.restart = method, tol, max_iter, restart
        try:
            code = L.lib().tt_parameter_walk(ctx.handle, C.byref(self.rp), C.byref(self.fwp), C.byref(self.pw),
                                             self.pot._h, dptr(sets), n_params, n_sets, iptr(chans), Parr, n_chn,
                                             dptr(T), len(T), dptr(U), cap, iptr(n_out), iptr(status), C.byref(o))
        finally:
            self.pot.update(self.pot.params)
        nds = [self.channel(int(c), T).nd for c in chans]
        Uc = U.view(np.complex128).reshape(n_sets, n_chn, cap)
        out = [[None] * n_chn for _ in range(n_sets)]
        for s_ in range(n_sets):
            for c in range(n_chn):
                k = s_ * n_chn + c
                if status[k] == L.TT_OK and n_out[k] > 0:
                    nd = nds[c]
                    out[s_][c] = Uc[s_, c, :n_out[k]].reshape(nd, nd, n_out[k] // (nd * nd)).copy()
        return out, status.reshape(n_sets, n_chn), code

    def basis_matrices(self, n_theta=12):
        """V_i = V_WP(theta = e_i) for the potential's linear strengths (tt_potential_basis_matrices):
        arrays (n_theta, nu*Np*Np) and (n_theta, nc*4*Np*Np)."""
        Np = self.rp.Np_WP
        su, sc = self.nu * Np * Np, self.nc * 4 * Np * Np
        Vb_u = np.zeros((n_theta, max(su, 1)))
        Vb_c = np.zeros((n_theta, max(sc, 1)))
        if su == 0 or (sc and Vb_c.shape[1] != sc):
            raise ValueError("unexpected 2N sizes")
        Vb_u = np.zeros(n_theta * su)
        Vb_c = np.zeros(max(n_theta * sc, 1))
        p = np.ascontiguousarray(self.pot.params, np.float64)
        check(L.lib().tt_potential_basis_matrices(C.byref(self.rp), C.byref(self.fwp), C.byref(self.pw), self.pot._h,
                                                  dptr(p), len(p), n_theta, dptr(Vb_u), dptr(Vb_c)))
        return Vb_u.reshape(n_theta, su), Vb_c[:n_theta * sc].reshape(n_theta, sc)

    def gradient(self, ctx, channel, P, theta, T_lab, basis=None, grad=True, tol=1e-12):
        """U(theta) of channel `channel` at the on-shell bins of T_lab and, with grad, dU/dtheta
        (tt_solve_gradient).  Returns (U[nd, nd, nE], dU[n_theta, nd, nd, nE] or None, stats)."""
        theta = np.ascontiguousarray(theta, np.float64)
        n_theta = len(theta)
        Vb_u, Vb_c = basis if basis is not None else self.basis_matrices(n_theta)
        Vb_u = np.ascontiguousarray(Vb_u, np.float64).ravel()
        Vb_c = np.ascontiguousarray(Vb_c, np.float64).ravel()
        if Vb_c.size == 0:
            Vb_c = np.zeros(1)
        T = np.ascontiguousarray(T_lab, np.float64)
        cap = 3 * 3 * 2 * (self.rp.Nq_WP + 1)
        U = np.zeros(2 * cap)
        dU = np.zeros(2 * cap * n_theta) if grad else None
        n_out = C.c_int(0)
        o = L.SolveOpts()
        L.lib().tt_solve_opts_default(C.byref(o))
        o.tol = tol
        st = L.GradStats()
        check(L.lib().tt_solve_gradient(ctx.handle, C.byref(self.rp), C.byref(self.fwp), C.byref(self.pw), int(channel),
                                        C.byref(P.csc), dptr(Vb_u), dptr(Vb_c), n_theta, dptr(theta), dptr(T), len(T),
                                        dptr(U), dptr(dU) if grad else None, cap, C.byref(n_out), C.byref(o),
                                        C.byref(st)))
        n = n_out.value
        ch = self.channel(channel, T_lab)
        shape = (ch.nd, ch.nd, ch.nE)
        assert n == ch.nd * ch.nd * ch.nE
        Uc = U[:2 * n].view(np.complex128).reshape(shape)
        dUc = None
        if grad:
            dUc = dU.view(np.complex128)[:n_theta * n].reshape((n_theta,) + shape)
        return Uc, dUc, st

    def pw_arrays(self):
        n = self.pw.Nalpha
        names = ["L_2N", "S_2N", "J_2N", "T_2N", "L_1N", "two_J_1N", "two_J_3N", "two_T_3N", "P_3N"]
        d = {k: _arr(getattr(self.pw, k), n, np.int64) for k in names}
        d["chn_3N_idx"] = _arr(self.pw.chn_3N_idx, self.pw.N_chn_3N + 1, np.int64)
        return d

    def __del__(self):
        try:
            L.lib().tt_pw_states_free(C.byref(self.pw))
            L.lib().tt_fwp_states_free(C.byref(self.fwp))
        except Exception:
            pass


class P123:
    """Device-resident P123 in the reference's CSC layout (tt_csc)."""

    def __init__(self, csc, ctx, stats=None):
        self.csc = csc
        self.ctx = ctx
        self.stats = stats

    @property
    def dim(self):
        return self.csc.dim

    @property
    def nnz(self):
        return self.csc.nnz

    @classmethod
    def from_host(cls, ctx, dim, val, row, colptr):
        csc = L.CSC()
        val = np.ascontiguousarray(val, np.float64)
        row = np.ascontiguousarray(row, np.int32)
        colptr = np.ascontiguousarray(colptr, np.uint64)
        check(L.lib().tt_csc_upload(ctx.handle, dim, len(val), dptr(val), row.ctypes.data_as(C.POINTER(C.c_int32)),
                                    colptr.ctypes.data_as(C.POINTER(C.c_uint64)), C.byref(csc)))
        return cls(csc, ctx)

    def to_host(self):
        check(L.lib().tt_csc_download(self.ctx.handle, C.byref(self.csc)))
        n, d = self.csc.nnz, self.csc.dim
        return (_arr(self.csc.val, n, np.float64), _arr(self.csc.row, n, np.int32),
                _arr(self.csc.colptr, d + 1, np.uint64))

    def __del__(self):
        try:
            L.lib().tt_csc_free(C.byref(self.csc))
        except Exception:
            pass


class Channel:
    """One 3N channel J^Pi with its on-shell indexing (elastic, and with breakup=True the
    breakup cells of tt_find_breakup_bins)."""

    def __init__(self, prob, c, T_lab, breakup=False):
        lib = L.lib()
        self.prob = prob
        self.index = c
        self.view = L.PwStates()
        check(lib.tt_pw_channel(C.byref(prob.pw), c, C.byref(self.view)))
        self.Nalpha = self.view.Nalpha
        self.dim = self.Nalpha * prob.rp.Np_WP * prob.rp.Nq_WP
        self.os = L.ChannelOS()
        T = np.ascontiguousarray(T_lab, np.float64)
        check(lib.tt_find_on_shell(C.byref(prob.swp), C.byref(self.view), dptr(T), len(T), C.byref(self.os)))
        nE = self.os.num_T_lab
        self.nE = nE
        self.q_idx = _arr(self.os.q_com_idx, nE, np.int64)
        self.T_lab = _arr(self.os.T_lab, nE, np.float64)
        self.E_com = _arr(self.os.E_com, nE, np.float64)
        self.nd = self.os.num_deuteron_states
        self.deut = _arr(self.os.deuteron_idx, self.nd, np.int64)
        self.two_J = int(self.view.two_J_3N[0])
        self.parity = int(self.view.P_3N[0])
        self.n_bu = 0
        if breakup:
            check(lib.tt_find_breakup_bins(C.byref(prob.rp), C.byref(prob.swp), C.byref(self.view),
                                           C.byref(self.os)))
            self.n_bu = self.os.num_BU_chns
            self.bu_ptr = _arr(self.os.q_com_BU_idx, nE + 1, np.int64)
            self.bu_cells = _arr(self.os.alphapq_idx, 3 * self.n_bu, np.int64).reshape(self.n_bu, 3)

    def pw_arrays(self):
        n = self.Nalpha
        names = ["L_2N", "S_2N", "J_2N", "T_2N", "L_1N", "two_J_1N", "two_J_3N", "two_T_3N", "P_3N"]
        return {k: _arr(getattr(self.view, k), n, np.int64) for k in names}

    def energies(self):
        """Total energies E_com + E_bound of the on-shell bins (CPP/main.cpp:391)."""
        return self.E_com + self.prob.E_bound

    def build_p123(self, ctx, production_run=True):
        csc = L.CSC()
        st = L.P123Stats()
        check(L.lib().tt_fill_P123_arrays(ctx.handle, C.byref(self.prob.rp), C.byref(self.prob.fwp),
                                          C.byref(self.view), int(production_run), C.byref(csc), C.byref(st)))
        return P123(csc, ctx, st)

    def resolvent(self, ctx, E=None):
        E = np.ascontiguousarray(self.energies() if E is None else E, np.float64)
        G = np.zeros(len(E) * self.dim * 2)
        check(L.lib().tt_calculate_resolvent(ctx.handle, C.byref(self.prob.rp), C.byref(self.prob.swp),
                                             C.byref(self.view), dptr(E), len(E), dptr(G)))
        return G.view(np.complex128).reshape(len(E), self.dim)

    def solve(self, ctx, P, G=None, method=L.SOLVE_PADE, tol=1e-12, max_iter=2000, restart=40, NM_max=14,
              neumann=False, p123_format=L.P123_CELLS, graphs=None, device_resolvent=False):
        """On-shell U[d_row, d_col, j] (CPP/solve_faddeev.cpp:424 layout) and solve statistics.
        graphs: True / False force replayed step graphs on / off, None leaves the library's choice.

        With ``neumann=True`` (Pade) also returns the Neumann terms a[d_row, d_col, j, n].
        With ``device_resolvent=True`` G is computed on the device inside the call (G ignored).
        A channel made with breakup=True also fills ``self.U_BU[d_row, b]``, the breakup
        amplitudes of its breakup cells."""
        if device_resolvent:
            G = None
        else:
            if G is None:
                G = self.resolvent(ctx)
            G = np.ascontiguousarray(G, np.complex128)
            assert G.shape == (self.nE, self.dim)
        o = L.SolveOpts()
        L.lib().tt_solve_opts_default(C.byref(o))
        o.method, o.tol, o.max_iter, o.restart, o.NM_max = method, tol, max_iter, restart, NM_max
        o.p123_format = p123_format
        o.use_graphs = -1 if graphs is None else int(bool(graphs))
        o.resolvent_on_device = int(bool(device_resolvent))
        nterms = 2 * NM_max + 1
        a = np.zeros(self.nd * self.nd * self.nE * nterms * 2)
        if neumann:
            o.neumann_terms = dptr(a)
        st = L.SolveStats()
        U = np.zeros(self.nd * self.nd * self.nE * 2)
        Ubu = np.zeros(max(self.nd * self.n_bu, 1) * 2)
        if self.n_bu:
            o.U_BU = dptr(Ubu)
        Gf = None if G is None else G.view(np.float64).ravel()
        check(L.lib().tt_solve_faddeev_equations(ctx.handle, C.byref(self.prob.rp), dptr(U),
                                                 dptr(Gf) if Gf is not None else None,
                                                 C.byref(P.csc), dptr(self.prob.V_unco), dptr(self.prob.V_coup),
                                                 C.byref(self.prob.swp), C.byref(self.os), C.byref(self.view),
                                                 C.byref(o), C.byref(st)))
        U = U.view(np.complex128).reshape(self.nd, self.nd, self.nE)
        if self.n_bu:
            self.U_BU = Ubu.view(np.complex128)[:self.nd * self.n_bu].reshape(self.nd, self.n_bu).copy()
        if neumann:
            return U, st, a.view(np.complex128).reshape(self.nd, self.nd, self.nE, nterms)
        return U, st

    def apply_rows(self, ctx, P, X, G=None, g_index=None, fp32=False, csc=False, column=False, noalloc=False):
        """Y = (X .* G_sel) A for rows X (r, dim) complex; with column=True Y_k = A (G_k .* X_k).
        noalloc=True calls tt_apply_rows_noalloc (FP64 row form, no device allocation)."""
        X = np.ascontiguousarray(X, np.complex128)
        r = X.shape[0]
        Y = np.zeros_like(X)
        nE = 0 if G is None else G.shape[0]
        Gf = np.zeros(2) if G is None else np.ascontiguousarray(G, np.complex128).view(np.float64).ravel()
        gi = None
        if g_index is not None:
            gi = np.ascontiguousarray(g_index, np.int32)
        fn = L.lib().tt_apply_rows_noalloc if noalloc else L.lib().tt_apply_rows
        check(fn(ctx.handle, C.byref(self.prob.rp), C.byref(P.csc), dptr(self.prob.V_unco), dptr(self.prob.V_coup),
                 C.byref(self.prob.swp), C.byref(self.view), dptr(Gf), nE, iptr(gi) if gi is not None else None, r,
                 dptr(X.view(np.float64)), dptr(Y.view(np.float64)),
                 int(fp32) | (2 if csc else 0) | (4 if column else 0)))
        return Y

    def __del__(self):
        try:
            L.lib().tt_channel_os_free(C.byref(self.os))
        except Exception:
            pass
