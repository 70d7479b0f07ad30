"""Oracle: kernel matrix A = C^T (2 P123) V C and the on-shell U-matrix solves
(test infrastructure only).

Restates:
  CPP/solve_faddeev.cpp:1588-1750   C^T block table (row-major, coupled quarter blocks)
  CPP/solve_faddeev.cpp:1754-1927   V.C block table (column-major, coupled quarter blocks)
  CPP/solve_faddeev.cpp:4-48        PVC column (P = P123 + P132 = 2 P123, line 39)
  CPP/solve_faddeev.cpp:50-269      C^T.PVC columns and the on-shell rows of A
  CPP/solve_faddeev.cpp:271-296     Pade approximant [N/M] at z via determinants
  CPP/solve_faddeev.cpp:305-438     dense solve (1 - A G_j) U = A
  CPP/solve_faddeev.cpp:440-1444    Neumann rows a_n = e_r^T A (G A)^n e_c and Pade selection
  CPP/General_functions/matrix_routines.cpp:82-101   determinant with pivot signs
  CPP/solve_faddeev.cpp:867-891,1177-1206,1300-1365,1388-1409
                                    breakup terms a_n at the breakup cells' columns and
                                    their Pade selection (same rules as elastic)

The block tables are written as embeddings of each 2N-channel matrix into the
(alpha, p) space: the reference's pointer tables select, for a pair
(alpha_r, alpha_c), the quarter block [h(alpha_r), h(alpha_c)] of the 2Np x 2Np
coupled matrix, with h = 0 for L = J-1 (or 2T3 = 1) and h = 1 for L = J+1 (or 2T3 = 3).

Pade selection at NM = 0 reads element NM-1 of the approximant array in the
reference (lines 1264, 1273: out of bounds).  The array is value-initialised
to zero, so the oracle takes PA_{-1} = 0; this choice is "parity unpinned"
(no reference test or run can pin undefined behaviour).

Breakup amplitudes are U[row, cell] at the breakup cells' flat indices
alpha*Nq*Np + q*Np + p (formed as the reference forms them, also for a neighbour one
bin outside the grid; a flat index outside the state space, where the reference reads
outside the row, gives NaN).  The reference fills them in its Pade path only; the oracle
also gives the converged value, U = (1 - A G_j)^-1 A read at those columns.

Pins (tests/test_oracle_pins.py): matrix-free row apply against the explicit
dense A, and elements of A by explicit sums over the per-alpha-pair blocks and P;
tests/test_oracle_recoupling.py: the embedded C_op is orthogonal, diagonalises the embedded
H0 + V with E_bound at the deuteron state (alpha_d, p = 0), carries the deuteron mostly in its
S-wave rows, and VC_op = V_op C_op (a swapped half or a transposed off-diagonal quarter fails); dense solve residual; Pade of a geometric series a_n = r^n equals
1/(1-r) exactly for N >= 1 and agrees with scipy.interpolate.pade; breakup values
of the dense solve against A (1 - G A)^-1 by explicit inversion, and their Pade
resummation against the dense values.
"""
import numpy as np
import scipy.sparse as sp

from . import states as st


def _half(pw, a, via_L, via_T):
    if via_L:
        return 0 if pw.L[a] < pw.J[a] else 1
    return 0 if pw.two_T3[a] == 1 else 1


def embed_2N(pw, swp, M_unco, M_coup, tensor_force, isospin_breaking_1S0):
    """Dense operator on (alpha, p): B[(ar, pr), (ac, pc)] = M_2N[h(ar) Np + pr, h(ac) Np + pc]
    for pairs that couple through the 2N interaction (CPP/solve_faddeev.cpp:1685-1749)."""
    Na, Np = pw.Nalpha, swp.Np
    B = np.zeros((Na, Np, Na, Np))
    for ar in range(Na):
        Lr, Sr, Jr, Tr, lr, jr = (int(pw.L[ar]), int(pw.S[ar]), int(pw.J[ar]), int(pw.T[ar]),
                                  int(pw.l[ar]), int(pw.two_j[ar]))
        for ac in range(Na):
            Lc, Sc, Jc, Tc, lc, jc = (int(pw.L[ac]), int(pw.S[ac]), int(pw.J[ac]), int(pw.T[ac]),
                                      int(pw.l[ac]), int(pw.two_j[ac]))
            checkL = (tensor_force and abs(Lr - Lc) <= 2) or Lr == Lc
            if not (Tr == Tc and Jr == Jc and Sr == Sc and checkL and lr == lc and jr == jc):
                continue
            via_L, via_T = st.coupling_kind(Lr, Sr, Jr, Tr, Lc, tensor_force, isospin_breaking_1S0)
            if via_L or via_T:
                idx = st.unique_2N_idx(Lr, Sr, Jr, Tr, True, tensor_force, isospin_breaking_1S0)
                hr, hc = _half(pw, ar, via_L, via_T), _half(pw, ac, via_L, via_T)
                B[ar, :, ac, :] = M_coup[idx][hr * Np:(hr + 1) * Np, hc * Np:(hc + 1) * Np]
            else:
                idx = st.unique_2N_idx(Lr, Sr, Jr, Tr, False, tensor_force, isospin_breaking_1S0)
                B[ar, :, ac, :] = M_unco[idx]
    return B


def c_and_vc(pw, swp, Vu, Vc, tensor_force, isospin_breaking_1S0):
    """C_op and (V C)_op on (alpha, p); C^T_op = C_op^T."""
    VCu = np.einsum("kij,kjl->kil", np.nan_to_num(Vu), np.nan_to_num(swp.C_unco))
    VCc = np.einsum("kij,kjl->kil", np.nan_to_num(Vc), np.nan_to_num(swp.C_coup)) if Vc.shape[0] else Vc
    C = embed_2N(pw, swp, np.nan_to_num(swp.C_unco), np.nan_to_num(swp.C_coup), tensor_force, isospin_breaking_1S0)
    VC = embed_2N(pw, swp, VCu, VCc, tensor_force, isospin_breaking_1S0)
    return C, VC


def block_full(B, Nq):
    """Sparse matrix on (alpha, q, p) of the q-diagonal block operator B[(a,p),(a',p')]."""
    Na, Np = B.shape[0], B.shape[1]
    ar, pr, ac, pc = np.nonzero(B)
    v = B[ar, pr, ac, pc]
    q = np.arange(Nq)
    rows = (ar[:, None] * Nq * Np + q[None, :] * Np + pr[:, None]).ravel()
    cols = (ac[:, None] * Nq * Np + q[None, :] * Np + pc[:, None]).ravel()
    vals = np.repeat(v, Nq)
    d = Na * Nq * Np
    return sp.csr_matrix((vals, (rows, cols)), shape=(d, d))


def p_matrix(val, row, colptr, dim):
    """Sparse P123 from the CSC arrays of CPP/main.cpp:226-244."""
    return sp.csc_matrix((val, row.astype(np.int64), colptr.astype(np.int64)), shape=(dim, dim))


class KernelA:
    """A = C^T (2 P123) VC on the full (alpha, q, p) space."""

    def __init__(self, pw, swp, Vu, Vc, P, tensor_force, isospin_breaking_1S0):
        C, VC = c_and_vc(pw, swp, Vu, Vc, tensor_force, isospin_breaking_1S0)
        self.Cb, self.VCb = C, VC
        Nq = swp.Nq
        self.CT = block_full(C, Nq).T.tocsr()
        self.VC = block_full(VC, Nq)
        self.P2 = (2.0 * P).tocsr()
        self.dim = P.shape[0]

    def dense(self):
        return (self.CT @ self.P2 @ self.VC).toarray()

    def row_apply(self, X):
        """X (r, dim) -> X A, the reference's rows x A (CPP/solve_faddeev.cpp:1072-1088)."""
        Y = (self.VC.T @ (self.P2.T @ (self.CT.T @ X.T))).T
        return Y

    def col_apply(self, X):
        return self.CT @ (self.P2 @ (self.VC @ X))


def onshell_indices(deut, q_idx, Nq, Np):
    """Row/column index of (alpha_d, q_os, p=0) (CPP/solve_faddeev.cpp:1134-1175)."""
    return {(d, j): int(a) * Nq * Np + int(q) * Np for d, a in enumerate(deut) for j, q in enumerate(q_idx)}


def dense_solve(A, G, deut, q_idx, Nq, Np):
    """U[d_row, d_col, j] from (1 - A G_j) X = A (CPP/solve_faddeev.cpp:305-438)."""
    nd, nE = len(deut), len(q_idx)
    U = np.zeros((nd, nd, nE), dtype=complex)
    dim = A.shape[0]
    for j in range(nE):
        L = np.eye(dim) - A * G[j][None, :]
        cols = [int(deut[dc]) * Nq * Np + int(q_idx[j]) * Np for dc in range(nd)]
        X = np.linalg.solve(L, A[:, cols].astype(complex))
        for dr in range(nd):
            r = int(deut[dr]) * Nq * Np + int(q_idx[j]) * Np
            for dc in range(nd):
                U[dr, dc, j] = X[r, dc]
    return U


def neumann_terms(Kop, G, deut, q_idx, Nq, Np, nterms):
    """a[d_row, d_col, j, n] = e_r^T A (G A)^n e_c via rows <- (rows . G_j) A
    (CPP/solve_faddeev.cpp:793-1175)."""
    nd, nE = len(deut), len(q_idx)
    dim = Kop.dim
    rows = np.zeros((nd * nE, dim), dtype=complex)
    # n = 0: on-shell rows of A, row index idx_d_row*num_q_com + idx_q_com
    E = np.zeros((nd * nE, dim))
    for dr in range(nd):
        for j in range(nE):
            E[dr * nE + j, int(deut[dr]) * Nq * Np + int(q_idx[j]) * Np] = 1.0
    rows = Kop.row_apply(E).astype(complex)
    a = np.zeros((nd, nd, nE, nterms), dtype=complex)
    for n in range(nterms):
        if n > 0:
            Gr = np.stack([G[j] for dr in range(nd) for j in range(nE)])
            rows = Kop.row_apply(rows * Gr)
        for dr in range(nd):
            for dc in range(nd):
                for j in range(nE):
                    a[dr, dc, j, n] = rows[dr * nE + j, int(deut[dc]) * Nq * Np + int(q_idx[j]) * Np]
    return a


def determinant(M):
    """LU with partial pivoting; det = prod diag * pivot signs (matrix_routines.cpp:82-101)."""
    return np.linalg.det(M)


def pade_approximant(a, N, M, z):
    """CPP/solve_faddeev.cpp:271-296; a must hold N+M+1 coefficients."""
    P = np.zeros((M + 1, M + 1), dtype=complex)
    Q = np.zeros((M + 1, M + 1), dtype=complex)
    for r in range(M):
        for c in range(M + 1):
            P[r, c] = a[N - M + 1 + r + c]
            Q[r, c] = a[N - M + 1 + r + c]
    for c in range(M + 1):
        Q[M, c] = z ** (M - c)
        s = 0
        for j in range(M - c, N + 1):
            s += a[j - (M - c)] * z ** j
        P[M, c] = s
    return determinant(P) / determinant(Q)


def pade_select(a_elem, NM_max=14):
    """Pade approximants PA[NM] = [NM/NM](1) and the reference's stopping and
    best-approximant rules (CPP/solve_faddeev.cpp:1236-1298), with PA_{-1} = 0.
    Returns (U, PA list, index of the chosen approximant)."""
    PA = []
    for NM in range(NM_max + 1):
        PA.append(pade_approximant(a_elem, NM, NM, 1.0))
        prev = lambda k: PA[k] if k >= 0 else 0.0
        best, mindiff = 0, 1.0
        for k in range(NM):
            d = abs(PA[k] - prev(k - 1))
            if d < mindiff and d > 1e-15:
                best, mindiff = k, d
        d = abs(PA[NM] - prev(NM - 1))
        if d < mindiff and d > 1e-15:
            best, mindiff = NM, d
        c0 = NM == NM_max
        c1 = NM - best > 4
        c2 = mindiff < 1e-6 * abs(PA[best])
        c3 = mindiff < 1e-7
        if c0 or c1 or c2 or c3:
            return PA[best], PA, best
    raise AssertionError("unreachable")


def pade_solve(a):
    """U[d_row, d_col, j] from Neumann terms a[..., n] by the reference's Pade rules."""
    nd, _, nE, _ = a.shape
    U = np.zeros((nd, nd, nE), dtype=complex)
    for dr in range(nd):
        for dc in range(nd):
            for j in range(nE):
                U[dr, dc, j] = pade_select(a[dr, dc, j])[0]
    return U


def breakup_flat(cell, Nq, Np):
    """Flat state index of a breakup cell (CPP/solve_faddeev.cpp:879)."""
    a, q, p = (int(x) for x in cell)
    return a * Nq * Np + q * Np + p


def dense_solve_breakup(A, G, deut, q_idx, Nq, Np, bu_ptr, bu_cells):
    """U_BU[d_row, b] = U[(alpha_d, q_j, 0), cell_b] with (1 - A G_j) X = A[:, cell_b columns]."""
    nd = len(deut)
    n_bu = len(bu_cells)
    U = np.zeros((nd, n_bu), dtype=complex)
    dim = A.shape[0]
    U[:] = np.nan
    for j in range(len(q_idx)):
        bs = [b for b in range(int(bu_ptr[j]), int(bu_ptr[j + 1])) if 0 <= breakup_flat(bu_cells[b], Nq, Np) < dim]
        if not len(bs):
            continue
        cols = [breakup_flat(bu_cells[b], Nq, Np) for b in bs]
        X = np.linalg.solve(np.eye(dim) - A * G[j][None, :], A[:, cols].astype(complex))
        for dr in range(nd):
            r = int(deut[dr]) * Nq * Np + int(q_idx[j]) * Np
            for k, b in enumerate(bs):
                U[dr, b] = X[r, k]
    return U


def neumann_terms_breakup(Kop, G, deut, q_idx, Nq, Np, bu_ptr, bu_cells, nterms):
    """a_BU[d_row, b, n] = e_r^T A (G A)^n at cell_b's column, r = (deuteron d_row, energy of b)
    (CPP/solve_faddeev.cpp:867-891,1177-1206)."""
    nd, nE = len(deut), len(q_idx)
    dim = Kop.dim
    E = np.zeros((nd * nE, dim))
    for dr in range(nd):
        for j in range(nE):
            E[dr * nE + j, int(deut[dr]) * Nq * Np + int(q_idx[j]) * Np] = 1.0
    rows = Kop.row_apply(E).astype(complex)
    n_bu = len(bu_cells)
    a = np.full((nd, n_bu, nterms), np.nan, dtype=complex)
    Gr = np.stack([G[j] for dr in range(nd) for j in range(nE)])
    for n in range(nterms):
        if n > 0:
            rows = Kop.row_apply(rows * Gr)
        for dr in range(nd):
            for j in range(nE):
                for b in range(int(bu_ptr[j]), int(bu_ptr[j + 1])):
                    f = breakup_flat(bu_cells[b], Nq, Np)
                    if 0 <= f < dim:
                        a[dr, b, n] = rows[dr * nE + j, f]
    return a
