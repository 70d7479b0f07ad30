"""ctypes front end of the oracle's C++/OpenMP restatement (oracle/cpu_solve.cpp, built by
oracle/Makefile into oracle/libttcpu.so) -- TEST INFRASTRUCTURE ONLY: the CPU baseline of
bench.py and checks in tests/.

  gmres_U        converged on-shell U[d_row, d_col, j] by GMRES on the row systems
                 (matrix-free, CSC gather), all host threads
  neumann_terms  the reference's Neumann terms a[d_row, d_col, j, n] with every A column
                 regenerated per term (CPP/solve_faddeev.cpp:440-1444), with a wall-time cap

Pins: tests/test_oracle_cpu.py compares both with the numpy oracle (oracle.solve) on the same
inputs: dense solve and Neumann terms to rounding.
"""
import ctypes as C
import os

import numpy as np

from . import solve as so

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libttcpu.so")
        if not os.path.exists(path):
            raise ImportError("oracle/libttcpu.so is not built (make -C oracle)")
        L = C.CDLL(path)
        dp, ip, i64p, u64p, i32p = (C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_int32))
        L.ttcpu_threads.restype = C.c_int
        L.ttcpu_gmres_rows.restype = C.c_int
        L.ttcpu_gmres_rows.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, dp, i32p, u64p, dp, dp, dp, C.c_int,
                                       C.c_int, i64p, ip, C.c_double, C.c_int, C.c_int, dp, ip, dp]
        L.ttcpu_neumann_terms.restype = C.c_int
        L.ttcpu_neumann_terms.argtypes = [C.c_int64, C.c_int, C.c_int, C.c_int, dp, i32p, u64p, dp, dp, dp, C.c_int,
                                          C.c_int, i64p, ip, C.c_int, i64p, C.c_int, C.c_double, dp, dp]
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Problem:
    """The operator inputs of one channel: P123 CSC arrays and the embedded C_op / VC_op."""

    def __init__(self, pw, swp, Vu, Vc, val, row, colptr, tensor_force, isospin_breaking_1S0):
        Cb, VCb = so.c_and_vc(pw, swp, Vu, Vc, tensor_force, isospin_breaking_1S0)
        self.Na, self.Np, self.Nq = pw.Nalpha, swp.Np, swp.Nq
        self.dim = self.Na * self.Nq * self.Np
        self.Cb = np.ascontiguousarray(Cb, np.float64)
        self.VCb = np.ascontiguousarray(VCb, np.float64)
        self.val = np.ascontiguousarray(val, np.float64)
        self.row = np.ascontiguousarray(row, np.int32)
        self.colptr = np.ascontiguousarray(colptr, np.uint64)

    def _rows(self, deut, q_idx):
        nd, nE = len(deut), len(q_idx)
        unit = np.array([int(deut[d]) * self.Nq * self.Np + int(q_idx[j]) * self.Np
                         for d in range(nd) for j in range(nE)], np.int64)
        g = np.array([j for d in range(nd) for j in range(nE)], np.int32)
        return unit, g

    def gmres_U(self, G, deut, q_idx, tol=1e-12, restart=40, max_iter=2000):
        nd, nE = len(deut), len(q_idx)
        unit, g = self._rows(deut, q_idx)
        r = len(unit)
        Gf = np.ascontiguousarray(G, np.complex128).view(np.float64)
        Y = np.zeros(2 * r * self.dim)
        its = np.zeros(r, np.int32)
        res = np.zeros(r)
        rc = lib().ttcpu_gmres_rows(self.dim, self.Na, self.Nq, self.Np, _p(self.val, C.c_double),
                                    _p(self.row, C.c_int32), _p(self.colptr, C.c_uint64), _p(self.Cb, C.c_double),
                                    _p(self.VCb, C.c_double), _p(Gf, C.c_double), G.shape[0], r,
                                    _p(unit, C.c_int64), _p(g, C.c_int), tol, restart, max_iter,
                                    _p(Y, C.c_double), _p(its, C.c_int), _p(res, C.c_double))
        Y = Y.view(np.complex128).reshape(r, self.dim)
        U = np.zeros((nd, nd, nE), complex)
        for dr in range(nd):
            for j in range(nE):
                for dc in range(nd):
                    U[dr, dc, j] = Y[dr * nE + j, int(deut[dc]) * self.Nq * self.Np + int(q_idx[j]) * self.Np]
        return U, rc, its, res

    def neumann_terms(self, G, deut, q_idx, n_terms, cap_s=1e30):
        """a[d_row, d_col, j, n] for n < the terms done; returns (a, terms done, seconds per term)."""
        nd, nE = len(deut), len(q_idx)
        unit, g = self._rows(deut, q_idx)
        r = len(unit)
        cols = np.array([int(deut[dc]) * self.Nq * self.Np + int(q_idx[j]) * self.Np
                         for dc in range(nd) for j in range(nE)], np.int64)
        Gf = np.ascontiguousarray(G, np.complex128).view(np.float64)
        a = np.zeros(2 * r * len(cols) * n_terms)
        spt = C.c_double(0.0)
        done = lib().ttcpu_neumann_terms(self.dim, self.Na, self.Nq, self.Np, _p(self.val, C.c_double),
                                         _p(self.row, C.c_int32), _p(self.colptr, C.c_uint64),
                                         _p(self.Cb, C.c_double), _p(self.VCb, C.c_double), _p(Gf, C.c_double),
                                         G.shape[0], r, _p(unit, C.c_int64), _p(g, C.c_int), len(cols),
                                         _p(cols, C.c_int64), n_terms, cap_s, _p(a, C.c_double), C.byref(spt))
        a = a.view(np.complex128).reshape(r, len(cols), n_terms)
        out = np.zeros((nd, nd, nE, n_terms), complex)
        for dr in range(nd):
            for j in range(nE):
                for dc in range(nd):
                    out[dr, dc, j] = a[dr * nE + j, dc * nE + j]
        return out[..., :done], done, spt.value
