"""ctypes binding of include/tictac.h (the C-ABI of libtictac.so).

The library is built in-tree (``make -C tic-tac_amd``); importing this module
without it raises, there is no fallback implementation.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libtictac.so")

TT_OK, TT_ERR_ARG, TT_ERR_HIP, TT_ERR_ALLOC, TT_ERR_PHYSICS, TT_ERR_NOT_CONVERGED, TT_ERR_UNSUPPORTED = range(7)
SOLVE_PADE, SOLVE_DENSE, SOLVE_GMRES, SOLVE_MIXED = range(4)
P123_CELLS, P123_CSC = range(2)

_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


class RunParams(C.Structure):
    _fields_ = [("two_J_3N_max", C.c_int), ("J_2N_max", C.c_int), ("Np_WP", C.c_int), ("Nq_WP", C.c_int),
                ("Np_per_WP", C.c_int), ("Nphi", C.c_int), ("Nx", C.c_int), ("chebyshev_s", C.c_double),
                ("chebyshev_t", C.c_double), ("tensor_force", C.c_int), ("isospin_breaking_1S0", C.c_int),
                ("midpoint_approx", C.c_int), ("swave_only", C.c_int)]


class PwStates(C.Structure):
    _fields_ = [("Nalpha", C.c_int), ("J_2N_max", C.c_int), ("L_2N", _ip), ("S_2N", _ip), ("J_2N", _ip),
                ("T_2N", _ip), ("L_1N", _ip), ("two_J_1N", _ip), ("two_J_3N", _ip), ("two_T_3N", _ip),
                ("P_3N", _ip), ("chn_3N_idx", _ip), ("N_chn_3N", C.c_int)]


class FwpStates(C.Structure):
    _fields_ = [("Np_WP", C.c_int), ("Nq_WP", C.c_int), ("Np_per_WP", C.c_int), ("Nq_per_WP", C.c_int),
                ("p_WP", _dp), ("q_WP", _dp), ("p", _dp), ("wp", _dp), ("q", _dp), ("wq", _dp),
                ("norm_p", _dp), ("norm_q", _dp)]


class SwpStates(C.Structure):
    _fields_ = [("Np_WP", C.c_int), ("Nq_WP", C.c_int), ("num_2N_unco_states", C.c_int),
                ("num_2N_coup_states", C.c_int), ("E_bound", C.c_double), ("e_SWP_unco", _dp),
                ("e_SWP_coup", _dp), ("C_SWP_unco", _dp), ("C_SWP_coup", _dp), ("q_WP", _dp)]


class ChannelOS(C.Structure):
    _fields_ = [("num_T_lab", C.c_int), ("q_com_idx", _ip), ("T_lab", _dp), ("q_com", _dp), ("E_com", _dp),
                ("num_deuteron_states", C.c_int), ("deuteron_idx", _ip), ("num_BU_chns", C.c_int),
                ("q_com_BU_idx", _ip), ("alphapq_idx", _ip)]


class CSC(C.Structure):
    _fields_ = [("dim", C.c_int64), ("nnz", C.c_int64), ("val", _dp), ("row", C.POINTER(C.c_int32)),
                ("colptr", C.POINTER(C.c_uint64)), ("dev", C.c_void_p)]


class P123Stats(C.Structure):
    _fields_ = [("mask_density", C.c_double), ("n_cells", C.c_int64), ("n_tuples", C.c_int),
                ("ms_total", C.c_double)]


class SolveOpts(C.Structure):
    _fields_ = [("method", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int), ("restart", C.c_int),
                ("NM_max", C.c_int), ("neumann_terms", _dp), ("p123_format", C.c_int), ("use_graphs", C.c_int),
                ("resolvent_on_device", C.c_int), ("U_BU", _dp)]


class SolveStats(C.Structure):
    _fields_ = [("n_apply", C.c_int), ("n_apply_fp32", C.c_int), ("iterations", C.c_int),
                ("residual", C.c_double), ("ms_total", C.c_double), ("ms_apply", C.c_double),
                ("n_spmm", C.c_int), ("ms_spmm", C.c_double), ("spmm_bytes", C.c_double),
                ("n_spmm_fp32", C.c_int), ("ms_spmm_fp32", C.c_double), ("spmm_bytes_fp32", C.c_double)]


class GradStats(C.Structure):
    _fields_ = [("ms_forward", C.c_double), ("ms_total", C.c_double), ("n_apply_forward", C.c_int),
                ("n_apply_adjoint", C.c_int), ("residual_forward", C.c_double), ("residual_adjoint", C.c_double)]


class BatchItem(C.Structure):
    _fields_ = [("chn", C.POINTER(PwStates)), ("os", C.POINTER(ChannelOS)), ("P123", C.POINTER(CSC)),
                ("U", _dp), ("stats", C.POINTER(SolveStats)), ("status", C.c_int), ("U_BU", _dp)]


# exported symbols and their prototypes (restype, argtypes)
PROTOTYPES = {
    "tt_last_error": (C.c_char_p, []),
    "tt_version": (C.c_char_p, []),
    "tt_ctx_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "tt_ctx_destroy": (None, [C.c_void_p]),
    "tt_ctx_stream": (C.c_void_p, [C.c_void_p]),
    "tt_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "tt_run_params_default": (None, [C.POINTER(RunParams)]),
    "tt_construct_pw_states": (C.c_int, [C.POINTER(RunParams), C.POINTER(PwStates)]),
    "tt_pw_states_free": (None, [C.POINTER(PwStates)]),
    "tt_pw_channel": (C.c_int, [C.POINTER(PwStates), C.c_int, C.POINTER(PwStates)]),
    "tt_make_fwp_statespace": (C.c_int, [C.POINTER(RunParams), C.POINTER(FwpStates)]),
    "tt_fwp_states_free": (None, [C.POINTER(FwpStates)]),
    "tt_potential_create": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "tt_potential_destroy": (None, [C.c_void_p]),
    "tt_potential_update_parameters": (C.c_int, [C.c_void_p, _dp, C.c_int]),
    "tt_potential_V": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int, C.c_int,
                                 C.c_int, _dp]),
    "tt_num_2N_states": (None, [C.POINTER(RunParams), _ip, _ip]),
    "tt_calculate_potential_matrices": (C.c_int, [C.POINTER(RunParams), C.POINTER(FwpStates),
                                                  C.POINTER(PwStates), C.c_void_p, _dp, _dp]),
    "tt_make_swp_states": (C.c_int, [C.POINTER(RunParams), C.POINTER(FwpStates), C.POINTER(PwStates), _dp, _dp,
                                     C.POINTER(SwpStates)]),
    "tt_find_on_shell": (C.c_int, [C.POINTER(SwpStates), C.POINTER(PwStates), _dp, C.c_int,
                                   C.POINTER(ChannelOS)]),
    "tt_find_breakup_bins": (C.c_int, [C.POINTER(RunParams), C.POINTER(SwpStates), C.POINTER(PwStates),
                                       C.POINTER(ChannelOS)]),
    "tt_channel_p_edges": (C.c_int, [C.POINTER(RunParams), C.POINTER(SwpStates), C.POINTER(PwStates), _dp]),
    "tt_channel_os_free": (None, [C.POINTER(ChannelOS)]),
    "tt_csc_upload": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, _dp, C.POINTER(C.c_int32),
                                C.POINTER(C.c_uint64), C.POINTER(CSC)]),
    "tt_csc_download": (C.c_int, [C.c_void_p, C.POINTER(CSC)]),
    "tt_csc_free": (None, [C.POINTER(CSC)]),
    "tt_fill_P123_arrays": (C.c_int, [C.c_void_p, C.POINTER(RunParams), C.POINTER(FwpStates),
                                      C.POINTER(PwStates), C.c_int, C.POINTER(CSC), C.POINTER(P123Stats)]),
    "tt_calculate_resolvent": (C.c_int, [C.c_void_p, C.POINTER(RunParams), C.POINTER(SwpStates),
                                         C.POINTER(PwStates), _dp, C.c_int, _dp]),
    "tt_solve_opts_default": (None, [C.POINTER(SolveOpts)]),
    "tt_solve_faddeev_equations": (C.c_int, [C.c_void_p, C.POINTER(RunParams), _dp, _dp, C.POINTER(CSC), _dp, _dp,
                                             C.POINTER(SwpStates), C.POINTER(ChannelOS), C.POINTER(PwStates),
                                             C.POINTER(SolveOpts), C.POINTER(SolveStats)]),
    "tt_solve_batch": (C.c_int, [C.c_void_p, C.POINTER(RunParams), _dp, _dp, C.POINTER(SwpStates),
                                 C.POINTER(BatchItem), C.c_int, C.POINTER(SolveOpts)]),
    "tt_potential_basis_matrices": (C.c_int, [C.POINTER(RunParams), C.POINTER(FwpStates), C.POINTER(PwStates),
                                              C.c_void_p, _dp, C.c_int, C.c_int, _dp, _dp]),
    "tt_solve_gradient": (C.c_int, [C.c_void_p, C.POINTER(RunParams), C.POINTER(FwpStates), C.POINTER(PwStates),
                                    C.c_int, C.POINTER(CSC), _dp, _dp, C.c_int, _dp, _dp, C.c_int, _dp, _dp,
                                    C.c_int, _ip, C.POINTER(SolveOpts), C.POINTER(GradStats)]),
    "tt_apply_rows": (C.c_int, [C.c_void_p, C.POINTER(RunParams), C.POINTER(CSC), _dp, _dp, C.POINTER(SwpStates),
                                C.POINTER(PwStates), _dp, C.c_int, _ip, C.c_int, _dp, _dp, C.c_int]),
    "tt_apply_rows_noalloc": (C.c_int, [C.c_void_p, C.POINTER(RunParams), C.POINTER(CSC), _dp, _dp,
                                        C.POINTER(SwpStates), C.POINTER(PwStates), _dp, C.c_int, _ip, C.c_int, _dp,
                                        _dp, C.c_int]),
    "tt_cell_tuning": (C.c_int, [C.c_char_p, C.c_int]),
    "tt_parameter_walk": (C.c_int, [C.c_void_p, C.POINTER(RunParams), C.POINTER(FwpStates), C.POINTER(PwStates),
                                    C.c_void_p, _dp, C.c_int, C.c_int, _ip, C.POINTER(C.POINTER(CSC)), C.c_int, _dp,
                                    C.c_int, _dp, C.c_int, _ip, _ip, C.POINTER(SolveOpts)]),
}


def cell_tuning():
    """The cell-product autotuner's choices so far (tt_cell_tuning), one line per shape."""
    n = lib().tt_cell_tuning(None, 0)
    buf = C.create_string_buffer(n)
    lib().tt_cell_tuning(buf, n)
    return [ln for ln in buf.value.decode().splitlines() if ln]

_lib = None


class TictacError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("tictac error %d: %s" % (code, msg))
        self.code = code


def lib():
    """The loaded C library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libtictac.so not built: run `make -C tic-tac_amd` (or __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code):
    if code != TT_OK:
        raise TictacError(code, lib().tt_last_error().decode())
    return code


def dptr(a):
    return a.ctypes.data_as(_dp)


def iptr(a):
    return a.ctypes.data_as(_ip)
