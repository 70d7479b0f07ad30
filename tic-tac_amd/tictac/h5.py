"""ctypes binding of include/tictac_h5.h: the P123 HDF5 cache (libtictac_h5.so).

The reference keeps each channel's P123 in an HDF5 file of its own
(``P123_sparse_JP_<2J>_<P>_Np_<Np>_Nq_<Nq>_J2max_<J2max>.h5``,
CPP/make_permutation_matrix.cpp:1320-1323) and reads it back instead of rebuilding it
(CPP/make_permutation_matrix.cpp:1297-1403).  These functions write and read those files in the
reference's layout; ``fill_p123`` is fill_P123_arrays with that cache.
"""
import ctypes as C
import os

import numpy as np

from . import _lib as L
from .problem import P123, _arr

H5_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtictac_h5.so")

_i32p = C.POINTER(C.c_int32)
_u64p = C.POINTER(C.c_uint64)
_pp_d = C.POINTER(L._dp)
_pp_i = C.POINTER(_i32p)

PROTOTYPES = {
    "tt_h5_last_error": (C.c_char_p, []),
    "tt_h5_free": (None, [C.c_void_p]),
    "tt_p123_h5_filename": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                      C.c_int]),
    "tt_p123_h5_subfile_name": (C.c_int, [C.c_char_p, C.c_size_t, C.c_char_p, C.c_int, C.c_int, C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_int]),
    "tt_store_P123_h5": (C.c_int, [C.c_char_p, L._dp, _i32p, _i32p, C.c_size_t, C.POINTER(L.FwpStates),
                                   C.POINTER(L.PwStates)]),
    "tt_read_P123_h5": (C.c_int, [C.c_char_p, C.POINTER(L.FwpStates), C.POINTER(L.PwStates), _pp_d, _pp_i, _pp_i,
                                  C.POINTER(C.c_size_t)]),
    "tt_merge_P123_thread_files_h5": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.POINTER(L.FwpStates),
                                                C.POINTER(L.PwStates), _pp_d, _pp_i, _pp_i, C.POINTER(C.c_size_t)]),
    "tt_coo_to_csc": (C.c_int, [C.c_int64, C.c_size_t, L._dp, _i32p, _i32p, L._dp, _i32p, _u64p]),
    "tt_csc_to_coo_row_major": (C.c_int, [C.c_int64, C.c_int64, L._dp, _i32p, _u64p, L._dp, _i32p, _i32p]),
    "tt_fill_P123_arrays_h5": (C.c_int, [C.c_void_p, C.POINTER(L.RunParams), C.POINTER(L.FwpStates),
                                         C.POINTER(L.PwStates), C.c_int, C.c_char_p, C.c_int, C.c_int, C.c_int,
                                         C.POINTER(L.CSC), C.POINTER(L.P123Stats)]),
}

_h5 = None


def h5lib():
    """The loaded cache library (raises if it has not been built)."""
    global _h5
    if _h5 is None:
        if not os.path.exists(H5_PATH):
            raise ImportError("libtictac_h5.so not built: run `make -C tic-tac_amd` (needs libhdf5)")
        L.lib()  # libtictac.so first: the cache library calls into it
        lib = C.CDLL(H5_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _h5 = lib
    return _h5


def _check(code):
    if code != L.TT_OK:
        raise L.TictacError(code, h5lib().tt_h5_last_error().decode())
    return code


def _i32(a):
    return np.ascontiguousarray(a, np.int32)


def p123_filename(folder, chan):
    """The channel's P123 file in folder (CPP/make_permutation_matrix.cpp:1320-1323)."""
    rp = chan.prob.rp
    args = (folder.encode(), chan.two_J, chan.parity, rp.Np_WP, rp.Nq_WP, chan.view.J_2N_max)
    n = h5lib().tt_p123_h5_filename(None, 0, *args)
    buf = C.create_string_buffer(n + 1)
    h5lib().tt_p123_h5_filename(buf, n + 1, *args)
    return buf.value.decode()


def p123_subfile_name(folder, chan, thread_idx, tfc):
    """One thread's spill file (CPP/make_permutation_matrix.cpp:143-156)."""
    rp = chan.prob.rp
    args = (folder.encode(), chan.two_J, chan.parity, rp.Np_WP, rp.Nq_WP, chan.view.J_2N_max, thread_idx, tfc)
    n = h5lib().tt_p123_h5_subfile_name(None, 0, *args)
    buf = C.create_string_buffer(n + 1)
    h5lib().tt_p123_h5_subfile_name(buf, n + 1, *args)
    return buf.value.decode()


def store_p123(path, chan, val, row, col):
    """Writes COO triplets (in the given order) with the channel's sizes, edges and partial waves."""
    val = np.ascontiguousarray(val, np.float64)
    row, col = _i32(row), _i32(col)
    assert len(val) == len(row) == len(col)
    _check(h5lib().tt_store_P123_h5(path.encode(), L.dptr(val), row.ctypes.data_as(_i32p),
                                    col.ctypes.data_as(_i32p), len(val), C.byref(chan.prob.fwp),
                                    C.byref(chan.view)))


def _take_triplets(v, r, c, n):
    out = (_arr(v, n, np.float64), _arr(r, n, np.int32), _arr(c, n, np.int32))
    for p in (v, r, c):
        h5lib().tt_h5_free(C.cast(p, C.c_void_p))
    return out


def read_p123(path, chan):
    """COO triplets (val, row, col) of a P123 file, checked against the channel."""
    v, r, c, n = L._dp(), _i32p(), _i32p(), C.c_size_t()
    _check(h5lib().tt_read_P123_h5(path.encode(), C.byref(chan.prob.fwp), C.byref(chan.view), C.byref(v),
                                   C.byref(r), C.byref(c), C.byref(n)))
    return _take_triplets(v, r, c, n.value)


def merge_thread_files(folder, chan, n_threads, max_tfc=0):
    """The P123_recovery merge of the channel's spill files: row-major sorted COO triplets."""
    v, r, c, n = L._dp(), _i32p(), _i32p(), C.c_size_t()
    _check(h5lib().tt_merge_P123_thread_files_h5(folder.encode(), n_threads, max_tfc, C.byref(chan.prob.fwp),
                                                 C.byref(chan.view), C.byref(v), C.byref(r), C.byref(c),
                                                 C.byref(n)))
    return _take_triplets(v, r, c, n.value)


def coo_to_csc(dim, val, row, col):
    """COO triplets in any order -> CSC (val, row, colptr), rows ascending in a column."""
    val = np.ascontiguousarray(val, np.float64)
    row, col = _i32(row), _i32(col)
    cv = np.zeros(len(val))
    cr = np.zeros(len(val), np.int32)
    cp = np.zeros(dim + 1, np.uint64)
    _check(h5lib().tt_coo_to_csc(dim, len(val), L.dptr(val), row.ctypes.data_as(_i32p), col.ctypes.data_as(_i32p),
                                 L.dptr(cv), cr.ctypes.data_as(_i32p), cp.ctypes.data_as(_u64p)))
    return cv, cr, cp


def csc_to_coo(dim, val, row, colptr):
    """CSC -> COO triplets sorted by (row, column)."""
    val = np.ascontiguousarray(val, np.float64)
    row = _i32(row)
    colptr = np.ascontiguousarray(colptr, np.uint64)
    v = np.zeros(len(val))
    r = np.zeros(len(val), np.int32)
    c = np.zeros(len(val), np.int32)
    _check(h5lib().tt_csc_to_coo_row_major(dim, len(val), L.dptr(val), row.ctypes.data_as(_i32p),
                                           colptr.ctypes.data_as(_u64p), L.dptr(v), r.ctypes.data_as(_i32p),
                                           c.ctypes.data_as(_i32p)))
    return v, r, c


def fill_p123(ctx, chan, folder, calculate_and_store=True, recovery=False, n_threads=1, production_run=True):
    """fill_P123_arrays with the reference's file cache: build on the GPU (or merge spill files) and
    store, or read the stored file.  Returns a device-resident P123 with its host arrays."""
    csc = L.CSC()
    st = L.P123Stats()
    _check(h5lib().tt_fill_P123_arrays_h5(ctx.handle, C.byref(chan.prob.rp), C.byref(chan.prob.fwp),
                                          C.byref(chan.view), int(production_run), folder.encode(),
                                          int(calculate_and_store), int(recovery), n_threads, C.byref(csc),
                                          C.byref(st)))
    return P123(csc, ctx, st)
