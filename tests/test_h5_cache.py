"""The P123 HDF5 cache (libtictac_h5.so): the reference's per-channel P123 files.

The file layout is checked with h5dump, the HDF5 distribution's own reader, against what the
reference's writer produces (CPP/disk_io_routines.cpp:1078-1164,1286-1302,1336-1352,1386-1432,
1474-1545,1625-1721): dataset names, element types, shapes, table titles and field names.  The
sorts are checked against scipy.sparse, the recovery merge against the sorted union of the spill
files, and on the GPU a solve on a P123 read back from its file against the solve on the P123
the GPU built.
"""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest
import scipy.sparse as sp

from helpers import mt_config, synthetic_config

import tictac as tt
from tictac import h5

H5DUMP = shutil.which("h5dump") or "/opt/conda/bin/h5dump"


@pytest.fixture(scope="module")
def mt_channel():
    rp, pot, T = mt_config(N=16)
    prob = tt.Problem(rp, pot)
    return prob, prob.channel(0, T)


def _random_coo(dim, n, seed, distinct=True):
    rng = np.random.default_rng(seed)
    flat = rng.choice(dim * dim, size=n, replace=False) if distinct else rng.integers(0, dim * dim, n)
    return rng.standard_normal(n), (flat // dim).astype(np.int32), (flat % dim).astype(np.int32)


def _h5dump(*args):
    out = subprocess.run([H5DUMP, *args], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stderr
    return out.stdout


def _dataset_header(text, name):
    """The DATATYPE and DATASPACE lines of one dataset in an h5dump -H listing."""
    m = re.search(r'DATASET "%s" \{\s*DATATYPE\s+(.*?)\s*DATASPACE\s+(SIMPLE \{[^}]*\})' % re.escape(name), text,
                  re.S)
    assert m, name
    return " ".join(m.group(1).split()), m.group(2)


def _data_numbers(text):
    body = text[text.index("DATA {"):]
    body = body.split("ATTRIBUTE")[0]  # a table's attributes follow its data
    return [float(x) for x in re.findall(r"[-+]?\d+(?:\.\d*)?(?:[eE][-+]?\d+)?", body)]


def test_file_name_follows_the_reference(mt_channel):
    # CPP/make_permutation_matrix.cpp:1320-1323 and 143-156
    prob, ch = mt_channel
    assert h5.p123_filename("out", ch) == "out/P123_sparse_JP_1_1_Np_16_Nq_16_J2max_1.h5"
    assert h5.p123_subfile_name("out", ch, 3, 7) == "out/P123_subsparse_JP_1_1_Np_16_Nq_16_J2max_1_TFC_3_7.h5"


def test_stored_file_has_the_reference_layout(mt_channel, tmp_path):
    prob, ch = mt_channel
    v, r, c = _random_coo(ch.dim, 300, 1)
    path = str(tmp_path / "p.h5")
    h5.store_p123(path, ch, v, r, c)
    hdr = _h5dump("-H", "-A", path)
    # scalars: write_integer_to_h5 / write_ULL_integer_to_h5 (H5T_NATIVE_INT, H5T_NATIVE_ULLONG, one element)
    for name in ("Nalpha", "Np_WP", "Nq_WP"):
        assert _dataset_header(hdr, name) == ("H5T_STD_I32LE", "SIMPLE { ( 1 ) / ( 1 ) }")
    assert _dataset_header(hdr, "P123_sparse_dim") == ("H5T_STD_U64LE", "SIMPLE { ( 1 ) / ( 1 ) }")
    # the COO arrays: little-endian int and double, nnz elements, fixed size
    for name, t in (("row indices", "H5T_STD_I32LE"), ("column indices", "H5T_STD_I32LE"),
                    ("values", "H5T_IEEE_F64LE")):
        assert _dataset_header(hdr, name) == (t, "SIMPLE { ( 300 ) / ( 300 ) }")
    # the tables: H5TBmake_table with the reference's titles and field names
    t, s = _dataset_header(hdr, "p boundaries")
    assert t == 'H5T_COMPOUND { H5T_STD_I32LE "index"; H5T_IEEE_F64LE "mesh point"; }'
    assert s == "SIMPLE { ( 17 ) / ( H5S_UNLIMITED ) }"
    fields = ["index", "L_12", "S_12", "J_12", "T_12", "l_3", "2*j_3", "2*T_total", "2*J_total", "PAR_total"]
    t, s = _dataset_header(hdr, "pw channels")
    assert t == "H5T_COMPOUND { " + " ".join('H5T_STD_I32LE "%s";' % f for f in fields) + " }"
    assert s == "SIMPLE { ( 2 ) / ( H5S_UNLIMITED ) }"
    for dset, title in (("p boundaries", "Info p boundaries"), ("q boundaries", "Info q boundaries"),
                        ("pw channels", "Partial wave info")):
        att = _h5dump("-a", "/%s/TITLE" % dset, path)
        assert '"%s"' % title in att
        assert '"TABLE"' in _h5dump("-a", "/%s/CLASS" % dset, path)
    for i, f in enumerate(fields):
        assert '"%s"' % f in _h5dump("-a", "/pw channels/FIELD_%d_NAME" % i, path)
    # values as h5dump reads them
    assert _data_numbers(_h5dump("-y", "-d", "/Nalpha", path)) == [2]
    assert _data_numbers(_h5dump("-y", "-d", "/P123_sparse_dim", path)) == [300]
    got = _data_numbers(_h5dump("-y", "-w", "0", "-m", "%.17g", "-d", "/values", path))
    np.testing.assert_array_equal(got, v)
    got = _data_numbers(_h5dump("-y", "-w", "0", "-d", "/row indices", path))
    np.testing.assert_array_equal(got, r)
    got = _data_numbers(_h5dump("-y", "-w", "0", "-m", "%.17g", "-d", "/p boundaries", path))
    np.testing.assert_array_equal(np.array(got).reshape(-1, 2), np.c_[np.arange(17), prob.p_edges])
    got = np.array(_data_numbers(_h5dump("-y", "-w", "0", "-d", "/pw channels", path))).reshape(-1, 10)
    pw = ch.pw_arrays()
    expect = np.c_[np.arange(ch.Nalpha), pw["L_2N"], pw["S_2N"], pw["J_2N"], pw["T_2N"], pw["L_1N"],
                   pw["two_J_1N"], pw["two_T_3N"], pw["two_J_3N"], pw["P_3N"]]
    np.testing.assert_array_equal(got, expect)


def test_read_back_and_mismatch_errors(mt_channel, tmp_path):
    prob, ch = mt_channel
    v, r, c = _random_coo(ch.dim, 500, 2)
    path = str(tmp_path / "p.h5")
    h5.store_p123(path, ch, v, r, c)
    vv, rr, cc = h5.read_p123(path, ch)
    assert np.array_equal(vv, v) and np.array_equal(rr, r) and np.array_equal(cc, c)
    # an empty P123 is a valid file
    h5.store_p123(str(tmp_path / "e.h5"), ch, [], [], [])
    assert all(len(x) == 0 for x in h5.read_p123(str(tmp_path / "e.h5"), ch))
    # another Np: state-space dimensions mismatch (CPP/disk_io_routines.cpp:1214-1216)
    rp2, pot2, T2 = mt_config(N=12)
    ch2 = tt.Problem(rp2, pot2).channel(0, T2)
    with pytest.raises(tt.TictacError, match=re.escape("(Nalpha, Nq_WP, Np_WP) mismatch")):
        h5.read_p123(path, ch2)
    # one p edge one ulp off: p-momentum boundaries mismatch (1219-1229)
    pe = np.ctypeslib.as_array(prob.fwp.p_WP, shape=(17,))
    keep = pe[5]
    pe[5] = np.nextafter(keep, np.inf)
    try:
        with pytest.raises(tt.TictacError, match="p-momentum boundaries mismatch"):
            h5.read_p123(path, ch)
    finally:
        pe[5] = keep
    qe = np.ctypeslib.as_array(prob.fwp.q_WP, shape=(17,))
    keep = qe[16]
    qe[16] = keep * 1.5
    try:
        with pytest.raises(tt.TictacError, match="q-momentum boundaries mismatch"):
            h5.read_p123(path, ch)
    finally:
        qe[16] = keep
    # one partial wave changed: PW state-space mismatch (1262-1276)
    L1 = np.ctypeslib.as_array(ch.view.L_1N, shape=(ch.Nalpha,))
    keep = L1[1]
    L1[1] = keep + 2
    try:
        with pytest.raises(tt.TictacError, match="PW state-space mismatch"):
            h5.read_p123(path, ch)
    finally:
        L1[1] = keep
    h5.read_p123(path, ch)  # restored
    # not an HDF5 file, and a missing one
    bad = tmp_path / "bad.h5"
    bad.write_text("not hdf5")
    with pytest.raises(tt.TictacError, match="cannot open"):
        h5.read_p123(str(bad), ch)
    with pytest.raises(tt.TictacError, match="cannot open"):
        h5.read_p123(str(tmp_path / "missing.h5"), ch)


def test_coo_csc_conversions_match_scipy():
    dim = 700
    v, r, c = _random_coo(dim, 20000, 3)
    cv, cr, cp = h5.coo_to_csc(dim, v, r, c)
    ref = sp.coo_matrix((v, (r, c)), shape=(dim, dim)).tocsc()
    ref.sort_indices()
    np.testing.assert_array_equal(cp, ref.indptr)
    np.testing.assert_array_equal(cr, ref.indices)
    np.testing.assert_array_equal(cv, ref.data)
    # back to COO sorted by (row, column): the order of the reference's files
    v2, r2, c2 = h5.csc_to_coo(dim, cv, cr, cp)
    order = np.lexsort((c, r))
    np.testing.assert_array_equal(r2, r[order])
    np.testing.assert_array_equal(c2, c[order])
    np.testing.assert_array_equal(v2, v[order])
    # indices out of range are errors
    with pytest.raises(tt.TictacError, match="row-index larger"):
        h5.coo_to_csc(dim, [1.0], [dim], [0])
    with pytest.raises(tt.TictacError, match="column-index larger"):
        h5.coo_to_csc(dim, [1.0], [0], [-1])


def test_recovery_merges_the_thread_files(mt_channel, tmp_path, capfd):
    """The P123_recovery merge (CPP/make_permutation_matrix.cpp:946-1141): spill files of several
    threads in random element order give one row-major sorted COO; a missing file below max_TFC is
    skipped with the reference's warning."""
    prob, ch = mt_channel
    v, r, c = _random_coo(ch.dim, 4000, 4)
    folder = str(tmp_path)
    parts = np.array_split(np.random.default_rng(5).permutation(len(v)), 6)
    layout = [(0, 0), (0, 1), (1, 0), (2, 0), (2, 1), (2, 2)]
    for (t, tfc), idx in zip(layout, parts):
        h5.store_p123(h5.p123_subfile_name(folder, ch, t, tfc), ch, v[idx], r[idx], c[idx])

    def sorted_union(keep):
        idx = np.concatenate([parts[i] for i in keep])
        o = np.lexsort((c[idx], r[idx]))
        return v[idx][o], r[idx][o], c[idx][o]

    for got, want in zip(h5.merge_thread_files(folder, ch, 3), sorted_union(range(6))):
        np.testing.assert_array_equal(got, want)
    # thread 1 has no TFC 1: with max_TFC = 3 it is looked for and skipped with a warning
    os.remove(h5.p123_subfile_name(folder, ch, 0, 0))
    capfd.readouterr()
    got = h5.merge_thread_files(folder, ch, 3, max_tfc=3)
    err = capfd.readouterr().err
    assert "WARNING file not found" in err and "_TFC_0_0.h5" in err and "_TFC_1_1.h5" in err
    for g, w in zip(got, sorted_union([1, 2, 3, 4, 5])):
        np.testing.assert_array_equal(g, w)
    # without max_TFC a thread's scan stops at its first missing file: thread 0 contributes nothing
    for g, w in zip(h5.merge_thread_files(folder, ch, 3), sorted_union([2, 3, 4, 5])):
        np.testing.assert_array_equal(g, w)
    # a spill file of another channel size is refused
    rp2, pot2, T2 = mt_config(N=12)
    ch2 = tt.Problem(rp2, pot2).channel(0, T2)
    h5.store_p123(h5.p123_subfile_name(folder, ch, 3, 0), ch2, v[:5], r[:5] % ch2.dim, c[:5] % ch2.dim)
    with pytest.raises(tt.TictacError, match="mismatch"):
        h5.merge_thread_files(folder, ch, 4)


# ------------------------------------------------------------------ GPU: the cache around the build
@pytest.mark.gpu
def test_stored_p123_reads_back_and_solves_identically(ctx, tmp_path):
    """C1 and the 1/2+ channel at Np = 30: the P123 built on the GPU is stored and read back
    element for element, and the solves on both agree."""
    for cfg, method, tol in ((mt_config(N=16), tt.SOLVE_GMRES, 1e-13), (synthetic_config(N=30), tt.SOLVE_MIXED, 1e-12)):
        rp, pot, T = cfg
        prob = tt.Problem(rp, pot)
        ch = prob.channel(0, T)
        folder = str(tmp_path)
        print("h5 cache: Np", rp.Np_WP, "dim", ch.dim, flush=True)
        P_built = h5.fill_p123(ctx, ch, folder, calculate_and_store=True)
        path = h5.p123_filename(folder, ch)
        assert os.path.exists(path)
        P_read = h5.fill_p123(ctx, ch, folder, calculate_and_store=False)
        for a, b in zip(P_built.to_host(), P_read.to_host()):
            np.testing.assert_array_equal(a, b)
        # the file holds the CSC's elements in row-major order
        v, r, c = h5.read_p123(path, ch)
        assert np.all(np.diff(r.astype(np.int64) * ch.dim + c) > 0)
        cv, cr, cp = h5.coo_to_csc(ch.dim, v, r, c)
        val, row, ptr = P_built.to_host()
        np.testing.assert_array_equal(cv, val)
        np.testing.assert_array_equal(cr, row)
        np.testing.assert_array_equal(cp, ptr)
        G = ch.resolvent(ctx)
        U1, _ = ch.solve(ctx, P_built, G, method=method)
        U2, _ = ch.solve(ctx, P_read, G, method=method)
        assert np.abs(U1 - U2).max() <= tol * np.abs(U1).max()


@pytest.mark.gpu
def test_recovery_rebuilds_the_channel_file(ctx, tmp_path):
    rp, pot, T = mt_config(N=16)
    prob = tt.Problem(rp, pot)
    ch = prob.channel(0, T)
    P = ch.build_p123(ctx)
    val, row, ptr = P.to_host()
    v, r, c = h5.csc_to_coo(ch.dim, val, row, ptr)
    folder = str(tmp_path)
    parts = np.array_split(np.random.default_rng(6).permutation(len(v)), 4)
    for t, idx in enumerate(parts):
        h5.store_p123(h5.p123_subfile_name(folder, ch, t, 0), ch, v[idx], r[idx], c[idx])
    P2 = h5.fill_p123(ctx, ch, folder, calculate_and_store=True, recovery=True, n_threads=4)
    for a, b in zip(P.to_host(), P2.to_host()):
        np.testing.assert_array_equal(a, b)
    for a, b in zip(h5.read_p123(h5.p123_filename(folder, ch), ch), (v, r, c)):
        np.testing.assert_array_equal(a, b)


def test_cache_library_exports_every_declared_symbol():
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                            "tictac_h5.h")).read()
    declared = set(re.findall(r"\b(tt_[A-Za-z0-9_]+)\s*\(", hdr))
    assert len(declared) == 10
    lib = h5.h5lib()
    assert not [n for n in sorted(declared) if not hasattr(lib, n)]
    assert set(h5.PROTOTYPES) == declared
