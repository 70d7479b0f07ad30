/* The P123 HDF5 cache (include/tictac_h5.h): the reference's per-channel P123 files, their
 * thread spill files and the recovery merge, over libhdf5 and its table API.
 *
 * Plain C99 so that the library needs nothing beyond libc, libhdf5(_hl) and libtictac.so:
 * it sits beside libtictac.so and only calls its C-ABI (the GPU build, tt_csc_download,
 * tt_csc_upload).  The file layout is the reference's (CPP/disk_io_routines.cpp:1078-1803):
 * dataset names, element types, the tables' titles, field names, record layouts and chunk
 * size are kept as they are there, so that a file written by one program reads in the other. */
#define _POSIX_C_SOURCE 200809L
#include <stdarg.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include <hdf5.h>
#include <hdf5_hl.h>

#include "tictac.h"
#include "tictac_h5.h"

static __thread char g_err[1024];

static tt_status fail(tt_status code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

const char *tt_h5_last_error(void) { return g_err; }

void tt_h5_free(void *p) { free(p); }

/* libhdf5 prints its error stack on every failed call unless told not to; each entry point
 * reports failures through its status instead. */
static void quiet_hdf5(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

/* ------------------------------------------------------------------ record layouts
 * The tables' records (CPP/type_defs.h:126-160).  The order of the members is part of the
 * files' compound types: the partial-wave record keeps 2*T_total after PAR_total in memory
 * while its field list names it before 2*J_total, as in the reference. */
typedef struct {
    int index;
    double mesh;
} bounds_record;

typedef struct {
    int alpha, L, S, J, T, l, twoj, twoJtotal, PARtotal, twoTtotal;
} pw_record;

static const char *const k_bounds_fields[2] = {"index", "mesh point"};
static const char *const k_pw_fields[10] = {"index", "L_12", "S_12", "J_12", "T_12",
                                            "l_3", "2*j_3", "2*T_total", "2*J_total", "PAR_total"};
static const size_t k_pw_offsets[10] = {
    offsetof(pw_record, alpha),     offsetof(pw_record, L),         offsetof(pw_record, S),
    offsetof(pw_record, J),         offsetof(pw_record, T),         offsetof(pw_record, l),
    offsetof(pw_record, twoj),      offsetof(pw_record, twoTtotal), offsetof(pw_record, twoJtotal),
    offsetof(pw_record, PARtotal)};
static const size_t k_bounds_offsets[2] = {offsetof(bounds_record, index), offsetof(bounds_record, mesh)};
/* the reference's chunk size for its tables (CPP/disk_io_routines.cpp:1411,1518) */
static const hsize_t k_table_chunk = 10;

/* ------------------------------------------------------------------ file names */
int tt_p123_h5_filename(char *buf, size_t n, const char *P123_folder, int two_J_3N, int P_3N,
                        int Np_WP, int Nq_WP, int J_2N_max) {
    return snprintf(buf, n, "%s/P123_sparse_JP_%d_%d_Np_%d_Nq_%d_J2max_%d.h5", P123_folder ? P123_folder : ".",
                    two_J_3N, P_3N, Np_WP, Nq_WP, J_2N_max);
}

int tt_p123_h5_subfile_name(char *buf, size_t n, const char *P123_folder, int two_J_3N, int P_3N,
                            int Np_WP, int Nq_WP, int J_2N_max, int thread_idx, int TFC) {
    return snprintf(buf, n, "%s/P123_subsparse_JP_%d_%d_Np_%d_Nq_%d_J2max_%d_TFC_%d_%d.h5",
                    P123_folder ? P123_folder : ".", two_J_3N, P_3N, Np_WP, Nq_WP, J_2N_max, thread_idx, TFC);
}

static char *channel_filename(const char *folder, const tt_fwp_states *fwp, const tt_pw_states *chn) {
    int len = tt_p123_h5_filename(NULL, 0, folder, chn->two_J_3N[0], chn->P_3N[0], fwp->Np_WP, fwp->Nq_WP,
                                  chn->J_2N_max);
    char *s = (char *)malloc((size_t)len + 1);
    if (s) tt_p123_h5_filename(s, (size_t)len + 1, folder, chn->two_J_3N[0], chn->P_3N[0], fwp->Np_WP, fwp->Nq_WP,
                               chn->J_2N_max);
    return s;
}

static tt_status check_channel_args(const tt_fwp_states *fwp, const tt_pw_states *chn) {
    if (!fwp || !chn || !fwp->p_WP || !fwp->q_WP || fwp->Np_WP <= 0 || fwp->Nq_WP <= 0)
        return fail(TT_ERR_ARG, "null or empty momentum state space");
    if (chn->Nalpha <= 0 || !chn->L_2N || !chn->S_2N || !chn->J_2N || !chn->T_2N || !chn->L_1N ||
        !chn->two_J_1N || !chn->two_T_3N || !chn->two_J_3N || !chn->P_3N)
        return fail(TT_ERR_ARG, "null or empty partial-wave channel");
    return TT_OK;
}

/* ------------------------------------------------------------------ writing */
static int write_one(hid_t file, const char *name, hid_t type, const void *value) {
    hsize_t dim = 1;
    hid_t space = H5Screate_simple(1, &dim, NULL);
    if (space < 0) return -1;
    hid_t ds = H5Dcreate2(file, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    herr_t st = ds < 0 ? -1 : H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, value);
    if (ds >= 0 && H5Dclose(ds) < 0) st = -1;
    H5Sclose(space);
    return st < 0 ? -1 : 0;
}

static int write_bounds(hid_t file, const char *title, const char *name, const double *edges, int n_bins) {
    bounds_record *rec = (bounds_record *)calloc((size_t)n_bins + 1, sizeof(bounds_record));
    if (!rec) return -1;
    for (int i = 0; i <= n_bins; ++i) {
        rec[i].index = i;
        rec[i].mesh = edges[i];
    }
    hid_t types[2] = {H5T_NATIVE_INT, H5T_NATIVE_DOUBLE};
    hid_t root = H5Gopen2(file, "/", H5P_DEFAULT);
    herr_t st = root < 0 ? -1
                         : H5TBmake_table(title, root, name, 2, (hsize_t)n_bins + 1, sizeof(bounds_record),
                                          (const char **)k_bounds_fields, k_bounds_offsets, types, k_table_chunk,
                                          NULL, 0, rec);
    if (root >= 0) H5Gclose(root);
    free(rec);
    return st < 0 ? -1 : 0;
}

static int write_pw_table(hid_t file, const tt_pw_states *chn) {
    int n = chn->Nalpha;
    pw_record *rec = (pw_record *)calloc((size_t)n, sizeof(pw_record));
    if (!rec) return -1;
    for (int i = 0; i < n; ++i) {
        rec[i].alpha = i;
        rec[i].L = chn->L_2N[i];
        rec[i].S = chn->S_2N[i];
        rec[i].J = chn->J_2N[i];
        rec[i].T = chn->T_2N[i];
        rec[i].l = chn->L_1N[i];
        rec[i].twoj = chn->two_J_1N[i];
        rec[i].twoTtotal = chn->two_T_3N[i];
        rec[i].twoJtotal = chn->two_J_3N[0];
        rec[i].PARtotal = chn->P_3N[0];
    }
    hid_t types[10];
    for (int f = 0; f < 10; ++f) types[f] = H5T_NATIVE_INT;
    hid_t root = H5Gopen2(file, "/", H5P_DEFAULT);
    herr_t st = root < 0 ? -1
                         : H5TBmake_table("Partial wave info", root, "pw channels", 10, (hsize_t)n, sizeof(pw_record),
                                          (const char **)k_pw_fields, k_pw_offsets, types, k_table_chunk, NULL, 0,
                                          rec);
    if (root >= 0) H5Gclose(root);
    free(rec);
    return st < 0 ? -1 : 0;
}

/* One 1-D little-endian array of n elements of the native type `native`. */
static int write_array(hid_t file, const char *name, hid_t native, const void *data, size_t n) {
    hsize_t dim = (hsize_t)n;
    hid_t space = H5Screate_simple(1, &dim, NULL);
    if (space < 0) return -1;
    hid_t type = H5Tcopy(native);
    herr_t st = type < 0 ? -1 : H5Tset_order(type, H5T_ORDER_LE);
    hid_t ds = st < 0 ? -1 : H5Dcreate2(file, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) st = -1;
    if (st >= 0 && n > 0) st = H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    if (ds >= 0 && H5Dclose(ds) < 0) st = -1;
    if (type >= 0) H5Tclose(type);
    H5Sclose(space);
    return st < 0 ? -1 : 0;
}

tt_status tt_store_P123_h5(const char *filename, const double *val, const int32_t *row, const int32_t *col,
                           size_t nnz, const tt_fwp_states *fwp, const tt_pw_states *chn) {
    quiet_hdf5();
    if (!filename || (nnz > 0 && (!val || !row || !col))) return fail(TT_ERR_ARG, "null argument");
    tt_status s = check_channel_args(fwp, chn);
    if (s != TT_OK) return s;
    hid_t file = H5Fcreate(filename, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (file < 0) return fail(TT_ERR_ARG, "cannot create %s", filename);
    unsigned long long nnz_ull = (unsigned long long)nnz;
    const char *what = NULL;
    if (write_one(file, "Nalpha", H5T_NATIVE_INT, &chn->Nalpha) < 0) what = "Nalpha";
    else if (write_one(file, "Np_WP", H5T_NATIVE_INT, &fwp->Np_WP) < 0) what = "Np_WP";
    else if (write_one(file, "Nq_WP", H5T_NATIVE_INT, &fwp->Nq_WP) < 0) what = "Nq_WP";
    else if (write_one(file, "P123_sparse_dim", H5T_NATIVE_ULLONG, &nnz_ull) < 0) what = "P123_sparse_dim";
    else if (write_bounds(file, "Info p boundaries", "p boundaries", fwp->p_WP, fwp->Np_WP) < 0) what = "p boundaries";
    else if (write_bounds(file, "Info q boundaries", "q boundaries", fwp->q_WP, fwp->Nq_WP) < 0) what = "q boundaries";
    else if (write_pw_table(file, chn) < 0) what = "pw channels";
    else if (write_array(file, "row indices", H5T_NATIVE_INT, row, nnz) < 0) what = "row indices";
    else if (write_array(file, "column indices", H5T_NATIVE_INT, col, nnz) < 0) what = "column indices";
    else if (write_array(file, "values", H5T_NATIVE_DOUBLE, val, nnz) < 0) what = "values";
    herr_t cl = H5Fclose(file);
    if (what) return fail(TT_ERR_ARG, "writing \"%s\" to %s failed", what, filename);
    if (cl < 0) return fail(TT_ERR_ARG, "closing %s failed", filename);
    return TT_OK;
}

/* ------------------------------------------------------------------ reading */
static int read_one(hid_t file, const char *name, hid_t native, void *value) {
    hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
    if (ds < 0) return -1;
    hid_t space = H5Dget_space(ds);
    hssize_t n = space < 0 ? -1 : H5Sget_simple_extent_npoints(space);
    herr_t st = n == 1 ? H5Dread(ds, native, H5S_ALL, H5S_ALL, H5P_DEFAULT, value) : -1;
    if (space >= 0) H5Sclose(space);
    H5Dclose(ds);
    return st < 0 ? -1 : 0;
}

/* Reads a 1-D dataset of exactly n elements into data (as the native type `native`). */
static int read_array(hid_t file, const char *name, hid_t native, void *data, size_t n) {
    hid_t ds = H5Dopen2(file, name, H5P_DEFAULT);
    if (ds < 0) return -1;
    hid_t space = H5Dget_space(ds);
    hssize_t m = space < 0 ? -1 : H5Sget_simple_extent_npoints(space);
    herr_t st = (m >= 0 && (size_t)m == n) ? 0 : -1;
    if (st >= 0 && n > 0) st = H5Dread(ds, native, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    if (space >= 0) H5Sclose(space);
    H5Dclose(ds);
    return st < 0 ? -1 : 0;
}

static tt_status read_bounds_checked(hid_t file, const char *name, const double *edges, int n_bins,
                                     const char *mismatch) {
    hsize_t nfields = 0, nrecords = 0;
    if (H5TBget_table_info(file, name, &nfields, &nrecords) < 0)
        return fail(TT_ERR_ARG, "Erroneous control return-value from calling H5TBread_table()");
    if (nrecords != (hsize_t)n_bins + 1) return fail(TT_ERR_ARG, "%s", mismatch);
    bounds_record *rec = (bounds_record *)calloc((size_t)n_bins + 1, sizeof(bounds_record));
    if (!rec) return fail(TT_ERR_ALLOC, "host allocation failed");
    bounds_record probe;
    size_t sizes[2] = {sizeof(probe.index), sizeof(probe.mesh)};
    herr_t st = H5TBread_table(file, name, sizeof(bounds_record), k_bounds_offsets, sizes, rec);
    tt_status s = TT_OK;
    if (st < 0) s = fail(TT_ERR_ARG, "Erroneous control return-value from calling H5TBread_table()");
    for (int i = 0; s == TT_OK && i <= n_bins; ++i)
        if (rec[i].mesh != edges[i]) s = fail(TT_ERR_ARG, "%s", mismatch);
    free(rec);
    return s;
}

static tt_status read_pw_checked(hid_t file, const tt_pw_states *chn) {
    const char *mismatch = "File-read P123 PW state-space mismatch.";
    int n = chn->Nalpha;
    hsize_t nfields = 0, nrecords = 0;
    if (H5TBget_table_info(file, "pw channels", &nfields, &nrecords) < 0)
        return fail(TT_ERR_ARG, "Erroneous control return-value from calling H5TBread_table()");
    if (nrecords != (hsize_t)n) return fail(TT_ERR_ARG, "%s", mismatch);
    pw_record *rec = (pw_record *)calloc((size_t)n, sizeof(pw_record));
    if (!rec) return fail(TT_ERR_ALLOC, "host allocation failed");
    size_t sizes[10];
    for (int f = 0; f < 10; ++f) sizes[f] = sizeof(int);
    herr_t st = H5TBread_table(file, "pw channels", sizeof(pw_record), k_pw_offsets, sizes, rec);
    tt_status s = TT_OK;
    if (st < 0) s = fail(TT_ERR_ARG, "Erroneous control return-value from calling H5TBread_table()");
    for (int i = 0; s == TT_OK && i < n; ++i) {
        const pw_record *r = &rec[i];
        if (r->L != chn->L_2N[i] || r->S != chn->S_2N[i] || r->J != chn->J_2N[i] || r->T != chn->T_2N[i] ||
            r->l != chn->L_1N[i] || r->twoj != chn->two_J_1N[i] || r->twoTtotal != chn->two_T_3N[i] ||
            r->twoJtotal != chn->two_J_3N[0] || r->PARtotal != chn->P_3N[0])
            s = fail(TT_ERR_ARG, "%s", mismatch);
    }
    free(rec);
    return s;
}

/* Opens a P123 file and checks its sizes, bin edges and partial waves against the caller's;
 * on success *file is open and *nnz holds its element count. */
static tt_status open_checked(const char *filename, const tt_fwp_states *fwp, const tt_pw_states *chn, hid_t *file,
                              size_t *nnz) {
    hid_t f = H5Fopen(filename, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail(TT_ERR_ARG, "cannot open %s as an HDF5 file", filename);
    int Nalpha = -1, Np = -1, Nq = -1;
    unsigned long long n = 0;
    tt_status s = TT_OK;
    if (read_one(f, "Nalpha", H5T_NATIVE_INT, &Nalpha) < 0 || read_one(f, "Np_WP", H5T_NATIVE_INT, &Np) < 0 ||
        read_one(f, "Nq_WP", H5T_NATIVE_INT, &Nq) < 0 || read_one(f, "P123_sparse_dim", H5T_NATIVE_ULLONG, &n) < 0)
        s = fail(TT_ERR_ARG, "Erroneous control return-value from calling H5Dread()");
    else if (Nalpha != chn->Nalpha || Np != fwp->Np_WP || Nq != fwp->Nq_WP)
        s = fail(TT_ERR_ARG, "File-read P123 state-space dimensions (Nalpha, Nq_WP, Np_WP) mismatch.");
    if (s == TT_OK)
        s = read_bounds_checked(f, "p boundaries", fwp->p_WP, fwp->Np_WP, "File-read P123 p-momentum boundaries mismatch.");
    if (s == TT_OK)
        s = read_bounds_checked(f, "q boundaries", fwp->q_WP, fwp->Nq_WP, "File-read P123 q-momentum boundaries mismatch.");
    if (s == TT_OK) s = read_pw_checked(f, chn);
    if (s != TT_OK) {
        H5Fclose(f);
        return s;
    }
    *file = f;
    *nnz = (size_t)n;
    return TT_OK;
}

static tt_status read_triplets(hid_t file, const char *filename, size_t nnz, double *val, int32_t *row, int32_t *col) {
    if (read_array(file, "row indices", H5T_NATIVE_INT, row, nnz) < 0 ||
        read_array(file, "column indices", H5T_NATIVE_INT, col, nnz) < 0 ||
        read_array(file, "values", H5T_NATIVE_DOUBLE, val, nnz) < 0)
        return fail(TT_ERR_ARG, "reading the P123 elements of %s failed (sizes differ from P123_sparse_dim?)", filename);
    return TT_OK;
}

static tt_status alloc_triplets(size_t nnz, double **val, int32_t **row, int32_t **col) {
    size_t m = nnz ? nnz : 1;
    *val = (double *)malloc(m * sizeof(double));
    *row = (int32_t *)malloc(m * sizeof(int32_t));
    *col = (int32_t *)malloc(m * sizeof(int32_t));
    if (!*val || !*row || !*col) {
        free(*val);
        free(*row);
        free(*col);
        *val = NULL;
        *row = *col = NULL;
        return fail(TT_ERR_ALLOC, "host allocation of %zu P123 elements failed", nnz);
    }
    return TT_OK;
}

tt_status tt_read_P123_h5(const char *filename, const tt_fwp_states *fwp, const tt_pw_states *chn, double **val,
                          int32_t **row, int32_t **col, size_t *nnz) {
    quiet_hdf5();
    if (!filename || !val || !row || !col || !nnz) return fail(TT_ERR_ARG, "null argument");
    tt_status s = check_channel_args(fwp, chn);
    if (s != TT_OK) return s;
    hid_t file;
    size_t n = 0;
    s = open_checked(filename, fwp, chn, &file, &n);
    if (s != TT_OK) return s;
    double *v;
    int32_t *r, *c;
    s = alloc_triplets(n, &v, &r, &c);
    if (s == TT_OK) s = read_triplets(file, filename, n, v, r, c);
    H5Fclose(file);
    if (s != TT_OK) {
        free(v);
        free(r);
        free(c);
        return s;
    }
    *val = v;
    *row = r;
    *col = c;
    *nnz = n;
    return TT_OK;
}

/* ------------------------------------------------------------------ sorting */
static tt_status check_indices(int64_t dim, size_t nnz, const int32_t *row, const int32_t *col) {
    for (size_t i = 0; i < nnz; ++i) {
        if (row[i] < 0 || row[i] >= dim) return fail(TT_ERR_ARG, "Encountered row-index larger than dense dimension!");
        if (col[i] < 0 || col[i] >= dim)
            return fail(TT_ERR_ARG, "Encountered column-index larger than dense dimension!");
    }
    return TT_OK;
}

tt_status tt_coo_to_csc(int64_t dim, size_t nnz, const double *val, const int32_t *row, const int32_t *col,
                        double *csc_val, int32_t *csc_row, uint64_t *colptr) {
    if (dim <= 0 || !colptr || (nnz > 0 && (!val || !row || !col || !csc_val || !csc_row)))
        return fail(TT_ERR_ARG, "null argument");
    tt_status s = check_indices(dim, nnz, row, col);
    if (s != TT_OK) return s;
    /* two stable counting sorts: by row, then by column (rows stay ascending in a column) */
    size_t *by_row = (size_t *)malloc((nnz ? nnz : 1) * sizeof(size_t));
    uint64_t *next = (uint64_t *)calloc((size_t)dim + 1, sizeof(uint64_t));
    if (!by_row || !next) {
        free(by_row);
        free(next);
        return fail(TT_ERR_ALLOC, "host allocation failed");
    }
    for (size_t i = 0; i < nnz; ++i) next[row[i] + 1]++;
    for (int64_t r = 0; r < dim; ++r) next[r + 1] += next[r];
    for (size_t i = 0; i < nnz; ++i) by_row[next[row[i]]++] = i;
    memset(colptr, 0, ((size_t)dim + 1) * sizeof(uint64_t));
    for (size_t i = 0; i < nnz; ++i) colptr[col[i] + 1]++;
    for (int64_t c = 0; c < dim; ++c) colptr[c + 1] += colptr[c];
    memcpy(next, colptr, ((size_t)dim + 1) * sizeof(uint64_t));
    for (size_t k = 0; k < nnz; ++k) {
        size_t i = by_row[k];
        uint64_t p = next[col[i]]++;
        csc_val[p] = val[i];
        csc_row[p] = row[i];
    }
    free(by_row);
    free(next);
    return TT_OK;
}

tt_status tt_csc_to_coo_row_major(int64_t dim, int64_t nnz, const double *csc_val, const int32_t *csc_row,
                                  const uint64_t *colptr, double *val, int32_t *row, int32_t *col) {
    if (dim <= 0 || nnz < 0 || !colptr || (nnz > 0 && (!csc_val || !csc_row || !val || !row || !col)))
        return fail(TT_ERR_ARG, "null argument");
    if (colptr[0] != 0 || (int64_t)colptr[dim] != nnz) return fail(TT_ERR_ARG, "colptr inconsistent with nnz");
    for (int64_t c = 0; c < dim; ++c)
        if (colptr[c] > colptr[c + 1]) return fail(TT_ERR_ARG, "colptr not monotone");
    for (int64_t k = 0; k < nnz; ++k)
        if (csc_row[k] < 0 || csc_row[k] >= dim) return fail(TT_ERR_ARG, "row index out of range");
    uint64_t *next = (uint64_t *)calloc((size_t)dim + 1, sizeof(uint64_t));
    if (!next) return fail(TT_ERR_ALLOC, "host allocation failed");
    for (int64_t k = 0; k < nnz; ++k) next[csc_row[k] + 1]++;
    for (int64_t r = 0; r < dim; ++r) next[r + 1] += next[r];
    /* columns in ascending order: inside a row the columns come out ascending */
    for (int64_t c = 0; c < dim; ++c)
        for (uint64_t k = colptr[c]; k < colptr[c + 1]; ++k) {
            uint64_t p = next[csc_row[k]]++;
            val[p] = csc_val[k];
            row[p] = csc_row[k];
            col[p] = (int32_t)c;
        }
    free(next);
    return TT_OK;
}

/* COO triplets in the file order -> sorted by (row, column), the reference's merge order. */
static tt_status sort_row_major(int64_t dim, size_t nnz, double *val, int32_t *row, int32_t *col) {
    double *cv = (double *)malloc((nnz ? nnz : 1) * sizeof(double));
    int32_t *cr = (int32_t *)malloc((nnz ? nnz : 1) * sizeof(int32_t));
    uint64_t *cp = (uint64_t *)malloc(((size_t)dim + 1) * sizeof(uint64_t));
    tt_status s = (cv && cr && cp) ? TT_OK : fail(TT_ERR_ALLOC, "host allocation failed");
    if (s == TT_OK) s = tt_coo_to_csc(dim, nnz, val, row, col, cv, cr, cp);
    if (s == TT_OK) s = tt_csc_to_coo_row_major(dim, (int64_t)nnz, cv, cr, cp, val, row, col);
    free(cv);
    free(cr);
    free(cp);
    return s;
}

/* ------------------------------------------------------------------ recovery merge */
static int file_exists(const char *name) { return access(name, F_OK) == 0; }

tt_status tt_merge_P123_thread_files_h5(const char *P123_folder, int n_threads, int max_TFC, const tt_fwp_states *fwp,
                                        const tt_pw_states *chn, double **val, int32_t **row, int32_t **col,
                                        size_t *nnz) {
    quiet_hdf5();
    if (!val || !row || !col || !nnz || n_threads <= 0) return fail(TT_ERR_ARG, "null argument or n_threads <= 0");
    tt_status s = check_channel_args(fwp, chn);
    if (s != TT_OK) return s;
    const int J = chn->two_J_3N[0], P = chn->P_3N[0], Np = fwp->Np_WP, Nq = fwp->Nq_WP, J2 = chn->J_2N_max;
    /* first pass: the files present and their element counts */
    size_t cap = 16, n_files = 0, total = 0;
    char **names = (char **)malloc(cap * sizeof(char *));
    if (!names) return fail(TT_ERR_ALLOC, "host allocation failed");
    for (int t = 0; t < n_threads && s == TT_OK; ++t) {
        for (int tfc = 0;; ++tfc) {
            int len = tt_p123_h5_subfile_name(NULL, 0, P123_folder, J, P, Np, Nq, J2, t, tfc);
            char *name = (char *)malloc((size_t)len + 1);
            if (!name) {
                s = fail(TT_ERR_ALLOC, "host allocation failed");
                break;
            }
            tt_p123_h5_subfile_name(name, (size_t)len + 1, P123_folder, J, P, Np, Nq, J2, t, tfc);
            if (!file_exists(name)) {
                if (tfc < max_TFC) {
                    fprintf(stderr, "     - WARNING file not found: %s \n"
                                    "       Are you certain input P123 thread files are complete? \n"
                                    "       Program will ignore file and continue ... \n", name);
                    free(name);
                    continue;
                }
                free(name);
                break;
            }
            hid_t file;
            size_t n = 0;
            s = open_checked(name, fwp, chn, &file, &n);
            if (s != TT_OK) {
                free(name);
                break;
            }
            H5Fclose(file);
            if (n_files == cap) {
                char **grown = (char **)realloc(names, 2 * cap * sizeof(char *));
                if (!grown) {
                    free(name);
                    s = fail(TT_ERR_ALLOC, "host allocation failed");
                    break;
                }
                names = grown;
                cap *= 2;
            }
            names[n_files++] = name;
            total += n;
        }
    }
    /* second pass: the elements, appended file by file */
    double *v = NULL;
    int32_t *r = NULL, *c = NULL;
    if (s == TT_OK) s = alloc_triplets(total, &v, &r, &c);
    size_t at = 0;
    for (size_t f = 0; f < n_files && s == TT_OK; ++f) {
        hid_t file;
        size_t n = 0;
        s = open_checked(names[f], fwp, chn, &file, &n);
        if (s != TT_OK) break;
        if (at + n > total) s = fail(TT_ERR_ARG, "%s changed while it was merged", names[f]);
        else s = read_triplets(file, names[f], n, v + at, r + at, c + at);
        H5Fclose(file);
        at += n;
    }
    for (size_t f = 0; f < n_files; ++f) free(names[f]);
    free(names);
    const int64_t dim = (int64_t)chn->Nalpha * Np * Nq;
    if (s == TT_OK) s = sort_row_major(dim, total, v, r, c);
    if (s != TT_OK) {
        free(v);
        free(r);
        free(c);
        return s;
    }
    *val = v;
    *row = r;
    *col = c;
    *nnz = total;
    return TT_OK;
}

/* ------------------------------------------------------------------ fill_P123_arrays */
static double now_ms(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec * 1e3 + ts.tv_nsec * 1e-6;
}

/* Uploads row-major (or any-order) triplets as the device CSC in out, keeping its host arrays. */
static tt_status upload_triplets(tt_ctx *ctx, int64_t dim, size_t nnz, const double *val, const int32_t *row,
                                 const int32_t *col, tt_csc *out) {
    double *cv = (double *)malloc((nnz ? nnz : 1) * sizeof(double));
    int32_t *cr = (int32_t *)malloc((nnz ? nnz : 1) * sizeof(int32_t));
    uint64_t *cp = (uint64_t *)malloc(((size_t)dim + 1) * sizeof(uint64_t));
    tt_status s = (cv && cr && cp) ? TT_OK : fail(TT_ERR_ALLOC, "host allocation failed");
    if (s == TT_OK) s = tt_coo_to_csc(dim, nnz, val, row, col, cv, cr, cp);
    if (s == TT_OK) {
        s = tt_csc_upload(ctx, dim, (int64_t)nnz, cv, cr, cp, out);
        if (s != TT_OK) fail(s, "%s", tt_last_error());
    }
    if (s != TT_OK) {
        free(cv);
        free(cr);
        free(cp);
        return s;
    }
    out->val = cv;
    out->row = cr;
    out->colptr = cp;
    return TT_OK;
}

tt_status tt_fill_P123_arrays_h5(tt_ctx *ctx, const tt_run_params *rp, const tt_fwp_states *fwp,
                                 const tt_pw_states *chn, int production_run, const char *P123_folder,
                                 int calculate_and_store, int recovery, int n_threads, tt_csc *out,
                                 tt_p123_stats *stats) {
    quiet_hdf5();
    if (!ctx || !rp || !out || !P123_folder) return fail(TT_ERR_ARG, "null argument");
    tt_status s = check_channel_args(fwp, chn);
    if (s != TT_OK) return s;
    if (fwp->Np_WP != rp->Np_WP || fwp->Nq_WP != rp->Nq_WP)
        return fail(TT_ERR_ARG, "run parameters and momentum state space differ in Np_WP or Nq_WP");
    const double t0 = now_ms();
    if (stats) memset(stats, 0, sizeof(*stats));
    memset(out, 0, sizeof(*out));
    const int64_t dim = (int64_t)chn->Nalpha * fwp->Np_WP * fwp->Nq_WP;
    char *filename = channel_filename(P123_folder, fwp, chn);
    if (!filename) return fail(TT_ERR_ALLOC, "host allocation failed");
    double *val = NULL;
    int32_t *row = NULL, *col = NULL;
    size_t nnz = 0;
    if (!calculate_and_store) {
        s = tt_read_P123_h5(filename, fwp, chn, &val, &row, &col, &nnz);
        if (s == TT_OK) s = upload_triplets(ctx, dim, nnz, val, row, col, out);
    } else if (recovery) {
        s = tt_merge_P123_thread_files_h5(P123_folder, n_threads, 0, fwp, chn, &val, &row, &col, &nnz);
        if (s == TT_OK) s = upload_triplets(ctx, dim, nnz, val, row, col, out);
        if (s == TT_OK) s = tt_store_P123_h5(filename, val, row, col, nnz, fwp, chn);
    } else {
        s = tt_fill_P123_arrays(ctx, rp, fwp, chn, production_run, out, stats);
        if (s == TT_OK) s = tt_csc_download(ctx, out);
        if (s != TT_OK) fail(s, "%s", tt_last_error());
        nnz = (size_t)out->nnz;
        if (s == TT_OK) s = alloc_triplets(nnz, &val, &row, &col);
        if (s == TT_OK) s = tt_csc_to_coo_row_major(dim, out->nnz, out->val, out->row, out->colptr, val, row, col);
        if (s == TT_OK) s = tt_store_P123_h5(filename, val, row, col, nnz, fwp, chn);
    }
    free(filename);
    free(val);
    free(row);
    free(col);
    if (s != TT_OK) {
        tt_csc_free(out);
        return s;
    }
    if (stats && (!calculate_and_store || recovery)) stats->ms_total = now_ms() - t0;
    return TT_OK;
}
