/* tictac_h5.h — the P123 HDF5 cache of the AGS/WPCD solve path (libtictac_h5.so).
 *
 * A plain-C library over libtictac.so and libhdf5 (with its table API, libhdf5_hl) that
 * writes and reads the reference's per-channel P123 files in the reference's layout, so
 * that files written by either program are read by the other:
 *
 *   <P123_folder>/P123_sparse_JP_<2J>_<P>_Np_<Np>_Nq_<Nq>_J2max_<J2max>.h5
 *     "Nalpha", "Np_WP", "Nq_WP"      int[1]
 *     "P123_sparse_dim"               unsigned long long[1]
 *     "p boundaries", "q boundaries"  tables {int "index", double "mesh point"}, N+1 records
 *     "pw channels"                   table of the channel's partial waves (10 int fields)
 *     "row indices", "column indices" int[nnz]     (little endian)
 *     "values"                        double[nnz]  (little endian)
 *
 * The nonzeros are stored as row-major sorted COO triplets, as the reference stores them
 * after its merge (CPP/make_permutation_matrix.cpp:1136-1141).
 *
 * Errors: every entry point returns a tt_status of include/tictac.h; tt_h5_last_error()
 * gives the message of the calling thread's last failure.  libhdf5 is not assumed to be
 * thread-safe: call these from one thread at a time. */
#ifndef TICTAC_H5_H
#define TICTAC_H5_H

#include <stddef.h>
#include <stdint.h>

#include "tictac.h"

#ifdef __cplusplus
extern "C" {
#endif

const char *tt_h5_last_error(void);
/* Frees arrays returned by the readers below. */
void tt_h5_free(void *p);

/* The channel's P123 file name, CPP/make_permutation_matrix.cpp:1320-1323.  Returns the
 * length of the full name (as snprintf; buf may be NULL when n is 0). */
int tt_p123_h5_filename(char *buf, size_t n, const char *P123_folder, int two_J_3N, int P_3N,
                        int Np_WP, int Nq_WP, int J_2N_max);
/* The name of one thread's spill file (thread file count TFC),
 * generate_subarray_file_name, CPP/make_permutation_matrix.cpp:143-156. */
int tt_p123_h5_subfile_name(char *buf, size_t n, const char *P123_folder, int two_J_3N, int P_3N,
                            int Np_WP, int Nq_WP, int J_2N_max, int thread_idx, int TFC);

/* Replaces store_sparse_permutation_matrix_for_3N_channel_h5
 * (CPP/disk_io_routines.cpp:1078-1164, with 1286-1302, 1336-1352, 1386-1432, 1474-1545,
 * 1625-1721): writes nnz COO triplets (in the order given) with the channel's sizes, bin
 * edges and partial waves; an existing file is replaced. */
tt_status tt_store_P123_h5(const char *filename, const double *val, const int32_t *row,
                           const int32_t *col, size_t nnz, const tt_fwp_states *fwp,
                           const tt_pw_states *chn);

/* Replaces read_sparse_permutation_matrix_for_3N_channel_h5
 * (CPP/disk_io_routines.cpp:1166-1284, with 1304-1334, 1354-1384, 1434-1472, 1547-1615,
 * 1723-1803).  The file's (Nalpha, Np_WP, Nq_WP), bin edges (compared exactly) and partial
 * waves must equal the caller's, or TT_ERR_ARG with the reference's message:
 *   "File-read P123 state-space dimensions (Nalpha, Nq_WP, Np_WP) mismatch."
 *   "File-read P123 p-momentum boundaries mismatch."  (and q-)
 *   "File-read P123 PW state-space mismatch."
 * The triplets are returned in the file's order in newly allocated arrays (tt_h5_free). */
tt_status tt_read_P123_h5(const char *filename, const tt_fwp_states *fwp, const tt_pw_states *chn,
                          double **val, int32_t **row, int32_t **col, size_t *nnz);

/* Replaces read_and_merge_thread_files_to_single_array
 * (CPP/make_permutation_matrix.cpp:946-1141), the P123_recovery mode: reads the spill files
 * of threads 0..n_threads-1 with TFC 0, 1, ... and merges them into one row-major sorted COO.
 * A file missing below max_TFC is skipped with a warning on stderr, as the reference does;
 * past max_TFC the scan of a thread goes on while files exist (max_TFC <= 0: only the files
 * that exist from TFC 0 on).  Each file is checked as by tt_read_P123_h5. */
tt_status tt_merge_P123_thread_files_h5(const char *P123_folder, int n_threads, int max_TFC,
                                        const tt_fwp_states *fwp, const tt_pw_states *chn,
                                        double **val, int32_t **row, int32_t **col, size_t *nnz);

/* COO triplets in any order -> CSC with rows ascending inside each column, the sort and
 * conversion of CPP/General_functions/matrix_routines.cpp:222-244,287-326 (called at
 * CPP/main.cpp:226-244).  Outputs are caller-allocated (nnz values and rows, dim+1 column
 * pointers).  An index outside [0, dim) gives TT_ERR_ARG. */
tt_status tt_coo_to_csc(int64_t dim, size_t nnz, const double *val, const int32_t *row,
                        const int32_t *col, double *csc_val, int32_t *csc_row, uint64_t *colptr);
/* CSC -> COO triplets sorted by (row, column), the order of the reference's files. */
tt_status tt_csc_to_coo_row_major(int64_t dim, int64_t nnz, const double *csc_val,
                                  const int32_t *csc_row, const uint64_t *colptr, double *val,
                                  int32_t *row, int32_t *col);

/* Replaces fill_P123_arrays with its file cache (CPP/make_permutation_matrix.cpp:1297-1403
 * and the recovery branch of 1158-1252), then the CSC conversion of CPP/main.cpp:226-244:
 *   calculate_and_store != 0: P123 is built on the GPU (tt_fill_P123_arrays) or, with
 *     recovery != 0, merged from the spill files of n_threads threads; it is then stored
 *     to the channel's file in P123_folder;
 *   calculate_and_store == 0: P123 is read from that file.
 * out receives the device CSC with its host arrays (tt_csc_free frees both).  stats may be
 * NULL; it is filled by the GPU build only (zeros otherwise, with ms_total the wall time). */
tt_status tt_fill_P123_arrays_h5(tt_ctx *ctx, const tt_run_params *rp, const tt_fwp_states *fwp,
                                 const tt_pw_states *chn, int production_run,
                                 const char *P123_folder, int calculate_and_store, int recovery,
                                 int n_threads, tt_csc *out, tt_p123_stats *stats);

#ifdef __cplusplus
}
#endif

#endif /* TICTAC_H5_H */
