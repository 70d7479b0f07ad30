/*
 * tictac.h -- C-ABI of the MI355X (gfx950) AGS/WPCD solve path.
 *
 * Drop-in boundary for the three solve-path entry points of seanbsm/Tic-tac
 * (all called from CPP/main.cpp) plus the state-space builders whose
 * conventions they consume.  Plain C: pointers, sizes, status codes; no C++ or
 * torch types.  Array layouts, index order (idx = alpha*Nq*Np + q*Np + p),
 * units (MeV) and on-shell conventions are the reference's.
 *
 * Errors: every entry point returns tt_status; tt_last_error() gives the
 * message of the last failure on the calling thread.  Nothing calls exit()
 * (the reference's raise_error exits with status 0, CPP/error_management.cpp:4-14).
 *
 * Complex arrays are interleaved (re, im) doubles, i.e. the memory layout of
 * std::complex<double> / cdouble in the reference.
 */
#ifndef TICTAC_H
#define TICTAC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TT_OK = 0,
    TT_ERR_ARG = 1,          /* invalid argument or inconsistent sizes          */
    TT_ERR_HIP = 2,          /* HIP runtime error                               */
    TT_ERR_ALLOC = 3,        /* host or device allocation failed                */
    TT_ERR_PHYSICS = 4,      /* e.g. unphysical bound states, T_lab not in space */
    TT_ERR_NOT_CONVERGED = 5,/* iterative solve hit its iteration limit          */
    TT_ERR_UNSUPPORTED = 6   /* option combination not implemented               */
} tt_status;

const char *tt_last_error(void);
const char *tt_version(void);

/* ------------------------------------------------------------------ context */
typedef struct tt_ctx tt_ctx;
/* Binds a device and creates one default-priority HIP stream. */
tt_status tt_ctx_create(int device, tt_ctx **out);
void tt_ctx_destroy(tt_ctx *ctx);
/* The stream all work of this context is issued on (a hipStream_t). */
void *tt_ctx_stream(tt_ctx *ctx);
tt_status tt_ctx_synchronize(tt_ctx *ctx);

/* ------------------------------------------------------------ run parameters
 * Subset of run_params (CPP/type_defs.h:87-123) used by the path. */
typedef struct {
    int two_J_3N_max;
    int J_2N_max;
    int Np_WP, Nq_WP;
    int Np_per_WP;          /* quadrature points per p bin for V_WP (default 8) */
    int Nphi, Nx;           /* P123 quadrature (default 48 x 48)               */
    double chebyshev_s;     /* MeV                                              */
    double chebyshev_t;
    int tensor_force;
    int isospin_breaking_1S0;
    int midpoint_approx;
    int swave_only;         /* the S-wave filter of CPP/make_pw_symm_states.cpp:176-179 */
} tt_run_params;

void tt_run_params_default(tt_run_params *rp);   /* CPP/set_run_parameters.cpp:577-615 */

/* ------------------------------------------------------------- state spaces */
/* Mirrors pw_3N_statespace (CPP/type_defs.h:15-29).  Arrays owned by the library. */
typedef struct {
    int Nalpha;
    int J_2N_max;
    int *L_2N, *S_2N, *J_2N, *T_2N, *L_1N, *two_J_1N, *two_J_3N, *two_T_3N, *P_3N;
    int *chn_3N_idx;        /* N_chn_3N + 1 entries */
    int N_chn_3N;
} tt_pw_states;

/* Replaces construct_symmetric_pw_states (CPP/make_pw_symm_states.cpp:62-284). */
tt_status tt_construct_pw_states(const tt_run_params *rp, tt_pw_states *out);
void tt_pw_states_free(tt_pw_states *pw);
/* View of one 3N channel (pointers into pw; nothing to free), as CPP/main.cpp:177-192. */
tt_status tt_pw_channel(const tt_pw_states *pw, int chn, tt_pw_states *view);

/* Mirrors fwp_statespace (CPP/type_defs.h:32-47); momentum WPs. */
typedef struct {
    int Np_WP, Nq_WP, Np_per_WP, Nq_per_WP;
    double *p_WP, *q_WP;        /* bin edges, N+1 each             */
    double *p, *wp;             /* per-bin quadrature (Np*Np_per_WP) or midpoints */
    double *q, *wq;
    double *norm_p, *norm_q;    /* sqrt(bin width)                 */
} tt_fwp_states;

/* Replaces make_fwp_statespace (CPP/make_wp_states.cpp:116-211), Chebyshev grids. */
tt_status tt_make_fwp_statespace(const tt_run_params *rp, tt_fwp_states *out);
void tt_fwp_states_free(tt_fwp_states *f);

/* ------------------------------------------------------ interaction plugin
 * Replaces the potential_model factory (CPP/Interactions/potential_model.cpp:18-62)
 * for the models of this path: "malfliet_tjon" and "synthetic".
 * V(...) follows CPP/Interactions/potential_model.h:20: six slots
 * [0] S=0, [1] S=1 L=J, [2] L=L'=J-1, [3] (+-), [4] (-+), [5] L=L'=J+1, in MeV^-2. */
typedef struct tt_potential tt_potential;
tt_status tt_potential_create(const char *name, tt_potential **out);
void tt_potential_destroy(tt_potential *pot);
/* update_parameters (potential_model.h:16).  "synthetic": 12 strengths theta
 * (3 Yukawa [MeV fm], C_1S0, C_3S1 [fm^2], C_3P0, C_1P1, C_3P1, C_SD [fm^4],
 * 3 separable [fm^2]) followed by the 3 separable ranges beta [fm^-1]. */
tt_status tt_potential_update_parameters(tt_potential *pot, const double *params, int n);
tt_status tt_potential_V(tt_potential *pot, double p_in, double p_out, int coupled,
                         int S, int J, int T, int Tz, double V[6]);

/* Number of uncoupled / coupled 2N channels (CPP/main.cpp:273-289). */
void tt_num_2N_states(const tt_run_params *rp, int *num_unco, int *num_coup);

/* Replaces calculate_potential_matrices_array_in_WP_basis (CPP/make_potential_matrix.cpp:51-430).
 * V_unco: num_unco*Np*Np, V_coup: num_coup*4*Np*Np (row-major), caller-allocated.
 * pw_full is the full state space (as the reference passes it, CPP/main.cpp:308-313). */
tt_status tt_calculate_potential_matrices(const tt_run_params *rp, const tt_fwp_states *fwp,
                                          const tt_pw_states *pw_full, tt_potential *pot,
                                          double *V_unco, double *V_coup);

/* Mirrors swp_statespace (CPP/type_defs.h:50-61); arrays caller-owned. */
typedef struct {
    int Np_WP, Nq_WP;
    int num_2N_unco_states, num_2N_coup_states;
    double E_bound;             /* deuteron energy, negative, MeV   */
    double *e_SWP_unco;         /* num_unco*(Np+1)                  */
    double *e_SWP_coup;         /* num_coup*2*(Np+1)                */
    double *C_SWP_unco;         /* num_unco*Np*Np, C[basis][state]  */
    double *C_SWP_coup;         /* num_coup*4*Np*Np                 */
    double *q_WP;               /* Nq+1 edges                       */
} tt_swp_states;

/* Replaces make_swp_states (CPP/make_swp_states.cpp:221-471); fills the caller's
 * arrays in swp (sizes as above) and E_bound.  TT_ERR_PHYSICS unless exactly one
 * 3S1 bound state and no other bound state. */
tt_status tt_make_swp_states(const tt_run_params *rp, const tt_fwp_states *fwp,
                             const tt_pw_states *pw_full, const double *V_unco,
                             const double *V_coup, tt_swp_states *swp);

/* Mirrors channel_os_indexing (CPP/type_defs.h:76-85). */
typedef struct {
    int num_T_lab;              /* number of on-shell q bins        */
    int *q_com_idx;             /* num_T_lab                        */
    double *T_lab, *q_com, *E_com;  /* num_T_lab                    */
    int num_deuteron_states;
    int *deuteron_idx;          /* indices into the channel's alphas */
    /* Breakup on-shell cells (tt_find_breakup_bins; 0 / NULL without): entries
     * q_com_BU_idx[j] .. q_com_BU_idx[j+1]-1 belong to on-shell energy j, entry b is the
     * state (alpha, q, p) = alphapq_idx[3b .. 3b+2]. */
    int num_BU_chns;
    int *q_com_BU_idx;          /* num_T_lab + 1                     */
    int *alphapq_idx;           /* 3 * num_BU_chns                   */
} tt_channel_os;

/* Replaces find_on_shell_bins (CPP/run_organizer.cpp:4-131, elastic) and
 * find_deuteron_channels (CPP/run_organizer.cpp:342-378) for one channel. */
tt_status tt_find_on_shell(const tt_swp_states *swp, const tt_pw_states *chn,
                           const double *T_lab_in, int n_T_lab_in, tt_channel_os *out);
/* Replaces the breakup part of find_on_shell_bins (CPP/run_organizer.cpp:136-339,
 * include_breakup_channels=true): for every on-shell energy E_com of os (E_com + E_bound
 * >= 0), every (alpha, q, p) cell whose energy range [e_p,lo + E_q,lo, e_p,hi + E_q,hi]
 * (SWP p-bin energies, E_q = q^2 / (2 mu1_free)) holds E_com with e_p,lo >= 0, followed by
 * the three neighbour cells of the quadrant that holds E_com, as the reference lists them
 * (a neighbour may lie one bin outside the grid: its flat index alpha*Nq*Np + q*Np + p is
 * what the solves read, as the reference does; one outside the state space, where the
 * reference reads outside the row, gets NaN in U_BU).  Fills the breakup
 * fields of os (replacing earlier ones); TT_ERR_PHYSICS when an energy at or above the
 * breakup threshold has no such cell. */
tt_status tt_find_breakup_bins(const tt_run_params *rp, const tt_swp_states *swp,
                               const tt_pw_states *chn, tt_channel_os *os);
/* The SWP p-bin energy edges of every alpha of a channel, edges[alpha*(Np+1) + p]: the
 * per-alpha choice of e_SWP_unco / e_SWP_coup (upper or lower half) that the reference makes
 * inline in CPP/make_resolvent.cpp:118-156, CPP/run_organizer.cpp:157-203 and
 * CPP/disk_io_routines.cpp:507-540 (the breakup output's p-bin boundaries). */
tt_status tt_channel_p_edges(const tt_run_params *rp, const tt_swp_states *swp, const tt_pw_states *chn,
                             double *edges);
void tt_channel_os_free(tt_channel_os *os);

/* ---------------------------------------------------------------- P123 CSC
 * The CSC of CPP/main.cpp:226-244: values (double), row indices (int32, rows
 * ascending inside a column), column pointers (uint64, dim+1).  Host arrays are
 * optional; the device mirror is owned by the library. */
typedef struct {
    int64_t dim, nnz;
    double *val;                /* host, may be NULL */
    int32_t *row;
    uint64_t *colptr;
    void *dev;                  /* library-owned device mirror (opaque) */
} tt_csc;

/* Wraps caller host arrays (copied to the device; caller keeps ownership of its arrays). */
tt_status tt_csc_upload(tt_ctx *ctx, int64_t dim, int64_t nnz, const double *val,
                        const int32_t *row, const uint64_t *colptr, tt_csc *out);
/* Copies the device CSC to newly malloc'ed host arrays in m (freed by tt_csc_free). */
tt_status tt_csc_download(tt_ctx *ctx, tt_csc *m);
void tt_csc_free(tt_csc *m);

/* Replaces fill_P123_arrays (CPP/make_permutation_matrix.h:79-87,
 * CPP/make_permutation_matrix.cpp:1297-1403) followed by the column-major sort and
 * CSC conversion of CPP/main.cpp:226-244.  Built on the GPU; out->val/row/colptr are
 * left NULL (call tt_csc_download for a host copy).  production_run = 0 gives the
 * placeholder values cos(row)*cos(col) of CPP/make_permutation_matrix.cpp:817-822. */
typedef struct {
    double mask_density;        /* fraction of (q',p',q,p) cells with bin overlap */
    int64_t n_cells;            /* overlapping cells                               */
    int n_tuples;               /* distinct angular functions                      */
    double ms_total;            /* wall time of the build                          */
} tt_p123_stats;
tt_status tt_fill_P123_arrays(tt_ctx *ctx, const tt_run_params *rp, const tt_fwp_states *fwp,
                              const tt_pw_states *chn, int production_run, tt_csc *out,
                              tt_p123_stats *stats);

/* ------------------------------------------------------------- resolvent G
 * Replaces calculate_resolvent_array_in_SWP_basis (CPP/make_resolvent.h:36-40),
 * batched over energies: G[j*dim + idx] at total energy E[j] (= E_com + E_bound,
 * CPP/main.cpp:391).  G is host memory (interleaved complex, nE*dim). */
tt_status tt_calculate_resolvent(tt_ctx *ctx, const tt_run_params *rp, const tt_swp_states *swp,
                                 const tt_pw_states *chn, const double *E, int nE, double *G);

/* ------------------------------------------------------------ Faddeev solve */
typedef enum {
    TT_SOLVE_PADE = 0,          /* reference semantics: Neumann terms + [N/N] Pade (default) */
    TT_SOLVE_DENSE = 1,         /* solve_dense=true: (1 - A G) U = A, dense LU (small dims)   */
    TT_SOLVE_GMRES = 2,         /* converged FP64 GMRES on y (1 - G A) = e_r^T A              */
    TT_SOLVE_MIXED = 3          /* FP32 products + FP64 iterative refinement                 */
} tt_solve_method;

/* How the product over P123 is formed (results agree to rounding).  The cell-block form covers
 * channels with Nalpha <= 64 (J_2N_max <= 3); a solve of a larger channel (J_2N_max >= 4 gives
 * Nalpha up to 139) asking for TT_P123_CELLS takes the CSC gather instead.  tt_solve_gradient and
 * the column operator of tt_apply_rows need the cell-block form: TT_ERR_UNSUPPORTED above 64. */
typedef enum {
    TT_P123_CELLS = 0,          /* dense Nalpha x Nalpha blocks per overlapping bin cell (default) */
    TT_P123_CSC = 1             /* gather over the reference's CSC arrays                          */
} tt_p123_format;

typedef struct {
    int method;                 /* tt_solve_method                         */
    double tol;                 /* relative residual target (GMRES/MIXED)   */
    int max_iter;               /* total Krylov iterations per system       */
    int restart;                /* GMRES restart length                     */
    int NM_max;                 /* Pade order limit (reference: 14)         */
    /* Optional (Pade): Neumann terms a_n, interleaved complex, laid out
     * [(d_row*nd + d_col)*nE + j][n] for n < 2*NM_max+1; terms not computed stay 0.
     * The reference appends them to neumann_terms*.txt (CPP/solve_faddeev.cpp:907-917). */
    double *neumann_terms;
    int p123_format;            /* tt_p123_format                           */
    int use_graphs;             /* 1: each Arnoldi step is captured once as a HIP graph and
                                 * replayed; 0: plain stream launches; -1 (default): graphs for
                                 * lockstep groups of 4 or more channels (tt_solve_batch),
                                 * where they save host launch time, plain launches otherwise
                                 * (for one channel a graph launch costs more than its
                                 * kernels' launches).  The context keeps the operator and
                                 * Krylov workspace of recent channels (keyed by the P123
                                 * object, the channel and the exact contents of V and C), so
                                 * repeated solves of a channel only upload G. */
    int resolvent_on_device;    /* 1: ignore G (may be NULL) and compute the resolvent of the
                                 * on-shell energies E_com + E_bound on the device, straight
                                 * into the solver's buffer (no host round trip); 0: use G. */
    /* Optional breakup amplitudes (U_BU_array of solve_faddeev_equations): when os carries
     * breakup cells, nd * num_BU_chns complex values U_BU[d_row*num_BU_chns + b] = the U row
     * of (deuteron d_row, energy of b) at b's state (CPP/solve_faddeev.cpp:867-891,
     * 1177-1206, 1300-1365, 1388-1409 for Pade; read from the same solution for the other
     * methods).  NULL: not computed. */
    double *U_BU;
} tt_solve_opts;
void tt_solve_opts_default(tt_solve_opts *o);

typedef struct {
    int n_apply;                /* A applications (FP64 + FP32)             */
    int n_apply_fp32;
    int iterations;             /* Krylov iterations (max over systems) or Neumann terms */
    double residual;            /* final relative residual (max over systems) */
    double ms_total;
    double ms_apply;            /* device time inside A applications        */
    /* The product over P123 (dominant kernel: the cell-block MFMA product, or the
     * gather over the CSC): device time summed over its launches (HIP events on the
     * solve stream around each launch) and the algorithmic bytes those launches must
     * move: per launch ncell*Na^2*s_v + 4*(ncell + NqNp + 1) + 2*dim*r*s_c for the
     * cell blocks, nnz*(s_v+4) + 8*(dim+1) + 2*dim*r*s_c for the CSC (s_v, s_c:
     * value and complex-vector element sizes of the precision used). */
    int n_spmm;
    double ms_spmm;
    double spmm_bytes;
    /* the FP32 launches among them (mixed precision); in graph replays only the first
     * Arnoldi step of each cycle is timed */
    int n_spmm_fp32;
    double ms_spmm_fp32;
    double spmm_bytes_fp32;
} tt_solve_stats;

/* Replaces solve_faddeev_equations (CPP/solve_faddeev.h:25-38).
 * U: host, interleaved complex, U[d_row*nd*nE + d_col*nE + j] (CPP/solve_faddeev.cpp:424).
 * G: host, nE*dim complex (from tt_calculate_resolvent), nE = os->num_T_lab.
 * V_unco/V_coup: as filled by tt_calculate_potential_matrices. */
tt_status tt_solve_faddeev_equations(tt_ctx *ctx, const tt_run_params *rp, double *U,
                                     const double *G, const tt_csc *P123,
                                     const double *V_unco, const double *V_coup,
                                     const tt_swp_states *swp, const tt_channel_os *os,
                                     const tt_pw_states *chn, const tt_solve_opts *opts,
                                     tt_solve_stats *stats);

/* Batched solve: every item is one 3N channel with all its on-shell bins (the energies
 * requested for it), solved as one block Krylov system (r = n_d * n_bins right-hand sides
 * share each A-application).  G is computed on the device for each item (E_com + E_bound,
 * os->E_com required).  With GMRES or MIXED all items' Krylov solves advance in lockstep,
 * spread over the context's stream lanes (TT_LANES, default 6; with TT_GROUPED=1 each stage of
 * a step is one grouped launch for the items of a lane, the lane's step captured as one graph
 * when step graphs are on); Pade and dense items are solved
 * one after another.  Each item goes through its channel's cached plan (operator, workspace,
 * captured step graphs), so a repeated batch over the same channels only recomputes G and
 * solves.  No reference counterpart: it replaces
 * the per-channel, per-energy loop of CPP/main.cpp:166-492 around solve_faddeev_equations.
 * Every item is attempted; each gets its own status; the return value is the first failure. */
typedef struct {
    const tt_pw_states *chn;    /* channel view (tt_pw_channel)                    */
    const tt_channel_os *os;    /* its on-shell bins (tt_find_on_shell)            */
    const tt_csc *P123;         /* its P123 (tt_fill_P123_arrays / tt_csc_upload)  */
    double *U;                  /* out: nd*nd*nE complex, layout as above          */
    tt_solve_stats *stats;      /* optional per-item statistics                    */
    tt_status status;           /* out                                             */
    double *U_BU;               /* optional out: breakup amplitudes as tt_solve_opts.U_BU */
} tt_batch_item;
tt_status tt_solve_batch(tt_ctx *ctx, const tt_run_params *rp, const double *V_unco,
                         const double *V_coup, const tt_swp_states *swp, tt_batch_item *items,
                         int n_items, const tt_solve_opts *opts);

/* Parameter walk: the PSI loop of CPP/main.cpp:126-136,254-466 (parameter_walk=true, parameter
 * sets read by read_parameter_sample_list, CPP/disk_io_routines.cpp:43-78: one set of n_params values per
 * row of params, n_sets rows) for the listed 3N channels at once.  For every set: the
 * potential's parameters are set (update_parameters), V_WP, the SWP states, each channel's
 * on-shell bins of T_lab and G are rebuilt, and all channels are solved in one lockstep batch
 * (as tt_solve_batch, with opts).  The host work of set s + 1 runs while the device solves set
 * s.  channels: indices into pw_full's 3N channels, P123[c] the P123 of channels[c] (parameter
 * independent, built once).  U: n_sets * n_chn blocks of `capacity` complex values, block
 * (s, c) at complex offset (s * n_chn + c) * capacity in the layout of
 * tt_solve_faddeev_equations; n_out / status: n_sets * n_chn entries (elements written, and
 * each item's status).  A set whose SWP states fail (e.g. no single 3S1 bound state) gets that
 * status on all its items; the walk goes on.  Returns the first failure.  The potential keeps
 * the last set's parameters.  No reference counterpart as one call: it replaces the PSI loop
 * around solve_faddeev_equations. */
tt_status tt_parameter_walk(tt_ctx *ctx, const tt_run_params *rp, const tt_fwp_states *fwp,
                            const tt_pw_states *pw_full, tt_potential *pot, const double *params,
                            int n_params, int n_sets, const int *channels, const tt_csc *const *P123,
                            int n_chn, const double *T_lab, int n_T_lab, double *U, int capacity,
                            int *n_out, tt_status *status, const tt_solve_opts *opts);

/* ------------------------------------------------------------ derivatives
 * No reference counterpart (the reference has no derivatives).  For a potential that is
 * linear in its first n_theta parameters (the "synthetic" model: its 12 strengths, with the
 * ranges beta fixed), V_WP(theta) = sum_i theta_i V_i with the basis matrices
 * V_i = V_WP(theta = e_i) filled here: Vb_unco n_theta*num_unco*Np*Np, Vb_coup
 * n_theta*num_coup*4*Np*Np.  params: the full parameter vector (its first n_theta entries
 * are ignored); the potential keeps params on return. */
tt_status tt_potential_basis_matrices(const tt_run_params *rp, const tt_fwp_states *fwp,
                                      const tt_pw_states *pw_full, tt_potential *pot, const double *params,
                                      int n_params, int n_theta, double *Vb_unco, double *Vb_coup);

typedef struct {
    double ms_forward;          /* V, SWP states, G and the FP64 row solve            */
    double ms_total;            /* forward + adjoint column solve + assembly          */
    int n_apply_forward;        /* row-operator applications                          */
    int n_apply_adjoint;        /* column-operator applications                       */
    double residual_forward, residual_adjoint;
} tt_grad_stats;

/* U(theta) of one 3N channel (index into pw_full) at the on-shell bins of T_lab, and, when
 * dU_dtheta is not NULL, dU/dtheta_i for i < n_theta by one adjoint (column) solve:
 * V = sum_i theta_i V_i, SWP states, on-shell bins and G are rebuilt for theta, the on-shell
 * U solved by FP64 GMRES (opts->tol).  The derivative includes the change of the SWP basis
 * (eigenvectors, bin edges, deuteron energy) and of G; the choice of on-shell bins is held
 * fixed.  U: nd*nd*nE complex ([d_row][d_col][j], as tt_solve_faddeev_equations);
 * dU_dtheta: n_theta consecutive blocks of that layout (block i at offset i*nd*nd*nE);
 * capacity: complex elements U can hold (dU_dtheta must hold n_theta*capacity); n_out:
 * nd*nd*nE, the block size written. */
tt_status tt_solve_gradient(tt_ctx *ctx, const tt_run_params *rp, const tt_fwp_states *fwp,
                            const tt_pw_states *pw_full, int channel, const tt_csc *P123,
                            const double *Vb_unco, const double *Vb_coup, int n_theta,
                            const double *theta, const double *T_lab, int n_T_lab, double *U,
                            double *dU_dtheta, int capacity, int *n_out, const tt_solve_opts *opts,
                            tt_grad_stats *stats);

/* The choices of the cell-product autotuner so far (no reference counterpart): one text line
 * per tuned shape, "dev D row|column fp32|fp64 Nalpha N R2 R ncell C -> item_len L un U (T us)".
 * Writes at most len bytes (NUL-terminated) into buf; returns the size the whole report needs. */
int tt_cell_tuning(char *buf, int len);

/* Raw kernel-matrix application for tests and profiling: Y = (X .* G_sel) A in
 * the reference's row form (CPP/solve_faddeev.cpp:956-1088), r rows.
 * X, Y: host, r*dim complex (row-major).  g_index[k] selects the energy of row k
 * (-1: no G factor).  G: host nE*dim complex.  flags: bit 0 = FP32 products,
 * bit 1 = gather over the CSC instead of the cell blocks, bit 2 = the column operator
 * Y_k = A (G_k .* X_k) (products with P123 itself) instead of the row form. */
tt_status tt_apply_rows(tt_ctx *ctx, const tt_run_params *rp, const tt_csc *P123,
                        const double *V_unco, const double *V_coup, const tt_swp_states *swp,
                        const tt_pw_states *chn, const double *G, int nE, const int *g_index,
                        int r, const double *X, double *Y, int flags);

/* No counterpart in the reference.  tt_apply_rows for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * the FP64 row form Y = (X .* G_sel) A (flags 0: cell blocks, or the CSC gather above 64 partial
 * waves; flags 2: the CSC gather) with every device buffer beyond the P123 object (operator
 * tables, work items, staged X, G and Y, the product's partial panels) in a static scratch array
 * of 128 MiB that each GPU the library runs on holds until the process ends; it calls no
 * hipMalloc or hipFree.  Requires a P123 that already has its cell-block form (tt_fill_P123_arrays
 * builds it) unless flags = 2.  Sizes needing more than 128 MiB give TT_ERR_ARG (roughly
 * 48 dim r + 8 items Nalpha 2r bytes); FP32 products or the column operator give
 * TT_ERR_UNSUPPORTED.  All calls on one GPU share that GPU's array: calls that overlap on one
 * GPU (several host threads or contexts) race and must be serialised by the caller.  The
 * context's stream is synchronised before the call returns. */
tt_status tt_apply_rows_noalloc(tt_ctx *ctx, const tt_run_params *rp, const tt_csc *P123,
                                const double *V_unco, const double *V_coup, const tt_swp_states *swp,
                                const tt_pw_states *chn, const double *G, int nE, const int *g_index,
                                int r, const double *X, double *Y, int flags);

#ifdef __cplusplus
}
#endif
#endif /* TICTAC_H */
