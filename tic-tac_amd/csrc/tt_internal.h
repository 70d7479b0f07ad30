// Internal declarations of the tictac library (host C++ and HIP).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <complex>
#include <cstdint>
#include <exception>
#include <mutex>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "../../include/tictac.h"

namespace tt {

// Physical constants, CPP/constants.h:19-35.
constexpr double kMp = 938.272;
constexpr double kMn = 939.565;
constexpr double kMN = 0.5 * (kMp + kMn);
constexpr double kMu23 = kMp * kMn / (kMn + kMp);
constexpr double kHbarc = 197.327;
constexpr int kDeuteronL = 0, kDeuteronS = 1, kDeuteronJ = 1, kDeuteronT = 0;

using cdouble = std::complex<double>;

// ------------------------------------------------------------------ errors
void set_error(const std::string &msg);
struct Error {
    tt_status code;
    std::string msg;
};
[[noreturn]] void fail(tt_status code, const std::string &msg);

#define TT_HIP(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess)                                                         \
            ::tt::fail(TT_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define TT_REQUIRE(cond, msg)                                  \
    do {                                                       \
        if (!(cond)) ::tt::fail(TT_ERR_ARG, std::string(msg)); \
    } while (0)

// ------------------------------------------------------------ host threads
// Worker threads for host-side tables: TT_HOST_THREADS if set (>= 1), else the hardware
// threads, at most 16.
int host_threads();

// f(r) for every r in [0, n), rows handed out one at a time to up to host_threads() threads
// (the caller's among them).  The first exception a row throws stops the hand-out and is
// rethrown on the caller once every worker has returned.
template <class F>
void parallel_rows(int64_t n, F &&f) {
    const int nt = (int)std::min<int64_t>(host_threads(), n);
    if (nt <= 1) {
        for (int64_t r = 0; r < n; ++r) f(r);
        return;
    }
    std::atomic<int64_t> next{0};
    std::atomic<bool> stop{false};
    std::exception_ptr err;
    std::mutex m;
    auto work = [&]() {
        try {
            while (!stop.load(std::memory_order_relaxed)) {
                const int64_t r = next.fetch_add(1);
                if (r >= n) return;
                f(r);
            }
        } catch (...) {
            std::lock_guard<std::mutex> g(m);
            if (!err) err = std::current_exception();
            stop = true;
        }
    };
    std::vector<std::thread> th;
    th.reserve(nt - 1);
    for (int i = 1; i < nt; ++i) {
        try {
            th.emplace_back(work);
        } catch (const std::system_error &) {
            break;   // fewer workers; the rows still all run
        }
    }
    work();
    for (auto &t : th) t.join();
    if (err) std::rethrow_exception(err);
}

// -------------------------------------------------------------- kinematics
double mu1(double Eb);
double com_momentum_to_lab_energy(double q_com, double Eb);
double lab_energy_to_com_momentum(double T_lab, double Eb);
double com_energy_to_com_q_momentum(double E_com);
double com_q_momentum_to_com_energy(double q_com);
void gauss_legendre(int n, double *x, double *w);                 // on [-1, 1], ascending
void gauss_legendre_ab(int n, double a, double b, double *x, double *w);

// ------------------------------------------------------------ 2N channels
int unique_2N_idx(int L, int S, int J, int T, bool coupled, const tt_run_params &rp);
void num_2N_states(const tt_run_params &rp, int &nu, int &nc);
// coupled_via_L_2N / coupled_via_T_3N of a (row, column) pair, CPP/solve_faddeev.cpp:1703-1713
void coupling_kind(int Lr, int Sr, int Jr, int Lc, const tt_run_params &rp, bool &via_L, bool &via_T);

// Block table of a 2N-channel operator embedded in the (alpha, p) space of one
// 3N channel: for every coupled pair (alpha_r, alpha_c) the Np x Np block
// M_2N[h(alpha_r) Np + p_r, h(alpha_c) Np + p_c] (see the pointer tables of
// CPP/solve_faddeev.cpp:1588-1927).  Dense (Nalpha*Np)^2 layout, row-major.
std::vector<double> embed_2N(const tt_pw_states &chn, int Np, const double *M_unco,
                             const double *M_coup, const tt_run_params &rp);
// The same blocks (all of them, zero or not) as positions in the table [M_unco | M_coup]
// (unco: n_unco Np x Np matrices, then the 2Np x 2Np coupled ones): block e of alpha a is
// element (p, pp) = table[src[e] + pr * ld[e] + pc], (pr, pc) = (pp, p) when transposed, else
// (p, pp).  For forming the blocks on the device from the 2N tables.
void embed_2N_layout(const tt_pw_states &chn, int Np, int n_unco, const tt_run_params &rp, bool transpose,
                     std::vector<int> &ptr, std::vector<int> &alpha, std::vector<int64_t> &src,
                     std::vector<int> &ld);
void embed_2N_blocks(const tt_pw_states &chn, int Np, const double *M_unco, const double *M_coup,
                     const tt_run_params &rp, bool transpose, std::vector<int> &ptr, std::vector<int> &alpha,
                     std::vector<double> &blocks, bool keep_zero = false);

// Per-alpha offset of the SWP energy edges used by the resolvent (CPP/make_resolvent.cpp:118-156):
// returns for each alpha a pointer offset into either e_SWP_unco (flag 0) or e_SWP_coup (flag 1).
void swp_edge_offsets(const tt_pw_states &chn, const tt_swp_states &swp, const tt_run_params &rp,
                      std::vector<int> &which, std::vector<int> &offset);

// --------------------------------------------------------- angular tables
double wigner3j(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3);
double wigner6j(int a, int b, int c, int d, int e, int f);
double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i);
double clebsch_gordan(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3);
double atilde(const tt_pw_states &chn, int a, int ap, int Ltot, int two_J, int two_T);

// Host description of the angular part of P123 for one channel:
//   G~_{a'a}(phi, x) = sum_e coef_e * F_{tuple_e}(phi, x)
//   F_t = sum_terms c * Y1[i1](x) * Y2[i2](cos th1) * Y3[i3](cos th2)
struct P123Angular {
    int lmax1 = 0, lmax2 = 0, lmax3 = 0;           // l3 (bra), L12 (ket), l3 (ket)
    int n_tuples = 0;
    std::vector<int> term_ptr;                      // n_tuples + 1
    std::vector<double> term_coef;                  // includes 8 pi^2 sqrt((2L'+1)/4pi), CGs, signs
    std::vector<int> term_i1, term_i2, term_i3;     // indices l(l+1)/2 + |m| into the tables
    std::vector<int> pair_ptr;                      // Nalpha^2 + 1, pair = a'*Nalpha + a
    std::vector<int> pair_tuple;                    // tuple index of each entry
    std::vector<double> pair_atilde;                // A~(a', a, Ltot) of each entry
};
P123Angular build_p123_angular(const tt_pw_states &chn);

// ------------------------------------------------------------ library ctx
// P123 on the device: the reference's CSC (the boundary format) and the cell-block form
// the A-application uses.  Every nonzero of P123 lies on an overlapping bin cell
// (bra (q',p'), ket (q,p)); on each cell the values form a dense Nalpha x Nalpha block
// blk[cell][alpha'][alpha] (structural zeros stored).  Cells are ordered ket-major,
// bras ascending: cells of ket k are ket_ptr[k] .. ket_ptr[k+1]-1.
// Zeroed elements allocated past the last cell of a block array: the MFMA product loads
// 16 * ceil(Na / 16) values from the start of a block row and so runs up to 15 past the end.
constexpr int64_t BLK_SLACK = 64;

struct DeviceCSC {
    // distinct for every DeviceCSC the process creates: cached solve plans key on it, so a
    // new matrix at a freed one's address never matches a stale plan
    uint64_t gen = next_gen();
    int64_t dim = 0, nnz = 0;
    double *val = nullptr;
    float *val32 = nullptr;          // lazily made FP32 copy for the mixed path
    int32_t *row = nullptr;
    uint64_t *colptr = nullptr;
    // cell-block form (Na == 0 until built)
    int Na = 0, NqNp = 0;
    int64_t ncell = 0;
    int max_ket_cells = 0;           // most cells of any one ket
    int *ket_ptr = nullptr;          // NqNp + 1
    int *cell_bra = nullptr;         // ncell
    double *blk = nullptr;           // ncell * Na * Na + BLK_SLACK (zeroed slack)
    float *blk32 = nullptr;          // lazily made FP32 copy, same length
    // transposed cell-block form for products with P123 itself (instead of its transpose):
    // cells grouped by bra (bras ascending, kets ascending inside), partner = ket, blocks
    // tblk[cell][alpha][alpha'] = blk[same cell][alpha'][alpha]; built on first use
    int *t_ptr = nullptr;            // NqNp + 1
    int *t_partner = nullptr;        // ncell
    double *t_blk = nullptr;         // ncell * Na * Na + BLK_SLACK
    float *t_blk32 = nullptr;
    // split-half copies of the blocks for the FP16 product (made on first use, per orientation):
    // each block column alpha as chunks of 8 consecutive input alphas (padded with zeros to
    // Kp = 8 * ceil(Na / 8)), every chunk its 8 high halves then its 8 low halves, of the block
    // values times blkh_scale (a power of two putting the largest magnitude at or below 2^14)
    uint16_t *blkh = nullptr, *t_blkh = nullptr;   // ncell * Na * (Kp / 8) * 16
    double blkh_scale = 0.0, t_blkh_scale = 0.0;
    ~DeviceCSC();
    static uint64_t next_gen();
};
// Builds the cell-block form of a CSC P123 (host pass over a device->host copy).
void build_cell_blocks(DeviceCSC &P, int Na, int Nq, int Np, hipStream_t s);
// Builds the transposed cell-block form from the cell-block form (device transpose).
void build_transposed_cells(DeviceCSC &P, hipStream_t s);

// SWP bin edges from the (reordered) eigenvalues of one 2N channel, CPP/make_swp_states.cpp:146-183.
void swp_edges(const std::vector<double> &ev, int Np, bool coupled, bool chn_3S1, double *e);

// First-order change of the SWP states under a change dV of the potential matrices (same
// layouts as V, C and e of tt_swp_states), with H = H0 + V diagonalised as in tt_make_swp_states.
struct SwpDerivative {
    std::vector<double> dC_unco, dC_coup, de_unco, de_coup;
    double dE_bound = 0.0;
};
void swp_derivative(const tt_run_params &rp, const tt_fwp_states &fwp, const tt_pw_states &pw_full,
                    const double *V_unco, const double *V_coup, const tt_swp_states &swp, const double *dV_unco,
                    const double *dV_coup, SwpDerivative &d);

// Solve plans cached per context (operator set-up, Krylov workspaces, captured graphs).
struct PlanCache;
void free_plan_cache(PlanCache *c);
// Device tables of grouped launches (one launch per stage for all channels of a lockstep group).
struct GroupTables;
void free_group_tables(GroupTables *g);

}  // namespace tt

#define TT_GUARD_BEGIN try {
#define TT_GUARD_END                                  \
    }                                                 \
    catch (const Error &e) {                          \
        set_error(e.msg);                             \
        return e.code;                                \
    }                                                 \
    catch (const std::bad_alloc &) {                  \
        set_error("host allocation failed");          \
        return TT_ERR_ALLOC;                          \
    }                                                 \
    return TT_OK;

namespace tt {
// LU with partial pivoting of the column-major n x n complex matrix M (device, leading dimension
// n; overwritten by the factors) and B <- M^-1 B for the column-major n x nrhs B (device, leading
// dimension n), on the stream.  Returns false when a pivot column is zero.  (k_dense.hip)
bool dense_lu_solve(hipStream_t s, double2 *M, int64_t n, double2 *B, int nrhs);
}  // namespace tt

struct tt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    tt::PlanCache *plans = nullptr;
    // extra streams (default priority) over which a lockstep group's channels are spread, so
    // that independent channels' kernels overlap; created on first use
    std::vector<hipStream_t> lanes;
    tt::GroupTables *groups = nullptr;   // created on first grouped solve
};
