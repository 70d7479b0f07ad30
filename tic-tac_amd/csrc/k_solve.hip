// gfx950 kernels and drivers of the solve path:
//   resolvent G                 CPP/make_resolvent.cpp:4-211
//   A = C^T (2 P123) V C        CPP/solve_faddeev.cpp:4-269 (applied matrix-free, row form)
//   Neumann rows + Pade         CPP/solve_faddeev.cpp:271-296,440-1444
//   dense solve                 CPP/solve_faddeev.cpp:305-438
//   converged GMRES / mixed     no reference counterpart (the reference's Pade stops at ~1e-6)
//
// Vectors of r rows are stored RHS-fastest: X[idx][k], idx = alpha*Nq*Np + q*Np + p,
// complex interleaved.  The reference's rows x A (cblas_dgemm, lines 1072-1088) becomes
//   S = C_op (X .* G)      block op over the 2N-coupled alphas, q-diagonal
//   U = 2 P123^T S         gather SpMM over the reference's own CSC arrays (CSC of P = CSR of P^T)
//   Y = VC_op^T U          block op
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <type_traits>

#include "tt_internal.h"

namespace tt {

// ------------------------------------------------------------------ device memory
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    explicit DevBuf(size_t count) { alloc(count); }
    void alloc(size_t count) {
        release();
        n = count;
        if (count) TT_HIP(hipMalloc(&p, sizeof(T) * count));
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    void upload(const T *h, size_t count, hipStream_t s) {
        if (count > n) alloc(count);
        if (count) TT_HIP(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, s));
    }
};

uint64_t DeviceCSC::next_gen() {
    static std::atomic<uint64_t> counter{1};
    return counter.fetch_add(1);
}

DeviceCSC::~DeviceCSC() {
    if (val) (void)hipFree(val);
    if (val32) (void)hipFree(val32);
    if (row) (void)hipFree(row);
    if (colptr) (void)hipFree(colptr);
    if (ket_ptr) (void)hipFree(ket_ptr);
    if (cell_bra) (void)hipFree(cell_bra);
    if (blk) (void)hipFree(blk);
    if (blk32) (void)hipFree(blk32);
    if (t_ptr) (void)hipFree(t_ptr);
    if (t_partner) (void)hipFree(t_partner);
    if (t_blk) (void)hipFree(t_blk);
    if (t_blk32) (void)hipFree(t_blk32);
    if (blkh) (void)hipFree(blkh);
    if (t_blkh) (void)hipFree(t_blkh);
}

// tblk[c'][i][j] = blk[src[c']][j][i]
__global__ void k_transpose_cells(const int *__restrict__ src, const double *__restrict__ blk,
                                  double *__restrict__ tblk, int64_t ncell, int Na) {
    const int64_t NaNa = (int64_t)Na * Na;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * NaNa) return;
    const int64_t c = t / NaNa;
    const int e = (int)(t - c * NaNa), i = e / Na, j = e - i * Na;
    tblk[t] = blk[(int64_t)src[c] * NaNa + (int64_t)j * Na + i];
}

void build_transposed_cells(DeviceCSC &P, hipStream_t s) {
    TT_REQUIRE(P.Na > 0 && P.ket_ptr && P.cell_bra && P.blk, "transposed cells need the cell-block form");
    if (P.t_blk) return;
    const int NqNp = P.NqNp;
    const int64_t ncell = P.ncell;
    std::vector<int> kp(NqNp + 1), bra(std::max<int64_t>(ncell, 1));
    TT_HIP(hipMemcpyAsync(kp.data(), P.ket_ptr, sizeof(int) * (NqNp + 1), hipMemcpyDeviceToHost, s));
    if (ncell) TT_HIP(hipMemcpyAsync(bra.data(), P.cell_bra, sizeof(int) * ncell, hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    // counting sort of the cells by bra; kets come out ascending inside each bra
    std::vector<int> tp(NqNp + 1, 0), partner(std::max<int64_t>(ncell, 1)), src(std::max<int64_t>(ncell, 1));
    for (int64_t c = 0; c < ncell; ++c) ++tp[bra[c] + 1];
    for (int b = 0; b < NqNp; ++b) tp[b + 1] += tp[b];
    std::vector<int> fill(tp.begin(), tp.end() - 1);
    for (int k = 0; k < NqNp; ++k)
        for (int c = kp[k]; c < kp[k + 1]; ++c) {
            const int pos = fill[bra[c]]++;
            partner[pos] = k;
            src[pos] = c;
        }
    const int64_t nval = ncell * P.Na * P.Na;
    TT_HIP(hipMalloc(&P.t_ptr, sizeof(int) * (NqNp + 1)));
    TT_HIP(hipMalloc(&P.t_partner, sizeof(int) * std::max<int64_t>(ncell, 1)));
    TT_HIP(hipMalloc(&P.t_blk, sizeof(double) * (nval + BLK_SLACK)));
    TT_HIP(hipMemcpyAsync(P.t_ptr, tp.data(), sizeof(int) * (NqNp + 1), hipMemcpyHostToDevice, s));
    TT_HIP(hipMemcpyAsync(P.t_partner, partner.data(), sizeof(int) * partner.size(), hipMemcpyHostToDevice, s));
    TT_HIP(hipMemsetAsync(P.t_blk + nval, 0, sizeof(double) * BLK_SLACK, s));
    int *d_src = nullptr;
    TT_HIP(hipMalloc(&d_src, sizeof(int) * src.size()));
    TT_HIP(hipMemcpyAsync(d_src, src.data(), sizeof(int) * src.size(), hipMemcpyHostToDevice, s));
    if (nval)
        hipLaunchKernelGGL(k_transpose_cells, dim3((unsigned)((nval + 255) / 256)), dim3(256), 0, s, d_src, P.blk,
                           P.t_blk, ncell, P.Na);
    TT_HIP(hipStreamSynchronize(s));
    (void)hipFree(d_src);
}

// CSC -> cell blocks: column c = (alpha, ket), row = (alpha', bra); every (bra, ket) pair
// that holds a nonzero becomes a cell.  One host pass over a device->host copy.
void build_cell_blocks(DeviceCSC &P, int Na, int Nq, int Np, hipStream_t s) {
    const int NqNp = Nq * Np;
    TT_REQUIRE((int64_t)Na * NqNp == P.dim, "P123 dimension does not match Nalpha*Nq*Np");
    std::vector<double> val(P.nnz);
    std::vector<int32_t> row(P.nnz);
    std::vector<uint64_t> ptr(P.dim + 1);
    if (P.nnz) {
        TT_HIP(hipMemcpyAsync(val.data(), P.val, sizeof(double) * P.nnz, hipMemcpyDeviceToHost, s));
        TT_HIP(hipMemcpyAsync(row.data(), P.row, sizeof(int32_t) * P.nnz, hipMemcpyDeviceToHost, s));
    }
    TT_HIP(hipMemcpyAsync(ptr.data(), P.colptr, sizeof(uint64_t) * (P.dim + 1), hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    std::vector<int> ket_ptr(NqNp + 1, 0), cell_bra;
    std::vector<int> mark(NqNp, -1);
    for (int ket = 0; ket < NqNp; ++ket) {
        std::vector<int> bras;
        for (int a = 0; a < Na; ++a) {
            const int64_t c = (int64_t)a * NqNp + ket;
            for (uint64_t nz = ptr[c]; nz < ptr[c + 1]; ++nz) {
                const int bra = row[nz] % NqNp;
                if (mark[bra] != ket) {
                    mark[bra] = ket;
                    bras.push_back(bra);
                }
            }
        }
        std::sort(bras.begin(), bras.end());
        cell_bra.insert(cell_bra.end(), bras.begin(), bras.end());
        ket_ptr[ket + 1] = (int)cell_bra.size();
    }
    const int64_t ncell = (int64_t)cell_bra.size();
    std::vector<double> blk((size_t)ncell * Na * Na, 0.0);
    std::vector<int> pos(NqNp, -1);
    for (int ket = 0; ket < NqNp; ++ket) {
        for (int c = ket_ptr[ket]; c < ket_ptr[ket + 1]; ++c) pos[cell_bra[c]] = c;
        for (int a = 0; a < Na; ++a) {
            const int64_t col = (int64_t)a * NqNp + ket;
            for (uint64_t nz = ptr[col]; nz < ptr[col + 1]; ++nz) {
                const int ap = row[nz] / NqNp, bra = row[nz] % NqNp;
                blk[((size_t)pos[bra] * Na + ap) * Na + a] = val[nz];
            }
        }
    }
    TT_HIP(hipMalloc(&P.ket_ptr, sizeof(int) * (NqNp + 1)));
    TT_HIP(hipMalloc(&P.cell_bra, sizeof(int) * std::max<int64_t>(ncell, 1)));
    blk.resize(blk.size() + BLK_SLACK, 0.0);
    TT_HIP(hipMalloc(&P.blk, sizeof(double) * blk.size()));
    TT_HIP(hipMemcpyAsync(P.ket_ptr, ket_ptr.data(), sizeof(int) * (NqNp + 1), hipMemcpyHostToDevice, s));
    if (ncell) TT_HIP(hipMemcpyAsync(P.cell_bra, cell_bra.data(), sizeof(int) * ncell, hipMemcpyHostToDevice, s));
    TT_HIP(hipMemcpyAsync(P.blk, blk.data(), sizeof(double) * blk.size(), hipMemcpyHostToDevice, s));
    TT_HIP(hipStreamSynchronize(s));
    P.Na = Na;
    P.NqNp = NqNp;
    P.ncell = ncell;
    P.max_ket_cells = 0;
    for (int k = 0; k < NqNp; ++k) P.max_ket_cells = std::max(P.max_ket_cells, ket_ptr[k + 1] - ket_ptr[k]);
}

// ------------------------------------------------------------------ resolvent

// One thread per (energy j, alpha, q, p); formulas of CPP/make_resolvent.cpp:14-80.
// Forward-mode dual number (value, derivative) for the derivative of the closed forms.
struct Dual {
    double v, d;
};
__host__ __device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__host__ __device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__host__ __device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__host__ __device__ inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__host__ __device__ inline double value(double a) { return a; }
__host__ __device__ inline double value(Dual a) { return a.v; }
__host__ __device__ inline double logabs(double a) { return log(fabs(a)); }
__host__ __device__ inline Dual logabs(Dual a) { return {log(fabs(a.v)), a.d / a.v}; }
template <typename T>
__host__ __device__ inline T cst(double c) {
    return T(c);
}
template <>
__host__ __device__ inline Dual cst<Dual>(double c) {
    return {c, 0.0};
}
// x H(x) (derivative H(x))
__host__ __device__ inline double xheav(double a) { return a < 0 ? 0.0 : a; }
__host__ __device__ inline Dual xheav(Dual a) { return a.v < 0 ? Dual{0.0, 0.0} : a; }

// One element of the channel resolvent at total energy E for the SWP bin [el, eu] and q bin
// [ql, qu] (CPP/make_resolvent.cpp:14-80): bound-continuum when el < 0 (Eb = el, which enters
// mu1), continuum-continuum otherwise (Eb = 0).
template <typename T>
__host__ __device__ inline void resolvent_element(T E, T el, T eu, double ql, double qu, T &re, T &im) {
    const T one = cst<T>(1.0);
    const T Mn = cst<T>(939.565), Mp = cst<T>(938.272);
    if (value(el) < 0) {
        const T Eb = el;
        const T m1 = Mn * (Mn + Mp + Eb) / (Mn + Mn + Mp + Eb);
        const T Eql = cst<T>(ql * ql) / (cst<T>(2.0) * m1), Equ = cst<T>(qu * qu) / (cst<T>(2.0) * m1);
        const T Dq = Equ - Eql;
        re = logabs((Eql + Eb - E) / (Equ + Eb - E)) / Dq;
        const double hh = ((value(Equ + Eb - E) < 0 ? 0.0 : 1.0) - (value(Eql + Eb - E) < 0 ? 0.0 : 1.0)) * (-M_PI);
        im = cst<T>(hh) / Dq;
    } else {
        const T m1 = Mn * (Mn + Mp) / (Mn + Mn + Mp);
        const T Eql = cst<T>(ql * ql) / (cst<T>(2.0) * m1), Equ = cst<T>(qu * qu) / (cst<T>(2.0) * m1);
        const T half = cst<T>(0.5);
        const T Ep = half * (el + eu), Eq = half * (Eql + Equ);
        const T Dq = Equ - Eql, Dp = eu - el;
        const T DM = half * (Dp - Dq), DP = half * (Dp + Dq);
        const T D = Ep + Eq - E;
        const T a = D + DM, b = D - DM, c = D + DP, d = D - DP;
        re = (a * logabs(a) + b * logabs(b) - c * logabs(c) - d * logabs(d)) / (Dp * Dq);
        im = cst<T>(M_PI) * (xheav(a) + xheav(b) - xheav(c) - xheav(d)) / (Dp * Dq);
    }
    (void)one;
}

__global__ void k_resolvent(double2 *__restrict__ G, const double *__restrict__ e_unco,
                            const double *__restrict__ e_coup, const int *__restrict__ which,
                            const int *__restrict__ offset, const double *__restrict__ qe,
                            const double *__restrict__ E, int Na, int Nq, int Np, int nE) {
    const int64_t dim = (int64_t)Na * Nq * Np;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dim * nE) return;
    const int j = (int)(t / dim);
    const int64_t idx = t - (int64_t)j * dim;
    const int a = (int)(idx / (Nq * Np));
    const int q = (int)((idx / Np) % Nq);
    const int p = (int)(idx % Np);
    const double *e = (which[a] ? e_coup : e_unco) + offset[a];
    double re, im;
    resolvent_element<double>(E[j], e[p], e[p + 1], qe[q], qe[q + 1], re, im);
    G[t] = make_double2(re, im);
}

// Partial derivatives of every G element with respect to its lower edge, upper edge and
// the energy: Gp[0][j][idx], Gp[1][j][idx], Gp[2][j][idx] (the lower edge of a bound bin is
// E_b, so its partial includes the change of mu1).
__global__ void k_resolvent_partials(double2 *__restrict__ Gp, const double *__restrict__ e_unco,
                                     const double *__restrict__ e_coup, const int *__restrict__ which,
                                     const int *__restrict__ offset, const double *__restrict__ qe,
                                     const double *__restrict__ E, int Na, int Nq, int Np, int nE) {
    const int64_t dim = (int64_t)Na * Nq * Np;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dim * nE) return;
    const int j = (int)(t / dim);
    const int64_t idx = t - (int64_t)j * dim;
    const int a = (int)(idx / (Nq * Np));
    const int q = (int)((idx / Np) % Nq);
    const int p = (int)(idx % Np);
    const double *e = (which[a] ? e_coup : e_unco) + offset[a];
    for (int w = 0; w < 3; ++w) {
        Dual re, im;
        resolvent_element<Dual>(Dual{E[j], w == 2 ? 1.0 : 0.0}, Dual{e[p], w == 0 ? 1.0 : 0.0},
                                Dual{e[p + 1], w == 1 ? 1.0 : 0.0}, qe[q], qe[q + 1], re, im);
        Gp[(int64_t)w * dim * nE + t] = make_double2(re.d, im.d);
    }
}

// dG[j][idx] = dG/del de(a, p) + dG/deu de(a, p+1) + dG/dE dE[j]
__global__ void k_resolvent_change(double2 *__restrict__ dG, const double2 *__restrict__ Gp,
                                   const double *__restrict__ de_unco, const double *__restrict__ de_coup,
                                   const int *__restrict__ which, const int *__restrict__ offset,
                                   const double *__restrict__ dE, int Na, int Nq, int Np, int nE) {
    const int64_t dim = (int64_t)Na * Nq * Np;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dim * nE) return;
    const int j = (int)(t / dim);
    const int64_t idx = t - (int64_t)j * dim;
    const int a = (int)(idx / (Nq * Np));
    const int p = (int)(idx % Np);
    const double *de = (which[a] ? de_coup : de_unco) + offset[a];
    const double2 gl = Gp[t], gu = Gp[dim * nE + t], gE = Gp[2 * dim * nE + t];
    const double dl = de[p], du = de[p + 1], dj = dE[j];
    dG[t] = make_double2(gl.x * dl + gu.x * du + gE.x * dj, gl.y * dl + gu.y * du + gE.y * dj);
}

// ------------------------------------------------------------------ MFMA 16x16x4 (FP64, FP32)
template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
    typedef double acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t fma(double a, double b, acc_t c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    // row of the 16x16 result held in accumulator register r of lane l (column = l % 16)
    static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <>
struct Mfma16<float> {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ acc_t fma(float a, float b, acc_t c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

// ------------------------------------------------------------------ block ops
// Y[(a,q,p)][k] = sum_{e in part(a)} sum_p' B_e[p][p'] * Xg[(alpha_e,q,p')][k]   (or Z - that sum)
// with Xg = X .* G_{gidx[k]} when G != nullptr.  One workgroup per output tile (a, q, k-chunk
// of kc columns): the partner blocks B_e (coalesced) and the partner rows Xg[(alpha_e, q, :)]
// are staged in LDS, then each thread forms outputs (p, k) from LDS only.
// in_cell / out_cell select the cell-major memory order [(q, p)][alpha][k] of the cell-block
// product for the input / output instead of the reference order [(alpha, q, p)][k].  With
// ket_item_ptr the input is the array of work-item partial panels of the cell-block product
// (cell-major per item) and Xg = in_scale * (sum of the ket's item panels, in item order).
// CT: the type the tile is staged and summed in (double; float for the FP32 vectors of the inner
// solves, whose blocks are then FP32 copies: half the LDS, so more workgroups per CU hide the
// op's load latency).  GT: the type of G (double2; float2, an FP32 copy of G, for the FP32
// vectors, whose tile is rounded to FP32 anyway: half the bytes of the op's largest input).
// words over which a block op spreads its max |output| atomics (the split-half product's scale)
constexpr int AMAX_SLOTS = 32;

template <typename CT>
struct Cplx;
template <>
struct Cplx<double> {
    typedef double2 type;
};
template <>
struct Cplx<float> {
    typedef float2 type;
};

// Rows p0 .. p0 + BP - 1 (those below Np) of the block op's output column kk, from the staged
// partner blocks Bs and inputs xs.
template <int BP, typename CT, typename CT2, typename TO>
__device__ __forceinline__ void block_rows(const CT *__restrict__ Bs, const CT2 *__restrict__ xs, int ne, int Np,
                                           int nk, int tile, int kk, int p0, TO *__restrict__ Y,
                                           const TO *__restrict__ Zsub, int out_cell, int a, int q, int Na, int Nq,
                                           int r, int k, float &vmax) {
    const int NpNp = Np * Np;
    CT sr[BP], si[BP];
    int prow[BP];
#pragma unroll
    for (int i = 0; i < BP; ++i) {
        sr[i] = si[i] = 0;
        prow[i] = min(p0 + i, Np - 1) * Np;   // a remainder row past Np is formed, not stored
    }
    for (int e = 0; e < ne; ++e) {
        const CT *B = Bs + e * NpNp;
        const CT2 *xe = xs + e * tile + kk;
#pragma unroll 4
        for (int pp = 0; pp < Np; ++pp) {
            const CT2 x = xe[pp * nk];
#pragma unroll
            for (int i = 0; i < BP; ++i) {
                const CT b = B[prow[i] + pp];
                sr[i] += b * x.x;
                si[i] += b * x.y;
            }
        }
    }
#pragma unroll
    for (int i = 0; i < BP; ++i) {
        const int p = p0 + i;
        if (p >= Np) continue;
        const int dst = (out_cell ? (q * Np + p) * Na + a : (a * Nq + q) * Np + p) * r + k;
        CT vr = sr[i], vi = si[i];
        if (Zsub) {  // Y = Z - (result): the operator y - (y .* G) A of the Krylov solves
            vr = (CT)Zsub[dst].x - vr;
            vi = (CT)Zsub[dst].y - vi;
        }
        Y[dst].x = vr;
        Y[dst].y = vi;
        vmax = fmaxf(vmax, fmaxf(fabsf((float)vr), fabsf((float)vi)));
    }
}

// The tile (bx, by) of the op: bx = a * Nq + q, by = the k-chunk (k_block's grid; a grouped
// launch passes each member's own tile indices).  Element indexes are 32-bit: the launches
// check that dim * r, items * Nalpha * r and n_E * dim stay below 2^31.
template <typename TI, typename TO, typename CT, typename GT>
__device__ __forceinline__ void block_tile(const TI *__restrict__ X, TO *__restrict__ Y,
                                           const int *__restrict__ part_ptr, const int *__restrict__ part_alpha,
                                           const CT *__restrict__ blocks, const GT *__restrict__ G,
                                           const int *__restrict__ gidx, int Na, int Nq, int Np, int r, int kc,
                                           int ne_max, const TO *__restrict__ Zsub, int in_cell, int out_cell,
                                           const int *__restrict__ ket_item_ptr, double in_scale,
                                           double2 *__restrict__ zero_buf, int zero_n, unsigned *__restrict__ amax,
                                           int bx, int by) {
    extern __shared__ double smem[];
    __shared__ int gk[256];   // gidx of this chunk's columns (kc <= 256)
    if (zero_buf && bx == 0 && by == 0)   // a buffer a later kernel accumulates into
        for (int i = threadIdx.x; i < zero_n; i += blockDim.x) zero_buf[i] = make_double2(0.0, 0.0);
    const int dim = Na * Nq * Np;
    const int a = bx / Nq, q = bx % Nq;
    const int k0 = by * kc, nk = min(kc, r - k0);
    const int e0 = part_ptr[a], ne = part_ptr[a + 1] - e0;
    const int NpNp = Np * Np;
    typedef typename Cplx<CT>::type CT2;
    CT *Bs = reinterpret_cast<CT *>(smem);                                    // ne * Np * Np
    CT2 *xs = reinterpret_cast<CT2 *>(Bs + ((ne_max * NpNp + 1) & ~1));       // ne * Np * nk
    // the partners' alphas and the item ranges of this q's kets, read once per workgroup so that
    // every staged element below costs a single round trip to memory
    int *s_alpha = reinterpret_cast<int *>(xs + ne_max * Np * kc);            // ne
    int *s_items = s_alpha + ne_max;                                          // Np + 1
    const CT *Bg = blocks + (size_t)e0 * NpNp;
    {   // the partner blocks: 16-byte loads, four in flight per thread before their LDS stores
        typedef typename std::conditional<sizeof(CT) == 8, double2, float4>::type V16;
        constexpr int per = 16 / (int)sizeof(CT);
        const int n = ne * NpNp;
        const int n16 = ((uintptr_t)Bg % 16 == 0) ? n / per : 0;
        const V16 *Bg16 = reinterpret_cast<const V16 *>(Bg);
        V16 *Bs16 = reinterpret_cast<V16 *>(Bs);
        for (int i0 = threadIdx.x; i0 < n16; i0 += 4 * blockDim.x) {
            V16 t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) t[u] = Bg16[min(i0 + u * (int)blockDim.x, n16 - 1)];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (i0 + u * (int)blockDim.x < n16) Bs16[i0 + u * blockDim.x] = t[u];
        }
        for (int i = n16 * per + threadIdx.x; i < n; i += blockDim.x) Bs[i] = Bg[i];
    }
    if (G)
        for (int i = threadIdx.x; i < nk; i += blockDim.x) gk[i] = gidx[k0 + i];
    for (int i = threadIdx.x; i < ne; i += blockDim.x) s_alpha[i] = part_alpha[e0 + i];
    if (ket_item_ptr)
        for (int i = threadIdx.x; i <= Np; i += blockDim.x) s_items[i] = ket_item_ptr[q * Np + i];
    __syncthreads();
    const int tile = Np * nk, total = ne * tile;
    // EU elements per thread and pass, their loads issued before any is used (the ket's first
    // two item panels, or the input element, and G): one wait for memory per EU elements; a
    // ket's further item panels are added after, in item order.  (EU = 4 was slower: 2858-2933
    // against 2953-2985 channel-solves/s on the batch, its registers cutting the occupancy.)
    constexpr int EU = 2;
    // idx -> (e, pp, kk) by reciprocal multiplication: floor((x + 1/2) / d) in single precision is
    // exact for x < 2^21 (the fractional part stays 1/(2d) away from an integer, far above the
    // rounding), and total = ne * Np * nk is far below that
    const float rtile = 1.0f / (float)tile, rnk = 1.0f / (float)nk;
    for (int base = threadIdx.x; base < total; base += EU * blockDim.x) {
        TI v[EU][2];
        GT gv[EU];
        int i0s[EU], i1s[EU], aes[EU], kks[EU];
#pragma unroll
        for (int u = 0; u < EU; ++u) {
            const int idx = min(base + u * (int)blockDim.x, total - 1);   // past total: loaded, not stored
            const int e = (int)(((float)idx + 0.5f) * rtile), rem = idx - e * tile;
            const int pp = (int)(((float)rem + 0.5f) * rnk), kk = rem - pp * nk, k = k0 + kk;
            const int ae = s_alpha[e];
            const int src = (ae * Nq + q) * Np + pp;  // (alpha, q, p) index
            aes[u] = ae;
            kks[u] = kk;
            if (ket_item_ptr) {
                const int i0 = s_items[pp], i1 = s_items[pp + 1];
                i0s[u] = i0;
                i1s[u] = i1;
#pragma unroll
                for (int w = 0; w < 2; ++w) v[u][w] = X[(max(0, min(i0 + w, i1 - 1)) * Na + ae) * r + k];
            } else {
                const int srcm = in_cell ? (q * Np + pp) * Na + ae : src;  // memory index of X
                v[u][0] = X[srcm * r + k];
            }
            if (G) gv[u] = G[max(0, gk[kk]) * dim + src];
        }
#pragma unroll
        for (int u = 0; u < EU; ++u) {
            const int idx = base + u * (int)blockDim.x;
            if (idx >= total) continue;
            double xr, xi;
            if (ket_item_ptr) {
                const int i0 = i0s[u], i1 = i1s[u];
                xr = 0.0;
                xi = 0.0;
#pragma unroll
                for (int w = 0; w < 2; ++w)
                    if (i0 + w < i1) {
                        xr += (double)v[u][w].x;
                        xi += (double)v[u][w].y;
                    }
                for (int it = i0 + 2; it < i1; ++it) {
                    const TI x = X[(it * Na + aes[u]) * r + k0 + kks[u]];
                    xr += (double)x.x;
                    xi += (double)x.y;
                }
                xr *= in_scale;
                xi *= in_scale;
            } else {
                xr = (double)v[u][0].x;
                xi = (double)v[u][0].y;
            }
            if (G && gk[kks[u]] >= 0) {
                const double2 g = make_double2((double)gv[u].x, (double)gv[u].y);
                const double tr = xr * g.x - xi * g.y;
                xi = xr * g.y + xi * g.x;
                xr = tr;
            }
            xs[idx].x = (CT)xr;
            xs[idx].y = (CT)xi;
        }
    }
    __syncthreads();
    // Thread (g, kk) forms the outputs (p, kk) of the bp consecutive rows p = g bp .. g bp + bp - 1,
    // with bp the fewest rows per thread that cover the Np rows with the block's threads (so no row
    // is formed twice and none past Np, but for the last group's remainder): each staged input
    // x[pp][kk] is read from LDS once per bp outputs, the block values B[p][pp] are the same for the
    // nk threads of a group (LDS broadcast).
    // (with more than 4 npg rows, groups take further runs of 4 rows, npg runs apart)
    const int npg = blockDim.x / nk, bp = min(4, (Np + npg - 1) / npg);
    const int g = threadIdx.x / nk, kk = threadIdx.x - g * nk;
    float vmax = 0.0f;   // largest magnitude this thread writes (NaNs are not counted)
    if (g < npg)
        for (int p0 = g * bp; p0 < Np; p0 += npg * bp) switch (bp) {
                case 1: block_rows<1>(Bs, xs, ne, Np, nk, tile, kk, p0, Y, Zsub, out_cell, a, q, Na, Nq, r, k0 + kk, vmax); break;
                case 2: block_rows<2>(Bs, xs, ne, Np, nk, tile, kk, p0, Y, Zsub, out_cell, a, q, Na, Nq, r, k0 + kk, vmax); break;
                case 3: block_rows<3>(Bs, xs, ne, Np, nk, tile, kk, p0, Y, Zsub, out_cell, a, q, Na, Nq, r, k0 + kk, vmax); break;
                default: block_rows<4>(Bs, xs, ne, Np, nk, tile, kk, p0, Y, Zsub, out_cell, a, q, Na, Nq, r, k0 + kk, vmax); break;
            }
    if (amax) {   // the output's largest magnitude, for the FP16 split of the product that follows
        __shared__ float wmax[4];
        for (int off = 32; off > 0; off >>= 1) vmax = fmaxf(vmax, __shfl_xor(vmax, off));
        if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = vmax;
        __syncthreads();
        // one atomic per workgroup, spread over AMAX_SLOTS words (a single word serialises the
        // launch's atomics); non-negative floats order as their bit patterns do
        if (threadIdx.x == 0)
            atomicMax(amax + (bx + by) % AMAX_SLOTS,
                      __float_as_uint(fmaxf(fmaxf(wmax[0], wmax[1]), fmaxf(wmax[2], wmax[3]))));
    }
}

template <typename TI, typename TO, typename CT = double, typename GT = double2>
__global__ __launch_bounds__(256) void k_block(const TI *__restrict__ X, TO *__restrict__ Y,
                                               const int *__restrict__ part_ptr, const int *__restrict__ part_alpha,
                                               const CT *__restrict__ blocks, const GT *__restrict__ G,
                                               const int *__restrict__ gidx, int Na, int Nq, int Np, int r, int kc,
                                               int ne_max, const TO *__restrict__ Zsub, int in_cell,
                                               int out_cell, const int *__restrict__ ket_item_ptr,
                                               double in_scale, double2 *__restrict__ zero_buf, int zero_n,
                                               unsigned *__restrict__ amax) {
    block_tile<TI, TO, CT, GT>(X, Y, part_ptr, part_alpha, blocks, G, gidx, Na, Nq, Np, r, kc, ne_max, Zsub, in_cell,
                               out_cell, ket_item_ptr, in_scale, zero_buf, zero_n, amax, blockIdx.x, blockIdx.y);
}

// Launch of k_block over all (a, q) and k-chunks of at most 24 columns, of balanced widths
// (every chunk stages all partner blocks, so a narrow remainder chunk costs a full one).
// G's type comes from the template arguments only (G may be passed as nullptr)
template <typename T>
struct NoDeduce {
    typedef T type;
};
// k-chunks of at most 24 columns, of balanced widths: kc columns each, gy chunks; the dynamic
// LDS of a workgroup (partner blocks, staged inputs, partner alphas and item ranges)
struct BlockGeom {
    int kc, gy;
    size_t lds;
};
template <typename CT>
static BlockGeom block_geom(int r, int ne_max, int Np) {
    BlockGeom g;
    const int nch = (r + 23) / 24;
    g.kc = (r + nch - 1) / nch;
    g.gy = (r + g.kc - 1) / g.kc;
    g.lds = (size_t)((ne_max * Np * Np + 1) & ~1) * sizeof(CT) +
            (size_t)ne_max * Np * g.kc * sizeof(typename Cplx<CT>::type) + (size_t)(ne_max + Np + 1) * sizeof(int);
    TT_REQUIRE(g.lds <= 160 * 1024, "block-op tile exceeds LDS (partner blocks too large)");
    return g;
}
template <typename TI, typename TO, typename CT = double, typename GT = double2>
static void launch_block(hipStream_t s, const TI *X, TO *Y, const int *part_ptr, const int *part_alpha,
                         const CT *blocks, int ne_max, const typename NoDeduce<GT>::type *G, const int *gidx, int Na,
                         int Nq, int Np,
                         int r, const TO *Zsub, int in_cell, int out_cell, const int *ket_item_ptr,
                         double in_scale, double2 *zero_buf = nullptr, int zero_n = 0, unsigned *amax = nullptr) {
    TT_REQUIRE((int64_t)Na * Nq * Np * r < INT32_MAX, "block op: dim * r exceeds its 32-bit element indexes");
    const BlockGeom bg = block_geom<CT>(r, ne_max, Np);
    const dim3 grid((unsigned)(Na * Nq), (unsigned)bg.gy), block(256);
    hipLaunchKernelGGL((k_block<TI, TO, CT, GT>), grid, block, bg.lds, s, X, Y, part_ptr, part_alpha, blocks, G, gidx,
                       Na, Nq, Np, r, bg.kc, ne_max, Zsub, in_cell, out_cell, ket_item_ptr, in_scale, zero_buf, zero_n,
                       amax);
}

// ------------------------------------------------------------------ gather SpMM
// U[c][v] = scale * sum_{nz in column c} val[nz] * S[row[nz]][v], v < R2 (real components of
// the r complex right-hand sides).  One wavefront per column of the CSC (= row of P^T).
// Lane = slot * NG + g: the 64/NG slots walk the column's nonzeros with one coalesced
// (val, row) load each, and every lane gathers VPL consecutive values of its row of S;
// UNROLL independent nonzeros per lane are in flight at once.  A cross-lane xor reduction
// over the slots ends the column.
template <typename T, int VPL, int UNROLL>
__global__ __launch_bounds__(256) void k_spmm_gather(const uint64_t *__restrict__ colptr,
                                                     const int32_t *__restrict__ row, const T *__restrict__ val,
                                                     const T *__restrict__ S, T *__restrict__ U, int64_t dim,
                                                     int R2, int NG, T scale) {
    const int lane = threadIdx.x & 63;
    const int64_t col = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (col >= dim) return;  // uniform per wavefront
    const uint64_t beg = colptr[col], end = colptr[col + 1];
    const int slots = 64 / NG;
    const int slot = lane / NG, g = lane % NG;
    const int v0 = g * VPL;
    T acc[VPL];
#pragma unroll
    for (int i = 0; i < VPL; ++i) acc[i] = 0;
    for (uint64_t base = beg + slot; base < end; base += (uint64_t)slots * UNROLL) {
        T pv[UNROLL];
        int pr[UNROLL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const uint64_t nz = base + (uint64_t)u * slots;
            const bool ok = nz < end;
            pv[u] = ok ? val[nz] : (T)0;
            pr[u] = ok ? row[nz] : 0;
        }
        T sv[UNROLL][VPL];
#pragma unroll
        for (int u = 0; u < UNROLL; ++u) {
            const T *src = S + (int64_t)pr[u] * R2 + v0;
#pragma unroll
            for (int i = 0; i < VPL; ++i) sv[u][i] = (v0 + i < R2) ? src[i] : (T)0;
        }
#pragma unroll
        for (int u = 0; u < UNROLL; ++u)
#pragma unroll
            for (int i = 0; i < VPL; ++i) acc[i] += pv[u] * sv[u][i];
    }
    for (int off = NG; off < 64; off <<= 1)
#pragma unroll
        for (int i = 0; i < VPL; ++i) acc[i] += __shfl_xor(acc[i], off);
    if (slot == 0)
#pragma unroll
        for (int i = 0; i < VPL; ++i)
            if (v0 + i < R2) U[col * R2 + v0 + i] = scale * acc[i];
}

// ------------------------------------------------------------------ cell-block product
// U[ket][a][v] = scale * sum_{cells c of ket} sum_{a'} blk[c][a'][a] * S[bra(c)][a'][v]
// (the same product as the gather SpMM: U = 2 P123^T S, in cell-major order).
//
// The blocks of the cells of one ket are contiguous, so for a run of n cells they form one
// (n*Na) x Na row-major matrix Bk with rows kg = c*Na + a'.  The ket's panel is then a single
// long-K product  U_ket (Na x R2) = Bk^T (Na x nNa) . Sk (nNa x R2)  with Sk[kg][v] =
// S[bra(c)][a'][v] gathered row by row: K is concatenated over the cells, so no cell is padded.
// It runs on the 16x16x4 MFMA (FP64 or FP32): M = a in MT tiles of 16, N = v in NT tiles of
// 16, K = kg in steps of 4.  The M tiles interleave the columns: row i of tile m is
// a = MT*i + m, so lane l = 16*kl + il needs the MT consecutive values Bk[kg][MT*il ...] for
// its A operands and loads them with one wide load; the four lane groups read four
// consecutive rows, i.e. one contiguous stretch of memory per k-step.  B[kl][il] =
// Sk[kg][16n + il].  Everything goes straight from global memory into the operand registers
// (no LDS staging: every block value is used exactly once).
//
// Workgroup = one work item, a run of at most CELL_ITEM cells of one ket (long kets are split
// so that no workgroup dominates the launch).  Its K range is cut into batches of 4*UN rows;
// wave w takes batches w, w+4, ..., loading the next batch while the MFMAs of the current one
// run.  The four waves' accumulators are added in LDS in a fixed order, the items of a ket by
// k_cell_reduce, so the result does not depend on timing.  blockIdx.y selects the v-tile.
// The item length and the batch depth UN are chosen per operator and shape by a timing
// autotuner (RowOperator::tune); the defaults below are its starting point (measured at
// Nalpha = 27, R2 = 8 on 22k cells).
constexpr int CELL_ITEM = 16;
// Largest N_alpha of the cell-block product (four 16-row M tiles).  Channels with more partial
// waves (J_2N_max >= 4 gives N_alpha up to 139) take the gather over the CSC instead.
constexpr int CELL_NA_MAX = 64;
template <typename T>
constexpr int cell_un() {
    return sizeof(T) == 8 ? 2 : 4;
}
// candidate batch depths of the autotuner: the deeper batches keep more bytes in flight per
// wave, which pays at large tiles (MT * NT), where the accumulators limit occupancy
template <typename T>
static const std::vector<int> &cell_un_candidates() {
    static const std::vector<int> c = sizeof(T) == 8 ? std::vector<int>{1, 2, 4} : std::vector<int>{2, 4, 8};
    return c;
}

template <typename T, int N>
struct Run {
    T v[N];
};

constexpr int XCD_CHUNK = 32;   // items per chunk of the XCD-aware order
// XCD-aware order of n workgroups: workgroups are dealt to the 8 XCDs round-robin, so
// workgroup b runs on XCD b % 8.  The items are cut into chunks of XCD_CHUNK consecutive items
// (neighbouring kets, which read largely the same bra rows of S), dealt to the XCDs in turn:
// each chunk is read through one XCD's L2, and every XCD gets work from the whole item range,
// so the XCDs finish together.  The last partial round keeps the launch order.  A bijection
// of [0, n).
__device__ __forceinline__ int xcd_order(int b, int n) {
    const int xcd = b % 8, w = b / 8;
    const int full = n / (8 * XCD_CHUNK) * (8 * XCD_CHUNK);
    return b < full ? ((w / XCD_CHUNK) * 8 + xcd) * XCD_CHUNK + w % XCD_CHUNK : b;
}

// Work item `item`, v-tile vt (columns 16 NT vt ...) of the product (k_cellmfma's grid; a
// grouped launch passes each member's own item and tile).
template <typename T, int MT, int NT, int UN>
__device__ __forceinline__ void cell_item(const int2 *__restrict__ items, const int *__restrict__ cell_bra,
                                          const T *__restrict__ B, const T *__restrict__ S, T *__restrict__ Upart,
                                          int Na, int R2, int item, int vt) {
    typedef typename Mfma16<T>::acc_t acc_t;
    // cross-wave sum of the accumulators in one round (FP32) or in two through half the LDS
    // (FP64, whose one-round buffer of 72 KB at MT x NT = 12 allows two workgroups per CU)
    constexpr bool two_rounds = sizeof(T) == 8;
    __shared__ acc_t red[two_rounds ? 2 : 3][MT * NT][64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int kl = lane >> 4, il = lane & 15;
    const int v0 = vt * 16 * NT;
    const int c_beg = items[item].x, ncell = items[item].y - c_beg;
    const int K = ncell * Na;
    const T *Bk = B + (int64_t)c_beg * Na * Na;
    // lane j < ncell (<= CELL_ITEM <= 64) holds the bra of cell j of the run
    const int my_bra = lane < ncell ? cell_bra[c_beg + lane] : 0;
    acc_t acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = acc_t{0, 0, 0, 0};

    T a0[UN][MT], b0[UN][NT], a1[UN][MT], b1[UN][NT];
    // Rows kg = base + 4u + kl of one batch.  Branch-free: every lane loads from a clamped,
    // in-range address (finite data; the block array carries BLK_SLACK zeros past its end for
    // the columns beyond Na of the last row).  Columns a >= Na of the block or v >= R2 of the
    // panel only reach result rows / columns that are never stored, so they need no mask; rows
    // kg >= K are zeroed in the panel operand when the batch is consumed, not when it is
    // loaded, so that the loads of the next batch stay in flight during the MFMAs.  The cell
    // of a row is (kg + 1/2) / Na in single precision, exact for kg < 4096 and Na <= 64 (the
    // fractional part stays 1/(2 Na) away from an integer, far above the rounding).
    const float rNa = 1.0f / (float)Na;
    auto load_batch = [&](int base, T (&a)[UN][MT], T (&b)[UN][NT]) {
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const int kc = min(base + 4 * u + kl, K - 1);
            const int c = (int)(((float)kc + 0.5f) * rNa);
            const int ap = kc - c * Na;
            const int bra = __shfl(my_bra, c);
            const Run<T, MT> ra = *reinterpret_cast<const Run<T, MT> *>(Bk + kc * Na + MT * il);
            const T *srow = S + ((int64_t)bra * Na + ap) * R2;
#pragma unroll
            for (int m = 0; m < MT; ++m) a[u][m] = ra.v[m];
#pragma unroll
            for (int n = 0; n < NT; ++n) b[u][n] = srow[min(v0 + 16 * n + il, R2 - 1)];
        }
    };
    auto mma_batch = [&](int base, const T (&a)[UN][MT], const T (&b)[UN][NT]) {
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const T mk = (base + 4 * u + kl < K) ? (T)1 : (T)0;
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                const T bv = b[u][n] * mk;
#pragma unroll
                for (int m = 0; m < MT; ++m) acc[m][n] = Mfma16<T>::fma(a[u][m], bv, acc[m][n]);
            }
        }
    };
    // Wave w consumes batches w, w+4, ... (nbw of them, rounded up to even); two register
    // buffers alternate (a copy would wait for the loads in flight).  Nothing in the loop is
    // conditional: past the end the loads re-read the clamped last row and the row mask makes
    // the MFMAs of a missing odd batch add zeros.  A load behind a branch would be sunk into
    // it and the compiler's wait counts would assume the worst case at the join.  A scheduling
    // barrier after each prefetch keeps its loads above the MFMAs of the batch in hand.
    const int step = 4 * UN, stride = 4 * step;
    int base = wave * step;
    const int nbw = base < K ? (K - base + stride - 1) / stride : 0;
    load_batch(base, a0, b0);
    for (int i = 0; i < nbw; i += 2) {
        load_batch(base + stride, a1, b1);
        __builtin_amdgcn_sched_barrier(0);
        mma_batch(base, a0, b0);
        load_batch(base + 2 * stride, a0, b0);
        __builtin_amdgcn_sched_barrier(0);
        mma_batch(base + stride, a1, b1);
        base += 2 * stride;
    }
    // the four waves' tiles are added in a fixed order.  One round (FP32): waves 1-3 leave all
    // their tiles at once, wave 0 adds ((w0 + w1) + w2) + w3.  Two rounds (FP64): waves 2 and 3
    // leave their tiles, waves 0 and 1 add them, wave 1 leaves its sum and wave 0 adds
    // (w0 + w2) + (w1 + w3).  Wave 0 stores.
    if constexpr (two_rounds) {
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n)
                if (wave >= 2) red[wave - 2][m * NT + n][lane] = acc[m][n];
        __syncthreads();
        if (wave < 2)
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int n = 0; n < NT; ++n) acc[m][n] += red[wave][m * NT + n][lane];
        __syncthreads();
        if (wave == 1)
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int n = 0; n < NT; ++n) red[0][m * NT + n][lane] = acc[m][n];
        __syncthreads();
    } else {
#pragma unroll
        for (int m = 0; m < MT; ++m)
#pragma unroll
            for (int n = 0; n < NT; ++n)
                if (wave > 0) red[wave - 1][m * NT + n][lane] = acc[m][n];
        __syncthreads();
    }
    if (wave != 0) return;
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int tt = m * NT + n;
            acc_t t = acc[m][n] + red[0][tt][lane];
            if constexpr (!two_rounds) t = (t + red[1][tt][lane]) + red[2][tt][lane];
            const int v = v0 + 16 * n + il;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = MT * Mfma16<T>::row(lane, r) + m;
                if (a < Na && v < R2) Upart[((int64_t)item * Na + a) * R2 + v] = t[r];
            }
        }
}

// FP32 instances of at most 12 tiles with batches of up to 2 steps are held to 128 registers,
// four waves per SIMD (instead of three at 74 + 72 for the 4 x 3 instance): with no spills,
// the Nalpha = 63, R2 = 36 product alone took 102 instead of 115-119 us on one box
template <typename T, int MT, int NT, int UN>
constexpr int cell_waves_per_eu() {
    return sizeof(T) == 4 && UN <= 2 && MT * NT <= 12 ? 4 : 1;
}
template <typename T, int MT, int NT, int UN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(cell_waves_per_eu<T, MT, NT, UN>())))
void k_cellmfma(const int2 *__restrict__ items, const int *__restrict__ cell_bra, const T *__restrict__ B,
                const T *__restrict__ S, T *__restrict__ Upart, int Na, int R2) {
    cell_item<T, MT, NT, UN>(items, cell_bra, B, S, Upart, Na, R2, xcd_order((int)blockIdx.x, (int)gridDim.x),
                             (int)blockIdx.y);
}

// U[ket][a][v] = scale * sum of the partial panels of the ket's work items, in item order
template <typename T>
__global__ void k_cell_reduce(const int *__restrict__ ket_item_ptr, const T *__restrict__ Upart, T *__restrict__ U,
                              int NqNp, int Na, int R2, T scale) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)Na * R2;
    if (t >= (int64_t)NqNp * per) return;
    const int ket = (int)(t / per);
    const int64_t o = t - (int64_t)ket * per;
    T acc = 0;
    for (int i = ket_item_ptr[ket]; i < ket_item_ptr[ket + 1]; ++i) acc += Upart[(int64_t)i * per + o];
    U[t] = scale * acc;
}

// ------------------------------------------------------------------ split-half cell product
// The FP32 product on the FP16 MFMA.  Every operand x (a block value, or an element of S) is
// held as two halves, hi = fp16(x) and lo = fp16(x - hi), after scaling by a power of two that
// puts the operand's largest magnitude at or below 2^14 (so neither half overflows and the low
// halves stay normal for all but tiny elements).  hi + lo carries 22 significant bits, and
//   A B ~ Ahi Bhi + Alo Bhi + Ahi Blo
// (the dropped Alo Blo is below 2^-22 of |A B|) with the products exact and the sums in FP32,
// so the result matches the FP32 product to its own accumulation error.  The 16x16x32 FP16
// MFMA takes 16 cycles for 32 K entries where the FP32 16x16x4 takes 32 cycles for 4: three of
// them per tile cost a fifth of the FP32 product's matrix time, so the product is left bound
// by the block bytes it streams (4 bytes per value, as in FP32).
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
struct __align__(16) HalfPair {   // the split halves of 8 consecutive K entries (32 bytes)
    half8_t hi, lo;
};
constexpr int SPLIT_HEADROOM_LOG2 = 14;   // largest scaled magnitude: 2^14 (fp16 max: 65504)

__device__ __forceinline__ void split8(const float (&x)[8], HalfPair &hp) {
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const _Float16 h = (_Float16)x[j];
        hp.hi[j] = h;
        hp.lo[j] = (_Float16)(x[j] - (float)h);   // x - h is exact in FP32
    }
}

// 2^e with m * 2^e <= 2^SPLIT_HEADROOM_LOG2 (1 for m = 0 or not finite)
__host__ __device__ inline float split_scale(float m) {
    if (!(m > 0.0f) || !(m <= 3.0e38f)) return 1.0f;
    int e = 0;
    (void)frexpf(m, &e);   // m = f 2^e, 0.5 <= f < 1
    return ldexpf(1.0f, SPLIT_HEADROOM_LOG2 - e);
}

__global__ void k_fill_u32(unsigned *__restrict__ p, int n, unsigned v) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// out = bits of max |x| over n values (atomicMax of non-negative floats' bit patterns)
__global__ void k_absmax_f64(const double *__restrict__ x, int64_t n, unsigned *__restrict__ out) {
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        m = fmaxf(m, (float)fabs(x[i]));
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, __float_as_uint(m));
}

// Split-half blocks: out[(c * Na + a) * Kp8 + k8] = halves of scale * blk[c][8 k8 + j][a], j < 8
// (zero past Na).  blk rows are input alphas, columns output alphas.
__global__ void k_blk_split(const double *__restrict__ blk, HalfPair *__restrict__ out, int64_t ncell, int Na,
                            int Kp8, double scale) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * Na * Kp8) return;
    const int k8 = (int)(t % Kp8);
    const int64_t ca = t / Kp8;
    const int a = (int)(ca % Na);
    const int64_t c = ca / Na;
    HalfPair hp;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int ap = 8 * k8 + j;
        const double x = ap < Na ? blk[(c * Na + ap) * Na + a] * scale : 0.0;
        const _Float16 h = (_Float16)(float)x;
        hp.hi[j] = h;
        hp.lo[j] = (_Float16)(float)(x - (double)h);
    }
    out[t] = hp;
}

// Split-half panels of S: St[(bra * R2 + v) * Kp8 + k8] = halves of sc * S[bra][8 k8 + j][v]
// (zero past Na), sc = split_scale(max |S|) from the block op that wrote S; sc is also left in
// *s_scale for the product.  One workgroup per bra bin: its Na x R2 panel goes through LDS
// (rows padded by one word against bank conflicts) and leaves transposed.
__global__ __launch_bounds__(256) void k_split_s(const float *__restrict__ S, HalfPair *__restrict__ St,
                                                 const unsigned *__restrict__ amax, float *__restrict__ s_scale,
                                                 int Na, int R2, int Kp8) {
    extern __shared__ float panel[];   // Na * (R2 + 1)
    const int bra = blockIdx.x, ld = R2 + 1;
    unsigned mbits = 0;
    for (int i = 0; i < AMAX_SLOTS; ++i) mbits = max(mbits, amax[i]);
    const float sc = split_scale(__uint_as_float(mbits));
    if (bra == 0 && threadIdx.x == 0) *s_scale = sc;
    const float *src = S + (int64_t)bra * Na * R2;
    for (int t = threadIdx.x; t < Na * R2; t += blockDim.x) {
        const int ap = t / R2;
        panel[ap * ld + (t - ap * R2)] = src[t];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < R2 * Kp8; t += blockDim.x) {
        const int v = t / Kp8, k8 = t - v * Kp8;
        float x[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int ap = 8 * k8 + j;
            x[j] = ap < Na ? panel[ap * ld + v] * sc : 0.0f;
        }
        HalfPair hp;
        split8(x, hp);
        St[((int64_t)bra * R2 + v) * Kp8 + k8] = hp;
    }
}

// U[ket][a][v] partial panels of the work items, as k_cellmfma, from the split halves:
// Upart = (sum_K A B) / (p_scale * s_scale).  M = a in MT tiles of 16 (a = 16 m + row), N = v in
// NT tiles of 16, K = the item's cells' input alphas, 8-entry chunks (Kp8 per cell) of which
// lane group g = lane / 16 takes chunk 4 s + g of step s (32 entries): its A operand is the 8
// entries of block column a = 16 m + lane % 16, its B operand those of column v of the bra's
// panel, both one 32-byte load (hi, lo) from a layout made for it.  Steps are dealt to the four
// waves in batches of UN; two register buffers alternate, loads of the next batch in flight
// during the MFMAs of the current one, as in k_cellmfma.  Chunks past the item's end are
// loaded from the clamped last chunk and zeroed in the B operand.  The first workgroup clears
// the AMAX_SLOTS words at amax_reset, the block op's maximum that k_split_s has consumed, for
// the next application.
template <int MT, int NT, int UN>
__global__ __launch_bounds__(256) void k_cellh(const int2 *__restrict__ items, const int *__restrict__ cell_bra,
                                               const HalfPair *__restrict__ A, const HalfPair *__restrict__ St,
                                               const float *__restrict__ s_scale, float p_scale,
                                               unsigned *__restrict__ amax_reset, float *__restrict__ Upart, int Na,
                                               int R2, int Kp8) {
    typedef float acc_t __attribute__((ext_vector_type(4)));
    __shared__ acc_t red[3][MT * NT][64];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int g = lane >> 4, il = lane & 15;
    if (amax_reset && blockIdx.x == 0 && blockIdx.y == 0 && tid < AMAX_SLOTS) amax_reset[tid] = 0u;
    const int item = xcd_order((int)blockIdx.x, (int)gridDim.x);   // as k_cellmfma
    const int v0 = blockIdx.y * 16 * NT;
    const int c_beg = items[item].x, ncell = items[item].y - c_beg;
    const int nch = ncell * Kp8;                       // 8-entry chunks of K
    const int nb = ((nch + 3) / 4 + UN - 1) / UN;      // batches of UN steps of 4 chunks
    const int my_bra = lane < ncell ? cell_bra[c_beg + lane] : 0;
    const float rKp8 = 1.0f / (float)Kp8;
    acc_t acc[MT][NT];
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) acc[m][n] = acc_t{0, 0, 0, 0};
    int arow[MT], bcol[NT];   // clamped block columns and panel columns of this lane
#pragma unroll
    for (int m = 0; m < MT; ++m) arow[m] = min(16 * m + il, Na - 1) * Kp8;
#pragma unroll
    for (int n = 0; n < NT; ++n) bcol[n] = min(v0 + 16 * n + il, R2 - 1) * Kp8;

    HalfPair a0[UN][MT], b0[UN][NT], a1[UN][MT], b1[UN][NT];
    auto load_batch = [&](int b, HalfPair (&a)[UN][MT], HalfPair (&bb)[UN][NT]) {
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const int ch = min((b * UN + u) * 4 + g, nch - 1);
            const int c = (int)(((float)ch + 0.5f) * rKp8);   // exact: ch < 4096, Kp8 <= 8
            const int k8 = ch - c * Kp8;
            const int bra = __shfl(my_bra, c);
            const HalfPair *ab = A + (int64_t)(c_beg + c) * Na * Kp8 + k8;
            const HalfPair *sb = St + (int64_t)bra * R2 * Kp8 + k8;
#pragma unroll
            for (int m = 0; m < MT; ++m) a[u][m] = ab[arow[m]];
#pragma unroll
            for (int n = 0; n < NT; ++n) bb[u][n] = sb[bcol[n]];
        }
    };
    auto mma_batch = [&](int b, const HalfPair (&a)[UN][MT], const HalfPair (&bb)[UN][NT]) {
#pragma unroll
        for (int u = 0; u < UN; ++u) {
            const bool ok = (b * UN + u) * 4 + g < nch;
            half8_t bh[NT], bl[NT];
#pragma unroll
            for (int n = 0; n < NT; ++n) {
                bh[n] = ok ? bb[u][n].hi : half8_t{};
                bl[n] = ok ? bb[u][n].lo : half8_t{};
            }
            // the three terms in turn over all tiles: no MFMA waits on the one just issued
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int n = 0; n < NT; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[u][m].hi, bh[n], acc[m][n], 0, 0, 0);
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int n = 0; n < NT; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[u][m].lo, bh[n], acc[m][n], 0, 0, 0);
#pragma unroll
            for (int m = 0; m < MT; ++m)
#pragma unroll
                for (int n = 0; n < NT; ++n)
                    acc[m][n] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[u][m].hi, bl[n], acc[m][n], 0, 0, 0);
        }
    };
    int b = wave;
    const int nbw = b < nb ? (nb - b + 3) / 4 : 0;
    load_batch(b, a0, b0);
    for (int i = 0; i < nbw; i += 2) {
        load_batch(b + 4, a1, b1);
        __builtin_amdgcn_sched_barrier(0);
        mma_batch(b, a0, b0);
        load_batch(b + 8, a0, b0);
        __builtin_amdgcn_sched_barrier(0);
        mma_batch(b + 4, a1, b1);
        b += 8;
    }
    // waves 1-3 leave all their tiles in LDS at once (one barrier), wave 0 adds them in a fixed
    // order ((w0 + w1) + w2) + w3 and stores
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n)
            if (wave > 0) red[wave - 1][m * NT + n][lane] = acc[m][n];
    __syncthreads();
    if (wave != 0) return;
    const float inv = 1.0f / (p_scale * *s_scale);   // powers of two: exact
#pragma unroll
    for (int m = 0; m < MT; ++m)
#pragma unroll
        for (int n = 0; n < NT; ++n) {
            const int t = m * NT + n;
            const acc_t sum = ((acc[m][n] + red[0][t][lane]) + red[1][t][lane]) + red[2][t][lane];
            const int v = v0 + 16 * n + il;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int a = 16 * m + 4 * g + r;
                if (a < Na && v < R2) Upart[((int64_t)item * Na + a) * R2 + v] = sum[r] * inv;
            }
        }
}

// One orientation of the cell-block form: cells grouped by output bin (kets for the product
// with P123^T, bras for the product with P123), the partner bin of each cell and the blocks
// (rows: input alpha, columns: output alpha).
struct CellView {
    int Na = 0, NqNp = 0;
    int64_t ncell = 0;
    const int *grp_ptr = nullptr;
    const int *partner = nullptr;
    const double *blk = nullptr;
    const float *blk32 = nullptr;
};
static CellView cells_of(const DeviceCSC &P, bool transposed) {
    CellView v;
    v.Na = P.Na;
    v.NqNp = P.NqNp;
    v.ncell = P.ncell;
    v.grp_ptr = transposed ? P.t_ptr : P.ket_ptr;
    v.partner = transposed ? P.t_partner : P.cell_bra;
    v.blk = transposed ? P.t_blk : P.blk;
    v.blk32 = transposed ? P.t_blk32 : P.blk32;
    return v;
}

// The device arrays of a set of work items (CellItems owns them; the allocation-free path carves
// them from its scratch).
struct ItemsRef {
    int n = 0;
    const int2 *items = nullptr;
    const int *ket_item_ptr = nullptr;
};

// Work items of the cell-block product: runs of at most item_len (<= 64) cells of one group.
struct CellItems {
    int n = 0;
    DevBuf<int2> items;
    DevBuf<int> ket_item_ptr;
    ItemsRef ref() const { return ItemsRef{n, items.p, ket_item_ptr.p}; }
    // runs of at most item_len cells of each group (host side): it[] and the per-group item
    // pointers kip[].  A group of n cells takes ceil(n / item_len) runs of as equal lengths as
    // possible (lengths differ by at most one), so that no short remainder run stands beside
    // full ones: the launch's workgroups do more equal work and its tail is shorter.
    static void runs(const CellView &P, hipStream_t s, int item_len, std::vector<int2> &it, std::vector<int> &kip) {
        TT_REQUIRE(item_len >= 1 && item_len <= 64, "work item length must be in [1, 64]");
        std::vector<int> kp(P.NqNp + 1);
        TT_HIP(hipMemcpyAsync(kp.data(), P.grp_ptr, sizeof(int) * (P.NqNp + 1), hipMemcpyDeviceToHost, s));
        TT_HIP(hipStreamSynchronize(s));
        it.clear();
        kip.assign(P.NqNp + 1, 0);
        for (int k = 0; k < P.NqNp; ++k) {
            const int n = kp[k + 1] - kp[k], m = (n + item_len - 1) / item_len;
            for (int i = 0, c = kp[k]; i < m; ++i) {
                const int len = n / m + (i < n % m ? 1 : 0);
                it.push_back(make_int2(c, c + len));
                c += len;
            }
            kip[k + 1] = (int)it.size();
        }
    }
    void build(const CellView &P, hipStream_t s, int item_len = CELL_ITEM) {
        std::vector<int2> it;
        std::vector<int> kip;
        runs(P, s, item_len, it, kip);
        n = (int)it.size();
        items.upload(it.data(), std::max<size_t>(it.size(), 1), s);
        ket_item_ptr.upload(kip.data(), kip.size(), s);
    }
    // Items over only the cells whose partner bin (q', p') has q' marked in qmask: runs of
    // consecutive such cells of one group, at most item_len long.  For inputs that are zero
    // outside those q bins (unit rows at on-shell states) the other cells add exact zeros.
    // n = 0 when no cell qualifies.
    void build_on_q(const CellView &P, hipStream_t s, int item_len, int Np, const std::vector<char> &qmask) {
        TT_REQUIRE(item_len >= 1 && item_len <= 64, "work item length must be in [1, 64]");
        std::vector<int> kp(P.NqNp + 1), partner((size_t)std::max<int64_t>(P.ncell, 1));
        TT_HIP(hipMemcpyAsync(kp.data(), P.grp_ptr, sizeof(int) * (P.NqNp + 1), hipMemcpyDeviceToHost, s));
        if (P.ncell)
            TT_HIP(hipMemcpyAsync(partner.data(), P.partner, sizeof(int) * P.ncell, hipMemcpyDeviceToHost, s));
        TT_HIP(hipStreamSynchronize(s));
        auto on = [&](int c) { return qmask[partner[c] / Np] != 0; };
        std::vector<int2> it;
        std::vector<int> kip(P.NqNp + 1, 0);
        for (int k = 0; k < P.NqNp; ++k) {
            int c = kp[k];
            while (c < kp[k + 1]) {
                if (!on(c)) {
                    ++c;
                    continue;
                }
                const int c0 = c;
                while (c < kp[k + 1] && c - c0 < item_len && on(c)) ++c;
                it.push_back(make_int2(c0, c));
            }
            kip[k + 1] = (int)it.size();
        }
        n = (int)it.size();
        items.upload(it.data(), std::max<size_t>(it.size(), 1), s);
        ket_item_ptr.upload(kip.data(), kip.size(), s);
    }
};

// With ev0/ev1 the launch records the kernel's own start and end in them (hipExtLaunchKernel:
// the dispatch's timestamps, which leave out the time the launch waits in its queue behind
// other streams' kernels).
template <typename T, int MT, int NT, int UN>
static void launch_cellmfma_un(hipStream_t s, const CellView *P, const ItemsRef &ci, const T *blk, const T *S,
                               T *Upart, int R2, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
    const dim3 grid((unsigned)ci.n, (unsigned)((R2 + 16 * NT - 1) / (16 * NT))), block(256);
    if (ev0)
        hipExtLaunchKernelGGL((k_cellmfma<T, MT, NT, UN>), grid, block, 0, s, ev0, ev1, 0, ci.items, P->partner, blk,
                              S, Upart, P->Na, R2);
    else
        hipLaunchKernelGGL((k_cellmfma<T, MT, NT, UN>), grid, block, 0, s, ci.items, P->partner, blk, S, Upart, P->Na,
                           R2);
}

// un: batch depth (one of cell_un_candidates<T>(); 0 = the default)
template <typename T, int MT, int NT>
static void launch_cellmfma(hipStream_t s, const CellView *P, const ItemsRef &ci, const T *blk, const T *S,
                            T *Upart, int R2, int un, hipEvent_t ev0, hipEvent_t ev1) {
    if (un == 0) un = cell_un<T>();
    if constexpr (sizeof(T) == 8) {
        if (un == 1)
            launch_cellmfma_un<T, MT, NT, 1>(s, P, ci, blk, S, Upart, R2, ev0, ev1);
        else if (un == 4)
            launch_cellmfma_un<T, MT, NT, 4>(s, P, ci, blk, S, Upart, R2, ev0, ev1);
        else
            launch_cellmfma_un<T, MT, NT, 2>(s, P, ci, blk, S, Upart, R2, ev0, ev1);
    } else {
        if (un == 2)
            launch_cellmfma_un<T, MT, NT, 2>(s, P, ci, blk, S, Upart, R2, ev0, ev1);
        else if (un == 8)
            launch_cellmfma_un<T, MT, NT, 8>(s, P, ci, blk, S, Upart, R2, ev0, ev1);
        else
            launch_cellmfma_un<T, MT, NT, 4>(s, P, ci, blk, S, Upart, R2, ev0, ev1);
    }
}

// N tiles per pass: up to 64 real components (32 complex right-hand sides) stream the blocks
// once; wider batches take ceil(R2 / 64) passes.  Within that number of passes the narrowest
// tile that covers R2 is taken: padded columns cost MFMA issue slots, and at Nalpha ~ 64 with
// 36 components a 64-wide tile makes the FP64 product MFMA-bound (a 48-wide one covers 75%
// useful columns instead of 56%).
static int cell_nt(int R2) {
    const int passes = (R2 + 63) / 64;
    return std::min(4, (R2 + 16 * passes - 1) / (16 * passes));
}

template <typename T, int MT>
static void launch_cellmfma_nt(hipStream_t s, const CellView *P, const ItemsRef &ci, const T *blk, const T *S,
                               T *Upart, int R2, int un, hipEvent_t ev0, hipEvent_t ev1) {
    switch (cell_nt(R2)) {
        case 1:
            launch_cellmfma<T, MT, 1>(s, P, ci, blk, S, Upart, R2, un, ev0, ev1);
            break;
        case 2:
            launch_cellmfma<T, MT, 2>(s, P, ci, blk, S, Upart, R2, un, ev0, ev1);
            break;
        case 3:
            launch_cellmfma<T, MT, 3>(s, P, ci, blk, S, Upart, R2, un, ev0, ev1);
            break;
        default:
            launch_cellmfma<T, MT, 4>(s, P, ci, blk, S, Upart, R2, un, ev0, ev1);
            break;
    }
}

// Upart = the work items' partial panels (items * Nalpha * R2); when U is given also
// U = scale * (sum of each ket's item panels), the full product in cell-major order.
template <typename T>
static void launch_cellmm(hipStream_t s, const CellView *P, const ItemsRef &ci, const T *blk, const T *S, T *U,
                          T *Upart, int R2, T scale, int un = 0, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr) {
    TT_REQUIRE(P->Na >= 1 && P->Na <= CELL_NA_MAX, "cell-block product needs 1 <= Nalpha <= 64 (use the CSC gather)");
    if (ci.n > 0) {
        if (P->Na <= 16)
            launch_cellmfma_nt<T, 1>(s, P, ci, blk, S, Upart, R2, un, ev0, ev1);
        else if (P->Na <= 32)
            launch_cellmfma_nt<T, 2>(s, P, ci, blk, S, Upart, R2, un, ev0, ev1);
        else if (P->Na <= 48)
            launch_cellmfma_nt<T, 3>(s, P, ci, blk, S, Upart, R2, un, ev0, ev1);
        else
            launch_cellmfma_nt<T, 4>(s, P, ci, blk, S, Upart, R2, un, ev0, ev1);
    } else if (ev0) {   // nothing to launch: the interval is empty
        TT_HIP(hipEventRecord(ev0, s));
        TT_HIP(hipEventRecord(ev1, s));
    }
    if (U) {
        const int64_t nout = (int64_t)P->NqNp * P->Na * R2;
        hipLaunchKernelGGL((k_cell_reduce<T>), dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, s,
                           ci.ket_item_ptr, Upart, U, P->NqNp, P->Na, R2, scale);
    }
}

// The split-half product over the split blocks Ah and panels St (k_split_s), un = 1 or 2 steps
// per batch; Upart as launch_cellmm's.
template <int MT, int NT>
static void launch_cellh_un(hipStream_t s, const CellView *P, const ItemsRef &ci, const HalfPair *Ah,
                            const HalfPair *St, const float *s_scale, float p_scale, unsigned *amax_reset,
                            float *Upart, int R2, int un) {
    const int Kp8 = (P->Na + 7) / 8;
    const dim3 grid((unsigned)ci.n, (unsigned)((R2 + 16 * NT - 1) / (16 * NT))), block(256);
    if (un == 2)
        hipLaunchKernelGGL((k_cellh<MT, NT, 2>), grid, block, 0, s, ci.items, P->partner, Ah, St, s_scale, p_scale,
                           amax_reset, Upart, P->Na, R2, Kp8);
    else
        hipLaunchKernelGGL((k_cellh<MT, NT, 1>), grid, block, 0, s, ci.items, P->partner, Ah, St, s_scale, p_scale,
                           amax_reset, Upart, P->Na, R2, Kp8);
}
template <int MT>
static void launch_cellh_nt(hipStream_t s, const CellView *P, const ItemsRef &ci, const HalfPair *Ah,
                            const HalfPair *St, const float *s_scale, float p_scale, unsigned *amax_reset,
                            float *Upart, int R2, int un) {
    switch (cell_nt(R2)) {
        case 1: launch_cellh_un<MT, 1>(s, P, ci, Ah, St, s_scale, p_scale, amax_reset, Upart, R2, un); break;
        case 2: launch_cellh_un<MT, 2>(s, P, ci, Ah, St, s_scale, p_scale, amax_reset, Upart, R2, un); break;
        case 3: launch_cellh_un<MT, 3>(s, P, ci, Ah, St, s_scale, p_scale, amax_reset, Upart, R2, un); break;
        default: launch_cellh_un<MT, 4>(s, P, ci, Ah, St, s_scale, p_scale, amax_reset, Upart, R2, un); break;
    }
}
// S (FP32, cell-major [bra][alpha][v], written by a block op that left max |S| in *amax) ->
// split panels St -> the work items' partial panels Upart of 2 P^T S / 2 (the factor 2 is the
// V.C op's, as for launch_cellmm).
static void launch_cellh(hipStream_t s, const CellView *P, const ItemsRef &ci, const HalfPair *Ah, float p_scale,
                         const float *S, HalfPair *St, float *s_scale, unsigned *amax, float *Upart, int R2, int un) {
    TT_REQUIRE(P->Na >= 1 && P->Na <= CELL_NA_MAX, "cell-block product needs 1 <= Nalpha <= 64 (use the CSC gather)");
    const int Kp8 = (P->Na + 7) / 8;
    const size_t lds = sizeof(float) * (size_t)P->Na * (R2 + 1);
    TT_REQUIRE(lds <= 64 * 1024, "split-half panel exceeds LDS");
    hipLaunchKernelGGL(k_split_s, dim3((unsigned)P->NqNp), dim3(256), lds, s, S, St, amax, s_scale, P->Na, R2, Kp8);
    if (ci.n <= 0) {   // nothing reads the maximum: clear it for the next application here
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, amax, AMAX_SLOTS, 0u);
        return;
    }
    if (P->Na <= 16)
        launch_cellh_nt<1>(s, P, ci, Ah, St, s_scale, p_scale, amax, Upart, R2, un);
    else if (P->Na <= 32)
        launch_cellh_nt<2>(s, P, ci, Ah, St, s_scale, p_scale, amax, Upart, R2, un);
    else if (P->Na <= 48)
        launch_cellh_nt<3>(s, P, ci, Ah, St, s_scale, p_scale, amax, Upart, R2, un);
    else
        launch_cellh_nt<4>(s, P, ci, Ah, St, s_scale, p_scale, amax, Upart, R2, un);
}

// Launch shape of the gather SpMM for R2 real components: NG value groups (a power of two)
// of VPL <= 16 values each.
struct SpmmShape {
    int NG, VPL;
};
static SpmmShape spmm_shape(int R2) {
    int NG = 1;
    while ((R2 + NG - 1) / NG > 16) NG <<= 1;
    int need = (R2 + NG - 1) / NG;
    int VPL = need <= 2 ? 2 : need <= 4 ? 4 : need <= 8 ? 8 : need <= 12 ? 12 : 16;
    return {NG, VPL};
}

template <typename T>
static void launch_spmm(hipStream_t s, const DeviceCSC *P, const T *val, const T *S, T *U, int64_t dim, int R2,
                        T scale) {
    const SpmmShape sh = spmm_shape(R2);
    const dim3 grid((unsigned)((dim + 3) / 4)), block(256);
    switch (sh.VPL) {
        case 2:
            hipLaunchKernelGGL((k_spmm_gather<T, 2, 4>), grid, block, 0, s, P->colptr, P->row, val, S, U, dim, R2, sh.NG, scale);
            break;
        case 4:
            hipLaunchKernelGGL((k_spmm_gather<T, 4, 4>), grid, block, 0, s, P->colptr, P->row, val, S, U, dim, R2, sh.NG, scale);
            break;
        case 8:
            hipLaunchKernelGGL((k_spmm_gather<T, 8, 4>), grid, block, 0, s, P->colptr, P->row, val, S, U, dim, R2, sh.NG, scale);
            break;
        case 12:
            hipLaunchKernelGGL((k_spmm_gather<T, 12, 2>), grid, block, 0, s, P->colptr, P->row, val, S, U, dim, R2, sh.NG, scale);
            break;
        default:
            hipLaunchKernelGGL((k_spmm_gather<T, 16, 2>), grid, block, 0, s, P->colptr, P->row, val, S, U, dim, R2, sh.NG, scale);
            break;
    }
}

__global__ void k_to_f32(const double *__restrict__ a, float *__restrict__ b, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = (float)a[t];
}

// ------------------------------------------------------------------ Krylov vector ops
// accum[i][k] += sum over block b's slice of conj(V_i[e]) W[e] in column k (accum zeroed
// beforehand), one workgroup per (slice, vector), the blocks' sums added by device atomics (the
// order of the additions then varies from run to run, at the level of rounding).  Few slices
// (each thread walking several batches of DOT_UN rows) run faster than one batch per thread:
// on MI355X the start and end of a workgroup (its reduction, its atomics) cost more than the
// loads it saves.  A single-launch "last block finishes" scheme would need a
// device-scope fence per block, which writes back the whole L2 of the block's XCD and costs
// several times the dot products themselves; atomics need none.
constexpr int DOT_UN = 8;

// Slice b of nb of vector i (k_dots' grid; a grouped launch passes each member's own).  Threads
// past the last whole column group (blockDim.x % r of them) only take part in the reduction.
template <typename VT>
__device__ __forceinline__ void dots_slice(const VT *__restrict__ V, int64_t vstride, const VT *__restrict__ W,
                                           int64_t n_elem, int r, double2 *__restrict__ accum, int b, int nb, int i) {
    extern __shared__ double2 sh[];
    const VT *Vi = V + (int64_t)i * vstride;
    const int tpr = blockDim.x / r;  // threads per column k
    const int k = threadIdx.x % r;
    const int sub = threadIdx.x / r;
    double sr = 0.0, si = 0.0;
    if (sub < tpr) {
        // DOT_UN independent loads per batch, all issued before use: clamped to the last row
        // and masked, so no load sits behind a branch
        const int64_t rows = n_elem / r, step = (int64_t)nb * tpr;
        for (int64_t e = (int64_t)b * tpr + sub; e < rows; e += DOT_UN * step) {
            VT v[DOT_UN], w[DOT_UN];
#pragma unroll
            for (int u = 0; u < DOT_UN; ++u) {
                const int64_t ee = min(e + u * step, rows - 1);
                v[u] = Vi[ee * r + k];
                w[u] = W[ee * r + k];
            }
            __builtin_amdgcn_sched_barrier(0);  // keep the loads together, ahead of their use
#pragma unroll
            for (int u = 0; u < DOT_UN; ++u) {
                const double mk = e + u * step < rows ? 1.0 : 0.0;
                const double vx = v[u].x, vy = v[u].y, wx = w[u].x, wy = w[u].y;
                sr += mk * (vx * wx + vy * wy);
                si += mk * (vx * wy - vy * wx);
            }
        }
    }
    sh[threadIdx.x] = make_double2(sr, si);
    __syncthreads();
    if (threadIdx.x < r) {
        double ar = 0.0, ai = 0.0;
        for (int s2 = 0; s2 < tpr; ++s2) {
            ar += sh[s2 * r + threadIdx.x].x;
            ai += sh[s2 * r + threadIdx.x].y;
        }
        atomicAdd(&accum[(int64_t)i * r + threadIdx.x].x, ar);
        atomicAdd(&accum[(int64_t)i * r + threadIdx.x].y, ai);
    }
}

template <typename VT>
__global__ void k_dots(const VT *__restrict__ V, int64_t vstride, const VT *__restrict__ W,
                       int64_t n_elem, int r, double2 *__restrict__ accum) {
    dots_slice<VT>(V, vstride, W, n_elem, r, accum, (int)blockIdx.x, (int)gridDim.x, (int)blockIdx.y);
}

// out[i][k] = sum_b partial[i][b][k]; one workgroup per vector i.  G = 256 / r threads (a power
// of two, at most 64, so a group is inside one wavefront) share one output: thread sub of a
// group adds partials sub, sub + G, ... in order and the group combines them by an xor
// butterfly, so the result does not depend on timing.
__global__ __launch_bounds__(256) void k_dots_reduce(const double2 *__restrict__ partial, int nb, int r,
                                                     double2 *__restrict__ out) {
    const int i = blockIdx.x;
    int G = 1;
    while (G < 64 && 2 * G * r <= (int)blockDim.x) G *= 2;
    const int gsub = threadIdx.x % G;
    for (int t0 = 0; t0 < r; t0 += (int)blockDim.x / G) {
        const int kk = t0 + (int)threadIdx.x / G;
        const bool act = kk < r;
        const double2 *pp = partial + (int64_t)i * nb * r + (act ? kk : 0);
        double ar[4] = {0.0, 0.0, 0.0, 0.0}, ai[4] = {0.0, 0.0, 0.0, 0.0};
        if (act) {
            int bb = gsub;
            for (; bb + 3 * G < nb; bb += 4 * G) {
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const double2 v = pp[(int64_t)(bb + u * G) * r];
                    ar[u] += v.x;
                    ai[u] += v.y;
                }
            }
            for (; bb < nb; bb += G) {
                const double2 v = pp[(int64_t)bb * r];
                ar[0] += v.x;
                ai[0] += v.y;
            }
        }
        double xr = (ar[0] + ar[1]) + (ar[2] + ar[3]), xi = (ai[0] + ai[1]) + (ai[2] + ai[3]);
        for (int off = 1; off < G; off <<= 1) {
            xr += __shfl_xor(xr, off);
            xi += __shfl_xor(xi, off);
        }
        if (act && gsub == 0) out[i * r + kk] = make_double2(xr, xi);
    }
}

// W[e] -= sum_i h[i][k] V_i[e]   (sign = -1) or X[e] += sum_i z[i][k] V_i[e]   (sign = +1).
// The vectors are taken in batches of COMBO_UN with all loads issued before use (clamped and
// masked past nvec), so each thread keeps COMBO_UN loads in flight.
constexpr int COMBO_UN = 8;
template <typename TV>
__device__ __forceinline__ double2 combo_element(double2 w, const TV *__restrict__ V, int64_t vstride, int nvec,
                                                 const double2 *__restrict__ h, int64_t t, int k, int r,
                                                 double sign, double &hh) {
    for (int i0 = 0; i0 < nvec; i0 += COMBO_UN) {
        double2 c[COMBO_UN];
        TV v[COMBO_UN];
#pragma unroll
        for (int u = 0; u < COMBO_UN; ++u) {
            const int i = min(i0 + u, nvec - 1);
            c[u] = h[i * r + k];
            v[u] = V[(int64_t)i * vstride + t];
        }
        __builtin_amdgcn_sched_barrier(0);  // keep the loads together, ahead of their use
#pragma unroll
        for (int u = 0; u < COMBO_UN; ++u) {
            const double f = i0 + u < nvec ? sign : 0.0;
            const double vx = v[u].x, vy = v[u].y;
            w.x += f * (c[u].x * vx - c[u].y * vy);
            w.y += f * (c[u].x * vy + c[u].y * vx);
            hh += f * f * (c[u].x * c[u].x + c[u].y * c[u].y);
        }
    }
    return w;
}

template <typename TW, typename TV>
__device__ __forceinline__ void combo_at(TW *__restrict__ W, const TV *__restrict__ V, int64_t vstride, int nvec,
                                         const double2 *__restrict__ h, int64_t n_elem, int r, double sign,
                                         int64_t t) {
    if (t >= n_elem) return;
    double hh = 0.0;
    const TW w0 = W[t];
    const double2 w = combo_element(make_double2(w0.x, w0.y), V, vstride, nvec, h, t, (int)(t % r), r, sign, hh);
    W[t].x = w.x;
    W[t].y = w.y;
}
template <typename TW, typename TV>
__global__ void k_combo(TW *__restrict__ W, const TV *__restrict__ V, int64_t vstride, int nvec,
                        const double2 *__restrict__ h, int64_t n_elem, int r, double sign) {
    combo_at<TW, TV>(W, V, vstride, nvec, h, n_elem, r, sign, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// The last pass of a reorthogonalised Gram-Schmidt step, with the normalisation folded in:
// W[e] <- (W[e] - sum_{i<nvec} h[i][k] V_i[e]) / nu_k.  h = V^H W from the pass before, whose
// entry h[nvec][k] is <W, W>_k (W was its last vector); with the columns of V orthonormal,
// nu_k^2 = <W, W>_k - sum_i |h[i][k]|^2 (Pythagoras), so no pass over the result is needed.
// nu_k^2 goes to n2[k] (written by the first r threads) for the Hessenberg matrix.  After the
// first projection h is of the order of the rounding of that projection, so the difference
// cancels only when W itself has collapsed into span(V), where GMRES has converged.
template <typename VT>
__device__ __forceinline__ void combo_normalize_at(VT *__restrict__ W, const VT *__restrict__ V, int64_t vstride,
                                                   int nvec, const double2 *__restrict__ h, int64_t n_elem, int r,
                                                   double2 *__restrict__ n2, int64_t t) {
    if (t >= n_elem) return;
    const int k = (int)(t % r);
    double hh = 0.0;
    const VT w0 = W[t];
    const double2 w = combo_element(make_double2(w0.x, w0.y), V, vstride, nvec, h, t, k, r, -1.0, hh);
    const double nu2 = fmax(h[nvec * r + k].x - hh, 0.0);
    const double f = nu2 > 0.0 ? 1.0 / sqrt(nu2) : 0.0;
    W[t].x = f * w.x;
    W[t].y = f * w.y;
    if (t < r) n2[t] = make_double2(nu2, 0.0);
}
template <typename VT>
__global__ void k_combo_normalize(VT *__restrict__ W, const VT *__restrict__ V, int64_t vstride,
                                  int nvec, const double2 *__restrict__ h, int64_t n_elem, int r,
                                  double2 *__restrict__ n2) {
    combo_normalize_at<VT>(W, V, vstride, nvec, h, n_elem, r, n2, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_axpby(double2 *__restrict__ Y, const double2 *__restrict__ A, const double2 *__restrict__ B,
                        double alpha, double beta, int64_t n_elem) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_elem) return;
    const double2 a = A[t], b = B[t];
    Y[t] = make_double2(alpha * a.x + beta * b.x, alpha * a.y + beta * b.y);
}

// out[m] = X[src_idx[m]][src_k[m]]
__global__ void k_gather(const double2 *__restrict__ X, const int64_t *__restrict__ src_idx,
                         const int *__restrict__ src_k, double2 *__restrict__ out, int n, int r) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = X[src_idx[t] * r + src_k[t]];
}

// X[dst_idx[m]][m] = 1 (unit rows), rest zero beforehand
__global__ void k_unit_rows(double2 *__restrict__ X, const int64_t *__restrict__ dst_idx, int r) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < r) X[dst_idx[t] * r + t] = make_double2(1.0, 0.0);
}

// Y[e] = X[e] / sqrt(n2[k].x), zero where the squared norm is zero
template <typename TY>
__global__ void k_scale_by_norm(TY *__restrict__ Y, const double2 *__restrict__ X, const double2 *__restrict__ n2,
                                int64_t n_elem, int r) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_elem) return;
    const double nn = n2[t % r].x;
    const double f = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
    const double2 x = X[t];
    Y[t].x = f * x.x;
    Y[t].y = f * x.y;
}

// W[t] /= sqrt(n2[k]) in place (a zero norm leaves zeros)
template <typename VT>
__device__ __forceinline__ void normalize_at(VT *__restrict__ W, const double2 *__restrict__ n2, int64_t n_elem,
                                             int r, int64_t t) {
    if (t >= n_elem) return;
    const double nn = n2[t % r].x;
    const double f = nn > 0.0 ? 1.0 / sqrt(nn) : 0.0;
    const VT w = W[t];
    W[t].x = f * w.x;
    W[t].y = f * w.y;
}
template <typename VT>
__global__ void k_normalize(VT *__restrict__ W, const double2 *__restrict__ n2, int64_t n_elem, int r) {
    normalize_at<VT>(W, n2, n_elem, r, (int64_t)blockIdx.x * blockDim.x + threadIdx.x);
}

// ------------------------------------------------------------------ grouped launches
// One launch of a stage for all members of a lockstep group (the channels of a batch, whose
// shapes differ).  A table in device memory gives each member's arguments and the first
// workgroup of its share of the grid; a workgroup finds its member by binary search over
// those starts and runs the member's tile through the same body as the member's own launch.
constexpr int GROUP_MAX = 64;   // members per grouped launch (larger groups take several)
struct GroupHead {
    int n;                        // members (>= 1)
    int start[GROUP_MAX + 1];     // first workgroup of each member; start[n] = the grid
};
// the member whose share holds workgroup b (0 <= b < start[n]): the last m with start[m] <= b,
// so a member with an empty share is never chosen
__device__ __forceinline__ int group_member(const GroupHead *__restrict__ h, int b) {
    int lo = 0, hi = h->n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (b >= h->start[mid])
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// k_block's arguments; the member's grid is gx * gy tiles (tile b: bx = b % gx, by = b / gx)
struct BlockMember {
    const void *X;
    void *Y;
    const int *part_ptr, *part_alpha;
    const void *blocks, *G;
    const int *gidx;
    const void *Zsub;
    const int *ket_item_ptr;
    double2 *zero_buf;
    unsigned *amax;
    double in_scale;
    int Na, Nq, Np, r, kc, ne_max, in_cell, out_cell, zero_n, gx;
};
template <typename TI, typename TO, typename CT, typename GT>
__global__ __launch_bounds__(256) void k_block_grp(const GroupHead *__restrict__ h,
                                                   const BlockMember *__restrict__ mb) {
    const int m = group_member(h, (int)blockIdx.x);
    const int b = (int)blockIdx.x - h->start[m];
    const BlockMember &d = mb[m];
    block_tile<TI, TO, CT, GT>(static_cast<const TI *>(d.X), static_cast<TO *>(d.Y), d.part_ptr, d.part_alpha,
                               static_cast<const CT *>(d.blocks), static_cast<const GT *>(d.G), d.gidx, d.Na, d.Nq,
                               d.Np, d.r, d.kc, d.ne_max, static_cast<const TO *>(d.Zsub), d.in_cell, d.out_cell,
                               d.ket_item_ptr, d.in_scale, d.zero_buf, d.zero_n, d.amax, b % d.gx, b / d.gx);
}

// k_cellmfma's arguments; the member's grid is nit items times the v-tiles
struct ProdMember {
    const int2 *items;
    const int *cell_bra;
    const void *B, *S;
    void *Upart;
    int Na, R2, nit;
};
// The XCD-aware order runs over the whole grouped grid, so each member's items are dealt to
// all XCDs in chunks of consecutive items, as in its own launch.
template <typename T, int MT, int NT, int UN>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(cell_waves_per_eu<T, MT, NT, UN>())))
void k_cellmfma_grp(const GroupHead *__restrict__ h, const ProdMember *__restrict__ mb) {
    const int g = xcd_order((int)blockIdx.x, (int)gridDim.x);
    const int m = group_member(h, g);
    const int b = g - h->start[m];
    const ProdMember &d = mb[m];
    cell_item<T, MT, NT, UN>(d.items, d.cell_bra, static_cast<const T *>(d.B), static_cast<const T *>(d.S),
                             static_cast<T *>(d.Upart), d.Na, d.R2, b % d.nit, b / d.nit);
}

// The vector kernels' arguments (k_dots: V, W, vstride, n_elem, r, out = accum, nvec, nb; the
// combinations: W, V, vstride, nvec, h, n_elem, r, sign, out = n2; k_normalize: W, h = n2)
struct VecMember {
    const void *V;
    void *W;
    const double2 *h;
    double2 *out;
    int64_t vstride, n_elem;
    double sign;
    int r, nvec, nb;
};
// 256 threads per workgroup: blockDim.x / r column groups, as k_dots' own launch with
// 256 - 256 % r threads
template <typename VT>
__global__ __launch_bounds__(256) void k_dots_grp(const GroupHead *__restrict__ h, const VecMember *__restrict__ mb) {
    const int m = group_member(h, (int)blockIdx.x);
    const int b = (int)blockIdx.x - h->start[m];
    const VecMember &d = mb[m];
    dots_slice<VT>(static_cast<const VT *>(d.V), d.vstride, static_cast<const VT *>(d.W), d.n_elem, d.r, d.out,
                   b % d.nb, d.nb, b / d.nb);
}
template <typename TW, typename TV>
__global__ __launch_bounds__(256) void k_combo_grp(const GroupHead *__restrict__ h, const VecMember *__restrict__ mb) {
    const int m = group_member(h, (int)blockIdx.x);
    const int b = (int)blockIdx.x - h->start[m];
    const VecMember &d = mb[m];
    combo_at<TW, TV>(static_cast<TW *>(d.W), static_cast<const TV *>(d.V), d.vstride, d.nvec, d.h, d.n_elem, d.r,
                     d.sign, (int64_t)b * 256 + threadIdx.x);
}
template <typename VT>
__global__ __launch_bounds__(256) void k_combo_normalize_grp(const GroupHead *__restrict__ h,
                                                             const VecMember *__restrict__ mb) {
    const int m = group_member(h, (int)blockIdx.x);
    const int b = (int)blockIdx.x - h->start[m];
    const VecMember &d = mb[m];
    combo_normalize_at<VT>(static_cast<VT *>(d.W), static_cast<const VT *>(d.V), d.vstride, d.nvec, d.h, d.n_elem,
                           d.r, d.out, (int64_t)b * 256 + threadIdx.x);
}
template <typename VT>
__global__ __launch_bounds__(256) void k_normalize_grp(const GroupHead *__restrict__ h,
                                                       const VecMember *__restrict__ mb) {
    const int m = group_member(h, (int)blockIdx.x);
    const int b = (int)blockIdx.x - h->start[m];
    const VecMember &d = mb[m];
    normalize_at<VT>(static_cast<VT *>(d.W), d.h, d.n_elem, d.r, (int64_t)b * 256 + threadIdx.x);
}

// out[idx][k] = in[idx][k] * G[gk][idx] + (idx == unit[k])   (e + x .* G for the adjoint)
__global__ void k_unit_plus_G(const double2 *__restrict__ in, double2 *__restrict__ out, const double2 *__restrict__ G,
                              const int *__restrict__ gidx, const int64_t *__restrict__ unit, int64_t dim, int r) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dim * r) return;
    const int64_t idx = t / r;
    const int k = (int)(t - idx * r);
    const double2 x = in[t], g = G[(int64_t)gidx[k] * dim + idx];
    double2 y = make_double2(x.x * g.x - x.y * g.y, x.x * g.y + x.y * g.x);
    if (idx == unit[k]) y.x += 1.0;
    out[t] = y;
}

// partial[pair][b] = sum over block b's slice of x of A[x][k] B[x][k'] (W[j][x]), the bilinear
// (unconjugated) products of the adjoint; x runs over the reference index idx = (a, qp), and A
// or B is read in the cell-major order [(qp)][a] when its flag is set.  Summed over b by
// k_dots_reduce (fixed order).
__global__ void k_pair_dots(const double2 *__restrict__ A, int a_cell, const double2 *__restrict__ B, int b_cell,
                            const double2 *__restrict__ W, const int4 *__restrict__ pairs, int Na, int NqNp, int r,
                            double2 *__restrict__ partial) {
    __shared__ double2 sh[256];
    const int pr = blockIdx.y, b = blockIdx.x, nb = gridDim.x;
    const int4 pq = pairs[pr];
    const int64_t dim = (int64_t)Na * NqNp;
    double sr = 0.0, si = 0.0;
    for (int64_t x = (int64_t)b * blockDim.x + threadIdx.x; x < dim; x += (int64_t)nb * blockDim.x) {
        const int a = (int)(x / NqNp), qp = (int)(x - (int64_t)a * NqNp);
        const int64_t xc = (int64_t)qp * Na + a;
        const double2 u = A[(a_cell ? xc : x) * r + pq.x], v = B[(b_cell ? xc : x) * r + pq.y];
        double pr_ = u.x * v.x - u.y * v.y, pi_ = u.x * v.y + u.y * v.x;
        if (W) {
            const double2 w = W[(int64_t)pq.z * dim + x];
            const double tr = pr_ * w.x - pi_ * w.y;
            pi_ = pr_ * w.y + pi_ * w.x;
            pr_ = tr;
        }
        sr += pr_;
        si += pi_;
    }
    sh[threadIdx.x] = make_double2(sr, si);
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) {
            sh[threadIdx.x].x += sh[threadIdx.x + st].x;
            sh[threadIdx.x].y += sh[threadIdx.x + st].y;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)pr * nb + b] = sh[0];
}

static inline unsigned nblk(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

// ------------------------------------------------------------------ the operator
// Matrix-free row-form application Y = (X .* G_sel) A for r rows at once.
// Process-wide table of the cell-product autotuner: the chosen item length and batch depth per
// device, orientation (row / column operator), precision, Nalpha, R2 and cell count, with the
// winning launch time.
struct TuneKey {
    int dev, col, fp32, Na, R2;
    int64_t ncell;
    int hmode;
    bool operator<(const TuneKey &o) const {
        return std::tie(dev, col, fp32, Na, R2, ncell, hmode) <
               std::tie(o.dev, o.col, o.fp32, o.Na, o.R2, o.ncell, o.hmode);
    }
};
struct TuneEntry {
    int len, un, h;
    float ms;
};
// Which FP32 cell product the autotuner may pick: the FP32 MFMA one (default, TT_CELL_H unset
// or 0), only the split-half FP16 one (TT_CELL_H=1), or whichever times faster
// (TT_CELL_H=auto).  On the 18-channel batch the two measure the same (2992-3020 against
// 3023-3025 channel-solves/s, alternating runs on one box), so the default keeps the product
// that needs no split kernel and no split copy of the blocks.
static int cell_h_mode() {
    const char *e = std::getenv("TT_CELL_H");
    if (!e) return 0;
    if (std::strcmp(e, "1") == 0) return 1;
    if (std::strcmp(e, "auto") == 0) return -1;
    return 0;
}
struct TuneTable {
    std::mutex mu;
    std::map<TuneKey, TuneEntry> map;
};
static TuneTable &tune_table() {
    static TuneTable t;
    return t;
}

// Device copies of the resolvent's inputs (SWP edges, q edges, energies, per-alpha edge
// offsets), kept by an operator so that computing its G allocates nothing and needs no wait.
__global__ void k_g_to_f32(const double2 *__restrict__ in, float2 *__restrict__ out, int64_t n) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = make_float2((float)in[t].x, (float)in[t].y);
}

struct ResolventStaging {
    DevBuf<double> eu, ec, qe, E;
    DevBuf<int> which, offset;
    // host copies of what the buffers hold: inputs equal to them are not sent again (the SWP
    // edges and the per-alpha tables stay the same from one solve of a channel to the next)
    std::vector<double> h_eu, h_ec, h_qe;
    std::vector<int> h_which, h_offset;
    template <typename T>
    static void put(DevBuf<T> &d, std::vector<T> &last, const T *h, size_t n, hipStream_t s) {
        if (d.p && d.n >= n && last.size() == n && std::equal(h, h + n, last.begin())) return;
        d.upload(h, n, s);
        last.assign(h, h + n);
    }
};
// d <- h[0..n) unless d already holds exactly that (last: host copy of what d holds)
template <typename T>
static void upload_if_changed(DevBuf<T> &d, std::vector<T> &last, const T *h, size_t n, hipStream_t s) {
    ResolventStaging::put(d, last, h, n, s);
}

// Exact identity of the inputs a plan's block contents are built from (V and the SWP
// eigenvectors): a copy of their words, with a hash that mixes every bit of every word into the
// whole key, so that most mismatches are rejected without comparing the words.  Two contents
// are the same only when their words are equal bit for bit.
struct Content {
    std::vector<double> words;
    uint64_t hash = 0;
};
typedef std::shared_ptr<const Content> ContentRef;

static inline uint64_t mix64(uint64_t z) {   // splitmix64 finaliser (a bijection of 64-bit words)
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

static bool same_content(const ContentRef &a, const ContentRef &b) {
    if (a == b) return true;
    if (!a || !b || a->hash != b->hash || a->words.size() != b->words.size()) return false;
    return a->words.empty() || std::memcmp(a->words.data(), b->words.data(), sizeof(double) * a->words.size()) == 0;
}

// V.C of every 2N channel (dot_MM in create_VC_col_maj_3N_pointer_array,
// CPP/solve_faddeev.cpp:1828-1855): row-major Np x Np (uncoupled) and 2Np x 2Np (coupled).
struct VCProducts {
    ContentRef content;     // the V and C they were made from
    bool valid = false;
    std::vector<double> u, c;
    // on the device, for forming every channel's partner blocks there: [C_unco | C_coup] and
    // [VC_unco | VC_coup] (uploaded once per V and C, shared by all channel plans)
    DevBuf<double> dC, dVC;
    bool dev_valid = false;
    void upload(const tt_swp_states &swp, int nu, int nc, hipStream_t s) {
        const size_t Np = (size_t)swp.Np_WP, nu2 = (size_t)nu * Np * Np, nc2 = (size_t)nc * 4 * Np * Np;
        std::vector<double> t(std::max<size_t>(nu2 + nc2, 1), 0.0);
        std::copy(swp.C_SWP_unco, swp.C_SWP_unco + nu2, t.begin());
        if (nc) std::copy(swp.C_SWP_coup, swp.C_SWP_coup + nc2, t.begin() + nu2);
        dC.upload(t.data(), t.size(), s);
        std::copy(u.begin(), u.begin() + nu2, t.begin());
        if (nc) std::copy(c.begin(), c.begin() + nc2, t.begin() + nu2);
        dVC.upload(t.data(), t.size(), s);
        dev_valid = true;
    }
};

// blocks[e][p][pp] = T[src[e] + pr * ld[e] + pc] with (pr, pc) = (pp, p) transposed, else (p, pp);
// its FP32 copy beside (the partner blocks of a channel from the shared 2N tables)
__global__ void k_embed(double *__restrict__ blocks, float *__restrict__ blocks32, const double *__restrict__ T,
                        const int64_t *__restrict__ src, const int *__restrict__ ld, int nb, int Np, int transpose) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t NpNp = (int64_t)Np * Np;
    if (t >= nb * NpNp) return;
    const int e = (int)(t / NpNp), i = (int)(t - e * NpNp), p = i / Np, pp = i - p * Np;
    const int pr = transpose ? pp : p, pc = transpose ? p : pp;
    const double v = T[src[e] + (int64_t)pr * ld[e] + pc];
    blocks[t] = v;
    blocks32[t] = (float)v;
}
static void vc_products(const tt_run_params &rp, const double *V_unco, const double *V_coup,
                        const tt_swp_states &swp, VCProducts &out) {
    int nu, nc;
    num_2N_states(rp, nu, nc);
    const int Np = swp.Np_WP;
    out.u.assign((size_t)nu * Np * Np, 0.0);
    out.c.assign((size_t)nc * 4 * Np * Np, 0.0);
    auto mm = [](const double *A, const double *B, double *C, int n) {
        for (int i = 0; i < n; ++i)
            for (int k = 0; k < n; ++k) {
                const double a = A[(size_t)i * n + k];
                if (a == 0.0) continue;
                for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * B[(size_t)k * n + j];
            }
    };
    // one 2N channel per row: the products do not depend on the thread count
    parallel_rows(nu + nc, [&](int64_t c) {
        if (c < nu)
            mm(V_unco + (size_t)c * Np * Np, swp.C_SWP_unco + (size_t)c * Np * Np, &out.u[(size_t)c * Np * Np], Np);
        else
            mm(V_coup + (size_t)(c - nu) * 4 * Np * Np, swp.C_SWP_coup + (size_t)(c - nu) * 4 * Np * Np,
               &out.c[(size_t)(c - nu) * 4 * Np * Np], 2 * Np);
    });
    out.valid = true;
}

// ------------------------------------------------------------------ grouped launches (host)
// Stages of a grouped Arnoldi step (or of a grouped operator application: stages 0-2), in the
// order they run; the members of one stage share a kernel instance.
enum GroupPhase { GP_BLOCK_IN = 0, GP_PRODUCT = 1, GP_BLOCK_OUT = 2, GP_DOTS = 3, GP_COMBO = 4, GP_DOTS2 = 5,
                  GP_FINAL = 6 };
struct GroupStage {
    bool f32 = false;
    int key = 0;          // product instance: MT * 100 + NT * 10 + UN
    size_t lds = 0;       // dynamic LDS of the kernel (the largest member's)
    size_t msize = 0;     // bytes of one member record
    std::vector<char> mem;
    std::vector<int> nblk;   // workgroups of each member
    int count() const { return (int)nblk.size(); }
};
// What one grouped call launches: stages by (phase, precision, product instance).
struct GroupPlan {
    std::map<std::tuple<int, int, int>, GroupStage> st;
    template <typename M>
    void add(int phase, bool f32, int key, const M &m, int64_t nblk, size_t lds = 0) {
        TT_REQUIRE(nblk >= 0 && nblk < (1 << 30), "grouped launch: member grid out of range");
        if (nblk == 0) return;   // nothing to run (no work items)
        GroupStage &g = st[std::make_tuple(phase, f32 ? 1 : 0, key)];
        TT_REQUIRE(g.msize == 0 || g.msize == sizeof(M), "grouped launch: mixed member records");
        g.f32 = f32;
        g.key = key;
        g.msize = sizeof(M);
        g.lds = std::max(g.lds, lds);
        const char *b = reinterpret_cast<const char *>(&m);
        g.mem.insert(g.mem.end(), b, b + sizeof(M));
        g.nblk.push_back((int)nblk);
    }
    bool empty() const { return st.empty(); }
};

// Device copies of the tables: a ring of slots, each a pinned host staging area and its
// device copy.  One grouped call fills a slot, sends it with one copy ahead of its launches and
// records the slot's event after them; a slot is refilled only after that event.
struct GroupTables {
    // two grouped calls per lane and step (the step and the one ahead), several steps in flight
    static constexpr int NSLOT = 64;
    size_t cap = 0;
    char *host[NSLOT] = {};
    char *dev[NSLOT] = {};
    hipEvent_t ev[NSLOT] = {};
    bool pending[NSLOT] = {};
    int slot = 0;
    size_t used = 0;
    ~GroupTables() {
        clear_graphs();
        for (int i = 0; i < NSLOT; ++i) {
            if (ev[i]) {
                if (pending[i]) (void)hipEventSynchronize(ev[i]);
                (void)hipEventDestroy(ev[i]);
            }
            if (host[i]) (void)hipHostFree(host[i]);
            if (dev[i]) (void)hipFree(dev[i]);
        }
    }
    void begin(size_t need) {
        slot = (slot + 1) % NSLOT;
        if (pending[slot]) {
            TT_HIP(hipEventSynchronize(ev[slot]));
            pending[slot] = false;
        }
        if (need > cap) {   // every slot is replaced: wait for all of them
            for (int i = 0; i < NSLOT; ++i) {
                if (pending[i]) TT_HIP(hipEventSynchronize(ev[i]));
                pending[i] = false;
                if (host[i]) TT_HIP(hipHostFree(host[i]));
                if (dev[i]) TT_HIP(hipFree(dev[i]));
                host[i] = dev[i] = nullptr;
            }
            cap = std::max<size_t>({need, 2 * cap, (size_t)64 << 10});
            for (int i = 0; i < NSLOT; ++i) {
                TT_HIP(hipHostMalloc(reinterpret_cast<void **>(&host[i]), cap, hipHostMallocDefault));
                TT_HIP(hipMalloc(reinterpret_cast<void **>(&dev[i]), cap));
                if (!ev[i]) TT_HIP(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming));
            }
        }
        used = 0;
    }
    // n bytes at the next 256-byte boundary of the slot; returns their device address
    void *put(const void *src, size_t n) {
        const size_t off = (used + 255) & ~(size_t)255;
        TT_REQUIRE(off + n <= cap, "grouped launch tables exceed their slot");
        std::memcpy(host[slot] + off, src, n);
        used = off + n;
        return dev[slot] + off;
    }
    void send(hipStream_t s) {
        if (used) TT_HIP(hipMemcpyAsync(dev[slot], host[slot], used, hipMemcpyHostToDevice, s));
    }
    void end(hipStream_t s) {
        TT_HIP(hipEventRecord(ev[slot], s));
        pending[slot] = true;
    }
    // Captured grouped steps, keyed by (stream, basis precision, step, the members' Krylov
    // objects), each with its tables in a device buffer of its own.  A graph is kept while every
    // member's operator layout is the one it was captured with (gens) and no Krylov object has
    // been made or freed since (kgen: the members' buffers are then unchanged).
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        void *tab = nullptr;
        std::vector<uint64_t> gens;
    };
    std::map<std::vector<uintptr_t>, StepGraph> graphs;
    uint64_t graphs_kgen = 0;
    static void drop(StepGraph &g) {
        if (g.exec || g.tab) (void)hipDeviceSynchronize();   // its last replay may still run
        if (g.exec) (void)hipGraphExecDestroy(g.exec);
        if (g.tab) (void)hipFree(g.tab);
        g.exec = nullptr;
        g.tab = nullptr;
    }
    void clear_graphs() {
        for (auto &kv : graphs) drop(kv.second);
        graphs.clear();
    }
};
void free_group_tables(GroupTables *g) { delete g; }

// the grouped product instance of a stage key (MT * 100 + NT * 10 + UN)
template <typename T, int MT, int NT>
static void launch_prod_grp_un(hipStream_t s, int UN, unsigned grid, const GroupHead *h, const ProdMember *m) {
    if constexpr (sizeof(T) == 8) {
        if (UN == 1)
            hipLaunchKernelGGL((k_cellmfma_grp<T, MT, NT, 1>), dim3(grid), dim3(256), 0, s, h, m);
        else if (UN == 4)
            hipLaunchKernelGGL((k_cellmfma_grp<T, MT, NT, 4>), dim3(grid), dim3(256), 0, s, h, m);
        else
            hipLaunchKernelGGL((k_cellmfma_grp<T, MT, NT, 2>), dim3(grid), dim3(256), 0, s, h, m);
    } else {
        if (UN == 2)
            hipLaunchKernelGGL((k_cellmfma_grp<T, MT, NT, 2>), dim3(grid), dim3(256), 0, s, h, m);
        else if (UN == 8)
            hipLaunchKernelGGL((k_cellmfma_grp<T, MT, NT, 8>), dim3(grid), dim3(256), 0, s, h, m);
        else
            hipLaunchKernelGGL((k_cellmfma_grp<T, MT, NT, 4>), dim3(grid), dim3(256), 0, s, h, m);
    }
}
template <typename T, int MT>
static void launch_prod_grp_nt(hipStream_t s, int NT, int UN, unsigned grid, const GroupHead *h, const ProdMember *m) {
    switch (NT) {
        case 1: launch_prod_grp_un<T, MT, 1>(s, UN, grid, h, m); break;
        case 2: launch_prod_grp_un<T, MT, 2>(s, UN, grid, h, m); break;
        case 3: launch_prod_grp_un<T, MT, 3>(s, UN, grid, h, m); break;
        default: launch_prod_grp_un<T, MT, 4>(s, UN, grid, h, m); break;
    }
}
template <typename T>
static void launch_prod_grp(hipStream_t s, int key, unsigned grid, const GroupHead *h, const ProdMember *m) {
    const int MT = key / 100, NT = key / 10 % 10, UN = key % 10;
    switch (MT) {
        case 1: launch_prod_grp_nt<T, 1>(s, NT, UN, grid, h, m); break;
        case 2: launch_prod_grp_nt<T, 2>(s, NT, UN, grid, h, m); break;
        case 3: launch_prod_grp_nt<T, 3>(s, NT, UN, grid, h, m); break;
        default: launch_prod_grp_nt<T, 4>(s, NT, UN, grid, h, m); break;
    }
}
// the product's instance key for Nalpha, R2 and batch depth un, as launch_cellmm chooses it
template <typename T>
static int prod_key(int Na, int R2, int un) {
    const int MT = Na <= 16 ? 1 : Na <= 32 ? 2 : Na <= 48 ? 3 : 4;
    if (un == 0) un = cell_un<T>();
    if (sizeof(T) == 8)
        un = (un == 1 || un == 4) ? un : 2;
    else
        un = (un == 2 || un == 8) ? un : 4;
    return MT * 100 + cell_nt(R2) * 10 + un;
}

// The plan's launches: each stage in order, one launch per GROUP_MAX members, with the byte
// offsets of its head and member records in the plan's packed tables (256-byte aligned).
struct GroupLaunch {
    int phase;
    const GroupStage *g;
    size_t head_off, mem_off;
    unsigned grid;
};
static size_t pack_group_plan(const GroupPlan &gp, std::vector<char> &bytes, std::vector<GroupLaunch> &ls) {
    bytes.clear();
    ls.clear();
    auto put = [&](const void *src, size_t n) {
        const size_t off = (bytes.size() + 255) & ~(size_t)255;
        bytes.resize(off + n);
        std::memcpy(bytes.data() + off, src, n);
        return off;
    };
    for (const auto &kv : gp.st) {
        const GroupStage &g = kv.second;
        for (int c0 = 0; c0 < g.count(); c0 += GROUP_MAX) {
            const int n = std::min(GROUP_MAX, g.count() - c0);
            GroupHead head{};
            head.n = n;
            int64_t tot = 0;
            for (int i = 0; i < n; ++i) {
                head.start[i] = (int)tot;
                tot += g.nblk[c0 + i];
            }
            TT_REQUIRE(tot > 0 && tot < (1ll << 31), "grouped launch: grid out of range");
            head.start[n] = (int)tot;
            const size_t ho = put(&head, sizeof(head));
            const size_t mo = put(g.mem.data() + (size_t)c0 * g.msize, (size_t)n * g.msize);
            ls.push_back(GroupLaunch{std::get<0>(kv.first), &g, ho, mo, (unsigned)tot});
        }
    }
    return bytes.size();
}

// The launches on s, their tables at dev (the packed bytes' device copy).
static void launch_group_plan(const std::vector<GroupLaunch> &ls, const char *dev, hipStream_t s) {
    for (const GroupLaunch &l : ls) {
        const dim3 grid(l.grid), blk(256);
        const bool f = l.g->f32;
        const GroupHead *h = reinterpret_cast<const GroupHead *>(dev + l.head_off);
        const void *m = dev + l.mem_off;
        switch (l.phase) {
            case GP_BLOCK_IN:
                if (f)
                    hipLaunchKernelGGL((k_block_grp<float2, float2, float, float2>), grid, blk, l.g->lds, s, h,
                                       static_cast<const BlockMember *>(m));
                else
                    hipLaunchKernelGGL((k_block_grp<double2, double2, double, double2>), grid, blk, l.g->lds, s, h,
                                       static_cast<const BlockMember *>(m));
                break;
            case GP_BLOCK_OUT:
                if (f)
                    hipLaunchKernelGGL((k_block_grp<float2, float2, float, double2>), grid, blk, l.g->lds, s, h,
                                       static_cast<const BlockMember *>(m));
                else
                    hipLaunchKernelGGL((k_block_grp<double2, double2, double, double2>), grid, blk, l.g->lds, s, h,
                                       static_cast<const BlockMember *>(m));
                break;
            case GP_PRODUCT:
                if (f)
                    launch_prod_grp<float>(s, l.g->key, l.grid, h, static_cast<const ProdMember *>(m));
                else
                    launch_prod_grp<double>(s, l.g->key, l.grid, h, static_cast<const ProdMember *>(m));
                break;
            case GP_DOTS:
            case GP_DOTS2:
                if (f)
                    hipLaunchKernelGGL(k_dots_grp<float2>, grid, blk, 256 * sizeof(double2), s, h,
                                       static_cast<const VecMember *>(m));
                else
                    hipLaunchKernelGGL(k_dots_grp<double2>, grid, blk, 256 * sizeof(double2), s, h,
                                       static_cast<const VecMember *>(m));
                break;
            case GP_COMBO:
                if (f)
                    hipLaunchKernelGGL((k_combo_grp<float2, float2>), grid, blk, 0, s, h,
                                       static_cast<const VecMember *>(m));
                else
                    hipLaunchKernelGGL((k_combo_grp<double2, double2>), grid, blk, 0, s, h,
                                       static_cast<const VecMember *>(m));
                break;
            case GP_FINAL:   // FP32: normalisation by the norm of the projected vector; FP64: the second pass
                if (f)
                    hipLaunchKernelGGL(k_normalize_grp<float2>, grid, blk, 0, s, h, static_cast<const VecMember *>(m));
                else
                    hipLaunchKernelGGL(k_combo_normalize_grp<double2>, grid, blk, 0, s, h,
                                       static_cast<const VecMember *>(m));
                break;
            default:
                fail(TT_ERR_ARG, "grouped launch: unknown stage");
        }
    }
    TT_HIP(hipGetLastError());
}

// The plan's stages on s, their tables sent from a slot of gt's ring.
static void run_group_plan(GroupTables &gt, const GroupPlan &gp, hipStream_t s) {
    if (gp.empty()) return;
    std::vector<char> bytes;
    std::vector<GroupLaunch> ls;
    const size_t n = pack_group_plan(gp, bytes, ls);
    gt.begin(n);
    const char *dev = static_cast<const char *>(gt.put(bytes.data(), n));
    gt.send(s);
    launch_group_plan(ls, dev, s);
    gt.end(s);
}

struct RowOperator {
    hipStream_t s;
    ResolventStaging rstage;
    int Na, Nq, Np, nE;
    int64_t dim;
    const DeviceCSC *P;
    DevBuf<int> c_ptr, c_alpha, vc_ptr, vc_alpha;
    DevBuf<int64_t> c_src, vc_src;   // where each partner block sits in the 2N tables
    DevBuf<int> c_ld, vc_ld;
    int c_nb = 0, vc_nb = 0;         // partner blocks (0: layout not made yet)
    DevBuf<double> c_blocks, vc_blocks;
    DevBuf<float> c_blocks32, vc_blocks32;   // FP32 copies for the block ops on FP32 vectors
    DevBuf<double2> G;
    DevBuf<float2> G32;                 // FP32 copy of G (the FP32 vectors' block op)
    uint64_t g_gen = 0, g32_gen = ~0ull;
    DevBuf<double2> S, U;      // FP64 scratch (dim * r)
    DevBuf<float2> S32, U32;   // FP32 scratch
    DevBuf<int> gidx;
    std::vector<int> h_gidx;   // what gidx holds (set_gidx sends only changes)
    // Event pool: four events per timed application (before the first block op, around the
    // SpMM, after the last block op); read back once in collect(), no per-apply synchronisation.
    std::vector<hipEvent_t> evs;
    size_t ev_used = 0;
    std::vector<double> spmm_bytes_of;   // algorithmic bytes of each timed SpMM launch
    std::vector<char> spmm_fp32_of;       // and whether it was an FP32 launch
    double ms_apply = 0.0, ms_spmm = 0.0, spmm_bytes = 0.0;
    double ms_spmm32 = 0.0, spmm_bytes32 = 0.0;
    int n_apply = 0, n_apply32 = 0, n_spmm = 0, n_spmm32 = 0;
    int r_cap = 0;
    int c_npart_max = 0, vc_npart_max = 0;
    bool cells = true;   // cell-block product (default) or gather over the CSC
    // col: the column operator z -> A z = C_op^T (2 P123) VC_op z (products with P123 itself,
    // through the transposed cells) instead of the row form y -> y A
    bool col = false;
    CellView cv;
    // work items by item length; the autotuner picks the length and batch depth per
    // (precision, R2), and every application of that shape uses its choice
    std::map<int, std::unique_ptr<CellItems>> items_by_len;
    struct CellTune {
        int len = CELL_ITEM, un = 0, h = 0;   // h: the split-half FP16 product (FP32 only)
    };
    std::map<int, CellTune> tuned;   // key tune_key(fp32, R2)
    DevBuf<double> upart;  // partial panels of the work items (FP64 and FP32 share it)
    // the split-half FP16 product: its split panels of S, [0] the S scale (float bits, written by
    // k_split_s) and [1 .. AMAX_SLOTS] the block op's max |S| (float bits, cleared by the
    // product), the split blocks and their scale
    DevBuf<HalfPair> Sth;
    DevBuf<unsigned> hsc;
    const HalfPair *Ah = nullptr;
    float Ah_scale = 1.0f;
    // bumped whenever a buffer the kernels address is (re)allocated: captured graphs hold
    // these addresses and are rebuilt when it changes
    uint64_t layout_gen = 0;
    // a buffer the first block op of the next application zeroes (for accumulations that follow)
    double2 *zero_next = nullptr;
    // work items for the next application only (items_on_q), else the tuned ones
    const CellItems *next_items = nullptr;
    std::map<std::vector<int>, std::unique_ptr<CellItems>> items_q;   // key: item length, then q bins
    int zero_next_n = 0;

    ~RowOperator() {
        for (hipEvent_t e : evs) (void)hipEventDestroy(e);
    }

    hipEvent_t next_event() {
        if (ev_used == evs.size()) {
            hipEvent_t e;
            TT_HIP(hipEventCreate(&e));
            evs.push_back(e);
        }
        return evs[ev_used++];
    }

    // Sum the device times of the applications recorded since the last collect().
    void collect() {
        if (ev_used == 0) return;
        TT_HIP(hipEventSynchronize(evs[ev_used - 1]));
        for (size_t i = 0; i + 3 < ev_used; i += 4) {
            float a = 0.f, b = 0.f;
            TT_HIP(hipEventElapsedTime(&a, evs[i], evs[i + 3]));
            TT_HIP(hipEventElapsedTime(&b, evs[i + 1], evs[i + 2]));
            ms_apply += a;
            add_product_time(b, spmm_bytes_of[i / 4], spmm_fp32_of[i / 4] != 0);
        }
        ev_used = 0;
        spmm_bytes_of.clear();
        spmm_fp32_of.clear();
    }

    void setup(hipStream_t stream, const tt_run_params &rp, const DeviceCSC *csc, const double *V_unco,
               const double *V_coup, const tt_swp_states &swp, const tt_pw_states &chn,
               int p123_format = TT_P123_CELLS, bool col_ = false, VCProducts *vc = nullptr) {
        s = stream;
        cells = p123_format == TT_P123_CELLS;
        col = col_;
        TT_REQUIRE(!col || cells, "the column operator uses the cell-block form");
        Na = chn.Nalpha;
        Nq = swp.Nq_WP;
        Np = swp.Np_WP;
        nE = 0;
        dim = (int64_t)Na * Nq * Np;
        P = csc;
        TT_REQUIRE(csc && csc->dim == dim, "P123 dimension does not match Nalpha*Nq*Np");
        if (cells && csc->Na == 0) build_cell_blocks(*const_cast<DeviceCSC *>(csc), Na, Nq, Np, s);
        TT_REQUIRE(!cells || (csc->Na == Na && csc->NqNp == Nq * Np), "P123 cell blocks do not match the channel");
        if (cells) {
            if (col) build_transposed_cells(*const_cast<DeviceCSC *>(csc), s);
            cv = cells_of(*csc, col);
            items_of_len(CELL_ITEM);
        }
        set_blocks(rp, V_unco, V_coup, swp, chn, vc);
    }

    // C_op / VC_op partner blocks from V and the SWP eigenvectors.  The partner lists are
    // structural (zero blocks kept), so new V or C (another theta) only rewrite the block
    // contents in the same buffers and captured graphs stay valid.  The layout (which 2N block
    // each partner block reads) is made once; the blocks are then formed on the device from the
    // 2N tables of C and V.C (k_embed), which the caller's vc holds on the device for every
    // channel of these V and C (else they are made and uploaded here).
    void set_blocks(const tt_run_params &rp, const double *V_unco, const double *V_coup, const tt_swp_states &swp,
                    const tt_pw_states &chn, VCProducts *vc = nullptr) {
        int nu, nc;
        num_2N_states(rp, nu, nc);
        std::unique_ptr<VCProducts> own;
        if (!vc) {
            own.reset(new VCProducts);
            vc_products(rp, V_unco, V_coup, swp, *own);
            vc = own.get();
        }
        if (!vc->dev_valid) vc->upload(swp, nu, nc, s);
        if (!c_nb) {
            std::vector<int> p1, a1, p2, a2, l1, l2;
            std::vector<int64_t> s1, s2;
            // row form: S = C_op (X G), Y = VC_op^T U; column form: S = VC_op (X G), Y = C_op^T U
            embed_2N_layout(chn, Np, nu, rp, false, p1, a1, s1, l1);
            embed_2N_layout(chn, Np, nu, rp, true, p2, a2, s2, l2);
            int m1 = 1, m2 = 1;
            for (int a = 0; a < Na; ++a) {
                m1 = std::max(m1, p1[a + 1] - p1[a]);
                m2 = std::max(m2, p2[a + 1] - p2[a]);
            }
            ++layout_gen;   // new buffers and launch parameters of the graphs
            c_npart_max = m1;
            vc_npart_max = m2;
            c_nb = (int)a1.size();
            vc_nb = (int)a2.size();
            a1.push_back(0);
            a2.push_back(0);
            s1.push_back(0);
            s2.push_back(0);
            l1.push_back(0);
            l2.push_back(0);
            c_ptr.upload(p1.data(), p1.size(), s);
            c_alpha.upload(a1.data(), a1.size(), s);
            c_src.upload(s1.data(), s1.size(), s);
            c_ld.upload(l1.data(), l1.size(), s);
            vc_ptr.upload(p2.data(), p2.size(), s);
            vc_alpha.upload(a2.data(), a2.size(), s);
            vc_src.upload(s2.data(), s2.size(), s);
            vc_ld.upload(l2.data(), l2.size(), s);
            const size_t n1 = (size_t)c_nb * Np * Np + 1, n2 = (size_t)vc_nb * Np * Np + 1;
            c_blocks.alloc(n1);
            c_blocks32.alloc(n1);
            vc_blocks.alloc(n2);
            vc_blocks32.alloc(n2);
        }
        const int64_t nn1 = (int64_t)c_nb * Np * Np, nn2 = (int64_t)vc_nb * Np * Np;
        const double *T1 = col ? vc->dVC.p : vc->dC.p, *T2 = col ? vc->dC.p : vc->dVC.p;
        if (nn1)
            hipLaunchKernelGGL(k_embed, dim3(nblk(nn1, 256)), dim3(256), 0, s, c_blocks.p, c_blocks32.p, T1, c_src.p,
                               c_ld.p, c_nb, Np, 0);
        if (nn2)
            hipLaunchKernelGGL(k_embed, dim3(nblk(nn2, 256)), dim3(256), 0, s, vc_blocks.p, vc_blocks32.p, T2,
                               vc_src.p, vc_ld.p, vc_nb, Np, 1);
        TT_HIP(hipGetLastError());
        if (own) TT_HIP(hipStreamSynchronize(s));   // own tables are freed on return
    }

    // Device buffer for the resolvent of nE energies (contents written by the caller).
    double2 *G_buffer(int nE_) {
        ++g_gen;
        nE = nE_;
        const size_t n = (size_t)std::max(nE, 1) * dim;
        if (n > G.n) {
            G.alloc(n);
            ++layout_gen;
        }
        return G.p;
    }

    // Resolvent of nE energies (host, [j][idx] complex), copied into the device buffer.
    void set_G(const double *Gh, int nE_) {
        ++g_gen;
        nE = nE_;
        if (nE <= 0 || !Gh) return;
        const size_t n = (size_t)nE * dim;
        if (n > G.n) ++layout_gen;
        G.upload(reinterpret_cast<const double2 *>(Gh), n, s);
    }

    // (the panels start zeroed: the autotuner times products over them before any data)
    void reserve(int r, bool fp32) {
        if (r > r_cap) {
            S.alloc((size_t)dim * r);
            U.alloc((size_t)dim * r);
            gidx.alloc(r);
            h_gidx.clear();
            r_cap = r;
            S32.release();
            U32.release();
            TT_HIP(hipMemsetAsync(S.p, 0, sizeof(double2) * S.n, s));
            ++layout_gen;
        }
        if (fp32 && !S32.p) {
            S32.alloc((size_t)dim * r_cap);
            U32.alloc((size_t)dim * r_cap);
            TT_HIP(hipMemsetAsync(S32.p, 0, sizeof(float2) * S32.n, s));
            ++layout_gen;
        }
    }

    // Every allocation an application with r rows needs, done ahead (graph capture may not
    // allocate).
    void prepare(int r, bool fp32) {
        reserve(r, fp32);
        if (fp32) ensure_f32();
        if (fp32 && G.p && g32_gen != g_gen) {   // FP32 copy of G for the FP32 block ops
            if (G32.n < G.n) {
                G32.alloc(G.n);
                ++layout_gen;
            }
            hipLaunchKernelGGL(k_g_to_f32, dim3(nblk((int64_t)G.n, 256)), dim3(256), 0, s, G.p, G32.p, (int64_t)G.n);
            g32_gen = g_gen;
        }
        if (cells) {
            tune(r, fp32);
            if (fp32 && choice(true, 2 * r).h) prepare_h(r);
            const size_t need = (size_t)std::max(items_for(fp32, 2 * r).n, 1) * Na * 2 * r;
            if (upart.n < need) {
                upart.alloc(need);
                ++layout_gen;
            }
        }
        // the block ops index X, Y, the item panels and G with 32-bit element indexes
        TT_REQUIRE((int64_t)dim * r < INT32_MAX && (int64_t)upart.n < INT32_MAX && (int64_t)G.n < INT32_MAX,
                   "block op: a vector, the item panels or G exceed its 32-bit element indexes");
    }

    static int tune_key(bool fp32, int R2) { return (fp32 ? (1 << 20) + ((cell_h_mode() + 1) << 16) : 0) + R2; }

    CellItems &items_of_len(int len) {
        auto it = items_by_len.find(len);
        if (it != items_by_len.end()) return *it->second;
        std::unique_ptr<CellItems> ci(new CellItems);
        ci->build(cv, s, len);
        CellItems &ref = *ci;
        items_by_len[len] = std::move(ci);
        return ref;
    }
    const CellTune &choice(bool fp32, int R2) const {
        static const CellTune dflt;
        auto it = tuned.find(tune_key(fp32, R2));
        return it == tuned.end() ? dflt : it->second;
    }
    CellItems &items_for(bool fp32, int R2) { return items_of_len(choice(fp32, R2).len); }
    // Work items over the cells whose input bin lies in one of the q bins (row operator), or
    // nullptr when none does.
    const CellItems *items_on_q(const std::vector<int> &qbins, int len) {
        std::vector<int> key(1, len);
        key.insert(key.end(), qbins.begin(), qbins.end());
        auto it = items_q.find(key);
        if (it == items_q.end()) {
            std::vector<char> mask(Nq, 0);
            for (int q : qbins) {
                TT_REQUIRE(q >= 0 && q < Nq, "q bin out of range");
                mask[q] = 1;
            }
            std::unique_ptr<CellItems> ci(new CellItems);
            ci->build_on_q(cv, s, len, Np, mask);
            it = items_q.emplace(key, std::move(ci)).first;
        }
        return it->second->n > 0 ? it->second.get() : nullptr;
    }

    // Autotuner of the cell-block product: on the first application of a shape (precision, r),
    // each candidate item length and batch depth is timed, together with the block op that
    // sums the work items' partial panels, on this operator's own blocks (a
    // warm-up launch, then the mean of three launches between events, the device otherwise
    // idle) and the fastest kept.  Choices are shared through a process-wide table keyed by
    // device, orientation, precision, Nalpha, R2 and the cell count, so channels of equal shape
    // tune once.  Runs eagerly from prepare(), never inside a graph capture.
    void tune(int r, bool fp32) {
        const int R2 = 2 * r;
        const int key = tune_key(fp32, R2);
        if (tuned.count(key)) return;
        int dev = 0;
        TT_HIP(hipGetDevice(&dev));
        TuneTable &tt_ = tune_table();
        const int hmode = fp32 ? cell_h_mode() : 0;
        const TuneKey gk{dev, col ? 1 : 0, fp32 ? 1 : 0, Na, R2, cv.ncell, hmode};
        {
            std::lock_guard<std::mutex> lock(tt_.mu);
            auto it = tt_.map.find(gk);
            if (it != tt_.map.end()) {
                tuned[key] = {it->second.len, it->second.un, it->second.h};
                if (it->second.h) prepare_h(r);
                ++layout_gen;
                return;
            }
        }
        const int lens[4] = {8, 16, 32, 64};
        size_t need = 0;
        for (int len : lens) need = std::max(need, (size_t)std::max(items_of_len(len).n, 1) * Na * R2);
        if (upart.n < need) {
            upart.alloc(need);
            ++layout_gen;
        }
        // the trials run alone on the device (other streams of a lockstep group would share it)
        TT_HIP(hipDeviceSynchronize());
        hipEvent_t e0, e1;
        TT_HIP(hipEventCreate(&e0));
        TT_HIP(hipEventCreate(&e1));
        CellTune best;
        float best_ms = 1e30f;
        if (hmode != 0) prepare_h(r);
        try {
            for (int len : lens) {
                const CellItems &ci = items_of_len(len);
                // candidates: (un, h); the split-half product takes 1 or 2 steps per batch
                std::vector<std::pair<int, int>> cands;
                if (hmode != 1)
                    for (int un : fp32 ? cell_un_candidates<float>() : cell_un_candidates<double>())
                        cands.emplace_back(un, 0);
                if (hmode != 0)
                    for (int un : {1, 2}) cands.emplace_back(un, 1);
                for (const auto &cand : cands) {
                    const int un = cand.first, h = cand.second;
                    // the product and the block op that sums its item panels (the item length
                    // trades the product's parallelism against that sum)
                    auto launch = [&]() {
                        if (h) {
                            launch_cellh(s, &cv, ci.ref(), Ah, Ah_scale, reinterpret_cast<const float *>(S32.p), Sth.p,
                                         reinterpret_cast<float *>(hsc.p), hsc.p + 1,
                                         reinterpret_cast<float *>(upart.p), R2, un);
                            launch_block<float2, float2, float>(s, reinterpret_cast<const float2 *>(upart.p), U32.p,
                                                                vc_ptr.p, vc_alpha.p, vc_blocks32.p, vc_npart_max,
                                                                nullptr, gidx.p, Na, Nq, Np, r, nullptr, 1, 0,
                                                                ci.ket_item_ptr.p, 2.0);
                        } else if (fp32) {
                            launch_cellmm<float>(s, &cv, ci.ref(), cv.blk32, reinterpret_cast<const float *>(S32.p), nullptr,
                                                 reinterpret_cast<float *>(upart.p), R2, 2.0f, un);
                            launch_block<float2, float2, float>(s, reinterpret_cast<const float2 *>(upart.p), U32.p,
                                                                vc_ptr.p, vc_alpha.p, vc_blocks32.p, vc_npart_max,
                                                                nullptr, gidx.p, Na, Nq, Np, r, nullptr, 1, 0,
                                                                ci.ket_item_ptr.p, 2.0);
                        } else {
                            launch_cellmm<double>(s, &cv, ci.ref(), cv.blk, reinterpret_cast<const double *>(S.p), nullptr,
                                                  upart.p, R2, 2.0, un);
                            launch_block<double2, double2>(s, reinterpret_cast<const double2 *>(upart.p), U.p, vc_ptr.p,
                                                           vc_alpha.p, vc_blocks.p, vc_npart_max, nullptr, gidx.p, Na,
                                                           Nq, Np, r, nullptr, 1, 0, ci.ket_item_ptr.p, 2.0);
                        }
                    };
                    launch();
                    TT_HIP(hipEventRecord(e0, s));
                    for (int rep = 0; rep < 3; ++rep) launch();
                    TT_HIP(hipEventRecord(e1, s));
                    TT_HIP(hipEventSynchronize(e1));
                    float ms = 0.f;
                    TT_HIP(hipEventElapsedTime(&ms, e0, e1));
                    if (ms < best_ms) {
                        best_ms = ms;
                        best.len = len;
                        best.un = un;
                        best.h = h;
                    }
                }
            }
        } catch (...) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            throw;
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        tuned[key] = best;
        ++layout_gen;
        std::lock_guard<std::mutex> lock(tt_.mu);
        tt_.map[gk] = {best.len, best.un, best.h, best_ms / 3.0f};
    }

    void set_gidx(const std::vector<int> &g) {
        reserve((int)g.size(), false);
        upload_if_changed(gidx, h_gidx, g.data(), g.size(), s);
    }

    void ensure_f32() {
        DeviceCSC *m = const_cast<DeviceCSC *>(P);
        if (cells) {
            float *&b32 = col ? m->t_blk32 : m->blk32;
            if (!b32) {
                const int64_t n = m->ncell * m->Na * m->Na;
                TT_HIP(hipMalloc(&b32, sizeof(float) * (n + BLK_SLACK)));
                hipLaunchKernelGGL(k_to_f32, dim3(nblk(n + BLK_SLACK, 256)), dim3(256), 0, s, col ? m->t_blk : m->blk,
                                   b32, n + BLK_SLACK);
            }
            cv.blk32 = b32;
        } else if (!m->val32) {
            TT_HIP(hipMalloc(&m->val32, sizeof(float) * std::max<int64_t>(m->nnz, 1)));
            if (m->nnz) hipLaunchKernelGGL(k_to_f32, dim3(nblk(m->nnz, 256)), dim3(256), 0, s, m->val, m->val32, m->nnz);
        }
    }

    // Split-half blocks of this orientation (made once per P123 from the FP64 blocks) and the
    // buffers of the split-half product for r right-hand sides.  Not inside a graph capture.
    void prepare_h(int r) {
        DeviceCSC *m = const_cast<DeviceCSC *>(P);
        uint16_t *&bh = col ? m->t_blkh : m->blkh;
        double &sc = col ? m->t_blkh_scale : m->blkh_scale;
        const int Kp8 = (Na + 7) / 8;
        if (!bh) {
            const double *src = col ? m->t_blk : m->blk;
            DevBuf<unsigned> mx(1);
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, mx.p, 1, 0u);
            const int64_t n = m->ncell * Na * Na;
            hipLaunchKernelGGL(k_absmax_f64, dim3(std::min<unsigned>(nblk(n, 256), 2048)), dim3(256), 0, s, src, n, mx.p);
            unsigned bits = 0;
            TT_HIP(hipMemcpyAsync(&bits, mx.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
            TT_HIP(hipStreamSynchronize(s));
            float mf;
            std::memcpy(&mf, &bits, sizeof(mf));
            sc = split_scale(mf);
            const int64_t nh = m->ncell * Na * Kp8;
            TT_HIP(hipMalloc(&bh, sizeof(HalfPair) * std::max<int64_t>(nh, 1)));
            if (nh)
                hipLaunchKernelGGL(k_blk_split, dim3(nblk(nh, 256)), dim3(256), 0, s, src,
                                   reinterpret_cast<HalfPair *>(bh), m->ncell, Na, Kp8, sc);
            TT_HIP(hipGetLastError());
        }
        Ah = reinterpret_cast<const HalfPair *>(bh);
        Ah_scale = (float)sc;
        reserve(r, true);
        const size_t need = (size_t)m->NqNp * 2 * r_cap * Kp8;
        if (Sth.n < need) {
            Sth.alloc(need);
            ++layout_gen;
        }
        if (!hsc.p) {
            hsc.alloc(1 + AMAX_SLOTS);
            hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, hsc.p, 1 + AMAX_SLOTS, 0u);
            ++layout_gen;
        }
    }

    // Bytes the product over P123 must move per launch (one pass over P, S read and U written once).
    double product_bytes(int r, bool fp32) const {
        const double sv = fp32 ? 4.0 : 8.0, sc = fp32 ? 8.0 : 16.0;
        const double vec = 2.0 * (double)dim * r * sc;
        if (cells) return (double)P->ncell * P->Na * P->Na * sv + 4.0 * (double)(P->ncell + P->NqNp + 1) + vec;
        return (double)P->nnz * (sv + 4.0) + 8.0 * (double)(dim + 1) + vec;
    }

    // Y = (X .* G_{gidx}) A ; X, Y device, dim*r complex; use_G selects whether G multiplies X.
    // Y = (X .* G_{gidx}) A, or Y = Zsub - (X .* G_{gidx}) A when Zsub is given.
    // Timed applications take four events of the pool, read back in collect().
    // VT: element type of X, Y and Zsub (double2, or float2 for the FP32 vectors of the inner
    // solves of iterative refinement, which take the FP32 product).
    template <typename VT>
    void apply(const VT *X, VT *Y, int r, bool use_G, bool fp32, bool timed = true, const VT *Zsub = nullptr) {
        constexpr bool v64 = std::is_same<VT, double2>::value;
        TT_REQUIRE(v64 || fp32, "FP32 vectors take the FP32 product");
        prepare(r, fp32);
        if (timed && ev_used + 4 > 4096) collect();
        hipEvent_t e0 = nullptr, e1 = nullptr, e2 = nullptr, e3 = nullptr;
        if (timed) {
            e0 = next_event();
            e1 = next_event();
            e2 = next_event();
            e3 = next_event();
        }
        auto record = [&](hipEvent_t e) { TT_HIP(hipEventRecord(e, s)); };
        if (timed) record(e0);
        const int R2 = 2 * r;
        const int cm = cells ? 1 : 0;
        // cell path: S -> work-item partial panels; the item sum and the factor 2 are folded
        // into the input staging of the V.C block op
        const CellItems *ci = cells ? (next_items ? next_items : &items_for(fp32, R2)) : nullptr;
        next_items = nullptr;
        if (ci && upart.n < (size_t)std::max(ci->n, 1) * Na * R2) {
            upart.alloc((size_t)std::max(ci->n, 1) * Na * R2);
            ++layout_gen;
        }
        const int un = cells ? choice(fp32, R2).un : 0;
        const int *kip = cells ? ci->ket_item_ptr.p : nullptr;
        const double in_scale = cells ? 2.0 : 1.0;
        double2 *zb = zero_next;
        const int zn = zero_next_n;
        zero_next = nullptr;
        zero_next_n = 0;
        if (!fp32) {
            if constexpr (v64) {
                launch_block<double2, double2>(s, X, S.p, c_ptr.p, c_alpha.p, c_blocks.p, c_npart_max,
                                               use_G ? G.p : nullptr, gidx.p, Na, Nq, Np, r, nullptr, 0, cm, nullptr,
                                               1.0, zb, zn);
                if (cells)   // timed: e1 and e2 hold the kernel's own start and end
                    launch_cellmm<double>(s, &cv, ci->ref(), cv.blk, reinterpret_cast<const double *>(S.p), nullptr,
                                          upart.p, R2, 2.0, un, timed ? e1 : nullptr, timed ? e2 : nullptr);
                else {
                    if (timed) record(e1);
                    launch_spmm<double>(s, P, P->val, reinterpret_cast<const double *>(S.p),
                                        reinterpret_cast<double *>(U.p), dim, R2, 2.0);
                    if (timed) record(e2);
                }
                const double2 *Uin = cells ? reinterpret_cast<const double2 *>(upart.p) : U.p;
                launch_block<double2, double2>(s, Uin, Y, vc_ptr.p, vc_alpha.p, vc_blocks.p, vc_npart_max, nullptr,
                                               gidx.p, Na, Nq, Np, r, Zsub, cm, 0, kip, in_scale);
                ++n_apply;
            }
        } else {
            ensure_f32();
            // FP32 vectors stage and sum the block ops in FP32 too (FP32 block copies); FP64
            // vectors keep FP64 block ops around the FP32 product
            // the split-half product needs the block op's max |S|
            const bool h = cells && choice(true, R2).h;
            unsigned *amax = h ? hsc.p + 1 : nullptr;
            if constexpr (v64)
                launch_block<VT, float2>(s, X, S32.p, c_ptr.p, c_alpha.p, c_blocks.p, c_npart_max,
                                         use_G ? G.p : nullptr, gidx.p, Na, Nq, Np, r, nullptr, 0, cm, nullptr, 1.0,
                                         zb, zn, amax);
            else
                launch_block<VT, float2, float, float2>(s, X, S32.p, c_ptr.p, c_alpha.p, c_blocks32.p, c_npart_max,
                                                use_G ? G32.p : nullptr, gidx.p, Na, Nq, Np, r, nullptr, 0, cm, nullptr,
                                                1.0, zb, zn, amax);
            if (h || !cells) {
                if (timed) record(e1);
                if (h)
                    launch_cellh(s, &cv, ci->ref(), Ah, Ah_scale, reinterpret_cast<const float *>(S32.p), Sth.p,
                                 reinterpret_cast<float *>(hsc.p), amax, reinterpret_cast<float *>(upart.p), R2, un);
                else
                    launch_spmm<float>(s, P, P->val32, reinterpret_cast<const float *>(S32.p),
                                       reinterpret_cast<float *>(U32.p), dim, R2, 2.0f);
                if (timed) record(e2);
            } else {   // timed: e1 and e2 hold the kernel's own start and end
                launch_cellmm<float>(s, &cv, ci->ref(), cv.blk32, reinterpret_cast<const float *>(S32.p), nullptr,
                                     reinterpret_cast<float *>(upart.p), R2, 2.0f, un, timed ? e1 : nullptr,
                                     timed ? e2 : nullptr);
            }
            const float2 *Uin = cells ? reinterpret_cast<const float2 *>(upart.p) : U32.p;
            if constexpr (v64)
                launch_block<float2, VT>(s, Uin, Y, vc_ptr.p, vc_alpha.p, vc_blocks.p, vc_npart_max, nullptr, gidx.p,
                                         Na, Nq, Np, r, Zsub, cm, 0, kip, in_scale);
            else
                launch_block<float2, VT, float>(s, Uin, Y, vc_ptr.p, vc_alpha.p, vc_blocks32.p, vc_npart_max, nullptr,
                                                gidx.p, Na, Nq, Np, r, Zsub, cm, 0, kip, in_scale);
            ++n_apply;
            ++n_apply32;
        }
        if (timed) {
            record(e3);
            spmm_bytes_of.push_back(product_bytes(r, fp32));
            spmm_fp32_of.push_back(fp32 ? 1 : 0);
        }
    }

    // Whether this operator's applications can join a grouped launch: the cell-block product
    // without the split-half form (FP64 vectors with the FP64 product, FP32 with FP32).
    bool groupable(int r, bool fp32) const { return cells && !(fp32 && choice(true, 2 * r).h); }

    // apply() as member records of a grouped call (nothing is launched here): the first block
    // op, the product and the second block op of this operator join stages 0-2 of gp, with the
    // same arguments apply() gives its own launches.  Untimed.
    template <typename VT>
    void group_apply(const VT *X, VT *Y, int r, bool use_G, bool fp32, const VT *Zsub, GroupPlan &gp) {
        constexpr bool v64 = std::is_same<VT, double2>::value;
        TT_REQUIRE(v64 != fp32, "grouped applications: FP64 vectors take the FP64 product, FP32 the FP32 one");
        prepare(r, fp32);
        TT_REQUIRE(groupable(r, fp32), "grouped applications need the cell-block product (not the split-half one)");
        const int R2 = 2 * r;
        const CellItems *ci = next_items ? next_items : &items_for(fp32, R2);
        next_items = nullptr;
        if (upart.n < (size_t)std::max(ci->n, 1) * Na * R2) {
            upart.alloc((size_t)std::max(ci->n, 1) * Na * R2);
            ++layout_gen;
        }
        const int un = choice(fp32, R2).un;
        BlockMember b{};
        b.part_ptr = c_ptr.p;
        b.part_alpha = c_alpha.p;
        b.gidx = gidx.p;
        b.Na = Na;
        b.Nq = Nq;
        b.Np = Np;
        b.r = r;
        b.gx = Na * Nq;
        b.in_scale = 1.0;
        b.X = X;
        b.Y = fp32 ? (void *)S32.p : (void *)S.p;
        b.blocks = fp32 ? (const void *)c_blocks32.p : (const void *)c_blocks.p;
        b.G = use_G ? (fp32 ? (const void *)G32.p : (const void *)G.p) : nullptr;
        b.ne_max = c_npart_max;
        b.in_cell = 0;
        b.out_cell = 1;
        b.zero_buf = zero_next;
        b.zero_n = zero_next_n;
        zero_next = nullptr;
        zero_next_n = 0;
        BlockGeom g1 = fp32 ? block_geom<float>(r, c_npart_max, Np) : block_geom<double>(r, c_npart_max, Np);
        b.kc = g1.kc;
        gp.add(GP_BLOCK_IN, fp32, 0, b, (int64_t)b.gx * g1.gy, g1.lds);
        ProdMember pm{ci->items.p, cv.partner, fp32 ? (const void *)cv.blk32 : (const void *)cv.blk,
                      fp32 ? (const void *)S32.p : (const void *)S.p, upart.p, Na, R2, ci->n};
        const int key = fp32 ? prod_key<float>(Na, R2, un) : prod_key<double>(Na, R2, un);
        const int ny = (R2 + 16 * cell_nt(R2) - 1) / (16 * cell_nt(R2));
        TT_REQUIRE(Na >= 1 && Na <= CELL_NA_MAX, "cell-block product needs 1 <= Nalpha <= 64");
        gp.add(GP_PRODUCT, fp32, key, pm, (int64_t)ci->n * ny);
        BlockMember o = b;
        o.X = upart.p;
        o.Y = Y;
        o.part_ptr = vc_ptr.p;
        o.part_alpha = vc_alpha.p;
        o.blocks = fp32 ? (const void *)vc_blocks32.p : (const void *)vc_blocks.p;
        o.G = nullptr;
        o.ne_max = vc_npart_max;
        o.Zsub = Zsub;
        o.in_cell = 1;
        o.out_cell = 0;
        o.ket_item_ptr = ci->ket_item_ptr.p;
        o.in_scale = 2.0;
        o.zero_buf = nullptr;
        o.zero_n = 0;
        BlockGeom g3 = fp32 ? block_geom<float>(r, vc_npart_max, Np) : block_geom<double>(r, vc_npart_max, Np);
        o.kc = g3.kc;
        gp.add(GP_BLOCK_OUT, fp32, 0, o, (int64_t)o.gx * g3.gy, g3.lds);
        ++n_apply;
        if (fp32) ++n_apply32;
    }

    // Ucell = 2 P^T (C_op X) for the row operator, 2 P (VC_op X) for the column operator: the
    // first block op (no G) and the cell product, as a full cell-major panel [(q,p)][alpha][k].
    void stage_product(const double2 *X, double2 *Ucell, int r) {
        TT_REQUIRE(cells, "stage products use the cell-block form");
        prepare(r, false);
        launch_block<double2, double2>(s, X, S.p, c_ptr.p, c_alpha.p, c_blocks.p, c_npart_max, nullptr, gidx.p, Na,
                                       Nq, Np, r, nullptr, 0, 1, nullptr, 1.0);
        launch_cellmm<double>(s, &cv, items_for(false, 2 * r).ref(), cv.blk, reinterpret_cast<const double *>(S.p),
                              reinterpret_cast<double *>(Ucell), upart.p, 2 * r, 2.0, choice(false, 2 * r).un);
    }

    void add_product_time(double ms, double bytes, bool fp32) {
        ms_spmm += ms;
        spmm_bytes += bytes;
        ++n_spmm;
        if (fp32) {
            ms_spmm32 += ms;
            spmm_bytes32 += bytes;
            ++n_spmm32;
        }
    }
};

// ------------------------------------------------------------------ Pade (host)
static cdouble det_lu(std::vector<cdouble> A, int n) {
    // LU with partial pivoting; det = prod(diag) * (-1)^swaps (matrix_routines.cpp:82-101)
    cdouble det = 1.0;
    for (int k = 0; k < n; ++k) {
        int piv = k;
        double best = std::abs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i)
            if (std::abs(A[(size_t)i * n + k]) > best) {
                best = std::abs(A[(size_t)i * n + k]);
                piv = i;
            }
        if (piv != k) {
            for (int j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)piv * n + j]);
            det = -det;
        }
        cdouble d = A[(size_t)k * n + k];
        det *= d;
        if (d == cdouble(0.0)) return 0.0;
        for (int i = k + 1; i < n; ++i) {
            cdouble f = A[(size_t)i * n + k] / d;
            for (int j = k + 1; j < n; ++j) A[(size_t)i * n + j] -= f * A[(size_t)k * n + j];
        }
    }
    return det;
}

// CPP/solve_faddeev.cpp:271-296
static cdouble pade_approximant(const cdouble *a, int N, int M, cdouble z) {
    std::vector<cdouble> P((size_t)(M + 1) * (M + 1)), Q((size_t)(M + 1) * (M + 1));
    for (int r = 0; r < M; ++r)
        for (int c = 0; c <= M; ++c) P[(size_t)r * (M + 1) + c] = Q[(size_t)r * (M + 1) + c] = a[N - M + 1 + r + c];
    for (int c = 0; c <= M; ++c) {
        Q[(size_t)M * (M + 1) + c] = std::pow(z, M - c);
        cdouble s = 0.0;
        for (int j = M - c; j <= N; ++j) s += a[j - (M - c)] * std::pow(z, j);
        P[(size_t)M * (M + 1) + c] = s;
    }
    return det_lu(P, M + 1) / det_lu(Q, M + 1);
}

// sqrt of a squared norm: rounding below zero gives 0, NaN stays NaN (std::max(0.0, NaN) is 0)
static inline double sqrt_nonneg(double n2) { return std::isnan(n2) ? n2 : std::sqrt(n2 > 0.0 ? n2 : 0.0); }

// ------------------------------------------------------------------ Krylov engine
// GMRES on r independent systems M y_k = b_k, M y = y - (y .* G_k) A, that share every
// operator application: one Arnoldi basis per system (column k of the basis vectors).  VT is
// the storage type of the basis: double2 for the FP64 solves, float2 for the inner solves of
// iterative refinement (FP32 products and basis vectors; dot products, norms and the
// Hessenberg arithmetic stay in FP64).
// The host drives a cycle one step at a time: launch(j) enqueues Arnoldi step j (operator, two
// classical Gram-Schmidt passes with the normalisation folded into the second, Hessenberg
// column to pinned host memory); after its event, take_column(j) applies the Givens rotations
// and updates the residual estimates.  The device runs one step ahead of the host except where
// the residual history predicts that the step being waited for converges.  Several Krylov
// objects (the channels of a batch) advance in lockstep on one stream (lockstep_cycle), so one
// host wait serves all of them and the stream always holds every channel's next step.
// With use_graphs each step (j) is captured once into a HIP graph and replayed afterwards.
// The part of a Krylov object the lockstep driver sees.
struct KrylovBase {
    virtual ~KrylovBase() = default;
    bool in_cycle = false;
    int jmax = 0, used = 0;
    int launched = 0;    // steps of the cycle enqueued
    bool hold = false;   // the next step is not launched ahead (predicted to be past convergence)
    std::vector<hipEvent_t> step_ev;
    virtual void advance(int j) = 0;       // enqueue step j (and one ahead)
    virtual void take_column(int j) = 0;   // after step_ev[j]: Givens rotations, convergence
    virtual void finish() = 0;             // enqueue x += V z (the x of begin_cycle)
    // grouped steps: step j's launches as member records of gp, then (after gp has been
    // launched) the Hessenberg column's copy and step_ev[j]
    virtual bool fp32_basis() const = 0;
    virtual hipStream_t stream() const = 0;
    virtual void add_step(int j, GroupPlan &gp) = 0;
    virtual void post_step(int j) = 0;
    // for captured grouped steps: the operator's allocations and layout generation, the
    // application count of a replayed step, and the two halves of post_step
    virtual void prepare_step() = 0;
    virtual uint64_t layout_gen() const = 0;
    virtual void count_step() = 0;
    virtual void copy_column(int j) = 0;
    virtual void record_step(int j) = 0;
};
// Bumped whenever a Krylov object is made or freed (captured grouped steps address their buffers).
static std::atomic<uint64_t> g_krylov_gen{1};

template <typename VT>
struct Krylov : KrylovBase {
    RowOperator &op;
    int r, m;
    int64_t ne;
    bool use_graphs;
    double2 *x_out = nullptr;            // solution the cycle's update is added to
    int nb_dots = 1;                    // row slices of k_dots per vector (at most 64)
    int slot_len;                       // (2(m+1)+1)*r double2 per Arnoldi step
    DevBuf<VT> V;                       // (m+1) basis vectors of ne elements
    DevBuf<double2> zdev, Hd, n2d;
    double2 *Hh = nullptr;              // pinned mirror of Hd
    double2 *zh = nullptr;              // pinned staging of the update coefficients
    double2 *n2h = nullptr;             // pinned copy of n2d (squared norms)
    struct StepGraph {
        hipGraphExec_t exec = nullptr;
        uint64_t gen = 0;
    };
    std::vector<StepGraph> graphs;
    static constexpr bool fp32 = std::is_same<VT, float2>::value;

    // state of the current cycle (host)
    std::vector<cdouble> H, cs, sn, g;
    std::vector<double> ref, prev_rel;  // residual norms are relative to ref[k]
    double tol = 0.0;
    bool conv = false;
    double est = 0.0;                   // largest relative residual estimate at the end of the cycle

    Krylov(RowOperator &o, int r_, int m_, bool graphs_) : op(o), r(r_), m(m_), use_graphs(graphs_) {
        ++g_krylov_gen;
        TT_REQUIRE(r <= 256, "at most 256 simultaneous systems");
        ne = o.dim * r;
        slot_len = (2 * (m + 1) + 1) * r;
        V.alloc((size_t)(m + 1) * ne);
        zdev.alloc((size_t)(m + 1) * r);
        const int tpr = (256 - 256 % r) / r;
        nb_dots = (int)std::min<int64_t>(64, std::max<int64_t>(1, (o.dim + (int64_t)tpr * DOT_UN - 1) / ((int64_t)tpr * DOT_UN)));
        Hd.alloc((size_t)m * slot_len);
        n2d.alloc(r);
        TT_HIP(hipHostMalloc(&Hh, sizeof(double2) * (size_t)m * slot_len, hipHostMallocDefault));
        TT_HIP(hipHostMalloc(&zh, sizeof(double2) * (size_t)(m + 1) * r, hipHostMallocDefault));
        TT_HIP(hipHostMalloc(&n2h, sizeof(double2) * r, hipHostMallocDefault));
        step_ev.resize(m);
        for (auto &e : step_ev) TT_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        graphs.resize(m);
        H.resize((size_t)(m + 1) * m * r);
        cs.resize((size_t)m * r);
        sn.resize((size_t)m * r);
        g.resize((size_t)(m + 1) * r);
        prev_rel.resize(r);
    }
    ~Krylov() {
        ++g_krylov_gen;
        for (auto &gr : graphs) {
            if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
        }
        for (auto e : step_ev) (void)hipEventDestroy(e);
        if (Hh) (void)hipHostFree(Hh);
        if (zh) (void)hipHostFree(zh);
        if (n2h) (void)hipHostFree(n2h);
    }

    // out (zeroed beforehand) += <Vb_i, Wv>_k for i < nvec, on the device
    template <typename T>
    void dots_dev(const T *Vb, int nvec, const T *Wv, double2 *out) {
        int bs = 256 - 256 % r;
        hipLaunchKernelGGL(k_dots<T>, dim3(nb_dots, nvec), dim3(bs), bs * sizeof(double2), op.s, Vb, ne, Wv, ne, r,
                           out);
    }

    // Squared column norms of an FP64 vector into n2d and its pinned copy; read_norms() after
    // the stream has passed (so that several Krylov objects share one wait).
    void enqueue_norms(const double2 *Wv) {
        TT_HIP(hipMemsetAsync(n2d.p, 0, sizeof(double2) * r, op.s));   // eager, never captured
        dots_dev(Wv, 1, Wv, n2d.p);
        TT_HIP(hipMemcpyAsync(n2h, n2d.p, sizeof(double2) * r, hipMemcpyDeviceToHost, op.s));
    }
    // (a NaN squared norm stays NaN, so that a NaN residual or right-hand side is never taken
    // for a converged one)
    std::vector<double> read_norms() const {
        std::vector<double> n(r);
        for (int k = 0; k < r; ++k) n[k] = sqrt_nonneg(n2h[k].x);
        return n;
    }
    std::vector<double> norms(const double2 *Wv) {
        enqueue_norms(Wv);
        TT_HIP(hipStreamSynchronize(op.s));
        return read_norms();
    }

    // The work of Arnoldi step j: v_{j+1} and Hessenberg column j (slot j of Hd, mirrored into Hh).
    // The product's launch time is sampled in the first step of a cycle only: timing events
    // are not free (a few microseconds of gap each).
    void enqueue_step(int j, bool capturing = false) {
        VT *vj = V.p + (size_t)j * ne, *vj1 = V.p + (size_t)(j + 1) * ne;
        double2 *h1 = Hd.p + (size_t)j * slot_len, *h2 = h1 + (m + 1) * r, *n2 = h2 + (m + 1) * r;
        op.zero_next = h1;   // h1, h2 and n2 accumulate by atomics: the step's first kernel zeroes them
        op.zero_next_n = slot_len;
        // the new vector w = M v_j is formed in place, in column j + 1 of the basis
        op.apply(vj, vj1, r, true, fp32, !capturing && j == 0, vj);
        // FP64: two classical Gram-Schmidt passes, the norm of the second folded in (one pass with
        // the Pythagorean norm ||w||^2 - ||h||^2 loses the normalisation of the basis)
        dots_dev(V.p, j + 1, vj1, h1);
        hipLaunchKernelGGL((k_combo<VT, VT>), dim3(nblk(ne, 256)), dim3(256), 0, op.s, vj1, V.p, ne, j + 1, h1, ne, r,
                           -1.0);
        if (fp32) {
            // the FP32 inner solves of the refinement: one classical pass, the norm of the projected
            // vector computed from the vector itself (not from ||w||^2 - ||h||^2, which cancels as
            // the solve converges).  Their basis needs orthogonality only to the inner tolerance
            // (1e-4 .. 1e-6): on the 18-channel batch they take the same steps as with two passes
            // and read the basis twice per step instead of four times.
            dots_dev(vj1, 1, vj1, n2);
            hipLaunchKernelGGL(k_normalize<VT>, dim3(nblk(ne, 256)), dim3(256), 0, op.s, vj1, n2, ne, r);
            TT_HIP(hipMemcpyAsync(Hh + (size_t)j * slot_len, h1, sizeof(double2) * slot_len, hipMemcpyDeviceToHost, op.s));
            return;
        }
        dots_dev(V.p, j + 2, vj1, h2);   // h2[j + 1] = <w, w>
        hipLaunchKernelGGL(k_combo_normalize<VT>, dim3(nblk(ne, 256)), dim3(256), 0, op.s, vj1, V.p, ne, j + 1, h2, ne,
                           r, n2);
        TT_HIP(hipMemcpyAsync(Hh + (size_t)j * slot_len, h1, sizeof(double2) * slot_len, hipMemcpyDeviceToHost, op.s));
    }

    // enqueue_step(j) as member records of a grouped call: the same kernels with the same
    // arguments, each stage launched once for all members of the group
    bool fp32_basis() const override { return fp32; }
    hipStream_t stream() const override { return op.s; }
    void add_step(int j, GroupPlan &gp) override {
        VT *vj = V.p + (size_t)j * ne, *vj1 = V.p + (size_t)(j + 1) * ne;
        double2 *h1 = Hd.p + (size_t)j * slot_len, *h2 = h1 + (m + 1) * r, *n2 = h2 + (m + 1) * r;
        op.zero_next = h1;
        op.zero_next_n = slot_len;
        op.group_apply(vj, vj1, r, true, fp32, vj, gp);
        const int64_t nbv = nblk(ne, 256);
        gp.add(GP_DOTS, fp32, 0, VecMember{V.p, vj1, nullptr, h1, ne, ne, 0.0, r, j + 1, nb_dots},
               (int64_t)nb_dots * (j + 1));
        gp.add(GP_COMBO, fp32, 0, VecMember{V.p, vj1, h1, nullptr, ne, ne, -1.0, r, j + 1, 0}, nbv);
        if (fp32) {
            gp.add(GP_DOTS2, fp32, 0, VecMember{vj1, vj1, nullptr, n2, ne, ne, 0.0, r, 1, nb_dots}, nb_dots);
            gp.add(GP_FINAL, fp32, 0, VecMember{nullptr, vj1, n2, nullptr, 0, ne, 0.0, r, 0, 0}, nbv);
        } else {
            gp.add(GP_DOTS2, fp32, 0, VecMember{V.p, vj1, nullptr, h2, ne, ne, 0.0, r, j + 2, nb_dots},
                   (int64_t)nb_dots * (j + 2));
            gp.add(GP_FINAL, fp32, 0, VecMember{V.p, vj1, h2, n2, ne, ne, 0.0, r, j + 1, 0}, nbv);
        }
    }
    void post_step(int j) override {
        copy_column(j);
        record_step(j);
    }
    void prepare_step() override { op.prepare(r, fp32); }
    uint64_t layout_gen() const override { return op.layout_gen; }
    void count_step() override {
        ++op.n_apply;
        if (fp32) ++op.n_apply32;
    }
    void copy_column(int j) override {
        TT_HIP(hipMemcpyAsync(Hh + (size_t)j * slot_len, Hd.p + (size_t)j * slot_len, sizeof(double2) * slot_len,
                              hipMemcpyDeviceToHost, op.s));
    }
    void record_step(int j) override { TT_HIP(hipEventRecord(step_ev[j], op.s)); }

    // Launch Arnoldi step j (a replay of its graph, captured on first use) and record step_ev[j].
    // The first step of a cycle always runs as plain launches: the product's launch time is
    // sampled there, by events on the stream (event nodes inside a graph add their own gaps).
    void step(int j) {
        if (!use_graphs || j == 0) {
            enqueue_step(j);
            TT_HIP(hipEventRecord(step_ev[j], op.s));
            return;
        }
        StepGraph &gr = graphs[j];
        op.prepare(r, fp32);
        if (!gr.exec || gr.gen != op.layout_gen) {
            if (gr.exec) TT_HIP(hipGraphExecDestroy(gr.exec));
            gr.exec = nullptr;
            const int na = op.n_apply, na32 = op.n_apply32;  // capture runs nothing: count launches below
            TT_HIP(hipStreamBeginCapture(op.s, hipStreamCaptureModeThreadLocal));
            hipGraph_t graph = nullptr;
            try {
                enqueue_step(j, true);
            } catch (...) {
                (void)hipStreamEndCapture(op.s, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            op.n_apply = na;
            op.n_apply32 = na32;
            TT_HIP(hipStreamEndCapture(op.s, &graph));
            hipError_t e = hipGraphInstantiate(&gr.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                gr.exec = nullptr;
                fail(TT_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            }
            gr.gen = op.layout_gen;
        }
        TT_HIP(hipGraphLaunch(gr.exec, op.s));
        TT_HIP(hipEventRecord(step_ev[j], op.s));
        ++op.n_apply;
        if (fp32) ++op.n_apply32;
    }

    // Starts a cycle on the residual R (FP64, device) whose column norms beta are in n2d
    // (enqueue_norms(R) ran last on this object): v_0 = R / beta.  Residual estimates are
    // relative to ref; the cycle stops when all are <= tol (with a margin of 2) or after
    // jmax_ steps.
    void begin_cycle(const double2 *R, const std::vector<double> &beta, const std::vector<double> &ref_, double tol_,
                     int jmax_, double2 *x) {
        x_out = x;
        hipLaunchKernelGGL(k_scale_by_norm<VT>, dim3(nblk(ne, 256)), dim3(256), 0, op.s, V.p, R, n2d.p, ne, r);
        ref = ref_;
        tol = tol_;
        jmax = std::min(jmax_, m);
        std::fill(g.begin(), g.end(), 0.0);
        for (int k = 0; k < r; ++k) {
            g[k] = beta[k];
            prev_rel[k] = beta[k] / std::max(ref[k], 1e-300);
        }
        used = launched = 0;
        hold = false;
        conv = false;
        in_cycle = jmax > 0;
    }

    // Enqueue step j if not yet launched, and step j + 1 ahead unless held back.
    void advance(int j) override {
        if (launched <= j) step(launched++);
        if (!hold && launched == j + 1 && launched < jmax) step(launched++);
    }

    // Hessenberg column j (the step has completed): Givens rotations and residual estimates.
    void take_column(int j) override {
        const double2 *hj = Hh + (size_t)j * slot_len;
        bool conv_all = true, hold_next = true;
        est = 0.0;
        for (int k = 0; k < r; ++k) {
            auto Hk = [&](int i, int jj) -> cdouble & { return H[((size_t)i * m + jj) * r + k]; };
            for (int i = 0; i <= j; ++i) {
                const double2 a1 = hj[(size_t)i * r + k], a2 = hj[(size_t)(m + 1 + i) * r + k];
                Hk(i, j) = cdouble(a1.x + a2.x, a1.y + a2.y);
            }
            Hk(j + 1, j) = sqrt_nonneg(hj[(size_t)2 * (m + 1) * r + k].x);
            for (int i = 0; i < j; ++i) {
                cdouble c = cs[(size_t)i * r + k], sv = sn[(size_t)i * r + k];
                cdouble t0 = std::conj(c) * Hk(i, j) + std::conj(sv) * Hk(i + 1, j);
                cdouble t1 = -sv * Hk(i, j) + c * Hk(i + 1, j);
                Hk(i, j) = t0;
                Hk(i + 1, j) = t1;
            }
            cdouble a = Hk(j, j), bb = Hk(j + 1, j);
            double den = std::sqrt(std::norm(a) + std::norm(bb));
            cdouble c = den > 0 ? a / den : 1.0, sv = den > 0 ? bb / den : 0.0;
            cs[(size_t)j * r + k] = c;
            sn[(size_t)j * r + k] = sv;
            Hk(j, j) = den;
            Hk(j + 1, j) = 0.0;
            cdouble gj = g[(size_t)j * r + k];
            g[(size_t)j * r + k] = std::conj(c) * gj;
            g[(size_t)(j + 1) * r + k] = -sv * gj;
            double rel = std::abs(g[(size_t)(j + 1) * r + k]) / std::max(ref[k], 1e-300);
            est = std::max(est, rel);
            if (rel > 0.5 * tol) conv_all = false;
            // geometric extrapolation of this system's residual to the next step
            const double pred = prev_rel[k] > 0.0 ? rel * std::min(1.0, rel / prev_rel[k]) : 0.0;
            if (pred > 0.5 * tol) hold_next = false;
            prev_rel[k] = rel;
        }
        used = j + 1;
        hold = hold_next;
        if (conv_all) conv = true;
        if (conv_all || used >= jmax) in_cycle = false;
    }

    // x += V z from the triangular least-squares solution of the cycle (enqueued).  Steps
    // launched ahead past the end of the cycle write basis vectors this does not read.
    void finish() override {
        for (int k = 0; k < r; ++k) {
            auto Hk = [&](int i, int jj) { return H[((size_t)i * m + jj) * r + k]; };
            for (int i = used - 1; i >= 0; --i) {
                cdouble sacc = g[(size_t)i * r + k];
                for (int l = i + 1; l < used; ++l)
                    sacc -= Hk(i, l) * cdouble(zh[(size_t)l * r + k].x, zh[(size_t)l * r + k].y);
                const cdouble zi = Hk(i, i) != cdouble(0.0) ? sacc / Hk(i, i) : 0.0;
                zh[(size_t)i * r + k] = make_double2(zi.real(), zi.imag());
            }
        }
        in_cycle = false;
        if (used == 0) return;
        // zh is rewritten by the next cycle only after that cycle's steps, which follow this copy
        TT_HIP(hipMemcpyAsync(zdev.p, zh, sizeof(double2) * (size_t)used * r, hipMemcpyHostToDevice, op.s));
        hipLaunchKernelGGL((k_combo<double2, VT>), dim3(nblk(ne, 256)), dim3(256), 0, op.s, x_out, V.p, ne, used,
                           zdev.p, ne, r, 1.0);
    }
};

// One lockstep cycle over the Krylov objects ks (each begun with begin_cycle): step j of every
// object still in its cycle is enqueued (with one step of lookahead), one wait covers them all,
// then the host takes each column; an object whose cycle ends adds its update to its x at once.
// Returns when every cycle has ended.
static void lockstep_cycle(const std::vector<KrylovBase *> &ks) {
    int jmax = 0;
    for (auto *k : ks) jmax = std::max(jmax, k->in_cycle ? k->jmax : 0);
    for (int j = 0; j < jmax; ++j) {
        bool any = false;
        for (auto *k : ks)
            if (k->in_cycle) {
                k->advance(j);
                any = true;
            }
        if (!any) break;
        for (auto *k : ks)
            if (k->in_cycle) TT_HIP(hipEventSynchronize(k->step_ev[j]));
        for (auto *k : ks) {
            if (!k->in_cycle) continue;
            k->take_column(j);
            if (!k->in_cycle) k->finish();
        }
    }
    for (auto *k : ks)
        if (k->in_cycle) k->finish();
}

// Step j of the Krylov objects ks as grouped calls, one per (stream, basis precision): each
// stage of the step is one launch for the objects of that stream.  With graphs, each such call
// is captured once (its launches and the members' Hessenberg copies, its tables sent once to a
// buffer of its own) and replayed afterwards.
constexpr size_t GROUP_GRAPHS_MAX = 4096;
static void grouped_step_graph(GroupTables &gt, const std::vector<KrylovBase *> &mem, int j, bool f32,
                               hipStream_t s) {
    if (gt.graphs_kgen != g_krylov_gen.load()) {
        gt.clear_graphs();
        gt.graphs_kgen = g_krylov_gen.load();
    }
    for (auto *k : mem) k->prepare_step();   // allocations first: a capture may not allocate
    std::vector<uintptr_t> key{(uintptr_t)s, (uintptr_t)f32, (uintptr_t)j};
    std::vector<uint64_t> gens;
    for (auto *k : mem) {
        key.push_back((uintptr_t)k);
        gens.push_back(k->layout_gen());
    }
    auto it = gt.graphs.find(key);
    if (it != gt.graphs.end() && it->second.gens == gens) {
        for (auto *k : mem) k->count_step();
    } else {
        if (it != gt.graphs.end()) {
            GroupTables::drop(it->second);
            gt.graphs.erase(it);
        }
        if (gt.graphs.size() >= GROUP_GRAPHS_MAX) gt.clear_graphs();
        GroupPlan gp;
        for (auto *k : mem) k->add_step(j, gp);
        gens.clear();
        for (auto *k : mem) gens.push_back(k->layout_gen());
        std::vector<char> bytes;
        std::vector<GroupLaunch> ls;
        const size_t n = pack_group_plan(gp, bytes, ls);
        GroupTables::StepGraph g;
        g.gens = gens;
        TT_HIP(hipMalloc(&g.tab, std::max<size_t>(n, 1)));
        try {
            TT_HIP(hipMemcpyAsync(g.tab, bytes.data(), n, hipMemcpyHostToDevice, s));
            TT_HIP(hipStreamSynchronize(s));   // bytes is pageable and local
            TT_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            hipGraph_t graph = nullptr;
            try {
                launch_group_plan(ls, static_cast<const char *>(g.tab), s);
                for (auto *k : mem) k->copy_column(j);
            } catch (...) {
                (void)hipStreamEndCapture(s, &graph);
                if (graph) (void)hipGraphDestroy(graph);
                throw;
            }
            TT_HIP(hipStreamEndCapture(s, &graph));
            const hipError_t e = hipGraphInstantiate(&g.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (e != hipSuccess) {
                g.exec = nullptr;
                fail(TT_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            }
        } catch (...) {
            GroupTables::drop(g);
            throw;
        }
        it = gt.graphs.emplace(key, g).first;
    }
    TT_HIP(hipGraphLaunch(it->second.exec, s));
    for (auto *k : mem) k->record_step(j);
}
static void grouped_steps(GroupTables &gt, const std::vector<KrylovBase *> &ks, int j, bool graphs) {
    std::vector<hipStream_t> streams;
    for (auto *k : ks)
        if (std::find(streams.begin(), streams.end(), k->stream()) == streams.end()) streams.push_back(k->stream());
    for (hipStream_t s : streams)
        for (int f = 0; f < 2; ++f) {
            std::vector<KrylovBase *> mem;
            for (auto *k : ks)
                if (k->stream() == s && (int)k->fp32_basis() == f) mem.push_back(k);
            if (mem.empty()) continue;
            if (graphs) {
                grouped_step_graph(gt, mem, j, f == 1, s);
                continue;
            }
            GroupPlan gp;
            for (auto *k : mem) k->add_step(j, gp);
            run_group_plan(gt, gp, s);
            for (auto *k : mem) k->post_step(j);
        }
}

// lockstep_cycle with grouped launches: the objects' step j (and the step ahead, for those not
// held back) is one grouped call per stream instead of one launch sequence per object.
static void lockstep_cycle_grouped(const std::vector<KrylovBase *> &ks, GroupTables &gt, bool graphs) {
    int jmax = 0;
    for (auto *k : ks) jmax = std::max(jmax, k->in_cycle ? k->jmax : 0);
    for (int j = 0; j < jmax; ++j) {
        std::vector<KrylovBase *> now, ahead;
        for (auto *k : ks)
            if (k->in_cycle && k->launched <= j) now.push_back(k);
        grouped_steps(gt, now, j, graphs);
        for (auto *k : now) k->launched = j + 1;
        bool any = false;
        for (auto *k : ks)
            if (k->in_cycle) {
                any = true;
                if (!k->hold && k->launched == j + 1 && k->launched < k->jmax) ahead.push_back(k);
            }
        if (!any) break;
        grouped_steps(gt, ahead, j + 1, graphs);
        for (auto *k : ahead) k->launched = j + 2;
        for (auto *k : ks)
            if (k->in_cycle) TT_HIP(hipEventSynchronize(k->step_ev[j]));
        for (auto *k : ks) {
            if (!k->in_cycle) continue;
            k->take_column(j);
            if (!k->in_cycle) k->finish();
        }
    }
    for (auto *k : ks)
        if (k->in_cycle) k->finish();
}

// ------------------------------------------------------------------ solve driver
struct OnshellLayout {
    int nd, nE, r;
    std::vector<int64_t> row_idx;   // r rows: (d_row, j) -> idx of (alpha_d, q_j, 0)
    std::vector<int> row_g;         // energy of each row
};

static OnshellLayout onshell_layout(const tt_channel_os &os, int Nq, int Np) {
    OnshellLayout L;
    L.nd = os.num_deuteron_states;
    L.nE = os.num_T_lab;
    L.r = L.nd * L.nE;
    for (int d = 0; d < L.nd; ++d)
        for (int j = 0; j < L.nE; ++j) {
            L.row_idx.push_back(((int64_t)os.deuteron_idx[d] * Nq + os.q_com_idx[j]) * Np);
            L.row_g.push_back(j);
        }
    return L;
}

// Flat state index of breakup cell b as the reference forms it (CPP/solve_faddeev.cpp:879):
// a neighbour one bin outside the grid lands on another state or outside the state space.
static inline int64_t breakup_flat(const tt_channel_os &os, int b, int Nq, int Np) {
    const int *t = os.alphapq_idx + 3 * (size_t)b;
    return (int64_t)t[0] * Nq * Np + (int64_t)t[1] * Np + t[2];
}

// Elements read from the row solutions x[state][k] (k = d_row * nE + j): first the elastic
// U[dr][dc][j] = x[(alpha_dc, q_j, 0)][dr * nE + j], then with breakup
// U_BU[dr][b] = x[(alpha_b, q_b, p_b)][dr * nE + j(b)] (CPP/solve_faddeev.cpp:867-891), the
// state's flat index alpha * Nq * Np + q * Np + p as the reference forms it.  A neighbour cell
// whose flat index falls outside the state space (where the reference reads outside the row)
// is left out; bu_pos gets the U_BU position (dr * num_BU_chns + b) of each breakup element
// listed, and the positions left out are set to NaN by the callers.
static void readout_list(const tt_channel_os &os, int Nq, int Np, int64_t dim, bool breakup,
                         std::vector<int64_t> &idx, std::vector<int> &kk, std::vector<int> *bu_pos = nullptr) {
    const int nd = os.num_deuteron_states, nE = os.num_T_lab;
    idx.clear();
    kk.clear();
    for (int dr = 0; dr < nd; ++dr)
        for (int dc = 0; dc < nd; ++dc)
            for (int j = 0; j < nE; ++j) {
                idx.push_back(((int64_t)os.deuteron_idx[dc] * Nq + os.q_com_idx[j]) * Np);
                kk.push_back(dr * nE + j);
            }
    if (bu_pos) bu_pos->clear();
    if (!breakup) return;
    TT_REQUIRE(os.q_com_BU_idx && os.alphapq_idx, "breakup cells missing (tt_find_breakup_bins)");
    for (int dr = 0; dr < nd; ++dr)
        for (int j = 0; j < nE; ++j)
            for (int b = os.q_com_BU_idx[j]; b < os.q_com_BU_idx[j + 1]; ++b) {
                const int64_t flat = breakup_flat(os, b, Nq, Np);
                if (flat < 0 || flat >= dim) continue;
                idx.push_back(flat);
                kk.push_back(dr * nE + j);
                if (bu_pos) bu_pos->push_back(dr * os.num_BU_chns + b);
            }
}

static inline bool wants_breakup(const tt_channel_os &os, const double *U_BU) {
    return U_BU && os.num_BU_chns > 0;
}

// values of the readout list of rows (elastic, then breakup) -> out
static void gather_onshell(RowOperator &op, const double2 *rows, const OnshellLayout &L, const tt_channel_os &os,
                           std::vector<cdouble> &out, bool breakup = false, std::vector<int> *bu_pos = nullptr) {
    std::vector<int64_t> idx;
    std::vector<int> kk;
    readout_list(os, op.Nq, op.Np, op.dim, breakup, idx, kk, bu_pos);
    int n = (int)idx.size();
    DevBuf<int64_t> di;
    DevBuf<int> dk;
    DevBuf<double2> dout(n);
    di.upload(idx.data(), n, op.s);
    dk.upload(kk.data(), n, op.s);
    hipLaunchKernelGGL(k_gather, dim3(nblk(n, 64)), dim3(64), 0, op.s, rows, di.p, dk.p, dout.p, n, L.r);
    std::vector<double2> h(n);
    TT_HIP(hipMemcpyAsync(h.data(), dout.p, sizeof(double2) * n, hipMemcpyDeviceToHost, op.s));
    TT_HIP(hipStreamSynchronize(op.s));
    out.resize(n);
    for (int i = 0; i < n; ++i) out[i] = cdouble(h[i].x, h[i].y);
}

// Bout = the on-shell rows e_k^T A (k = (d_row, j)); with scratch (dim * r elements) and idx the
// unit rows are formed there instead of in buffers of their own (h_idx: what idx holds).
static void on_shell_rows_of_A(RowOperator &op, const OnshellLayout &L, double2 *Bout, double2 *scratch = nullptr,
                               DevBuf<int64_t> *idx = nullptr, std::vector<int64_t> *h_idx = nullptr) {
    DevBuf<double2> own;
    DevBuf<int64_t> own_idx;
    if (!scratch) {
        own.alloc((size_t)op.dim * L.r);
        scratch = own.p;
    }
    if (!idx) idx = &own_idx;
    TT_HIP(hipMemsetAsync(scratch, 0, sizeof(double2) * (size_t)op.dim * L.r, op.s));
    if (h_idx)
        upload_if_changed(*idx, *h_idx, L.row_idx.data(), (size_t)L.r, op.s);
    else
        idx->upload(L.row_idx.data(), L.r, op.s);
    hipLaunchKernelGGL(k_unit_rows, dim3(nblk(L.r, 64)), dim3(64), 0, op.s, scratch, idx->p, L.r);
    // The unit rows, and so the input of the cell product, are zero outside the on-shell q
    // bins: only the cells reading those bins take part (untimed, as it moves fewer bytes than
    // a full product)
    bool timed = true;
    if (op.cells && !op.col) {
        op.prepare(L.r, false);
        std::vector<int> qb;
        for (int64_t ri : L.row_idx) qb.push_back((int)((ri / op.Np) % op.Nq));
        std::sort(qb.begin(), qb.end());
        qb.erase(std::unique(qb.begin(), qb.end()), qb.end());
        op.next_items = op.items_on_q(qb, op.choice(false, 2 * L.r).len);
        timed = op.next_items == nullptr;
    }
    op.apply(scratch, Bout, L.r, false, false, timed);
    if (own.p) TT_HIP(hipStreamSynchronize(op.s));   // own buffers are freed on return
}

// Elastic and breakup elements alike get Neumann terms and Pade selection; the series runs
// until every element has converged.  (The reference stops once the count of converged elastic
// plus breakup elements equals the elastic count, lines 920-924, which can leave elastic
// elements unselected; here all of them are selected.)
static void solve_pade(RowOperator &op, const tt_channel_os &os, const tt_solve_opts &o, cdouble *U,
                       tt_solve_stats &st, cdouble *U_BU = nullptr) {
    OnshellLayout L = onshell_layout(os, op.Nq, op.Np);
    const bool bu = U_BU && os.num_BU_chns > 0;
    const int n_el = L.nd * L.nd * L.nE;
    std::vector<int> bu_pos;
    {
        std::vector<int64_t> idx;
        std::vector<int> kk;
        readout_list(os, op.Nq, op.Np, op.dim, bu, idx, kk, &bu_pos);
    }
    const int nel = n_el + (int)bu_pos.size();
    const int NM_max = o.NM_max > 0 ? o.NM_max : 14;
    const int nterms = 2 * NM_max + 1;
    op.set_gidx(L.row_g);
    DevBuf<double2> rows((size_t)op.dim * L.r), rows2((size_t)op.dim * L.r);
    on_shell_rows_of_A(op, L, rows.p);
    std::vector<cdouble> a((size_t)nel * nterms, 0.0), an;
    gather_onshell(op, rows.p, L, os, an, bu);
    for (int e = 0; e < nel; ++e) a[(size_t)e * nterms] = an[e];
    std::vector<cdouble> PA((size_t)nel * (NM_max + 1), 0.0);
    std::vector<int> best_idx(nel, 0);
    std::vector<char> conv(nel, 0);
    int nconv = 0, n_done = 0;
    for (int NM = 0; NM <= NM_max; ++NM) {
        if (nconv == nel) break;
        for (int n = 2 * NM - 1; n < 2 * NM + 1; ++n) {
            if (n <= 0) continue;
            op.apply(rows.p, rows2.p, L.r, true, false);
            std::swap(rows.p, rows2.p);
            gather_onshell(op, rows.p, L, os, an, bu);
            for (int e = 0; e < nel; ++e)
                if (!conv[e]) a[(size_t)e * nterms + n] = an[e];
            n_done = n;
        }
        for (int e = 0; e < nel; ++e) {
            if (conv[e]) continue;
            cdouble *PAe = &PA[(size_t)e * (NM_max + 1)];
            PAe[NM] = pade_approximant(&a[(size_t)e * nterms], NM, NM, 1.0);
            // selection rules of lines 1255-1295; the approximant "before" NM = 0 is taken as 0
            auto prev = [&](int k) { return k >= 0 ? PAe[k] : cdouble(0.0); };
            int best = 0;
            double mind = 1.0;
            for (int k = 0; k < NM; ++k) {
                double d = std::abs(PAe[k] - prev(k - 1));
                if (d < mind && d > 1e-15) {
                    best = k;
                    mind = d;
                }
            }
            double d = std::abs(PAe[NM] - prev(NM - 1));
            if (d < mind && d > 1e-15) {
                best = NM;
                mind = d;
            }
            bool c0 = NM == NM_max, c1 = NM - best > 4, c2 = mind < 1e-6 * std::abs(PAe[best]), c3 = mind < 1e-7;
            if (c0 || c1 || c2 || c3) {
                conv[e] = 1;
                best_idx[e] = best;
                ++nconv;
            }
        }
    }
    for (int e = 0; e < n_el; ++e) U[e] = PA[(size_t)e * (NM_max + 1) + best_idx[e]];
    if (bu) {
        std::fill(U_BU, U_BU + (size_t)L.nd * os.num_BU_chns, cdouble(NAN, NAN));
        for (int e = n_el; e < nel; ++e) U_BU[bu_pos[e - n_el]] = PA[(size_t)e * (NM_max + 1) + best_idx[e]];
    }
    if (o.neumann_terms) std::memcpy(o.neumann_terms, a.data(), sizeof(cdouble) * n_el * nterms);
    st.iterations = n_done + 1;
}

// G[j][idx] at the total energies E[0..nE) (host), written to device memory dG (nE*dim);
// with partials, the three partial-derivative arrays of k_resolvent_partials (3*nE*dim).
// With stage the inputs are uploaded into the caller's buffers, which outlive the kernel (no
// wait); without it into buffers of this call, and the call waits for the kernel.
static void resolvent_on_device(hipStream_t s, const tt_run_params &rp, const tt_swp_states &swp,
                                const tt_pw_states &chn, const double *E, int nE, double2 *dG,
                                bool partials = false, ResolventStaging *stage = nullptr) {
    std::vector<int> which, offset;
    swp_edge_offsets(chn, swp, rp, which, offset);
    const int Na = chn.Nalpha, Nq = swp.Nq_WP, Np = swp.Np_WP;
    const int64_t dim = (int64_t)Na * Nq * Np;
    const size_t nu = (size_t)swp.num_2N_unco_states * (Np + 1), nc = (size_t)swp.num_2N_coup_states * 2 * (Np + 1);
    ResolventStaging own;
    ResolventStaging &st = stage ? *stage : own;
    DevBuf<double> &eu = st.eu, &ec = st.ec, &qe = st.qe, &dE = st.E;
    DevBuf<int> &dw = st.which, &doff = st.offset;
    ResolventStaging::put(eu, st.h_eu, swp.e_SWP_unco, std::max<size_t>(nu, 1), s);
    if (nc) ResolventStaging::put(ec, st.h_ec, swp.e_SWP_coup, nc, s);
    ResolventStaging::put(qe, st.h_qe, swp.q_WP, (size_t)Nq + 1, s);
    dE.upload(E, nE, s);
    ResolventStaging::put(dw, st.h_which, which.data(), (size_t)Na, s);
    ResolventStaging::put(doff, st.h_offset, offset.data(), (size_t)Na, s);
    if (partials)
        hipLaunchKernelGGL(k_resolvent_partials, dim3(nblk(dim * nE, 256)), dim3(256), 0, s, dG, eu.p, ec.p, dw.p,
                           doff.p, qe.p, dE.p, Na, Nq, Np, nE);
    else
        hipLaunchKernelGGL(k_resolvent, dim3(nblk(dim * nE, 256)), dim3(256), 0, s, dG, eu.p, ec.p, dw.p, doff.p,
                           qe.p, dE.p, Na, Nq, Np, nE);
    TT_HIP(hipGetLastError());
    // the call's own staging buffers are freed on return: let the kernel finish with them first
    if (!stage) TT_HIP(hipStreamSynchronize(s));
}

// ------------------------------------------------------------------ solve plans
// Everything a channel's solves reuse: the operator (V.C and C^T blocks, work items, P123
// in cell form), the Krylov workspace with its captured step graphs, the right-hand-side and
// solution vectors.  Cached per context, keyed by the P123 object and a hash of the inputs
// the operator is built from; only G (the energies) changes between reuses.
struct SolvePlan {
    uint64_t p_gen = 0;
    std::vector<int> key;   // plan_struct_key
    ContentRef content;     // the V and C its blocks hold
    int format = TT_P123_CELLS;
    bool col = false;   // column operator (adjoint solves)
    uint64_t last_use = 0;
    RowOperator op;
    std::unique_ptr<Krylov<double2>> k64;   // FP64 GMRES
    std::unique_ptr<Krylov<float2>> k32;    // inner solves of iterative refinement
    DevBuf<double2> b, x, rres;
    // on-shell index lists and the gathered U elements (pinned host copy)
    DevBuf<int64_t> unit_idx, os_idx;
    DevBuf<int> os_k;
    std::vector<int64_t> h_os_idx, h_unit_idx;   // what os_idx, os_k and unit_idx hold
    std::vector<int> h_os_k;
    DevBuf<double2> os_out;
    double2 *os_out_h = nullptr;
    size_t os_out_cap = 0;
    int os_out_n = 0;      // elements of the last enqueue_onshell
    int os_bu_total = 0;   // its U_BU size and the U_BU position of each breakup element
    std::vector<int> os_bu_pos;
    uint64_t in_use = 0;   // mark of the lockstep solve holding this plan
    // work buffers of the gradient assembly (row plans of tt_solve_gradient)
    struct GradWork {
        DevBuf<int64_t> unit;
        DevBuf<double2> z, w, Uw, bz, Dw, Dz, part, dG, Gp, tri;
        DevBuf<int4> pairs;
        DevBuf<int> which, offset, ints;
        DevBuf<double> dbls;
    };
    std::unique_ptr<GradWork> gw;

    ~SolvePlan() {
        if (os_out_h) (void)hipHostFree(os_out_h);
    }

    template <typename VT>
    Krylov<VT> &krylov(int r, int m, bool graphs) {
        std::unique_ptr<Krylov<VT>> *k;
        if constexpr (std::is_same<VT, double2>::value)
            k = &k64;
        else
            k = &k32;
        if (!*k || (*k)->r != r || (*k)->m != m || (*k)->use_graphs != graphs) {
            k->reset();
            k->reset(new Krylov<VT>(op, r, m, graphs));
        }
        return **k;
    }

    // b, x and the residual vector hold dim * r elements
    void fit(int r) {
        const size_t ne = (size_t)op.dim * r;
        for (DevBuf<double2> *v : {&b, &x, &rres})
            if (v->n < ne) v->alloc(ne);
    }

    // Enqueue the on-shell elements of x (rows (d_row, j), columns (alpha_dc, q_j, 0)) into
    // pinned host memory, read by read_onshell once the stream has passed:
    // out[dr][dc][j] = x[(alpha_dc, q_j, 0)][dr * nE + j].
    // With breakup the breakup elements follow (readout_list).
    void enqueue_onshell(const double2 *xv, const OnshellLayout &L, const tt_channel_os &os, bool breakup = false) {
        std::vector<int64_t> idx;
        std::vector<int> kk;
        readout_list(os, op.Nq, op.Np, op.dim, breakup, idx, kk, &os_bu_pos);
        const int n = (int)idx.size();
        os_out_n = n;
        os_bu_total = breakup ? L.nd * os.num_BU_chns : 0;
        if ((size_t)n > os_out_cap) {
            if (os_out_h) TT_HIP(hipHostFree(os_out_h));
            os_out_h = nullptr;
            TT_HIP(hipHostMalloc(&os_out_h, sizeof(double2) * n, hipHostMallocDefault));
            os_out_cap = n;
            os_out.alloc(n);
        }
        upload_if_changed(os_idx, h_os_idx, idx.data(), (size_t)n, op.s);
        upload_if_changed(os_k, h_os_k, kk.data(), (size_t)n, op.s);
        hipLaunchKernelGGL(k_gather, dim3(nblk(n, 64)), dim3(64), 0, op.s, xv, os_idx.p, os_k.p, os_out.p, n, L.r);
        TT_HIP(hipMemcpyAsync(os_out_h, os_out.p, sizeof(double2) * n, hipMemcpyDeviceToHost, op.s));
    }
    void read_onshell(const OnshellLayout &L, cdouble *U, cdouble *U_BU = nullptr) const {
        const int n = L.nd * L.nd * L.nE;
        for (int i = 0; i < n; ++i) U[i] = cdouble(os_out_h[i].x, os_out_h[i].y);
        if (U_BU) {
            std::fill(U_BU, U_BU + os_bu_total, cdouble(NAN, NAN));
            for (int i = n; i < os_out_n; ++i) U_BU[os_bu_pos[i - n]] = cdouble(os_out_h[i].x, os_out_h[i].y);
        }
    }
};

struct PlanCache {
    std::vector<std::unique_ptr<SolvePlan>> plans;
    uint64_t tick = 0;
    static constexpr size_t max_plans = 64;
    VCProducts vc;   // of the last V and C seen: shared by every channel of a batch
};

void free_plan_cache(PlanCache *c) { delete c; }

// Structure of a plan: everything that fixes the operator's layout (flags, sizes, the
// channel's quantum numbers), compared word for word.  Its contents (V and the SWP
// eigenvectors) are compared apart (Content).
static std::vector<int> plan_struct_key(const tt_run_params &rp, const tt_swp_states &swp, const tt_pw_states &chn,
                                        int format) {
    std::vector<int> k = {rp.tensor_force, rp.isospin_breaking_1S0, swp.Np_WP, swp.Nq_WP, chn.Nalpha, format};
    const int *arrs[] = {chn.L_2N, chn.S_2N, chn.J_2N, chn.T_2N, chn.L_1N, chn.two_J_1N, chn.two_J_3N,
                         chn.two_T_3N, chn.P_3N};
    for (const int *a : arrs) k.insert(k.end(), a, a + chn.Nalpha);
    return k;
}

static ContentRef plan_content(const tt_run_params &rp, const double *V_unco, const double *V_coup,
                               const tt_swp_states &swp) {
    int nu, nc;
    num_2N_states(rp, nu, nc);
    const size_t Np = (size_t)swp.Np_WP, nu2 = nu * Np * Np, nc2 = nc * 4 * Np * Np;
    std::shared_ptr<Content> c(new Content);
    std::vector<double> &w = c->words;
    w.reserve(2 * (nu2 + nc2));
    w.insert(w.end(), V_unco, V_unco + nu2);
    if (nc) w.insert(w.end(), V_coup, V_coup + nc2);
    w.insert(w.end(), swp.C_SWP_unco, swp.C_SWP_unco + nu2);
    if (nc) w.insert(w.end(), swp.C_SWP_coup, swp.C_SWP_coup + nc2);
    uint64_t h = mix64(w.size());
    for (size_t i = 0; i < w.size(); ++i) {
        uint64_t b;
        std::memcpy(&b, &w[i], 8);
        h = mix64(h ^ mix64(b + 0x9e3779b97f4a7c15ull * (i + 1)));
    }
    c->hash = h;
    return c;
}

// The context's plan for these inputs: found by the P123 object and the structure, its block
// contents refreshed when V or C changed (parameter sweeps, gradients), else built (evicting
// the least recently used plan beyond max_plans).
// content_key: plan_content of V and C when the caller has it (one per batch call).
static SolvePlan &get_plan(tt_ctx *ctx, const tt_run_params &rp, const DeviceCSC *P, const double *V_unco,
                           const double *V_coup, const tt_swp_states &swp, const tt_pw_states &chn, int format,
                           bool col = false, const ContentRef *content_key = nullptr) {
    if (!ctx->plans) ctx->plans = new PlanCache;
    PlanCache &c = *ctx->plans;
    const std::vector<int> key = plan_struct_key(rp, swp, chn, format);
    const ContentRef content = content_key ? *content_key : plan_content(rp, V_unco, V_coup, swp);
    auto products = [&]() -> VCProducts & {
        if (!c.vc.valid || !same_content(c.vc.content, content)) {
            c.vc.valid = false;
            c.vc.dev_valid = false;
            vc_products(rp, V_unco, V_coup, swp, c.vc);
            c.vc.content = content;
        }
        return c.vc;
    };
    for (auto &pl : c.plans)
        if (pl->p_gen == P->gen && pl->key == key && pl->format == format && pl->col == col) {
            pl->last_use = ++c.tick;
            if (!same_content(pl->content, content)) {
                pl->op.set_blocks(rp, V_unco, V_coup, swp, chn, &products());
                pl->content = content;
            }
            return *pl;
        }
    if (c.plans.size() >= PlanCache::max_plans) {
        // the least recently used plan that no running lockstep group holds
        auto lru = c.plans.end();
        for (auto it = c.plans.begin(); it != c.plans.end(); ++it)
            if (!(*it)->in_use && (lru == c.plans.end() || (*it)->last_use < (*lru)->last_use)) lru = it;
        if (lru != c.plans.end()) {
            TT_HIP(hipStreamSynchronize(ctx->stream));
            c.plans.erase(lru);
        }
    }
    std::unique_ptr<SolvePlan> pl(new SolvePlan);
    pl->op.setup(ctx->stream, rp, P, V_unco, V_coup, swp, chn, format, col, &products());
    pl->p_gen = P->gen;
    pl->key = key;
    pl->content = content;
    pl->format = format;
    pl->col = col;
    pl->last_use = ++c.tick;
    c.plans.push_back(std::move(pl));
    return *c.plans.back();
}

// One channel of a lockstep Krylov solve: its plan, on-shell layout and right-hand sides b
// (the on-shell rows of its operator), solution x and residual in the plan's buffers.
// mixed: iterative refinement (FP64 residual, FP32 inner GMRES) instead of FP64 GMRES.
struct KMember {
    SolvePlan *pl = nullptr;
    const tt_channel_os *os = nullptr;
    bool mixed = false;
    bool breakup = false;   // also read the breakup elements of the solution
    OnshellLayout L;
    std::vector<double> bnorm;
    double res = 0.0;
    int iters = 0;
    bool done = false;
};

// Every member solved to o.tol, all members advancing together on the one stream of their
// plans: every host wait (residual norms, Arnoldi steps) serves all members.  An outer pass is
// one cycle per member (a restart cycle of FP64 GMRES, or the inner FP32 solve of a refinement
// step) followed by the true FP64 residual.  G must be in each operator; the solutions stay in
// the plans' x and their on-shell elements are fetched (read_onshell) before returning.
// With gt, each stage of an Arnoldi step, and of the FP64 residual application, is one grouped
// launch for the members of one stream (tables in gt).
constexpr int GRAPH_MIN_MEMBERS = 4;
static void solve_lockstep(std::vector<KMember> &ms, const tt_solve_opts &o, GroupTables *gt = nullptr) {
    if (ms.empty()) return;
    const int m = o.restart > 0 ? o.restart : 40;
    const size_t n = ms.size();
    // automatic (use_graphs < 0): replayed step graphs for groups of GRAPH_MIN_MEMBERS or more,
    // where the host's launch rate limits the lanes (measured on the 18-channel batch: +5-8%);
    // plain launches for smaller groups, where a graph launch costs more than its kernels' own
    // launches (one channel: -9%)
    const bool want_graphs = o.use_graphs > 0 || (o.use_graphs < 0 && n >= (size_t)GRAPH_MIN_MEMBERS);
    const bool graphs = !gt && want_graphs;   // the members' own step graphs
    const bool group_graphs = gt && want_graphs;   // captured grouped steps
    // the members' streams (one, or the context's lanes): every host wait covers all of them
    std::vector<hipStream_t> streams;
    for (const auto &mb : ms)
        if (std::find(streams.begin(), streams.end(), mb.pl->op.s) == streams.end()) streams.push_back(mb.pl->op.s);
    auto sync_all = [&]() {
        for (hipStream_t st : streams) TT_HIP(hipStreamSynchronize(st));
    };
    std::vector<Krylov<double2> *> k64(n, nullptr);
    std::vector<Krylov<float2> *> k32(n, nullptr);
    auto enqueue_norms = [&](size_t i, const double2 *v) {
        if (ms[i].mixed)
            k32[i]->enqueue_norms(v);
        else
            k64[i]->enqueue_norms(v);
    };
    auto read_norms = [&](size_t i) { return ms[i].mixed ? k32[i]->read_norms() : k64[i]->read_norms(); };
    for (size_t i = 0; i < n; ++i) {
        KMember &mb = ms[i];
        SolvePlan &pl = *mb.pl;
        mb.L = onshell_layout(*mb.os, pl.op.Nq, pl.op.Np);
        const int r = mb.L.r;
        pl.op.set_gidx(mb.L.row_g);
        pl.fit(r);
        on_shell_rows_of_A(pl.op, mb.L, pl.b.p, pl.rres.p, &pl.unit_idx, &pl.h_unit_idx);
        TT_HIP(hipMemsetAsync(pl.x.p, 0, sizeof(double2) * (size_t)pl.op.dim * r, pl.op.s));
        if (mb.mixed)
            k32[i] = &pl.krylov<float2>(r, m, graphs);
        else
            k64[i] = &pl.krylov<double2>(r, m, graphs);
        enqueue_norms(i, pl.b.p);
        mb.res = 1.0;
        mb.iters = 0;
        mb.done = false;
    }
    sync_all();
    for (size_t i = 0; i < n; ++i) ms[i].bnorm = read_norms(i);
    // largest relative residual over the systems; NaN when any ratio is not finite (a NaN or
    // infinite norm), which no tolerance accepts
    auto rel_max = [](const std::vector<double> &a, const std::vector<double> &ref) {
        double v = 0.0;
        for (size_t k = 0; k < a.size(); ++k) {
            const double q = a[k] / (ref[k] > 1e-300 ? ref[k] : 1e-300);
            if (!std::isfinite(q)) return std::numeric_limits<double>::quiet_NaN();
            v = std::max(v, q);
        }
        return v;
    };
    std::vector<std::vector<double>> beta(n);
    for (int outer = 0;; ++outer) {
        // members still solving; the first pass takes R = b (x = 0), whose norms are in hand
        std::vector<size_t> act;
        for (size_t i = 0; i < n; ++i)
            if (!ms[i].done) act.push_back(i);
        if (act.empty()) break;
        if (outer == 0) {
            for (size_t i : act) beta[i] = ms[i].bnorm;
        } else {
            if (gt)   // the members' FP64 applications M x as one grouped call per stream
                for (hipStream_t st : streams) {
                    GroupPlan gp;
                    for (size_t i : act)
                        if (ms[i].pl->op.s == st)
                            ms[i].pl->op.group_apply(ms[i].pl->x.p, ms[i].pl->rres.p, ms[i].L.r, true, false,
                                                     (const double2 *)ms[i].pl->x.p, gp);
                    run_group_plan(*gt, gp, st);
                }
            for (size_t i : act) {   // R = b - M x (FP64) and its column norms
                SolvePlan &pl = *ms[i].pl;
                const int64_t ne = pl.op.dim * ms[i].L.r;
                if (!gt) pl.op.apply(pl.x.p, pl.rres.p, ms[i].L.r, true, false, true, pl.x.p);
                hipLaunchKernelGGL(k_axpby, dim3(nblk(ne, 256)), dim3(256), 0, pl.op.s, pl.rres.p, pl.b.p, pl.rres.p,
                                   1.0, -1.0, ne);
                enqueue_norms(i, pl.rres.p);
            }
            sync_all();
            for (size_t i : act) beta[i] = read_norms(i);
        }
        std::vector<KrylovBase *> ks;
        for (size_t i : act) {
            KMember &mb = ms[i];
            mb.res = rel_max(beta[i], mb.bnorm);
            if (mb.res <= o.tol || std::isnan(mb.res) || mb.iters >= o.max_iter || (mb.mixed && outer >= 20)) {
                mb.done = true;
                continue;
            }
            SolvePlan &pl = *mb.pl;
            const double2 *R0 = outer == 0 ? pl.b.p : pl.rres.p;
            if (!mb.mixed) {
                k64[i]->begin_cycle(R0, beta[i], mb.bnorm, o.tol, o.max_iter - mb.iters, pl.x.p);
                ks.push_back(k64[i]);
            } else {   // inner FP32 GMRES on the residual, to a tolerance relative to it
                const double inner_tol = std::min(1e-4, std::max(1e-6, o.tol / std::max(mb.res, 1e-300)));
                k32[i]->begin_cycle(R0, beta[i], beta[i], inner_tol, o.max_iter - mb.iters, pl.x.p);
                ks.push_back(k32[i]);
            }
        }
        if (ks.empty()) break;
        if (gt)
            lockstep_cycle_grouped(ks, *gt, group_graphs);
        else
            lockstep_cycle(ks);
        for (size_t i : act)
            if (!ms[i].done) ms[i].iters += ms[i].mixed ? k32[i]->used : k64[i]->used;
    }
    for (auto &mb : ms) mb.pl->enqueue_onshell(mb.pl->x.p, mb.L, *mb.os, mb.breakup);
    sync_all();
}

// solve_lockstep with the members spread over the context's stream lanes (round robin), so that
// independent channels' kernels overlap on the device.  Every lane first waits for the work
// queued on the context stream (G, the set-up); afterwards the lanes are drained and the plans
// go back to the context stream, also when the solve throws.
// Lanes: TT_LANES (1..16) if set, else 6.  With HIP's default of four hardware queues per
// process, the streams share them; measured on the 18-channel batch with plain launches, 8 lanes
// gave 2.7% more than 4 (16 the same as 8, and 8 lanes on 8 queues less than 4 on 4); with the
// replayed step graphs, 6 lanes gave 1.5-3% more than 5, 7 or 8 (alternating runs on one box).
static int lanes_wanted() {   // read at every batch call
    const char *e = std::getenv("TT_LANES");
    const int v = e ? std::atoi(e) : 0;
    return v >= 1 && v <= 16 ? v : 6;
}
// TT_GROUPED=1: every stage of the steps is one grouped launch for the members of each lane
// (solve_lockstep with tables; TT_LANES=1 puts all members in one group), where all of them
// take the cell-block product and the split-half form is off.
static bool grouped_wanted() {
    const char *e = std::getenv("TT_GROUPED");
    return e && std::atoi(e) == 1;
}
static void solve_on_lanes(tt_ctx *ctx, std::vector<KMember> &ms, const tt_solve_opts &o) {
    GroupTables *gt = nullptr;
    if (grouped_wanted() && ms.size() > 1 && cell_h_mode() == 0 &&
        std::all_of(ms.begin(), ms.end(), [](const KMember &mb) { return mb.pl->op.cells; })) {
        if (!ctx->groups) ctx->groups = new GroupTables;
        gt = ctx->groups;
    }
    const int L = std::min<int>((int)ms.size(), lanes_wanted());
    if (L <= 1) {
        solve_lockstep(ms, o, gt);
        return;
    }
    while ((int)ctx->lanes.size() < L) {
        hipStream_t st = nullptr;
        TT_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        ctx->lanes.push_back(st);
    }
    hipEvent_t ready = nullptr;
    TT_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
    struct Restore {
        tt_ctx *c;
        std::vector<KMember> *m;
        hipEvent_t ev;
        int L;
        ~Restore() {
            for (int l = 0; l < L; ++l) (void)hipStreamSynchronize(c->lanes[l]);
            for (auto &mb : *m) mb.pl->op.s = c->stream;
            (void)hipEventDestroy(ev);
        }
    } restore{ctx, &ms, ready, L};
    TT_HIP(hipEventRecord(ready, ctx->stream));
    for (int l = 0; l < L; ++l) TT_HIP(hipStreamWaitEvent(ctx->lanes[l], ready, 0));
    for (size_t i = 0; i < ms.size(); ++i) ms[i].pl->op.s = ctx->lanes[i % L];
    solve_lockstep(ms, o, gt);
}

// M[c][r0 + k] (column-major, leading dim dim) = delta - A[r0 + k][c] G[c] from the rows
// r0 .. r0 + rc - 1 of A, Y[c][k] (RHS-fastest: Y = E A for unit rows E)
__global__ void k_dense_rows(const double2 *__restrict__ Y, int rc, int64_t r0, int64_t dim,
                             const double2 *__restrict__ G, double2 *__restrict__ M) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dim * rc) return;
    const int64_t c = t / rc, k = t - c * rc, r = r0 + k;
    const double2 y = Y[t], g = G[c];
    M[c * dim + r] = make_double2((r == c ? 1.0 : 0.0) - (y.x * g.x - y.y * g.y), -(y.x * g.y + y.y * g.x));
}

// B[q][r0 + k] = A[r0 + k][cols[q]] (the right-hand-side columns of A) from the same rows
__global__ void k_dense_rhs(const double2 *__restrict__ Y, int rc, int64_t r0, int64_t dim,
                            const int64_t *__restrict__ cols, int nrhs, double2 *__restrict__ B) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nrhs * rc) return;
    const int q = (int)(t / rc), k = (int)(t - (int64_t)q * rc);
    B[(int64_t)q * dim + r0 + k] = Y[cols[q] * rc + k];
}

// Dense path (CPP/solve_faddeev.cpp:305-438): for every on-shell energy j, L = 1 - A diag(G_j)
// is formed on the device from the rows of A (dim unit rows through the operator, in chunks),
// with the on-shell columns of A (and, with U_BU, the breakup states' columns) as right-hand
// sides, and solved by the device LU (dense_lu_solve).  The reference solves for all dim columns
// of A; only the columns whose elements are read are solved here.
static void solve_dense(RowOperator &op, const tt_channel_os &os, cdouble *U, tt_solve_stats &st,
                        cdouble *U_BU = nullptr) {
    const int64_t dim = op.dim;
    const int nd = os.num_deuteron_states, nE = os.num_T_lab;
    const bool bu = U_BU && os.num_BU_chns > 0;
    {
        size_t fr = 0, tot = 0;
        TT_HIP(hipMemGetInfo(&fr, &tot));
        const double need = 16.0 * (double)dim * dim + 16.0 * (double)dim * (2 * 64 + 130 + 3 * 64);
        if (need > 0.95 * (double)fr)
            fail(TT_ERR_ALLOC, "dense solve needs " + std::to_string((long long)(need / (1 << 20))) +
                                   " MiB of device memory for dim " + std::to_string((long long)dim) + ", " +
                                   std::to_string((long long)(fr >> 20)) + " MiB free");
    }
    const int chunk = 64;
    DevBuf<double2> E((size_t)dim * chunk), Y((size_t)dim * chunk), M((size_t)dim * dim);
    DevBuf<int64_t> di, dcols;
    op.set_gidx(std::vector<int>(chunk, -1));
    if (bu) std::fill(U_BU, U_BU + (size_t)nd * os.num_BU_chns, cdouble(NAN, NAN));
    for (int j = 0; j < nE; ++j) {
        // right-hand sides: the on-shell columns of A, then the breakup cells of energy j inside the
        // state space (CPP/solve_faddeev.cpp:372-388)
        std::vector<int64_t> cols;
        for (int dc = 0; dc < nd; ++dc) cols.push_back(((int64_t)os.deuteron_idx[dc] * op.Nq + os.q_com_idx[j]) * op.Np);
        std::vector<int> bsel;
        if (bu)
            for (int b = os.q_com_BU_idx[j]; b < os.q_com_BU_idx[j + 1]; ++b) {
                const int64_t flat = breakup_flat(os, b, op.Nq, op.Np);
                if (flat < 0 || flat >= dim) continue;
                bsel.push_back(b);
                cols.push_back(flat);
            }
        const int nrhs = (int)cols.size();
        dcols.upload(cols.data(), cols.size(), op.s);
        DevBuf<double2> B((size_t)dim * nrhs);
        const double2 *Gj = op.G.p + (size_t)j * dim;
        for (int64_t r0 = 0; r0 < dim; r0 += chunk) {
            const int rc = (int)std::min<int64_t>(chunk, dim - r0);
            std::vector<int64_t> idx(rc);
            for (int k = 0; k < rc; ++k) idx[k] = r0 + k;
            TT_HIP(hipMemsetAsync(E.p, 0, sizeof(double2) * dim * rc, op.s));
            di.upload(idx.data(), rc, op.s);
            hipLaunchKernelGGL(k_unit_rows, dim3(nblk(rc, 64)), dim3(64), 0, op.s, E.p, di.p, rc);
            op.apply(E.p, Y.p, rc, false, false, false);
            hipLaunchKernelGGL(k_dense_rows, dim3(nblk(dim * rc, 256)), dim3(256), 0, op.s, Y.p, rc, r0, dim, Gj, M.p);
            hipLaunchKernelGGL(k_dense_rhs, dim3(nblk((int64_t)nrhs * rc, 256)), dim3(256), 0, op.s, Y.p, rc, r0, dim,
                               dcols.p, nrhs, B.p);
        }
        if (!dense_lu_solve(op.s, M.p, dim, B.p, nrhs))
            fail(TT_ERR_PHYSICS, "dense solve: 1 - A G has a zero pivot column (singular)");
        std::vector<double2> Bh((size_t)dim * nrhs);
        TT_HIP(hipMemcpyAsync(Bh.data(), B.p, sizeof(double2) * Bh.size(), hipMemcpyDeviceToHost, op.s));
        TT_HIP(hipStreamSynchronize(op.s));
        for (int dr = 0; dr < nd; ++dr) {
            const int64_t rr = ((int64_t)os.deuteron_idx[dr] * op.Nq + os.q_com_idx[j]) * op.Np;
            for (int dc = 0; dc < nd; ++dc) {
                const double2 v = Bh[(size_t)dc * dim + rr];
                U[(size_t)dr * nd * nE + (size_t)dc * nE + j] = cdouble(v.x, v.y);
            }
            for (size_t k = 0; k < bsel.size(); ++k) {
                const double2 v = Bh[(size_t)(nd + k) * dim + rr];
                U_BU[(size_t)dr * os.num_BU_chns + bsel[k]] = cdouble(v.x, v.y);
            }
        }
    }
    st.iterations = 0;
}

static DeviceCSC *dev_of(const tt_csc *m) {
    TT_REQUIRE(m && m->dev, "P123 has no device mirror (use tt_csc_upload or tt_fill_P123_arrays)");
    return reinterpret_cast<DeviceCSC *>(m->dev);
}

}  // namespace tt

using namespace tt;

namespace tt {
// ------------------------------------------------------------------ static scratch
// Device memory of tt_apply_rows_noalloc: every buffer its kernels use beyond the caller's
// inputs and outputs (operator tables, work items, the staged rows, G and the product's partial
// panels) is carved from this array, so that call allocates nothing.  Each GPU the library runs
// on holds its own copy, at its own address, until the process ends.  In host code the array's
// name is not a device address: kernels and copies take scratch_ptr(device) instead.
constexpr size_t SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

static std::mutex scratch_mu;
static std::vector<void *> scratch_addr;   // per device, from hipGetSymbolAddress

// Called once per device with that device current (tt_ctx_create).
static void scratch_setup(int device) {
    std::lock_guard<std::mutex> lock(scratch_mu);
    if ((int)scratch_addr.size() <= device) scratch_addr.resize(device + 1, nullptr);
    if (scratch_addr[device]) return;
    void *ptr = nullptr;
    TT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    TT_REQUIRE(ptr, "no device address for the static scratch array");
    scratch_addr[device] = ptr;
}

static unsigned char *scratch_ptr(int device) {
    std::lock_guard<std::mutex> lock(scratch_mu);
    TT_REQUIRE(device >= 0 && device < (int)scratch_addr.size() && scratch_addr[device],
               "static scratch not set up on this device (tt_ctx_create does it)");
    return static_cast<unsigned char *>(scratch_addr[device]);
}

// Consecutive 256-byte aligned pieces of one region, sized first (base == nullptr) and then
// placed.
struct Carve {
    unsigned char *base = nullptr;
    size_t used = 0;
    template <typename T>
    T *take(size_t count) {
        const size_t at = (used + 255) & ~(size_t)255;
        used = at + sizeof(T) * std::max<size_t>(count, 1);
        return base ? reinterpret_cast<T *>(base + at) : nullptr;
    }
};

}  // namespace tt

extern "C" {

tt_status tt_ctx_create(int device, tt_ctx **out) {
    TT_GUARD_BEGIN
    TT_REQUIRE(out, "null argument");
    int n = 0;
    TT_HIP(hipGetDeviceCount(&n));
    TT_REQUIRE(device >= 0 && device < n, "no such HIP device");
    TT_HIP(hipSetDevice(device));
    scratch_setup(device);   // the static scratch's address on this device (once per device)
    auto *c = new tt_ctx;
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        fail(TT_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
    }
    *out = c;
    TT_GUARD_END
}

void tt_ctx_destroy(tt_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (hipStream_t st : c->lanes) (void)hipStreamSynchronize(st);
    free_plan_cache(c->plans);
    c->plans = nullptr;
    free_group_tables(c->groups);
    c->groups = nullptr;
    for (hipStream_t st : c->lanes) (void)hipStreamDestroy(st);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

void *tt_ctx_stream(tt_ctx *c) { return c ? (void *)c->stream : nullptr; }

tt_status tt_ctx_synchronize(tt_ctx *c) {
    TT_GUARD_BEGIN
    TT_REQUIRE(c, "null context");
    TT_HIP(hipStreamSynchronize(c->stream));
    for (hipStream_t st : c->lanes) TT_HIP(hipStreamSynchronize(st));
    TT_GUARD_END
}

void tt_solve_opts_default(tt_solve_opts *o) {
    o->method = TT_SOLVE_PADE;
    o->tol = 1e-12;
    o->max_iter = 2000;
    o->restart = 40;
    o->NM_max = 14;
    o->neumann_terms = nullptr;
    o->p123_format = TT_P123_CELLS;
    o->use_graphs = -1;
    o->resolvent_on_device = 0;
    o->U_BU = nullptr;
}

tt_status tt_csc_upload(tt_ctx *ctx, int64_t dim, int64_t nnz, const double *val, const int32_t *row,
                        const uint64_t *colptr, tt_csc *out) {
    TT_GUARD_BEGIN
    TT_REQUIRE(ctx && out && colptr && (nnz == 0 || (val && row)) && dim > 0, "null argument");
    TT_REQUIRE(colptr[0] == 0 && (int64_t)colptr[dim] == nnz, "colptr inconsistent with nnz");
    for (int64_t c = 0; c < dim; ++c) TT_REQUIRE(colptr[c] <= colptr[c + 1], "colptr not monotone");
    for (int64_t i = 0; i < nnz; ++i) TT_REQUIRE(row[i] >= 0 && row[i] < dim, "row index out of range");
    TT_HIP(hipSetDevice(ctx->device));
    auto *d = new DeviceCSC;
    try {
        d->dim = dim;
        d->nnz = nnz;
        TT_HIP(hipMalloc(&d->val, sizeof(double) * std::max<int64_t>(nnz, 1)));
        TT_HIP(hipMalloc(&d->row, sizeof(int32_t) * std::max<int64_t>(nnz, 1)));
        TT_HIP(hipMalloc(&d->colptr, sizeof(uint64_t) * (dim + 1)));
        if (nnz) {
            TT_HIP(hipMemcpyAsync(d->val, val, sizeof(double) * nnz, hipMemcpyHostToDevice, ctx->stream));
            TT_HIP(hipMemcpyAsync(d->row, row, sizeof(int32_t) * nnz, hipMemcpyHostToDevice, ctx->stream));
        }
        TT_HIP(hipMemcpyAsync(d->colptr, colptr, sizeof(uint64_t) * (dim + 1), hipMemcpyHostToDevice, ctx->stream));
        TT_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        delete d;
        throw;
    }
    std::memset(out, 0, sizeof(*out));
    out->dim = dim;
    out->nnz = nnz;
    out->dev = d;
    TT_GUARD_END
}

tt_status tt_csc_download(tt_ctx *ctx, tt_csc *m) {
    TT_GUARD_BEGIN
    TT_REQUIRE(ctx && m, "null argument");
    DeviceCSC *d = dev_of(m);
    std::free(m->val);
    std::free(m->row);
    std::free(m->colptr);
    m->val = (double *)std::malloc(sizeof(double) * std::max<int64_t>(d->nnz, 1));
    m->row = (int32_t *)std::malloc(sizeof(int32_t) * std::max<int64_t>(d->nnz, 1));
    m->colptr = (uint64_t *)std::malloc(sizeof(uint64_t) * (d->dim + 1));
    if (!m->val || !m->row || !m->colptr) throw std::bad_alloc();
    if (d->nnz) {
        TT_HIP(hipMemcpyAsync(m->val, d->val, sizeof(double) * d->nnz, hipMemcpyDeviceToHost, ctx->stream));
        TT_HIP(hipMemcpyAsync(m->row, d->row, sizeof(int32_t) * d->nnz, hipMemcpyDeviceToHost, ctx->stream));
    }
    TT_HIP(hipMemcpyAsync(m->colptr, d->colptr, sizeof(uint64_t) * (d->dim + 1), hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    m->dim = d->dim;
    m->nnz = d->nnz;
    TT_GUARD_END
}

void tt_csc_free(tt_csc *m) {
    if (!m) return;
    std::free(m->val);
    std::free(m->row);
    std::free(m->colptr);
    delete reinterpret_cast<DeviceCSC *>(m->dev);
    std::memset(m, 0, sizeof(*m));
}

tt_status tt_calculate_resolvent(tt_ctx *ctx, const tt_run_params *rp, const tt_swp_states *swp,
                                 const tt_pw_states *chn, const double *E, int nE, double *G) {
    TT_GUARD_BEGIN
    TT_REQUIRE(ctx && rp && swp && chn && E && G && nE > 0, "null argument");
    TT_HIP(hipSetDevice(ctx->device));
    const int64_t dim = (int64_t)chn->Nalpha * swp->Nq_WP * swp->Np_WP;
    DevBuf<double2> dG((size_t)dim * nE);
    resolvent_on_device(ctx->stream, *rp, *swp, *chn, E, nE, dG.p);
    TT_HIP(hipMemcpyAsync(G, dG.p, sizeof(double2) * dim * nE, hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    TT_GUARD_END
}

}  // extern "C"

namespace tt {
// The plan of one channel's solve (checked arguments).
static SolvePlan &find_plan(tt_ctx *ctx, const tt_run_params *rp, const double *G, const tt_csc *P123,
                            const double *V_unco, const double *V_coup, const tt_swp_states *swp,
                            const tt_channel_os *os, const tt_pw_states *chn, const tt_solve_opts &o,
                            const ContentRef *content_key = nullptr) {
    TT_REQUIRE(ctx && rp && P123 && V_unco && swp && os && chn, "null argument");
    TT_REQUIRE(G || o.resolvent_on_device, "G required unless opts->resolvent_on_device is set");
    TT_REQUIRE(os->num_deuteron_states > 0 && os->num_T_lab > 0, "no on-shell elements to solve for");
    TT_HIP(hipSetDevice(ctx->device));
    const DeviceCSC *P = dev_of(P123);
    TT_REQUIRE(P->dim == (int64_t)chn->Nalpha * swp->Nq_WP * swp->Np_WP, "P123 dimension does not match Nalpha*Nq*Np");
    // the cell-block form covers N_alpha <= CELL_NA_MAX; larger channels take the CSC gather
    const int format = o.p123_format == TT_P123_CELLS && chn->Nalpha > CELL_NA_MAX ? TT_P123_CSC : o.p123_format;
    return get_plan(ctx, *rp, P, V_unco, V_coup, *swp, *chn, format, false, content_key);
}

// G of the solve in place in the plan (computed on the device or uploaded) and the plan's
// statistics counters cleared.
static void load_plan(tt_ctx *ctx, SolvePlan &pl, const tt_run_params *rp, const double *G,
                      const tt_swp_states *swp, const tt_channel_os *os, const tt_pw_states *chn,
                      const tt_solve_opts &o) {
    RowOperator &op = pl.op;
    op.collect();
    op.ms_apply = op.ms_spmm = op.spmm_bytes = op.ms_spmm32 = op.spmm_bytes32 = 0.0;
    op.n_apply = op.n_apply32 = op.n_spmm = op.n_spmm32 = 0;
    if (o.resolvent_on_device) {
        // E_com + E_bound of the on-shell bins (CPP/main.cpp:391), straight into the plan's G
        TT_REQUIRE(os->E_com, "resolvent_on_device needs the on-shell energies (os->E_com)");
        std::vector<double> E(os->num_T_lab);
        for (int j = 0; j < os->num_T_lab; ++j) E[j] = os->E_com[j] + swp->E_bound;
        resolvent_on_device(ctx->stream, *rp, *swp, *chn, E.data(), os->num_T_lab, op.G_buffer(os->num_T_lab), false,
                            &op.rstage);
    } else {
        op.set_G(G, os->num_T_lab);
    }
}

// find_plan + load_plan: the set-up shared by the single and the batched solves.
static SolvePlan &channel_plan(tt_ctx *ctx, const tt_run_params *rp, const double *G, const tt_csc *P123,
                               const double *V_unco, const double *V_coup, const tt_swp_states *swp,
                               const tt_channel_os *os, const tt_pw_states *chn, const tt_solve_opts &o) {
    SolvePlan &pl = find_plan(ctx, rp, G, P123, V_unco, V_coup, swp, os, chn, o);
    load_plan(ctx, pl, rp, G, swp, os, chn, o);
    return pl;
}

// Statistics of the plan's operator since channel_plan (the stream has passed).
static void channel_stats(SolvePlan &pl, tt_solve_stats &st) {
    RowOperator &op = pl.op;
    op.collect();
    st.n_apply = op.n_apply;
    st.n_apply_fp32 = op.n_apply32;
    st.ms_apply = op.ms_apply;
    st.n_spmm = op.n_spmm;
    st.ms_spmm = op.ms_spmm;
    st.spmm_bytes = op.spmm_bytes;
    st.n_spmm_fp32 = op.n_spmm32;
    st.ms_spmm_fp32 = op.ms_spmm32;
    st.spmm_bytes_fp32 = op.spmm_bytes32;
}

static void require_converged(const tt_solve_opts &o, double res) {
    if ((o.method == TT_SOLVE_GMRES || o.method == TT_SOLVE_MIXED) && !(res <= o.tol))
        fail(TT_ERR_NOT_CONVERGED, "Krylov solve stopped at relative residual " + std::to_string(res) +
                                       " above the requested tolerance (U holds the last iterate)");
}

// One channel's on-shell solve (the body of tt_solve_faddeev_equations; throws on error).
static void solve_channel(tt_ctx *ctx, const tt_run_params *rp, double *U, const double *G, const tt_csc *P123,
                          const double *V_unco, const double *V_coup, const tt_swp_states *swp,
                          const tt_channel_os *os, const tt_pw_states *chn, const tt_solve_opts &o,
                          tt_solve_stats *stats) {
    TT_REQUIRE(U, "null argument");
    auto t0 = std::chrono::steady_clock::now();
    tt_solve_stats st;
    std::memset(&st, 0, sizeof(st));
    SolvePlan &pl = channel_plan(ctx, rp, G, P123, V_unco, V_coup, swp, os, chn, o);
    RowOperator &op = pl.op;
    std::vector<cdouble> Uc((size_t)os->num_deuteron_states * os->num_deuteron_states * os->num_T_lab);
    const bool bu = wants_breakup(*os, o.U_BU);
    std::vector<cdouble> Ubu(bu ? (size_t)os->num_deuteron_states * os->num_BU_chns : 0);
    cdouble *bu_out = bu ? Ubu.data() : nullptr;
    switch (o.method) {
        case TT_SOLVE_PADE:
            solve_pade(op, *os, o, Uc.data(), st, bu_out);
            break;
        case TT_SOLVE_DENSE:
            solve_dense(op, *os, Uc.data(), st, bu_out);
            break;
        case TT_SOLVE_GMRES:
        case TT_SOLVE_MIXED: {
            std::vector<KMember> ms(1);
            ms[0].pl = &pl;
            ms[0].os = os;
            ms[0].mixed = o.method == TT_SOLVE_MIXED;
            ms[0].breakup = bu;
            solve_lockstep(ms, o);
            pl.read_onshell(ms[0].L, Uc.data(), bu_out);
            st.iterations = ms[0].iters;
            st.residual = ms[0].res;
            break;
        }
        default:
            fail(TT_ERR_ARG, "unknown solve method");
    }
    TT_HIP(hipStreamSynchronize(ctx->stream));
    std::memcpy(U, Uc.data(), sizeof(cdouble) * Uc.size());
    if (bu) std::memcpy(o.U_BU, Ubu.data(), sizeof(cdouble) * Ubu.size());
    channel_stats(pl, st);
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = st;
    require_converged(o, st.residual);
}
// ------------------------------------------------------------------ batch
// Every item is set up (plan, G on the device); the Krylov methods then solve all items'
// channels in lockstep (solve_lockstep: one stream, every wait shared), at most
// PlanCache::max_plans channels at a time; a plan in use by the running group is never evicted.
// Pade and dense solves go item by item.  Every item is attempted and gets its own
// status; the first failure is raised at the end.
static void solve_batch(tt_ctx *ctx, const tt_run_params *rp, const double *V_unco, const double *V_coup,
                        const tt_swp_states *swp, tt_batch_item *items, int n_items, const tt_solve_opts &o) {
    tt_status first = TT_OK;
    std::string first_msg;
    auto record = [&](int i, tt_status code, const std::string &msg) {
        items[i].status = code;
        if (code != TT_OK && first == TT_OK) {
            first = code;
            first_msg = "batch item " + std::to_string(i) + ": " + msg;
        }
    };
    for (int i = 0; i < n_items; ++i) items[i].status = TT_OK;
    const bool krylov = o.method == TT_SOLVE_GMRES || o.method == TT_SOLVE_MIXED;
    if (!krylov) {
        for (int i = 0; i < n_items; ++i) {
            tt_batch_item &it = items[i];
            try {
                tt_solve_opts oi = o;
                oi.U_BU = it.U_BU;
                solve_channel(ctx, rp, it.U, nullptr, it.P123, V_unco, V_coup, swp, it.os, it.chn, oi, it.stats);
            } catch (const Error &e) {
                record(i, e.code, e.msg);
            }
        }
        if (first != TT_OK) fail(first, first_msg);
        return;
    }
    const int group_max = (int)PlanCache::max_plans;
    // V and C are the same for every item: take their snapshot once
    const ContentRef content = plan_content(*rp, V_unco, V_coup, *swp);
    int i0 = 0;
    while (i0 < n_items) {
        auto t0 = std::chrono::steady_clock::now();
        std::vector<KMember> ms;
        std::vector<int> idx;
        struct Release {   // plans of the group may be evicted again afterwards
            std::vector<KMember> *m;
            ~Release() {
                for (auto &mb : *m) mb.pl->in_use = 0;
            }
        } release{&ms};
        int i = i0;
        for (; i < n_items && (int)ms.size() < group_max; ++i) {
            tt_batch_item &it = items[i];
            try {
                TT_REQUIRE(it.U, "null U");
                SolvePlan &pl = find_plan(ctx, rp, nullptr, it.P123, V_unco, V_coup, swp, it.os, it.chn, o, &content);
                if (pl.in_use) break;   // the same channel twice: it starts the next group
                load_plan(ctx, pl, rp, nullptr, swp, it.os, it.chn, o);
                pl.in_use = 1;
                KMember mb;
                mb.pl = &pl;
                mb.os = it.os;
                mb.mixed = o.method == TT_SOLVE_MIXED;
                mb.breakup = wants_breakup(*it.os, it.U_BU);
                ms.push_back(std::move(mb));
                idx.push_back(i);
            } catch (const Error &e) {
                record(i, e.code, e.msg);
            }
        }
        i0 = i;
        try {
            solve_on_lanes(ctx, ms, o);
        } catch (const Error &e) {
            for (int k : idx) record(k, e.code, e.msg);
            continue;
        }
        const double ms_group = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        for (size_t k = 0; k < ms.size(); ++k) {
            tt_batch_item &it = items[idx[k]];
            SolvePlan &pl = *ms[k].pl;
            std::vector<cdouble> Uc((size_t)ms[k].L.nd * ms[k].L.nd * ms[k].L.nE);
            std::vector<cdouble> Ubu(ms[k].breakup ? (size_t)ms[k].L.nd * it.os->num_BU_chns : 0);
            pl.read_onshell(ms[k].L, Uc.data(), ms[k].breakup ? Ubu.data() : nullptr);
            std::memcpy(it.U, Uc.data(), sizeof(cdouble) * Uc.size());
            if (ms[k].breakup) std::memcpy(it.U_BU, Ubu.data(), sizeof(cdouble) * Ubu.size());
            tt_solve_stats st;
            std::memset(&st, 0, sizeof(st));
            channel_stats(pl, st);
            st.iterations = ms[k].iters;
            st.residual = ms[k].res;
            st.ms_total = ms_group;   // the group's wall time (its channels are solved together)
            if (it.stats) *it.stats = st;
            try {
                require_converged(o, st.residual);
            } catch (const Error &e) {
                record(idx[k], e.code, e.msg);
            }
        }
    }
    if (first != TT_OK) fail(first, first_msg);
}

// ------------------------------------------------------------------ gradient
// U(theta) and dU/dtheta for one channel, V = sum_i theta_i V_i (no reference counterpart).
//   forward:  V, SWP states, on-shell bins, G on the device, row GMRES y_k (1 - G_j A) = e_k^T A
//   adjoint:  column GMRES (1 - A G_j) t_k = A e_k, z_k = e_k + G_j t_k, w_k = e_k + y_k G_j, and
//     dU[dr][dc][j] = (dC_op w_k).(2P VC_op z_k') + (2P^T C_op w_k).(dVC_op z_k') + y_k.(dG_j t_k')
//   with k = (dr, j), k' = (dc, j) and dVC = dV C + V dC per 2N channel; dC, the SWP edge and
//   deuteron-energy changes come from first-order perturbation of the SWP states, dG from the
//   closed forms.  The on-shell bins are held fixed (their choice is piecewise constant).
// Cost: the forward solve, one column solve, two stage products and per-theta block ops.
static void matmul_add(const double *A, const double *B, double *C, int n) {
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) {
            const double a = A[(size_t)i * n + k];
            if (a == 0.0) continue;
            for (int j = 0; j < n; ++j) C[(size_t)i * n + j] += a * B[(size_t)k * n + j];
        }
}

static void solve_gradient(tt_ctx *ctx, const tt_run_params &rp, const tt_fwp_states &fwp,
                           const tt_pw_states &pw_full, int channel, const tt_csc *P123, const double *Vb_unco,
                           const double *Vb_coup, int n_theta, const double *theta, const double *T_lab, int nT,
                           double *U, double *dU, int capacity, int *n_out, const tt_solve_opts &o_in,
                           tt_grad_stats *gs) {
    auto t0 = std::chrono::steady_clock::now();
    auto ms_since = [&]() { return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); };
    int nu, nc;
    num_2N_states(rp, nu, nc);
    const int Np = fwp.Np_WP, Nq = fwp.Nq_WP;
    const size_t su = (size_t)nu * Np * Np, sc = (size_t)nc * 4 * Np * Np;
    std::vector<double> Vu(std::max<size_t>(su, 1), 0.0), Vc(std::max<size_t>(sc, 1), 0.0);
    for (int i = 0; i < n_theta; ++i) {
        for (size_t e = 0; e < su; ++e) Vu[e] += theta[i] * Vb_unco[(size_t)i * su + e];
        for (size_t e = 0; e < sc; ++e) Vc[e] += theta[i] * Vb_coup[(size_t)i * sc + e];
    }
    std::vector<double> eu(std::max<size_t>((size_t)nu * (Np + 1), 1)), ec(std::max<size_t>((size_t)nc * 2 * (Np + 1), 1));
    std::vector<double> Cu(std::max<size_t>(su, 1)), Cc(std::max<size_t>(sc, 1));
    tt_swp_states swp;
    std::memset(&swp, 0, sizeof(swp));
    swp.e_SWP_unco = eu.data();
    swp.e_SWP_coup = ec.data();
    swp.C_SWP_unco = Cu.data();
    swp.C_SWP_coup = Cc.data();
    auto check = [](tt_status st) {
        if (st != TT_OK) fail(st, tt_last_error());
    };
    check(tt_make_swp_states(&rp, &fwp, &pw_full, Vu.data(), Vc.data(), &swp));
    tt_pw_states chn;
    check(tt_pw_channel(&pw_full, channel, &chn));
    if (dU && chn.Nalpha > CELL_NA_MAX)
        fail(TT_ERR_UNSUPPORTED, "the adjoint's column operator uses the cell-block form, which covers N_alpha <= 64 "
                                 "(this channel has " + std::to_string(chn.Nalpha) + ")");
    tt_channel_os os;
    check(tt_find_on_shell(&swp, &chn, T_lab, nT, &os));
    struct OsGuard {
        tt_channel_os *o;
        ~OsGuard() { tt_channel_os_free(o); }
    } os_guard{&os};
    const int nd = os.num_deuteron_states, nE = os.num_T_lab, n = nd * nd * nE;
    if (n_out) *n_out = n;
    TT_REQUIRE(nd > 0 && nE > 0, "no on-shell elements to solve for");
    TT_REQUIRE(capacity >= n, "U / dU capacity is smaller than nd*nd*nE");
    tt_solve_opts o = o_in;
    o.method = TT_SOLVE_GMRES;
    o.resolvent_on_device = 1;
    o.p123_format = TT_P123_CELLS;
    tt_grad_stats g;
    std::memset(&g, 0, sizeof(g));
    if (!dU) {
        tt_solve_stats st_f;
        solve_channel(ctx, &rp, U, nullptr, P123, Vu.data(), Vc.data(), &swp, &os, &chn, o, &st_f);
        g.ms_forward = g.ms_total = ms_since();
        g.n_apply_forward = st_f.n_apply;
        g.residual_forward = st_f.residual;
        if (gs) *gs = g;
        return;
    }
    // The forward row solve and the adjoint column solve are independent: they run in lockstep
    // on the device while host threads compute the SWP perturbation of every theta.
    const DeviceCSC *P = dev_of(P123);
    SolvePlan &rowp = channel_plan(ctx, &rp, nullptr, P123, Vu.data(), Vc.data(), &swp, &os, &chn, o);
    rowp.in_use = 1;   // not evicted by the lookup of the column plan
    struct Release {
        SolvePlan *p;
        ~Release() { p->in_use = 0; }
    } release{&rowp};
    SolvePlan &colp = get_plan(ctx, rp, P, Vu.data(), Vc.data(), swp, chn, TT_P123_CELLS, true);
    RowOperator &R = rowp.op, &Cl = colp.op;
    hipStream_t s = ctx->stream;
    const OnshellLayout L = onshell_layout(os, Nq, Np);
    const int r = L.r, Na = chn.Nalpha;
    const int64_t dim = R.dim, ne = dim * r;
    std::vector<double> E(nE), fE(nE);
    const double Md = kMn + kMp + swp.E_bound;
    for (int j = 0; j < nE; ++j) {
        E[j] = os.E_com[j] + swp.E_bound;
        // E_com = (q_l^2 + q_u^2) / (4 mu1(E_bound)) at the bin midpoint: dE/dE_bound
        fE[j] = 1.0 - os.E_com[j] * kMn / (Md * (kMn + Md));
    }
    // host phase, one theta per task on a few threads: SWP perturbation, dVC = dV C + V dC,
    // and the partner lists of dC_op and dVC_op
    struct ThetaWork {
        std::vector<int> p1, a1, p2, a2;
        std::vector<double> b1, b2, de_u, de_c, dE;
        int ne1 = 0, ne2 = 0;
    };
    std::vector<ThetaWork> tw(n_theta);
    auto work = [&](int i) {
        ThetaWork &t = tw[i];
        const double *dVu = Vb_unco + (size_t)i * su, *dVc = Vb_coup + (size_t)i * sc;
        SwpDerivative d;
        swp_derivative(rp, fwp, pw_full, Vu.data(), Vc.data(), swp, dVu, dVc, d);
        std::vector<double> dVCu(std::max<size_t>(su, 1), 0.0), dVCc(std::max<size_t>(sc, 1), 0.0);
        for (int c = 0; c < nu; ++c) {
            const size_t o1 = (size_t)c * Np * Np;
            matmul_add(dVu + o1, Cu.data() + o1, dVCu.data() + o1, Np);
            matmul_add(Vu.data() + o1, d.dC_unco.data() + o1, dVCu.data() + o1, Np);
        }
        for (int c = 0; c < nc; ++c) {
            const size_t o1 = (size_t)c * 4 * Np * Np;
            matmul_add(dVc + o1, Cc.data() + o1, dVCc.data() + o1, 2 * Np);
            matmul_add(Vc.data() + o1, d.dC_coup.data() + o1, dVCc.data() + o1, 2 * Np);
        }
        embed_2N_blocks(chn, Np, d.dC_unco.data(), d.dC_coup.data(), rp, false, t.p1, t.a1, t.b1);
        embed_2N_blocks(chn, Np, dVCu.data(), dVCc.data(), rp, false, t.p2, t.a2, t.b2);
        for (int a = 0; a < Na; ++a) {
            t.ne1 = std::max(t.ne1, t.p1[a + 1] - t.p1[a]);
            t.ne2 = std::max(t.ne2, t.p2[a + 1] - t.p2[a]);
        }
        t.a1.push_back(0);
        t.a2.push_back(0);
        t.b1.push_back(0.0);
        t.b2.push_back(0.0);
        t.de_u = d.de_unco;
        t.de_c = d.de_coup;
        t.de_u.push_back(0.0);
        t.de_c.push_back(0.0);
        t.dE.resize(nE);
        for (int j = 0; j < nE; ++j) t.dE[j] = d.dE_bound * fE[j];
    };
    const int nthr = (int)std::max(1u, std::min<unsigned>((unsigned)n_theta, std::min(8u, std::thread::hardware_concurrency())));
    std::vector<std::exception_ptr> errs(nthr);   // declared before the workers: outlives them
    struct Workers {   // joined on every exit path
        std::vector<std::thread> th;
        ~Workers() {
            for (auto &x : th)
                if (x.joinable()) x.join();
        }
    } workers;
    for (int t = 0; t < nthr; ++t)
        workers.th.emplace_back([&, t]() {
            try {
                for (int i = t; i < n_theta; i += nthr) work(i);
            } catch (...) {
                errs[t] = std::current_exception();
            }
        });
    // the column operator takes the same G and energy of each system
    double2 *Gc = Cl.G_buffer(nE);
    TT_HIP(hipMemcpyAsync(Gc, R.G.p, sizeof(double2) * (size_t)nE * dim, hipMemcpyDeviceToDevice, s));
    Cl.collect();
    Cl.n_apply = Cl.n_apply32 = 0;
    std::vector<KMember> ms(2);
    ms[0].pl = &rowp;   // y_k (1 - G A) = e_k^T A
    ms[0].os = &os;
    ms[1].pl = &colp;   // (1 - A G) t_k = A e_k
    ms[1].os = &os;
    solve_on_lanes(ctx, ms, o);
    {
        std::vector<cdouble> Uc((size_t)n);
        rowp.read_onshell(ms[0].L, Uc.data());
        std::memcpy(U, Uc.data(), sizeof(cdouble) * n);
    }
    g.ms_forward = ms_since();   // V, SWP states, G and both solves (run together)
    g.n_apply_forward = R.n_apply;
    g.residual_forward = ms[0].res;
    g.residual_adjoint = ms[1].res;
    g.n_apply_adjoint = Cl.n_apply;
    if (!(ms[0].res <= o.tol))
        fail(TT_ERR_NOT_CONVERGED, "forward row solve stopped at relative residual " + std::to_string(ms[0].res));
    if (!(ms[1].res <= o.tol))
        fail(TT_ERR_NOT_CONVERGED, "adjoint column solve stopped at relative residual " + std::to_string(ms[1].res));
    // device phase in the row plan's gradient buffers (allocated on first use, then reused)
    if (!rowp.gw) rowp.gw.reset(new SolvePlan::GradWork);
    SolvePlan::GradWork &gw = *rowp.gw;
    for (DevBuf<double2> *v : {&gw.z, &gw.w, &gw.Uw, &gw.bz, &gw.Dw, &gw.Dz})
        if (v->n < (size_t)ne) v->alloc(ne);
    gw.unit.upload(L.row_idx.data(), r, s);
    hipLaunchKernelGGL(k_unit_plus_G, dim3(nblk(ne, 256)), dim3(256), 0, s, colp.x.p, gw.z.p, R.G.p, R.gidx.p,
                       gw.unit.p, dim, r);
    hipLaunchKernelGGL(k_unit_plus_G, dim3(nblk(ne, 256)), dim3(256), 0, s, rowp.x.p, gw.w.p, R.G.p, R.gidx.p,
                       gw.unit.p, dim, r);
    R.stage_product(gw.w.p, gw.Uw.p, r);    // 2 P^T C_op w   (cell-major)
    Cl.stage_product(gw.z.p, gw.bz.p, r);   // 2 P VC_op z    (cell-major)
    std::vector<int4> pairs;
    for (int dr = 0; dr < nd; ++dr)
        for (int dc = 0; dc < nd; ++dc)
            for (int j = 0; j < nE; ++j) pairs.push_back(make_int4(dr * nE + j, dc * nE + j, j, 0));
    const int npairs = (int)pairs.size();
    gw.pairs.upload(pairs.data(), npairs, s);
    const int nbp = (int)std::min<int64_t>(256, nblk(dim, 256));
    if (gw.part.n < (size_t)npairs * nbp) gw.part.alloc((size_t)npairs * nbp);
    if (gw.dG.n < (size_t)nE * dim) gw.dG.alloc((size_t)nE * dim);
    if (gw.Gp.n < (size_t)3 * nE * dim) gw.Gp.alloc((size_t)3 * nE * dim);   // dG/d(lower edge), d(upper), dG/dE
    resolvent_on_device(s, rp, swp, chn, E.data(), nE, gw.Gp.p, true, &R.rstage);
    std::vector<int> which, offset;
    swp_edge_offsets(chn, swp, rp, which, offset);
    gw.which.upload(which.data(), Na, s);
    gw.offset.upload(offset.data(), Na, s);
    for (auto &x : workers.th) x.join();
    for (auto &e : errs)
        if (e) std::rethrow_exception(e);
    // every theta's lists packed into one int and one double array (one upload each)
    struct Off {
        size_t p1, a1, p2, a2, b1, b2, du, dc, dE;
    };
    std::vector<Off> off(n_theta);
    std::vector<int> ints;
    std::vector<double> dbls;
    auto put_i = [&](const std::vector<int> &v) {
        const size_t o1 = ints.size();
        ints.insert(ints.end(), v.begin(), v.end());
        return o1;
    };
    auto put_d = [&](const std::vector<double> &v) {
        const size_t o1 = dbls.size();
        dbls.insert(dbls.end(), v.begin(), v.end());
        return o1;
    };
    for (int i = 0; i < n_theta; ++i) {
        const ThetaWork &t = tw[i];
        off[i] = {put_i(t.p1), put_i(t.a1), put_i(t.p2), put_i(t.a2), put_d(t.b1), put_d(t.b2), put_d(t.de_u),
                  put_d(t.de_c), put_d(t.dE)};
    }
    gw.ints.upload(ints.data(), ints.size(), s);
    gw.dbls.upload(dbls.data(), dbls.size(), s);
    if (gw.tri.n < (size_t)3 * npairs * n_theta) gw.tri.alloc((size_t)3 * npairs * n_theta);
    const int *I = gw.ints.p;
    const double *D = gw.dbls.p;
    for (int i = 0; i < n_theta; ++i) {
        const ThetaWork &t = tw[i];
        const Off &f = off[i];
        launch_block<double2, double2>(s, gw.w.p, gw.Dw.p, I + f.p1, I + f.a1, D + f.b1, t.ne1, nullptr, R.gidx.p,
                                       Na, Nq, Np, r, nullptr, 0, 1, nullptr, 1.0);   // dC_op w
        launch_block<double2, double2>(s, gw.z.p, gw.Dz.p, I + f.p2, I + f.a2, D + f.b2, t.ne2, nullptr, R.gidx.p,
                                       Na, Nq, Np, r, nullptr, 0, 1, nullptr, 1.0);   // dVC_op z
        hipLaunchKernelGGL(k_resolvent_change, dim3(nblk((int64_t)nE * dim, 256)), dim3(256), 0, s, gw.dG.p, gw.Gp.p,
                           D + f.du, D + f.dc, gw.which.p, gw.offset.p, D + f.dE, Na, Nq, Np, nE);
        double2 *tri_i = gw.tri.p + (size_t)3 * npairs * i;
        const dim3 pg((unsigned)nbp, (unsigned)npairs);
        hipLaunchKernelGGL(k_pair_dots, pg, dim3(256), 0, s, gw.Dw.p, 1, gw.bz.p, 1, nullptr, gw.pairs.p, Na, Nq * Np,
                           r, gw.part.p);
        hipLaunchKernelGGL(k_dots_reduce, dim3(npairs), dim3(256), 0, s, gw.part.p, nbp, 1, tri_i);
        hipLaunchKernelGGL(k_pair_dots, pg, dim3(256), 0, s, gw.Uw.p, 1, gw.Dz.p, 1, nullptr, gw.pairs.p, Na, Nq * Np,
                           r, gw.part.p);
        hipLaunchKernelGGL(k_dots_reduce, dim3(npairs), dim3(256), 0, s, gw.part.p, nbp, 1, tri_i + npairs);
        hipLaunchKernelGGL(k_pair_dots, pg, dim3(256), 0, s, rowp.x.p, 0, colp.x.p, 0, gw.dG.p, gw.pairs.p, Na,
                           Nq * Np, r, gw.part.p);
        hipLaunchKernelGGL(k_dots_reduce, dim3(npairs), dim3(256), 0, s, gw.part.p, nbp, 1, tri_i + 2 * npairs);
    }
    std::vector<double2> tri_h((size_t)3 * npairs * n_theta);
    TT_HIP(hipMemcpyAsync(tri_h.data(), gw.tri.p, sizeof(double2) * tri_h.size(), hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n_theta; ++i) {
        const double2 *tr = tri_h.data() + (size_t)3 * npairs * i;
        for (int q = 0; q < npairs; ++q) {
            const int dr = pairs[q].x / nE, dc = pairs[q].y / nE, j = pairs[q].z;
            const size_t e = (size_t)i * n + ((size_t)dr * nd + dc) * nE + j;
            dU[2 * e] = tr[q].x + tr[npairs + q].x + tr[2 * npairs + q].x;
            dU[2 * e + 1] = tr[q].y + tr[npairs + q].y + tr[2 * npairs + q].y;
        }
    }
    Cl.collect();
    g.ms_total = ms_since();
    if (gs) *gs = g;
}
}  // namespace tt

extern "C" {

tt_status tt_solve_faddeev_equations(tt_ctx *ctx, const tt_run_params *rp, double *U, const double *G,
                                     const tt_csc *P123, const double *V_unco, const double *V_coup,
                                     const tt_swp_states *swp, const tt_channel_os *os, const tt_pw_states *chn,
                                     const tt_solve_opts *opts, tt_solve_stats *stats) {
    TT_GUARD_BEGIN
    tt_solve_opts o;
    tt_solve_opts_default(&o);
    if (opts) o = *opts;
    solve_channel(ctx, rp, U, G, P123, V_unco, V_coup, swp, os, chn, o, stats);
    TT_GUARD_END
}

tt_status tt_solve_batch(tt_ctx *ctx, const tt_run_params *rp, const double *V_unco, const double *V_coup,
                         const tt_swp_states *swp, tt_batch_item *items, int n_items, const tt_solve_opts *opts) {
    TT_GUARD_BEGIN
    TT_REQUIRE(ctx && rp && V_unco && swp && (items || n_items == 0) && n_items >= 0, "null argument");
    tt_solve_opts o;
    tt_solve_opts_default(&o);
    if (opts) o = *opts;
    o.resolvent_on_device = 1;
    o.U_BU = nullptr;   // per item (tt_batch_item.U_BU)
    tt::solve_batch(ctx, rp, V_unco, V_coup, swp, items, n_items, o);
    TT_GUARD_END
}

tt_status tt_solve_gradient(tt_ctx *ctx, const tt_run_params *rp, const tt_fwp_states *fwp,
                            const tt_pw_states *pw_full, int channel, const tt_csc *P123, const double *Vb_unco,
                            const double *Vb_coup, int n_theta, const double *theta, const double *T_lab, int n_T_lab,
                            double *U, double *dU_dtheta, int capacity, int *n_out, const tt_solve_opts *opts,
                            tt_grad_stats *stats) {
    TT_GUARD_BEGIN
    TT_REQUIRE(ctx && rp && fwp && pw_full && P123 && Vb_unco && theta && T_lab && U && n_theta > 0 && n_T_lab > 0,
               "null argument");
    TT_REQUIRE(channel >= 0 && channel < pw_full->N_chn_3N, "channel index out of range");
    tt_solve_opts o;
    tt_solve_opts_default(&o);
    if (opts) o = *opts;
    solve_gradient(ctx, *rp, *fwp, *pw_full, channel, P123, Vb_unco, Vb_coup, n_theta, theta, T_lab, n_T_lab, U,
                   dU_dtheta, capacity, n_out, o, stats);
    TT_GUARD_END
}

int tt_cell_tuning(char *buf, int len) {
    tt::TuneTable &t = tt::tune_table();
    std::string out;
    {
        std::lock_guard<std::mutex> lock(t.mu);
        for (const auto &kv : t.map) {
            char line[192];
            std::snprintf(line, sizeof(line), "dev %d %s %s Nalpha %d R2 %d ncell %lld -> item_len %d un %d%s (%.1f us)\n",
                          kv.first.dev, kv.first.col ? "column" : "row", kv.first.fp32 ? "fp32" : "fp64",
                          kv.first.Na, kv.first.R2, (long long)kv.first.ncell, kv.second.len, kv.second.un,
                          kv.second.h ? " fp16-split" : "", 1e3 * kv.second.ms);
            out += line;
        }
    }
    if (buf && len > 0) {
        const int n = std::min<int>(len - 1, (int)out.size());
        std::memcpy(buf, out.data(), n);
        buf[n] = 0;
    }
    return (int)out.size() + 1;
}

tt_status tt_apply_rows(tt_ctx *ctx, const tt_run_params *rp, const tt_csc *P123, const double *V_unco,
                        const double *V_coup, const tt_swp_states *swp, const tt_pw_states *chn, const double *G,
                        int nE, const int *g_index, int r, const double *X, double *Y, int flags) {
    TT_GUARD_BEGIN
    TT_REQUIRE(ctx && rp && P123 && V_unco && swp && chn && X && Y && r > 0, "null argument");
    TT_HIP(hipSetDevice(ctx->device));
    RowOperator op;
    const bool csc = (flags & 2) != 0 || chn->Nalpha > CELL_NA_MAX;   // cells cover N_alpha <= 64
    if ((flags & 4) && chn->Nalpha > CELL_NA_MAX)
        fail(TT_ERR_UNSUPPORTED, "the column operator uses the cell-block form, which covers N_alpha <= 64");
    op.setup(ctx->stream, *rp, dev_of(P123), V_unco, V_coup, *swp, *chn, csc ? TT_P123_CSC : TT_P123_CELLS,
             (flags & 4) != 0);
    op.set_G(G, nE);
    std::vector<int> g(r, -1);
    bool useG = false;
    if (g_index)
        for (int k = 0; k < r; ++k) {
            TT_REQUIRE(g_index[k] < nE, "g_index out of range");
            g[k] = g_index[k];
            if (g[k] >= 0) useG = true;
        }
    TT_REQUIRE(!useG || G, "G required when g_index selects energies");
    op.set_gidx(g);
    const int64_t dim = op.dim;
    // host rows (r x dim) -> device RHS-fastest layout
    std::vector<double2> xt((size_t)dim * r), yt((size_t)dim * r);
    for (int k = 0; k < r; ++k)
        for (int64_t i = 0; i < dim; ++i) xt[(size_t)i * r + k] = make_double2(X[2 * ((size_t)k * dim + i)], X[2 * ((size_t)k * dim + i) + 1]);
    DevBuf<double2> dx, dy((size_t)dim * r);
    dx.upload(xt.data(), xt.size(), ctx->stream);
    op.apply(dx.p, dy.p, r, useG, (flags & 1) != 0);
    TT_HIP(hipMemcpyAsync(yt.data(), dy.p, sizeof(double2) * yt.size(), hipMemcpyDeviceToHost, ctx->stream));
    TT_HIP(hipStreamSynchronize(ctx->stream));
    for (int k = 0; k < r; ++k)
        for (int64_t i = 0; i < dim; ++i) {
            Y[2 * ((size_t)k * dim + i)] = yt[(size_t)i * r + k].x;
            Y[2 * ((size_t)k * dim + i) + 1] = yt[(size_t)i * r + k].y;
        }
    TT_GUARD_END
}

// tt_apply_rows (FP64 row form) with every device buffer beyond the P123 object carved from
// the static scratch array: the same block-op, cell-product (or CSC gather) and block-op kernels,
// the work items at the default length, no allocation, no autotuning, no timing events.
tt_status tt_apply_rows_noalloc(tt_ctx *ctx, const tt_run_params *rp, const tt_csc *P123, const double *V_unco,
                                const double *V_coup, const tt_swp_states *swp, const tt_pw_states *chn,
                                const double *G, int nE, const int *g_index, int r, const double *X, double *Y,
                                int flags) {
    TT_GUARD_BEGIN
    TT_REQUIRE(ctx && rp && P123 && V_unco && swp && chn && X && Y && r > 0, "null argument");
    if (flags & ~2)
        fail(TT_ERR_UNSUPPORTED, "tt_apply_rows_noalloc takes the FP64 row form only (flags 0 or 2)");
    TT_HIP(hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const DeviceCSC *P = dev_of(P123);
    const int Na = chn->Nalpha, Nq = swp->Nq_WP, Np = swp->Np_WP;
    const int64_t dim = (int64_t)Na * Nq * Np;
    TT_REQUIRE(P->dim == dim, "P123 dimension does not match Nalpha*Nq*Np");
    const bool cells = (flags & 2) == 0 && Na <= CELL_NA_MAX;
    if (cells && (P->Na != Na || P->NqNp != Nq * Np))
        fail(TT_ERR_UNSUPPORTED, "this P123 has no cell-block form yet (tt_fill_P123_arrays builds it; a "
                                 "solve builds it for an uploaded CSC): build it first or pass flags = 2 "
                                 "(gather over the CSC)");
    std::vector<int> g(r, -1);
    bool useG = false;
    if (g_index)
        for (int k = 0; k < r; ++k) {
            TT_REQUIRE(g_index[k] < nE, "g_index out of range");
            g[k] = g_index[k];
            if (g[k] >= 0) useG = true;
        }
    TT_REQUIRE(!useG || (G && nE > 0), "G required when g_index selects energies");
    // operator tables (host): C_op and VC_op partner blocks, as RowOperator::set_blocks
    VCProducts vc;
    vc_products(*rp, V_unco, V_coup, *swp, vc);
    std::vector<int> p1, a1, p2, a2;
    std::vector<double> b1, b2;
    embed_2N_blocks(*chn, Np, swp->C_SWP_unco, swp->C_SWP_coup, *rp, false, p1, a1, b1, true);
    embed_2N_blocks(*chn, Np, vc.u.data(), vc.c.data(), *rp, true, p2, a2, b2, true);
    int m1 = 1, m2 = 1;
    for (int a = 0; a < Na; ++a) {
        m1 = std::max(m1, p1[a + 1] - p1[a]);
        m2 = std::max(m2, p2[a + 1] - p2[a]);
    }
    a1.push_back(0);
    a2.push_back(0);
    b1.push_back(0.0);
    b2.push_back(0.0);
    CellView cv;
    std::vector<int2> it;
    std::vector<int> kip;
    if (cells) {
        cv = cells_of(*P, false);
        CellItems::runs(cv, s, CELL_ITEM, it, kip);
    }
    const int R2 = 2 * r;
    const size_t n_items = std::max<size_t>(it.size(), 1);
    // the pieces of the scratch: sized first, then placed
    unsigned char *base = nullptr;
    for (int pass = 0; pass < 2; ++pass) {
        Carve cut;
        cut.base = base;
        int *d_p1 = cut.take<int>(p1.size()), *d_a1 = cut.take<int>(a1.size());
        double *d_b1 = cut.take<double>(b1.size());
        int *d_p2 = cut.take<int>(p2.size()), *d_a2 = cut.take<int>(a2.size());
        double *d_b2 = cut.take<double>(b2.size());
        int *d_g = cut.take<int>(r);
        int2 *d_it = cut.take<int2>(n_items);
        int *d_kip = cut.take<int>(kip.size());
        double2 *d_G = cut.take<double2>(useG ? (size_t)nE * dim : 1);
        double2 *d_X = cut.take<double2>((size_t)dim * r), *d_S = cut.take<double2>((size_t)dim * r);
        double2 *d_Y = cut.take<double2>((size_t)dim * r);
        double *d_U = cut.take<double>(cells ? n_items * Na * R2 : (size_t)dim * R2);
        if (pass == 0) {
            if (cut.used > SCRATCH_BYTES)
                fail(TT_ERR_ARG, "tt_apply_rows_noalloc needs " + std::to_string(cut.used) +
                                     " bytes of scratch for these sizes, above its maximum of " +
                                     std::to_string(SCRATCH_BYTES) + " (use tt_apply_rows)");
            base = scratch_ptr(ctx->device);
            continue;
        }
        std::vector<double2> xt((size_t)dim * r), yt((size_t)dim * r);
        for (int k = 0; k < r; ++k)
            for (int64_t i = 0; i < dim; ++i)
                xt[(size_t)i * r + k] = make_double2(X[2 * ((size_t)k * dim + i)], X[2 * ((size_t)k * dim + i) + 1]);
        auto h2d = [&](void *d, const void *h, size_t bytes) {
            if (bytes) TT_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, s));
        };
        h2d(d_p1, p1.data(), sizeof(int) * p1.size());
        h2d(d_a1, a1.data(), sizeof(int) * a1.size());
        h2d(d_b1, b1.data(), sizeof(double) * b1.size());
        h2d(d_p2, p2.data(), sizeof(int) * p2.size());
        h2d(d_a2, a2.data(), sizeof(int) * a2.size());
        h2d(d_b2, b2.data(), sizeof(double) * b2.size());
        h2d(d_g, g.data(), sizeof(int) * r);
        if (cells) {
            h2d(d_it, it.data(), sizeof(int2) * it.size());
            h2d(d_kip, kip.data(), sizeof(int) * kip.size());
        }
        if (useG) h2d(d_G, G, sizeof(double2) * (size_t)nE * dim);
        h2d(d_X, xt.data(), sizeof(double2) * xt.size());
        // S = C_op (X .* G), then U = 2 P123^T S, then Y = VC_op^T U (RowOperator::apply, FP64)
        const int cm = cells ? 1 : 0;
        launch_block<double2, double2>(s, d_X, d_S, d_p1, d_a1, d_b1, m1, useG ? d_G : nullptr, d_g, Na, Nq, Np, r,
                                       nullptr, 0, cm, nullptr, 1.0);
        if (cells) {
            const ItemsRef ci{(int)it.size(), d_it, d_kip};
            launch_cellmm<double>(s, &cv, ci, cv.blk, reinterpret_cast<const double *>(d_S), nullptr, d_U, R2, 2.0);
            launch_block<double2, double2>(s, reinterpret_cast<const double2 *>(d_U), d_Y, d_p2, d_a2, d_b2, m2,
                                           nullptr, d_g, Na, Nq, Np, r, nullptr, cm, 0, d_kip, 2.0);
        } else {
            launch_spmm<double>(s, P, P->val, reinterpret_cast<const double *>(d_S), d_U, dim, R2, 2.0);
            launch_block<double2, double2>(s, reinterpret_cast<const double2 *>(d_U), d_Y, d_p2, d_a2, d_b2, m2,
                                           nullptr, d_g, Na, Nq, Np, r, nullptr, 0, 0, nullptr, 1.0);
        }
        TT_HIP(hipGetLastError());
        TT_HIP(hipMemcpyAsync(yt.data(), d_Y, sizeof(double2) * yt.size(), hipMemcpyDeviceToHost, s));
        TT_HIP(hipStreamSynchronize(s));
        for (int k = 0; k < r; ++k)
            for (int64_t i = 0; i < dim; ++i) {
                Y[2 * ((size_t)k * dim + i)] = yt[(size_t)i * r + k].x;
                Y[2 * ((size_t)k * dim + i) + 1] = yt[(size_t)i * r + k].y;
            }
    }
    TT_GUARD_END
}

}  // extern "C"
