// gfx950 dense solve of the debug path solve_dense=true (CPP/solve_faddeev.cpp:305-438,
// CPP/General_functions/matrix_routines.cpp:73-80): LU factorisation with partial pivoting of
// the complex FP64 matrix L = 1 - A diag(G_j) and the solve for the on-shell right-hand sides,
// on the device.
//
// Right-looking blocked LU on a column-major n x n matrix (leading dimension n), panels of
// NB = 64 columns:
//   * panel: for each column, the pivot (first row of largest |re| + |im| on or below the
//     diagonal, as izamax) from per-block maxima (k_colmax / k_eliminate) reduced by one
//     workgroup (k_pivot), which also swaps the two rows inside the panel; then the column is
//     scaled and the panel's later columns updated (k_eliminate, which also forms the next
//     column's block maxima);
//   * the panel's row interchanges on every other column (k_laswp);
//   * U12 = inv(L11) A12 and the trailing update A22 -= L21 U12 as complex FP64 GEMMs on the
//     MFMA (rocBLAS zgemm; inv(L11) by one workgroup, k_trinv);
//   * solve: the interchanges on the right-hand sides, then forward and backward substitution by
//     blocks with the inverted diagonal blocks and GEMM updates.
// The triangular diagonal blocks are inverted explicitly (unit lower L11; upper U11), the
// blocked form GPU LU codes use for their triangular solves: with partial pivoting |l_ij| <= 1.
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <vector>

#include "tt_internal.h"

namespace tt {

namespace {

constexpr int NB = 64;          // panel width
constexpr int MAXPARTS = 1024;  // block maxima per column search

__device__ inline double cabs1(double2 z) { return fabs(z.x) + fabs(z.y); }
__device__ inline double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ inline double2 crecip(double2 b) {
    const double d = b.x * b.x + b.y * b.y;
    return make_double2(b.x / d, -b.y / d);
}

// Larger value wins; equal values: the smaller row (the first maximum, as izamax).
__device__ inline void better(double &v, int64_t &i, double v2, int64_t i2) {
    if (v2 > v || (v2 == v && i2 < i)) {
        v = v2;
        i = i2;
    }
}

// Block reduction of (value, row) pairs over 256 threads into pv[blockIdx.x], pi[blockIdx.x].
__device__ void block_argmax(double v, int64_t i, double *pv, int64_t *pi) {
    __shared__ double sv[256];
    __shared__ int64_t si[256];
    sv[threadIdx.x] = v;
    si[threadIdx.x] = i;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = sv[threadIdx.x];
            int64_t ai = si[threadIdx.x];
            better(a, ai, sv[threadIdx.x + w], si[threadIdx.x + w]);
            sv[threadIdx.x] = a;
            si[threadIdx.x] = ai;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        pv[blockIdx.x] = sv[0];
        pi[blockIdx.x] = si[0];
    }
}

// Block maxima of |M[i][c]|_1 over rows i in [c, n).
__global__ __launch_bounds__(256) void k_colmax(const double2 *__restrict__ M, int64_t ld, int64_t n, int64_t c,
                                                double *pv, int64_t *pi) {
    double v = -1.0;
    int64_t idx = n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = c + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        better(v, idx, cabs1(M[c * ld + i]), i);
    block_argmax(v, idx, pv, pi);
}

// One workgroup: the pivot row p of column c from nparts block maxima; ipiv[c] = p; rows c and p
// swapped in the panel's columns [k0, k1).  A zero pivot column sets *singular.
__global__ __launch_bounds__(256) void k_pivot(double2 *__restrict__ M, int64_t ld, int64_t c, int64_t k0, int64_t k1,
                                               const double *pv, const int64_t *pi, int nparts, int64_t *ipiv,
                                               int *singular) {
    __shared__ double sv[256];
    __shared__ int64_t si[256];
    double v = -1.0;
    int64_t idx = INT64_MAX;
    for (int t = threadIdx.x; t < nparts; t += blockDim.x) better(v, idx, pv[t], pi[t]);
    sv[threadIdx.x] = v;
    si[threadIdx.x] = idx;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            double a = sv[threadIdx.x];
            int64_t ai = si[threadIdx.x];
            better(a, ai, sv[threadIdx.x + w], si[threadIdx.x + w]);
            sv[threadIdx.x] = a;
            si[threadIdx.x] = ai;
        }
        __syncthreads();
    }
    int64_t p = si[0];
    if (!(sv[0] > 0.0)) {   // zero (or NaN) column: no pivot
        p = c;
        if (threadIdx.x == 0) *singular = 1;
    }
    if (threadIdx.x == 0) ipiv[c] = p;
    if (p != c)
        for (int64_t j = k0 + threadIdx.x; j < k1; j += blockDim.x) {
            const double2 a = M[j * ld + c];
            M[j * ld + c] = M[j * ld + p];
            M[j * ld + p] = a;
        }
}

// Column c of the panel [k0, k1): rows i > c scaled by the pivot, the panel's columns (c, k1)
// updated by the rank-1 term; then block maxima of column c + 1 (when it is in the panel).
__global__ __launch_bounds__(256) void k_eliminate(double2 *__restrict__ M, int64_t ld, int64_t n, int64_t c,
                                                   int64_t k1, double *pv, int64_t *pi) {
    __shared__ double2 urow[NB];
    const int w = (int)(k1 - c);   // columns c .. k1-1, at most NB
    for (int j = threadIdx.x; j < w; j += blockDim.x) urow[j] = M[(c + j) * ld + c];
    __syncthreads();
    const double2 pinv = cabs1(urow[0]) > 0.0 ? crecip(urow[0]) : make_double2(1.0, 0.0);
    double v = -1.0;
    int64_t idx = n;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = c + 1 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double2 l = cmul(M[c * ld + i], pinv);
        M[c * ld + i] = l;
        for (int j = 1; j < w; ++j) {
            double2 m = M[(c + j) * ld + i];
            const double2 u = urow[j];
            m.x -= l.x * u.x - l.y * u.y;
            m.y -= l.x * u.y + l.y * u.x;
            M[(c + j) * ld + i] = m;
            if (j == 1) better(v, idx, cabs1(m), i);
        }
    }
    if (w > 1) block_argmax(v, idx, pv, pi);
}

// The row interchanges of the panel [k0, k1) on every column outside it, in order.
__global__ __launch_bounds__(256) void k_laswp(double2 *__restrict__ M, int64_t ld, int64_t n, int64_t k0,
                                               int64_t k1, const int64_t *__restrict__ ipiv) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n || (j >= k0 && j < k1)) return;
    double2 *col = M + j * ld;
    for (int64_t t = k0; t < k1; ++t) {
        const int64_t p = ipiv[t];
        if (p != t) {
            const double2 a = col[t];
            col[t] = col[p];
            col[p] = a;
        }
    }
}

// All row interchanges on the nrhs columns of B (leading dimension ld), in order.
__global__ void k_laswp_all(double2 *__restrict__ B, int64_t ld, int64_t n, int nrhs, const int64_t *__restrict__ ipiv) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nrhs) return;
    double2 *col = B + j * ld;
    for (int64_t t = 0; t < n; ++t) {
        const int64_t p = ipiv[t];
        if (p != t) {
            const double2 a = col[t];
            col[t] = col[p];
            col[p] = a;
        }
    }
}

// Inverse of the w x w diagonal block at (k0, k0) of M: unit lower triangular (lower = 1, the
// strictly lower part) or upper triangular (lower = 0), into W (column-major, leading dim NB).
// One workgroup of NB threads; thread j forms column j of the inverse by substitution.
__global__ __launch_bounds__(NB) void k_trinv(const double2 *__restrict__ M, int64_t ld, int64_t k0, int w, int lower,
                                              double2 *__restrict__ W) {
    __shared__ double2 T[NB * NB];   // T[r * NB + t] = M[k0 + r][k0 + t]
    __shared__ double2 X[NB * NB];   // X[t * NB + j] = row t of the inverse's column j
    for (int e = threadIdx.x; e < w * w; e += blockDim.x) {
        const int r = e % w, t = e / w;
        T[r * NB + t] = M[(k0 + t) * ld + k0 + r];
    }
    __syncthreads();
    // thread j forms column j by substitution, its partial results in LDS (the threads of a row t
    // touch consecutive words; the T values they read are the same word: a broadcast)
    const int j = threadIdx.x;
    if (j < w) {
        if (lower) {
            for (int r = 0; r < w; ++r) {
                double2 s = make_double2(r == j ? 1.0 : 0.0, 0.0);
                if (r > j)
                    for (int t = j; t < r; ++t) {
                        const double2 a = T[r * NB + t], b = X[t * NB + j];
                        s.x -= a.x * b.x - a.y * b.y;
                        s.y -= a.x * b.y + a.y * b.x;
                    }
                X[r * NB + j] = r < j ? make_double2(0.0, 0.0) : s;
            }
        } else {
            for (int r = w - 1; r >= 0; --r) {
                if (r > j) {
                    X[r * NB + j] = make_double2(0.0, 0.0);
                    continue;
                }
                double2 s = make_double2(r == j ? 1.0 : 0.0, 0.0);
                for (int t = r + 1; t <= j; ++t) {
                    const double2 a = T[r * NB + t], b = X[t * NB + j];
                    s.x -= a.x * b.x - a.y * b.y;
                    s.y -= a.x * b.y + a.y * b.x;
                }
                X[r * NB + j] = cmul(s, crecip(T[r * NB + r]));
            }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < w * w; e += blockDim.x) {   // W[j * NB + r] = column j, row r
        const int r = e % w, jj = e / w;
        W[(int64_t)jj * NB + r] = X[r * NB + jj];
    }
}

struct Blas {
    rocblas_handle h = nullptr;
    explicit Blas(hipStream_t s) {
        if (rocblas_create_handle(&h) != rocblas_status_success) fail(TT_ERR_HIP, "rocblas_create_handle failed");
        rocblas_set_stream(h, s);
    }
    ~Blas() {
        if (h) rocblas_destroy_handle(h);
    }
    // C = alpha A B + beta C (column-major, no transposes)
    void gemm(int64_t m, int64_t nn, int64_t k, double2 alpha, const double2 *A, int64_t lda, const double2 *B,
              int64_t ldb, double2 beta, double2 *C, int64_t ldc) {
        if (m <= 0 || nn <= 0 || k <= 0) return;
        const rocblas_status st = rocblas_zgemm(
            h, rocblas_operation_none, rocblas_operation_none, (rocblas_int)m, (rocblas_int)nn, (rocblas_int)k,
            reinterpret_cast<const rocblas_double_complex *>(&alpha), reinterpret_cast<const rocblas_double_complex *>(A),
            (rocblas_int)lda, reinterpret_cast<const rocblas_double_complex *>(B), (rocblas_int)ldb,
            reinterpret_cast<const rocblas_double_complex *>(&beta), reinterpret_cast<rocblas_double_complex *>(C),
            (rocblas_int)ldc);
        if (st != rocblas_status_success) fail(TT_ERR_HIP, "rocblas_zgemm failed: " + std::to_string((int)st));
    }
};

inline unsigned nblocks(int64_t n, int per) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

template <typename T>
struct Dev {
    T *p = nullptr;
    explicit Dev(size_t n) { TT_HIP(hipMalloc(&p, sizeof(T) * std::max<size_t>(n, 1))); }
    ~Dev() {
        if (p) (void)hipFree(p);
    }
    Dev(const Dev &) = delete;
    Dev &operator=(const Dev &) = delete;
};

}  // namespace

// M (n x n, column-major, leading dim n) is overwritten by its LU factors; B (n x nrhs,
// column-major, leading dim n) by M^-1 B.  Returns false when a pivot column is zero.
bool dense_lu_solve(hipStream_t s, double2 *M, int64_t n, double2 *B, int nrhs) {
    TT_REQUIRE(n > 0 && n < (int64_t)1 << 31 && nrhs > 0, "dense solve: bad sizes");
    const int64_t ld = n;
    const int64_t nblk_ = (n + NB - 1) / NB;
    Dev<int64_t> ipiv(n), pi(MAXPARTS);
    Dev<double> pv(MAXPARTS);
    Dev<int> singular(1);
    Dev<double2> Linv((size_t)nblk_ * NB * NB), Uinv((size_t)nblk_ * NB * NB), W((size_t)NB * n),
        T((size_t)NB * nrhs);
    TT_HIP(hipMemsetAsync(singular.p, 0, sizeof(int), s));
    Blas blas(s);
    const double2 one = make_double2(1.0, 0.0), zero = make_double2(0.0, 0.0), mone = make_double2(-1.0, 0.0);
    for (int64_t b = 0; b < nblk_; ++b) {
        const int64_t k0 = b * NB, k1 = std::min(n, k0 + NB);
        const int w = (int)(k1 - k0);
        {   // pivot of the panel's first column
            const unsigned g = std::min<unsigned>(MAXPARTS, nblocks(n - k0, 256));
            hipLaunchKernelGGL(k_colmax, dim3(g), dim3(256), 0, s, M, ld, n, k0, pv.p, pi.p);
            hipLaunchKernelGGL(k_pivot, dim3(1), dim3(256), 0, s, M, ld, k0, k0, k1, pv.p, pi.p, (int)g, ipiv.p,
                               singular.p);
        }
        for (int64_t c = k0; c < k1; ++c) {
            const unsigned g = std::min<unsigned>(MAXPARTS, nblocks(n - c - 1, 256));
            hipLaunchKernelGGL(k_eliminate, dim3(g), dim3(256), 0, s, M, ld, n, c, k1, pv.p, pi.p);
            if (c + 1 < k1)
                hipLaunchKernelGGL(k_pivot, dim3(1), dim3(256), 0, s, M, ld, c + 1, k0, k1, pv.p, pi.p, (int)g, ipiv.p,
                                   singular.p);
        }
        hipLaunchKernelGGL(k_laswp, dim3(nblocks(n, 256)), dim3(256), 0, s, M, ld, n, k0, k1, ipiv.p);
        double2 *Lb = Linv.p + (size_t)b * NB * NB, *Ub = Uinv.p + (size_t)b * NB * NB;
        hipLaunchKernelGGL(k_trinv, dim3(1), dim3(NB), 0, s, M, ld, k0, w, 1, Lb);
        if (k1 < n) {
            // U12 = inv(L11) A12 (via W), then A22 -= L21 U12
            blas.gemm(w, n - k1, w, one, Lb, NB, M + k1 * ld + k0, ld, zero, W.p, NB);
            TT_HIP(hipMemcpy2DAsync(M + k1 * ld + k0, sizeof(double2) * ld, W.p, sizeof(double2) * NB,
                                    sizeof(double2) * w, n - k1, hipMemcpyDeviceToDevice, s));
            blas.gemm(n - k1, n - k1, w, mone, M + k0 * ld + k1, ld, M + k1 * ld + k0, ld, one, M + k1 * ld + k1, ld);
        }
        hipLaunchKernelGGL(k_trinv, dim3(1), dim3(NB), 0, s, M, ld, k0, w, 0, Ub);
    }
    // B <- P B, then L y = B and U x = y by blocks
    hipLaunchKernelGGL(k_laswp_all, dim3(nblocks(nrhs, 64)), dim3(64), 0, s, B, ld, n, nrhs, ipiv.p);
    for (int64_t b = 0; b < nblk_; ++b) {
        const int64_t k0 = b * NB, k1 = std::min(n, k0 + NB);
        const int w = (int)(k1 - k0);
        blas.gemm(w, nrhs, w, one, Linv.p + (size_t)b * NB * NB, NB, B + k0, ld, zero, T.p, NB);
        TT_HIP(hipMemcpy2DAsync(B + k0, sizeof(double2) * ld, T.p, sizeof(double2) * NB, sizeof(double2) * w, nrhs,
                                hipMemcpyDeviceToDevice, s));
        blas.gemm(n - k1, nrhs, w, mone, M + k0 * ld + k1, ld, B + k0, ld, one, B + k1, ld);
    }
    for (int64_t b = nblk_ - 1; b >= 0; --b) {
        const int64_t k0 = b * NB, k1 = std::min(n, k0 + NB);
        const int w = (int)(k1 - k0);
        blas.gemm(w, nrhs, w, one, Uinv.p + (size_t)b * NB * NB, NB, B + k0, ld, zero, T.p, NB);
        TT_HIP(hipMemcpy2DAsync(B + k0, sizeof(double2) * ld, T.p, sizeof(double2) * NB, sizeof(double2) * w, nrhs,
                                hipMemcpyDeviceToDevice, s));
        blas.gemm(k0, nrhs, w, mone, M + k0 * ld, ld, B + k0, ld, one, B, ld);
    }
    int sing = 0;
    TT_HIP(hipMemcpyAsync(&sing, singular.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    return sing == 0;
}

}  // namespace tt
