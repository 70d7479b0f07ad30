// gfx950 build of the permutation matrix P123 in the wave-packet basis.
//
// Replaces fill_P123_arrays (CPP/make_permutation_matrix.cpp:1297-1403) and the
// column-major sort + CSC conversion of CPP/main.cpp:226-244:
//   k_mask_bra  bin-overlap test per (q',p',q,p) cell        (lines 160-253)
//               (k_mask, the per-cell form, where the bra-bin tables exceed 64 KB of LDS)
//   k_angular   angular basis F_t(phi, x) per bra bin        (CPP/auxiliary.cpp:950-1067)
//   k_quad      H[cell][t] = sum_{phi,x} F_t W_cell           (CPP/auxiliary.cpp:660-746)
//               W = w_x w_phi cos sin (kmax^2-kmin^2)/(2 zeta1 zeta2) / norm, staged through LDS
//   k_expand    P[cell][a'][a] = sum_Ltot A~ H[cell][t(a',a,Ltot)]
//   k_count / k_fill   CSC with rows ascending, exact zeros dropped (line 824)
// G-tilde is linear in the A-tilde table, so the (phi, x) quadrature is done once
// per distinct angular tuple (L', l', L, l, Ltot) instead of once per alpha pair.
// Quadrature is FP64 throughout.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cstring>

#include "tt_internal.h"

namespace tt {

// Y_l^m(theta, 0), 0 <= m <= l <= lmax, GSL_SF_LEGENDRE_SPHARM normalisation with the
// Condon-Shortley phase, at out[(l(l+1)/2 + m) * stride].
__host__ __device__ inline void spharm_table(int lmax, double x, double *out, int stride = 1) {
    const double s = sqrt(fmax(0.0, 1.0 - x * x));
    double ymm = 0.28209479177387814;  // 1/sqrt(4 pi)
    for (int m = 0; m <= lmax; ++m) {
        if (m > 0) ymm = -sqrt((2.0 * m + 1) / (2.0 * m)) * s * ymm;
        out[(m * (m + 1) / 2 + m) * stride] = ymm;
        if (m + 1 <= lmax) {
            double y1 = sqrt(2.0 * m + 3) * x * ymm;
            out[((m + 1) * (m + 2) / 2 + m) * stride] = y1;
            double ylm2 = ymm, ylm1 = y1;
            for (int l = m + 2; l <= lmax; ++l) {
                double a = sqrt((4.0 * l * l - 1) / ((double)l * l - (double)m * m));
                double b = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1) * (l - 1) - 1));
                double yl = a * (x * ylm1 - b * ylm2);
                out[(l * (l + 1) / 2 + m) * stride] = yl;
                ylm2 = ylm1;
                ylm1 = yl;
            }
        }
    }
}

__device__ inline double pi1t(double p, double q, double x) { return sqrt(0.25 * p * p + 9.0 / 16.0 * q * q + 0.75 * p * q * x); }
__device__ inline double pi2t(double p, double q, double x) { return sqrt(p * p + 0.25 * q * q - p * q * x); }

// mask[((qp*Np + pp)*Nq + q)*Np + p]: does any (phi, x) give a non-empty k interval
__global__ void k_mask(unsigned char *__restrict__ mask, const double *__restrict__ pe, const double *__restrict__ qe,
                       const double *__restrict__ x, const double *__restrict__ phi, int Np, int Nq, int Nx, int Nphi) {
    const int64_t ncell = (int64_t)Nq * Np * Nq * Np;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell) return;
    const int p = (int)(t % Np), q = (int)((t / Np) % Nq);
    const int pp = (int)((t / ((int64_t)Np * Nq)) % Np), qp = (int)(t / ((int64_t)Np * Nq * Np));
    const double ppl = pe[pp], ppu = pe[pp + 1], qpl = qe[qp], qpu = qe[qp + 1];
    const double pl = pe[p], pu = pe[p + 1], ql = qe[q], qu = qe[q + 1];
    const double *ph = phi + ((int64_t)qp * Np + pp) * Nphi;
    bool hit = false;
    for (int i = 0; i < Nphi && !hit; ++i) {
        double sp, cp;
        sincos(ph[i], &sp, &cp);
        const double kmin = fmax(ppl / sp, qpl / cp), kmax = fmin(ppu / sp, qpu / cp);
        if (!(kmin < kmax)) continue;
        for (int j = 0; j < Nx; ++j) {
            const double z1 = pi1t(sp, cp, x[j]), z2 = pi2t(sp, cp, x[j]);
            const double hl = fmax(pl / z1, ql / z2), hu = fmin(pu / z1, qu / z2);
            if (hl < hu && fmax(kmin, hl) < fmin(kmax, hu)) {
                hit = true;
                break;
            }
        }
    }
    mask[t] = hit ? 1 : 0;
}

// k_mask with the work that depends on the bra bin only done once per bra bin: one workgroup per
// (q', p'), whose k interval per phi node and pi1 / pi2 per (phi, x) node sit in LDS (the same
// expressions as k_mask, so the same bits), its threads over the (q, p) kets.  Dynamic LDS:
// 4 Nphi + 2 Nphi Nx doubles.
__global__ __launch_bounds__(256) void k_mask_bra(unsigned char *__restrict__ mask, const double *__restrict__ pe,
                                                  const double *__restrict__ qe, const double *__restrict__ x,
                                                  const double *__restrict__ phi, int Np, int Nq, int Nx, int Nphi) {
    extern __shared__ double lds[];
    double *sps = lds, *cps = sps + Nphi, *kmn = cps + Nphi, *kmx = kmn + Nphi;
    double *z1s = kmx + Nphi, *z2s = z1s + Nphi * Nx;   // [i][j]
    const int bra = blockIdx.x, qp = bra / Np, pp = bra - qp * Np;
    const double ppl = pe[pp], ppu = pe[pp + 1], qpl = qe[qp], qpu = qe[qp + 1];
    const double *ph = phi + (int64_t)bra * Nphi;
    for (int i = threadIdx.x; i < Nphi; i += blockDim.x) {
        double sp, cp;
        sincos(ph[i], &sp, &cp);
        sps[i] = sp;
        cps[i] = cp;
        kmn[i] = fmax(ppl / sp, qpl / cp);
        kmx[i] = fmin(ppu / sp, qpu / cp);
    }
    __syncthreads();
    for (int ij = threadIdx.x; ij < Nphi * Nx; ij += blockDim.x) {
        const int i = ij / Nx, j = ij - i * Nx;
        z1s[ij] = pi1t(sps[i], cps[i], x[j]);
        z2s[ij] = pi2t(sps[i], cps[i], x[j]);
    }
    __syncthreads();
    const int nket = Nq * Np;
    for (int c = threadIdx.x; c < nket; c += blockDim.x) {
        const int q = c / Np, p = c - q * Np;
        const double pl = pe[p], pu = pe[p + 1], ql = qe[q], qu = qe[q + 1];
        bool hit = false;
        for (int i = 0; i < Nphi && !hit; ++i) {
            const double kmin = kmn[i], kmax = kmx[i];
            if (!(kmin < kmax)) continue;
            for (int j = 0; j < Nx; ++j) {
                const double z1 = z1s[i * Nx + j], z2 = z2s[i * Nx + j];
                const double hl = fmax(pl / z1, ql / z2), hu = fmin(pu / z1, qu / z2);
                if (hl < hu && fmax(kmin, hl) < fmin(kmax, hu)) {
                    hit = true;
                    break;
                }
            }
        }
        mask[(int64_t)bra * nket + c] = hit ? 1 : 0;
    }
}

// F[b][t][k] for the bra bins rb0 + b, k = iphi*Nx + ix (points fastest, so a wave's stores are
// one contiguous row).  Block: (chunk of kc = 64 or 128 points, b), 16 waves.  In LDS, point-fastest
// so that a wave's reads of one table entry are conflict-free: the chunk's tables Y2[i][kk],
// Y3[i][kk], and Y1[i][ix] of all Nx x nodes.
constexpr int ANG_THREADS = 1024;
template <bool PAIR>
__global__ __launch_bounds__(ANG_THREADS) void k_angular(double *__restrict__ F, int rb0, int kc,
                                                         const double *__restrict__ phi, const double *__restrict__ x,
                                                         const double *__restrict__ Y1tab, int n1, int lmax2, int n2,
                                                         int lmax3, int n3, int K, int Nx, int Nphi, int T,
                                                         const int *__restrict__ term_ptr,
                                                         const double *__restrict__ coef,
                                                         const int4 *__restrict__ tidx) {
    extern __shared__ double lds[];
    double *Y2 = lds;                    // n2 * kc
    double *Y3 = lds + n2 * kc;          // n3 * kc
    double *Y1 = Y3 + n3 * kc;           // n1 * Nx
    const int b = blockIdx.y;
    const int rb = rb0 + b;
    const int k0 = blockIdx.x * kc;
    for (int i = threadIdx.x; i < n1 * Nx; i += blockDim.x) {
        const int i1 = i / Nx, ix = i - i1 * Nx;
        Y1[i] = Y1tab[ix * n1 + i1];
    }
    if (threadIdx.x < 2 * kc) {
        const int kk = threadIdx.x % kc, which = threadIdx.x / kc;
        const int k = min(k0 + kk, K - 1);   // points past K get tables too (their F is not stored)
        const int iphi = k / Nx, ix = k % Nx;
        double sp, cp;
        sincos(phi[(int64_t)rb * Nphi + iphi], &sp, &cp);
        const double xx = x[ix];
        if (which == 0) {
            const double z1 = pi1t(sp, cp, xx);
            const double ct1 = fmin(1.0, fmax(-1.0, -(0.5 * sp + 0.75 * cp * xx) / z1));
            spharm_table(lmax2, ct1, Y2 + kk, kc);
        } else {
            const double z2 = pi2t(sp, cp, xx);
            const double ct2 = fmin(1.0, fmax(-1.0, (sp - 0.5 * cp * xx) / z2));
            spharm_table(lmax3, ct2, Y3 + kk, kc);
        }
    }
    __syncthreads();
    // one tuple per wave at a time, its lanes on 64 of the chunk's points: every lane of a wave runs
    // the same term list, whose data (coefficient, Y1 row offset pre-scaled by Nx, Y2 / Y3 offsets
    // pre-scaled by kc) are wave-uniform scalar loads; eight terms' table reads are issued before
    // their products are added (in term order, as a plain loop would)
    const int lane = threadIdx.x & 63, nh = kc >> 6;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // uniform: term data by scalar loads
    if constexpr (PAIR) {
        // Two adjacent points per lane (even Nx, chunks of 128 points), one tuple per wave over the
        // whole chunk: each table read fetches both points' entries (16 bytes per lane, the pair
        // never straddles a Y1 row since k and Nx are even), so the LDS instructions and the scalar
        // term loads per point are halved.  Each point's sum runs over the terms in the same order
        // and form as below, so F is the same to the bit.
        const int kk = 2 * lane, k = k0 + kk;
        const double *y1 = Y1 + min(k, K - 2) % Nx;
        const double *y2 = Y2 + kk, *y3 = Y3 + kk;
        for (int t = wave; t < T; t += blockDim.x >> 6) {
            const int j0 = term_ptr[t], j1 = term_ptr[t + 1];
            double acc0 = 0.0, acc1 = 0.0;
            int j = j0;
            for (; j + 4 <= j1; j += 4) {
                double2 a[4], bb[4], d[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int4 q = tidx[j + u];
                    a[u] = *reinterpret_cast<const double2 *>(y1 + q.x);
                    bb[u] = *reinterpret_cast<const double2 *>(y2 + q.y);
                    d[u] = *reinterpret_cast<const double2 *>(y3 + q.z);
                }
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    acc0 += coef[j + u] * a[u].x * bb[u].x * d[u].x;
                    acc1 += coef[j + u] * a[u].y * bb[u].y * d[u].y;
                }
            }
            for (; j < j1; ++j) {
                const int4 q = tidx[j];
                const double2 a = *reinterpret_cast<const double2 *>(y1 + q.x);
                const double2 bb = *reinterpret_cast<const double2 *>(y2 + q.y);
                const double2 d = *reinterpret_cast<const double2 *>(y3 + q.z);
                acc0 += coef[j] * a.x * bb.x * d.x;
                acc1 += coef[j] * a.y * bb.y * d.y;
            }
            if (k < K) *reinterpret_cast<double2 *>(F + ((int64_t)b * T + t) * K + k) = make_double2(acc0, acc1);
        }
        return;
    }
    const int kk = (wave % nh) * 64 + lane, k = k0 + kk;
    const double *y1 = Y1 + min(k, K - 1) % Nx;
    const double *y2 = Y2 + kk, *y3 = Y3 + kk;
    for (int t = wave / nh; t < T; t += (blockDim.x >> 6) / nh) {
        const int j0 = term_ptr[t], j1 = term_ptr[t + 1];
        double acc = 0.0;
        int j = j0;
        for (; j + 8 <= j1; j += 8) {
            double a[8], bb[8], d[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int4 q = tidx[j + u];
                a[u] = y1[q.x];
                bb[u] = y2[q.y];
                d[u] = y3[q.z];
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += coef[j + u] * a[u] * bb[u] * d[u];
        }
        for (; j < j1; ++j) {
            const int4 q = tidx[j];
            acc += coef[j] * y1[q.x] * y2[q.y] * y3[q.z];
        }
        if (k < K) F[((int64_t)b * T + t) * K + k] = acc;
    }
}

// Per quadrature point of every bra bin: the bra's k range max(p'_lo/sin, q'_lo/cos) ..
// min(p'_hi/sin, q'_hi/cos), 1/zeta1, 1/zeta2 and the weight w_x w_phi cos sin / (2 zeta1 zeta2),
// so that the cell weights of k_quad are products and comparisons only.
struct QPoint {
    double klo, khi, iz1, iz2, f;
};
__global__ void k_qpoints(QPoint *__restrict__ Q, int NqNp, int K, int Nx, int Nphi, int Np,
                          const double *__restrict__ pe, const double *__restrict__ qe, const double *__restrict__ x,
                          const double *__restrict__ wx, const double *__restrict__ phi,
                          const double *__restrict__ wphi) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)NqNp * K) return;
    const int rb = (int)(t / K), k = (int)(t % K);
    const int qp = rb / Np, pp = rb % Np, iphi = k / Nx, ix = k % Nx;
    double sp, cp;
    sincos(phi[(int64_t)rb * Nphi + iphi], &sp, &cp);
    const double z1 = pi1t(sp, cp, x[ix]), z2 = pi2t(sp, cp, x[ix]);
    QPoint o;
    o.klo = fmax(pe[pp] / sp, qe[qp] / cp);
    o.khi = fmin(pe[pp + 1] / sp, qe[qp + 1] / cp);
    o.iz1 = 1.0 / z1;
    o.iz2 = 1.0 / z2;
    o.f = wx[ix] * wphi[(int64_t)rb * Nphi + iphi] * cp * sp / (2 * z1 * z2);
    Q[t] = o;
}

// H[cell][t] = sum_k F[b][t][k] W[cell][k] / norm(cell) for the cells of bra bin rb0 + b, with
// W[cell][k] = f_k (kmax^2 - kmin^2) where the bra's and the ket's k ranges overlap (else 0),
// kmin = max(klo_k, p_lo/zeta1, q_lo/zeta2), kmax = min(khi_k, p_hi/zeta1, q_hi/zeta2).
// On the FP64 MFMA (16x16x4): wave w of the block takes the 16 tuples t0 = 128 blockIdx.x + 16 w
// (the M tile: A = F rows, each lane loading four consecutive points per 16-point group, which
// the four MFMA steps of the group take one at a time), against the bra's cells in groups of up to
// 64 (four N tiles: B = W from LDS, staged per chunk of QKC points by the whole block).  The sum
// over k runs in any order that pairs each F value with its own W value: lane slot j of step s
// in group g takes point 16 g + 4 j + s on both sides.
constexpr int QKC = 32;   // quadrature points per LDS stage
constexpr int QCG = 64;   // cells per group (four 16-wide N tiles)
constexpr int QUAD_THREADS = 256;
typedef double v4d __attribute__((ext_vector_type(4)));
// The chunk's MFMA steps against NNT tiles of 16 cells: step s of group g pairs the lane's F value
// a[g][s] (tuple row, point 16 g + 4 slot + s) with W of cell 16 nt + row at the same point.
template <int NNT>
__device__ __forceinline__ void quad_mfma(v4d (&acc)[4], const double (&a)[QKC / 16][4],
                                          const double (*__restrict__ Wl)[QKC + 1], int row, int slot) {
#pragma unroll
    for (int g = 0; g < QKC / 16; ++g)
#pragma unroll
        for (int s = 0; s < 4; ++s)
#pragma unroll
            for (int nt = 0; nt < NNT; ++nt)
                acc[nt] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[g][s], Wl[16 * nt + row][16 * g + 4 * slot + s],
                                                               acc[nt], 0, 0, 0);
}
__global__ __launch_bounds__(QUAD_THREADS) void k_quad(double *__restrict__ H, const double *__restrict__ F, int rb0,
                                                       int T, int K, int Np, const int *__restrict__ rb_ptr,
                                                       const int *__restrict__ rb_cells,
                                                       const int *__restrict__ cell_ket, const double *__restrict__ pe,
                                                       const double *__restrict__ qe, const QPoint *__restrict__ Q) {
    __shared__ double Wl[QCG][QKC + 1];
    __shared__ double sk[5][QKC];
    __shared__ double cb[4][QCG];
    const int b = blockIdx.y, rb = rb0 + b;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int row = lane & 15, slot = lane >> 4;
    const int t0 = blockIdx.x * (QUAD_THREADS / 4) + wave * 16;
    const bool active = t0 < T;   // waves past T take part in the staging only
    const double *Fr = F + ((int64_t)b * T + min(t0 + row, T - 1)) * K;   // rows past T: not stored
    const double pl = pe[rb % Np], pu = pe[rb % Np + 1], ql = qe[rb / Np], qu = qe[rb / Np + 1];
    const QPoint *Qb = Q + (int64_t)rb * K;
    const bool even_K = (K & 1) == 0;   // rows of F start 16-byte aligned
    const int c_beg = rb_ptr[rb], c_end = rb_ptr[rb + 1];
    for (int cg0 = c_beg; cg0 < c_end; cg0 += QCG) {
        const int ncg = min(QCG, c_end - cg0);
        const int nnt = (ncg + 15) >> 4;
        __syncthreads();   // the previous group's cell bounds are no longer read
        if (threadIdx.x < QCG) {
            const int c = threadIdx.x;
            if (c < ncg) {
                const int ket = cell_ket[rb_cells[cg0 + c]];
                const int q = ket / Np, p = ket % Np;
                cb[0][c] = pe[p];
                cb[1][c] = pe[p + 1];
                cb[2][c] = qe[q];
                cb[3][c] = qe[q + 1];
            } else {   // an empty range: W = 0
                cb[0][c] = 1.0;
                cb[1][c] = 0.0;
                cb[2][c] = 1.0;
                cb[3][c] = 0.0;
            }
        }
        v4d acc[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) acc[i] = v4d{0.0, 0.0, 0.0, 0.0};
        for (int k0 = 0; k0 < K; k0 += QKC) {
            // this chunk's F values of the lane (in flight while W is formed)
            double a[QKC / 16][4];
#pragma unroll
            for (int g = 0; g < QKC / 16; ++g) {
                const int kb = k0 + 16 * g + 4 * slot;
                if (!active) {
#pragma unroll
                    for (int s = 0; s < 4; ++s) a[g][s] = 0.0;
                } else if (even_K && kb + 3 < K) {   // 16-byte aligned pairs
                    const double2 lo = *reinterpret_cast<const double2 *>(Fr + kb);
                    const double2 hi = *reinterpret_cast<const double2 *>(Fr + kb + 2);
                    a[g][0] = lo.x;
                    a[g][1] = lo.y;
                    a[g][2] = hi.x;
                    a[g][3] = hi.y;
                } else {
#pragma unroll
                    for (int s = 0; s < 4; ++s) a[g][s] = kb + s < K ? Fr[kb + s] : 0.0;
                }
            }
            __syncthreads();   // the previous chunk's W is no longer read
            if (threadIdx.x < QKC) {
                const int k = k0 + threadIdx.x;
                if (k < K) {
                    const QPoint qp = Qb[k];
                    sk[0][threadIdx.x] = qp.klo;
                    sk[1][threadIdx.x] = qp.khi;
                    sk[2][threadIdx.x] = qp.iz1;
                    sk[3][threadIdx.x] = qp.iz2;
                    sk[4][threadIdx.x] = qp.f;
                } else {
                    sk[0][threadIdx.x] = 1.0;
                    sk[1][threadIdx.x] = 0.0;
                    sk[2][threadIdx.x] = sk[3][threadIdx.x] = sk[4][threadIdx.x] = 0.0;
                }
            }
            __syncthreads();
            for (int e = threadIdx.x; e < QCG * QKC; e += blockDim.x) {
                const int c = e / QKC, kk = e % QKC;
                const double iz1 = sk[2][kk], iz2 = sk[3][kk];
                const double kmin = fmax(fmax(sk[0][kk], cb[0][c] * iz1), cb[2][c] * iz2);
                const double kmax = fmin(fmin(sk[1][kk], cb[1][c] * iz1), cb[3][c] * iz2);
                Wl[c][kk] = kmin > kmax ? 0.0 : sk[4][kk] * (kmax * kmax - kmin * kmin);
            }
            __syncthreads();
            if (active) switch (nnt) {   // branch-free MFMA chains per tile count
                case 1: quad_mfma<1>(acc, a, Wl, row, slot); break;
                case 2: quad_mfma<2>(acc, a, Wl, row, slot); break;
                case 3: quad_mfma<3>(acc, a, Wl, row, slot); break;
                default: quad_mfma<4>(acc, a, Wl, row, slot); break;
            }
        }
        // the FP64 MFMA's lane holds D[slot + 4 i][row]: tuple t0 + slot + 4 i, cell 16 nt + row
        if (active) {
#pragma unroll
            for (int nt = 0; nt < 4; ++nt) {
                const int c = 16 * nt + row;
                if (nt >= nnt || c >= ncg) continue;
                const int cell = rb_cells[cg0 + c];
                const double norm = sqrt(cb[3][c] - cb[2][c]) * sqrt(cb[1][c] - cb[0][c]) * sqrt(qu - ql) * sqrt(pu - pl);
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    const int t = t0 + slot + 4 * i;
                    if (t < T) H[(int64_t)cell * T + t] = acc[nt][i] / norm;
                }
            }
        }
    }
}

// V[cell][a'][a] = sum_e A~_e H[cell][t_e]; placeholder cos(row) cos(col) when production == 0.
__global__ void k_expand(double *__restrict__ V, const double *__restrict__ H, int64_t ncell, int Na, int T,
                         const int *__restrict__ pair_ptr, const int *__restrict__ pair_tuple,
                         const double *__restrict__ pair_at, const int *__restrict__ cell_bra,
                         const int *__restrict__ cell_ket, int NqNp, int production) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t npair = (int64_t)Na * Na;
    if (t >= ncell * npair) return;
    const int64_t cell = t / npair;
    const int pr = (int)(t % npair);
    double v = 0.0;
    if (production) {
        for (int e = pair_ptr[pr]; e < pair_ptr[pr + 1]; ++e) v += pair_at[e] * H[cell * T + pair_tuple[e]];
    } else {
        const int ap = pr / Na, a = pr % Na;
        const double row = (double)((int64_t)ap * NqNp + cell_bra[cell]);
        const double col = (double)((int64_t)a * NqNp + cell_ket[cell]);
        v = cos(row) * cos(col);
    }
    V[t] = v;
}

// Column c = a * NqNp + ket: rows (a', bra) for a' ascending, bras ascending (cells of a ket are
// contiguous).  One thread per column with a fastest across threads (t = ket * Na + a), so that a
// wave reads each (cell, a') row of V in one contiguous pass.
__global__ void k_count(uint64_t *__restrict__ cnt, const double *__restrict__ V, const int *__restrict__ ket_ptr,
                        int Na, int NqNp) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Na * NqNp) return;
    const int ket = (int)(t / Na), a = (int)(t % Na);
    const int64_t c = (int64_t)a * NqNp + ket;
    uint64_t n = 0;
    for (int ap = 0; ap < Na; ++ap)
        for (int cell = ket_ptr[ket]; cell < ket_ptr[ket + 1]; ++cell)
            n += V[((int64_t)cell * Na + ap) * Na + a] != 0.0;
    cnt[c] = n;
}

__global__ void k_fill(double *__restrict__ val, int32_t *__restrict__ row, const uint64_t *__restrict__ colptr,
                       const double *__restrict__ V, const int *__restrict__ ket_ptr, const int *__restrict__ cell_bra,
                       int Na, int NqNp) {
    // one thread per column c with ket fastest: neighbouring threads write neighbouring columns'
    // ranges of the CSC (k_count's a-fastest order reads better, but with it this kernel's
    // scattered writes took 62 against 52 ms for the 18 channels of 2J <= 17 at Np = 30)
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int64_t)Na * NqNp) return;
    const int a = (int)(c / NqNp), ket = (int)(c % NqNp);
    uint64_t o = colptr[c];
    for (int ap = 0; ap < Na; ++ap)
        for (int cell = ket_ptr[ket]; cell < ket_ptr[ket + 1]; ++cell) {
            const double v = V[((int64_t)cell * Na + ap) * Na + a];
            if (v != 0.0) {
                val[o] = v;
                row[o] = ap * NqNp + cell_bra[cell];
                ++o;
            }
        }
}

template <typename T>
struct DBuf {
    T *p = nullptr;
    explicit DBuf(size_t n) {
        if (n) TT_HIP(hipMalloc(&p, sizeof(T) * n));
    }
    ~DBuf() {
        if (p) (void)hipFree(p);
    }
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
};

template <typename T>
static void h2d(T *d, const std::vector<T> &h, hipStream_t s) {
    if (!h.empty()) TT_HIP(hipMemcpyAsync(d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

static inline unsigned nb(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

}  // namespace tt

using namespace tt;

extern "C" tt_status tt_fill_P123_arrays(tt_ctx *ctx, const tt_run_params *rp, const tt_fwp_states *fwp,
                                         const tt_pw_states *chn, int production_run, tt_csc *out,
                                         tt_p123_stats *stats) {
    TT_GUARD_BEGIN
    TT_REQUIRE(ctx && rp && fwp && chn && out && chn->Nalpha > 0, "null argument");
    TT_HIP(hipSetDevice(ctx->device));
    auto t0 = std::chrono::steady_clock::now();
    hipStream_t s = ctx->stream;
    const int Np = fwp->Np_WP, Nq = fwp->Nq_WP, Nx = rp->Nx, Nphi = rp->Nphi, Na = chn->Nalpha;
    const int NqNp = Nq * Np, K = Nphi * Nx;
    const int64_t dim = (int64_t)Na * NqNp;
    TT_REQUIRE(dim < (int64_t)INT32_MAX, "dimension exceeds int32 row indices");
    // quadrature meshes: x = gauss(Nx) on [-1,1] (line 1330); phi per bra bin (lines 176-183)
    std::vector<double> x(Nx), wx(Nx), g(Nphi), wg(Nphi), phi((size_t)NqNp * Nphi), wphi((size_t)NqNp * Nphi);
    gauss_legendre(Nx, x.data(), wx.data());
    gauss_legendre(Nphi, g.data(), wg.data());
    for (int qp = 0; qp < Nq; ++qp)
        for (int pp = 0; pp < Np; ++pp) {
            double lo = std::atan(fwp->p_WP[pp] / fwp->q_WP[qp + 1]);
            double hi = fwp->q_WP[qp] > 0 ? std::atan(fwp->p_WP[pp + 1] / fwp->q_WP[qp]) : M_PI / 2;
            for (int i = 0; i < Nphi; ++i) {
                phi[((size_t)qp * Np + pp) * Nphi + i] = 0.5 * (hi - lo) * g[i] + 0.5 * (hi + lo);
                wphi[((size_t)qp * Np + pp) * Nphi + i] = 0.5 * (hi - lo) * wg[i];
            }
        }
    std::vector<double> pe(fwp->p_WP, fwp->p_WP + Np + 1), qe(fwp->q_WP, fwp->q_WP + Nq + 1);
    DBuf<double> d_pe(Np + 1), d_qe(Nq + 1), d_x(Nx), d_wx(Nx), d_phi(phi.size()), d_wphi(wphi.size());
    h2d(d_pe.p, pe, s);
    h2d(d_qe.p, qe, s);
    h2d(d_x.p, x, s);
    h2d(d_wx.p, wx, s);
    h2d(d_phi.p, phi, s);
    h2d(d_wphi.p, wphi, s);

    // (a1) overlap mask
    const int64_t ncell_all = (int64_t)NqNp * NqNp;
    DBuf<unsigned char> d_mask(ncell_all);
    const size_t mask_lds = sizeof(double) * ((size_t)4 * Nphi + (size_t)2 * Nphi * Nx);
    if (mask_lds <= 64 * 1024)
        hipLaunchKernelGGL(k_mask_bra, dim3(NqNp), dim3(256), mask_lds, s, d_mask.p, d_pe.p, d_qe.p, d_x.p, d_phi.p,
                           Np, Nq, Nx, Nphi);
    else   // angular grids too fine for the bra-bin tables: the per-cell form
        hipLaunchKernelGGL(k_mask, dim3(nb(ncell_all, 256)), dim3(256), 0, s, d_mask.p, d_pe.p, d_qe.p, d_x.p, d_phi.p,
                           Np, Nq, Nx, Nphi);
    TT_HIP(hipGetLastError());
    std::vector<unsigned char> mask(ncell_all);
    TT_HIP(hipMemcpyAsync(mask.data(), d_mask.p, ncell_all, hipMemcpyDeviceToHost, s));
    TT_HIP(hipStreamSynchronize(s));
    // cells ordered ket-major, bra-minor; per-bra lists of global cell ids
    std::vector<int> cell_bra, cell_ket, ket_ptr(NqNp + 1, 0), rb_ptr(NqNp + 1, 0);
    for (int ket = 0; ket < NqNp; ++ket) {
        for (int bra = 0; bra < NqNp; ++bra)
            if (mask[(size_t)bra * NqNp + ket]) {
                cell_bra.push_back(bra);
                cell_ket.push_back(ket);
                rb_ptr[bra + 1]++;
            }
        ket_ptr[ket + 1] = (int)cell_bra.size();
    }
    const int64_t ncell = (int64_t)cell_bra.size();
    for (int i = 0; i < NqNp; ++i) rb_ptr[i + 1] += rb_ptr[i];
    std::vector<int> rb_cells(ncell), fillpos(rb_ptr.begin(), rb_ptr.end() - 1);
    for (int64_t c = 0; c < ncell; ++c) rb_cells[fillpos[cell_bra[c]]++] = (int)c;

    // angular tuples (a2, a3)
    P123Angular ang = build_p123_angular(*chn);
    const int T = std::max(ang.n_tuples, 1);
    const int n1 = (ang.lmax1 + 1) * (ang.lmax1 + 2) / 2, n2 = (ang.lmax2 + 1) * (ang.lmax2 + 2) / 2,
              n3 = (ang.lmax3 + 1) * (ang.lmax3 + 2) / 2;
    std::vector<double> Y1((size_t)Nx * n1);
    for (int i = 0; i < Nx; ++i) spharm_table(ang.lmax1, x[i], &Y1[(size_t)i * n1]);
    DBuf<double> d_Y1(Y1.size()), d_coef(std::max<size_t>(ang.term_coef.size(), 1));
    // k_angular's chunk of points: 128 where its LDS tables fit, else 64
    auto ang_lds = [&](int kc) { return (size_t)(kc * (n2 + n3) + n1 * Nx) * sizeof(double); };
    const int kc = ang_lds(128) <= 160 * 1024 ? 128 : 64;
    TT_REQUIRE(ang_lds(kc) <= 160 * 1024, "angular tables exceed LDS");
    // term data of k_angular: offsets in the point-fastest LDS tables
    std::vector<int4> tidx(std::max<size_t>(ang.term_i1.size(), 1));
    for (size_t j = 0; j < ang.term_i1.size(); ++j)
        tidx[j] = make_int4(ang.term_i1[j] * Nx, ang.term_i2[j] * kc, ang.term_i3[j] * kc, 0);
    DBuf<int> d_tptr(ang.term_ptr.size());
    DBuf<int4> d_tidx(tidx.size());
    h2d(d_Y1.p, Y1, s);
    h2d(d_coef.p, ang.term_coef, s);
    h2d(d_tptr.p, ang.term_ptr, s);
    h2d(d_tidx.p, tidx, s);
    DBuf<int> d_rbptr(rb_ptr.size()), d_rbcells(std::max<int64_t>(ncell, 1)), d_cbra(std::max<int64_t>(ncell, 1)),
        d_cket(std::max<int64_t>(ncell, 1)), d_ketptr(ket_ptr.size());
    h2d(d_rbptr.p, rb_ptr, s);
    h2d(d_rbcells.p, rb_cells, s);
    h2d(d_cbra.p, cell_bra, s);
    h2d(d_cket.p, cell_ket, s);
    h2d(d_ketptr.p, ket_ptr, s);
    DBuf<double> d_H(std::max<int64_t>(ncell * T, 1));
    TT_HIP(hipMemsetAsync(d_H.p, 0, sizeof(double) * std::max<int64_t>(ncell * T, 1), s));

    // (a3, a4) angular basis and quadrature, bra bins in batches bounded to ~512 MiB of F
    if (production_run && ang.n_tuples > 0 && ncell > 0) {
        int batch = (int)std::max<int64_t>(1, std::min<int64_t>(NqNp, (512ll << 20) / ((int64_t)K * T * 8)));
        DBuf<double> d_F((size_t)batch * K * T);
        DBuf<QPoint> d_Q((size_t)NqNp * K);
        hipLaunchKernelGGL(k_qpoints, dim3(nb((int64_t)NqNp * K, 256)), dim3(256), 0, s, d_Q.p, NqNp, K, Nx, Nphi, Np,
                           d_pe.p, d_qe.p, d_x.p, d_wx.p, d_phi.p, d_wphi.p);
        const int tpb = QUAD_THREADS / 4;   // tuples per k_quad block
        // two points per lane where the pairs stay inside a Y1 row and the chunk is one wave's
        const bool ang_pair = Nx % 2 == 0 && kc == 128;
        for (int rb0 = 0; rb0 < NqNp; rb0 += batch) {
            const int nrb = std::min(batch, NqNp - rb0);
            hipLaunchKernelGGL(ang_pair ? k_angular<true> : k_angular<false>, dim3(nb(K, kc), nrb), dim3(ANG_THREADS),
                               ang_lds(kc), s, d_F.p, rb0, kc, d_phi.p, d_x.p, d_Y1.p, n1, ang.lmax2, n2, ang.lmax3,
                               n3, K, Nx, Nphi, T, d_tptr.p, d_coef.p, d_tidx.p);
            hipLaunchKernelGGL(k_quad, dim3(nb(T, tpb), nrb), dim3(QUAD_THREADS), 0, s, d_H.p, d_F.p, rb0, T, K, Np,
                               d_rbptr.p, d_rbcells.p, d_cket.p, d_pe.p, d_qe.p, d_Q.p);
        }
        TT_HIP(hipGetLastError());
    }
    // (a4) alpha-pair expansion
    DBuf<int> d_pptr(ang.pair_ptr.size()), d_ptup(std::max<size_t>(ang.pair_tuple.size(), 1));
    DBuf<double> d_pat(std::max<size_t>(ang.pair_atilde.size(), 1));
    h2d(d_pptr.p, ang.pair_ptr, s);
    h2d(d_ptup.p, ang.pair_tuple, s);
    h2d(d_pat.p, ang.pair_atilde, s);
    const int64_t nval = ncell * Na * Na;
    // the block array becomes the cell-block form of the result (owned by the DeviceCSC)
    double *d_Vraw = nullptr;
    TT_HIP(hipMalloc(&d_Vraw, sizeof(double) * (nval + BLK_SLACK)));
    struct Own {
        double *p;
        ~Own() {
            if (p) (void)hipFree(p);
        }
    } d_V{d_Vraw};
    TT_HIP(hipMemsetAsync(d_V.p + nval, 0, sizeof(double) * BLK_SLACK, s));
    if (nval)
        hipLaunchKernelGGL(k_expand, dim3(nb(nval, 256)), dim3(256), 0, s, d_V.p, d_H.p, ncell, Na, T, d_pptr.p,
                           d_ptup.p, d_pat.p, d_cbra.p, d_cket.p, NqNp, production_run);
    // (a5) CSC: count, scan, fill
    auto *csc = new DeviceCSC;
    try {
        csc->dim = dim;
        TT_HIP(hipMalloc(&csc->colptr, sizeof(uint64_t) * (dim + 1)));
        DBuf<uint64_t> d_cnt(dim + 1);
        TT_HIP(hipMemsetAsync(d_cnt.p, 0, sizeof(uint64_t) * (dim + 1), s));
        hipLaunchKernelGGL(k_count, dim3(nb(dim, 256)), dim3(256), 0, s, d_cnt.p, d_V.p, d_ketptr.p, Na, NqNp);
        size_t tmp_bytes = 0;
        TT_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, d_cnt.p, csc->colptr, dim + 1, s));
        DBuf<unsigned char> d_tmp(std::max<size_t>(tmp_bytes, 1));
        TT_HIP(hipcub::DeviceScan::ExclusiveSum(d_tmp.p, tmp_bytes, d_cnt.p, csc->colptr, dim + 1, s));
        uint64_t nnz = 0;
        TT_HIP(hipMemcpyAsync(&nnz, csc->colptr + dim, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
        TT_HIP(hipStreamSynchronize(s));
        csc->nnz = (int64_t)nnz;
        TT_HIP(hipMalloc(&csc->val, sizeof(double) * std::max<uint64_t>(nnz, 1)));
        TT_HIP(hipMalloc(&csc->row, sizeof(int32_t) * std::max<uint64_t>(nnz, 1)));
        hipLaunchKernelGGL(k_fill, dim3(nb(dim, 256)), dim3(256), 0, s, csc->val, csc->row, csc->colptr, d_V.p,
                           d_ketptr.p, d_cbra.p, Na, NqNp);
        TT_HIP(hipGetLastError());
        // cell-block form for the A-application
        csc->Na = Na;
        csc->NqNp = NqNp;
        csc->ncell = ncell;
        for (int k = 0; k < NqNp; ++k) csc->max_ket_cells = std::max(csc->max_ket_cells, ket_ptr[k + 1] - ket_ptr[k]);
        TT_HIP(hipMalloc(&csc->ket_ptr, sizeof(int) * (NqNp + 1)));
        TT_HIP(hipMalloc(&csc->cell_bra, sizeof(int) * std::max<int64_t>(ncell, 1)));
        TT_HIP(hipMemcpyAsync(csc->ket_ptr, d_ketptr.p, sizeof(int) * (NqNp + 1), hipMemcpyDeviceToDevice, s));
        if (ncell) TT_HIP(hipMemcpyAsync(csc->cell_bra, d_cbra.p, sizeof(int) * ncell, hipMemcpyDeviceToDevice, s));
        csc->blk = d_V.p;
        d_V.p = nullptr;
        TT_HIP(hipStreamSynchronize(s));
    } catch (...) {
        delete csc;
        throw;
    }
    std::memset(out, 0, sizeof(*out));
    out->dim = dim;
    out->nnz = csc->nnz;
    out->dev = csc;
    if (stats) {
        stats->mask_density = (double)ncell / (double)ncell_all;
        stats->n_cells = ncell;
        stats->n_tuples = ang.n_tuples;
        stats->ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    TT_GUARD_END
}
