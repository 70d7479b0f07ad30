// gfx950 build of the permutation matrix P123 in the wave-packet basis.
//
// Replaces fill_P123_arrays (CPP/make_permutation_matrix.cpp:1297-1403) and the
// column-major sort + CSC conversion of CPP/main.cpp:226-244:
//   k_mask      bin-overlap test per (q',p',q,p) cell        (lines 160-253)
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
// Condon-Shortley phase, index l(l+1)/2 + m.
__host__ __device__ inline void spharm_table(int lmax, double x, double *out) {
    const double s = sqrt(fmax(0.0, 1.0 - x * x));
    double ymm = 0.28209479177387814;  // 1/sqrt(4 pi)
    for (int m = 0; m <= lmax; ++m) {
        if (m > 0) ymm = -sqrt((2.0 * m + 1) / (2.0 * m)) * s * ymm;
        out[m * (m + 1) / 2 + m] = ymm;
        if (m + 1 <= lmax) {
            double y1 = sqrt(2.0 * m + 3) * x * ymm;
            out[(m + 1) * (m + 2) / 2 + m] = y1;
            double ylm2 = ymm, ylm1 = y1;
            for (int l = m + 2; l <= lmax; ++l) {
                double a = sqrt((4.0 * l * l - 1) / ((double)l * l - (double)m * m));
                double b = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1) * (l - 1) - 1));
                double yl = a * (x * ylm1 - b * ylm2);
                out[l * (l + 1) / 2 + m] = yl;
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

constexpr int KCHUNK = 64;

// F[b][k][t] for the bra bins rb0 + b, k = iphi*Nx + ix.  Block: (b, chunk of KCHUNK points).
__global__ __launch_bounds__(256) void k_angular(double *__restrict__ F, int rb0, int nrb, const double *__restrict__ phi,
                                                 const double *__restrict__ x, const double *__restrict__ Y1tab, int n1,
                                                 int lmax2, int n2, int lmax3, int n3, int K, int Nx, int Nphi,
                                                 int T, const int *__restrict__ term_ptr,
                                                 const double *__restrict__ coef, const int *__restrict__ i1,
                                                 const int *__restrict__ i2, const int *__restrict__ i3) {
    extern __shared__ double lds[];
    double *Y2 = lds;                    // KCHUNK * n2
    double *Y3 = lds + KCHUNK * n2;      // KCHUNK * n3
    const int b = blockIdx.y;
    const int rb = rb0 + b;
    const int k0 = blockIdx.x * KCHUNK;
    if (threadIdx.x < KCHUNK) {
        const int k = k0 + threadIdx.x;
        if (k < K) {
            const int iphi = k / Nx, ix = k % Nx;
            double sp, cp;
            sincos(phi[(int64_t)rb * Nphi + iphi], &sp, &cp);
            const double xx = x[ix];
            const double z1 = pi1t(sp, cp, xx), z2 = pi2t(sp, cp, xx);
            double ct1 = -(0.5 * sp + 0.75 * cp * xx) / z1;
            double ct2 = (sp - 0.5 * cp * xx) / z2;
            ct1 = fmin(1.0, fmax(-1.0, ct1));
            ct2 = fmin(1.0, fmax(-1.0, ct2));
            spharm_table(lmax2, ct1, Y2 + threadIdx.x * n2);
            spharm_table(lmax3, ct2, Y3 + threadIdx.x * n3);
        }
    }
    __syncthreads();
    const int kmax = min(KCHUNK, K - k0);
    // one tuple per wave at a time, its lanes on the points: every lane of a wave runs the
    // same term list (no divergence on the tuples' different term counts) and the term data
    // (coefficients and table indexes) are wave-uniform scalar loads
    for (int e = threadIdx.x; e < KCHUNK * T; e += blockDim.x) {
        const int t = __builtin_amdgcn_readfirstlane(e / KCHUNK), kk = e % KCHUNK;
        if (kk >= kmax) continue;
        const int ix = (k0 + kk) % Nx;
        const double *y1 = Y1tab + (int64_t)ix * n1;
        const double *y2 = Y2 + kk * n2;
        const double *y3 = Y3 + kk * n3;
        double acc = 0.0;
        for (int j = term_ptr[t]; j < term_ptr[t + 1]; ++j) acc += coef[j] * y1[i1[j]] * y2[i2[j]] * y3[i3[j]];
        F[((int64_t)b * K + k0 + kk) * T + t] = acc;
    }
    (void)nrb;
}

constexpr int CG = 8;     // cells per register group
constexpr int KC = 128;   // quadrature points per LDS stage

// H[cell][t] for all cells of bra bins rb0 .. rb0+nrb-1.  Block (b, tuple block of 256).
__global__ __launch_bounds__(256) void k_quad(double *__restrict__ H, const double *__restrict__ F, int rb0, int T,
                                              int K, int Nx, int Nphi, int Np, const int *__restrict__ rb_ptr,
                                              const int *__restrict__ rb_cells, const int *__restrict__ cell_ket,
                                              const double *__restrict__ pe, const double *__restrict__ qe,
                                              const double *__restrict__ x, const double *__restrict__ wx,
                                              const double *__restrict__ phi, const double *__restrict__ wphi) {
    __shared__ double Wl[CG][KC];
    const int b = blockIdx.y;
    const int rb = rb0 + b;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int qp = rb / Np, pp = rb % Np;
    const double ppl = pe[pp], ppu = pe[pp + 1], qpl = qe[qp], qpu = qe[qp + 1];
    const double *Fb = F + (int64_t)b * K * T;
    const int c_beg = rb_ptr[rb], c_end = rb_ptr[rb + 1];
    for (int cg0 = c_beg; cg0 < c_end; cg0 += CG) {
        const int ncg = min(CG, c_end - cg0);
        double acc[CG];
#pragma unroll
        for (int c = 0; c < CG; ++c) acc[c] = 0.0;
        for (int k0 = 0; k0 < K; k0 += KC) {
            __syncthreads();
            for (int e = threadIdx.x; e < CG * KC; e += blockDim.x) {
                const int c = e / KC, kk = e % KC, k = k0 + kk;
                double w = 0.0;
                if (c < ncg && k < K) {
                    const int ket = cell_ket[rb_cells[cg0 + c]];
                    const int q = ket / Np, p = ket % Np;
                    const int iphi = k / Nx, ix = k % Nx;
                    double sp, cp;
                    sincos(phi[(int64_t)rb * Nphi + iphi], &sp, &cp);
                    const double z1 = pi1t(sp, cp, x[ix]), z2 = pi2t(sp, cp, x[ix]);
                    const double kmin = fmax(fmax(ppl / sp, qpl / cp), fmax(pe[p] / z1, qe[q] / z2));
                    const double kmax = fmin(fmin(ppu / sp, qpu / cp), fmin(pe[p + 1] / z1, qe[q + 1] / z2));
                    if (!(kmin > kmax))
                        w = wx[ix] * wphi[(int64_t)rb * Nphi + iphi] * cp * sp * (kmax * kmax - kmin * kmin) /
                            (2 * z1 * z2);
                }
                Wl[c][kk] = w;
            }
            __syncthreads();
            if (t < T) {
                const int kn = min(KC, K - k0);
                for (int kk = 0; kk < kn; ++kk) {
                    const double f = Fb[(int64_t)(k0 + kk) * T + t];
#pragma unroll
                    for (int c = 0; c < CG; ++c) acc[c] += f * Wl[c][kk];
                }
            }
        }
        if (t < T) {
            for (int c = 0; c < ncg; ++c) {
                const int cell = rb_cells[cg0 + c];
                const int ket = cell_ket[cell];
                const int q = ket / Np, p = ket % Np;
                const double norm = sqrt(qe[q + 1] - qe[q]) * sqrt(pe[p + 1] - pe[p]) * sqrt(qpu - qpl) * sqrt(ppu - ppl);
                H[(int64_t)cell * T + t] = acc[c] / norm;
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

// Column c = (a, ket): rows (a', bra) for a' ascending, bras ascending (cells of a ket are contiguous).
__global__ void k_count(uint64_t *__restrict__ cnt, const double *__restrict__ V, const int *__restrict__ ket_ptr,
                        int Na, int NqNp) {
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= (int64_t)Na * NqNp) return;
    const int a = (int)(c / NqNp), ket = (int)(c % NqNp);
    uint64_t n = 0;
    for (int ap = 0; ap < Na; ++ap)
        for (int cell = ket_ptr[ket]; cell < ket_ptr[ket + 1]; ++cell)
            n += V[((int64_t)cell * Na + ap) * Na + a] != 0.0;
    cnt[c] = n;
}

__global__ void k_fill(double *__restrict__ val, int32_t *__restrict__ row, const uint64_t *__restrict__ colptr,
                       const double *__restrict__ V, const int *__restrict__ ket_ptr, const int *__restrict__ cell_bra,
                       int Na, int NqNp) {
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
    hipLaunchKernelGGL(k_mask, dim3(nb(ncell_all, 256)), dim3(256), 0, s, d_mask.p, d_pe.p, d_qe.p, d_x.p, d_phi.p, Np,
                       Nq, Nx, Nphi);
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
    DBuf<int> d_tptr(ang.term_ptr.size()), d_i1(std::max<size_t>(ang.term_i1.size(), 1)),
        d_i2(std::max<size_t>(ang.term_i2.size(), 1)), d_i3(std::max<size_t>(ang.term_i3.size(), 1));
    h2d(d_Y1.p, Y1, s);
    h2d(d_coef.p, ang.term_coef, s);
    h2d(d_tptr.p, ang.term_ptr, s);
    h2d(d_i1.p, ang.term_i1, s);
    h2d(d_i2.p, ang.term_i2, s);
    h2d(d_i3.p, ang.term_i3, s);
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
        const size_t lds = (size_t)KCHUNK * (n2 + n3) * sizeof(double);
        TT_REQUIRE(lds <= 160 * 1024, "angular tables exceed LDS");
        for (int rb0 = 0; rb0 < NqNp; rb0 += batch) {
            const int nrb = std::min(batch, NqNp - rb0);
            hipLaunchKernelGGL(k_angular, dim3(nb(K, KCHUNK), nrb), dim3(256), lds, s, d_F.p, rb0, nrb, d_phi.p, d_x.p,
                               d_Y1.p, n1, ang.lmax2, n2, ang.lmax3, n3, K, Nx, Nphi, T, d_tptr.p, d_coef.p, d_i1.p,
                               d_i2.p, d_i3.p);
            hipLaunchKernelGGL(k_quad, dim3(nb(T, 256), nrb), dim3(256), 0, s, d_H.p, d_F.p, rb0, T, K, Nx, Nphi, Np,
                               d_rbptr.p, d_rbcells.p, d_cket.p, d_pe.p, d_qe.p, d_x.p, d_wx.p, d_phi.p, d_wphi.p);
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
