// CPU restatement (C++/OpenMP, FP64) of the on-shell row solves -- TEST INFRASTRUCTURE ONLY.
//
// Used by the cpu_baseline leg of bench.py as the timed CPU baseline, and by tests/ as a checker;
// the product (tic-tac_amd/) never links or runs it.  Two variants of the reference's path:
//
//  ttcpu_neumann_terms   the reference's semantics (CPP/solve_faddeev.cpp:440-1444): every
//                        Neumann term regenerates the columns of A = C^T (2 P123) VC one by one
//                        (calculate_PVC_col, lines 4-48, and calculate_CPVC_col, lines 50-181)
//                        and multiplies the on-shell rows (row .* G_j) into them (the cblas_dgemm
//                        of lines 1072-1088, real and imaginary parts), giving
//                        a_n = e_r^T A (G A)^n e_c at the on-shell columns.
//  ttcpu_gmres_rows      a matrix-free restatement: GMRES (restarted, classical Gram-Schmidt
//                        with one reorthogonalisation) on the row systems y (1 - G_k A) = e_k^T A
//                        of every on-shell row k, A applied as (((y .* G) C^T_op^T) 2 P123) VC_op
//                        over the reference's CSC arrays (CSC of P123 = CSR of P123^T: a gather,
//                        no atomics), all rows sharing each application.
//
// Operator input: the CSC of P123 (CPP/main.cpp:226-244) and the embedded 2N blocks
// C_op[a][p][a'][p'] and VC_op[a][p][a'][p'] (Nalpha x Np x Nalpha x Np, as oracle.solve.c_and_vc
// gives them; the reference's pointer tables of CPP/solve_faddeev.cpp:1588-1927 hold the same
// blocks).  State index idx = alpha*Nq*Np + q*Np + p.  Complex arrays are interleaved (re, im).
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>

typedef std::complex<double> cd;

namespace {

// Nonzero partner blocks of an embedded operator: for output alpha a, the partners e with
// their Np x Np block M[a][.][e][.] (row-major over (p, p')).
struct Partners {
    int Na = 0, Np = 0;
    std::vector<int> ptr, alpha;
    std::vector<double> blk;
    // transpose: blocks of M^T (output alpha a takes M[e][.][a][.]^T)
    Partners(const double *M, int Na_, int Np_, bool transpose) : Na(Na_), Np(Np_) {
        ptr.assign(Na + 1, 0);
        for (int a = 0; a < Na; ++a) {
            for (int e = 0; e < Na; ++e) {
                bool nz = false;
                for (int p = 0; p < Np && !nz; ++p)
                    for (int q = 0; q < Np && !nz; ++q) {
                        const double v = transpose ? M[(((size_t)e * Np + q) * Na + a) * Np + p]
                                                   : M[(((size_t)a * Np + p) * Na + e) * Np + q];
                        nz = v != 0.0;
                    }
                if (!nz) continue;
                alpha.push_back(e);
                for (int p = 0; p < Np; ++p)
                    for (int q = 0; q < Np; ++q)
                        blk.push_back(transpose ? M[(((size_t)e * Np + q) * Na + a) * Np + p]
                                                : M[(((size_t)a * Np + p) * Na + e) * Np + q]);
            }
            ptr[a + 1] = (int)alpha.size();
        }
    }
};

struct Operator {
    int64_t dim;
    int Na, Nq, Np;
    const double *val;
    const int32_t *row;
    const uint64_t *colptr;
    Partners C, VCt;   // C_op, and VC_op^T

    Operator(int64_t dim_, int Na_, int Nq_, int Np_, const double *v, const int32_t *r, const uint64_t *cp,
             const double *Cop, const double *VCop)
        : dim(dim_), Na(Na_), Nq(Nq_), Np(Np_), val(v), row(r), colptr(cp), C(Cop, Na_, Np_, false),
          VCt(VCop, Na_, Np_, true) {}

    // Y[(a,q,p)][k] = sum_e sum_p' B_e[p][p'] X[(e,q,p')][k] for r columns (RHS-fastest layout)
    void block(const Partners &B, const cd *X, cd *Y, int r) const {
#pragma omp parallel for collapse(2) schedule(static)
        for (int a = 0; a < Na; ++a)
            for (int q = 0; q < Nq; ++q) {
                cd *y = Y + (((size_t)a * Nq + q) * Np) * r;
                std::fill(y, y + (size_t)Np * r, cd(0.0));
                for (int t = B.ptr[a]; t < B.ptr[a + 1]; ++t) {
                    const int e = B.alpha[t];
                    const double *M = &B.blk[(size_t)t * Np * Np];
                    const cd *x = X + (((size_t)e * Nq + q) * Np) * r;
                    for (int p = 0; p < Np; ++p)
                        for (int pp = 0; pp < Np; ++pp) {
                            const double m = M[(size_t)p * Np + pp];
                            if (m == 0.0) continue;
                            for (int k = 0; k < r; ++k) y[(size_t)p * r + k] += m * x[(size_t)pp * r + k];
                        }
                }
            }
    }

    // U[c][k] = 2 sum_{nz of CSC column c} val * S[row][k]  (U = 2 P123^T S)
    void gather(const cd *S, cd *U, int r) const {
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t c = 0; c < dim; ++c) {
            cd *u = U + (size_t)c * r;
            for (int k = 0; k < r; ++k) u[k] = 0.0;
            for (uint64_t z = colptr[c]; z < colptr[c + 1]; ++z) {
                const double v = 2.0 * val[z];
                const cd *s = S + (size_t)row[z] * r;
                for (int k = 0; k < r; ++k) u[k] += v * s[k];
            }
        }
    }

    // Y = (X .* G_sel) A in the row form; Gk[k] points at the resolvent row of column k (or null)
    void apply(const cd *X, cd *Y, int r, const std::vector<const cd *> &Gk, std::vector<cd> &t1,
               std::vector<cd> &t2) const {
        t1.resize((size_t)dim * r);
        t2.resize((size_t)dim * r);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < dim; ++i)
            for (int k = 0; k < r; ++k) t2[(size_t)i * r + k] = Gk[k] ? X[(size_t)i * r + k] * Gk[k][i] : X[(size_t)i * r + k];
        block(C, t2.data(), t1.data(), r);   // S = C_op (X .* G)
        gather(t1.data(), t2.data(), r);     // U = 2 P123^T S
        block(VCt, t2.data(), Y, r);         // Y = VC_op^T U
    }
};

double seconds_since(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

int ttcpu_threads(void) { return omp_get_max_threads(); }

// GMRES(restart) on the r row systems y_k (1 - G_{g[k]} A) = e_{unit[k]}^T A (FP64), all sharing
// each application of A.  Y: r x dim complex solutions (row-major), its[k] iterations, res[k] final
// relative residual estimate.  Returns 0 when every system reached tol.
int ttcpu_gmres_rows(int64_t dim, int Na, int Nq, int Np, const double *val, const int32_t *row,
                     const uint64_t *colptr, const double *Cop, const double *VCop, const double *G, int nE,
                     int r, const int64_t *unit, const int *g, double tol, int restart, int max_iter, double *Y,
                     int *its, double *res) {
    Operator op(dim, Na, Nq, Np, val, row, colptr, Cop, VCop);
    const cd *Gc = reinterpret_cast<const cd *>(G);
    std::vector<const cd *> none(r, nullptr), Gk(r);
    for (int k = 0; k < r; ++k) Gk[k] = (g[k] >= 0 && g[k] < nE) ? Gc + (size_t)g[k] * dim : nullptr;
    const size_t n = (size_t)dim * r;
    std::vector<cd> t1, t2, b(n, 0.0), x(n, 0.0), w(n), e(n, 0.0);
    for (int k = 0; k < r; ++k) e[(size_t)unit[k] * r + k] = 1.0;
    op.apply(e.data(), b.data(), r, none, t1, t2);   // b_k = e_k^T A
    auto norms = [&](const std::vector<cd> &v, std::vector<double> &out) {
        out.assign(r, 0.0);
        for (size_t i = 0; i < (size_t)dim; ++i)
            for (int k = 0; k < r; ++k) out[k] += std::norm(v[i * r + k]);
        for (auto &o : out) o = std::sqrt(o);
    };
    std::vector<double> bn, rn;
    norms(b, bn);
    const int m = restart;
    std::vector<cd> V((size_t)(m + 1) * n), H((size_t)(m + 1) * m * r), cs((size_t)m * r), sn((size_t)m * r),
        gg((size_t)(m + 1) * r);
    for (int k = 0; k < r; ++k) its[k] = 0;
    int total = 0;
    while (true) {
        // residual R = b - (x - (x .* G) A)
        op.apply(x.data(), w.data(), r, Gk, t1, t2);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < (int64_t)n; ++i) w[i] = b[i] - (x[i] - w[i]);
        norms(w, rn);
        bool all = true;
        for (int k = 0; k < r; ++k) {
            res[k] = rn[k] / std::max(bn[k], 1e-300);
            if (!(res[k] <= tol)) all = false;
        }
        if (all || total >= max_iter) {
            for (int k = 0; k < r; ++k)
                for (int64_t i = 0; i < dim; ++i) {
                    Y[2 * ((size_t)k * dim + i)] = x[(size_t)i * r + k].real();
                    Y[2 * ((size_t)k * dim + i) + 1] = x[(size_t)i * r + k].imag();
                }
            return all ? 0 : 1;
        }
        for (size_t i = 0; i < n; ++i) V[i] = rn[i % r] > 0 ? w[i] / rn[i % r] : 0.0;
        std::fill(gg.begin(), gg.end(), cd(0.0));
        for (int k = 0; k < r; ++k) gg[k] = rn[k];
        int j = 0;
        for (; j < m && total < max_iter; ++j, ++total) {
            cd *vj = &V[(size_t)j * n], *vn = &V[(size_t)(j + 1) * n];
            op.apply(vj, w.data(), r, Gk, t1, t2);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < (int64_t)n; ++i) w[i] = vj[i] - w[i];   // w = v - (v .* G) A
            std::vector<cd> h((size_t)(j + 1) * r, 0.0);
            for (int pass = 0; pass < 2; ++pass) {   // classical Gram-Schmidt, reorthogonalised
                std::vector<cd> hp((size_t)(j + 1) * r, 0.0);
#pragma omp parallel
                {
                    std::vector<cd> loc((size_t)(j + 1) * r, 0.0);
#pragma omp for schedule(static)
                    for (int64_t i = 0; i < dim; ++i)
                        for (int l = 0; l <= j; ++l)
                            for (int k = 0; k < r; ++k)
                                loc[(size_t)l * r + k] += std::conj(V[(size_t)l * n + i * r + k]) * w[i * r + k];
#pragma omp critical
                    for (size_t q = 0; q < loc.size(); ++q) hp[q] += loc[q];
                }
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < dim; ++i)
                    for (int l = 0; l <= j; ++l)
                        for (int k = 0; k < r; ++k) w[i * r + k] -= hp[(size_t)l * r + k] * V[(size_t)l * n + i * r + k];
                for (size_t q = 0; q < h.size(); ++q) h[q] += hp[q];
            }
            std::vector<double> wn;
            norms(w, wn);
            for (size_t i = 0; i < n; ++i) vn[i] = wn[i % r] > 0 ? w[i] / wn[i % r] : 0.0;
            bool conv = true;
            for (int k = 0; k < r; ++k) {
                auto Hk = [&](int a, int c) -> cd & { return H[((size_t)a * m + c) * r + k]; };
                for (int l = 0; l <= j; ++l) Hk(l, j) = h[(size_t)l * r + k];
                Hk(j + 1, j) = wn[k];
                for (int l = 0; l < j; ++l) {
                    const cd c = cs[(size_t)l * r + k], s = sn[(size_t)l * r + k];
                    const cd t0 = std::conj(c) * Hk(l, j) + std::conj(s) * Hk(l + 1, j);
                    const cd t1v = -s * Hk(l, j) + c * Hk(l + 1, j);
                    Hk(l, j) = t0;
                    Hk(l + 1, j) = t1v;
                }
                const cd a = Hk(j, j), bb = Hk(j + 1, j);
                const double den = std::sqrt(std::norm(a) + std::norm(bb));
                const cd c = den > 0 ? a / den : 1.0, s = den > 0 ? bb / den : 0.0;
                cs[(size_t)j * r + k] = c;
                sn[(size_t)j * r + k] = s;
                Hk(j, j) = den;
                Hk(j + 1, j) = 0.0;
                const cd gj = gg[(size_t)j * r + k];
                gg[(size_t)j * r + k] = std::conj(c) * gj;
                gg[(size_t)(j + 1) * r + k] = -s * gj;
                if (!(std::abs(gg[(size_t)(j + 1) * r + k]) / std::max(bn[k], 1e-300) <= 0.5 * tol)) conv = false;
                its[k] += 1;
            }
            if (conv) {
                ++j;
                ++total;
                break;
            }
        }
        // x += V z, z from the triangular systems
        for (int k = 0; k < r; ++k) {
            std::vector<cd> z(j, 0.0);
            for (int a = j - 1; a >= 0; --a) {
                cd s = gg[(size_t)a * r + k];
                for (int c = a + 1; c < j; ++c) s -= H[((size_t)a * m + c) * r + k] * z[c];
                const cd d = H[((size_t)a * m + a) * r + k];
                z[a] = d != cd(0.0) ? s / d : 0.0;
            }
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < dim; ++i)
                for (int a = 0; a < j; ++a) x[i * r + k] += z[a] * V[(size_t)a * n + i * r + k];
        }
    }
}

// The reference's Neumann rows: a[(k, c_idx)][t] = (rows_t)_k at column cols[c_idx], with
// rows_0 = e_k^T A and rows_{t+1} = (rows_t .* G_{g[k]}) A, every A column regenerated for every
// term as CPP/solve_faddeev.cpp:1019-1091 does (calculate_PVC_col + calculate_CPVC_col, then the
// row products).  Stops after n_terms terms or when the wall time exceeds cap_s; returns the
// number of terms done, their mean time in *sec_per_term.  a: r * ncols * n_terms complex.
int ttcpu_neumann_terms(int64_t dim, int Na, int Nq, int Np, const double *val, const int32_t *row,
                        const uint64_t *colptr, const double *Cop, const double *VCop, const double *G, int nE,
                        int r, const int64_t *unit, const int *g, int ncols, const int64_t *cols, int n_terms,
                        double cap_s, double *a, double *sec_per_term) {
    const Partners CT(Cop, Na, Np, true);    // C^T_op: output alpha_r takes C_op[e][.][r][.]^T
    const Partners VC(VCop, Na, Np, true);   // column c = (alpha_c, q, p_c) takes VC_op[.][.][alpha_c][p_c]
    const cd *Gc = reinterpret_cast<const cd *>(G);
    cd *A = reinterpret_cast<cd *>(a);
    // rows as separate real and imaginary parts (CPP/solve_faddeev.cpp:503-510)
    std::vector<double> re((size_t)r * dim, 0.0), im((size_t)r * dim, 0.0), nre(re.size()), nim(im.size());
    for (int k = 0; k < r; ++k) re[(size_t)k * dim + unit[k]] = 1.0;   // e_k; the first term is e_k^T A
    auto t0 = std::chrono::steady_clock::now();
    int done = 0;
    const int64_t NqNp = (int64_t)Nq * Np;
    for (int t = 0; t < n_terms; ++t) {
        // rows .* G (not for the first product, whose rows are the unit rows)
        if (t > 0)
            for (int k = 0; k < r; ++k) {
                if (g[k] < 0 || g[k] >= nE) continue;
                const cd *gk = Gc + (size_t)g[k] * dim;
                for (int64_t i = 0; i < dim; ++i) {
                    const cd v = cd(re[(size_t)k * dim + i], im[(size_t)k * dim + i]) * gk[i];
                    re[(size_t)k * dim + i] = v.real();
                    im[(size_t)k * dim + i] = v.imag();
                }
            }
#pragma omp parallel
        {
            std::vector<double> pvc(dim), col(dim);
#pragma omp for schedule(dynamic, 16)
            for (int64_t c = 0; c < dim; ++c) {
                const int ac = (int)(c / NqNp), qc = (int)((c / Np) % Nq), pc = (int)(c % Np);
                // PVC column (calculate_PVC_col): sum over VC partners alpha_j and p_j of
                // 2 P123[:, (alpha_j, q_c, p_j)] VC_{alpha_j alpha_c}[p_j][p_c]
                std::fill(pvc.begin(), pvc.end(), 0.0);
                for (int tq = VC.ptr[ac]; tq < VC.ptr[ac + 1]; ++tq) {
                    const int aj = VC.alpha[tq];
                    const double *M = &VC.blk[(size_t)tq * Np * Np];   // M[pc][pj] = VC_op[aj][pj][ac][pc]
                    for (int pj = 0; pj < Np; ++pj) {
                        const double vcv = M[(size_t)pc * Np + pj];
                        if (vcv == 0.0) continue;
                        const int64_t cj = (int64_t)aj * NqNp + (int64_t)qc * Np + pj;
                        for (uint64_t z = colptr[cj]; z < colptr[cj + 1]; ++z) pvc[row[z]] += 2.0 * val[z] * vcv;
                    }
                }
                // A column (calculate_CPVC_col): C^T blocks per (alpha_r, alpha_i), q-diagonal
                std::fill(col.begin(), col.end(), 0.0);
                for (int ar = 0; ar < Na; ++ar)
                    for (int tq = CT.ptr[ar]; tq < CT.ptr[ar + 1]; ++tq) {
                        const int ai = CT.alpha[tq];
                        const double *M = &CT.blk[(size_t)tq * Np * Np];   // M[pr][pi] = C_op[ai][pi][ar][pr]
                        for (int q = 0; q < Nq; ++q) {
                            const double *src = &pvc[(size_t)ai * NqNp + (size_t)q * Np];
                            double *dst = &col[(size_t)ar * NqNp + (size_t)q * Np];
                            for (int pi = 0; pi < Np; ++pi) {
                                const double s = src[pi];
                                if (s == 0.0) continue;
                                for (int pr = 0; pr < Np; ++pr) dst[pr] += M[(size_t)pr * Np + pi] * s;
                            }
                        }
                    }
                // the row products (cblas_dgemm, real and imaginary parts)
                for (int k = 0; k < r; ++k) {
                    double sr = 0.0, si = 0.0;
                    const double *xr = &re[(size_t)k * dim], *xi = &im[(size_t)k * dim];
                    for (int64_t i = 0; i < dim; ++i) {
                        sr += xr[i] * col[i];
                        si += xi[i] * col[i];
                    }
                    nre[(size_t)k * dim + c] = sr;
                    nim[(size_t)k * dim + c] = si;
                }
            }
        }
        std::swap(re, nre);
        std::swap(im, nim);
        for (int k = 0; k < r; ++k)
            for (int ci = 0; ci < ncols; ++ci)
                A[((size_t)k * ncols + ci) * n_terms + t] = cd(re[(size_t)k * dim + cols[ci]], im[(size_t)k * dim + cols[ci]]);
        ++done;
        if (seconds_since(t0) > cap_s) break;
    }
    *sec_per_term = done ? seconds_since(t0) / done : 0.0;
    return done;
}

}  // extern "C"
