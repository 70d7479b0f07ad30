// Host side of dU/dtheta (no reference counterpart; the reference has no derivatives): the
// first-order change of the SWP basis (C, bin edges, deuteron energy) and of the channel
// resolvent G under a change of the potential matrices.  The device side (column solve,
// products, pair dots) is in k_solve.hip.
#include <algorithm>
#include <cmath>

#include "tt_internal.h"

namespace tt {

// ------------------------------------------------------------------ SWP perturbation
// H = H0 + V (upper triangle, as tt_make_swp_states), C[basis][state] its eigenvectors in the
// stored order (ascending, coupled spectra de-interleaved), lambda_k = c_k^T H c_k.
//   d lambda_k = c_k^T dH c_k
//   d c_k      = sum_{l != k} c_l (c_l^T dH c_k) / (lambda_k - lambda_l)
// (non-degenerate first-order perturbation theory; the sign convention of c_k is kept).
// Edges are linear in the eigenvalues, so d e = swp_edges(d lambda).
void swp_derivative(const tt_run_params &rp, const tt_fwp_states &fwp, const tt_pw_states &pw,
                    const double *V_unco, const double *V_coup, const tt_swp_states &swp, const double *dV_unco,
                    const double *dV_coup, SwpDerivative &d) {
    int nu, nc;
    num_2N_states(rp, nu, nc);
    const int Np = fwp.Np_WP;
    d.dC_unco.assign((size_t)nu * Np * Np, 0.0);
    d.dC_coup.assign((size_t)nc * 4 * Np * Np, 0.0);
    d.de_unco.assign((size_t)nu * (Np + 1), 0.0);
    d.de_coup.assign((size_t)nc * 2 * (Np + 1), 0.0);
    d.dE_bound = 0.0;
    std::vector<double> H0(Np);
    for (int i = 0; i < Np; ++i) {
        const double p1 = fwp.p_WP[i], p2 = fwp.p_WP[i + 1];
        H0[i] = (p2 * p2 + p2 * p1 + p1 * p1) / (6 * kMu23);
    }
    std::vector<char> done_u(nu, 0), done_c(nc, 0);
    for (int ar = 0; ar < pw.Nalpha; ++ar) {
        const int Lr = pw.L_2N[ar], Sr = pw.S_2N[ar], Jr = pw.J_2N[ar], Tr = pw.T_2N[ar];
        for (int ac = 0; ac < pw.Nalpha; ++ac) {
            const int Lc = pw.L_2N[ac];
            if (!(Tr == pw.T_2N[ac] && Jr == pw.J_2N[ac] && Sr == pw.S_2N[ac] && std::abs(Lr - Lc) <= 2)) continue;
            bool via_L, via_T;
            coupling_kind(Lr, Sr, Jr, Lc, rp, via_L, via_T);
            const bool cm = via_L || via_T;
            const int idx = unique_2N_idx(Lr, Sr, Jr, Tr, cm, rp);
            const int n = cm ? 2 * Np : Np;
            const double *V, *dV, *C;
            double *dC, *de;
            if (cm) {
                if (done_c[idx]) continue;
                done_c[idx] = 1;
                V = V_coup + (size_t)idx * n * n;
                dV = dV_coup + (size_t)idx * n * n;
                C = swp.C_SWP_coup + (size_t)idx * n * n;
                dC = d.dC_coup.data() + (size_t)idx * n * n;
                de = d.de_coup.data() + (size_t)idx * 2 * (Np + 1);
            } else {
                if (done_u[idx]) continue;
                done_u[idx] = 1;
                V = V_unco + (size_t)idx * n * n;
                dV = dV_unco + (size_t)idx * n * n;
                C = swp.C_SWP_unco + (size_t)idx * n * n;
                dC = d.dC_unco.data() + (size_t)idx * n * n;
                de = d.de_unco.data() + (size_t)idx * (Np + 1);
            }
            std::vector<double> H((size_t)n * n), dH((size_t)n * n);
            for (int i = 0; i < n; ++i)
                for (int j = i; j < n; ++j) {
                    const double v = V[(size_t)i * n + j] + (i == j ? H0[i % Np] : 0.0), w = dV[(size_t)i * n + j];
                    H[(size_t)i * n + j] = H[(size_t)j * n + i] = v;
                    dH[(size_t)i * n + j] = dH[(size_t)j * n + i] = w;
                }
            // HC = H C, dHC = dH C, then lambda_k = (C^T H C)_kk and M = C^T dH C
            std::vector<double> HC((size_t)n * n, 0.0), dHC((size_t)n * n, 0.0);
            for (int i = 0; i < n; ++i)
                for (int m = 0; m < n; ++m) {
                    const double h = H[(size_t)i * n + m], g = dH[(size_t)i * n + m];
                    if (h == 0.0 && g == 0.0) continue;
                    const double *Cm = C + (size_t)m * n;
                    double *hc = &HC[(size_t)i * n], *gc = &dHC[(size_t)i * n];
                    for (int k = 0; k < n; ++k) {
                        hc[k] += h * Cm[k];
                        gc[k] += g * Cm[k];
                    }
                }
            std::vector<double> lam(n, 0.0), M((size_t)n * n, 0.0);
            for (int i = 0; i < n; ++i) {
                const double *Ci = C + (size_t)i * n;
                for (int k = 0; k < n; ++k) lam[k] += Ci[k] * HC[(size_t)i * n + k];
                for (int l = 0; l < n; ++l) {
                    const double cil = Ci[l];
                    if (cil == 0.0) continue;
                    const double *g = &dHC[(size_t)i * n];
                    double *Ml = &M[(size_t)l * n];
                    for (int k = 0; k < n; ++k) Ml[k] += cil * g[k];
                }
            }
            std::vector<double> dl(n), K((size_t)n * n, 0.0);
            for (int k = 0; k < n; ++k) {
                dl[k] = M[(size_t)k * n + k];
                for (int l = 0; l < n; ++l)
                    if (l != k) K[(size_t)l * n + k] = M[(size_t)l * n + k] / (lam[k] - lam[l]);
            }
            for (int j = 0; j < n; ++j)           // dC = C K
                for (int l = 0; l < n; ++l) {
                    const double c = C[(size_t)j * n + l];
                    if (c == 0.0) continue;
                    for (int k = 0; k < n; ++k) dC[(size_t)j * n + k] += c * K[(size_t)l * n + k];
                }
            const bool chn_3S1 = (Lr == 0 && Sr == 1 && Jr == 1 && Tr == 0);
            swp_edges(dl, Np, cm, chn_3S1, de);
            if (chn_3S1) d.dE_bound = dl[0];
        }
    }
}

// ------------------------------------------------------------------ resolvent derivative
namespace {
struct Dual {
    double v, d;
};
inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
inline Dual operator*(double c, Dual a) { return {c * a.v, c * a.d}; }
inline Dual logabs(Dual a) { return {std::log(std::fabs(a.v)), a.d / a.v}; }
inline double heav(Dual a) { return a.v < 0 ? 0.0 : 1.0; }
// x ln|x| (its derivative ln|x| + 1 is finite except at 0, where the value is not used)
inline Dual xlogx(Dual a) { return a * logabs(a); }
// x H(x), derivative H(x)
inline Dual xheav(Dual a) { return {a.v * heav(a), a.d * heav(a)}; }
inline Dual mu1_dual(Dual Eb) {
    // mu1 = Mn (Mn + Mp + Eb) / (2 Mn + Mp + Eb)
    const Dual Md = Dual{kMn + kMp, 0.0} + Eb;
    return Dual{kMn, 0.0} * Md / (Dual{kMn, 0.0} + Md);
}
}  // namespace

// The closed forms of k_resolvent (CPP/make_resolvent.cpp:14-80) in forward-mode duals.
void resolvent_derivative(const tt_run_params &rp, const tt_swp_states &swp, const tt_pw_states &chn,
                          const SwpDerivative &d, const double *E, const double *dE, int nE, double *dG) {
    std::vector<int> which, offset;
    swp_edge_offsets(chn, swp, rp, which, offset);
    const int Na = chn.Nalpha, Nq = swp.Nq_WP, Np = swp.Np_WP;
    const int64_t dim = (int64_t)Na * Nq * Np;
    for (int j = 0; j < nE; ++j) {
        const Dual Ej{E[j], dE[j]};
        for (int a = 0; a < Na; ++a) {
            const double *e = (which[a] ? swp.e_SWP_coup : swp.e_SWP_unco) + offset[a];
            const double *de = (which[a] ? d.de_coup.data() : d.de_unco.data()) + offset[a];
            for (int q = 0; q < Nq; ++q) {
                const double ql = swp.q_WP[q], qu = swp.q_WP[q + 1];
                for (int p = 0; p < Np; ++p) {
                    const Dual el{e[p], de[p]}, eu{e[p + 1], de[p + 1]};
                    Dual re, im;
                    if (el.v < 0) {
                        const Dual Eb = el;
                        const Dual m1 = mu1_dual(Eb);
                        const Dual Eql = Dual{ql * ql, 0.0} / (2.0 * m1), Equ = Dual{qu * qu, 0.0} / (2.0 * m1);
                        const Dual Dq = Equ - Eql;
                        re = logabs((Eql + Eb - Ej) / (Equ + Eb - Ej)) / Dq;
                        const Dual hh{(heav(Equ + Eb - Ej) - heav(Eql + Eb - Ej)) * (-M_PI), 0.0};
                        im = hh / Dq;
                    } else {
                        const Dual m1 = mu1_dual(Dual{0.0, 0.0});
                        const Dual Eql = Dual{ql * ql, 0.0} / (2.0 * m1), Equ = Dual{qu * qu, 0.0} / (2.0 * m1);
                        const Dual Ep = 0.5 * (el + eu), Eq = 0.5 * (Eql + Equ);
                        const Dual Dq = Equ - Eql, Dp = eu - el;
                        const Dual DM = 0.5 * (Dp - Dq), DP = 0.5 * (Dp + Dq);
                        const Dual D = Ep + Eq - Ej;
                        re = (xlogx(D + DM) + xlogx(D - DM) - xlogx(D + DP) - xlogx(D - DP)) / (Dp * Dq);
                        im = M_PI * (xheav(D + DM) + xheav(D - DM) - xheav(D + DP) - xheav(D - DP)) / (Dp * Dq);
                    }
                    const int64_t idx = ((int64_t)a * Nq + q) * Np + p;
                    dG[2 * ((size_t)j * dim + idx)] = re.d;
                    dG[2 * ((size_t)j * dim + idx) + 1] = im.d;
                }
            }
        }
    }
}

}  // namespace tt
