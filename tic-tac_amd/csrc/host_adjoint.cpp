// Host side of dU/dtheta (no reference counterpart; the reference has no derivatives): the
// first-order change of the SWP basis (C, bin edges, deuteron energy) under a change of the
// potential matrices.  The device side (resolvent partials, column solve, products, pair
// dots) is in k_solve.hip.
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
            bool any = false;
            for (size_t e = 0; e < (size_t)n * n && !any; ++e) any = dV[e] != 0.0;
            if (!any) continue;   // untouched channel: dC, d lambda and d e stay zero
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

}  // namespace tt
