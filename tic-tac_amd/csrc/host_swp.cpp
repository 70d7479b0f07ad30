// Scattering wave packets: diagonalise H = H0 + V_WP per 2N channel,
// CPP/make_swp_states.cpp:21-183,221-471.  The reference calls LAPACKE_dspevd;
// here a cyclic Jacobi solver (matrices are at most 2Np x 2Np).
#include <algorithm>
#include <numeric>

#include "tt_internal.h"

namespace tt {

// Symmetric eigen-decomposition by cyclic Jacobi rotations.  On return ev is
// ascending and Z[i*n + k] is component i of eigenvector k (column-wise), the
// layout dspevd returns in row-major mode.
static void jacobi_eigh(int n, std::vector<double> A, std::vector<double> &ev, std::vector<double> &Z) {
    std::vector<double> V((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i) V[(size_t)i * n + i] = 1.0;
    auto a = [&](int i, int j) -> double & { return A[(size_t)i * n + j]; };
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += a(i, j) * a(i, j);
                if (i != j) off += a(i, j) * a(i, j);
            }
        if (off <= 1e-30 * tot) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = a(p, q);
                if (apq == 0.0) continue;
                double app = a(p, p), aqq = a(q, q);
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = a(k, p), akq = a(k, q);
                    a(k, p) = c * akp - s * akq;
                    a(k, q) = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = a(p, k), aqk = a(q, k);
                    a(p, k) = c * apk - s * aqk;
                    a(q, k) = s * apk + c * aqk;
                }
                for (int k = 0; k < n; ++k) {
                    double vkp = V[(size_t)k * n + p], vkq = V[(size_t)k * n + q];
                    V[(size_t)k * n + p] = c * vkp - s * vkq;
                    V[(size_t)k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::sort(order.begin(), order.end(), [&](int i, int j) { return a(i, i) < a(j, j); });
    ev.resize(n);
    Z.assign((size_t)n * n, 0.0);
    for (int k = 0; k < n; ++k) {
        ev[k] = a(order[k], order[k]);
        for (int i = 0; i < n; ++i) Z[(size_t)i * n + k] = V[(size_t)i * n + order[k]];
    }
}

// CPP/make_swp_states.cpp:146-183.  Linear in ev (all constants are zero), so applied to
// eigenvalue derivatives it gives the edge derivatives.
void swp_edges(const std::vector<double> &ev, int Np, bool coupled, bool chn_3S1, double *e) {
    e[0] = 0;
    if (coupled) e[Np + 1] = 0;
    for (int i = 0; i < Np - 1; ++i) {
        e[i + 1] = 0.5 * (ev[i + 1] + ev[i]);
        if (coupled) e[i + 1 + Np + 1] = 0.5 * (ev[i + 1 + Np] + ev[i + Np]);
    }
    e[Np] = ev[Np - 1] + 0.5 * (e[Np - 1] - e[Np - 2]);
    if (coupled) e[2 * Np + 1] = ev[2 * Np - 1] + 0.5 * (e[2 * Np] - e[2 * Np - 1]);
    if (chn_3S1) {
        e[0] = ev[0];
        e[1] = 0;
    }
}

}  // namespace tt

using namespace tt;

extern "C" tt_status tt_make_swp_states(const tt_run_params *rp, const tt_fwp_states *fwp, const tt_pw_states *pw,
                                        const double *V_unco, const double *V_coup, tt_swp_states *swp) {
    TT_GUARD_BEGIN
    TT_REQUIRE(rp && fwp && pw && V_unco && swp, "null argument");
    int nu, nc;
    num_2N_states(*rp, nu, nc);
    const int Np = fwp->Np_WP;
    TT_REQUIRE(swp->e_SWP_unco && swp->C_SWP_unco && (nc == 0 || (swp->e_SWP_coup && swp->C_SWP_coup && V_coup)),
               "swp arrays must be allocated by the caller");
    swp->Np_WP = fwp->Np_WP;
    swp->Nq_WP = fwp->Nq_WP;
    swp->num_2N_unco_states = nu;
    swp->num_2N_coup_states = nc;
    swp->q_WP = fwp->q_WP;
    std::fill(swp->e_SWP_unco, swp->e_SWP_unco + (size_t)nu * (Np + 1), 0.0);
    if (nc) std::fill(swp->e_SWP_coup, swp->e_SWP_coup + (size_t)nc * 2 * (Np + 1), 0.0);
    std::vector<double> H0(Np);
    for (int i = 0; i < Np; ++i) {
        double p1 = fwp->p_WP[i], p2 = fwp->p_WP[i + 1];
        H0[i] = (p2 * p2 + p2 * p1 + p1 * p1) / (6 * kMu23);
    }
    std::vector<char> done_u(nu, 0), done_c(nc, 0);
    double E_bound = 0.0;
    for (int ar = 0; ar < pw->Nalpha; ++ar) {
        int Lr = pw->L_2N[ar], Sr = pw->S_2N[ar], Jr = pw->J_2N[ar], Tr = pw->T_2N[ar];
        for (int ac = 0; ac < pw->Nalpha; ++ac) {
            int Lc = pw->L_2N[ac];
            if (!(Tr == pw->T_2N[ac] && Jr == pw->J_2N[ac] && Sr == pw->S_2N[ac] && std::abs(Lr - Lc) <= 2)) continue;
            bool via_L, via_T;
            coupling_kind(Lr, Sr, Jr, Lc, *rp, via_L, via_T);
            bool cm = via_L || via_T;
            int idx = unique_2N_idx(Lr, Sr, Jr, Tr, cm, *rp);
            int n = cm ? 2 * Np : Np;
            const double *V;
            double *C, *e;
            if (cm) {
                TT_REQUIRE(idx >= 0 && idx < nc, "coupled 2N index out of range");
                if (done_c[idx]) continue;
                done_c[idx] = 1;
                V = V_coup + (size_t)idx * n * n;
                C = swp->C_SWP_coup + (size_t)idx * n * n;
                e = swp->e_SWP_coup + (size_t)idx * 2 * (Np + 1);
            } else {
                TT_REQUIRE(idx >= 0 && idx < nu, "uncoupled 2N index out of range");
                if (done_u[idx]) continue;
                done_u[idx] = 1;
                V = V_unco + (size_t)idx * n * n;
                C = swp->C_SWP_unco + (size_t)idx * n * n;
                e = swp->e_SWP_unco + (size_t)idx * (Np + 1);
            }
            // H from the upper triangle of V (construct_full_hamiltonian, lines 45-81)
            std::vector<double> H((size_t)n * n);
            for (int i = 0; i < n; ++i)
                for (int j = i; j < n; ++j) {
                    double v = V[(size_t)i * n + j] + (i == j ? H0[i % Np] : 0.0);
                    H[(size_t)i * n + j] = v;
                    H[(size_t)j * n + i] = v;
                }
            std::vector<double> ev, Z;
            jacobi_eigh(n, H, ev, Z);
            bool chn_3S1 = (Lr == 0 && Sr == 1 && Jr == 1 && Tr == 0);
            int nb = 0;
            for (double x : ev)
                if (x < 0) {
                    ++nb;
                    E_bound = x;
                }
            if ((nb == 0 && !chn_3S1) || (nb == 1 && chn_3S1)) {
            } else if (nb != 0) {
                fail(TT_ERR_PHYSICS, "Found unphysical bound states in NN-pair Hamiltonian eigenspectrum!");
            } else {
                fail(TT_ERR_PHYSICS, "Didn't find any bound states in NN-pair Hamiltonian eigenspectrum!");
            }
            if (cm) {
                // reorder_coupled_eigenspectrum, lines 85-112
                std::vector<double> ev2(n), Z2((size_t)n * n);
                for (int i = 0; i < Np; ++i) {
                    ev2[i] = ev[2 * i];
                    ev2[i + Np] = ev[2 * i + 1];
                    for (int j = 0; j < n; ++j) {
                        Z2[(size_t)j * n + i] = Z[(size_t)j * n + 2 * i];
                        Z2[(size_t)j * n + i + Np] = Z[(size_t)j * n + 2 * i + 1];
                    }
                }
                ev.swap(ev2);
                Z.swap(Z2);
            }
            std::copy(Z.begin(), Z.end(), C);
            swp_edges(ev, Np, cm, chn_3S1, e);
        }
    }
    swp->E_bound = E_bound;
    TT_GUARD_END
}
