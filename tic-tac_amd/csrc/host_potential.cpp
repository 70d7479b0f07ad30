// Interaction plugin and V in the WP basis:
//   plugin API            CPP/Interactions/potential_model.h:11-22
//   Malfliet-Tjon I/III   CPP/Interactions/malfliet_tjon.cpp:27-66
//   synthetic potential   no reference counterpart: seeded Yukawas + contacts + separables,
//                         regulated by exp(-(p/Lambda)^6), linear in its 12 strengths
//   V_WP blocks           CPP/make_potential_matrix.cpp:4-430
#include <cstring>
#include <memory>

#include "tt_internal.h"

namespace tt {

// Q_L(z) for z > 1: Q_0 = 1/2 ln((z+1)/(z-1)) times the ratios Q_l/Q_{l-1} from the
// backward continued fraction r_l = l / ((2l+1) z - (l+1) r_{l+1}).
static double legendre_Q_from_q0(int L, double z, double q0) {
    if (L == 0) return q0;
    double rho = z + std::sqrt(z * z - 1.0);
    double lr = std::log(rho);
    int nstart = L + 20 + (int)(40.0 / std::max(lr, 1e-3));
    nstart = std::min(nstart, L + 20000);
    double r = 0.0, ratios[32];
    for (int l = nstart; l >= 1; --l) {
        r = l / ((2 * l + 1) * z - (l + 1) * r);
        if (l <= L) ratios[l] = r;
    }
    double q = q0;
    for (int l = 1; l <= L; ++l) q *= ratios[l];
    return q;
}

// Partial-wave Yukawa e^{-mu r}/r per unit (V0/hbarc): Q_L(z)/(pi p p').
static double yukawa_pw(int L, double p_out, double p_in, double mu) {
    double q0 = 0.5 * std::log(((p_out + p_in) * (p_out + p_in) + mu * mu) /
                               ((p_out - p_in) * (p_out - p_in) + mu * mu));
    double z = (p_out * p_out + p_in * p_in + mu * mu) / (2.0 * p_out * p_in);
    return legendre_Q_from_q0(L, z, q0) / (M_PI * p_out * p_in);
}

// V() is called from several host threads at once (the V_WP rows): it must not write state.
struct Potential {
    virtual ~Potential() = default;
    virtual void update(const double *p, int n) = 0;
    virtual void V(double p_in, double p_out, bool coupled, int S, int J, int T, int Tz, double *v) const = 0;
};

struct MalflietTjon final : Potential {
    void update(const double *, int) override {}
    void V(double qi, double qo, bool, int, int J, int, int, double *v) const override {
        for (int i = 0; i < 6; ++i) v[i] = 0.0;
        const double VA_I = 513.968, VA_III = 626.885, VR = 1438.720;
        const double muA = 1.550 * kHbarc, muR = 3.110 * kHbarc;
        if (J == 0 || J == 1) {
            double num_A = (qo + qi) * (qo + qi) + muA * muA;
            double num_R = (qo + qi) * (qo + qi) + muR * muR;
            double den_A = (qo - qi) * (qo - qi) + muA * muA;
            double den_R = (qo - qi) * (qo - qi) + muR * muR;
            double VA = J == 0 ? VA_I : VA_III;
            double val = ((VR / kHbarc) * std::log(num_R / den_R) - (VA / kHbarc) * std::log(num_A / den_A)) /
                         (2 * M_PI * qi * qo);
            v[J == 0 ? 0 : 2] = val;
        }
    }
};

// theta = [Y_0.7, Y_2.0, Y_4.0 (MeV fm), C_1S0, C_3S1 (fm^2), C_3P0, C_1P1, C_3P1, C_SD (fm^4),
//          lam_1S0, lam_3S1, lam_1P1 (fm^2)], beta = 3 separable ranges (fm^-1).
struct Synthetic final : Potential {
    double th[12] = {0};
    double beta[3] = {1.0, 1.0, 1.0};
    double mu[3] = {0.7 * kHbarc, 2.0 * kHbarc, 4.0 * kHbarc};
    double Lam = 2.5 * kHbarc;
    double cd_nn = 0.02;
    void update(const double *p, int n) override {
        if (n != 15) fail(TT_ERR_ARG, "synthetic potential takes 12 strengths + 3 ranges");
        for (int i = 0; i < 12; ++i) th[i] = p[i];
        for (int i = 0; i < 3; ++i) beta[i] = p[12 + i];
    }
    static double g(int l, double b, double p) {
        double x = p / kHbarc;
        return (l ? x : 1.0) * std::exp(-(x / b) * (x / b));
    }
    void V(double pi, double po, bool, int S, int J, int, int Tz, double *v) const override {
        for (int i = 0; i < 6; ++i) v[i] = 0.0;
        double ri = pi / Lam, ro = po / Lam;
        double f = std::exp(-(ri * ri * ri * ri * ri * ri)) * std::exp(-(ro * ro * ro * ro * ro * ro));
        const double hc2 = kHbarc * kHbarc, hc4 = hc2 * hc2;
        auto yuk = [&](int L) {
            double s = 0.0;
            for (int i = 0; i < 3; ++i) s += th[i] / kHbarc * yukawa_pw(L, po, pi, mu[i]);
            return s;
        };
        if (S == 0) {
            v[0] = yuk(J);
        } else {
            v[1] = yuk(J);
            if (J >= 1) v[2] = yuk(J - 1);
            v[5] = yuk(J + 1);
        }
        if (S == 0 && J == 0) {
            double c = th[3] * (Tz == -1 ? 1.0 + cd_nn : 1.0);
            v[0] += c / hc2 + th[9] / hc2 * g(0, beta[0], pi) * g(0, beta[0], po);
        }
        if (S == 1 && J == 1) {
            v[2] += th[4] / hc2 + th[10] / hc2 * g(0, beta[1], pi) * g(0, beta[1], po);
            v[4] += th[8] * pi * pi / hc4;
            v[3] += th[8] * po * po / hc4;
            v[1] += th[7] * pi * po / hc4;
        }
        if (S == 1 && J == 0) v[5] += th[5] * pi * po / hc4;
        if (S == 0 && J == 1) {
            v[0] += th[6] * pi * po / hc4;
            v[0] += th[11] / hc2 * g(1, beta[2], pi) * g(1, beta[2], po);
        }
        for (int i = 0; i < 6; ++i) v[i] *= f;
    }
};

// CPP/make_potential_matrix.cpp:4-48
static double extract_element(int L, int Lp, int J, int S, bool coupled, const double *v) {
    if (J < std::abs(S - L) || J > S + L || J < std::abs(S - Lp) || J > S + Lp)
        fail(TT_ERR_PHYSICS, "Encountered unphysical state in LS-solver");
    const double sgn = -1.0;
    if (coupled) {
        if (L == Lp && L < J) return v[2];
        if (L == Lp && L > J) return v[5];
        if (L < Lp) return sgn * v[4];
        return sgn * v[3];
    }
    if (L == J + 1) return v[5];
    if (L == J - 1) return v[2];
    if (S == 0) return v[0];
    if (S == 1 && L == J) return v[1];
    return 0.0;
}

}  // namespace tt

using namespace tt;

struct tt_potential {
    std::unique_ptr<Potential> impl;
};

extern "C" {

tt_status tt_potential_create(const char *name, tt_potential **out) {
    TT_GUARD_BEGIN
    TT_REQUIRE(name && out, "null argument");
    std::string n(name);
    auto *p = new tt_potential;
    if (n == "malfliet_tjon")
        p->impl.reset(new MalflietTjon);
    else if (n == "synthetic")
        p->impl.reset(new Synthetic);
    else {
        delete p;
        fail(TT_ERR_ARG, "Invalid potential model entered: " + n);
    }
    *out = p;
    TT_GUARD_END
}

void tt_potential_destroy(tt_potential *p) { delete p; }

tt_status tt_potential_update_parameters(tt_potential *p, const double *params, int n) {
    TT_GUARD_BEGIN
    TT_REQUIRE(p && (n == 0 || params), "null argument");
    p->impl->update(params, n);
    TT_GUARD_END
}

tt_status tt_potential_V(tt_potential *p, double p_in, double p_out, int coupled, int S, int J, int T, int Tz,
                         double V[6]) {
    TT_GUARD_BEGIN
    TT_REQUIRE(p && V, "null argument");
    p->impl->V(p_in, p_out, coupled != 0, S, J, T, Tz, V);
    TT_GUARD_END
}

tt_status tt_calculate_potential_matrices(const tt_run_params *rp, const tt_fwp_states *fwp,
                                          const tt_pw_states *pw, tt_potential *pot, double *V_unco,
                                          double *V_coup) {
    TT_GUARD_BEGIN
    TT_REQUIRE(rp && fwp && pw && pot && V_unco, "null argument");
    int nu, nc;
    num_2N_states(*rp, nu, nc);
    TT_REQUIRE(nc == 0 || V_coup, "V_coup required");
    const int Np = fwp->Np_WP, npw = rp->midpoint_approx ? 1 : fwp->Np_per_WP;
    std::vector<char> done_u(nu, 0), done_c(nc, 0);
    const Potential &P = *pot->impl;
    // One task per unique 2N block: the (alpha, alpha') pair that first reaches it.
    struct Block {
        int Lr, Lc, Sr, Jr, Tr, iso, idx;
        bool via_L, via_T, cm, coupled_model;
    };
    std::vector<Block> blocks;
    for (int ar = 0; ar < pw->Nalpha; ++ar) {
        int Lr = pw->L_2N[ar], Sr = pw->S_2N[ar], Jr = pw->J_2N[ar], Tr = pw->T_2N[ar], t3r = pw->two_T_3N[ar];
        for (int ac = 0; ac < pw->Nalpha; ++ac) {
            int Lc = pw->L_2N[ac], t3c = pw->two_T_3N[ac];
            bool checkL = (rp->tensor_force && std::abs(Lr - Lc) <= 2) || Lr == Lc;
            if (!(Tr == pw->T_2N[ac] && Jr == pw->J_2N[ac] && Sr == pw->S_2N[ac] && checkL)) continue;
            bool s1S0 = (Sr == 0 && Jr == 0 && Lr == 0);
            int iso;  // CPP/make_potential_matrix.cpp:138-158, precedence of line 147 kept
            if (s1S0 && t3r == 3 && t3c == 3)
                iso = 2;
            else if ((s1S0 && (t3r == 3 && t3c == 1)) || (t3r == 1 && t3c == 3))
                iso = 3;
            else if (Tr == 0 && t3r == t3c && t3r == 1)
                iso = 0;
            else if (Tr == 1 && t3r == t3c && t3r == 1)
                iso = 1;
            else
                fail(TT_ERR_PHYSICS, "Unknown isoscalar-potential encountered in potential-matrix construction!");
            bool via_L, via_T;
            coupling_kind(Lr, Sr, Jr, Lc, *rp, via_L, via_T);
            bool cm = via_L || via_T;
            bool coupled_model = (Lr != Lc) || (Lr == Lc && Lr != Jr);
            int idx = unique_2N_idx(Lr, Sr, Jr, Tr, cm, *rp);
            if (cm) {
                TT_REQUIRE(idx >= 0 && idx < nc, "coupled 2N index out of range");
                if (done_c[idx]) continue;
                done_c[idx] = 1;
            } else {
                TT_REQUIRE(idx >= 0 && idx < nu, "uncoupled 2N index out of range");
                if (done_u[idx]) continue;
                done_u[idx] = 1;
            }
            if (!via_T && iso != 0 && iso != 1) fail(TT_ERR_PHYSICS, "Unknown isoscalar potential encountered.");
            blocks.push_back({Lr, Lc, Sr, Jr, Tr, iso, idx, via_L, via_T, cm, coupled_model});
        }
    }
    // Row br of block b: Np elements, each a npw x npw quadrature over the two bins.
    auto row = [&](const Block &B, int br) {
        const double Nr = fwp->norm_p[br];
        for (int bc = 0; bc < Np; ++bc) {
            double Nc = fwp->norm_p[bc];
            double acc[6] = {0, 0, 0, 0, 0, 0};
            for (int ir = 0; ir < npw; ++ir) {
                double pr, fr;
                if (rp->midpoint_approx) {
                    pr = fwp->p[br];
                    fr = pr * (fwp->p_WP[br + 1] - fwp->p_WP[br]);
                } else {
                    pr = fwp->p[br * npw + ir];
                    fr = fwp->wp[br * npw + ir] * pr;
                }
                for (int ic = 0; ic < npw; ++ic) {
                    double pc, fc;
                    if (rp->midpoint_approx) {
                        pc = fwp->p[bc];
                        fc = pc * (fwp->p_WP[bc + 1] - fwp->p_WP[bc]);
                    } else {
                        pc = fwp->p[bc * npw + ic];
                        fc = fwp->wp[bc * npw + ic] * pc;
                    }
                    double vis[6] = {0, 0, 0, 0, 0, 0}, vnn[6], vnp[6];
                    if (!B.via_T && B.iso == 0) {
                        P.V(pc, pr, B.coupled_model, B.Sr, B.Jr, B.Tr, 0, vis);
                    } else if (!B.via_T) {
                        P.V(pc, pr, B.coupled_model, B.Sr, B.Jr, B.Tr, -1, vnn);
                        P.V(pc, pr, B.coupled_model, B.Sr, B.Jr, B.Tr, 0, vnp);
                        for (int k = 0; k < 6; ++k) vis[k] = (2. / 3) * vnn[k] + (1. / 3) * vnp[k];
                    } else {
                        P.V(pc, pr, B.coupled_model, B.Sr, B.Jr, B.Tr, -1, vnn);
                        P.V(pc, pr, B.coupled_model, B.Sr, B.Jr, B.Tr, 0, vnp);
                        double enn = extract_element(B.Lr, B.Lc, B.Jr, B.Sr, false, vnn);
                        double enp = extract_element(B.Lr, B.Lc, B.Jr, B.Sr, false, vnp);
                        vis[2] = (2. / 3) * enn + (1. / 3) * enp;
                        vis[3] = (std::sqrt(2.0) / 3.) * (enn - enp);
                        vis[4] = (std::sqrt(2.0) / 3.) * (enn - enp);
                        vis[5] = (1. / 3) * enn + (2. / 3) * enp;
                    }
                    double fac = fr * fc;
                    for (int k = 0; k < 6; ++k) acc[k] += fac * vis[k] / (Nr * Nc);
                }
            }
            if (B.cm) {
                double *M = V_coup + (size_t)B.idx * 4 * Np * Np;
                size_t ul = (size_t)br * 2 * Np + bc, ur = ul + Np, ll = (size_t)(br + Np) * 2 * Np + bc,
                       lr = ll + Np;
                const int Jr = B.Jr, Sr = B.Sr;
                if (B.via_L) {
                    M[ul] = extract_element(Jr - 1, Jr - 1, Jr, Sr, true, acc);
                    M[ur] = extract_element(Jr - 1, Jr + 1, Jr, Sr, true, acc);
                    M[ll] = extract_element(Jr + 1, Jr - 1, Jr, Sr, true, acc);
                    M[lr] = extract_element(Jr + 1, Jr + 1, Jr, Sr, true, acc);
                } else {
                    M[ul] = acc[2];
                    M[ur] = acc[3];
                    M[ll] = acc[4];
                    M[lr] = acc[5];
                }
            } else {
                V_unco[(size_t)B.idx * Np * Np + (size_t)br * Np + bc] =
                    extract_element(B.Lr, B.Lc, B.Jr, B.Sr, false, acc);
            }
        }
    };
    // Rows are independent and each element is summed in the same order as a serial pass, so
    // the matrices do not depend on the thread count.
    parallel_rows((int64_t)blocks.size() * Np, [&](int64_t r) { row(blocks[r / Np], (int)(r % Np)); });
    TT_GUARD_END
}

tt_status tt_potential_basis_matrices(const tt_run_params *rp, const tt_fwp_states *fwp, const tt_pw_states *pw,
                                      tt_potential *pot, const double *params, int n_params, int n_theta,
                                      double *Vb_unco, double *Vb_coup) {
    TT_GUARD_BEGIN
    TT_REQUIRE(rp && fwp && pw && pot && params && Vb_unco && n_theta > 0 && n_theta <= n_params, "null argument");
    int nu, nc;
    num_2N_states(*rp, nu, nc);
    TT_REQUIRE(nc == 0 || Vb_coup, "null argument");
    const size_t Np = (size_t)fwp->Np_WP, su = (size_t)nu * Np * Np, sc = (size_t)nc * 4 * Np * Np;
    std::vector<double> p(params, params + n_params);
    try {
        for (int i = 0; i < n_theta; ++i) {
            std::fill(p.begin(), p.begin() + n_theta, 0.0);
            p[i] = 1.0;
            pot->impl->update(p.data(), n_params);
            tt_status st = tt_calculate_potential_matrices(rp, fwp, pw, pot, Vb_unco + (size_t)i * su,
                                                           nc ? Vb_coup + (size_t)i * sc : nullptr);
            if (st != TT_OK) fail(st, tt_last_error());
        }
    } catch (...) {
        pot->impl->update(params, n_params);
        throw;
    }
    pot->impl->update(params, n_params);
    TT_GUARD_END
}

}  // extern "C"
