// Host-side state spaces and conventions of the solve path:
//   partial-wave enumeration     CPP/make_pw_symm_states.cpp:4-283
//   FWP grids, normalisation     CPP/make_wp_states.cpp:3-211
//   Gauss-Legendre               CPP/General_functions/gauss_legendre.cpp:36-95
//   kinematics                   CPP/General_functions/kinetic_conversion.cpp:5-53
//   on-shell bins, deuterons     CPP/run_organizer.cpp:4-131,342-378
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "tt_internal.h"

namespace tt {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

void fail(tt_status code, const std::string &msg) { throw Error{code, msg}; }

int host_threads() {
    static const int n = [] {
        if (const char *e = std::getenv("TT_HOST_THREADS")) {
            const int v = std::atoi(e);
            if (v >= 1) return v;
        }
        const unsigned hw = std::thread::hardware_concurrency();
        return (int)std::max(1u, std::min(16u, hw));
    }();
    return n;
}

// ----------------------------------------------------------------- kinematics
static double deuteron_mass(double Eb) { return kMn + kMp + Eb; }

double mu1(double Eb) {
    double Md = deuteron_mass(Eb);
    return kMn * Md / (kMn + Md);
}

double com_momentum_to_lab_energy(double q_com, double Eb) {
    double Md = deuteron_mass(Eb);
    double b = 2 * (kMp - q_com * q_com / Md);
    double c = -q_com * q_com * (kMp + Md) * (kMp + Md) / (Md * Md);
    return 0.5 * (-b + std::sqrt(b * b - 4 * c));
}

double lab_energy_to_com_momentum(double T_lab, double Eb) {
    double Md = deuteron_mass(Eb);
    return std::sqrt(Md * Md * T_lab * (T_lab + 2 * kMp) / ((kMp + Md) * (kMp + Md) + 2 * T_lab * Md));
}

double com_energy_to_com_q_momentum(double E_com) { return std::sqrt(4 * E_com * kMN / 3); }
double com_q_momentum_to_com_energy(double q_com) { return 3 * q_com * q_com / (4 * kMN); }

// Newton iteration on the Legendre recurrence, nodes ascending.
void gauss_legendre(int n, double *x, double *w) {
    int m = (n + 1) / 2;
    for (int i = 1; i <= m; ++i) {
        double t = std::cos(M_PI * (i - 0.25) / (n + 0.5));
        double t_old = 1.0, dp = 0.0;
        for (int it = 0; it < 100 && std::fabs(t - t_old) >= 3e-16; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((2 * j - 1) * t * p2 - (j - 1) * p3) / j;
            }
            dp = n * (t * p1 - p2) / (t * t - 1);
            t_old = t;
            t = t_old - p1 / dp;
        }
        x[i - 1] = -t;
        x[n - i] = t;
        w[i - 1] = 2.0 / ((1 - t * t) * dp * dp);
        w[n - i] = w[i - 1];
    }
}

void gauss_legendre_ab(int n, double a, double b, double *x, double *w) {
    gauss_legendre(n, x, w);
    for (int i = 0; i < n; ++i) {
        x[i] = 0.5 * (b - a) * x[i] + 0.5 * (b + a);
        w[i] = 0.5 * (b - a) * w[i];
    }
}

// ----------------------------------------------------------------- 2N indexing
int unique_2N_idx(int L, int S, int J, int T, bool coupled, const tt_run_params &rp) {
    if (coupled && !rp.tensor_force) fail(TT_ERR_ARG, "coupled 2N state without a tensor force");
    bool s1S0 = (S == 0 && J == 0 && L == 0 && T == 1);
    int u;
    if (rp.tensor_force) {
        if (coupled) {
            u = J - 1;
            if (!s1S0 && rp.isospin_breaking_1S0) u += 1;
            if (s1S0 && rp.isospin_breaking_1S0) u = 0;
        } else {
            u = 2 * J + S;
            if (rp.isospin_breaking_1S0) u -= 1;
        }
    } else {
        u = J * (4 * (J > 1) + 2) + (S == 1) + (S == 1 && J != 0) * (J - L + 1);
        if (rp.isospin_breaking_1S0) u -= 1;
    }
    return u;
}

void num_2N_states(const tt_run_params &rp, int &nu, int &nc) {
    int J = rp.J_2N_max;
    if (rp.tensor_force) {
        nu = 2 * (J + 1);
        nc = J;
        if (rp.isospin_breaking_1S0) {
            nu -= 1;
            nc += 1;
        }
    } else {
        nu = 4 * J + 2;
        nc = 0;
        if (rp.isospin_breaking_1S0) {
            nu -= 1;
            nc = 1;
        }
    }
}

void coupling_kind(int Lr, int Sr, int Jr, int Lc, const tt_run_params &rp, bool &via_L, bool &via_T) {
    bool s1S0 = (Sr == 0 && Jr == 0 && Lr == 0);
    via_L = rp.tensor_force && (Lr != Lc || (Lr == Lc && Lr != Jr && Jr != 0));
    via_T = s1S0 && rp.isospin_breaking_1S0;
    if (via_L && via_T) fail(TT_ERR_UNSUPPORTED, "isospin breaking in a coupled channel");
}

static int half_of(const tt_pw_states &chn, int a, bool via_L) {
    if (via_L) return chn.L_2N[a] < chn.J_2N[a] ? 0 : 1;
    return chn.two_T_3N[a] == 1 ? 0 : 1;
}

std::vector<double> embed_2N(const tt_pw_states &chn, int Np, const double *M_unco, const double *M_coup,
                             const tt_run_params &rp) {
    const int Na = chn.Nalpha;
    const int64_t n = (int64_t)Na * Np;
    std::vector<double> B((size_t)(n * n), 0.0);
    for (int ar = 0; ar < Na; ++ar) {
        for (int ac = 0; ac < Na; ++ac) {
            int Lr = chn.L_2N[ar], Sr = chn.S_2N[ar], Jr = chn.J_2N[ar], Tr = chn.T_2N[ar];
            int Lc = chn.L_2N[ac];
            bool checkL = (rp.tensor_force && std::abs(Lr - Lc) <= 2) || Lr == Lc;
            if (!(Tr == chn.T_2N[ac] && Jr == chn.J_2N[ac] && Sr == chn.S_2N[ac] && checkL &&
                  chn.L_1N[ar] == chn.L_1N[ac] && chn.two_J_1N[ar] == chn.two_J_1N[ac]))
                continue;
            bool via_L, via_T;
            coupling_kind(Lr, Sr, Jr, Lc, rp, via_L, via_T);
            for (int pr = 0; pr < Np; ++pr)
                for (int pc = 0; pc < Np; ++pc) {
                    double v;
                    if (via_L || via_T) {
                        int idx = unique_2N_idx(Lr, Sr, Jr, Tr, true, rp);
                        int hr = half_of(chn, ar, via_L), hc = half_of(chn, ac, via_L);
                        v = M_coup[(size_t)idx * 4 * Np * Np + (size_t)(hr * Np + pr) * 2 * Np + hc * Np + pc];
                    } else {
                        int idx = unique_2N_idx(Lr, Sr, Jr, Tr, false, rp);
                        v = M_unco[(size_t)idx * Np * Np + (size_t)pr * Np + pc];
                    }
                    B[(size_t)(ar * Np + pr) * n + ac * Np + pc] = v;
                }
        }
    }
    return B;
}

// Partner lists of the embedded operator B = embed_2N(M) without forming it: for each alpha
// a the partners a' whose Np x Np block is not all zero (keep_zero: every structurally
// coupled partner, so the lists depend on the channel only), block[p][p'] = B[(a,p),(a',p')]
// (transpose: B[(a',p'),(a,p)]).
void embed_2N_blocks(const tt_pw_states &chn, int Np, const double *M_unco, const double *M_coup,
                     const tt_run_params &rp, bool transpose, std::vector<int> &ptr, std::vector<int> &alpha,
                     std::vector<double> &blocks, bool keep_zero) {
    const int Na = chn.Nalpha;
    ptr.assign(Na + 1, 0);
    alpha.clear();
    blocks.clear();
    std::vector<double> blk((size_t)Np * Np);
    for (int a = 0; a < Na; ++a) {
        for (int b = 0; b < Na; ++b) {
            const int ar = transpose ? b : a, ac = transpose ? a : b;  // element B[(ar,.),(ac,.)]
            int Lr = chn.L_2N[ar], Sr = chn.S_2N[ar], Jr = chn.J_2N[ar], Tr = chn.T_2N[ar];
            int Lc = chn.L_2N[ac];
            bool checkL = (rp.tensor_force && std::abs(Lr - Lc) <= 2) || Lr == Lc;
            if (!(Tr == chn.T_2N[ac] && Jr == chn.J_2N[ac] && Sr == chn.S_2N[ac] && checkL &&
                  chn.L_1N[ar] == chn.L_1N[ac] && chn.two_J_1N[ar] == chn.two_J_1N[ac]))
                continue;
            bool via_L, via_T;
            coupling_kind(Lr, Sr, Jr, Lc, rp, via_L, via_T);
            // the 2N block this pair reads: element (pr, pc) at base[pr * ld + pc]
            const double *base;
            size_t ld;
            if (via_L || via_T) {
                const int idx = unique_2N_idx(Lr, Sr, Jr, Tr, true, rp);
                const int hr = half_of(chn, ar, via_L), hc = half_of(chn, ac, via_L);
                base = M_coup + (size_t)idx * 4 * Np * Np + (size_t)hr * Np * 2 * Np + (size_t)hc * Np;
                ld = 2 * (size_t)Np;
            } else {
                const int idx = unique_2N_idx(Lr, Sr, Jr, Tr, false, rp);
                base = M_unco + (size_t)idx * Np * Np;
                ld = Np;
            }
            bool nz = false;
            for (int p = 0; p < Np; ++p)
                for (int pp = 0; pp < Np; ++pp) {
                    const int pr = transpose ? pp : p, pc = transpose ? p : pp;
                    const double v = base[(size_t)pr * ld + pc];
                    blk[(size_t)p * Np + pp] = v;
                    nz = nz || v != 0.0;
                }
            if (!nz && !keep_zero) continue;
            alpha.push_back(b);
            blocks.insert(blocks.end(), blk.begin(), blk.end());
        }
        ptr[a + 1] = (int)alpha.size();
    }
}

void embed_2N_layout(const tt_pw_states &chn, int Np, int n_unco, const tt_run_params &rp, bool transpose,
                     std::vector<int> &ptr, std::vector<int> &alpha, std::vector<int64_t> &src,
                     std::vector<int> &ld) {
    const int Na = chn.Nalpha;
    ptr.assign(Na + 1, 0);
    alpha.clear();
    src.clear();
    ld.clear();
    for (int a = 0; a < Na; ++a) {
        for (int b = 0; b < Na; ++b) {
            const int ar = transpose ? b : a, ac = transpose ? a : b;
            int Lr = chn.L_2N[ar], Sr = chn.S_2N[ar], Jr = chn.J_2N[ar], Tr = chn.T_2N[ar];
            int Lc = chn.L_2N[ac];
            bool checkL = (rp.tensor_force && std::abs(Lr - Lc) <= 2) || Lr == Lc;
            if (!(Tr == chn.T_2N[ac] && Jr == chn.J_2N[ac] && Sr == chn.S_2N[ac] && checkL &&
                  chn.L_1N[ar] == chn.L_1N[ac] && chn.two_J_1N[ar] == chn.two_J_1N[ac]))
                continue;
            bool via_L, via_T;
            coupling_kind(Lr, Sr, Jr, Lc, rp, via_L, via_T);
            if (via_L || via_T) {
                const int idx = unique_2N_idx(Lr, Sr, Jr, Tr, true, rp);
                const int hr = half_of(chn, ar, via_L), hc = half_of(chn, ac, via_L);
                src.push_back((int64_t)n_unco * Np * Np + (int64_t)idx * 4 * Np * Np + (int64_t)hr * Np * 2 * Np +
                              (int64_t)hc * Np);
                ld.push_back(2 * Np);
            } else {
                src.push_back((int64_t)unique_2N_idx(Lr, Sr, Jr, Tr, false, rp) * Np * Np);
                ld.push_back(Np);
            }
            alpha.push_back(b);
        }
        ptr[a + 1] = (int)alpha.size();
    }
}

void swp_edge_offsets(const tt_pw_states &chn, const tt_swp_states &swp, const tt_run_params &rp,
                      std::vector<int> &which, std::vector<int> &offset) {
    const int Np = swp.Np_WP;
    which.assign(chn.Nalpha, 0);
    offset.assign(chn.Nalpha, 0);
    for (int a = 0; a < chn.Nalpha; ++a) {
        int L = chn.L_2N[a], S = chn.S_2N[a], J = chn.J_2N[a], T = chn.T_2N[a];
        bool s1S0 = (S == 0 && J == 0 && L == 0);
        bool via_L = rp.tensor_force && L != J && J != 0;
        bool via_T = s1S0 && rp.isospin_breaking_1S0;
        if (via_L && via_T) fail(TT_ERR_UNSUPPORTED, "isospin breaking in a coupled channel");
        if (via_L || via_T) {
            int idx = unique_2N_idx(L, S, J, T, true, rp);
            bool upper = via_L ? (L >= J) : (chn.two_T_3N[a] != 1);
            which[a] = 1;
            offset[a] = idx * 2 * (Np + 1) + (upper ? Np + 1 : 0);
            TT_REQUIRE(idx < swp.num_2N_coup_states, "coupled 2N index out of range");
        } else {
            int idx = unique_2N_idx(L, S, J, T, false, rp);
            which[a] = 0;
            offset[a] = idx * (Np + 1);
            TT_REQUIRE(idx < swp.num_2N_unco_states, "uncoupled 2N index out of range");
        }
    }
}

}  // namespace tt

using namespace tt;

extern "C" {

const char *tt_last_error(void) { return g_last_error.c_str(); }
const char *tt_version(void) { return "tictac-mi355x 0.1"; }

void tt_run_params_default(tt_run_params *rp) {
    // CPP/set_run_parameters.cpp:577-615
    rp->two_J_3N_max = 1;
    rp->J_2N_max = 3;
    rp->Np_WP = 50;
    rp->Nq_WP = 50;
    rp->Np_per_WP = 8;
    rp->Nphi = 48;
    rp->Nx = 48;
    rp->chebyshev_s = 100;
    rp->chebyshev_t = 1;
    rp->tensor_force = 1;
    rp->isospin_breaking_1S0 = 1;
    rp->midpoint_approx = 0;
    rp->swave_only = 0;
}

void tt_num_2N_states(const tt_run_params *rp, int *nu, int *nc) { num_2N_states(*rp, *nu, *nc); }

static int *dup_ints(const std::vector<int> &v) {
    int *p = (int *)std::malloc(sizeof(int) * std::max<size_t>(v.size(), 1));
    if (!p) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(int) * v.size());
    return p;
}

static double *dup_doubles(const std::vector<double> &v) {
    double *p = (double *)std::malloc(sizeof(double) * std::max<size_t>(v.size(), 1));
    if (!p) throw std::bad_alloc();
    if (!v.empty()) std::memcpy(p, v.data(), sizeof(double) * v.size());
    return p;
}

tt_status tt_construct_pw_states(const tt_run_params *rp, tt_pw_states *out) {
    TT_GUARD_BEGIN
    TT_REQUIRE(rp && out, "null argument");
    TT_REQUIRE(rp->J_2N_max >= 0, "Cannot have negative J_2N!");
    TT_REQUIRE(rp->two_J_3N_max > 0 && rp->two_J_3N_max % 2 == 1, "Cannot have even two_J_3N_max!");
    int two_T3_max = rp->isospin_breaking_1S0 ? 3 : 1;
    std::vector<int> L, S, J, T, l, tj, tJ3, tT3, P3, chn;
    int n_prev = 0;
    // loop order of CPP/make_pw_symm_states.cpp:122-228
    for (int two_J3 = 1; two_J3 <= rp->two_J_3N_max; two_J3 += 2) {
        for (int Prem = 0; Prem < 2; ++Prem) {
            int n_in = 0;
            for (int two_T3 = 1; two_T3 <= two_T3_max; two_T3 += 2)
                for (int J2 = 0; J2 <= rp->J_2N_max; ++J2)
                    for (int S2 = 0; S2 < 2; ++S2)
                        for (int L2 = std::abs(J2 - S2); L2 <= J2 + S2; ++L2) {
                            int Tmin = std::abs(two_T3 - 1) / 2;
                            int Tmax = std::min((two_T3 + 1) / 2, 1);
                            for (int T2 = Tmin; T2 <= Tmax; ++T2) {
                                if (!((L2 + S2 + T2) % 2)) continue;
                                for (int two_j = std::abs(two_J3 - 2 * J2); two_j <= two_J3 + 2 * J2; two_j += 2)
                                    for (int l1 = std::abs(two_j - 1) / 2; l1 <= (two_j + 1) / 2; ++l1) {
                                        if ((L2 + l1) % 2 != Prem) continue;
                                        if (two_T3 == 3 && !(S2 == 0 && L2 == 0 && J2 == 0)) continue;
                                        if (rp->swave_only && (L2 != 0 || l1 != 0)) continue;
                                        L.push_back(L2);
                                        S.push_back(S2);
                                        J.push_back(J2);
                                        T.push_back(T2);
                                        l.push_back(l1);
                                        tj.push_back(two_j);
                                        tJ3.push_back(two_J3);
                                        tT3.push_back(two_T3);
                                        P3.push_back(1 - 2 * Prem);
                                        ++n_in;
                                    }
                            }
                        }
            if (n_in) chn.push_back(n_prev);
            n_prev += n_in;
        }
    }
    chn.push_back((int)L.size());
    out->Nalpha = (int)L.size();
    out->J_2N_max = rp->J_2N_max;
    out->L_2N = dup_ints(L);
    out->S_2N = dup_ints(S);
    out->J_2N = dup_ints(J);
    out->T_2N = dup_ints(T);
    out->L_1N = dup_ints(l);
    out->two_J_1N = dup_ints(tj);
    out->two_J_3N = dup_ints(tJ3);
    out->two_T_3N = dup_ints(tT3);
    out->P_3N = dup_ints(P3);
    out->chn_3N_idx = dup_ints(chn);
    out->N_chn_3N = (int)chn.size() - 1;
    TT_GUARD_END
}

void tt_pw_states_free(tt_pw_states *pw) {
    if (!pw) return;
    int **arrs[] = {&pw->L_2N, &pw->S_2N, &pw->J_2N, &pw->T_2N, &pw->L_1N, &pw->two_J_1N,
                    &pw->two_J_3N, &pw->two_T_3N, &pw->P_3N, &pw->chn_3N_idx};
    for (int **a : arrs) {
        std::free(*a);
        *a = nullptr;
    }
    pw->Nalpha = 0;
    pw->N_chn_3N = 0;
}

tt_status tt_pw_channel(const tt_pw_states *pw, int c, tt_pw_states *v) {
    TT_GUARD_BEGIN
    TT_REQUIRE(pw && v && c >= 0 && c < pw->N_chn_3N, "channel index out of range");
    int lo = pw->chn_3N_idx[c], hi = pw->chn_3N_idx[c + 1];
    v->Nalpha = hi - lo;
    v->J_2N_max = pw->J_2N_max;
    v->L_2N = pw->L_2N + lo;
    v->S_2N = pw->S_2N + lo;
    v->J_2N = pw->J_2N + lo;
    v->T_2N = pw->T_2N + lo;
    v->L_1N = pw->L_1N + lo;
    v->two_J_1N = pw->two_J_1N + lo;
    v->two_J_3N = pw->two_J_3N + lo;
    v->two_T_3N = pw->two_T_3N + lo;
    v->P_3N = pw->P_3N + lo;
    v->chn_3N_idx = nullptr;
    v->N_chn_3N = 1;
    TT_GUARD_END
}

tt_status tt_make_fwp_statespace(const tt_run_params *rp, tt_fwp_states *f) {
    TT_GUARD_BEGIN
    TT_REQUIRE(rp && f && rp->Np_WP > 1 && rp->Nq_WP > 1, "bad FWP dimensions");
    std::memset(f, 0, sizeof(*f));
    f->Np_WP = rp->Np_WP;
    f->Nq_WP = rp->Nq_WP;
    f->Np_per_WP = rp->Np_per_WP;
    f->Nq_per_WP = rp->Np_per_WP;
    auto edges = [&](int N) {
        std::vector<double> e(N + 1, 0.0);
        for (int i = 1; i <= N; ++i)
            e[i] = rp->chebyshev_s * std::pow(std::tan((2 * i - 1) * M_PI / (4 * N)), rp->chebyshev_t);
        return e;
    };
    std::vector<double> pe = edges(rp->Np_WP), qe = edges(rp->Nq_WP);
    auto mesh = [&](const std::vector<double> &e, int npw, std::vector<double> &x, std::vector<double> &w) {
        int N = (int)e.size() - 1;
        if (rp->midpoint_approx) {
            x.resize(N);
            w.assign(N, 0.0);
            for (int i = 0; i < N; ++i) x[i] = 0.5 * (e[i] + e[i + 1]);
        } else {
            x.resize((size_t)N * npw);
            w.resize((size_t)N * npw);
            for (int b = 0; b < N; ++b) gauss_legendre_ab(npw, e[b], e[b + 1], &x[(size_t)b * npw], &w[(size_t)b * npw]);
        }
    };
    std::vector<double> px, pw_, qx, qw;
    mesh(pe, rp->Np_per_WP, px, pw_);
    mesh(qe, rp->Np_per_WP, qx, qw);
    std::vector<double> np_(rp->Np_WP), nq_(rp->Nq_WP);
    for (int i = 0; i < rp->Np_WP; ++i) np_[i] = std::sqrt(pe[i + 1] - pe[i]);
    for (int i = 0; i < rp->Nq_WP; ++i) nq_[i] = std::sqrt(qe[i + 1] - qe[i]);
    f->p_WP = dup_doubles(pe);
    f->q_WP = dup_doubles(qe);
    f->p = dup_doubles(px);
    f->wp = dup_doubles(pw_);
    f->q = dup_doubles(qx);
    f->wq = dup_doubles(qw);
    f->norm_p = dup_doubles(np_);
    f->norm_q = dup_doubles(nq_);
    TT_GUARD_END
}

void tt_fwp_states_free(tt_fwp_states *f) {
    if (!f) return;
    double **arrs[] = {&f->p_WP, &f->q_WP, &f->p, &f->wp, &f->q, &f->wq, &f->norm_p, &f->norm_q};
    for (double **a : arrs) {
        std::free(*a);
        *a = nullptr;
    }
}

tt_status tt_find_on_shell(const tt_swp_states *swp, const tt_pw_states *chn, const double *T_in, int nT,
                           tt_channel_os *out) {
    TT_GUARD_BEGIN
    TT_REQUIRE(swp && chn && out && (nT == 0 || T_in), "null argument");
    std::memset(out, 0, sizeof(*out));
    const int Nq = swp->Nq_WP;
    const double *qe = swp->q_WP;
    const double Eb = swp->E_bound;
    // bin-midpoint lab energies, CPP/run_organizer.cpp:41-48
    std::vector<double> Tmid(Nq);
    for (int k = 0; k < Nq; ++k) {
        double Eql = 0.5 * qe[k] * qe[k] / mu1(Eb);
        double Equ = 0.5 * qe[k + 1] * qe[k + 1] / mu1(Eb);
        double Ec = 0.5 * (Equ + Eql);
        Tmid[k] = com_momentum_to_lab_energy(com_energy_to_com_q_momentum(Ec), Eb);
    }
    std::vector<char> mid(Nq, 0), onbin(Nq, 0);
    for (int k = 0; k + 1 < Nq; ++k)
        for (int i = 0; i < nT; ++i)
            if (Tmid[k] <= T_in[i] && T_in[i] <= Tmid[k + 1]) mid[k] = 1;
    for (int k = 0; k + 1 < Nq; ++k)
        if (mid[k]) onbin[k] = onbin[k + 1] = 1;
    std::vector<int> bins;
    for (int k = 0; k < Nq; ++k)
        if (onbin[k]) bins.push_back(k);
    std::vector<double> T(bins.size()), qc(bins.size()), Ec(bins.size());
    std::vector<int> qidx(bins.size());
    for (size_t i = 0; i < bins.size(); ++i) {
        T[i] = Tmid[bins[i]];
        qc[i] = lab_energy_to_com_momentum(T[i], Eb);
        Ec[i] = com_q_momentum_to_com_energy(qc[i]);
        int found = -1;
        for (int k = 0; k < Nq; ++k)
            if (qe[k] < qc[i] && qc[i] < qe[k + 1]) {
                found = k;
                break;
            }
        if (found < 0) fail(TT_ERR_PHYSICS, "on-shell T_lab does not exist in the WP state space");
        qidx[i] = found;
    }
    std::vector<int> deut;
    for (int a = 0; a < chn->Nalpha; ++a)
        if (chn->L_2N[a] == kDeuteronL && chn->S_2N[a] == kDeuteronS && chn->J_2N[a] == kDeuteronJ &&
            chn->T_2N[a] == kDeuteronT)
            deut.push_back(a);
    out->num_T_lab = (int)bins.size();
    out->q_com_idx = dup_ints(qidx);
    out->T_lab = dup_doubles(T);
    out->q_com = dup_doubles(qc);
    out->E_com = dup_doubles(Ec);
    out->num_deuteron_states = (int)deut.size();
    out->deuteron_idx = dup_ints(deut);
    TT_GUARD_END
}

tt_status tt_find_breakup_bins(const tt_run_params *rp, const tt_swp_states *swp, const tt_pw_states *chn,
                               tt_channel_os *os) {
    TT_GUARD_BEGIN
    TT_REQUIRE(rp && swp && chn && os && (os->num_T_lab == 0 || os->E_com), "null argument");
    const int Na = chn->Nalpha, Nq = swp->Nq_WP, Np = swp->Np_WP, nE = os->num_T_lab;
    const double *qe = swp->q_WP;
    // the p-bin energy edges of every alpha (the same choice as the resolvent's)
    std::vector<int> which, offset;
    swp_edge_offsets(*chn, *swp, *rp, which, offset);
    // E_q of the q-bin edges with the free spectator mass (mu1_free, kinetic_conversion.cpp:16-19)
    const double M12 = kMp + kMn, mu_free = kMn * M12 / (kMn + M12);
    std::vector<int> apq, ptr(1, 0);
    for (int j = 0; j < nE; ++j) {
        const double E = os->E_com[j];
        if (E + swp->E_bound < 0) {   // below the breakup threshold: no entries
            ptr.push_back((int)(apq.size() / 3));
            continue;
        }
        bool found = false;
        auto add = [&](int a, int q, int p) {
            apq.push_back(a);
            apq.push_back(q);
            apq.push_back(p);
        };
        for (int a = 0; a < Na; ++a) {
            const double *ep = (which[a] ? swp->e_SWP_coup : swp->e_SWP_unco) + offset[a];
            for (int q = 0; q < Nq; ++q) {
                const double Eql = 0.5 * qe[q] * qe[q] / mu_free, Equ = 0.5 * qe[q + 1] * qe[q + 1] / mu_free;
                for (int p = 0; p < Np; ++p) {
                    const double Epl = ep[p], Epu = ep[p + 1];
                    const double lo = Epl + Eql, hi = Epu + Equ;
                    if (!(lo <= E && E <= hi && Epl >= 0 && Epu >= 0)) continue;
                    found = true;
                    add(a, q, p);
                    // the quadrant of the cell (split at its bin midpoints) whose energy range
                    // holds E, tried in the order upper-upper, lower-p/upper-q, lower-lower,
                    // upper-p/lower-q; its three neighbours follow (run_organizer.cpp:250-315)
                    const double Epm = 0.5 * (Epu + Epl), Eqm = 0.5 * (Equ + Eql);
                    int dp = 0, dq = 0;
                    if (Eqm + Epm <= E && E <= hi) {
                        dp = 1;
                        dq = 1;
                    } else if (Eqm + Epl <= E && E <= Epm + Equ) {
                        dp = -1;
                        dq = 1;
                    } else if (lo <= E && E <= Eqm + Epm) {
                        dp = -1;
                        dq = -1;
                    } else if (Eql + Epm <= E && E <= Epu + Eqm) {
                        dp = 1;
                        dq = -1;
                    }
                    if (dp == 0) continue;
                    // the 2x2 block spanned by the cell and (p + dp, q + dq), p-major, q inner,
                    // both ascending, without the cell itself
                    const int p0 = std::min(p, p + dp), q0 = std::min(q, q + dq);
                    for (int pi = p0; pi <= p0 + 1; ++pi)
                        for (int qj = q0; qj <= q0 + 1; ++qj)
                            if (pi != p || qj != q) add(a, qj, pi);
                }
            }
        }
        if (!found) fail(TT_ERR_PHYSICS, "on-shell breakup energy E_com=" + std::to_string(E) +
                                             " MeV has no cell in the WP state space");
        ptr.push_back((int)(apq.size() / 3));
    }
    std::free(os->q_com_BU_idx);
    std::free(os->alphapq_idx);
    os->q_com_BU_idx = dup_ints(ptr);
    os->alphapq_idx = dup_ints(apq.empty() ? std::vector<int>(3, 0) : apq);
    os->num_BU_chns = (int)(apq.size() / 3);
    TT_GUARD_END
}

tt_status tt_channel_p_edges(const tt_run_params *rp, const tt_swp_states *swp, const tt_pw_states *chn,
                             double *edges) {
    TT_GUARD_BEGIN
    TT_REQUIRE(rp && swp && chn && edges, "null argument");
    std::vector<int> which, offset;
    swp_edge_offsets(*chn, *swp, *rp, which, offset);
    const int Np = swp->Np_WP;
    for (int a = 0; a < chn->Nalpha; ++a) {
        const double *e = (which[a] ? swp->e_SWP_coup : swp->e_SWP_unco) + offset[a];
        std::memcpy(edges + (size_t)a * (Np + 1), e, sizeof(double) * (Np + 1));
    }
    TT_GUARD_END
}

void tt_channel_os_free(tt_channel_os *os) {
    if (!os) return;
    std::free(os->q_com_idx);
    std::free(os->T_lab);
    std::free(os->q_com);
    std::free(os->E_com);
    std::free(os->deuteron_idx);
    std::free(os->q_com_BU_idx);
    std::free(os->alphapq_idx);
    std::memset(os, 0, sizeof(*os));
}

}  // extern "C"
