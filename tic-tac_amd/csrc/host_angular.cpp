// Angular algebra of P123 on the host:
//   Wigner 3j/6j/9j and Clebsch-Gordan   (the GSL calls of CPP/auxiliary.cpp:127-166)
//   A-tilde                              CPP/auxiliary.cpp:863-884
//   G-tilde as a sum of angular tuples   CPP/auxiliary.cpp:950-1067
// Wigner symbols use Racah's single-sum formulas in long double with tabulated
// factorials; the spectator spins here are 1/2, so every sum has only a few terms.
#include <map>
#include <tuple>

#include "tt_internal.h"

namespace tt {

static const long double *factorials() {
    static long double f[171];
    static bool init = false;
    if (!init) {
        f[0] = 1.0L;
        for (int i = 1; i <= 170; ++i) f[i] = f[i - 1] * (long double)i;
        init = true;
    }
    return f;
}

static inline bool tri_ok(int a, int b, int c) { return ((a + b + c) % 2 == 0) && std::abs(a - b) <= c && c <= a + b; }

static long double delta2(int a, int b, int c) {
    const long double *F = factorials();
    return F[(a + b - c) / 2] * F[(a - b + c) / 2] * F[(-a + b + c) / 2] / F[(a + b + c) / 2 + 1];
}

double wigner3j(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3) {
    if (tm1 + tm2 + tm3 != 0 || !tri_ok(tj1, tj2, tj3)) return 0.0;
    if (std::abs(tm1) > tj1 || std::abs(tm2) > tj2 || std::abs(tm3) > tj3) return 0.0;
    if ((tj1 + tm1) % 2 || (tj2 + tm2) % 2 || (tj3 + tm3) % 2) return 0.0;
    const long double *F = factorials();
    int j1p = (tj1 + tm1) / 2, j1m = (tj1 - tm1) / 2, j2p = (tj2 + tm2) / 2, j2m = (tj2 - tm2) / 2;
    int j3p = (tj3 + tm3) / 2, j3m = (tj3 - tm3) / 2;
    long double pref = delta2(tj1, tj2, tj3) * F[j1p] * F[j1m] * F[j2p] * F[j2m] * F[j3p] * F[j3m];
    int a1 = (tj3 - tj2 + tm1) / 2, a2 = (tj3 - tj1 - tm2) / 2, b1 = (tj1 + tj2 - tj3) / 2, b2 = j1m, b3 = j2p;
    int kmin = std::max(0, std::max(-a1, -a2)), kmax = std::min(b1, std::min(b2, b3));
    long double s = 0.0L;
    for (int k = kmin; k <= kmax; ++k) {
        long double den = F[k] * F[b1 - k] * F[b2 - k] * F[b3 - k] * F[a1 + k] * F[a2 + k];
        s += ((k & 1) ? -1.0L : 1.0L) / den;
    }
    int ph = ((tj1 - tj2 - tm3) / 2) & 1;
    return (double)((ph ? -1.0L : 1.0L) * std::sqrt(pref) * s);
}

double clebsch_gordan(int tj1, int tj2, int tj3, int tm1, int tm2, int tm3) {
    if (std::abs(tm1) > tj1 || std::abs(tm2) > tj2) return 0.0;
    double w = wigner3j(tj1, tj2, tj3, tm1, tm2, -tm3);
    if (w == 0.0) return 0.0;
    int ph = (((tj1 - tj2 + tm3) / 2) % 2 + 2) % 2;
    return (ph ? -1.0 : 1.0) * std::sqrt((double)tj3 + 1.0) * w;
}

double wigner6j(int a, int b, int c, int d, int e, int f) {
    if (!tri_ok(a, b, c) || !tri_ok(a, e, f) || !tri_ok(d, b, f) || !tri_ok(d, e, c)) return 0.0;
    const long double *F = factorials();
    long double pref = delta2(a, b, c) * delta2(a, e, f) * delta2(d, b, f) * delta2(d, e, c);
    int t1 = (a + b + c) / 2, t2 = (a + e + f) / 2, t3 = (d + b + f) / 2, t4 = (d + e + c) / 2;
    int u1 = (a + b + d + e) / 2, u2 = (a + c + d + f) / 2, u3 = (b + c + e + f) / 2;
    long double s = 0.0L;
    for (int k = std::max(std::max(t1, t2), std::max(t3, t4)); k <= std::min(u1, std::min(u2, u3)); ++k) {
        long double term = F[k + 1] / (F[k - t1] * F[k - t2] * F[k - t3] * F[k - t4] * F[u1 - k] * F[u2 - k] * F[u3 - k]);
        s += (k & 1) ? -term : term;
    }
    return (double)(std::sqrt(pref) * s);
}

double wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i) {
    int xmin = std::max(std::abs(a - i), std::max(std::abs(h - d), std::abs(b - f)));
    int xmax = std::min(a + i, std::min(h + d, b + f));
    long double tot = 0.0L;
    for (int x = xmin; x <= xmax; x += 2) {
        double w1 = wigner6j(a, b, c, f, i, x);
        if (w1 == 0.0) continue;
        double w2 = wigner6j(d, e, f, b, x, h);
        if (w2 == 0.0) continue;
        double w3 = wigner6j(g, h, i, x, a, d);
        tot += ((x & 1) ? -1.0L : 1.0L) * (long double)(x + 1) * w1 * w2 * w3;
    }
    return (double)tot;
}

// CPP/auxiliary.cpp:863-884 with alpha = a, alpha' = ap
double atilde(const tt_pw_states &p, int a, int ap, int Ltot, int two_J, int two_T) {
    double ret = 0.0;
    for (int two_S = 1; two_S <= 3; two_S += 2) {
        double pre = std::sqrt((double)(2 * p.J_2N[a] + 1) * (p.two_J_1N[a] + 1) * (2 * p.S_2N[a] + 1) * (2 * p.T_2N[a] + 1) *
                               (2 * p.J_2N[ap] + 1) * (p.two_J_1N[ap] + 1) * (2 * p.S_2N[ap] + 1) *
                               (2 * p.T_2N[ap] + 1));
        double ph = ((p.S_2N[ap] + p.T_2N[ap]) & 1) ? -1.0 : 1.0;
        ret += pre * ph * (two_S + 1) *
               wigner9j(2 * p.L_2N[a], 2 * p.S_2N[a], 2 * p.J_2N[a], 2 * p.L_1N[a], 1, p.two_J_1N[a], 2 * Ltot, two_S, two_J) *
               wigner9j(2 * p.L_2N[ap], 2 * p.S_2N[ap], 2 * p.J_2N[ap], 2 * p.L_1N[ap], 1, p.two_J_1N[ap], 2 * Ltot, two_S,
                        two_J) *
               wigner6j(1, 1, 2 * p.S_2N[a], 1, two_S, 2 * p.S_2N[ap]) * wigner6j(1, 1, 2 * p.T_2N[a], 1, two_T, 2 * p.T_2N[ap]);
    }
    return ret;
}

P123Angular build_p123_angular(const tt_pw_states &chn) {
    P123Angular G;
    const int Na = chn.Nalpha;
    const int two_J = chn.two_J_3N[0];
    int maxL = 0, maxl = 0;
    for (int a = 0; a < Na; ++a) {
        maxL = std::max(maxL, chn.L_2N[a]);
        maxl = std::max(maxl, chn.L_1N[a]);
    }
    G.lmax1 = maxl;
    G.lmax2 = maxL;
    G.lmax3 = maxl;
    std::map<std::tuple<int, int, int, int, int>, int> tuple_id;
    std::vector<std::tuple<int, int, int, int, int>> tuples;
    G.pair_ptr.assign((size_t)Na * Na + 1, 0);
    for (int ap = 0; ap < Na; ++ap) {          // bra (row)
        for (int a = 0; a < Na; ++a) {         // ket (column)
            if (chn.two_T_3N[ap] == chn.two_T_3N[a]) {
                const int L12 = chn.L_2N[ap], l3 = chn.L_1N[ap], L12p = chn.L_2N[a], l3p = chn.L_1N[a];
                const int lo = std::max(std::abs(L12 - l3), std::abs(L12p - l3p));
                const int hi = std::min((two_J + 5) / 2, std::min(L12 + l3, L12p + l3p));
                for (int Lt = lo; Lt <= hi; ++Lt) {
                    // Atilde_store[alpha' * ... + alpha * ... + Ltot] = Atilde(alpha', alpha) (lines 444-465, 771-788)
                    double A = atilde(chn, ap, a, Lt, two_J, chn.two_T_3N[ap]);
                    if (A == 0.0) continue;
                    auto key = std::make_tuple(L12, l3, L12p, l3p, Lt);
                    auto it = tuple_id.find(key);
                    int id;
                    if (it == tuple_id.end()) {
                        id = (int)tuples.size();
                        tuple_id[key] = id;
                        tuples.push_back(key);
                    } else {
                        id = it->second;
                    }
                    G.pair_tuple.push_back(id);
                    G.pair_atilde.push_back(A);
                }
            }
            G.pair_ptr[(size_t)ap * Na + a + 1] = (int)G.pair_tuple.size();
        }
    }
    G.n_tuples = (int)tuples.size();
    G.term_ptr.assign(G.n_tuples + 1, 0);
    auto sgn = [](int m) { return (m < 0 && ((-m) & 1)) ? -1.0 : 1.0; };
    for (int t = 0; t < G.n_tuples; ++t) {
        int L12, l3, L12p, l3p, Lt;
        std::tie(L12, l3, L12p, l3p, Lt) = tuples[t];
        const double pref = 8.0 * M_PI * M_PI * std::sqrt((2.0 * L12 + 1) / (4 * M_PI));
        for (int M = -std::min(l3, Lt); M <= std::min(l3, Lt); ++M) {
            double cg1 = clebsch_gordan(2 * L12, 2 * l3, 2 * Lt, 0, 2 * M, 2 * M);
            if (cg1 == 0.0) continue;
            double ph = (M & 1) ? -1.0 : 1.0;  // (-1)^M
            for (int M12 = -L12p; M12 <= L12p; ++M12) {
                if (std::abs(M - M12) > l3p) continue;
                double cg2 = clebsch_gordan(2 * L12p, 2 * l3p, 2 * Lt, 2 * M12, 2 * M - 2 * M12, 2 * M);
                if (cg2 == 0.0) continue;
                const int Md = M - M12;
                G.term_coef.push_back(pref * cg1 * ph * cg2 * sgn(M) * sgn(M12) * sgn(Md));
                G.term_i1.push_back(l3 * (l3 + 1) / 2 + std::abs(M));
                G.term_i2.push_back(L12p * (L12p + 1) / 2 + std::abs(M12));
                G.term_i3.push_back(l3p * (l3p + 1) / 2 + std::abs(Md));
            }
        }
        G.term_ptr[t + 1] = (int)G.term_coef.size();
    }
    return G;
}

}  // namespace tt
