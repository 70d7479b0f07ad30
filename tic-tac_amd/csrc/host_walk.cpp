// Parameter walk on the GPU: the PSI loop of CPP/main.cpp:126-136,254-466 for many 3N channels
// at once.  For every parameter set the potential's parameters are set, V_WP, the SWP states,
// the on-shell bins and G are rebuilt and every listed channel's on-shell U is solved, all
// channels of a set in one lockstep batch (tt_solve_batch).  The host work of set s + 1 (V_WP,
// SWP eigenproblems, on-shell bins) runs on a second thread while the device solves set s.
// Host orchestration over the public C-ABI only.
#include <cstring>
#include <exception>
#include <string>
#include <thread>
#include <vector>

#include "tt_internal.h"

namespace {

// Everything one parameter set needs before its solve, built on the host.
struct SetData {
    std::vector<double> V_unco, V_coup, e_unco, e_coup, C_unco, C_coup;
    tt_swp_states swp;
    std::vector<tt_channel_os> os;   // per listed channel
    std::vector<char> os_ok;
    tt_status status = TT_OK;
    std::string msg;

    SetData() { std::memset(&swp, 0, sizeof(swp)); }
    ~SetData() { clear_os(); }
    SetData(const SetData &) = delete;
    SetData &operator=(const SetData &) = delete;
    void clear_os() {
        for (size_t i = 0; i < os.size(); ++i)
            if (os_ok[i]) tt_channel_os_free(&os[i]);
        os.clear();
        os_ok.clear();
    }
};

struct WalkInputs {
    const tt_run_params *rp;
    const tt_fwp_states *fwp;
    const tt_pw_states *pw_full;
    tt_potential *pot;
    const double *params;
    int n_params;
    const std::vector<tt_pw_states> *chn;
    const double *T_lab;
    int n_T_lab;
};

// Set `set` of the walk: update_parameters (potential_model.h:16), V_WP, SWP states and the
// on-shell bins of every channel.  Failures are recorded in d.status (the set is skipped).
void prepare_set(const WalkInputs &in, int set, SetData &d) {
    d.clear_os();
    d.status = TT_OK;
    d.msg.clear();
    auto check = [&](tt_status st) {
        if (st != TT_OK && d.status == TT_OK) {
            d.status = st;
            d.msg = tt_last_error();
        }
        return st == TT_OK;
    };
    int nu = 0, nc = 0;
    tt_num_2N_states(in.rp, &nu, &nc);
    const size_t Np = (size_t)in.fwp->Np_WP;
    d.V_unco.assign(std::max<size_t>(nu * Np * Np, 1), 0.0);
    d.V_coup.assign(std::max<size_t>(nc * 4 * Np * Np, 1), 0.0);
    d.e_unco.assign(std::max<size_t>(nu * (Np + 1), 1), 0.0);
    d.e_coup.assign(std::max<size_t>(nc * 2 * (Np + 1), 1), 0.0);
    d.C_unco.assign(d.V_unco.size(), 0.0);
    d.C_coup.assign(d.V_coup.size(), 0.0);
    std::memset(&d.swp, 0, sizeof(d.swp));
    d.swp.e_SWP_unco = d.e_unco.data();
    d.swp.e_SWP_coup = d.e_coup.data();
    d.swp.C_SWP_unco = d.C_unco.data();
    d.swp.C_SWP_coup = d.C_coup.data();
    if (!check(tt_potential_update_parameters(in.pot, in.params + (size_t)set * in.n_params, in.n_params))) return;
    if (!check(tt_calculate_potential_matrices(in.rp, in.fwp, in.pw_full, in.pot, d.V_unco.data(), d.V_coup.data())))
        return;
    if (!check(tt_make_swp_states(in.rp, in.fwp, in.pw_full, d.V_unco.data(), d.V_coup.data(), &d.swp))) return;
    const size_t nch = in.chn->size();
    d.os.resize(nch);
    d.os_ok.assign(nch, 0);
    for (size_t c = 0; c < nch; ++c) {
        std::memset(&d.os[c], 0, sizeof(tt_channel_os));
        if (!check(tt_find_on_shell(&d.swp, &(*in.chn)[c], in.T_lab, in.n_T_lab, &d.os[c]))) return;
        d.os_ok[c] = 1;
    }
}

}  // namespace

using namespace tt;

extern "C" tt_status tt_parameter_walk(tt_ctx *ctx, const tt_run_params *rp, const tt_fwp_states *fwp,
                                       const tt_pw_states *pw_full, tt_potential *pot, const double *params,
                                       int n_params, int n_sets, const int *channels, const tt_csc *const *P123,
                                       int n_chn, const double *T_lab, int n_T_lab, double *U, int capacity,
                                       int *n_out, tt_status *status, const tt_solve_opts *opts) {
    TT_GUARD_BEGIN
    TT_REQUIRE(ctx && rp && fwp && pw_full && pot && params && channels && P123 && T_lab && U && n_out && status,
               "null argument");
    TT_REQUIRE(n_params > 0 && n_sets >= 0 && n_chn > 0 && n_T_lab > 0 && capacity > 0, "empty walk");
    std::vector<tt_pw_states> chn(n_chn);
    for (int c = 0; c < n_chn; ++c) {
        TT_REQUIRE(channels[c] >= 0 && channels[c] < pw_full->N_chn_3N, "channel index out of range");
        TT_REQUIRE(P123[c], "null P123");
        const tt_status st = tt_pw_channel(pw_full, channels[c], &chn[c]);
        if (st != TT_OK) fail(st, tt_last_error());
    }
    tt_solve_opts o;
    tt_solve_opts_default(&o);
    if (opts) o = *opts;
    const WalkInputs in{rp, fwp, pw_full, pot, params, n_params, &chn, T_lab, n_T_lab};
    SetData buf[2];
    const size_t items_total = (size_t)n_sets * n_chn;
    for (size_t i = 0; i < items_total; ++i) {
        n_out[i] = 0;
        status[i] = TT_OK;
    }
    tt_status first = TT_OK;
    std::string first_msg;
    auto record = [&](size_t i, tt_status st, const std::string &msg) {
        status[i] = st;
        if (st != TT_OK && first == TT_OK) {
            first = st;
            first_msg = "parameter set " + std::to_string(i / n_chn) + ", channel item " +
                        std::to_string(i % n_chn) + ": " + msg;
        }
    };
    if (n_sets > 0) prepare_set(in, 0, buf[0]);
    for (int s = 0; s < n_sets; ++s) {
        SetData &d = buf[s & 1];
        // the next set's host work on a second thread (it alone uses the potential meanwhile)
        std::thread next;
        std::exception_ptr next_err;
        if (s + 1 < n_sets)
            next = std::thread([&, s]() {
                try {
                    prepare_set(in, s + 1, buf[(s + 1) & 1]);
                } catch (...) {
                    next_err = std::current_exception();
                }
            });
        struct Join {   // on every exit path
            std::thread *t;
            ~Join() {
                if (t->joinable()) t->join();
            }
        } join{&next};
        const size_t i0 = (size_t)s * n_chn;
        if (d.status != TT_OK) {
            for (int c = 0; c < n_chn; ++c) record(i0 + c, d.status, d.msg);
        } else {
            std::vector<tt_batch_item> items(n_chn);
            std::vector<int> need(n_chn);
            for (int c = 0; c < n_chn; ++c) {
                const tt_channel_os &os = d.os[c];
                need[c] = os.num_deuteron_states * os.num_deuteron_states * os.num_T_lab;
                items[c].chn = &chn[c];
                items[c].os = &os;
                items[c].P123 = P123[c];
                items[c].U = U + 2 * (i0 + c) * (size_t)capacity;
                items[c].stats = nullptr;
                items[c].status = TT_OK;
            }
            std::vector<int> run;   // items whose U fits
            std::vector<tt_batch_item> fit;
            for (int c = 0; c < n_chn; ++c) {
                if (need[c] > capacity) {
                    record(i0 + c, TT_ERR_ARG, "U capacity is smaller than nd*nd*nE");
                    continue;
                }
                run.push_back(c);
                fit.push_back(items[c]);
            }
            (void)tt_solve_batch(ctx, rp, d.V_unco.data(), d.V_coup.data(), &d.swp, fit.data(), (int)fit.size(), &o);
            for (size_t k = 0; k < run.size(); ++k) {
                const int c = run[k];
                n_out[i0 + c] = need[c];
                record(i0 + c, fit[k].status, fit[k].status == TT_OK ? std::string() : std::string(tt_last_error()));
            }
        }
        if (next.joinable()) next.join();
        if (next_err) std::rethrow_exception(next_err);
    }
    if (first != TT_OK) fail(first, first_msg);
    TT_GUARD_END
}
