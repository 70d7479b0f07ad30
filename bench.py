"""Benchmark of the AGS/WPCD solve path on one MI355X.

Step = one pass over the batched workload: every 3N channel 2J = 1..17 of both parities (18
channels, N_alpha 27..63, dim 24300..56700) of the seeded synthetic NN potential at
Np = Nq = 30, at the 8 lab energies linspace(1, 200, 8) MeV, i.e. 144 channel-solves (one
channel-solve = every on-shell U element of one channel at one requested T_lab).  A step computes
the resolvent of every on-shell bin on the device and solves all channels through tt_solve_batch;
P123, V and the SWP basis are built once before timing and stay resident in HBM.

The headline value is that batched solve in mixed precision (FP32 P.V.G products, FP64 iterative
refinement).  Beside it, on the same 144 channel-solves: the batched solve in FP64, and three
plain paths: FP64 GMRES with one tt_solve_faddeev_equations call per channel and T_lab, FP64
GMRES with one call per channel holding all 8 T_lab (the reference driver's granularity), and
the reference's own method (Neumann series + Pade) at that granularity; with the relative error
of every mixed U element against the plain FP64 one (target <= 1e-8).  Further lines: the
single-channel J^Pi = 1/2+ workload, Np = Nq = 40 precision, the gradient cost and the CPU
baseline.  One JSON line on stdout.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "tic-tac_amd"))

import numpy as np  # noqa: E402

import tictac as tt  # noqa: E402

HBM_PEAK_GBS = 8000.0
MFMA_PEAK_TFLOPS = {"fp32": 157.3, "fp64": 78.6}   # dense MFMA peaks of MI355X
# streams a batched lockstep group is spread over: the library reads TT_LANES (1..16), else 6
LANES = int(os.environ["TT_LANES"]) if os.environ.get("TT_LANES", "").isdigit() and \
    1 <= int(os.environ["TT_LANES"]) <= 16 else 6
METRIC = "U-matrix channel-solves/s batched-mixed vs plain FP64 @|ΔU/U|≤1e-8; grad/fwd time"
SOLVE_TOL = 1e-12
BATCH_T = [float(t) for t in np.linspace(1.0, 200.0, 8)]
BATCH_WORKLOAD = ("synthetic NN (3 Yukawas + contacts + 3 separables, seed 7), all 3N channels 2J=1..17 both "
                  "parities (18 channels, N_alpha 27..63, dim 24300..56700), Np=Nq=30, J2max=3, T_lab = "
                  "linspace(1, 200, 8) MeV -> 144 channel-solves (6 distinct on-shell bins per channel)")
SINGLE_WORKLOAD = ("same synthetic NN, J^Pi=1/2+, Np=Nq=30, J2max=3, N_alpha=27, dim=24300, T_lab=10 MeV -> "
                   "2 on-shell bins x 2 deuteron channels")


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


def wp_params(N, two_J_max):
    return tt.run_params(two_J_3N_max=two_J_max, J_2N_max=3, Np_WP=N, Nq_WP=N, chebyshev_s=tt.chebyshev_scale(N))


class BatchWorkload:
    """The 18 channels at the 8 energies: one Channel per channel holding all energies (the batched
    solve) and one per (channel, T_lab) (the plain path), with one P123 per channel."""

    def __init__(self, ctx):
        self.prob = tt.Problem(wp_params(30, 17), tt.Potential("synthetic", tt.synthetic_params(7)))
        t0 = time.perf_counter()
        self.chans = [self.prob.channel(c, BATCH_T) for c in range(self.prob.n_channels)]
        self.Ps = [ch.build_p123(ctx) for ch in self.chans]
        ctx.synchronize()
        self.build_s = time.perf_counter() - t0
        self.plain = [[self.prob.channel(c, [T]) for T in BATCH_T] for c in range(self.prob.n_channels)]
        self.n_solves = len(self.chans) * len(BATCH_T)

    graphs = None   # the library's choice

    def solve(self, ctx, method):
        return self.prob.solve_batch(ctx, self.chans, self.Ps, method=method, tol=SOLVE_TOL, graphs=self.graphs)

    def solve_plain(self, ctx, method=tt.SOLVE_GMRES):
        return [[ch.solve(ctx, P, None, method=method, tol=SOLVE_TOL, device_resolvent=True)[0] for ch in row]
                for row, P in zip(self.plain, self.Ps)]

    def solve_per_channel(self, ctx, method=tt.SOLVE_GMRES):
        """The reference driver's call granularity (CPP/main.cpp:386-435): one
        tt_solve_faddeev_equations call per channel holding all 8 T_lab, channels one after
        another on one stream."""
        return [ch.solve(ctx, P, None, method=method, tol=SOLVE_TOL, device_resolvent=True)[0]
                for ch, P in zip(self.chans, self.Ps)]

    def channel_table(self):
        """Per channel: J^Pi, N_alpha, dim, overlapping bin cells, nonzeros of P123 and the
        number of right-hand sides r = n_d x on-shell bins of the batched solve."""
        return [{"JP": "%d/2%s" % (ch.two_J, "+" if ch.parity > 0 else "-"), "Nalpha": ch.Nalpha, "dim": ch.dim,
                 "ncell": int(P.stats.n_cells), "nnz": int(P.nnz), "r": int(ch.nd * ch.nE)}
                for ch, P in zip(self.chans, self.Ps)]

    def accuracy(self, U_batch, U_plain):
        """Relative error of every on-shell element of the batched U against the plain one of the same
        channel and bin: max |dU|/|U| over elements, and max |dU| / max |U| per channel."""
        elem = norm = 0.0
        for c, (Ub, row) in enumerate(zip(U_batch, U_plain)):
            qb = list(self.chans[c].q_idx)
            for ch, Up in zip(self.plain[c], row):
                for jp, q in enumerate(ch.q_idx):
                    a, b = Ub[:, :, qb.index(q)], Up[:, :, jp]
                    elem = max(elem, float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))))
                    norm = max(norm, float(np.max(np.abs(a - b)) / np.max(np.abs(b))))
        return elem, norm


def time_batches(ctx, wl, method, steps, warmup):
    for _ in range(warmup):
        out = wl.solve(ctx, method)
    ctx.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        out = wl.solve(ctx, method)
        stats.extend(st for (_, st, _) in out)
    ctx.synchronize()
    return time.perf_counter() - t0, [u for (u, _, _) in out], stats


def s8d_bytes(nnz, dim, r, fp32):
    """Algorithmic bytes of one product over P123 for r complex right-hand sides, as a CSR
    SpMM must move them: values and int32 indices of the nonzeros, the row pointers, the
    input panel read and the output panel written once (nnz (s_v + 4) + 4 (dim + 1) +
    2 r dim s_c, s_v / s_c the value and complex element sizes)."""
    sv, sc = (4.0, 8.0) if fp32 else (8.0, 16.0)
    return nnz * (sv + 4.0) + 4.0 * (dim + 1) + 2.0 * r * dim * sc


TRACE_FILE = os.path.join(ROOT, "profiles", "kernels_batch_synthetic_np30.md")


def traced_product_us(fp32):
    """Mean kernel duration of the cell products of one precision in the batch's rocprofv3 kernel
    trace summarised in profiles/ (the same bench.py workload; the mixed batch's table for FP32,
    the FP64 batch's for FP64), launch-weighted over the instances, or None."""
    head = "Per mixed batch" if fp32 else "Per FP64 batch"
    prefix = "| `k_cellmfma<%s" % ("float" if fp32 else "double")
    n = tot = 0.0
    try:
        lines = open(TRACE_FILE).read().split(head, 1)[1].split("\n\n", 2)[1].splitlines()
    except (OSError, IndexError):
        return None
    for line in lines:
        if not line.startswith("|"):
            break
        if line.startswith(prefix):
            cells = [c.strip() for c in line.strip("|").split("|")]
            n += float(cells[1])
            tot += float(cells[2]) * 1e3
    return tot / n if n else None


def product_roofline(stats, shapes, kernel, fp32=False, traffic=None, wall_s=None, lanes=1, traced_us=None):
    """HBM roofline of the product over P123 (the dominant kernel of the A-application).

    achieved = the algorithmic bytes per launch by s8d_bytes (per-channel nonzeros, dim and r)
    over the launch time from the HIP events that bracket each sampled launch on its stream
    during the timed steps.  achieved_cell_bytes uses the bytes the cell-block form itself must
    move instead (every block value once, structural zeros included, no index per value, the
    cell indexes and the two panels), which is fewer.  traffic = measured HBM bytes per launch
    (rocprofv3 FETCH_SIZE / WRITE_SIZE, copy-calibrated), or None.  mfma: the same launches
    against the MFMA roof of their precision, at 4 real flops per nonzero and complex
    right-hand side (a real value times a complex element: two fused multiply-adds).
    The sampled cell-product launches carry events that the runtime sets at the kernel's own
    start and end (hipExtLaunchKernel).  With lanes > 1 the channels' streams run concurrently,
    and the sampled launches (the first step of each cycle) start when every lane starts its
    cycle, so they share the device with the most other kernels: their mean is above the mean
    over all launches.  traced_us (traced_product_us) is the kernels' mean duration over all
    their launches in the rocprofv3 trace of the same workload, and frac_kernel_trace the
    fraction with it.  With
    wall_s, aggregate_achieved = the s8d bytes of every product launch of the timed steps
    (sampled or not) over their wall time: a lower bound on the products' HBM rate.
    stats[i] belongs to the channel of shapes[i % len(shapes)] (nnz, dim, r)."""
    ms = nb = b8 = fl = every = 0.0
    n = 0
    for i, s in enumerate(stats):
        nnz, dim, r = shapes[i % len(shapes)]
        if fp32:   # only the FP32 launches of the mixed solves (its FP64 residual applications are apart)
            k, t, cb, tot = s.n_spmm_fp32, s.ms_spmm_fp32, s.spmm_bytes_fp32, s.n_apply_fp32
        else:
            k, t, cb = s.n_spmm - s.n_spmm_fp32, s.ms_spmm - s.ms_spmm_fp32, s.spmm_bytes - s.spmm_bytes_fp32
            tot = s.n_apply - s.n_apply_fp32
        n += k
        ms += t
        nb += cb
        b8 += k * s8d_bytes(nnz, dim, r, fp32)
        fl += k * 4.0 * nnz * r
        every += tot * s8d_bytes(nnz, dim, r, fp32)
    sec = ms * 1e-3
    achieved = b8 / sec / 1e9 if ms > 0 else 0.0
    peak_tf = MFMA_PEAK_TFLOPS["fp32" if fp32 else "fp64"]
    out = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
           "kernel": kernel, "launches": n, "avg_us": round(1e3 * ms / max(n, 1), 2),
           "bytes_per_launch": round(b8 / max(n, 1)),
           "achieved_cell_bytes": round(nb / sec / 1e9, 1) if ms > 0 else 0.0,
           "cell_bytes_per_launch": round(nb / max(n, 1)),
           "mfma": {"flops_per_launch": round(fl / max(n, 1)),
                    "achieved_tflops": round(fl / sec / 1e12, 2) if ms > 0 else 0.0,
                    "peak_tflops": peak_tf,
                    "frac": round(fl / sec / 1e12 / peak_tf, 4) if ms > 0 else 0.0},
           "concurrent_streams": lanes}
    if traced_us:
        out["avg_us_kernel_trace"] = round(traced_us, 2)
        out["frac_kernel_trace"] = round(b8 / max(n, 1) / (traced_us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4)
    if wall_s:
        out["aggregate_achieved"] = round(every / wall_s / 1e9, 1)
    return out


def iteration_summary(stats, n_items):
    """Krylov steps per channel in the last timed batch: mean and max."""
    last = [s.iterations for s in stats[-n_items:]]
    return {"mean": round(sum(last) / max(len(last), 1), 1), "max": max(last) if last else 0}


# rocprofv3 summaries made by this bench.py's own options (each file gives its commands)
TRAFFIC_FILES = [os.path.join(ROOT, "profiles", f) for f in
                 ("pmc_cellmfma_synthetic_np30_batch.md", "pmc_cellmfma_synthetic_np30_single.md")]


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes summarised
    in profiles/ (scaled by a copy-kernel calibration), or None."""
    for path in TRAFFIC_FILES:
        try:
            for line in open(path):
                if line.startswith("traffic_bytes_per_launch " + kernel + " = "):
                    return int(line.rsplit("=", 1)[1])
        except OSError:
            pass
    return None


def single_channel(ctx, steps, warmup):
    """The J^Pi = 1/2+ channel at 10 MeV alone: mixed, FP64 and PCIe-inclusive channel-solves/s."""
    prob = tt.Problem(wp_params(30, 1), tt.Potential("synthetic", tt.synthetic_params(7)))
    ch = prob.channel(0, [10.0])
    P = ch.build_p123(ctx)
    ctx.synchronize()

    def run(method, host_G, n, w):
        def one():
            if host_G:
                return ch.solve(ctx, P, ch.resolvent(ctx), method=method, tol=SOLVE_TOL)
            return ch.solve(ctx, P, None, method=method, tol=SOLVE_TOL, device_resolvent=True)
        for _ in range(w):
            one()
        ctx.synchronize()
        st, t0 = [], time.perf_counter()
        for _ in range(n):
            U, s = one()
            st.append(s)
        ctx.synchronize()
        return n / (time.perf_counter() - t0), U, st

    v_mix, U_mix, st_mix = run(tt.SOLVE_MIXED, False, steps, warmup)
    v_64, U_64, st_64 = run(tt.SOLVE_GMRES, False, steps, warmup)
    v_pcie, _, _ = run(tt.SOLVE_MIXED, True, steps, 1)
    shape = [(int(P.nnz), ch.dim, ch.nd * ch.nE)]
    return {"workload": SINGLE_WORKLOAD, "unit": "channel-solves/s", "nnz_P123": int(P.nnz),
            "p123_build_ms": round(P.stats.ms_total, 2), "mixed": round(v_mix, 2), "fp64": round(v_64, 2),
            "mixed_pcie_inclusive": round(v_pcie, 2),
            "mixed_vs_fp64_max_rel": float(np.max(np.abs(U_mix - U_64)) / np.max(np.abs(U_64))),
            "roofline_fp32": product_roofline(st_mix, shape, "k_cellmfma<float, 2, 1, UN>", True,
                                              measured_traffic("k_cellmfma<float, 2, 1, UN>")),
            "roofline_fp64": product_roofline(st_64, shape, "k_cellmfma<double, 2, 1, UN>", False,
                                              measured_traffic("k_cellmfma<double, 2, 1, UN>"))}


def dominant_alone(ctx, wl, reps=4):
    """The batch's dominant product shape measured without the other lanes: the first N_alpha = 63
    channel (7/2+, dim 56,700) with all 8 T_lab (r = 18 systems, R2 = 36) solved alone on one
    stream, mixed and FP64, reps times after a warm-up; the product's roofline over those
    launches (the same instances the batch runs, k_cellmfma<*, 4, 3, UN>)."""
    i = next(k for k, ch in enumerate(wl.chans) if ch.Nalpha == 63)
    ch, P = wl.chans[i], wl.Ps[i]
    row = wl.channel_table()[i]
    shape = [(row["nnz"], row["dim"], row["r"])]
    res = {"channel": row["JP"], "dim": row["dim"], "r": row["r"]}
    for method, fp32, key in ((tt.SOLVE_MIXED, True, "roofline_fp32"), (tt.SOLVE_GMRES, False, "roofline_fp64")):
        ch.solve(ctx, P, None, method=method, tol=SOLVE_TOL, device_resolvent=True)
        st = [ch.solve(ctx, P, None, method=method, tol=SOLVE_TOL, device_resolvent=True)[1] for _ in range(reps)]
        k = "k_cellmfma<%s, 4, 3, UN>" % ("float" if fp32 else "double")
        res[key] = product_roofline(st, shape, k, fp32, measured_traffic(k))
    return res


def dense_leg(ctx):
    """The dense path (solve_dense=true) on the single-channel workload: per on-shell energy, L = 1 -
    A G_j (dim 24,300, complex FP64) formed and LU-factorised with partial pivoting on the device;
    seconds per energy and the LU's rate at (8/3) dim^3 real flops per factorisation against the
    FP64 MFMA peak, with the difference to the GMRES solve."""
    prob = tt.Problem(wp_params(30, 1), tt.Potential("synthetic", tt.synthetic_params(7)))
    ch = prob.channel(0, [10.0])
    P = ch.build_p123(ctx)
    G = ch.resolvent(ctx)
    U64, _ = ch.solve(ctx, P, G, method=tt.SOLVE_GMRES, tol=SOLVE_TOL)
    ctx.synchronize()
    t0 = time.perf_counter()
    Ud, _ = ch.solve(ctx, P, G, method=tt.SOLVE_DENSE)
    dt = (time.perf_counter() - t0) / ch.nE
    flops = 8.0 / 3.0 * float(ch.dim) ** 3
    return {"workload": SINGLE_WORKLOAD + "; dense LU of 1 - A G_j per on-shell bin", "dim": ch.dim,
            "seconds_per_energy": round(dt, 3), "lu_tflops": round(flops / dt / 1e12, 2),
            "fp64_mfma_peak_tflops": MFMA_PEAK_TFLOPS["fp64"],
            "frac_of_peak_incl_forming_L": round(flops / dt / 1e12 / MFMA_PEAK_TFLOPS["fp64"], 4),
            "max_rel_vs_gmres": float(np.max(np.abs(Ud - U64)) / np.max(np.abs(U64)))}


GRAD_WORKLOAD = ("same synthetic NN, channels 2J=1,3 both parities (N_alpha 27, 47; dim 24300, 42300), Np=Nq=30, "
                 "T_lab 5, 50, 150 MeV, 12 strengths theta (3 Yukawas, 6 contacts, 3 separables); FP64")


def gradient_ratio(ctx, reps=3):
    """Cost of U(theta) with dU/dtheta (tt_solve_gradient: forward row solve + adjoint column
    solve + assembly) over the forward alone (same call without the derivative: V, SWP
    states, G and the row solve), summed over the 4 channels, best of reps each."""
    T = [5.0, 50.0, 150.0]
    prob = tt.Problem(wp_params(30, 3), tt.Potential("synthetic", tt.synthetic_params(7)))
    basis = prob.basis_matrices(12)
    th = tt.synthetic_params()[:12]
    t_f = t_g = 0.0
    for c in range(prob.n_channels):
        P = prob.channel(c, T).build_p123(ctx)
        prob.gradient(ctx, c, P, th, T, basis=basis, grad=True)   # warm-up: plans and Krylov workspaces
        bf = bg = None
        for rep in range(reps):
            theta = th * (1.0 + 1e-3 * (rep + 1))                  # a new theta each time
            t0 = time.perf_counter()
            prob.gradient(ctx, c, P, theta, T, basis=basis, grad=False)
            t1 = time.perf_counter()
            prob.gradient(ctx, c, P, theta, T, basis=basis, grad=True)
            t2 = time.perf_counter()
            bf = (t1 - t0) if bf is None else min(bf, t1 - t0)
            bg = (t2 - t1) if bg is None else min(bg, t2 - t1)
        t_f += bf
        t_g += bg
    return {"workload": GRAD_WORKLOAD, "grad_over_fwd": round(t_g / t_f, 3),
            "adjoint_over_fwd": round((t_g - t_f) / t_f, 3), "ms_forward": round(1e3 * t_f, 2),
            "ms_forward_plus_gradient": round(1e3 * t_g, 2)}


N40_WORKLOAD = ("same synthetic NN at Np=Nq=40: channels J^Pi=1/2+ (N_alpha 27, dim 43200) and 7/2+ (N_alpha 63, "
                "dim 100800), T_lab = 10 MeV (2 on-shell bins)")


def precision_n40(ctx, reps=3):
    """Mixed precision at Np=Nq=40: channel-solves/s (mixed and FP64, best of reps, G on the
    device) and the relative error of every on-shell U element of the mixed solve against the
    converged FP64 solve, per channel and energy (target <= 1e-8)."""
    prob = tt.Problem(wp_params(40, 7), tt.Potential("synthetic", tt.synthetic_params(7)))
    out = []
    for c in range(prob.n_channels):
        ch = prob.channel(c, [10.0])
        if not ((ch.two_J == 1 and ch.parity == 1) or (ch.two_J == 7 and ch.parity == 1)):
            continue
        P = ch.build_p123(ctx)
        res = {"channel": "%d/2%s" % (ch.two_J, "+" if ch.parity > 0 else "-"), "dim": ch.dim,
               "p123_build_ms": round(P.stats.ms_total, 1)}
        U = {}
        for method, name in ((tt.SOLVE_MIXED, "mixed"), (tt.SOLVE_GMRES, "fp64")):
            ch.solve(ctx, P, None, method=method, tol=SOLVE_TOL, device_resolvent=True)
            best = None
            for _ in range(reps):
                t0 = time.perf_counter()
                U[name], st = ch.solve(ctx, P, None, method=method, tol=SOLVE_TOL, device_resolvent=True)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            res[name] = round(1.0 / best, 2)
        err = np.abs(U["mixed"] - U["fp64"]) / np.abs(U["fp64"])
        res["mixed_rel_err_per_energy"] = [float(err[..., j].max()) for j in range(ch.nE)]
        out.append(res)
        del P
    return {"workload": N40_WORKLOAD, "unit": "channel-solves/s", "channels": out}


def cpu_baseline(wl, ctx, U_plain, budget_s=30.0):
    """The oracle's C++/OpenMP restatement (oracle/cpu_solve.cpp) on the host's cores, on the same
    inputs as the GPU (its P123, V and SWP basis):

    * value: matrix-free FP64 GMRES (CSC gather, all rows of a channel sharing each application)
      to rtol 1e-12, one call per channel with all 8 T_lab, over the whole workload (18 channels,
      144 channel-solves; about 15 s on 16 host threads), with the largest difference of its U to
      the GPU's plain FP64 solves.
    * reference_semantics: the reference's own algorithm on the 1/2+ channel at the 8 T_lab: every
      Neumann term regenerates the columns of A (CPP/solve_faddeev.cpp:1019-1091), timed for up
      to budget_s / 2 and extrapolated to the Neumann terms the same Pade solve takes on the GPU.
    U_plain: the GPU's plain FP64 U per channel (all 8 T_lab), or None."""
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle import cpu
    from helpers import oracle_pw, oracle_swp, oracle_V, flags
    prob = wl.prob
    tf, ib = flags(prob)
    Vu, Vc = oracle_V(prob)
    swp = oracle_swp(prob)
    threads = cpu.lib().ttcpu_threads()
    secs, diff = {}, 0.0
    op0 = None
    for i, (ch, P) in enumerate(zip(wl.chans, wl.Ps)):
        val, row, ptr = P.to_host()
        op = cpu.Problem(oracle_pw(ch), swp, Vu, Vc, val, row, ptr, tf, ib)
        del val, row, ptr
        G = ch.resolvent(ctx)
        t0 = time.perf_counter()
        U, rc, its, res = op.gmres_U(G, ch.deut, ch.q_idx, tol=SOLVE_TOL)
        secs[wl.channel_table()[i]["JP"]] = round(time.perf_counter() - t0, 3)
        if U_plain is not None:
            diff = max(diff, float(np.max(np.abs(U - U_plain[i])) / np.max(np.abs(U_plain[i]))))
        if i == 0:
            op0 = op
    total = sum(secs.values())
    out = {"value": round(wl.n_solves / total, 3), "unit": "channel-solves/s", "cores": threads, "kind": "port",
           "sample": "the whole workload: oracle/cpu_solve.cpp (C++/OpenMP, FP64) matrix-free GMRES to rtol 1e-12, "
                     "one call per channel with the 8 T_lab, 18 channels = %d channel-solves in %.1f s on %d threads"
                     % (wl.n_solves, total, threads),
           "seconds_per_channel": secs,
           "max_rel_vs_gpu_plain_fp64": diff,
           "host": host_info()}
    # the reference's own algorithm (columns of A regenerated per Neumann term), capped
    ch = wl.chans[0]
    G = ch.resolvent(ctx)
    n_need = ch.solve(ctx, wl.Ps[0], G, method=tt.SOLVE_PADE)[1].iterations
    a, n_done, spt = op0.neumann_terms(G, ch.deut, ch.q_idx, n_need, cap_s=budget_s / 2)
    out["reference_semantics"] = {
        "channel": "1/2+ (N_alpha 27, dim 24300), 8 T_lab -> 12 on-shell rows",
        "seconds_per_neumann_term": round(spt, 3), "terms_timed": n_done, "terms_needed": n_need,
        "extrapolated_s": round(spt * n_need, 2),
        "channel_solves_per_s": round(len(BATCH_T) / (spt * n_need), 4),
        "note": "every Neumann term regenerates all dim columns of A and multiplies the on-shell rows into "
                "them, as CPP/solve_faddeev.cpp:1019-1091; terms_needed is the Pade order the GPU's "
                "reference-semantics solve of this channel reached"}
    return out


def host_info():
    """The host the CPU baseline ran on: CPU model, the CPUs this process may use, and the
    OpenMP thread setting the restatement runs with."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"cpu_model": model, "nproc": nproc, "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


WALK_WORKLOAD = ("same 18 channels x 8 energies for 8 parameter sets (the 12 strengths scaled by 0.99 .. 1.02): "
                 "per set V_WP, SWP states, on-shell bins and G rebuilt, then one lockstep batch (tt_parameter_walk)")


def parameter_walk(ctx, wl, n_sets=8):
    """The reference's parameter walk on the batched workload: channel-solves/s over n_sets
    parameter sets, the host rebuild of each set included (overlapped with the previous set's
    solve), mixed precision to the same tolerance."""
    th = tt.synthetic_params(7)
    sets = []
    for i in range(n_sets):
        p = th.copy()
        p[:12] *= 0.99 + 0.03 * i / max(n_sets - 1, 1)
        sets.append(p)
    chans = list(range(len(wl.chans)))
    wl.prob.parameter_walk(ctx, chans, wl.Ps, sets[:2], BATCH_T, method=tt.SOLVE_MIXED, tol=SOLVE_TOL)  # warm-up
    t0 = time.perf_counter()
    U, status, code = wl.prob.parameter_walk(ctx, chans, wl.Ps, sets, BATCH_T, method=tt.SOLVE_MIXED, tol=SOLVE_TOL)
    dt = time.perf_counter() - t0
    ok = int((status == 0).sum())
    return {"workload": WALK_WORKLOAD, "unit": "channel-solves/s", "value": round(ok * len(BATCH_T) / dt, 1),
            "ms_per_set": round(1e3 * dt / n_sets, 2), "items_failed": int(status.size - ok)}


def copy_calibration(mib=256, reps=3):
    """A plain device-to-device copy of known size (the HIP runtime's copy kernel behind
    hipMemcpy, mib MiB each way, reps launches): the reference point that scales rocprofv3's
    FETCH_SIZE / WRITE_SIZE counters to true HBM bytes in the traffic summaries."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), C.c_size_t(n)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, C.c_size_t(n)) != 0:
            raise RuntimeError("hipMemset failed")
        for _ in range(reps):
            if hip.hipMemcpy(dst, src, C.c_size_t(n), 3) != 0:   # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    return {"bytes_each_way": n, "launches": reps}


LEGS = ("plain", "walk", "single", "n40", "grad", "dense", "cpu", "copy")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed batches (144 channel-solves each)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--legs", default=",".join(x for x in LEGS if x != "copy"),
                    help="comma-separated extra lines beside the headline batch: " + ", ".join(LEGS) + " (or none)")
    ap.add_argument("--graphs", choices=("auto", "on", "off"), default="auto",
                    help="replayed HIP graphs of the Arnoldi steps in the batched solves: the library's choice "
                    "(auto: graphs for groups of 4 or more channels), or forced on / off")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each channel-solve is "
              "independent; the design says how the work would shard)", file=sys.stderr)
        sys.exit(2)
    legs = set() if args.legs == "none" else set(x for x in args.legs.split(",") if x)
    bad = legs - set(LEGS)
    if bad:
        ap.error("unknown legs: " + ", ".join(sorted(bad)))

    ctx = tt.Context(0)
    log("building P123 of the 18 channels")
    wl = BatchWorkload(ctx)
    wl.graphs = {"auto": None, "on": True, "off": False}[args.graphs]
    log("P123 built in %.1f s; timing the batched mixed solve" % wl.build_s)
    t_mix, U_mix, st_mix = time_batches(ctx, wl, tt.SOLVE_MIXED, args.steps, args.warmup)
    log("batched FP64")
    n64 = max(2, args.steps // 2)
    t_64, U_64, st_64 = time_batches(ctx, wl, tt.SOLVE_GMRES, n64, 1)
    value = wl.n_solves * args.steps / t_mix
    fp64 = wl.n_solves * n64 / t_64
    k32 = "k_cellmfma<float, MT, NT, UN>"
    table = wl.channel_table()
    shapes = [(c["nnz"], c["dim"], c["r"]) for c in table]
    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "channel-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mix / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "complex fp32 products + fp64 refinement",
        "data": "synthetic",
        "config": {"workload": BATCH_WORKLOAD, "channel_solves_per_step": wl.n_solves,
                   "nnz_P123_total": int(sum(P.nnz for P in wl.Ps)),
                   "p123_build_s_all_channels": round(wl.build_s, 2), "solve_tol": SOLVE_TOL,
                   "step_graphs": args.graphs,
                   "channels": table},
        "batched_fp64": round(fp64, 2),
        "mixed_over_batched_fp64": round(value / fp64, 3),
        "plain_fp64": None,
        "mixed_over_plain_fp64": None,
        "grad_over_fwd": None,
        "roofline": product_roofline(st_mix, shapes, k32, True, measured_traffic(k32), t_mix, LANES,
                                     traced_product_us(True)),
        "roofline_fp64": product_roofline(st_64, shapes, "k_cellmfma<double, MT, NT, UN>", False,
                                          measured_traffic("k_cellmfma<double, MT, NT, UN>"), t_64, LANES,
                                          traced_product_us(False)),
        "krylov_steps_per_channel": {"mixed": iteration_summary(st_mix, len(wl.chans)),
                                     "fp64": iteration_summary(st_64, len(wl.chans))},
        "cell_autotune": tt.cell_tuning(),
    }
    log("dominant product shape alone")
    out["roofline_alone"] = dominant_alone(ctx, wl)
    U_plain = U_plain_ch = None
    if "plain" in legs or "cpu" in legs:
        log("plain FP64: one call per channel and T_lab")
        wl.solve_plain(ctx)   # warm-up
        ctx.synchronize()
        t0 = time.perf_counter()
        U_plain = wl.solve_plain(ctx)
        ctx.synchronize()
        t_plain = time.perf_counter() - t0
        elem, norm = wl.accuracy(U_mix, U_plain)
        elem64, norm64 = wl.accuracy(U_64, U_plain)
        out["plain_fp64"] = round(wl.n_solves / t_plain, 2)
        out["mixed_over_plain_fp64"] = round(value / out["plain_fp64"], 3)
        out["mixed_vs_plain_fp64_max_rel_elem"] = elem
        out["mixed_vs_plain_fp64_max_rel_norm"] = norm
        out["batched_fp64_vs_plain_fp64_max_rel_elem"] = elem64
        out["accuracy_target_met"] = bool(elem <= 1e-8)
        # the reference driver's granularity: one call per channel with all 8 T_lab
        log("plain FP64 and Pade: one call per channel with all energies")
        for method, key in ((tt.SOLVE_GMRES, "plain_fp64_per_channel"), (tt.SOLVE_PADE, "plain_pade_per_channel")):
            wl.solve_per_channel(ctx, method)   # warm-up
            ctx.synchronize()
            t0 = time.perf_counter()
            Uc = wl.solve_per_channel(ctx, method)
            if method == tt.SOLVE_GMRES:
                U_plain_ch = Uc
            ctx.synchronize()
            rate = wl.n_solves / (time.perf_counter() - t0)
            ref = U_64 if method == tt.SOLVE_PADE else U_mix
            err = max(float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))) for a, b in zip(ref, Uc))
            out[key] = round(rate, 2)
            out["mixed_over_" + key] = round(value / rate, 3)
            if method == tt.SOLVE_PADE:
                out["plain_pade_per_channel_max_rel_elem_vs_fp64"] = err
            else:
                out["mixed_vs_plain_fp64_per_channel_max_rel_elem"] = err
    if "walk" in legs:
        log("parameter walk")
        out["parameter_walk"] = parameter_walk(ctx, wl)
    if "single" in legs:
        log("single channel 1/2+")
        out["single_channel"] = single_channel(ctx, 20, 3)
    if "n40" in legs:
        log("Np=Nq=40")
        out["precision_n40"] = precision_n40(ctx)
    if "grad" in legs:
        log("gradient")
        out["gradient"] = gradient_ratio(ctx)
        out["grad_over_fwd"] = out["gradient"]["grad_over_fwd"]
    if "dense" in legs:
        log("dense LU")
        out["dense"] = dense_leg(ctx)
    if "copy" in legs:
        log("copy calibration")
        out["copy_calibration"] = copy_calibration()
    if "cpu" in legs:
        log("CPU baseline")
        out["cpu_baseline"] = cpu_baseline(wl, ctx, U_plain_ch)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
