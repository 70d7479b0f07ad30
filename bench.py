"""Benchmark of the AGS/WPCD solve path on one MI355X.

Step = one channel-solve: the resolvent G of the on-shell energies plus the
on-shell U-matrix solve, for the 3N channel J^Pi = 1/2+ of the seeded synthetic
NN potential at Np = Nq = 30 (N_alpha = 27, dim = 24300), T_lab = 10 MeV
(2 on-shell q bins x 2 deuteron channels = 4 coupled systems).  P123, V and the
SWP basis are built once before timing and stay resident in HBM.

The headline value is the mixed-precision solve (FP32 P.V.G products with FP64
iterative refinement to |dU/U| <= 1e-8); the plain FP64 GMRES path is timed on
the same inputs and reported beside it, with the accuracy of the mixed result
against it.  One JSON line on stdout.
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
WORKLOAD = ("synthetic NN (3 Yukawas + contacts, seed 7), J^Pi=1/2+, Np=Nq=30, J2max=3, N_alpha=27, "
            "dim=24300, T_lab=10 MeV -> 2 on-shell bins x 2 deuteron channels")


def build(N=30, seed=7, T_lab=(10.0,)):
    rp = tt.run_params(two_J_3N_max=1, J_2N_max=3, Np_WP=N, Nq_WP=N, chebyshev_s=tt.chebyshev_scale(N))
    prob = tt.Problem(rp, tt.Potential("synthetic", tt.synthetic_params(seed)))
    ch = prob.channel(0, list(T_lab))
    return prob, ch


def timed_steps(ctx, ch, P, method, steps, warmup, host_G=False):
    """One step = resolvent of the on-shell energies + the on-shell U solve.  By default G is
    computed on the device inside the solve (inputs resident in HBM); host_G routes it through
    host memory as the reference's G_array does (tt_calculate_resolvent, then the solve uploads
    it), i.e. the PCIe-inclusive rate."""
    def one():
        if host_G:
            return ch.solve(ctx, P, ch.resolvent(ctx), method=method)
        return ch.solve(ctx, P, None, method=method, device_resolvent=True)
    stats = []
    U = None
    for _ in range(warmup):
        U, st = one()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        U, st = one()
        stats.append(st)
    ctx.synchronize()
    return time.perf_counter() - t0, U, stats


BATCH_WORKLOAD = ("same synthetic NN, all 3N channels 2J=1..17 both parities (18 channels, N_alpha 27..63, "
                  "dim 24300..56700), Np=Nq=30, T_lab = linspace(1, 200, 8) MeV -> 6 on-shell bins per channel, "
                  "each channel one block Krylov system (tt_solve_batch)")


def batched(ctx, reps=2):
    """All channels x 8 energies through tt_solve_batch (G on the device, P123 resident):
    channel-solves/s = 18 * 8 / wall time of one batch, mixed and FP64, best of reps."""
    T = list(np.linspace(1.0, 200.0, 8))
    rp = tt.run_params(two_J_3N_max=17, J_2N_max=3, Np_WP=30, Nq_WP=30, chebyshev_s=tt.chebyshev_scale(30))
    prob = tt.Problem(rp, tt.Potential("synthetic", tt.synthetic_params(7)))
    t0 = time.perf_counter()
    chans = [prob.channel(c, T) for c in range(prob.n_channels)]
    Ps = [ch.build_p123(ctx) for ch in chans]
    ctx.synchronize()
    t_build = time.perf_counter() - t0
    res = {}
    U = {}
    for method, name in ((tt.SOLVE_MIXED, "mixed"), (tt.SOLVE_GMRES, "fp64")):
        prob.solve_batch(ctx, chans, Ps, method=method)   # warm-up: plans and Krylov workspaces
        best = None
        for _ in range(reps):
            ctx.synchronize()
            t1 = time.perf_counter()
            out = prob.solve_batch(ctx, chans, Ps, method=method)
            ctx.synchronize()
            dt = time.perf_counter() - t1
            best = dt if best is None else min(best, dt)
        res[name] = round(len(chans) * len(T) / best, 2)
        U[name] = [u for (u, _, _) in out]
    rel = max(float(np.abs(a - b).max() / np.abs(b).max()) for a, b in zip(U["mixed"], U["fp64"]))
    return {"workload": BATCH_WORKLOAD, "unit": "channel-solves/s", "channel_solves": len(chans) * len(T),
            "mixed": res["mixed"], "fp64": res["fp64"], "mixed_vs_fp64_max_rel": rel,
            "p123_build_s_all_channels": round(t_build, 2)}


GRAD_WORKLOAD = ("same synthetic NN, channels 2J=1,3 both parities (N_alpha 27, 47; dim 24300, 42300), Np=Nq=30, "
                 "T_lab 5, 50, 150 MeV, 12 strengths theta (3 Yukawas, 6 contacts, 3 separables); FP64")


def gradient_ratio(ctx, reps=3):
    """Cost of U(theta) with dU/dtheta (tt_solve_gradient: forward row solve + adjoint column
    solve + assembly) over the forward alone (same call without the derivative: V, SWP
    states, G and the row solve), summed over the 4 channels, best of reps each."""
    T = [5.0, 50.0, 150.0]
    rp = tt.run_params(two_J_3N_max=3, J_2N_max=3, Np_WP=30, Nq_WP=30, chebyshev_s=tt.chebyshev_scale(30))
    prob = tt.Problem(rp, tt.Potential("synthetic", tt.synthetic_params(7)))
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


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_cellmfma_synthetic_np30.md")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` at this workload from the rocprofv3 FETCH_SIZE /
    WRITE_SIZE passes summarised in profiles/ (scaled by a copy-kernel calibration), or None."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch " + kernel + " = "):
                return int(line.rsplit("=", 1)[1])
    except OSError:
        pass
    return None


N40_WORKLOAD = ("same synthetic NN at Np=Nq=40: channels J^Pi=1/2+ (N_alpha 27, dim 43200) and 7/2+ (N_alpha 63, "
                "dim 100800), T_lab = 10 MeV (2 on-shell bins)")


def precision_n40(ctx, reps=3):
    """Mixed precision at Np=Nq=40: channel-solves/s (mixed and FP64, best of reps, G on the
    device) and the relative error of every on-shell U element of the mixed solve against the
    converged FP64 solve, per channel and energy (target <= 1e-8)."""
    N = 40
    rp = tt.run_params(two_J_3N_max=7, J_2N_max=3, Np_WP=N, Nq_WP=N, chebyshev_s=tt.chebyshev_scale(N))
    prob = tt.Problem(rp, tt.Potential("synthetic", tt.synthetic_params(7)))
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
            ch.solve(ctx, P, None, method=method, device_resolvent=True)
            best = None
            for _ in range(reps):
                t0 = time.perf_counter()
                U[name], st = ch.solve(ctx, P, None, method=method, device_resolvent=True)
                dt = time.perf_counter() - t0
                best = dt if best is None else min(best, dt)
            res[name] = round(1.0 / best, 2)
        err = np.abs(U["mixed"] - U["fp64"]) / np.abs(U["fp64"])
        res["mixed_rel_err_per_energy"] = [float(err[..., j].max()) for j in range(ch.nE)]
        out.append(res)
        del P
    return {"workload": N40_WORKLOAD, "unit": "channel-solves/s", "channels": out}


def product_roofline(stats, kernel, fp32=False):
    """HBM roofline of the P123 cell-block product (the dominant kernel of the A-application):
    algorithmic bytes per launch (every block value once, the S panel read and the U panel
    written once, the cell indexes) over the launch time from the HIP events that bracket each
    launch on the solve stream during the timed steps; traffic = measured HBM bytes per launch."""
    if fp32:   # only the FP32 launches of the mixed solves (its FP64 residual applications are the other kernel)
        ms = sum(s.ms_spmm_fp32 for s in stats)
        nb = sum(s.spmm_bytes_fp32 for s in stats)
        n = sum(s.n_spmm_fp32 for s in stats)
    else:
        ms = sum(s.ms_spmm for s in stats)
        nb = sum(s.spmm_bytes for s in stats)
        n = sum(s.n_spmm for s in stats)
    achieved = nb / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kernel),
            "kernel": kernel, "launches": n, "avg_us": round(1e3 * ms / max(n, 1), 2),
            "bytes_per_launch": round(nb / max(n, 1))}


def cpu_baseline(prob, ch, P, G, budget_s=20.0):
    """Oracle (numpy/scipy, 1 thread) converged GMRES on the same inputs: channel-solves/s."""
    sys.path.insert(0, ROOT)
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from threadpoolctl import threadpool_limits
    import scipy.sparse.linalg as sla
    from oracle import solve as so
    from helpers import oracle_pw, oracle_swp, oracle_V, flags
    val, row, ptr = P.to_host()
    tf, ib = flags(prob)
    Vu, Vc = oracle_V(prob)
    N = prob.rp.Np_WP
    with threadpool_limits(limits=1):
        K = so.KernelA(oracle_pw(ch), oracle_swp(prob), Vu, Vc, so.p_matrix(val, row, ptr, ch.dim), tf, ib)
        done, t0 = 0, time.perf_counter()
        U = np.zeros((ch.nd, ch.nd, ch.nE), complex)
        while True:
            for dr in range(ch.nd):
                for j in range(ch.nE):
                    e = np.zeros(ch.dim)
                    e[ch.deut[dr] * N * N + ch.q_idx[j] * N] = 1.0
                    b = K.row_apply(e[None, :])[0].astype(complex)
                    g = G[j]
                    op = sla.LinearOperator((ch.dim, ch.dim), dtype=complex,
                                            matvec=lambda v, g=g: v - K.row_apply((v * g)[None, :])[0])
                    y, _ = sla.gmres(op, b, rtol=1e-12, restart=60, maxiter=50)
                    for dc in range(ch.nd):
                        U[dr, dc, j] = y[ch.deut[dc] * N * N + ch.q_idx[j] * N]
            done += 1
            el = time.perf_counter() - t0
            if el > budget_s / 2 or done >= 8:
                break
    return {"value": round(done / el, 4), "unit": "channel-solves/s", "cores": 1, "kind": "port",
            "sample": "%d full channel-solve(s) (4 GMRES systems each, rtol 1e-12) of the same workload, "
                      "scipy.sparse on 1 thread" % done}, U


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-batched", action="store_true", help="skip the all-channel batched line")
    ap.add_argument("--no-grad", action="store_true", help="skip the gradient line")
    ap.add_argument("--no-n40", action="store_true", help="skip the Np=Nq=40 precision line")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each channel-solve is "
              "independent; see DESIGN.md for how the work would shard)", file=sys.stderr)
        sys.exit(2)

    prob, ch = build()
    ctx = tt.Context(0)
    P = ch.build_p123(ctx)
    ctx.synchronize()

    t_mix, U_mix, st_mix = timed_steps(ctx, ch, P, tt.SOLVE_MIXED, args.steps, args.warmup)
    t_64, U_64, st_64 = timed_steps(ctx, ch, P, tt.SOLVE_GMRES, args.steps, args.warmup)
    t_pcie, _, _ = timed_steps(ctx, ch, P, tt.SOLVE_MIXED, args.steps, 1, host_G=True)
    rel = float(np.max(np.abs(U_mix - U_64)) / np.max(np.abs(U_64)))
    value = args.steps / t_mix
    fp64 = args.steps / t_64

    out = {
        "metric": "U-matrix channel-solves/s batched-mixed vs plain FP64 @|\u0394U/U|\u22641e-8; grad/fwd time",
        "value": round(value, 3),
        "unit": "channel-solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mix / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 products + fp64 refinement (complex)",
        "data": "synthetic",
        "config": {"workload": WORKLOAD, "nnz_P123": int(P.nnz), "p123_build_ms": round(P.stats.ms_total, 2)},
        "fp64_value": round(fp64, 3),
        "mixed_over_fp64": round(value / fp64, 3),
        "value_pcie_inclusive": round(args.steps / t_pcie, 3),
        "mixed_vs_fp64_max_rel": rel,
        "grad_over_fwd": None,
        "roofline": product_roofline(st_mix, "k_cellmfma<float, 2, 1, 4>", fp32=True),
        "roofline_fp64": product_roofline(st_64, "k_cellmfma<double, 2, 1, 2>"),
    }
    if not args.no_batched:
        out["batched"] = batched(ctx)
    if not args.no_n40:
        out["precision_n40"] = precision_n40(ctx)
    if not args.no_grad:
        out["gradient"] = gradient_ratio(ctx)
        out["grad_over_fwd"] = out["gradient"]["grad_over_fwd"]
    if not args.no_cpu:
        G = ch.resolvent(ctx)
        cb, U_cpu = cpu_baseline(prob, ch, P, G)
        cb["max_rel_vs_gpu_fp64"] = float(np.max(np.abs(U_cpu - U_64)) / np.max(np.abs(U_64)))
        out["cpu_baseline"] = cb
    print(json.dumps(out))


if __name__ == "__main__":
    main()
