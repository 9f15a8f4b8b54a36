"""Stokes saddle-point solve benchmark on one MI355X.

One step = one full solve, StokesLinearSystem::solve: numeric ILU(k)
factorisation of A and M_p, the learned initial guess, and the Krylov solve
to the FP64 true residual ||b - K x|| <= 1e-8 ||b||. rhs and x are resident in
HBM when the timed region starts.

Default workload: the 3D 48^3 ellipsoid system (2.74M velocity + 118k
pressure dofs) with ILU(2) on A, a held-out field (sample 512; the network was
trained on samples 0..255 at 32^3). The headline value is the mixed-precision
solve (FP32 preconditioner inside FP64 FGMRES) started from the network's
guess; beside it, on the same handle and factorisation: the mixed solve from a
zero guess and the plain FP64 solve of the same method from a zero guess.

    python bench.py [--steps K] [--warmup W] [--config C5] [--no-cpu] ...
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-rtol-1e-8 per Stokes solve, mixed+learned vs plain FP64 GPU; HBM GB/s"
WORKLOADS = {
    "C1": "2D inclusion, Q2-Q1 on 64x64 quads, viscosity contrast 1e3 disk r=0.1, simple shear, "
          "periodic x; ILU(0)",
    "C2": "2D Rayleigh-Taylor, Q2-Q1 on 256x256 quads, interface y=0.5+0.02cos(2 pi x), "
          "viscosity 1/100, density 1/1.1, unit gravity, free slip; ILU(1) on A, ILU(0) on M_p",
    "C3": "3D porphyroclast, Q2-Q1 on 32^3 hexes, sphere r=0.2 contrast 1e3, linear simple shear; ILU(1)",
    "C4": "2D 128x128 seeded ellipses (1-8, log10 contrast U[0,3]), simple shear; ILU(1)",
    "C5": "3D 48^3 seeded ellipsoids (1-4, semi-axes U[0.08,0.25], log10 contrast U[1,3]), simple shear; "
          "ILU(2) FP32 inner + FP64 outer + 3D neural-operator warm start trained on 256 seeded 32^3 solves",
}
HELD_OUT = {"C4": 512, "C5": 512}  # first sample outside the networks' training and held-out sets
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline_leg(lin, S, kA, plain_stats, inner_a, threads):
    """The oracle's FP64 restatement of the same method, timed on the host on
    bounded samples of this system and scaled by the GPU solve's own counts:
      numeric ILU(k): on the leading principal block of P A P^T that holds the
        first stages (its factor is exactly the leading part of the full
        factor's pattern rule), scaled by the ratio of IKJ work (sum over L
        entries of U-row lengths) full / sample;
      ILU application (serial, as the reference's vmult): on that block,
        scaled by nnz(full factor) / nnz(block factor);
      SpMV with K and with A: the full matrices, `threads` threads.
    Estimate = numeric + outer (max(1, inner_a) ILU apply + inner_a A SpMV +
    K SpMV). The Arnoldi vector work and the inner CG on M_p are left out (a
    lower bound)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py
    oracle_py.build()
    import scipy.sparse as sp
    perm, n_stages = lin.ordering(0)
    blk, stg = lin.schedule(0)
    fi = lin.factor_info(0)
    A = S.blocks["A"]
    nu = A.shape[0]
    iperm = np.empty(nu, np.int64)
    iperm[perm] = np.arange(nu)
    out = {}
    for n_st in (2, 3, 4):
        # the rows of the first n_st stages, as the factor orders them
        m = int(blk[stg[min(n_st, n_stages)]])
        R = A[perm[:m]].tocoo()
        keep = iperm[R.col] < m
        B = sp.csr_matrix((R.data[keep], (R.row[keep], iperm[R.col[keep]])), shape=(m, m))
        B.sort_indices()
        t = time.perf_counter()
        F = oracle_py.SparseILU(B, kA)
        t_sym = time.perf_counter() - t
        t = time.perf_counter()
        F.refactor(B)
        t_num = time.perf_counter() - t
        Fs = F.factors()
        rows = np.repeat(np.arange(m), np.diff(Fs.indptr))
        low = Fs.indices < rows
        diag = np.diff(Fs.indptr) - np.bincount(rows[Fs.indices <= rows], minlength=m)  # U entries per row
        upd = int(diag[Fs.indices[low]].sum())
        v = np.random.default_rng(0).standard_normal(m)
        t = time.perf_counter()
        for _ in range(3):
            F.vmult(v)
        t_app = (time.perf_counter() - t) / 3
        out = dict(m=m, stages=n_st, t_sym=t_sym, t_num=t_num, upd=upd, nnz=Fs.nnz, t_app=t_app)
        log("cpu sample: %d rows (%d of %d stages): symbolic %.2fs numeric %.2fs (%.3e IKJ updates), "
            "ILU application %.3fs (nnz %d)" % (m, n_st, n_stages, t_sym, t_num, upd, t_app, Fs.nnz))
        if t_num >= 2.0 or t_sym + t_num > 12.0:
            break
    K = S.K()
    x = np.random.default_rng(1).standard_normal(K.shape[1])
    oracle_py.spmv(K, x)
    t = time.perf_counter()
    for _ in range(3):
        oracle_py.spmv(K, x)
    t_k = (time.perf_counter() - t) / 3
    t_a = t_k * A.nnz / K.nnz
    num = out["t_num"] * fi["ikj_updates"] / max(1, out["upd"])
    app = out["t_app"] * fi["nnz"] / max(1, out["nnz"])
    outer = plain_stats["outer_iters"]
    est = num + outer * (max(1, inner_a) * app + inner_a * t_a + t_k)
    return {"value": round(est * 1e3, 1), "unit": "ms", "cores": threads, "kind": "port",
            "sample": ("oracle FP64 restatement of the same method (FGMRES, %d inner ILU(%d)-PCG steps on A; "
                       "0: one ILU application), "
                       "timed on bounded samples of this system and scaled: numeric ILU(%d) on the leading "
                       "%d rows of P A P^T (%d of %d colour stages) %.2f s for %.3e IKJ updates -> %.1f s for "
                       "the full factor's %.3e; one serial ILU application %.4f s on %d factor entries -> "
                       "%.3f s for %d; K SpMV %.3f s (%d threads); times the FP64 GPU solve's %d outer "
                       "iterations; Arnoldi vector work and the inner CG on M_p not counted (a lower bound)"
                       % (inner_a, kA, kA, out["m"], out["stages"], n_stages, out["t_num"], out["upd"], num,
                          fi["ikj_updates"], out["t_app"], out["nnz"], app, fi["nnz"], t_k, threads, outer)),
            "numeric_s": round(num, 2), "ilu_apply_s": round(app, 4), "spmv_K_s": round(t_k, 4)}


def assembly_leg(S, t_gen):
    """The same system's blocks and rhs re-assembled on the GPU from its
    quadrature samples (fs_assembler), checked bit for bit against the host
    generator's, with the device time beside the host generator's."""
    import torch
    from felspa_amd.system import BLOCK_NAMES, DeviceAssembler
    asm = DeviceAssembler(S)
    eta, rho = asm.sample_field()
    asm.assemble(eta, rho)
    ms = [asm.assemble(eta, rho) for _ in range(3)]
    same = True
    for b in BLOCK_NAMES:
        want = torch.from_numpy(S.blocks[b].data).to(asm.rhs.device)
        same &= bool(torch.equal(asm.values[b].view(torch.int64), want.view(torch.int64)))
        del want
    want = torch.from_numpy(S.rhs).to(asm.rhs.device)
    same &= bool(torch.equal(asm.rhs.view(torch.int64), want.view(torch.int64)))
    del want, eta, rho
    asm.close()
    torch.cuda.synchronize()
    out = {"gpu_ms": round(min(ms), 3), "gpu_ms_runs": [round(m, 3) for m in ms],
           "host_generator_s": round(t_gen, 2), "bit_identical_to_host": same,
           "note": "fs_assembler_assemble on the system's patterns from device quadrature samples; the host "
                   "generator time also covers the patterns and the Cuthill-McKee numbering"}
    log("device assembly %.2f ms (host generator %.1f s), bit-identical: %s" % (min(ms), t_gen, same))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C5")
    ap.add_argument("--sample", type=int, default=None, help="seeded field (default: held out from training)")
    ap.add_argument("--n-cells", type=int, default=None)
    ap.add_argument("--inner-a", type=int, default=None, help="inner ILU-PCG steps (default: the configuration's)")
    ap.add_argument("--restart", type=int, default=None, help="FGMRES basis (default: the configuration's)")
    ap.add_argument("--hint", type=int, default=None, help="lattice nodes per block edge (default: the configuration's)")
    ap.add_argument("--ilu-k", type=int, default=None, help="override the configuration's ILU level on A")
    ap.add_argument("--weights", default=None, help="network weights (default: the committed ones for the dimension)")
    ap.add_argument("--no-learned", action="store_true", help="headline from the zero guess")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-zero", action="store_true", help="skip the mixed zero-guess leg")
    ap.add_argument("--no-errors", action="store_true", help="skip the host residual / direct-solve checks")
    ap.add_argument("--no-assembly", action="store_true", help="skip the device-assembly leg")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--kernel-reps", type=int, default=10, help="launches per kernel class in the roofline timing")
    ap.add_argument("--copy-probe-mb", type=float, default=0.0,
                    help="also run a plain device copy of this many MB (HBM counter calibration for profiles)")
    ap.add_argument("--kernels-only", action="store_true",
                    help="no solves: factorise, time the kernel classes (and the copy probe) and print their "
                         "figures (for hardware-counter profiles)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the solve is not sharded across GPUs "
              "(see the sharding design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import felspa_amd as fa
    from felspa_amd import _lib as L
    from felspa_amd import guess as G

    cfg = args.config
    settings = fa.solver_settings(cfg)
    kA, kS = settings["ilu_k_A"], settings["ilu_k_S"]
    if args.ilu_k is not None:
        kA = args.ilu_k
    sample = args.sample if args.sample is not None else HELD_OUT.get(cfg, 0)
    t0 = time.time()
    S = fa.StokesSystem(fa.make_problem(cfg, n_cells=args.n_cells, sample=sample))
    t_gen = time.time() - t0
    log("assembled %s sample %d: n_u=%d n_p=%d nnz(A)=%d in %.1fs" % (cfg, sample, S.n_u, S.n_p,
                                                                       S.blocks["A"].nnz, t_gen))
    hint_nodes = args.hint or settings["hint_nodes"]
    inner_a = args.inner_a if args.inner_a is not None else settings["inner_a"]
    restart = args.restart or settings["restart"]
    t0 = time.time()
    # one handle: the mixed and the plain FP64 solves share the symbolic phase
    # and the factor (the precision is switched between solves); both stop the
    # inner CG on M_p at the configuration's relative tolerance and step cap
    cg_tol, cg_max = settings["inner_cg_tol"], settings["inner_cg_max"]
    lin = fa.StokesLinearSystem(S.blocks, A_lof=kA, S_lof=kS, precision=L.FS_PREC_MIXED, guess=L.FS_GUESS_ZERO,
                                block_hint=S.block_hint(hint_nodes), constraints=S.constrained,
                                restart=restart, inner_a_iters=inner_a, max_outer=3000,
                                inner_cg_tol=cg_tol, inner_cg_max=cg_max, inner_cg_tol_mixed=cg_tol,
                                inner_cg_max_mixed=cg_max)
    t_create = time.time() - t0
    log("handle: orderings + symbolic ILU(%d) + upload %.1fs" % (kA, t_create))
    model, weights = None, None
    if not args.no_learned:
        weights = args.weights or (G.DEFAULT_2D if S.dim == 2 else G.DEFAULT_3D)
        if os.path.exists(weights):
            model = G.GuessModel.load(weights)
            lin.set_guess_field(model, S)
        else:
            log("no network weights at %s: headline from the zero guess" % weights)

    dev = torch.device("cuda:0")
    d_rhs = torch.from_numpy(S.rhs).to(dev)
    d_x = torch.zeros(S.n, dtype=torch.float64, device=dev)

    def solve():
        rc, st = lin.solve_device(d_rhs.data_ptr(), d_x.data_ptr())
        if rc != 0:
            raise RuntimeError("solve did not converge: %s" % st)
        return st

    def timed(steps, warmup, label):
        for i in range(warmup):
            st = solve()
            log("%s warmup %d: %.1f ms, outer %d, res %.2e" % (label, i, st["t_total_ms"], st["outer_iters"],
                                                                st["true_rel_residual"]))
        torch.cuda.synchronize()
        t = time.perf_counter()
        sts = [solve() for _ in range(steps)]
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / steps
        st = sts[-1]
        log("%s: %.1f ms/solve over %d (setup %.1f + krylov %.1f ms; outer %d, inner CG %d, guess %s %.3e)" % (
            label, ms, steps, st["t_setup_ms"], st["t_krylov_ms"], st["outer_iters"], st["inner_iters"],
            {0: "user", 1: "zero", 2: "learned"}[st["guess_used"]], st["guess_rel_residual"]))
        return ms, st, d_x.cpu().numpy().copy()

    if args.kernels_only:
        kinds = {0: "A-block SpMV (k_spmv_blk_dot, FP32)",
                 2: "ILU(k) apply on A (k_ilu_ext_partials + k_ilu_stage_dense, FP32)",
                 1: "K SpMV (FP64 values, the Arnoldi product)"}
        out = {"config_id": cfg, "kernels": {}}
        for kind in (0, 2, 1):
            ms, by = lin.time_kernel(kind, precision=L.FS_PREC_MIXED,
                                     reps=args.kernel_reps)
            out["kernels"][kinds[kind]] = {"ms_per_unit": round(ms, 5), "bytes_per_unit_format": by,
                                           "reps": args.kernel_reps}
            log("kernel %-64s %.4f ms/unit  %.3e B" % (kinds[kind], ms, by))
        if args.copy_probe_mb > 0:
            nel = int(args.copy_probe_mb * 1e6 / 4)
            a = torch.ones(nel, dtype=torch.float32, device=dev)
            c = torch.empty_like(a)
            for _ in range(3):
                c.copy_(a)
            torch.cuda.synchronize()
            out["copy_probe_bytes_per_copy"] = 2 * 4 * nel
            del a, c
        lin.close()
        if model is not None:
            model.close()
        del d_rhs, d_x
        torch.cuda.synchronize()
        print(json.dumps(out), flush=True)
        return

    head_mode = L.FS_GUESS_LEARNED if model is not None else L.FS_GUESS_ZERO
    lin.set_guess_mode(head_mode)
    ms_head, st_head, x_head = timed(args.steps, args.warmup, "mixed+learned" if model else "mixed")
    extra = {}
    if model is not None and not args.no_zero:
        lin.set_guess_mode(L.FS_GUESS_ZERO)
        ms_zero, st_zero, _ = timed(1, 0, "mixed, zero guess")
        extra.update(mixed_zero_guess_ms=round(ms_zero, 3), mixed_zero_guess_outer=st_zero["outer_iters"],
                     speedup_learned_vs_zero_guess=round(ms_zero / ms_head, 3))
    x_plain, st_plain = None, None
    if not args.no_plain:
        lin.set_precision(L.FS_PREC_FP64)
        lin.set_guess_mode(L.FS_GUESS_ZERO)
        ms_plain, st_plain, x_plain = timed(1, 0, "plain FP64")
        lin.set_precision(L.FS_PREC_MIXED)
        extra.update(plain_fp64_ms=round(ms_plain, 3), plain_fp64_outer=st_plain["outer_iters"],
                     plain_fp64_setup_ms=round(st_plain["t_setup_ms"], 3),
                     plain_fp64_krylov_ms=round(st_plain["t_krylov_ms"], 3),
                     speedup_vs_plain_fp64=round(ms_plain / ms_head, 3))

    # accuracy, computed on the host with scipy from the caller's matrices
    # (independent of the GPU kernels): relative residuals, the difference
    # between the mixed and plain solutions, and below 250k dofs the error
    # against a refined sparse direct solve
    if not args.no_errors:
        K = S.K()
        bn = np.linalg.norm(S.rhs)
        extra["host_rel_residual"] = float(np.linalg.norm(S.rhs - K @ x_head) / bn)
        if x_plain is not None:
            extra["plain_host_rel_residual"] = float(np.linalg.norm(S.rhs - K @ x_plain) / bn)
            du = np.linalg.norm(x_head[:S.n_u] - x_plain[:S.n_u]) / np.linalg.norm(x_plain[:S.n_u])
            dp = x_head[S.n_u:] - x_plain[S.n_u:]
            tp = x_plain[S.n_u:]
            extra["diff_u_mixed_vs_plain"] = float(du)
            extra["diff_p_meanfree_mixed_vs_plain"] = float(np.linalg.norm(dp - dp.mean()) /
                                                            np.linalg.norm(tp - tp.mean()))
        if S.n <= 250000:
            import scipy.sparse.linalg as spla
            lu = spla.splu(K.tocsc())
            xt = lu.solve(S.rhs)
            for _ in range(2):
                xt = xt + lu.solve(S.rhs - K @ xt)

            def errs(x):
                eu = np.linalg.norm(x[:S.n_u] - xt[:S.n_u]) / np.linalg.norm(xt[:S.n_u])
                dp = x[S.n_u:] - xt[S.n_u:]
                tp = xt[S.n_u:]
                return float(eu), float(np.linalg.norm(dp - dp.mean()) / np.linalg.norm(tp - tp.mean()))
            extra["err_u_vs_direct"], extra["err_p_meanfree_vs_direct"] = errs(x_head)
            if x_plain is not None:
                extra["plain_err_u_vs_direct"], extra["plain_err_p_meanfree_vs_direct"] = errs(x_plain)
        log("accuracy:", {k: v for k, v in extra.items() if "resid" in k or "err" in k or "diff" in k})

    # roofline of the dominant kernel class of the headline solve, timed with
    # HIP events on the solver's own stream: the ILU(k) application on A (all
    # its stage and piece launches, the forward ones carrying the PCG update)
    # or the A-block SpMV, whichever holds the larger share of the solve
    fi = lin.factor_info(0)
    kinds = {0: "A-block SpMV (k_spmv_blk_dot, FP32)",
             2: "ILU(k) apply on A (k_ilu_ext_partials + k_ilu_stage_dense, FP32)",
             1: "K SpMV (FP64 values, the Arnoldi product)", 3: "inner CG on M_p (FP32)"}
    timing = {}
    for kind in (0, 2, 1, 3):
        ms, by = lin.time_kernel(kind, precision=L.FS_PREC_MIXED,
                                 reps=args.kernel_reps)
        timing[kind] = (ms, by)
        log("kernel %-64s %.4f ms/unit  %.3e B  %.1f GB/s" % (kinds[kind], ms, by, by / ms / 1e6))
    # per outer iteration: max(1, inner_a) ILU applications (inner_a = 0: one
    # application as the reference's A preconditioner) and inner_a SpMVs
    n_apply = st_head["outer_iters"] * max(1, inner_a)
    n_spmv = st_head["outer_iters"] * inner_a
    share = {0: n_spmv * timing[0][0] / st_head["t_total_ms"], 2: n_apply * timing[2][0] / st_head["t_total_ms"]}
    dom = 2 if share[2] >= share[0] else 0
    ms_dom, by_fmt = timing[dom]
    nu = S.n_u
    if dom == 2:
        # ILU apply bytes: nnz(ILU) (s_v + 4) + 2 (m+1) 8 + 3 m s_x, FP32 values and vectors
        by_alg = fi["nnz"] * (4 + 4) + 2 * (nu + 1) * 8 + 3 * nu * 4
    else:
        # CSR SpMV: nnz (s_v + 4) + (m+1) 8 + 2 m s_x
        by_alg = S.blocks["A"].nnz * (4 + 4) + (nu + 1) * 8 + 2 * nu * 4
    achieved = by_alg / (ms_dom * 1e-3) / 1e9
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath) and args.n_cells is None and args.ilu_k is None:
        with open(tpath) as f:
            tt = json.load(f).get(cfg, {}).get("ilu_apply" if dom == 2 else "spmv_A")
        if tt:
            traffic, traffic_src = tt["traffic_bytes"], tt["source"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": kinds[dom], "bytes_per_unit_alg": by_alg, "bytes_per_unit_format": by_fmt,
                "achieved_format_bytes": round(by_fmt / (ms_dom * 1e-3) / 1e9, 1),
                "ms_per_unit": round(ms_dom, 5), "share_of_solve": round(share[dom], 3),
                "factor_nnz": fi["nnz"], "colour_stages": fi["n_stages"]}
    # the numeric ILU(k) factorisation (FP64, recomputed in every solve as the
    # reference reinitialises its preconditioner): its share of the solve and
    # its rate in IKJ updates (one multiply-add each) against the FP64 vector
    # peak; it is bound by the dependent column steps of the elimination, not
    # by HBM
    t_fact = st_head["t_setup_ms"]
    extra["factorisation"] = {"ms": round(t_fact, 1), "share_of_solve": round(t_fact / st_head["t_total_ms"], 3),
                              "ikj_updates": fi["ikj_updates"],
                              "gflops": round(2.0 * fi["ikj_updates"] / (t_fact * 1e-3) / 1e9, 1),
                              "frac_of_fp64_vector_peak": round(2.0 * fi["ikj_updates"] / (t_fact * 1e-3) / 78.6e12, 4),
                              "note": "t_setup_ms of the headline solve: numeric ILU(k) of A (left-looking "
                                      "stage updates + in-block stage kernel), ILU(0) of M_p, block inverses"}
    extra["kernel_gbs_format_bytes"] = {kinds[k]: round(v[1] / (v[0] * 1e-3) / 1e9, 1) for k, v in timing.items()}
    extra["kernel_ms_per_unit"] = {kinds[k]: round(v[0], 5) for k, v in timing.items()}
    extra["solve_gbs_algorithmic"] = round(st_head["bytes_algorithmic"] / (st_head["t_total_ms"] * 1e-3) / 1e9, 1)

    if args.copy_probe_mb > 0:
        # a plain device-to-device copy of known size: the HBM byte counters of
        # a profile are scaled by what they read for it
        nel = int(args.copy_probe_mb * 1e6 / 4)
        a = torch.ones(nel, dtype=torch.float32, device=dev)
        c = torch.empty_like(a)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
        extra["copy_probe_bytes"] = 2 * 4 * nel
        del a, c

    cpu = None
    if not args.no_cpu and st_plain is not None:
        cpu = cpu_baseline_leg(lin, S, kA, st_plain, inner_a, args.cpu_threads)
        log("cpu baseline:", cpu["value"], "ms:", cpu["sample"])

    line = {
        "metric": METRIC, "value": round(ms_head, 3), "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_head, 3),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic (generated Q2-Q1 system, seeded field)",
        "config": {"workload": WORKLOADS.get(cfg, cfg), "config_id": cfg, "sample": sample,
                   "n_u": S.n_u, "n_p": S.n_p, "nnz_A": int(S.blocks["A"].nnz), "ilu_k_A": kA, "ilu_k_S": kS,
                   "restart": restart, "inner_a_pcg": inner_a, "inner_cg_tol": cg_tol, "inner_cg_max": cg_max,
                   "block_nodes": hint_nodes,
                   "guess": "learned" if model is not None else "zero", "learned_guess": model is not None,
                   "weights": os.path.relpath(weights, ROOT) if model is not None else None},
        "outer_iters": st_head["outer_iters"], "inner_cg_iters": st_head["inner_iters"],
        "guess_used": {0: "user", 1: "zero", 2: "learned"}[st_head["guess_used"]],
        "guess_rel_residual": st_head["guess_rel_residual"], "t_guess_ms": round(st_head["t_guess_ms"], 3),
        "true_rel_residual": st_head["true_rel_residual"],
        "t_setup_ms": round(st_head["t_setup_ms"], 3), "t_krylov_ms": round(st_head["t_krylov_ms"], 3),
        "t_create_s": round(t_create, 1),
        "roofline": roofline, "cpu_baseline": cpu,
    }
    line.update(extra)
    # explicit teardown before interpreter exit: device memory is released
    # while the runtime is still up
    lin.close()
    if model is not None:
        model.close()
    del d_rhs, d_x
    torch.cuda.synchronize()
    if not args.no_assembly:
        line["assembly"] = assembly_leg(S, t_gen)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
