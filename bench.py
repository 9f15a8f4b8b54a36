"""Stokes saddle-point solve benchmark on one MI355X.

One step = one full solve, StokesLinearSystem::solve: numeric ILU(k)
factorisation of A and M_p plus the Krylov solve to the FP64 true residual
||b - K x|| <= 1e-8 ||b||. The rhs and x are already resident in HBM when
the timed region starts. The headline value is the mixed-precision solve
(FP32 SpMV / ILU / inner Krylov inside the FP64 outer FGMRES); the plain FP64
GPU solve of the same algorithm is timed beside it.

    python bench.py [--steps K] [--warmup W] [--config C2] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("time-to-rtol-1e-8 per Stokes solve, mixed+learned vs plain FP64 GPU; HBM GB/s")
WORKLOADS = {
    "C2": "2D Rayleigh-Taylor, Q2-Q1 on 256x256 quads, interface y=0.5+0.02cos(2 pi x), "
          "viscosity 1/100, density 1/1.1, unit gravity, free slip; ILU(1) on A, ILU(0) on M_p",
    "C1": "2D inclusion, Q2-Q1 on 64x64 quads, viscosity contrast 1e3 disk r=0.1, simple shear, "
          "periodic x; ILU(0)",
    "C3": "3D porphyroclast, Q2-Q1 on 32^3 hexes, sphere r=0.2 contrast 1e3, linear simple shear",
    "C4": "2D 128x128 seeded ellipses (1-8, log10 contrast U[0,3]), simple shear",
    "C5": "3D 48^3 seeded ellipsoids (1-4, log10 contrast U[1,3]), simple shear",
}
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def learned_leg(args, log):
    """Held-out fields (samples 512+) of the 2D ellipse family at 128^2: one
    mixed-precision handle (numeric refactor per field) timed from the zero
    guess and from the network's guess (inference inside the solve), and the
    plain FP64 solve of the same method; inputs resident in HBM."""
    import torch
    import felspa_amd as fa
    from felspa_amd import _lib as L
    from felspa_amd import guess as G
    if not os.path.exists(G.DEFAULT_2D):
        log("learned leg skipped: %s not present" % G.DEFAULT_2D)
        return None
    model = G.GuessModel.load(G.DEFAULT_2D)
    st_cfg = fa.solver_settings("C4")
    rows = []
    mixed = plain = None
    dev = torch.device("cuda:0")
    for s in range(512, 512 + args.learned_samples):
        S = fa.StokesSystem(fa.make_problem("C4", sample=s))
        common = dict(A_lof=st_cfg["ilu_k_A"], S_lof=st_cfg["ilu_k_S"], block_hint=S.block_hint(5),
                      constraints=S.constrained, restart=args.restart, inner_a_iters=args.inner_a,
                      max_outer=3000)
        if mixed is None:
            mixed = fa.StokesLinearSystem(S.blocks, precision=L.FS_PREC_MIXED, guess=L.FS_GUESS_ZERO, **common)
            plain = fa.StokesLinearSystem(S.blocks, precision=L.FS_PREC_FP64, guess=L.FS_GUESS_ZERO,
                                          inner_cg_tol=1e-2, inner_cg_max=50, **common)
        else:
            mixed.refactor(S.blocks)
            plain.refactor(S.blocks)
        mixed.set_guess_field(model, S)
        d_rhs = torch.from_numpy(S.rhs).to(dev)
        d_x = torch.zeros(S.n, dtype=torch.float64, device=dev)
        res = {}
        for label, lin, mode in (("zero", mixed, L.FS_GUESS_ZERO), ("learned", mixed, L.FS_GUESS_LEARNED),
                                 ("plain_fp64", plain, L.FS_GUESS_ZERO)):
            lin.set_guess_mode(mode)
            lin.solve_device(d_rhs.data_ptr(), d_x.data_ptr())  # warm-up (numeric factorisation cached)
            torch.cuda.synchronize()
            t = time.perf_counter()
            rc, st = lin.solve_device(d_rhs.data_ptr(), d_x.data_ptr())
            torch.cuda.synchronize()
            ms = (time.perf_counter() - t) * 1e3
            if rc != 0:
                raise RuntimeError("learned leg: %s solve did not converge" % label)
            res[label] = dict(ms=round(ms, 3), outer=st["outer_iters"], res=st["true_rel_residual"],
                              guess_used=st["guess_used"], guess_rel_residual=st["guess_rel_residual"],
                              t_guess_ms=round(st["t_guess_ms"], 3))
        log("learned leg sample %d: zero %.1f ms (%d outer), learned %.1f ms (%d outer, guess res %.2e), "
            "plain fp64 %.1f ms (%d outer)" % (s, res["zero"]["ms"], res["zero"]["outer"], res["learned"]["ms"],
                                               res["learned"]["outer"], res["learned"]["guess_rel_residual"],
                                               res["plain_fp64"]["ms"], res["plain_fp64"]["outer"]))
        rows.append(dict(sample=s, **{k: v for k, v in res.items()}))
    mean = lambda k, f: float(np.mean([r[k][f] for r in rows]))
    out = {"workload": "C4: " + WORKLOADS["C4"] + "; held-out samples 512..%d" % (511 + args.learned_samples),
           "network": dict(model.cfg, n_params=model.n_params, weights=os.path.relpath(G.DEFAULT_2D, ROOT)),
           "ms_mixed_learned": round(mean("learned", "ms"), 3), "ms_mixed_zero": round(mean("zero", "ms"), 3),
           "ms_plain_fp64": round(mean("plain_fp64", "ms"), 3),
           "outer_mixed_learned": mean("learned", "outer"), "outer_mixed_zero": mean("zero", "outer"),
           "outer_plain_fp64": mean("plain_fp64", "outer"),
           "guess_rel_residual": mean("learned", "guess_rel_residual"),
           "t_guess_ms": round(mean("learned", "t_guess_ms"), 3)}
    out["speedup_mixed_learned_vs_plain_fp64"] = round(out["ms_plain_fp64"] / out["ms_mixed_learned"], 3)
    out["speedup_learned_vs_zero_guess"] = round(out["ms_mixed_zero"] / out["ms_mixed_learned"], 3)
    out["samples"] = rows
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="C2")
    ap.add_argument("--n-cells", type=int, default=None)
    ap.add_argument("--inner-a", type=int, default=60)
    ap.add_argument("--restart", type=int, default=60)
    ap.add_argument("--hint", type=int, default=None, help="lattice nodes per block edge (default 5 in 2D, 3 in 3D)")
    ap.add_argument("--ilu-k", type=int, default=None, help="override the configuration's ILU level on A")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-fp64-vectors", action="store_true",
                    help="skip the FP32-storage / FP64-vector leg (a third handle: host memory at 48^3)")
    ap.add_argument("--cpu-outer", type=int, default=2, help="outer iterations in the CPU sample")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-warm-start leg (C4 fields)")
    ap.add_argument("--learned-samples", type=int, default=4, help="held-out C4 fields in the learned leg")
    ap.add_argument("--no-ref", action="store_true", help="skip the tight FP64 reference solve (error columns)")
    ap.add_argument("--kernel-reps", type=int, default=30, help="launches per kernel class in the roofline timing")
    ap.add_argument("--copy-probe-mb", type=float, default=0.0,
                    help="also run a plain device copy of this many MB (HBM counter calibration for profiles)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the solve is not sharded across GPUs "
              "(see the sharding design notes). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    import felspa_amd as fa
    from felspa_amd import _lib as L

    cfg = args.config
    settings = fa.solver_settings(cfg)
    kA, kS = settings["ilu_k_A"], settings["ilu_k_S"]
    if args.ilu_k is not None:
        kA = args.ilu_k
    t0 = time.time()
    S = fa.StokesSystem(fa.make_problem(cfg, n_cells=args.n_cells))
    log("assembled %s: n_u=%d n_p=%d nnz(A)=%d in %.1fs" % (cfg, S.n_u, S.n_p, S.blocks["A"].nnz,
                                                            time.time() - t0))
    if args.hint is None:
        args.hint = 5 if S.dim == 2 else 3
    hint = S.block_hint(args.hint)
    common = dict(A_lof=kA, S_lof=kS, block_hint=hint, constraints=S.constrained,
                  restart=args.restart, inner_a_iters=args.inner_a, max_outer=3000)
    t0 = time.time()
    mixed = fa.StokesLinearSystem(S.blocks, precision=L.FS_PREC_MIXED, guess=L.FS_GUESS_ZERO, **common)
    log("mixed handle: setup (orderings + symbolic ILU + upload) %.1fs" % (time.time() - t0))
    plain = None
    if not args.no_plain:
        # the same method entirely in FP64 (same inner tolerances and counts)
        plain = fa.StokesLinearSystem(S.blocks, precision=L.FS_PREC_FP64, guess=L.FS_GUESS_ZERO,
                                      inner_cg_tol=1e-2, inner_cg_max=50, **common)

    dev = torch.device("cuda:0")
    d_rhs = torch.from_numpy(S.rhs).to(dev)
    d_x = torch.zeros(S.n, dtype=torch.float64, device=dev)

    def solve(lin):
        rc, st = lin.solve_device(d_rhs.data_ptr(), d_x.data_ptr())
        if rc != 0:
            raise RuntimeError("solve did not converge: %s" % st)
        return st

    def timed(lin, steps, warmup, label):
        for i in range(warmup):
            st = solve(lin)
            log("%s warmup %d: %.1f ms, outer %d, inner %d, res %.2e" % (
                label, i, st["t_total_ms"], st["outer_iters"], st["inner_iters"], st["true_rel_residual"]))
        torch.cuda.synchronize()
        t = time.perf_counter()
        sts = [solve(lin) for _ in range(steps)]
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t) * 1e3 / steps
        log("%s: %.1f ms/solve over %d solves (device setup %.1f + krylov %.1f ms)" % (
            label, ms, steps, sts[-1]["t_setup_ms"], sts[-1]["t_krylov_ms"]))
        return ms, sts[-1]

    ms_mixed, st_mixed = timed(mixed, args.steps, args.warmup, "mixed")
    x_mixed = d_x.cpu().numpy().copy()
    out_extra = {}
    if not args.no_plain and not args.no_fp64_vectors:
        # the same method with FP32-stored matrices and factors but the
        # preconditioner's vectors and arithmetic in FP64
        f64v = fa.StokesLinearSystem(S.blocks, precision=L.FS_PREC_MIXED_FP64_VECTORS, guess=L.FS_GUESS_ZERO,
                                     **common)
        ms_f64v, st_f64v = timed(f64v, 1, 1, "mixed-fp64-vectors")
        out_extra.update(mixed_fp64_vectors_ms=round(ms_f64v, 3),
                         mixed_fp64_vectors_outer_iters=st_f64v["outer_iters"])
        del f64v
    if plain is not None:
        ms_plain, st_plain = timed(plain, max(1, args.steps), min(args.warmup, 1), "plain-fp64")
        x_plain = d_x.cpu().numpy().copy()
        out_extra.update(plain_fp64_ms=round(ms_plain, 3), speedup_vs_plain_fp64=round(ms_plain / ms_mixed, 3),
                         plain_outer_iters=st_plain["outer_iters"], plain_inner_iters=st_plain["inner_iters"])
    # accuracy: against a tight FP64 GPU solve (rtol 1e-12) of the same system
    x_ref = None
    if not args.no_ref:
        ref_lin = fa.StokesLinearSystem(S.blocks, precision=L.FS_PREC_FP64, guess=L.FS_GUESS_ZERO, rtol=1e-12,
                                        **{k: v for k, v in common.items()})
        solve(ref_lin)
        x_ref = d_x.cpu().numpy().copy()
        del ref_lin

    def errs(x):
        nu = S.n_u
        eu = np.linalg.norm(x[:nu] - x_ref[:nu]) / np.linalg.norm(x_ref[:nu])
        dp = x[nu:] - x_ref[nu:]
        tp = x_ref[nu:]
        ep = np.linalg.norm(dp) / np.linalg.norm(tp)
        ep0 = np.linalg.norm(dp - dp.mean()) / np.linalg.norm(tp - tp.mean())
        return round(float(eu), 10), round(float(ep), 10), round(float(ep0), 10)

    if x_ref is not None:
        eu, ep, ep0 = errs(x_mixed)
        out_extra.update(err_u_vs_tight=eu, err_p_vs_tight=ep, err_p_meanfree_vs_tight=ep0)
        if plain is not None:
            e = errs(x_plain)
            out_extra.update(plain_err_u_vs_tight=e[0], plain_err_p_vs_tight=e[1])

    # roofline of the dominant kernel class of the mixed solve, timed with HIP
    # events on the solver's own stream. Per mixed solve every inner PCG step
    # runs one A-block SpMV (kind 0) and one ILU(k) application (kind 2, all
    # its stage launches); the class with the larger share of the solve is
    # reported, its unit being one SpMV launch or one ILU application.
    _, n_stages = mixed.ordering(0)
    kinds = {0: "A-block SpMV (k_spmv_blk_dot, FP32)",
             2: "ILU(k) apply on A (k_ilu_stage_dense fwd+bwd, FP32)",
             1: "K SpMV (k_spmv4, FP64)", 3: "inner CG on M_p (k_inner_cg, FP32)"}
    # (the last colour stage's forward and backward solves share one launch)
    launches = {0: 1, 2: 2 * n_stages - 1, 1: 1, 3: 1}
    timing = {}
    for kind in (0, 2, 1, 3):
        ms, by = mixed.time_kernel(kind, precision=L.FS_PREC_FP64 if kind == 1 else L.FS_PREC_MIXED,
                                   reps=args.kernel_reps)
        timing[kind] = (ms, by)
        log("kernel %-52s %.4f ms/unit  %.3e B  %.1f GB/s" % (kinds[kind], ms, by, by / ms / 1e6))
    n_inner = st_mixed["inner_iters"]
    pcg_steps = st_mixed["outer_iters"] * args.inner_a
    share = {k: pcg_steps * timing[k][0] / st_mixed["t_total_ms"] for k in (0, 2)}
    dom = 2 if share[2] >= share[0] else 0
    ms_dom, by_dom = timing[dom]
    achieved = by_dom / (ms_dom * 1e-3) / 1e9
    # HBM bytes per unit from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    # passes on this workload (scaled by a plain copy of known size)
    traffic, traffic_src = None, None
    tpath = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(tpath) and args.n_cells is None and args.ilu_k is None:
        with open(tpath) as f:
            tt = json.load(f).get(cfg, {}).get("ilu_apply" if dom == 2 else "spmv_A")
        if tt:
            traffic, traffic_src = tt["traffic_bytes"], tt["source"]
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                "kernel": kinds[dom],
                "bytes_per_unit": by_dom, "ms_per_unit": round(ms_dom, 5), "launches_per_unit": launches[dom],
                "share_of_solve": round(share[dom], 3)}
    out_extra["kernel_gbs"] = {kinds[k]: round(v[1] / (v[0] * 1e-3) / 1e9, 1) for k, v in timing.items()}
    out_extra["kernel_ms_per_unit"] = {kinds[k]: round(v[0], 5) for k, v in timing.items()}
    out_extra["solve_gbs_algorithmic"] = round(st_mixed["bytes_algorithmic"] / (st_mixed["t_total_ms"] * 1e-3) / 1e9, 1)

    if args.copy_probe_mb > 0:
        # a plain device-to-device copy of known size: the HBM byte counters of
        # a profile are scaled by what they read for it
        nel = int(args.copy_probe_mb * 1e6 / 4)
        a = torch.ones(nel, dtype=torch.float32, device=dev)
        c = torch.empty_like(a)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
        out_extra["copy_probe_bytes"] = 2 * 4 * nel
        del a, c

    # learned warm start: held-out fields of the ellipse family (2D 128^2),
    # mixed + learned guess vs mixed + zero guess vs the plain FP64 solve
    if not args.no_learned:
        out_extra["learned"] = learned_leg(args, log)

    # CPU baseline: the oracle's FP64 restatement of the same algorithm on the
    # host cores, on a bounded sample (a few outer iterations), scaled to one
    # solve by the plain FP64 GPU outer iteration count
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle_py
        threads = min(os.cpu_count() or 1, 16)
        outer = args.cpu_outer
        t = time.perf_counter()
        _, cst = oracle_py.stokes_solve(S.blocks, S.rhs, kA=kA, kS=kS, restart=args.restart, rtol=1e-8,
                                        method=2, threads=threads, max_steps=outer, inner_tol=1e-2,
                                        inner_a_iters=args.inner_a)
        wall = time.perf_counter() - t
        per_outer = cst["t_krylov_s"] / max(1, cst["outer_iters"])
        n_outer = (out_extra.get("plain_outer_iters") or st_mixed["outer_iters"])
        value = (cst["t_setup_s"] + per_outer * n_outer) * 1e3
        cpu = {"value": round(value, 1), "unit": "ms", "cores": threads, "kind": "port",
               "sample": "oracle FP64 FGMRES(%d) + inner ILU(%d)-PCG(%d) on A + CG/ILU(0) on M_p, same "
                         "matrices: numeric ILU (%.2f s) + %d outer iterations (%.2f s, %.1f s wall), "
                         "scaled to the %d outer iterations of the FP64 GPU solve; ILU application "
                         "serial as in the reference, SpMV/vector ops on %d threads" % (
                             args.restart, kA, args.inner_a, cst["t_setup_s"], cst["outer_iters"],
                             cst["t_krylov_s"], wall, n_outer, threads)}
        log("cpu baseline:", cpu)

    line = {
        "metric": METRIC, "value": round(ms_mixed, 3), "unit": "ms", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_mixed, 3),
        "higher_is_better": False, "scaling": "strong", "vs_baseline": None, "dtype": "fp32+fp64",
        "data": "synthetic (generated Q2-Q1 system)",
        "config": {"workload": WORKLOADS.get(cfg, cfg), "config_id": cfg, "n_u": S.n_u, "n_p": S.n_p,
                   "nnz_A": int(S.blocks["A"].nnz), "ilu_k_A": kA, "ilu_k_S": kS, "restart": args.restart, "inner_a_pcg": args.inner_a,
                   "block_nodes": args.hint, "guess": "zero", "learned_guess": False},
        "outer_iters": st_mixed["outer_iters"], "inner_iters": n_inner,
        "true_rel_residual": st_mixed["true_rel_residual"],
        "t_setup_ms": round(st_mixed["t_setup_ms"], 3), "t_krylov_ms": round(st_mixed["t_krylov_ms"], 3),
        "roofline": roofline, "cpu_baseline": cpu,
    }
    line.update(out_extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
