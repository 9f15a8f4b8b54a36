"""Train the learned warm start on solves of a seeded configuration.

    python -m felspa_amd.train_guess [--config C4] [--train 512] [--held-out 64]
                                     [--epochs 150] [--out felspa_amd/weights/guess_ellipses_2d.fsgm]
                                     [--summary summary.json]

    3D (48^3 ellipsoid family, trained at 32^3):
    python -m felspa_amd.train_guess --config C5 --n-cells 32 --train 256 --held-out 16 --eval 4 \
        --modes 8 --ilu-k 1 --inner-a 15 --epochs 120 --summary felspa_amd/weights/guess_ellipsoids_3d.json

1. Generates the data with the GPU solver itself: samples 0..train-1 (and
   train..train+held_out-1 held out) of the configuration, each solved by
   mixed-precision FGMRES to rtol 1e-8 from a zero guess (the method without
   the network), the solution mapped to the Q2 node lattice.
2. Normalises every output channel by its training-set mean and standard
   deviation (stored in the weight file).
3. Trains the network with Adam on the library's BF16-MFMA forward and
   backward kernels, on the GPU.
4. On the held-out fields, solves once from the zero guess and once from the
   learned guess and reports outer iterations, guess residuals and times.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", default="C4")
    ap.add_argument("--n-cells", type=int, default=None)
    ap.add_argument("--train", type=int, default=512)
    ap.add_argument("--held-out", type=int, default=64)
    ap.add_argument("--eval", type=int, default=16, help="held-out fields solved learned vs zero")
    ap.add_argument("--epochs", type=int, default=150)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--width", type=int, default=32)
    ap.add_argument("--modes", type=int, default=12)
    ap.add_argument("--layers", type=int, default=4)
    ap.add_argument("--proj", type=int, default=64)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--data-rtol", type=float, default=1e-6,
                    help="rtol of the training solves (targets need far less than the solve's 1e-8)")
    ap.add_argument("--ilu-k", type=int, default=None,
                    help="ILU level of the data solves and of the held-out evaluation (default: the configuration's)")
    ap.add_argument("--inner-a", type=int, default=60, help="inner ILU-PCG steps on A per outer iteration")
    ap.add_argument("--restart", type=int, default=60, help="FGMRES restart of the data and evaluation solves")
    ap.add_argument("--data-budget-s", type=float, default=0.0, help="stop generating after this long (0: none)")
    ap.add_argument("--train-budget-s", type=float, default=0.0, help="stop training after this long (0: none)")
    ap.add_argument("--out", default=None)
    ap.add_argument("--summary", default=None)
    args = ap.parse_args(argv)

    import torch
    import felspa_amd as fa
    from felspa_amd import _lib as L
    from felspa_amd import guess as G

    dev = torch.device("cuda:0")
    nc = args.n_cells or fa.CONFIGS[args.config]["n_cells"]
    dim = fa.CONFIGS[args.config]["dim"]
    t0 = time.time()
    # held-out fields first (samples train .. train+held_out-1), then the
    # training fields, so a data budget only shortens the training set
    held = list(range(args.train, args.train + args.held_out))
    order = held + list(range(args.train))
    le, T, stats = G.generate_solutions(args.config, order, n_cells=nc, rtol=args.data_rtol, log=log,
                                        budget_s=args.data_budget_s, ilu_k=args.ilu_k, restart=args.restart,
                                        inner_a=args.inner_a)
    t_data = time.time() - t0
    log("data: %d solves in %.1fs, mean outer %.1f" % (len(stats), t_data, np.mean([s["outer"] for s in stats])))
    nh = args.held_out
    # training set first, held-out last
    le = np.concatenate([le[nh:], le[:nh]])
    T = np.concatenate([T[nh:], T[:nh]])
    ntr = len(stats) - nh
    samples = list(range(ntr)) + held
    mean = T[:ntr].mean(axis=(0, 2))
    std = T[:ntr].std(axis=(0, 2)) + 1e-12
    Tn = (T - mean[None, :, None]) / std[None, :, None]

    model = G.GuessModel(dim=dim, width=args.width, modes=args.modes, layers=args.layers, proj=args.proj,
                         seed=args.seed)
    model.set_normalization(mean, std)
    d_le = torch.from_numpy(le).to(dev)
    d_X = model.prepare_input(d_le, nc)
    d_T = torch.from_numpy(Tn.astype(np.float32)).to(dev)
    del T, Tn
    t1 = time.time()
    hist = G.train(model, d_X[:ntr], d_T[:ntr], nc, epochs=args.epochs, batch=args.batch, lr=args.lr,
                   d_Xv=d_X[ntr:], d_Tv=d_T[ntr:], log=log, seed=args.seed, budget_s=args.train_budget_s)
    t_train = time.time() - t1
    out = args.out or (G.DEFAULT_2D if dim == 2 else os.path.join(G.WEIGHTS_DIR, "guess_ellipsoids_3d.fsgm"))
    model.save(out)
    log("saved", out)
    del d_X, d_T

    # held-out: learned vs zero guess, same tolerance
    from felspa_amd.system import StokesSystem
    st_cfg = fa.solver_settings(args.config)
    ev = []
    lin = None
    for s in samples[ntr:ntr + args.eval]:
        S = StokesSystem(fa.make_problem(args.config, n_cells=nc, sample=s))
        if lin is None:
            lin = fa.StokesLinearSystem(S.blocks, A_lof=st_cfg["ilu_k_A"] if args.ilu_k is None else args.ilu_k,
                                        S_lof=st_cfg["ilu_k_S"], precision=L.FS_PREC_MIXED, guess=L.FS_GUESS_ZERO,
                                        block_hint=S.block_hint(st_cfg["hint_nodes"]), constraints=S.constrained,
                                        restart=args.restart, inner_a_iters=args.inner_a)
        else:
            lin.refactor(S.blocks)
        lin.set_guess_field(model, S)
        lin.set_guess_mode(L.FS_GUESS_ZERO)
        x = np.zeros(S.n)
        z = lin.solve(x, S.rhs)
        lin.set_guess_mode(L.FS_GUESS_LEARNED)
        x = np.zeros(S.n)
        g = lin.solve(x, S.rhs)
        ev.append(dict(sample=s, outer_zero=z["outer_iters"], outer_learned=g["outer_iters"],
                       ms_zero=z["t_total_ms"], ms_learned=g["t_total_ms"], guess_ms=g["t_guess_ms"],
                       guess_rel_residual=g["guess_rel_residual"], guess_used=g["guess_used"],
                       res_zero=z["true_rel_residual"], res_learned=g["true_rel_residual"]))
        log("held-out %d: outer zero %d learned %d, ms %.1f -> %.1f, guess residual %.3e" % (
            s, z["outer_iters"], g["outer_iters"], z["t_total_ms"], g["t_total_ms"], g["guess_rel_residual"]))
    if lin is not None:
        lin.close()
    summary = dict(config=args.config, n_cells=nc, train=ntr, held_out=args.held_out, epochs=len(hist),
                   data_rtol=args.data_rtol, ilu_k=args.ilu_k, inner_a=args.inner_a, restart=args.restart, mean_outer_data=float(np.mean([s["outer"] for s in stats])),
                   net=model.cfg, n_params=model.n_params, t_data_s=t_data, t_train_s=t_train,
                   final_train_loss=hist[-1][0], final_held_out_loss=hist[-1][1],
                   mean_outer_zero=float(np.mean([e["outer_zero"] for e in ev])),
                   mean_outer_learned=float(np.mean([e["outer_learned"] for e in ev])),
                   mean_ms_zero=float(np.mean([e["ms_zero"] for e in ev])),
                   mean_ms_learned=float(np.mean([e["ms_learned"] for e in ev])),
                   mean_guess_rel_residual=float(np.mean([e["guess_rel_residual"] for e in ev])),
                   eval=ev, loss_history=hist)
    js = json.dumps(summary)
    if args.summary:
        with open(args.summary, "w") as f:
            f.write(js)
    print(js, flush=True)


if __name__ == "__main__":
    main()
