"""Train the learned block-Jacobi correction on the device and write its weights.

    python blocklinearsystems.jl_amd/train_precond.py [--seeds 32] [--scale 0.05] [--steps 600] [--out PATH]

Training data: the batched 64-problem generator (config5_union) with seeds 100 .. 100 + seeds - 1, i.e. seeds x 64
problems, landmark counts scaled by `scale` (at 1/20 block-Jacobi PCG needs tens to thousands of iterations, so
there is something to learn). Each union is reduced by the GPU Schur path; for every problem the group coupling
weight w in GRID (the same for all its groups of precond.GROUP cameras) that reaches ||S x - g|| <= 1e-8 ||g|| in the fewest PCG iterations is
the target of its cameras (w = 0, plain block-Jacobi, wins ties). The MLP regresses logit(w) from each camera's
features with the library's forward / backward / Adam kernels. Four further seeds are held out: the output scale
alpha in ALPHAS with the fewest mean PCG iterations on them is kept (alpha = 0 switches the correction off, i.e.
plain block-Jacobi when the network does not help).
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GRID = (0.0, 0.25, 0.5, 0.75, 1.0)
ALPHAS = (0.0, 0.5, 1.0)


def log(*a):
    print(*a, flush=True)


def reduced(b, torch, p):
    dev = b.device
    st = p.structure
    h = dev.SchurHandle(st, p.ws_index, pattern=p.pattern, r_to_s=p.r_to_s)
    ws = dev.SchurReductionWorkspace(h)
    T = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).cuda()
    S = torch.empty(p.pattern.nnz, dtype=torch.float64, device="cuda")
    g = torch.empty(st.nr, dtype=torch.float64, device="cuda")
    dev.reset_schur_sparse_(S, T(p.R_data), h, T(p.scaling_r), p.lam)
    if not dev.reduce_schur_sparse_(S, g, T(p.b_r), T(p.b_e), T(p.scaling_e), p.lam, h, T(p.E_data), T(p.C_data), ws):
        raise RuntimeError("reduction failed")
    return S, g, dev.PCGHandle(st, p.pattern, p.prob_block_ptr)


def dataset(b, torch, seeds, scale, maxit=20000):
    P = b.precond
    Xs, ys, table = [], [], []
    for seed in seeds:
        t0 = time.time()
        p = b.generators.config5_union(seed=seed, scale=scale)
        S, g, pcg = reduced(b, torch, p)
        F = P.features(pcg, S, p.lam)
        x = torch.zeros_like(g)
        its = []
        for w in GRID:
            M = P.group_blocks(pcg, S, torch.full((p.structure.nrblocks,), w, dtype=torch.float32, device="cuda"))
            x.zero_()
            ok, iters, _ = pcg.solve_groups(S, g, x, M, tol=1e-8, maxit=maxit)
            its.append(np.where(iters < 0, maxit, iters))
        its = np.array(its)                               # (grid, problems)
        best = np.array(GRID)[np.argmin(its, axis=0)]     # first minimum: the smallest weight
        prob_of_block = np.repeat(np.arange(p.nprob), np.diff(p.prob_block_ptr))
        Xs.append(F.cpu().numpy())
        ys.append(P.target_logit(best[prob_of_block]))
        table.append(its)
        log(f"seed {seed}: {p.structure.nrblocks} cameras, iterations plain max {its[0].max()} mean {its[0].mean():.1f}, "
            f"best-weight mean {its.min(0).mean():.1f} ({time.time() - t0:.1f}s)")
    return np.concatenate(Xs).astype(np.float32), np.concatenate(ys).astype(np.float32), table


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--seeds", type=int, default=32)
    ap.add_argument("--scale", type=float, default=0.05)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    import torch
    import blsload
    b = blsload.load()
    P = b.precond
    X, y, table = dataset(b, torch, range(100, 100 + args.seeds), args.scale)
    keep = (X != 0).any(axis=1)
    X, y = X[keep], y[keep]
    fmean, fstd = X.mean(0), np.maximum(X.std(0), 1e-3)
    model = P.MLP.initial(fmean, fstd, seed=0)
    losses = P.train(model, torch.as_tensor(X).cuda(), torch.as_tensor(y).cuda(), steps=args.steps, lr=1e-3, log=log)
    out = args.out or P.WEIGHTS
    # output scale on held-out seeds
    its = {a: [] for a in ALPHAS}
    for seed in range(100 + args.seeds, 104 + args.seeds):
        p = b.generators.config5_union(seed=seed, scale=args.scale)
        S, g, pcg = reduced(b, torch, p)
        x = torch.zeros_like(g)
        for a in ALPHAS:
            model.alpha = a
            ok, it, _, used = P.solve(pcg, S, g, x, p.lam, model=model if a > 0 else None, maxit=20000)
            its[a].append(it)
    mean_its = {a: float(np.concatenate(v).mean()) for a, v in its.items()}
    model.alpha = min(ALPHAS, key=lambda a: (mean_its[a], a))
    log("held-out mean PCG iterations by output scale: " +
        ", ".join(f"{a}: {m:.1f}" for a, m in mean_its.items()) + f" -> alpha {model.alpha}")
    model.save(out, losses=np.array(losses), seeds=np.array([100, args.seeds]), scale=np.array(args.scale),
               grid=np.array(GRID), alphas=np.array(ALPHAS), heldout_mean_iters=np.array([mean_its[a] for a in ALPHAS]))
    log(f"{len(y)} samples from {args.seeds * 64} problems; loss {losses[0]:.4f} -> {losses[-1]:.4f}; wrote {out}")


if __name__ == "__main__":
    main()
