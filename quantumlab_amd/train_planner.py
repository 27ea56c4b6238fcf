"""Train the learned Trotter planner on the GPU and write its planner file.

    python -m quantumlab_amd.train_planner [--count 20000] [--out quantumlab_amd/data/trotter_planner.txt]

Pipeline (every step on libquest_mi355x):
  1. instances of the Hamiltonian family (quantumlab_amd.hamiltonians) on small
     registers, n drawn from --sizes;
  2. labels: for orders 2 and 4, the batched in-LDS ladder (FP64) gives
     1 - |<psi_r|psi_4r>|^2 for r = 1, 2, 4, ... (the 4x-finer run as judge);
     the cheapest repetitions r*_o follow from the power law at the first rung
     that meets eps;
  3. size exponents beta_o: slope of the per-size mean log2 r*_o against
     log2(n); targets y_o = log2 r*_o - beta_o log2(n);
  4. the network (16 -> 64 -> 64 -> 2) trained with Adam by the library's
     forward/backward kernels on standardised features;
  5. the log2 margin: the --quantile quantile of the residuals (label minus
     prediction, both orders) on labelled instances of a larger register size
     (--calib-size) than any trained on, so it covers size extrapolation;
  6. evaluation on fresh instances of that size: the planner's (order, reps)
     judged by the 4x-finer FP64 run, and its cost in Pauli exponentials
     against plain step doubling (r = 1, 2, 4, ... until consecutive rungs
     agree to eps) at fixed order 2 and 4.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

from . import hamiltonians as H
from . import planner as PL

EPS = 1e-4


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def make_instances(seed: int, count: int, sizes):
    rng = np.random.default_rng([seed, 17])
    ns = rng.choice(np.asarray(sizes), size=count)
    return [H.trotter_instance(int(n), seed, i) for i, n in enumerate(ns)]


def label(qe, insts, max_log2=12, eps=EPS):
    """log2 of the cheapest repetitions (orders 2, 4) per instance; NaN if none."""
    inf2, _ = PL.batch_ladder(qe, insts, [2] * len(insts), max_log2, eps)
    inf4, _ = PL.batch_ladder(qe, insts, [4] * len(insts), max_log2, eps)
    return PL.log2_reps_from_ladders(inf2, inf4, eps)


def step_doubling_costs(qe, insts, order, eps=EPS, max_log2=12):
    """Repetitions spent by plain step doubling on each instance (sum over the
    rungs r = 1, 2, 4, ... it evolves until consecutive rungs agree to eps)
    and the reps it ends at (None if max_log2 ran out); rungs batched."""
    spent = np.zeros(len(insts))
    final = [None] * len(insts)
    prev = [None] * len(insts)
    active = list(range(len(insts)))
    for k in range(max_log2 + 1):
        if not active:
            break
        states = PL.batch_evolve(qe, [insts[j] for j in active], [order] * len(active), [1 << k] * len(active))
        still = []
        for j, st in zip(active, states):
            spent[j] += 1 << k
            if prev[j] is not None and 1 - abs(np.vdot(prev[j], st)) ** 2 <= eps:
                final[j] = 1 << k
            else:
                prev[j] = st
                still.append(j)
        active = still
    return spent, final


def judge(qe, insts, choices, eps=EPS):
    """1 - |<psi_(o,r)|psi_(o,4r)>|^2 of each choice (FP64 batched runs)."""
    orders = [o for o, _ in choices]
    a = PL.batch_evolve(qe, insts, orders, [r for _, r in choices])
    b = PL.batch_evolve(qe, insts, orders, [4 * r for _, r in choices])
    return np.array([1 - abs(np.vdot(x, y)) ** 2 for x, y in zip(a, b)])


def evaluate(qe, net, insts, eps=EPS):
    choices = [net.choose(i) for i in insts]
    inf = judge(qe, insts, choices, eps)
    learned = np.array([H.stages(o) * r for o, r in choices], dtype=float)
    sd2 = H.stages(2) * step_doubling_costs(qe, insts, 2, eps)[0]
    sd4 = H.stages(4) * step_doubling_costs(qe, insts, 4, eps)[0]
    terms = np.array([i.terms for i in insts], dtype=float)
    return {"instances": len(insts), "success_rate": float(np.mean(inf <= eps)),
            "max_infidelity": float(np.max(inf)),
            "order4_share": float(np.mean([o == 4 for o, _ in choices])),
            "pauli_exps_learned": float(np.sum(learned * terms)),
            "pauli_exps_step_doubling_o2": float(np.sum(sd2 * terms)),
            "pauli_exps_step_doubling_o4": float(np.sum(sd4 * terms)),
            "gain_vs_step_doubling_o2": float(np.sum(sd2 * terms) / np.sum(learned * terms)),
            "gain_vs_step_doubling_o4": float(np.sum(sd4 * terms) / np.sum(learned * terms))}


def fit(qe, X, L, sizes, Xc, Lc, sizes_c, epochs, seed, quantile, lr=2e-3, batch=256):
    """Train on (X, L) and calibrate the margin on (Xc, Lc)."""
    beta = PL.fit_size_exponents(L, sizes)
    T, M = PL.size_free_targets(L, sizes, beta)
    mean = X.mean(0)
    scale = np.maximum(X.std(0), 1e-3)
    net = PL.Planner(qe, seed=seed)
    net.set_normalisation(mean, scale)
    net.set_size_exponents(beta)
    losses = net.train(X, T, M, epochs=epochs, batch=batch, lr=lr, seed=seed)
    Tc, Mc = PL.size_free_targets(Lc, sizes_c, beta)
    resid = (Tc - net.predict(Xc))[Mc > 0]
    margin = float(np.quantile(resid, quantile)) if resid.size else 0.0
    net.set_margin(max(0.0, margin))
    return net, losses, {"size_exponents": [float(b) for b in beta],
                         "calib_rmse": float(np.sqrt(np.mean(resid ** 2))) if resid.size else None,
                         "calib_mean_residual": float(np.mean(resid)) if resid.size else None,
                         "margin_log2": net.margin(), "train_loss_first": float(losses[0]),
                         "train_loss_last": float(losses[-1])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=20000)
    ap.add_argument("--sizes", type=str, default="6,7,8,9,10,11")
    ap.add_argument("--calib-size", type=int, default=12)
    ap.add_argument("--calib-count", type=int, default=2000)
    ap.add_argument("--eval-count", type=int, default=300)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--quantile", type=float, default=0.99)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--out", type=str, default=os.path.join(os.path.dirname(__file__), "data", "trotter_planner.txt"))
    ap.add_argument("--summary", type=str, default=None)
    args = ap.parse_args()
    import quantumlab_amd
    qe = quantumlab_amd.load()
    sizes = [int(s) for s in args.sizes.split(",")]
    t0 = time.perf_counter()
    insts = make_instances(args.seed, args.count, sizes)
    calib = make_instances(args.seed + 1, args.calib_count, [args.calib_size])
    X = PL.feature_matrix(qe.lib, insts)
    Xc = PL.feature_matrix(qe.lib, calib)
    log(f"{len(insts)} + {len(calib)} instances, features in {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    L = label(qe, insts)
    Lc = label(qe, calib)
    t_label = time.perf_counter() - t1
    ns = np.array([i.n for i in insts])
    log(f"labels in {t_label:.1f} s; labelled fractions {np.isfinite(L).mean(0)}")
    for n in sizes + [args.calib_size]:
        Ls = L[ns == n] if n in sizes else Lc
        log(f"  n={n}: mean log2 r* o2 {np.nanmean(Ls[:, 0]):.3f} o4 {np.nanmean(Ls[:, 1]):.3f}")
    t2 = time.perf_counter()
    net, losses, fitinfo = fit(qe, X, L, ns, Xc, Lc, np.full(len(calib), args.calib_size), args.epochs, args.seed,
                               args.quantile)
    t_train = time.perf_counter() - t2
    log(f"trained in {t_train:.1f} s: {fitinfo}")
    ev_insts = [H.trotter_instance(args.calib_size, args.seed + 2, i) for i in range(args.eval_count)]
    t3 = time.perf_counter()
    ev = evaluate(qe, net, ev_insts)
    log(f"fresh n={args.calib_size}: {ev} ({time.perf_counter() - t3:.1f} s)")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    net.save(args.out)
    summary = {"instances": len(insts), "sizes": sizes, "calib_size": args.calib_size, "calib_instances": len(calib),
               "eps": EPS, "label_seconds": t_label, "train_seconds": t_train, "epochs": args.epochs, **fitinfo,
               "evaluation": ev}
    if args.summary:
        with open(args.summary, "w") as f:
            json.dump(summary, f, indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
