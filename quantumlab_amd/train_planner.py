"""Train the learned Trotter planner on the GPU and write its planner file.

    python -m quantumlab_amd.train_planner [--count 20000] [--out quantumlab_amd/data/trotter_planner.txt]

Pipeline (every step on libquest_mi355x):
  1. instances of the Hamiltonian family (quantumlab_amd.hamiltonians) on small
     registers, n drawn from --sizes;
  2. labels: for orders 2 and 4, the cheapest integer r*_o with
     1 - |<psi_r|psi_4r>|^2 <= eps (the 4x-finer FP64 run as judge): the
     batched in-LDS ladder r = 1, 2, 4, ... brackets it and batched bisection
     finds it;
  3. calibration labels at the deployment size (--calib-size, 24 qubits):
     the same criterion and search through the public API (applyTrotterCircuit
     on FP64 registers, fused sweeps) for --calib-count instances;
  4. the size curve: the mean log2 r*_o of every training size and of the
     calibration size, piecewise linear in log2(n); targets
     y_o = log2 r*_o - curve_o(n);
  5. the network (16 -> 64 -> 64 -> 2) trained with Adam by the library's
     forward/backward kernels on standardised features;
  6. the log2 margin: the larger of the --quantile quantiles of held-out
     training residuals and of the calibration residuals (label minus
     prediction, both orders), so it covers the deployment size too;
  7. evaluation on fresh 12-qubit instances (batched kernels): the planner's
     (order, reps) judged by the 4x-finer FP64 run, and its cost in Pauli
     exponentials against plain step doubling (r = 1, 2, 4, ... until
     consecutive rungs agree to eps) at fixed order 2 and 4.
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
    """log2 of the cheapest integer repetitions (orders 2, 4) per instance
    meeting 1 - F(psi_r, psi_4r) <= eps; NaN if none."""
    L = np.full((len(insts), 2), np.nan)
    for c, o in enumerate((2, 4)):
        for j, r in enumerate(PL.cheapest_reps(qe, insts, o, eps, max_log2)):
            if r is not None:
                L[j, c] = np.log2(r)
    return L


def _infid_large(qe, hb, inst, order, r, init, a, b):
    qe.lib.cloneQureg(a, init)
    qe.lib.applyTrotterCircuit(a, hb[0], inst.time, order, r)
    qe.lib.cloneQureg(b, init)
    qe.lib.applyTrotterCircuit(b, hb[0], inst.time, order, 4 * r)
    ip = qe.calcInnerProduct(a, b)
    return 1 - (ip.real ** 2 + ip.imag ** 2) / (qe.calcTotalProb(a) * qe.calcTotalProb(b))


def cheapest_reps_large(qe, inst, order, eps=EPS, max_log2=10):
    """The smallest integer r with normalised 1 - F(psi_r, psi_4r) <= eps on a
    full-size FP64 register through the public API: ladder r = 1, 2, 4, ...,
    then bisection inside the bracket; None if the ladder never passes."""
    from .hamiltonians import HamilBatch
    hb = HamilBatch([inst])
    init = qe.qureg(inst.n)
    H.prepare(qe, init, inst)
    a, b = qe.qureg(inst.n), qe.qureg(inst.n)
    hi = None
    for k in range(max_log2 + 1):
        if _infid_large(qe, hb, inst, order, 1 << k, init, a, b) <= eps:
            hi = 1 << k
            break
    if hi is not None:
        lo = hi // 2 if hi > 1 else 0
        while hi - lo > 1:
            mid = (lo + hi) // 2
            if _infid_large(qe, hb, inst, order, mid, init, a, b) <= eps:
                hi = mid
            else:
                lo = mid
    for q in (init, a, b):
        qe.destroy(q)
    return hi


def label_large(qe, insts, eps=EPS):
    L = np.full((len(insts), 2), np.nan)
    t0 = time.perf_counter()
    for j, inst in enumerate(insts):
        for c, o in enumerate((2, 4)):
            r = cheapest_reps_large(qe, inst, o, eps)
            if r is not None:
                L[j, c] = np.log2(r)
        log(f"  calibration instance {j + 1}/{len(insts)} ({inst.model}, {inst.terms} terms): "
            f"r* = {2 ** L[j, 0]:.0f} (order 2), {2 ** L[j, 1]:.0f} (order 4); {time.perf_counter() - t0:.1f} s")
    return L


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


def fit(qe, X, L, sizes, Xc, Lc, sizes_c, epochs, seed, quantile, lr=2e-3, batch=256, val_frac=0.1):
    """Train on (X, L); the size curve runs through the per-size means of both
    sets; the margin is the larger of the `quantile` quantile of held-out
    training residuals and of the calibration residuals (Xc, Lc)."""
    sizes = np.asarray(sizes)
    sizes_c = np.asarray(sizes_c)
    kx, ky = PL.fit_size_curve(np.vstack([L, Lc]), np.concatenate([sizes, sizes_c]))
    T, M = PL.size_free_targets(L, sizes, kx, ky)
    perm = np.random.default_rng(seed).permutation(len(X))
    nv = max(1, int(len(X) * val_frac))
    va, tr = perm[:nv], perm[nv:]
    mean = X[tr].mean(0)
    scale = np.maximum(X[tr].std(0), 1e-3)
    net = PL.Planner(qe, seed=seed)
    net.set_normalisation(mean, scale)
    net.set_size_curve(kx, ky)
    losses = net.train(X[tr], T[tr], M[tr], epochs=epochs, batch=batch, lr=lr, seed=seed)
    rv = (T[va] - net.predict(X[va]))[M[va] > 0]
    Tc, Mc = PL.size_free_targets(Lc, sizes_c, kx, ky)
    rc = (Tc - net.predict(Xc))[Mc > 0]
    margin = max(float(np.quantile(rv, quantile)) if rv.size else 0.0,
                 float(np.quantile(rc, quantile)) if rc.size else 0.0)
    net.set_margin(max(0.0, margin))
    return net, losses, {"size_curve_log2n": [float(v) for v in kx],
                         "size_curve_log2reps": [[float(a), float(b)] for a, b in ky],
                         "val_rmse": float(np.sqrt(np.mean(rv ** 2))) if rv.size else None,
                         "val_quantile": float(np.quantile(rv, quantile)) if rv.size else None,
                         "calib_rmse": float(np.sqrt(np.mean(rc ** 2))) if rc.size else None,
                         "calib_quantile": float(np.quantile(rc, quantile)) if rc.size else None,
                         "margin_log2": net.margin(), "train_loss_first": float(losses[0]),
                         "train_loss_last": float(losses[-1])}


def load_calibration_cache(args):
    """(labels, done) for at least calib_count instances of the calibration
    seed and size: the cache's entries where it matches, NaN / False after."""
    n = args.calib_count
    labels, done = np.full((n, 2), np.nan), np.zeros(n, dtype=bool)
    if args.calib_labels and os.path.exists(args.calib_labels):
        z = np.load(args.calib_labels, allow_pickle=False)
        if (int(z["seed"]) == args.seed + 1 and int(z["size"]) == args.calib_size
                and str(z["method"]) == "bisection"):
            zl = np.asarray(z["labels"], dtype=np.float64)
            zd = np.asarray(z["done"], dtype=bool) if "done" in z.files else np.ones(len(zl), dtype=bool)
            if len(zl) > n:
                labels, done = np.full((len(zl), 2), np.nan), np.zeros(len(zl), dtype=bool)
            labels[:len(zl)], done[:len(zl)] = zl, zd
    return labels, done


def save_calibration_cache(args, labels, done):
    os.makedirs(os.path.dirname(os.path.abspath(args.calib_labels)), exist_ok=True)
    np.savez(args.calib_labels, labels=labels, done=done, seed=args.seed + 1, size=args.calib_size,
             method="bisection")


def label_calibration_range(args, a, b):
    """Label calibration instances a..b-1 (those not yet in the cache), saving
    after each one."""
    import quantumlab_amd
    qe = quantumlab_amd.load()
    args.calib_count = max(args.calib_count, b)
    labels, done = load_calibration_cache(args)
    t0 = time.perf_counter()
    for i in range(a, b):
        if done[i]:
            continue
        inst = H.trotter_instance(args.calib_size, args.seed + 1, i)
        labels[i] = label_large(qe, [inst])[0] if args.calib_size > 12 else label(qe, [inst])[0]
        done[i] = True
        save_calibration_cache(args, labels, done)
        log(f"calibration instance {i}: log2 r* = {labels[i]}; {time.perf_counter() - t0:.1f} s")
    log(f"calibration labels done: {int(done.sum())} of {len(done)}")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=20000)
    ap.add_argument("--sizes", type=str, default="6,7,8,9,10,11,12")
    ap.add_argument("--calib-size", type=int, default=24)
    ap.add_argument("--calib-count", type=int, default=32)
    ap.add_argument("--eval-size", type=int, default=12)
    ap.add_argument("--eval-count", type=int, default=300)
    ap.add_argument("--epochs", type=int, default=200)
    ap.add_argument("--quantile", type=float, default=0.99)
    ap.add_argument("--seed", type=int, default=1234)
    ap.add_argument("--out", type=str, default=os.path.join(os.path.dirname(__file__), "data", "trotter_planner.txt"))
    ap.add_argument("--summary", type=str, default=None)
    ap.add_argument("--calib-labels", type=str,
                    default=os.path.join(os.path.dirname(__file__), "data", "trotter_calib_labels.npz"),
                    help="cache of the calibration labels (reused when its seed and size match; missing entries "
                         "are labelled and added)")
    ap.add_argument("--calib-range", type=str, default=None,
                    help="A:B — only label calibration instances A..B-1 into the cache (saved after each "
                         "instance, so the work can be split over several runs), then stop")
    args = ap.parse_args()
    if args.calib_range:
        a, b = (int(v) for v in args.calib_range.split(":"))
        label_calibration_range(args, a, b)
        return
    import quantumlab_amd
    qe = quantumlab_amd.load()
    sizes = [int(s) for s in args.sizes.split(",")]
    t0 = time.perf_counter()
    insts = make_instances(args.seed, args.count, sizes)
    calib = [H.trotter_instance(args.calib_size, args.seed + 1, i) for i in range(args.calib_count)]
    X = PL.feature_matrix(qe.lib, insts)
    Xc = PL.feature_matrix(qe.lib, calib)
    log(f"{len(insts)} + {len(calib)} instances, features in {time.perf_counter() - t0:.1f} s")
    t1 = time.perf_counter()
    L = label(qe, insts)
    t_label = time.perf_counter() - t1
    ns = np.array([i.n for i in insts])
    log(f"small-register labels in {t_label:.1f} s; labelled fractions {np.isfinite(L).mean(0)}")
    t1 = time.perf_counter()
    labels, done = load_calibration_cache(args)
    missing = [i for i in range(len(calib)) if not done[i]]
    if missing:
        Lm = label(qe, [calib[i] for i in missing]) if args.calib_size <= 12 else \
            label_large(qe, [calib[i] for i in missing])
        labels[missing], done[missing] = Lm, True
        save_calibration_cache(args, labels, done)
    else:
        log(f"calibration labels from {args.calib_labels}")
    Lc = labels[:len(calib)]
    t_calib = time.perf_counter() - t1
    log(f"{args.calib_size}-qubit labels in {t_calib:.1f} s; labelled fractions {np.isfinite(Lc).mean(0)}")
    for n in sizes:
        log(f"  n={n}: mean log2 r* o2 {np.nanmean(L[ns == n, 0]):.3f} o4 {np.nanmean(L[ns == n, 1]):.3f}")
    log(f"  n={args.calib_size}: mean log2 r* o2 {np.nanmean(Lc[:, 0]):.3f} o4 {np.nanmean(Lc[:, 1]):.3f}")
    t2 = time.perf_counter()
    net, losses, fitinfo = fit(qe, X, L, ns, Xc, Lc, np.full(len(calib), args.calib_size), args.epochs, args.seed,
                               args.quantile)
    t_train = time.perf_counter() - t2
    log(f"trained in {t_train:.1f} s: {fitinfo}")
    ev_insts = [H.trotter_instance(args.eval_size, args.seed + 2, i) for i in range(args.eval_count)]
    t3 = time.perf_counter()
    ev = evaluate(qe, net, ev_insts)
    log(f"fresh n={args.eval_size}: {ev} ({time.perf_counter() - t3:.1f} s)")
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)
    net.save(args.out)
    summary = {"instances": len(insts), "sizes": sizes, "calib_size": args.calib_size, "calib_instances": len(calib),
               "eps": EPS, "label_seconds": t_label, "calib_label_seconds": t_calib, "train_seconds": t_train,
               "epochs": args.epochs, **fitinfo,
               "mean_log2_reps": {str(n): [float(np.nanmean(L[ns == n, c])) for c in range(2)] for n in sizes}
               | {str(args.calib_size): [float(np.nanmean(Lc[:, c])) for c in range(2)]},
               "evaluation_size": args.eval_size, "evaluation": ev}
    if args.summary:
        with open(args.summary, "w") as f:
            json.dump(summary, f, indent=1)
    print(json.dumps(summary))


if __name__ == "__main__":
    main()
