"""Learned warm start: data from the plain method, training and evaluation.

The network (include/magritte_hip.h, mgr_mlp_*) maps per-cell features --
log10 collision-partner density, temperature, species density, neighbour
velocity gradient and log10 Jlin/B_nu(T) of the first four transitions -- to
log10 level-population fractions.  Jlin comes from a first pass from LTE, a
radiation field over every stride-th ray pair (mgr_first_pass; --stride,
default FIRST_PASS_STRIDE), or, with stride 0, the Jlin inputs are 0 and no
radiation field is paid before the ALI: the shipped weights are trained that
way (--stride 0), which converged as fast as first-pass networks on held-out
models and costs less.  Its training data are populations the plain
LTE-started ALI iteration converged on many small envelope models; training
and inference run in the project's own HIP kernels.

    python -m magritte_amd.warmstart train --models 2000 --epochs 80 --stride 0 --out weights.npz
    python -m magritte_amd.warmstart evaluate --weights weights.npz --models 8

A warm start replaces only the starting populations: the ALI loop after it is
the plain one, so the converged populations are those of the method.
"""
import argparse
import json
import sys
import time

import numpy as np

from .hip import MLP_IN, MLP_OUT, Context, Mlp
from .setup import synthetic

DEFAULT_WEIGHTS = __file__.replace("warmstart.py", "data/warmstart_mlp.npz")
CRITERION = "max"
MAX_ITER = 400
FIRST_PASS_STRIDE = 4


def training_model(rng):
    """An envelope of config-3 physics: N log-uniform in [1000, 32768] cells
    (n^3 grid, n = round(N^(1/3))), n0 log-uniform 1e12..1e16 m^-3, T0 uniform
    500..3000 K, 12 or 48 rays."""
    N = np.exp(rng.uniform(np.log(1000.0), np.log(32768.0)))
    n = int(np.clip(round(N ** (1.0 / 3.0)), 10, 32))
    n0 = 10.0 ** rng.uniform(12.0, 16.0)
    T0 = rng.uniform(500.0, 3000.0)
    nside = int(rng.choice([1, 2]))
    seed = int(rng.integers(0, 2 ** 31))
    return synthetic.envelope_model(n=n, n0=n0, T0=T0, nside=nside, seed=seed)


def first_pass(model, stride=FIRST_PASS_STRIDE):
    """Context after set-up, trace and the warm start's first pass from LTE
    (a radiation field over every stride-th ray pair, and J_eff); stride 0:
    no first pass (the Jlin features are 0)."""
    ctx = Context(model).setup()
    ctx.trace()
    if stride > 0:
        ctx.first_pass("fp32", stride)
    return ctx


def targets(model, pops):
    ntot = model.abundance[:, model.linedata.num]
    return np.log10(np.maximum(pops / ntot[:, None], 1e-30)).astype(np.float32)


TRAIN_MAX_ITER = 100


def build_dataset(n_models, seed=0, log=print, stride=FIRST_PASS_STRIDE):
    """Features and converged log10 fractions of `n_models` training models.
    A model whose LTE-started iteration does not meet the criterion within
    TRAIN_MAX_ITER iterations, or ends with a non-positive population, gives
    no usable target and is left out (counted in the returned `skipped`)."""
    rng = np.random.default_rng(seed)
    Xs, Ys, its = [], [], []
    skipped = 0
    t0 = time.perf_counter()
    for i in range(n_models):
        m = training_model(rng)
        ctx = first_pass(m, stride)
        X = ctx.features()
        it = ctx.compute_level_populations(True, TRAIN_MAX_ITER, CRITERION, "fp32")
        P = ctx.populations()
        ctx.close()
        if it >= TRAIN_MAX_ITER or not np.all(P > 0.0) or not np.all(np.isfinite(X)):
            skipped += 1
            continue
        Xs.append(X)
        Ys.append(targets(m, P))
        its.append(it)
        if (i + 1) % max(1, n_models // 20) == 0:
            log(f"[warmstart] {i + 1}/{n_models} models solved ({time.perf_counter() - t0:.1f} s, "
                f"{sum(x.shape[0] for x in Xs)} cells, {skipped} not converged)")
    return np.concatenate(Xs), np.concatenate(Ys), np.array(its), skipped


def fit(X, Y, epochs=30, batch=16384, lr=4e-3, seed=0, log=print):
    """Adam on the normalised data in three phases of falling step size
    (lr for `epochs`, lr/5 for epochs/2, lr/25 for epochs/4)."""
    mlp = Mlp(seed=seed)
    xm, xs = X.mean(0), X.std(0) + 1e-6
    ym, ys = Y.mean(0), Y.std(0) + 1e-6
    mlp.set_normalisation(xm, xs, ym, ys)
    t0 = time.perf_counter()
    losses = []
    for k, (ep, step) in enumerate(((epochs, lr), (max(1, epochs // 2), lr / 5), (max(1, epochs // 4), lr / 25))):
        losses.extend(mlp.train(X, Y, ep, batch, step, seed + k).tolist())
    losses = np.array(losses)
    log(f"[warmstart] trained {len(losses)} epochs on {X.shape[0]} cells in {time.perf_counter() - t0:.1f} s; "
        f"loss {losses[0]:.4g} -> {losses[-1]:.4g}")
    return mlp, dict(x_mean=xm, x_std=xs, y_mean=ym, y_std=ys), losses


def save(path, mlp, norm, meta):
    p = mlp.params()
    np.savez(path, **{k: v.astype(np.float32) for k, v in p.items()},
             **{k: np.asarray(v, np.float32) for k, v in norm.items()}, meta=json.dumps(meta))


def load(path, device=0):
    """The network of a weights file; its first-pass stride (1 for files
    without one: a full radiation field) is kept as mlp.first_pass_stride."""
    d = np.load(path, allow_pickle=False)
    mlp = Mlp(seed=0, device=device)
    mlp.set_params({k: d[k] for k in ("W1", "b1", "W2", "b2", "W3", "b3")})
    mlp.set_normalisation(d["x_mean"], d["x_std"], d["y_mean"], d["y_std"])
    meta = json.loads(str(d["meta"])) if "meta" in d.files else {}
    mlp.first_pass_stride = int(meta.get("first_pass_stride", 1))
    if mlp.first_pass_stride == 0:
        mlp.set_jlin_inputs(False)   # trained without the first pass
    return mlp


def converge(model, mlp=None, precision="fp32", max_iter=MAX_ITER):
    """Iterations and seconds to the max-change criterion from LTE (mlp None)
    or from the network's populations.  The clock starts after the one-off
    per-model set-up (trace, masks, far-end cache, launch tuning and the ALI
    graphs: mgr_prepare) and includes, for the warm start, the first pass,
    the features and the inference.  Also returns the cells whose prediction
    the warm start rejected (they kept LTE)."""
    ctx = Context(model).setup()
    ctx.trace()
    ctx.prepare(precision)
    t0 = time.perf_counter()
    fallbacks = 0
    if mlp is not None:
        stride = getattr(mlp, "first_pass_stride", 1)
        if stride > 0:
            ctx.first_pass(precision, stride)
        fallbacks = ctx.warm_start(mlp)
    it = ctx.compute_level_populations(True, max_iter, CRITERION, precision)
    dt = time.perf_counter() - t0
    pops = ctx.populations()
    ctx.close()
    return int(it), dt, pops, fallbacks


def evaluate_model(model, mlp, precision="fp32", max_iter=MAX_ITER):
    """LTE start against the warm start on one model.  `converged` is False
    when the LTE-started iteration (the reference for the comparison) does not
    meet the criterion within max_iter iterations."""
    it_l, t_l, p_l, _ = converge(model, None, precision, max_iter)
    it_n, t_n, p_n, fb = converge(model, mlp, precision, max_iter)
    ntot = model.abundance[:, model.linedata.num]
    mask = p_l > 1e-10 * ntot[:, None]
    rel = np.abs(p_n - p_l)[mask] / np.abs(p_l[mask])
    return {"converged": bool(it_l < max_iter and it_n < max_iter),
            "lte_iterations": it_l, "lte_seconds": round(t_l, 4), "nn_iterations": it_n,
            "nn_seconds": round(t_n, 4), "speedup": round(t_l / t_n, 3),
            "pops_max_rel_diff": float(rel.max()), "cells_kept_lte": int(fb)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    tr = sub.add_parser("train")
    tr.add_argument("--models", type=int, default=1000)
    tr.add_argument("--epochs", type=int, default=30)
    tr.add_argument("--batch", type=int, default=16384)
    tr.add_argument("--lr", type=float, default=4e-3)
    tr.add_argument("--stride", type=int, default=FIRST_PASS_STRIDE,
                    help="ray-pair stride of the first pass; 0: no first pass (the Jlin inputs are 0)")
    tr.add_argument("--seed", type=int, default=0)
    tr.add_argument("--out", required=True)
    ev = sub.add_parser("evaluate")
    ev.add_argument("--weights", default=DEFAULT_WEIGHTS)
    ev.add_argument("--models", type=int, default=8)
    ev.add_argument("--n", type=int, default=64, help="cells per side; 0: models of the training distribution")
    ev.add_argument("--seed", type=int, default=12345)
    ev.add_argument("--out", default=None)
    ev.add_argument("--max-iter", type=int, default=100)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if a.cmd == "train":
        t0 = time.perf_counter()
        X, Y, its, skipped = build_dataset(a.models, a.seed, log, a.stride)
        t_data = time.perf_counter() - t0
        t0 = time.perf_counter()
        mlp, norm, losses = fit(X, Y, a.epochs, a.batch, a.lr, a.seed, log)
        t_fit = time.perf_counter() - t0
        meta = {"models": a.models, "models_not_converged": int(skipped), "cells": int(X.shape[0]),
                "epochs": a.epochs, "batch": a.batch, "lr": a.lr, "first_pass_stride": a.stride,
                "seconds_data": round(t_data, 1), "seconds_training": round(t_fit, 1),
                "seed": a.seed, "mean_iterations_lte": float(its.mean()), "final_loss": float(losses[-1]),
                "features": ["log10 n_partner", "log10 T", "log10 n_species", "log10 max |dv/dx| (1/s)",
                             "log10 Jlin_0/B_0", "log10 Jlin_1/B_1", "log10 Jlin_2/B_2", "log10 Jlin_3/B_3"],
                "targets": "log10 n_i / n_tot, i = 0..6"}
        save(a.out, mlp, norm, meta)
        print(json.dumps({"losses": [float(x) for x in losses], **meta}))
        return 0
    mlp = load(a.weights)
    rng = np.random.default_rng(a.seed)
    rows = []
    for i in range(a.models):
        if a.n > 0:
            n0 = 10.0 ** rng.uniform(12.0, 16.0)
            T0 = rng.uniform(500.0, 3000.0)
            seed = int(rng.integers(0, 2 ** 31))
            m = synthetic.envelope_model(n=a.n, n0=n0, T0=T0, seed=seed)
            r = {"n0": n0, "T0": T0, "seed": seed}
        else:   # held-out draws from the training distribution
            m = training_model(rng)
            r = {"n0": m.meta["n0"], "T0": m.meta["T0"], "ncells": m.ncells, "nrays": m.nrays}
        r.update(evaluate_model(m, mlp, max_iter=a.max_iter))
        log(f"[warmstart] held-out {i}: {json.dumps(r)}")
        rows.append(r)
    ok = [r for r in rows if r["converged"]] or [{"speedup": float("nan"), "lte_iterations": float("nan"),
                                                  "nn_iterations": float("nan"), "pops_max_rel_diff": float("nan")}]
    summary = {"models": a.models, "ncells": a.n ** 3 if a.n > 0 else "training distribution",
               "converged_models": sum(r["converged"] for r in rows),
               "mean_speedup": float(np.mean([r["speedup"] for r in ok])),
               "mean_lte_iterations": float(np.mean([r["lte_iterations"] for r in ok])),
               "mean_nn_iterations": float(np.mean([r["nn_iterations"] for r in ok])),
               "models_fewer_iterations": sum(r["nn_iterations"] < r["lte_iterations"] for r in rows if r["converged"]),
               "models_more_iterations": sum(r["nn_iterations"] > r["lte_iterations"] for r in rows if r["converged"]),
               "max_pops_rel_diff": float(np.max([r["pops_max_rel_diff"] for r in ok])),
               "models_with_cells_kept_lte": sum(r["cells_kept_lte"] > 0 for r in rows),
               "cells_kept_lte": sum(r["cells_kept_lte"] for r in rows), "rows": rows}
    if a.out:
        with open(a.out, "w") as fh:
            json.dump(summary, fh, indent=1)
    print(json.dumps({k: v for k, v in summary.items() if k != "rows"}))
    return 0


if __name__ == "__main__":
    sys.exit(main())
