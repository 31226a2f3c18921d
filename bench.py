"""Benchmark of the radiation-field hot path on one MI355X.

Workload (the at-scale single-GPU configuration): 262,144-cell jittered 64^3
point cloud with 26 grid-neighbour lists, 48 HEALPix rays (nside 2),
spherical envelope n_H2 = 1e14 (r0/r)^2 m^-3, T = 2500 (r0/r)^0.6 K, radial
outflow 15 (1 - R*/r)^0.5 km/s, 7-level linear rotor, 6 lines x 31 bins.

A step is one compute_radiation_field over the whole model (every ray pair x
every frequency bin), inputs resident in HBM.  value = ray-pair.freq solves
per second over the timed steps.  The JSON line also carries:
  roofline      -- the radiation-field kernel, timed with HIP events on its
                   own stream.  It fuses the line profiles and the Feautrier
                   solve and is bound by vector-ALU issue, so "bound" is
                   "valu": achieved = the solve's algorithmic FP32 flops
                   (28 per point and bin) per second against MI355X's FP32
                   peak (157.3 TFLOP/s published; 78.6 without packed FP32,
                   which this build does not use, in "frac_non_packed").
                   "traffic" and "hbm" give the measured HBM bytes per launch
                   and their rate against the 8 TB/s peak; "equivalent_bytes"
                   are the dtau/S/u/Lambda* bytes an HBM-staged solver would
                   move, which this kernel never moves.  The counter figures
                   come from profiles/ and are reported only when the profiled
                   kernel sources hash to the benched ones;
  cpu_baseline  -- the FP64 OpenMP oracle (restatement of the reference's
                   RayPair path) timed on a bounded random sample of origins,
                   with the host CPU model and thread count, and the largest
                   relative difference of the GPU's J from the oracle's over
                   that sample;
  fp64_kernel   -- the FP64 radiation-field kernel on the same workload;
  time_to_converged -- the ALI loop to max relative change < 1e-6, from LTE
                   and from the learned warm start (trained weights in
                   magritte_amd/data), with the converged populations' match.

--workload regular measures the same on the 16^3 regular grid (4096 cells,
6-neighbour lists, 12 rays, one 2-level species, 39 bins at 20 K), where the
CPU leg also runs the oracle's whole ALI: its time to converged populations
and their difference from the GPU's.

--workload stress measures the same on the optically thick 64^3 envelope
(n0 = 1e17 m^-3, 20 % of cells with +-30 km/s radial velocity jumps).

--workload batch measures the same on 512 small envelopes of uneven sizes
and ray counts solved together in one context (one launch per kernel).

--workload mlp measures training of the warm-start network (8 -> 128 -> 128
-> 7, BF16 MFMA, FP32 accumulation) on synthetic rows: rows per second and
the algorithmic matrix-core flop rate against the 2.5 PFLOP/s dense BF16 peak.
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Feautrier ray-pair·freq solves/s; time-to-converged pops, NN vs LTE start"
# Algorithmic FP32 flops per (point, frequency bin) of the Feautrier
# elimination (dtau 3, the two coefficients and their inverses 8, F and Su 10,
# back-substitution share 7).
FLOPS_PER_POINT_BIN = 28
# Algorithmic bytes of an HBM-staged solve: per point and bin 4 B dtau + 4 B S
# read, per solve 4 B u + 4 B Lambda* written.
BYTES_PER_POINT_BIN = 8
BYTES_PER_SOLVE = 8
PEAK_HBM_GBS = 8000.0
# MI355X published dense peaks (TFLOP/s): FP32 with packed FP32 instructions,
# and the vector rate without them (one FMA per lane per cycle); FP64 vector.
PEAK_FP32_TFLOPS = 157.3
PEAK_FP32_NON_PACKED_TFLOPS = 78.6
PEAK_FP64_TFLOPS = 78.6
# Kernel sources whose hash the counter summary under profiles/ records: the
# counters describe the benched kernel only when the hashes agree.
KERNEL_SOURCES = ("magritte_amd/csrc/mgr_radiation.hip", "magritte_amd/csrc/mgr_ray.hpp",
                  "magritte_amd/csrc/mgr_common.hpp")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(n, kind="envelope"):
    from magritte_amd.setup import synthetic
    if kind == "regular":
        return synthetic.regular_grid_model(n=n)
    if kind == "stress":
        return synthetic.stress_model(n=n)
    return synthetic.envelope_model(n=n)


def kernel_source_sha1():
    h = hashlib.sha1()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(ROOT, rel), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()


def profile_counters(path):
    """Counter summary of the radiation-field kernel committed under profiles/
    (HBM bytes per launch from FETCH_SIZE + WRITE_SIZE, copy-kernel scaled;
    VALU issue fraction from SQ_INSTS_VALU net of transcendentals)."""
    try:
        with open(path) as fh:
            return json.load(fh)
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(model, target_s, J_gpu=None, seed=0):
    """The FP64 oracle on a random sample of origins (all ray pairs and bins),
    in solves per second; with the GPU's J of the same model, also the largest
    relative difference of J over the sample."""
    from oracle.py_oracle import Oracle
    orc = Oracle(model).setup()
    rng = np.random.default_rng(seed)
    order = rng.permutation(model.ncells)
    ant = orc.antipod()
    F = model.nfreqs

    def solves(origins):
        s = 0
        for o in origins:
            for r in range(model.nrays // 2):
                nd = orc.npoints_on_ray(int(o), r) + orc.npoints_on_ray(int(o), int(ant[r])) + 1
                s += F if nd > 1 else 0
        return s

    n = 16
    while True:
        sample = order[:n]
        t0 = time.perf_counter()
        orc.compute_radiation_field_list(sample)
        dt = time.perf_counter() - t0
        if dt >= 0.5 * target_s or n >= model.ncells:
            break
        n = min(model.ncells, int(n * max(2.0, min(8.0, target_s / max(dt, 1e-3)))))
    from oracle import py_oracle
    cores = py_oracle.lib().orc_nthreads_max()
    out = dict(value=solves(sample) / dt, unit="ray-pair·freq solves/s", cores=int(cores),
               cpu_model=cpu_model(), kind="port",
               sample=f"{len(sample)} random origins x all {model.nrays // 2} ray pairs x {F} bins, "
                      f"FP64 oracle, {dt:.1f} s")
    if J_gpu is not None:
        J_o = orc.J()[sample]
        Jg = J_gpu[sample]
        out["J_max_rel_err_gpu_vs_oracle"] = float(np.max(np.abs(Jg - J_o) / np.abs(J_o)))
    return out


def cpu_time_to_converged(model, P_gpu=None):
    """The oracle's whole ALI from LTE (max relative change < 1e-6, Ng), on
    all host cores, and the largest relative difference of the GPU's
    converged populations from the oracle's (levels above 1e-10 of the
    cell's total)."""
    from oracle import py_oracle
    orc = py_oracle.Oracle(model).setup()
    t0 = time.perf_counter()
    it = orc.compute_level_populations(True, 400, 1)
    dt = time.perf_counter() - t0
    out = {"seconds": round(dt, 3), "iterations": int(it), "cores": int(py_oracle.lib().orc_nthreads_max()),
           "kind": "port"}
    if P_gpu is not None:
        P_o = orc.populations()   # (N, V): one species
        mask = P_o > 1e-10 * P_o.sum(axis=1, keepdims=True)
        out["pops_max_rel_diff_gpu_vs_oracle"] = float(np.max(np.abs(P_gpu - P_o)[mask] / P_o[mask]))
    return out


def copy_probe(gib):
    import torch
    n = int(gib * (1 << 30)) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    log(f"[bench] copy probe: {4 * n} bytes read, {4 * n} bytes written")
    del src, dst


def time_radiation_field(ctx, steps, warmup, precision):
    """Wall time of `steps` radiation fields after `warmup` untimed ones, and
    the kernel's own HIP-event time per launch."""
    for _ in range(warmup):
        ctx.compute_radiation_field(precision)
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(steps):
        ctx.compute_radiation_field(precision)
        kernel_ms.append(ctx.last_radiation_field_ms())
        log(f"[bench] step {i}: kernel {kernel_ms[-1]:.2f} ms")
    return time.perf_counter() - t0, float(np.mean(kernel_ms))


# Algorithmic matrix flops of one training row of the warm-start network
# (8 -> 128 -> 128 -> 7): forward 2 (8*128 + 128*128 + 128*7), backward
# through W3 and W2 2 (7*128 + 128*128), weight gradients 2 (8*128 + 128*128 + 128*7).
MLP_FLOPS_PER_ROW = 2 * (8 * 128 + 128 * 128 + 128 * 7) + 2 * (7 * 128 + 128 * 128) + 2 * (8 * 128 + 128 * 128 + 128 * 7)
PEAK_BF16_TFLOPS = 2500.0


def bench_mlp(args):
    """Training throughput of the warm-start network on synthetic rows."""
    from magritte_amd.hip import Mlp
    rng = np.random.default_rng(0)
    n = args.mlp_rows
    X = rng.normal(0, 1, (n, 8)).astype(np.float32)
    W = rng.normal(0, 0.5, (8, 7))
    Y = (np.tanh(X @ W) + 0.1 * X[:, :7] ** 2).astype(np.float32)
    net = Mlp(seed=0)
    net.set_normalisation(X.mean(0), X.std(0), Y.mean(0), Y.std(0))
    net.train(X[: args.mlp_batch * 4], Y[: args.mlp_batch * 4], max(1, args.warmup), args.mlp_batch, 4e-3, 0)
    t0 = time.perf_counter()
    losses = net.train(X, Y, args.steps, args.mlp_batch, 4e-3, 1)
    dt = time.perf_counter() - t0
    net.close()
    rows = args.steps * n
    tflops = rows * MLP_FLOPS_PER_ROW / dt / 1e12
    out = {
        "metric": "warm-start network training rows/s", "value": rows / dt, "unit": "rows/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1e3 * dt / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": f"MLP 8-128-128-7 GELU, {n} rows per epoch, batch {args.mlp_batch}, Adam",
                   "epochs": args.steps, "loss_first_last": [float(losses[0]), float(losses[-1])]},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 2), "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(tflops / PEAK_BF16_TFLOPS, 5), "traffic": None,
                     "flops_per_row": MLP_FLOPS_PER_ROW},
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def roofline_entry(sum_ndep, nfreqs, solves, k_ms, precision, kernel, counters_workload):
    points = sum_ndep * nfreqs                 # sum over solves of their points
    equiv_bytes = BYTES_PER_POINT_BIN * points + BYTES_PER_SOLVE * solves
    flops = FLOPS_PER_POINT_BIN * points
    tflops = flops / (k_ms * 1e-3) / 1e12
    prof = profile_counters(os.path.join(ROOT, "profiles", "radiation_field_envelope64_counters.json"))
    sha = kernel_source_sha1()
    same = (prof is not None and prof.get("workload") == counters_workload and prof.get("kernel") == kernel
            and prof.get("source_sha1") == sha)
    traffic = prof.get("hbm_bytes_per_launch") if same else None
    fp32 = precision == "fp32"
    peak = PEAK_FP32_TFLOPS if fp32 else PEAK_FP64_TFLOPS
    out = {
        "bound": "valu",
        "kernel": kernel,
        "achieved": round(tflops, 2), "peak": peak, "unit": "TFLOP/s",
        "frac": round(tflops / peak, 4),
        "traffic": traffic,
        "kernel_ms": round(k_ms, 3),
        "flops_per_point_bin": FLOPS_PER_POINT_BIN,
        "hbm": ({"achieved": round(traffic / (k_ms * 1e-3) / 1e9, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                 "frac": round(traffic / (k_ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4)} if traffic else None),
        "valu_issue_frac": prof.get("valu_issue_frac") if same else None,
        "counters_source_sha1": prof.get("source_sha1") if prof else None,
        "kernel_source_sha1": sha,
        "equivalent_bytes": {"per_launch": int(equiv_bytes),
                             "rate_GBs": round(equiv_bytes / (k_ms * 1e-3) / 1e9, 1),
                             "note": "dtau and S per point and bin, u and Lambda* per solve, as an HBM-staged "
                                     "solver would move them; this fused kernel never moves these bytes"},
    }
    if fp32:
        out["peak_non_packed"] = PEAK_FP32_NON_PACKED_TFLOPS
        out["frac_non_packed"] = round(tflops / PEAK_FP32_NON_PACKED_TFLOPS, 4)
    return out


def batch_summary(ctx, pop_prec=1e-6):
    """Models of a batch that met the criterion, and their iteration counts."""
    its, conv = ctx.convergence()
    done = np.array([c[2] < pop_prec for c in conv])
    out = {"models_converged": int(done.sum()), "models": int(len(done))}
    if done.any():
        out["iterations_of_converged_models"] = {"median": float(np.median(its[done])), "max": int(its[done].max())}
    return out


def converged_cells(ctx, pop_prec=1e-6):
    """Cell mask of the batch models that met the criterion."""
    _, conv = ctx.convergence()
    mask = np.zeros(ctx.N, bool)
    for k, c in enumerate(conv):
        if c[2] < pop_prec:
            mask[ctx.cell_offsets[k]:ctx.cell_offsets[k + 1]] = True
    return mask


def time_to_converged(make_ctx, precision, weights, max_iter=400):
    """ALI from LTE, and from the learned warm start when trained weights are
    present; the clock starts after the one-off set-up and trace.  For a batch
    the call runs until every model has converged or max_iter; the models
    that converged are counted."""
    from magritte_amd.hip import BatchContext
    out = {"criterion": "max relative change < 1e-6, Ng every 5th", "max_iterations": max_iter,
           "clock": "after the per-model set-up (trace, masks, far-end cache, tuning, graphs)"}
    ctx = make_ctx()
    ctx.prepare(precision)
    t0 = time.perf_counter()
    it = ctx.compute_level_populations(True, max_iter, "max", precision)
    out["lte"] = {"seconds": round(time.perf_counter() - t0, 3), "iterations": int(it)}
    cells = np.ones(ctx.N, bool)
    if isinstance(ctx, BatchContext):
        out["lte"].update(batch_summary(ctx))
        cells = converged_cells(ctx)
    P_lte = ctx.populations()
    out["_populations"] = P_lte
    ctx.close()
    log(f"[bench] converged from LTE in {it} iterations, {out['lte']['seconds']:.2f} s")
    if weights and os.path.exists(weights):
        from magritte_amd import warmstart
        net = warmstart.load(weights)
        ctx = make_ctx()
        ctx.prepare(precision)
        t0 = time.perf_counter()
        if net.first_pass_stride > 0:
            ctx.first_pass(precision, net.first_pass_stride)
        fallbacks = ctx.warm_start(net)
        it = ctx.compute_level_populations(True, max_iter, "max", precision)
        dt = time.perf_counter() - t0
        P_nn = ctx.populations()
        summary = {}
        if isinstance(ctx, BatchContext):
            summary = batch_summary(ctx)
            cells &= converged_cells(ctx)   # the models that converged from both starts
        ctx.close()
        net.close()
        mask = (P_lte > 1e-10 * P_lte.sum(axis=1, keepdims=True)) & cells[:, None]
        diff = float(np.max(np.abs(P_nn - P_lte)[mask] / P_lte[mask])) if mask.any() else None
        out["nn"] = {"seconds": round(dt, 3), "iterations": int(it), "pops_max_rel_diff_vs_lte": diff,
                     "first_pass_ray_pair_stride": net.first_pass_stride, "cells_kept_lte": int(fallbacks),
                     **summary}
        out["speedup_nn_vs_lte"] = round(out["lte"]["seconds"] / dt, 3)
        log(f"[bench] converged from the warm start in {it} iterations, {dt:.2f} s")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="envelope", choices=["envelope", "regular", "stress", "batch", "mlp"],
                    help="envelope: one n^3 model; regular: the 16^3 regular grid; stress: the optically "
                         "thick envelope; batch: many small models in one context; "
                         "mlp: training of the warm-start network (a step is one epoch)")
    ap.add_argument("--mlp-rows", type=int, default=2_000_000)
    ap.add_argument("--mlp-batch", type=int, default=16384)
    ap.add_argument("--n", type=int, default=None,
                    help="cells per side (default 64 for the envelopes = full workload, 16 for the regular grid)")
    ap.add_argument("--models", type=int, default=512, help="models of the batch workload")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--fp64-steps", type=int, default=2, help="timed FP64 radiation fields beside the FP32 ones")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--max-iter", type=int, default=None,
                    help="ALI iteration cap (default 400; 60 for the batch, whose low-density models do not converge)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "magritte_amd", "data", "warmstart_mlp.npz"))
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="first run a plain device copy of this size (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        return 2
    if args.copy_probe_gib > 0:
        copy_probe(args.copy_probe_gib)
    if args.workload == "mlp":
        return bench_mlp(args)

    from magritte_amd.hip import BatchContext, Context
    t0 = time.perf_counter()
    single = args.workload in ("envelope", "regular", "stress")
    if args.n is None:
        args.n = 16 if args.workload == "regular" else 64
    if single:
        model = build_workload(args.n, args.workload)
        models = [model]
        make_ctx = lambda: Context(model).setup()   # noqa: E731
        if args.workload == "regular":
            workload = (f"regular grid {args.n}^3, 6-neighbour lists, {model.nrays} rays, 2-level species, "
                        f"{model.nfreqs} bins, n_H2 1e12, 20 K")
        else:
            workload = (f"{'stress ' if args.workload == 'stress' else ''}envelope point cloud {args.n}^3 "
                        f"jittered, 26-neighbour lists, {model.nrays} rays, 7-level rotor, 6 lines x 31 bins"
                        + (", n0 1e17, 20% +-30 km/s jumps" if args.workload == "stress" else ""))
        tag = f"{args.workload}{args.n}"
    else:
        from magritte_amd.warmstart import training_model
        rng = np.random.default_rng(2024)
        models = [training_model(rng) for _ in range(args.models)]
        model = models[0]
        make_ctx = lambda: BatchContext(models).setup()   # noqa: E731
        workload = (f"batch of {args.models} envelopes, cells log-uniform 1000-32768, n0 1e12-1e16, "
                    f"T0 500-3000 K, 12 or 48 rays, one context")
        tag = f"batch{args.models}"
    ncells = sum(m.ncells for m in models)
    log(f"[bench] {len(models)} model(s), {ncells} cells, {model.nfreqs} bins ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    ctx = make_ctx()
    ctx.trace()
    sum_ndep, pairs = ctx.trace_stats()
    log(f"[bench] set-up + trace {time.perf_counter() - t0:.1f} s; {pairs} ray pairs solved, sum ndep {sum_ndep}")
    solves_per_step = pairs * model.nfreqs
    wall, k_ms = time_radiation_field(ctx, args.steps, args.warmup, args.precision)
    kernel = ctx.last_radiation_field_kernel()
    shape, tried = ctx.radiation_field_shape(args.precision)
    log(f"[bench] {kernel}: launch shape (waves per workgroup, steps per chunk) {shape}; tried {tried}")
    ms_per_step = 1e3 * wall / args.steps
    value = solves_per_step * args.steps / wall
    J_gpu = ctx.J() if single and not args.no_cpu else None
    other = None
    if args.precision == "fp32" and args.fp64_steps > 0:
        # the FP64 kernel on the same workload, for comparison
        w64, k64 = time_radiation_field(ctx, args.fp64_steps, 1, "fp64")
        other = {"precision": "fp64", "kernel": ctx.last_radiation_field_kernel(), "kernel_ms": round(k64, 3),
                 "ms_per_step": round(1e3 * w64 / args.fp64_steps, 3),
                 "solves_per_s": solves_per_step * args.fp64_steps / w64}
        log(f"[bench] FP64 kernel {k64:.1f} ms")

    out = {
        "metric": METRIC, "value": value, "unit": "ray-pair·freq solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": workload, "models": len(models), "ncells": ncells,
                   "nrays": max(m.nrays for m in models), "nfreqs": model.nfreqs,
                   "solves_per_step": int(solves_per_step), "sum_ndep": int(sum_ndep),
                   "launch_shape": {"waves_per_workgroup": shape[0], "steps_per_chunk": shape[1]}},
        "roofline": roofline_entry(sum_ndep, model.nfreqs, solves_per_step, k_ms, args.precision, kernel, tag),
    }
    if other:
        out["fp64_kernel"] = other
    ctx.close()

    P_gpu = None
    if not args.no_converge:
        # the network predicts the 7 levels of the rotor: no warm start for the 2-level grid
        weights = None if args.workload == "regular" else args.weights
        max_iter = args.max_iter if args.max_iter else (60 if args.workload == "batch" else 400)
        out["time_to_converged"] = time_to_converged(make_ctx_traced(make_ctx), args.precision, weights, max_iter)
        P_gpu = out["time_to_converged"].pop("_populations")

    if not args.no_cpu and single:
        log("[bench] CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(model, args.cpu_seconds, J_gpu)
        log(f"[bench] GPU J against the oracle on the sample: max rel err "
            f"{out['cpu_baseline']['J_max_rel_err_gpu_vs_oracle']:.3e}")
        if args.workload == "regular":
            # small enough for the oracle's whole ALI
            out["cpu_baseline"]["time_to_converged"] = cpu_time_to_converged(model, P_gpu)
            log(f"[bench] oracle ALI: {out['cpu_baseline']['time_to_converged']}")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


def make_ctx_traced(make_ctx):
    def f():
        c = make_ctx()
        c.trace()
        return c
    return f


if __name__ == "__main__":
    sys.exit(main())
