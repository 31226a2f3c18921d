"""Benchmark of the radiation-field hot path on one MI355X.

Workload (the at-scale single-GPU configuration): 262,144-cell jittered 64^3
point cloud with 26 grid-neighbour lists, 48 HEALPix rays (nside 2),
spherical envelope n_H2 = 1e14 (r0/r)^2 m^-3, T = 2500 (r0/r)^0.6 K, radial
outflow 15 (1 - R*/r)^0.5 km/s, 7-level linear rotor, 6 lines x 31 bins.

A step is one compute_radiation_field over the whole model (every ray pair x
every frequency bin), inputs resident in HBM.  value = ray-pair.freq solves
per second over the timed steps.  The JSON line also carries:
  roofline      -- the radiation-field kernel, timed with HIP events on its
                   own stream: algorithmic bytes per launch (8 B per point and
                   bin for dtau and S, 8 B per solve for J and Lambda*) over the
                   kernel time, against the HBM peak.  The kernel fuses the
                   line profiles and the Feautrier solve, so those bytes never
                   reach HBM ("traffic" holds the measured HBM bytes per
                   launch); its limiter is VALU issue, reported beside it;
  cpu_baseline  -- the FP64 OpenMP oracle (restatement of the reference's
                   RayPair path) timed on a bounded random sample of origins;
  time_to_converged -- the ALI loop from LTE to max relative change < 1e-6.
"""
import argparse
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def build_workload(n):
    from magritte_amd.setup import synthetic
    return synthetic.envelope_model(n=n)


def profile_counters(path):
    """Counter summary of the radiation-field kernel committed under profiles/
    (HBM bytes per launch from FETCH_SIZE + WRITE_SIZE, copy-kernel scaled;
    VALU issue fraction from SQ_INSTS_VALU)."""
    try:
        with open(path) as fh:
            d = json.load(fh)
        return d, d.get("workload")
    except (OSError, ValueError):
        return None, None


def cpu_baseline(model, target_s, seed=0):
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
    return dict(value=solves(sample) / dt, unit="ray-pair·freq solves/s", cores=int(cores),
                kind="port",
                sample=f"{len(sample)} random origins x all {model.nrays // 2} ray pairs x {F} bins, "
                       f"FP64 oracle, {dt:.1f} s")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=64, help="cells per side (64 = full workload)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-converge", action="store_true")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="first run a plain device copy of this size (counter calibration)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        return 2

    if args.copy_probe_gib > 0:
        import torch
        n = int(args.copy_probe_gib * (1 << 30)) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy probe: {4 * n} bytes read, {4 * n} bytes written")
        del src, dst

    from magritte_amd.hip import Context
    t0 = time.perf_counter()
    model = build_workload(args.n)
    log(f"[bench] model: {model.ncells} cells, {model.nrays} rays, {model.nfreqs} bins "
        f"({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    ctx = Context(model).setup()
    ctx.trace()
    sum_ndep, pairs = ctx.trace_stats()
    log(f"[bench] set-up + trace {time.perf_counter() - t0:.1f} s; {pairs} ray pairs solved, "
        f"mean ndep {sum_ndep / (model.ncells * model.nrays // 2):.1f}")
    solves_per_step = pairs * model.nfreqs

    for _ in range(args.warmup):
        ctx.compute_radiation_field(args.precision)
    kernel_ms = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ctx.compute_radiation_field(args.precision)
        kernel_ms.append(ctx.last_radiation_field_ms())
        log(f"[bench] step {i}: kernel {kernel_ms[-1]:.2f} ms")
    wall = time.perf_counter() - t0
    ms_per_step = 1e3 * wall / args.steps
    value = solves_per_step * args.steps / wall
    k_ms = float(np.mean(kernel_ms))
    points = sum_ndep * model.nfreqs                 # sum over solves of their points
    alg_bytes = BYTES_PER_POINT_BIN * points + BYTES_PER_SOLVE * solves_per_step
    flops = FLOPS_PER_POINT_BIN * points
    achieved = alg_bytes / (k_ms * 1e-3) / 1e9
    prof, pw = profile_counters(os.path.join(ROOT, "profiles", "radiation_field_envelope64_counters.json"))
    same = prof is not None and pw == f"envelope{args.n}" and args.precision == "fp32"
    roofline = {
        "bound": "hbm",
        "kernel": "k_radiation_field<float>" if args.precision == "fp32" else "k_radiation_field<double>",
        "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
        "frac": round(achieved / PEAK_HBM_GBS, 4),
        "traffic": prof.get("hbm_bytes_per_launch") if same else None,
        "kernel_ms": round(k_ms, 3),
        "algorithmic_bytes_per_launch": int(alg_bytes),
        "limiter": "valu",
        "valu_issue_frac": prof.get("valu_issue_frac") if same else None,
        "fp32_tflops": round(flops / (k_ms * 1e-3) / 1e12, 3),
    }

    out = {
        "metric": METRIC, "value": value, "unit": "ray-pair·freq solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_per_step,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32" if args.precision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"envelope point cloud {args.n}^3 jittered, 26-neighbour lists, "
                               f"{model.nrays} rays, 7-level rotor, 6 lines x 31 bins",
                   "ncells": model.ncells, "nrays": model.nrays, "nfreqs": model.nfreqs,
                   "solves_per_step": int(solves_per_step), "sum_ndep": int(sum_ndep)},
        "roofline": roofline,
    }

    if not args.no_converge:
        ctx2 = Context(model).setup()
        ctx2.trace()
        t0 = time.perf_counter()
        it = ctx2.compute_level_populations(True, 100, "max", args.precision)
        tc = time.perf_counter() - t0
        out["time_to_converged"] = {"start": "LTE", "seconds": round(tc, 3), "iterations": int(it),
                                    "criterion": "max relative change < 1e-6, Ng every 5th"}
        log(f"[bench] converged from LTE in {it} iterations, {tc:.2f} s")
        ctx2.close()

    if not args.no_cpu:
        log("[bench] CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(model, args.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    ctx.close()
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
