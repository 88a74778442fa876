"""Benchmark of the pulse-profile path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload plain|batched]

Prints one JSON line.  A step is one pass of the hot path over one batch of
synthetic input with inputs and outputs resident in HBM; the timed region is
K steps between two device synchronisations.

Workloads (the configurations the headline metric is quoted on):
  plain    one oblate-star profile (single 0.25 rad cap at 20 deg, NSX-format
           atmosphere, kT 0.35 keV) at N_side 1024, 64 phase bins x 2 GL
           nodes, 360 energies, folded through a 300 x 360 response; one
           profile per launch sequence, as the reference's calc_one_star.
  batched  256 stars drawn from the batched-evaluation priors at the same
           resolution in one persistent launch (when built).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "profile evals/s, batched vs plain GPU, at ≤1e-4 max rel err; p50/p99 latency"
FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 peak
HBM_PEAK_GBS = 8000.0
FLOP_PER_EVAL = 32         # 4 x 4 (logEkT x mu) Lagrange stencil = 16 FMA


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def data_root():
    from gpu_ppm_amd import fixtures
    root = os.path.join(ROOT, "tests", "_data", "bench")
    fixtures.ensure_all(root, small_nsx=False, full_nsx=True, lensing=("std", "wide"), responses=("std",))
    return root


def cpu_baseline(root, star, grid, target_s=15.0):
    """The FP64 oracle (a C port of the reference CPU path) on a bounded
    sample of the same workload: all rings and phase bins, a subset of the
    energies; the per-profile time is scaled by the energy ratio."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from gpu_ppm_amd import abi
    from gpu_ppm_amd import scenarios as S
    from oracle_binding import Oracle
    threads = max(1, min(16, os.cpu_count() or 1))
    orc = Oracle(root, nsx_path=os.path.join(root, "ext", "model_data", "nsx_H_v200804.out"), responses=("std",))
    energies = S.geomspace(0.2, 12.0, 360)

    def run(n_e):
        idx = [int(round(k * 359 / max(n_e - 1, 1))) for k in range(n_e)]
        g = abi.make_grid(grid.n_side, grid.n_phase, n_fine=grid.n_fine, sampling=grid.sampling,
                          n_gl=grid.n_gl, energies=[energies[i] for i in idx])
        t0 = time.perf_counter()
        orc.profile(star, g, threads=threads)
        return time.perf_counter() - t0

    t2 = run(2)
    n_e = max(2, min(360, int(2 * target_s / max(t2, 1e-3))))
    t = run(n_e)
    orc.close()
    per_profile = t * 360.0 / n_e
    return {"value": 1.0 / per_profile, "unit": "profiles/s", "cores": threads, "kind": "port",
            "sample": "FP64 oracle, same star and grid, all rings and bins, %d of 360 energies "
                      "in %.1f s on %d threads; scaled by 360/%d" % (n_e, t, threads, n_e)}


def bench_plain(root, steps, warmup):
    from gpu_ppm_amd import abi, hip
    from gpu_ppm_amd import scenarios as S
    from gpu_ppm_amd.lib import Context
    star = S.cfg1_star()
    grid = abi.make_grid(1024, 64, sampling=abi.SAMPLE_GL, n_gl=2, response=0, method=abi.METHOD_PLAIN)
    ctx = Context(root, lensing="std", responses=("std",))
    n_ch, e = ctx.response_info(0)
    d_flux = hip.DeviceArray((1, 64, e.size))
    d_cnt = hip.DeviceArray((1, n_ch, 64))
    hip.synchronize()
    for _ in range(warmup):
        ctx.eval_device(star, grid, d_flux.ptr, d_cnt.ptr)
    hip.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(ctx.eval_device(star, grid, d_flux.ptr, d_cnt.ptr))
    hip.synchronize()
    dt = time.perf_counter() - t0
    integral_ms = sum(s.integral_ms for s in stats) / steps
    n_eval = stats[0].n_eval_done
    ctx.close()
    return dict(star=star, grid=grid, dt=dt, profiles=steps, integral_ms=integral_ms, n_eval=n_eval,
                n_eval_ref=stats[0].n_eval_ref, n_active=stats[0].n_active, kernel="k_integral_plain",
                kernel_ms=sum(s.kernel_ms for s in stats) / steps, fold_ms=sum(s.fold_ms for s in stats) / steps)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="plain", choices=["plain"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (profiles shard by "
              "star or ring range with one small reduction, see the design notes); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    root = data_root()
    log("bench: data ready; workload", args.workload)
    r = bench_plain(root, args.steps, args.warmup)
    log("bench: timed %d steps in %.3f s" % (args.steps, r["dt"]))
    achieved = r["n_eval"] * FLOP_PER_EVAL / (r["integral_ms"] * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": r["profiles"] / r["dt"],
        "unit": "profiles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["dt"] / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain single-profile: N_side 1024, 64 bins x 2 GL, 360 E, NSX, fold 300x360",
                   "n_side": 1024, "n_phase": 64, "n_energy": 360, "n_active": r["n_active"],
                   "method": "plain"},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS, "traffic": None, "kernel": r["kernel"],
                     "kernel_ms": r["integral_ms"], "roof": "FP32 vector (no packed math)",
                     "units": "%d evaluations x %d FLOP" % (r["n_eval"], FLOP_PER_EVAL)},
        "plain": {"profiles_per_s": r["profiles"] / r["dt"], "kernel_ms": r["kernel_ms"],
                  "integral_ms": r["integral_ms"], "fold_ms": r["fold_ms"],
                  "n_eval_ref": r["n_eval_ref"], "evals_per_s": r["n_eval"] / (r["integral_ms"] * 1e-3)},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(root, r["star"], r["grid"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
