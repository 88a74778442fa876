"""Benchmark of the pulse-profile path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload batched|plain]

Prints one JSON line.  A step is one pass of the hot path over one batch of
synthetic input with tables, inputs and outputs resident in HBM; the timed
region is K steps between two device synchronisations.

Workloads (the configurations the headline metric is quoted on):
  batched (default)  256 stars drawn (seed 0) from the batched-evaluation
           priors: M 1.0-2.2 Msun, R_eq 9-15 km, spin 200-700 Hz,
           inclination 5-90 deg, one cap of colatitude 0-180 deg and radius
           0.02-1.2 rad, kT 0.1-0.5 keV; NSX-format atmosphere; N_side 1024,
           64 phase bins x 2 GL nodes, 360 energies, 300 x 360 response fold.
           One step = the 256 profiles in one persistent launch (hipGraph).
           The plain path is timed on the same stars, one profile per launch.
  plain    one star (0.25 rad cap at 20 deg, M 1.4, R_eq 12 km, 600 Hz,
           i 60 deg, kT 0.35 keV) at the same resolution, one profile per step.

The batched workload is followed by the serving leg (bench_serving: p50/p99
request latency and sustained evaluations under open-loop arrivals) and the
CPU baseline leg (the FP64 oracle on a bounded sample).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "profile evals/s, batched vs plain GPU, at ≤1e-4 max rel err; p50/p99 latency"
FP32_PEAK_TFLOPS = 157.3   # MI355X published FP32 peak
FLOP_PER_EVAL = 32         # 4 x 4 (logEkT x mu) Lagrange stencil = 16 FMA
SERVE_FIXED_RATE = 37.0    # requests/s of the fixed-load serving stream
N_SIDE, N_PHASE, N_E = 1024, 64, 360


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def data_root():
    from gpu_ppm_amd import fixtures
    root = os.path.join(ROOT, "tests", "_data", "bench")
    fixtures.ensure_all(root, small_nsx=False, full_nsx=True, lensing=("std", "wide"), responses=("std", "std128"))
    return root


def grid(method):
    from gpu_ppm_amd import abi
    return abi.make_grid(N_SIDE, N_PHASE, sampling=abi.SAMPLE_GL, n_gl=2, response=0, method=method)


def host_cores():
    """CPU threads this job may use: the affinity mask, limited by the cgroup
    CPU quota when one is set (a container's share of a larger host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        q, per = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if q != "max":
            n = min(n, max(1, int(int(q) // int(per))))
    except (OSError, ValueError):
        pass
    return n


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def reference_os1_cpu(root):
    """The reference's own CPU program, cpu/os1_cpu.cpp built unmodified from
    its sources (oracle/_ref/os1_cpu), on its ten hard-coded os1 scenarios
    (N_side 256, 128 phase points, blackbody at 1 keV), timed by process wall
    clock (the program has no internal timer), one thread."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "oracle", "_ref", "os1_cpu")
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as d:
        os.makedirs(os.path.join(d, "tables", "lensing"))
        os.makedirs(os.path.join(d, "output"))
        os.symlink(os.path.join(root, "tables", "lensing", "std"), os.path.join(d, "tables", "lensing", "std"))
        t0 = time.perf_counter()
        r = subprocess.run([exe], cwd=d, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL)
        dt = time.perf_counter() - t0
    if r.returncode != 0:
        return None
    return {"kind": "reference", "s_per_profile": dt / 10, "profiles": 10, "cores": 1,
            "what": "unmodified cpu/os1_cpu.cpp (reference flags minus -march=native), os1 a-j: N_side 256, "
                    "128 phase points, blackbody at 1 keV; process wall time / 10"}


def cpu_config0(root, threads):
    """The FP64 oracle on config 0 at full size (cap 0.25 rad at 20 deg,
    blackbody kT 0.35 keV, N_side 128, 32 bins x 2 GL, 128 energies, fold
    300 x 128), on 1 thread and on `threads`, the reference's timing
    boundaries (integral + fold; tables loaded and the grid set up outside)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from gpu_ppm_amd import abi
    from gpu_ppm_amd import scenarios as S
    from oracle_binding import Oracle
    orc = Oracle(root, lensing="std", responses=("std", "std128"))
    star = S.cfg0_star()
    g = abi.make_grid(128, 32, sampling=abi.SAMPLE_GL, n_gl=2, response=1)
    out = {"config": "cap 0.25 rad at 20 deg, blackbody kT 0.35 keV, N_side 128, 32 bins x 2 GL, 128 E, fold 300x128"}
    for key, th in (("1_thread", 1), ("all_cores", threads)):
        best = 1e300
        for _ in range(3):
            t0 = time.perf_counter()
            f, info = orc.profile(star, g, threads=th)
            orc.fold(f, 1, star.exposure_s / g.n_phase)
            best = min(best, time.perf_counter() - t0)
        out[key] = {"threads": th, "s_per_profile": best, "evals_per_s": info.n_eval_ref / best}
    out["n_eval_ref"] = info.n_eval_ref
    orc.close()
    return out


def cpu_baseline(root, stars, lensing, batched_flux, target_s=20.0):
    """The FP64 oracle (a C port of the reference CPU path) on a bounded
    sample of the same workload: the first star(s), all rings and phase bins,
    a subset of the energies; the per-profile time is scaled by the energy
    ratio.  The oracle's columns are compared with the same columns of the
    batched and the plain device results.  Also: the oracle on config 0 at
    full size (1 thread and all cores) and the reference's own os1_cpu."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from gpu_ppm_amd import abi, hip
    from gpu_ppm_amd import scenarios as S
    from gpu_ppm_amd.lib import Context
    from oracle_binding import Oracle
    threads = host_cores()
    orc = Oracle(root, lensing=lensing, nsx_path=os.path.join(root, "ext", "model_data", "nsx_H_v200804.out"),
                 responses=("std",))
    energies = S.geomspace(0.2, 12.0, N_E)

    def columns(n_e):
        return [int(round(k * (N_E - 1) / max(n_e - 1, 1))) for k in range(n_e)]

    def run(n_e, sts):
        idx = columns(n_e)
        g = abi.make_grid(N_SIDE, N_PHASE, sampling=abi.SAMPLE_GL, n_gl=2, energies=[energies[i] for i in idx])
        t0 = time.perf_counter()
        out = [orc.profile(st, g, threads=threads)[0] for st in sts]
        return time.perf_counter() - t0, out

    sample = stars[:2]
    t2, _ = run(2, sample)
    n_e = max(2, min(N_E, int(2 * target_s / max(t2, 1e-3))))
    t, ref = run(n_e, sample)
    orc.close()
    idx = columns(n_e)
    # the same columns of the device results: the batched launch of the timed
    # workload, and the plain path on the same stars
    gp = grid(abi.METHOD_PLAIN)
    plain_ctx = Context(root, lensing=lensing, responses=("std",))
    n_ch = plain_ctx.response_info(0)[0]
    d_f = hip.DeviceArray((1, N_PHASE, N_E))
    d_c = hip.DeviceArray((1, n_ch, N_PHASE))
    err = {"batched_flux": 0.0, "plain_flux": 0.0, "stars": len(sample), "energy_columns": idx}
    for k, st in enumerate(sample):
        if batched_flux is not None:
            err["batched_flux"] = max(err["batched_flux"], rel_err(batched_flux[k][:, idx], ref[k]))
        plain_ctx.eval_device([st], gp, d_f.ptr, d_c.ptr)
        hip.synchronize()
        err["plain_flux"] = max(err["plain_flux"], rel_err(d_f.numpy()[0][:, idx], ref[k]))
    plain_ctx.close()
    log("bench: device vs FP64 oracle columns %s" % err)
    per_profile = t * N_E / n_e / len(sample)
    out = {"value": 1.0 / per_profile, "unit": "profiles/s", "cores": threads, "kind": "port",
           "cpu_model": cpu_model(),
           "sample": "FP64 oracle on %d of the workload's stars, all rings and bins, %d of %d energies, "
                     "%.1f s on %d threads; scaled by %d/%d" % (len(sample), n_e, N_E, t, threads, N_E, n_e),
           "max_rel_err_vs_fp64_oracle": err}
    log("bench: CPU config 0")
    out["config0"] = cpu_config0(root, threads)
    ref_run = reference_os1_cpu(root)
    if ref_run:
        out["reference_os1_cpu"] = ref_run
    return out


def rel_err(x, ref, floor=1e-6):
    """max |x - ref| / |ref| over the entries with |ref| > floor * max|ref|."""
    import numpy as np
    x, ref = np.asarray(x, dtype=np.float64), np.asarray(ref, dtype=np.float64)
    m = np.abs(ref) > floor * np.abs(ref).max()
    return float((np.abs(x[m] - ref[m]) / np.abs(ref[m])).max()) if m.any() else 0.0


def timed(ctx, stars, g, d_flux, d_cnt, steps, warmup):
    from gpu_ppm_amd import hip
    for _ in range(warmup):
        ctx.eval_device(stars, g, d_flux.ptr, d_cnt.ptr)
    hip.synchronize()
    stats = []
    t0 = time.perf_counter()
    for _ in range(steps):
        stats.append(ctx.eval_device(stars, g, d_flux.ptr, d_cnt.ptr))
    hip.synchronize()
    return time.perf_counter() - t0, stats


def bench_batched(root, steps, warmup, n_stars, plain_stars, tune=True):
    from gpu_ppm_amd import abi, hip
    from gpu_ppm_amd import scenarios as S
    from gpu_ppm_amd.lib import Context
    stars = S.sample_priors(n_stars, seed=0)
    ctx = Context(root, lensing="wide", responses=("std",))
    n_ch, e = ctx.response_info(0)
    d_flux = hip.DeviceArray((n_stars, N_PHASE, N_E))
    d_cnt = hip.DeviceArray((n_stars, n_ch, N_PHASE))
    gb = grid(abi.METHOD_BATCHED)
    shape = None
    if tune:  # the launch shape the autotuner keeps for this grid shape (outside the timed region)
        log("bench: autotuning the batched launch shape on the %d stars" % n_stars)
        tr = ctx.tune(stars, gb, repeats=1)
        shape = {"items_per_slot": tr.best.items_per_slot, "wg_per_cu": tr.best.wg_per_cu,
                 "energies_per_lane": tr.best.energies_per_lane, "tuned_integral_ms": tr.best_ms, "default_integral_ms": tr.default_ms, "shapes_tried": tr.n_tried}
        log("bench: tuned shape %s" % shape)
    log("bench: batched %d stars, warmup %d, steps %d" % (n_stars, warmup, steps))
    dt, st = timed(ctx, stars, gb, d_flux, d_cnt, steps, warmup)
    log("bench: batched %.3f s per step" % (dt / steps))
    ctx.close()
    # per-kernel time of the batched integral: one direct (non-graph) launch
    os.environ["PPM_NO_GRAPH"] = "1"
    ctx = Context(root, lensing="wide", responses=("std",))
    if shape:
        ctx.set_launch_shape(gb, n_stars, abi.ATM_NSX, shape["items_per_slot"], shape["wg_per_cu"],
                             shape["energies_per_lane"])
    _, sk = timed(ctx, stars, gb, d_flux, d_cnt, 1, 0)
    # plain path on the same stars, one profile per launch sequence
    gp = grid(abi.METHOD_PLAIN)
    sub = stars[:plain_stars]
    d_f1 = hip.DeviceArray((1, N_PHASE, N_E))
    d_c1 = hip.DeviceArray((1, n_ch, N_PHASE))
    ctx.eval_device([sub[0]], gp, d_f1.ptr, d_c1.ptr)
    hip.synchronize()
    t0 = time.perf_counter()
    ref_evals, plain_int_ms = 0, 0.0
    for s in sub:
        ps = ctx.eval_device([s], gp, d_f1.ptr, d_c1.ptr)
        ref_evals += ps.n_eval_ref
        plain_int_ms += ps.integral_ms
    hip.synchronize()
    t_plain = time.perf_counter() - t0
    log("bench: plain %d stars in %.2f s" % (len(sub), t_plain))
    # the batched path (phase-shift dedupe) on one profile per launch, the
    # same stars: separates the dedupe's gain from the batching's
    gs = grid(abi.METHOD_BATCHED)
    ctx.eval_device([sub[0]], gs, d_f1.ptr, d_c1.ptr)
    hip.synchronize()
    t0 = time.perf_counter()
    single_int_ms = 0.0
    for s in sub:
        single_int_ms += ctx.eval_device([s], gs, d_f1.ptr, d_c1.ptr).integral_ms
    hip.synchronize()
    t_single = time.perf_counter() - t0
    log("bench: dedupe, one profile per launch: %d stars in %.2f s" % (len(sub), t_single))
    # outside the timed regions: the batched results of the direct launch
    # against the plain path on every 8th timed star
    fb, cb = d_flux.numpy(), d_cnt.numpy()
    checked = list(range(0, len(sub), 8))
    err = {"flux": 0.0, "counts": 0.0, "stars_checked": len(checked)}
    for i in checked:
        ctx.eval_device([stars[i]], gp, d_f1.ptr, d_c1.ptr)
        hip.synchronize()
        err["flux"] = max(err["flux"], rel_err(fb[i], d_f1.numpy()[0]))
        err["counts"] = max(err["counts"], rel_err(cb[i], d_c1.numpy()[0]))
    log("bench: batched vs plain max rel err %s" % err)
    ctx.close()
    del os.environ["PPM_NO_GRAPH"]
    k = sk[0]
    return dict(stars=stars, lensing="wide", value=n_stars * steps / dt, dt=dt,
                kernel="k_integral_batched", kernel_ms=k.integral_ms, n_eval=k.n_eval_done,
                n_eval_ref=k.n_eval_ref, n_active=k.n_active,
                plain_profiles_per_s=len(sub) / t_plain, plain_stars=len(sub),
                plain_evals_per_s=ref_evals / (plain_int_ms * 1e-3), plain_integral_ms=plain_int_ms / len(sub),
                single_profiles_per_s=len(sub) / t_single, single_integral_ms=single_int_ms / len(sub),
                batched_flux=fb[:2].copy(),
                graph_ms=sum(s.total_ms for s in st) / steps, setup_reduce_fold_ms=k.total_ms - k.integral_ms,
                launch_shape=shape, max_rel_err=err)


def bench_serving(root, duration_s, burst, seed=7):
    """Continuous-batching serving loop under open-loop Poisson arrivals.

    Requests: a seeded reflecting random walk over the batched priors (2 %
    of each range per step), resolution drawn uniformly from N_side
    {256, 1024, 2048} x {32, 64, 128} bins (2 GL nodes) x {128, 360}
    energies (folded through the 300 x 128 or 300 x 360 response).  The
    capacity is first measured on a burst of `burst` requests submitted at
    once; the arrival rate is then 80 % of it for `duration_s` seconds, and
    a second stream arrives at the fixed rate SERVE_FIXED_RATE (latency at a
    load that does not move with the capacity).
    Latency is submit -> result on the host, reported by ppm_poll.  Every
    100th request of the timed stream is audited afterwards against a direct
    plain-path evaluation of the same star and grid, and against the FP64
    oracle on one energy column."""
    import numpy as np
    from gpu_ppm_amd import abi
    from gpu_ppm_amd import scenarios as S
    from gpu_ppm_amd.lib import Context
    ctx = Context(root, lensing="wide", responses=("std", "std128"))
    rng = np.random.default_rng(seed)
    combos = [(ns, nph, ne) for ns in (256, 1024, 2048) for nph in (32, 64, 128) for ne in (128, 360)]
    grids = {c: abi.make_grid(c[0], c[1], sampling=abi.SAMPLE_GL, n_gl=2, response=0 if c[2] == 360 else 1)
             for c in combos}
    n_ch = ctx.response_info(0)[0]
    bufs = {c: (np.zeros((c[1], c[2])), np.zeros((n_ch, c[1]))) for c in combos}
    walk = S.PriorWalk(seed)

    def draw(n):
        return [(walk.next(), combos[rng.integers(len(combos))]) for _ in range(n)]

    def run(reqs, arrivals, audit=None):
        pending, lat, i = {}, [], 0
        t0 = time.perf_counter()
        while i < len(reqs) or pending:
            now = time.perf_counter() - t0
            while i < len(reqs) and (arrivals is None or arrivals[i] <= now):
                star, c = reqs[i]
                pending[ctx.submit(star, grids[c])] = (c, i)
                i += 1
            for t in list(pending):
                c, k = pending[t]
                st, f, q, l = ctx.poll(t, grids[c], out=bufs[c])
                if st == abi.PENDING:
                    continue
                if st != abi.OK:
                    raise RuntimeError("serving request failed: status %d" % st)
                lat.append(l)
                if audit is not None and k % 100 == 0:
                    audit[k] = (np.array(f, copy=True), np.array(q, copy=True))
                del pending[t]
            time.sleep(0.0005)
        return time.perf_counter() - t0, np.array(lat)

    def stream(rate, n_req):
        arrivals = np.cumsum(rng.exponential(1.0 / rate, n_req))
        reqs = draw(n_req)
        return reqs, arrivals

    run(draw(len(combos)), None)  # warm-up: every grid once
    t_burst, _ = run(draw(burst), None)
    capacity = burst / t_burst
    rate = 0.8 * capacity
    n_fixed = max(1, int(SERVE_FIXED_RATE * duration_s))
    reqs_f, arr_f = stream(SERVE_FIXED_RATE, n_fixed)
    log("bench: serving %d requests at a fixed %.1f/s (capacity %.1f/s)" % (n_fixed, SERVE_FIXED_RATE, capacity))
    dt_f, lat_f = run(reqs_f, arr_f)
    fixed = {"requests": n_fixed, "offered_per_s": SERVE_FIXED_RATE, "completed_per_s": n_fixed / dt_f,
             "p50_ms": float(np.percentile(lat_f, 50)), "p99_ms": float(np.percentile(lat_f, 99)),
             "max_ms": float(lat_f.max())}
    n_req = max(1, int(rate * duration_s))
    reqs, arrivals = stream(rate, n_req)
    log("bench: serving %d requests at %.1f/s (capacity %.1f/s)" % (n_req, rate, capacity))
    audit = {}
    s1 = ctx.serve_stats()
    dt, lat = run(reqs, arrivals, audit)
    s2 = ctx.serve_stats()
    lanes = {"express_launches": s2.launches[0] - s1.launches[0], "bulk_launches": s2.launches[1] - s1.launches[1],
             "express_launches_overlapping_bulk": s2.overlapped - s1.overlapped,
             "device_bytes_held": s2.device_bytes}
    err = {"flux": 0.0, "counts": 0.0, "requests_audited": len(audit)}
    for k, (f, q) in audit.items():
        star, c = reqs[k]
        g = abi.make_grid(c[0], c[1], sampling=abi.SAMPLE_GL, n_gl=2, response=grids[c].response,
                          method=abi.METHOD_PLAIN)
        f0, q0, _ = ctx.eval([star], g)
        err["flux"] = max(err["flux"], rel_err(f, f0[0]))
        err["counts"] = max(err["counts"], rel_err(q, q0[0]))
    # the same audited requests against the FP64 oracle on one energy column
    # each (the middle of the grid; every column of a profile is independent),
    # within about a minute of CPU time
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_binding import Oracle
    orc = Oracle(root, lensing="wide", nsx_path=os.path.join(root, "ext", "model_data", "nsx_H_v200804.out"),
                 responses=("std", "std128"))
    err_o = {"flux": 0.0, "requests_audited": 0, "energy_column": "middle of each grid", "oracle_s": 0.0}
    t_o = time.perf_counter()
    for k, (f, q) in sorted(audit.items()):
        if time.perf_counter() - t_o > 60:
            break
        star, c = reqs[k]
        e_all = orc.response(grids[c].response)[0]
        col = len(e_all) // 2
        gc = abi.make_grid(c[0], c[1], sampling=abi.SAMPLE_GL, n_gl=2, energies=[float(e_all[col])])
        ref, _ = orc.profile(star, gc, threads=host_cores())
        err_o["flux"] = max(err_o["flux"], rel_err(f[:, [col]], ref))
        err_o["requests_audited"] += 1
    err_o["oracle_s"] = time.perf_counter() - t_o
    orc.close()
    log("bench: served vs plain %s, vs FP64 oracle %s" % (err, err_o))
    ref_evals = 0
    for star, c in reqs:
        ref_evals += ctx.count_active(star, grids[c]) * c[1] * 2 * c[2]
    ctx.close()
    return {"requests": n_req, "offered_per_s": rate, "burst_capacity_per_s": capacity,
            "completed_per_s": n_req / dt, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "max_ms": float(lat.max()),
            "evals_per_s_ref": ref_evals / dt, "audit_max_rel_err_vs_plain": err,
            "audit_max_rel_err_vs_fp64_oracle": err_o, "lanes": lanes, "fixed_rate": fixed,
            "workload": "prior random walk (2%% steps), N_side {256,1024,2048} x bins {32,64,128} x 2 GL "
                        "x E {128,360}, open-loop Poisson arrivals at 80%% of burst capacity for %.0f s; "
                        "fixed_rate: the same at %.1f/s" % (duration_s, SERVE_FIXED_RATE)}


def copy_calibration(mib):
    """Device-to-device copies of a known size: the reference point that
    scales the FETCH_SIZE / WRITE_SIZE counters into HBM bytes."""
    from gpu_ppm_amd import hip
    n = mib << 20
    src, dst = hip.DeviceArray((n,), "u1"), hip.DeviceArray((n,), "u1")
    for _ in range(4):  # the first is a warm-up
        hip.copy_device(dst, src, n)
    hip.synchronize()
    return {"calibration": "device-to-device copy", "bytes_per_copy": n, "copies": 4}


def committed_traffic(kernel):
    """HBM bytes per launch of `kernel` on the default batched workload, from
    the committed rocprofv3 counter summary (FETCH_SIZE / WRITE_SIZE passes
    scaled by a copy-kernel calibration), or None."""
    path = os.path.join(ROOT, "profiles", "k_integral_batched_256stars_counters.md")
    if not os.path.exists(path):
        return None
    key = "traffic_bytes_per_launch %s:" % kernel
    with open(path) as f:
        for line in f:
            if line.startswith(key):
                return float(line[len(key):])
    return None


def bench_plain(root, steps, warmup):
    from gpu_ppm_amd import abi, hip
    from gpu_ppm_amd import scenarios as S
    from gpu_ppm_amd.lib import Context
    star = S.cfg1_star()
    ctx = Context(root, lensing="std", responses=("std",))
    n_ch, e = ctx.response_info(0)
    d_flux = hip.DeviceArray((1, N_PHASE, N_E))
    d_cnt = hip.DeviceArray((1, n_ch, N_PHASE))
    dt, st = timed(ctx, [star], grid(abi.METHOD_PLAIN), d_flux, d_cnt, steps, warmup)
    ctx.close()
    return dict(stars=[star], lensing="std", value=steps / dt, dt=dt, kernel="k_integral_plain",
                kernel_ms=sum(s.integral_ms for s in st) / steps, n_eval=st[0].n_eval_done,
                n_eval_ref=st[0].n_eval_ref, n_active=st[0].n_active)


def phase_breakdown(root, n_stars):
    """Wavefront time of k_integral_batched by phase (libppm_phase.so: the
    same kernels with shader-clock reads at the phase boundaries), for one
    direct launch of the batched workload after an untimed one.  Fractions
    are of the summed wavefront time; the sweep includes the patch set-up
    and the window ends, and "evaluation" is the rest of the sweep."""
    from gpu_ppm_amd import abi, hip
    from gpu_ppm_amd import scenarios as S
    from gpu_ppm_amd.lib import Context
    os.environ["PPM_NO_GRAPH"] = "1"
    stars = S.sample_priors(n_stars, seed=0)
    ctx = Context(root, lensing="wide", responses=("std",))
    n_ch, _ = ctx.response_info(0)
    d_flux = hip.DeviceArray((n_stars, N_PHASE, N_E))
    d_cnt = hip.DeviceArray((n_stars, n_ch, N_PHASE))
    gb = grid(abi.METHOD_BATCHED)
    ctx.eval_device(stars, gb, d_flux.ptr, d_cnt.ptr)
    ctx.phase_timers(reset=True)
    st = ctx.eval_device(stars, gb, d_flux.ptr, d_cnt.ptr)
    ph = ctx.phase_timers()
    ctx.close()
    tot = float(ph["total"])
    frac = {k: ph[k] / tot for k in abi.PHASE_NAMES if k != "total"}
    frac["evaluation"] = (ph["sweep"] - ph["patch_setup"] - ph["window_ends"]) / tot
    return {"kernel": "k_integral_batched", "library": "libppm_phase.so", "n_stars": n_stars,
            "integral_ms": st.integral_ms, "wavefronts": ph["wavefronts"], "cycles_total": ph["total"],
            "fraction_of_wavefront_time": frac,
            "workload": "batched: %d prior-drawn stars, N_side 1024, 64 bins x 2 GL, 360 E, NSX" % n_stars}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="batched", choices=["batched", "plain"])
    ap.add_argument("--stars", type=int, default=256, help="stars per batched step")
    ap.add_argument("--plain-stars", type=int, default=256,
                    help="stars timed on the plain path (default: the whole batched workload)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="use the default batched launch shape")
    ap.add_argument("--serve-seconds", type=float, default=15.0,
                    help="duration of the open-loop serving leg (0 skips it)")
    ap.add_argument("--serve-burst", type=int, default=192, help="requests in the capacity burst")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run device-to-device copies of MIB MiB (counter calibration)")
    ap.add_argument("--phase-timers", action="store_true",
                    help="only report where k_integral_batched's wavefronts spend their time, from the "
                         "phase-timer build of the library (libppm_phase.so), on the batched workload")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (profiles shard by star or by "
              "ring range with one small reduction); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        print(json.dumps(copy_calibration(args.copy_calibration)), flush=True)
        return
    if args.phase_timers:
        os.environ["PPM_LIBRARY"] = "phase"  # before the library is loaded
        print(json.dumps(phase_breakdown(data_root(), args.stars)), flush=True)
        return
    root = data_root()
    log("bench: data ready; workload", args.workload)
    if args.workload == "batched":
        r = bench_batched(root, args.steps, args.warmup, args.stars, args.plain_stars, tune=not args.no_tune)
        cfg = {"workload": "batched: %d prior-drawn stars per launch, N_side 1024, 64 bins x 2 GL, 360 E, NSX, "
                           "fold 300x360" % args.stars, "n_stars": args.stars}
        unit_note = "%d distinct evaluations (phase-shift dedupe) x %d FLOP" % (r["n_eval"], FLOP_PER_EVAL)
    else:
        r = bench_plain(root, args.steps, args.warmup)
        cfg = {"workload": "plain single profile: N_side 1024, 64 bins x 2 GL, 360 E, NSX, fold 300x360"}
        unit_note = "%d evaluations x %d FLOP" % (r["n_eval"], FLOP_PER_EVAL)
    cfg.update({"n_side": N_SIDE, "n_phase": N_PHASE, "n_gl": 2, "n_energy": N_E, "n_active": r["n_active"]})
    achieved = r["n_eval"] * FLOP_PER_EVAL / (r["kernel_ms"] * 1e-3) / 1e12
    ref_eq = r["n_eval_ref"] * FLOP_PER_EVAL / (r["kernel_ms"] * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": r["value"],
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
        "config": cfg,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": committed_traffic(r["kernel"]) if args.workload == "batched" and args.stars == 256
                     else None,
                     "kernel": r["kernel"],
                     "kernel_ms": r["kernel_ms"], "roof": "FP32 vector ALU (no packed FP32)",
                     "units": unit_note, "ref_equivalent_tflops": ref_eq},
    }
    if args.workload == "batched":
        plain_tf = r["plain_evals_per_s"] * FLOP_PER_EVAL / 1e12
        out["plain"] = {"profiles_per_s": r["plain_profiles_per_s"], "stars_timed": r["plain_stars"],
                        "evals_per_s_ref": r["plain_evals_per_s"], "integral_ms_per_profile": r["plain_integral_ms"],
                        "kernel": "k_integral_plain", "achieved_tflops": plain_tf,
                        "frac": plain_tf / FP32_PEAK_TFLOPS}
        out["dedupe_single"] = {"profiles_per_s": r["single_profiles_per_s"],
                                "integral_ms_per_profile": r["single_integral_ms"],
                                "what": "the batched path (phase-shift dedupe) with one star per launch, same stars"}
        out["batched_vs_plain"] = r["value"] / r["plain_profiles_per_s"]
        out["batching_only_factor"] = r["value"] / r["single_profiles_per_s"]
        out["dedupe_single_vs_plain"] = r["single_profiles_per_s"] / r["plain_profiles_per_s"]
        out["batched_vs_plain_max_rel_err"] = r["max_rel_err"]
        out["batched"] = {"graph_ms_per_step": r["graph_ms"], "integral_ms": r["kernel_ms"],
                          "setup_reduce_fold_ms": r["setup_reduce_fold_ms"], "n_eval_ref": r["n_eval_ref"],
                          "dedupe_factor": r["n_eval_ref"] / max(r["n_eval"], 1),
                          "launch_shape": r["launch_shape"]}
    if args.workload == "batched" and args.serve_seconds > 0:
        out["serving"] = bench_serving(root, args.serve_seconds, args.serve_burst)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(root, r["stars"], r["lensing"], r.get("batched_flux"))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
