"""Benchmark of the MI355X r²HPDHG backend.

Default workload (`--workload lp1m`, one GPU): the random LP with a planted optimum, m = 1,000,000
rows, n = 2,000,000 columns, exactly 6 nonzeros per row (6,000,000 nnz), seed 1, fp64, solved
through the C API to relative KKT 1e-8 and verified (presolve, scaling, ‖Ã‖₂, r²HPDHG on the GPU,
finish and check). `time_to_1e-8_s` is the whole solve (the reference's definition); a step is
one r²HPDHG iteration of the solver's loop (two fused launches, the check every 64 iterations,
restarts), `steps` is the number the solve ran; `value` is the HBM GB/s of the iteration loop
(algorithmic bytes per iteration × iterations ÷ the loop's wall time, inputs resident in HBM).
The reference's restart schedule alone stalls near a relative gap of 3e-7 on this instance, so
epochs are capped at `--epoch-cap 4096` iterations while the gap leads (said in `config`);
`--compare-seconds S` also runs both schedules for S seconds and reports the gap each reaches.
Per-kernel durations come from HIP events on the solve's stream in a second, short run; `roofline`
reports the fused Ãᵀy + primal-update kernel against the 8 TB/s HBM peak, with the HBM bytes per
launch from the committed rocprofv3 counter summary. `cpu_baseline` times the reference CPU solver
(oracle/_ref, all host cores) for 256 iterations of the same model, in the same unit (one thread for
64 beside it). `reference_instances` solves the reference's own published benchmark instances
(rand-1e6-s1 and refinery-T8760-s1, made by its generators restated in lpgen) with default options
in fp64 and mixed, timed as the reference times them, beside its published RTX 4050 and 12-thread
CPU times. `serving` is the metric's other half: a short Poisson stream (`--line-requests`, m
log-uniform in [1e3, 2e4]) served with continuous batching, solves/s and p50 / p99 latency beside
the same requests solved one after another in plain fp64.

Other workloads (each prints one JSON line of the same shape):
  refinery   planning LP (T periods of 49 rows × 60 columns), mixed precision in increment form
      versus plain fp64: iteration rates at T = 8760, and both to 1e-8 end to end at a T that
      converges in bench time; value = fp64 seconds ÷ mixed seconds, with the objective difference.
  scenarios  64 scenarios sharing one 100,000 × 200,000 random matrix solved as one batch (SpMM);
      value = scenario-iterations per second (fp64; mixed beside it) against one LP alone, plus an
      end-to-end run at 10,000 rows with b, c and u scaled by lognormal(0, 0.05) factors (warm
      start, 1e-6) against sequential single solves, and the 64 cost scenarios at 10,000 rows to
      1e-6 and to 1e-8, fp64 batch against mixed batch.
  serving    Poisson arrivals of random LPs with m log-uniform in [m_lo, m_hi], half at 1e-6
      and half at 1e-8, admitted under an HBM byte budget, served in mixed precision (increment
      form; --serving-precision fp64 for plain); value = solves per second, with p50 / p99 latency,
      against sequential plain fp64 solves of the same requests.

  copy       a device copy of known size (torch), the calibration kernel of the counter summaries.
  stream     1 GiB copy / read / write rates of torch's kernels: the practical HBM ceiling.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload lp1m|refinery|scenarios|serving|copy|stream]
                    [--m M] [--precision fp64|mixed] [--time-limit S] [--no-cpu-baseline]
                    [--epoch-cap N] [--compare-seconds S] [--line-requests N]
                    [--scenario-epoch-cap N] [--serving-epoch-cap N] [--m-hi M] [--requests N]
                    [--iteration-limit N]   (profiling: same per-launch work, N iterations)
    (python bench.py --help lists every option)
"""
import argparse
import json
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path[:0] = [os.path.join(ROOT, "ps26119-solver_amd", "python"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

METRIC = "time-to-1e-8 KKT (s) + HBM GB/s on 1M-row LP; served solves/s + p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_bytes(m, n, nnz):
    """Algorithmic bytes of one launch of each fused fp64 kernel (int32 indices; the gathered
    vector counted once).
      primal: Ãᵀ (nnz·(8+4) + (n+1)·4), gather y (8m), read x, x0, c, l, u (40n), write x̄, x (16n)
      dual:   Ã  (nnz·(8+4) + (m+1)·4), gather x̄ (8n), read y, y0, rl, ru (32m), write y (8m)
    The x̂ / ŷ, ȳ stores happen only on the launches before a check or at an epoch start and are
    not counted (under 2% of launches)."""
    primal = nnz * 12 + (n + 1) * 4 + 8 * m + 56 * n
    dual = nnz * 12 + (m + 1) * 4 + 8 * n + 40 * m
    return primal, dual


def delta_kernel_bytes(m, n, nnz):
    """The same for the increment-form (mixed precision) kernels: fp32 values and increments.
      primal: Ãᵀ (nnz·(4+4) + (n+1)·4), gather δy (4m), read δx, G, LX, UX (16n), write δx̄, δx (8n)
      dual:   Ã  (nnz·(4+4) + (m+1)·4), gather δx̄ (4n), read δy, BL, BU, KY (16m), write δy (4m)"""
    primal = nnz * 8 + (n + 1) * 4 + 4 * m + 24 * n
    dual = nnz * 8 + (m + 1) * 4 + 4 * n + 20 * m
    return primal, dual


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(lp, gpu_iterations, gpu_solve, what, single_thread_iterations=0):
    """Reference CPU solver on all host cores, bounded sample: setup plus 256 iterations; with
    single_thread_iterations > 0 also a one-thread sample of that many iterations (the
    reference's own CPU rows pair its fastest thread count with one thread)."""
    import orc
    if not orc.available("ref"):
        return None
    ref = orc.load("ref")
    sample_it = 256
    t = time.time()
    r, _ = ref.solve(lp, presolve=True, iteration_limit=sample_it, threads=0)
    wall = time.time() - t
    per_it = (r["seconds"] - r["setup_seconds"]) / max(1, r["iterations"])
    est = r["setup_seconds"] + per_it * gpu_iterations + max(0.0, wall - r["seconds"])
    out = {"value": est, "unit": "s", "cores": host_cores(), "kind": "reference",
           "sample": f"reference r2HPDHG (oracle/_ref, threads=all) on the same {what}, {r['iterations']} iterations "
                     f"in {r['seconds']:.1f} s (setup {r['setup_seconds']:.1f} s, {per_it * 1e3:.2f} ms/iteration), "
                     f"extrapolated to the GPU run's {gpu_iterations} iterations",
           "ms_per_iteration": per_it * 1e3, "setup_seconds": r["setup_seconds"],
           "speedup_vs_cpu": est / gpu_solve if gpu_solve > 0 else None}
    if single_thread_iterations > 0:
        r1, _ = ref.solve(lp, presolve=True, iteration_limit=single_thread_iterations, threads=1)
        out["single_thread"] = {"iterations": r1["iterations"], "setup_seconds": r1["setup_seconds"],
                                "ms_per_iteration": (r1["seconds"] - r1["setup_seconds"]) / max(1, r1["iterations"]) * 1e3}
    return out


def base_line(args, value, unit, higher, dtype, config):
    return {"metric": METRIC, "value": value, "unit": unit, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": value * 1e3 if unit == "s" else None, "higher_is_better": higher, "scaling": "weak",
            "vs_baseline": None, "dtype": dtype, "data": "synthetic", "config": config}


def roofline_of(st, pb, name):
    launches, kms = st["kernel_launches"], st["kernel_ms"]
    avg_us = [kms[k] / launches[k] * 1e3 if launches[k] else float("nan") for k in range(3)]
    gbs = pb / (avg_us[0] * 1e-6) / 1e9 if launches[0] else float("nan")
    return {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
            "traffic": None, "kernel": name, "bytes_per_launch": pb, "avg_launch_us": avg_us[0]}, avg_us


def epoch_knobs(args, cap=None):
    """Engine knobs of --epoch-cap (or `cap`): end an epoch at that length while the relative gap
    is the largest KKT term (not in the reference; {} = its restart schedule alone)."""
    n = args.epoch_cap if cap is None else cap
    return {"hip_max_epoch": n, "hip_max_epoch_gap_only": 1} if n > 0 else {}


def serving_knobs(args):
    """Engine knobs of the served requests: --serving-epoch-cap N caps their epochs at N iterations
    while the gap leads ({} = the reference's restart schedule)."""
    n = args.serving_epoch_cap
    return {"hip_max_epoch": n, "hip_max_epoch_gap_only": 1} if n > 0 else {}


# ------------------------------------------------------------------ lp1m
PH_PRIMAL = "k_ph_primal (phased Ãᵀy + projection / reflection / Halpern update)"


def load_traffic(name):
    """HBM bytes per launch of the iteration kernels from the committed rocprofv3 counter summary
    profiles/<name>.json (read / write requests counted per request size), or {}."""
    path = os.path.join(ROOT, "profiles", name + ".json")
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f).get("traffic", {})


def bench_lp1m(args, ps26119, lpgen):
    """The 1M-row random LP solved to relative KKT 1e-8 through the C API, verified (presolve,
    device scaling, ‖Ã‖₂, r²HPDHG, finish and the in-process check, all inside `seconds`, the
    reference's time-to-1e-8 definition). A step is one r²HPDHG iteration of that solve (two
    fused launches, the check every 64 iterations, restarts): `steps` = the iterations it ran,
    `value` = algorithmic bytes per iteration × iterations ÷ the wall time of the iteration loop
    (inputs resident in HBM; setup and finish excluded). The reference's restart schedule alone
    plateaus near a relative gap of 3e-7 on this instance, so the solve ends an epoch at
    --epoch-cap iterations while the gap is the largest KKT term (engine knobs hip_max_epoch,
    hip_max_epoch_gap_only; in `config`); --compare-seconds S > 0 also runs both schedules for S
    seconds and reports the gap each reaches. --iteration-limit N instead stops the timed solve after N
    iterations (profiling runs)."""
    t = time.time()
    lp = lpgen.random_lp(args.m, seed=args.seed)
    log(f"generated m={lp.m} n={lp.n} nnz={lp.nnz} in {time.time() - t:.1f}s")
    mixed = args.precision == "mixed"
    knobs = epoch_knobs(args)
    profiling = args.iteration_limit > 0
    if args.warmup > 0:  # one untimed solve: --warmup iterations rounded up to whole 64-iteration blocks
        w_it = 64 * ((args.warmup + 63) // 64)
        r = ps26119.solve_model(lp, iteration_limit=w_it, precision=args.precision)
        log(f"warmup: {r.status_name} it={r.iterations} {r.seconds:.2f}s")
    res = ps26119.solve_model(lp, stats=True, time_limit=args.time_limit, iteration_limit=args.iteration_limit,
                              verbosity=1, precision=args.precision, params=knobs)
    st = res.stats
    log(f"timed: {res.status_name} it={res.iterations} solve={res.seconds:.2f}s setup={st['setup_seconds']:.2f}s "
        f"loop={st['iterate_seconds']:.3f}s obj={res.objective!r} planted={lp.optimum!r} gap={res.gap:.2e} "
        f"check={res.check} {res.message[:160]}")
    timed = ps26119.solve_model(lp, params={"hip_timing": 1, **knobs}, stats=True, iteration_limit=256,
                                precision=args.precision)
    pb, db = (delta_kernel_bytes if mixed else kernel_bytes)(lp.m, lp.n, lp.nnz)
    roof, avg_us = roofline_of(timed.stats, pb, "k_delta_primal (fp32 increments: Ãᵀδy + projection/reflection/Halpern)"
                               if mixed else PH_PRIMAL)
    traffic = load_traffic("k_ph_fp64_random1M_pmc_traffic")
    phase_traffic = None
    if not mixed and lp.m == 1_000_000 and lp.nnz == 6_000_000 and traffic:
        roof["traffic"] = traffic.get("primal")
        roof["traffic_source"] = "profiles/k_ph_fp64_random1M_pmc_traffic.md"
        phase_traffic = traffic.get("primal", 0) + traffic.get("dual", 0)
    loop_s = st["iterate_seconds"]
    per_it = loop_s / max(1, res.iterations)
    value = (pb + db) / per_it / 1e9
    reached = res.status == ps26119.OPTIMAL and res.check == "PASS"
    line = base_line(args, value, "GB/s", True, "f32+f64" if mixed else "f64",
                     {"workload": f"random LP with planted optimum, m={lp.m} n={lp.n} nnz={lp.nnz} (6 per row), seed "
                                  f"{args.seed}, {args.precision}, r2HPDHG solve to relative KKT 1e-8 through the C API"
                                  + (f", epochs capped at {args.epoch_cap} iterations while the gap leads" if knobs else "")
                                  + (f"; profiling run stopped at {args.iteration_limit} iterations" if profiling else ""),
                      "m": lp.m, "n": lp.n, "nnz": lp.nnz, "tolerance": 1e-8,
                      "engine_knobs": knobs, "reductions": "reference order for decisions, KKT sums as device trees"})
    line["steps"] = res.iterations
    line["ms_per_step"] = per_it * 1e3
    line.update({
        "time_to_1e-8_s": res.seconds if reached else None,
        "status": res.status_name, "check": res.check, "iterations": res.iterations,
        "objective_error_vs_planted": abs(res.objective - lp.optimum) / (1 + abs(lp.optimum)),
        "relative_gap_reached": res.gap, "setup_seconds": st["setup_seconds"], "loop_seconds": loop_s,
        "roofline": roof, "bytes_per_iteration": pb + db, "iterations_per_s": 1.0 / per_it,
        "dual_kernel": {"achieved_GBs": db / (avg_us[1] * 1e-6) / 1e9, "avg_launch_us": avg_us[1], "bytes_per_launch": db},
        "check_pass": {"launches": timed.stats["kernel_launches"][2], "avg_us": avg_us[2]},
        "spmv_phase_GBs": (pb + db) / ((avg_us[0] + avg_us[1]) * 1e-6) / 1e9,
        # HBM bytes the two iteration kernels move per launch (counters) over the same launch times
        "spmv_phase_traffic_GBs": (phase_traffic / ((avg_us[0] + avg_us[1]) * 1e-6) / 1e9
                                   if phase_traffic else None),
        "device_bytes": st["device_bytes"], "message": res.message[:200],
    })
    if not profiling and args.compare_seconds > 0 and knobs:
        cmp = {}
        for tag, prm in (("reference_schedule", {}), ("epoch_cap", knobs)):
            r = ps26119.solve_model(lp, time_limit=args.compare_seconds, params=prm, precision=args.precision)
            cmp[tag] = {"status": r.status_name, "iterations": r.iterations, "relative_gap": r.gap,
                        "relative_primal": r.primal_residual, "seconds": r.seconds}
            log(f"{tag} for {args.compare_seconds:.0f}s: {r.status_name} it={r.iterations} gap={r.gap:.2e}")
        line["equal_time_comparison"] = cmp
    if not profiling and not args.skip_reference_instances:
        line["reference_instances"] = reference_instances(ps26119, lpgen)
    if not profiling and args.line_requests > 0:  # the metric's serving half: solves/s and p99 of a short stream
        line["serving"] = serving_stream(ps26119, lpgen, args.line_requests, args.m_lo, args.m_hi, args.rate,
                                         args.budget_gb, args.serving_precision, args.seed, 60.0,
                                         params=serving_knobs(args))
    cb = None if args.no_cpu_baseline else cpu_baseline(lp, res.iterations, res.seconds, "1M-row LP",
                                                         single_thread_iterations=64)
    if cb:
        one = cb.get("single_thread")
        cpu_gbs = (pb + db) / (cb["ms_per_iteration"] * 1e-3) / 1e9
        sample = (f"reference r2HPDHG (oracle/_ref, threads=all) on the same 1M-row LP: 256 iterations after its setup, "
                  f"{cb['ms_per_iteration']:.2f} ms/iteration; value = the same algorithmic bytes per iteration ÷ that time; "
                  f"time to 1e-8 extrapolated to the GPU run's iterations: {cb['value']:.0f} s")
        cb = {"value": cpu_gbs, "unit": "GB/s", "cores": cb["cores"], "kind": "reference", "sample": sample,
              "cpu_model": cpu_model(), "ms_per_iteration": cb["ms_per_iteration"], "setup_seconds": cb["setup_seconds"],
              "time_to_1e-8_s_extrapolated": cb["value"], "speedup_per_iteration": cb["ms_per_iteration"] / (per_it * 1e3)}
        if one:
            cb["single_thread"] = {**one, "GB/s": (pb + db) / (one["ms_per_iteration"] * 1e-3) / 1e9,
                                   "sample": f"the same solver with threads=1: {one['iterations']} iterations after its setup"}
    line["cpu_baseline"] = cb
    return line


# The reference's published time-to-1e-8 runs on its own instances (whole solve() as there,
# bench/scale.py:74-83 of the reference; bench/results/scale-rtx4050-laptop-183c59c.csv lines):
# name -> {precision: (iterations, GPU seconds (RTX 4050 Laptop), CPU 12-thread seconds)}
REFERENCE_PUBLISHED = {
    "rand-1000000-s1": {"fp64": (17600, 115.7293, 489.3651), "mixed": (16576, 105.8289, 435.7844)},
    "refinery-T8760-s1": {"fp64": (2880, 6.1839, 18.7071), "mixed": (3136, 5.7836, 14.0595)},
}


def reference_instances(ps26119, lpgen):
    """The reference's own benchmark instances (its generators restated in lpgen, instance and
    planted optimum bit for bit) solved here with default options — presolve, 1e-8, the
    reference's restart schedule, the reference's reductions for decisions — in fp64, in mixed
    with the reference's precision policy, and in mixed as the increment form, each timed as the
    reference times it (Solution.seconds) beside its published GPU and 12-thread CPU times."""
    out = {}
    for name, make in (("rand-1000000-s1", lambda: lpgen.reference_bench_lp(1_000_000, seed=1)),
                       ("refinery-T8760-s1", lambda: lpgen.reference_refinery_lp(8760, seed=1))):
        lp = make()
        rows = {}
        for tag, prec, params in (("fp64", "fp64", None), ("mixed", "mixed", {"hip_refine": 0}),
                                  ("mixed_increment_form", "mixed", None)):
            r = ps26119.solve_model(lp, precision=prec, params=params)
            pub = REFERENCE_PUBLISHED[name]["fp64" if prec == "fp64" else "mixed"]
            rows[tag] = {"status": r.status_name, "check": r.check, "iterations": r.iterations,
                         "seconds": r.seconds, "objective_error_vs_planted": abs(r.objective - lp.optimum) / (1 + abs(lp.optimum)),
                         "published_iterations": pub[0], "published_gpu_seconds": pub[1],
                         "published_cpu12_seconds": pub[2], "speedup_vs_published_gpu": pub[1] / r.seconds}
            log(f"{name} {tag}: {r.status_name} it={r.iterations} (published {pub[0]}) {r.seconds:.2f}s "
                f"(published GPU {pub[1]:.2f}s) check={r.check}")
        out[name] = {"m": lp.m, "n": lp.n, "nnz": lp.nnz, **rows}
    return out


# ------------------------------------------------------------------ copy (counter calibration)
def bench_copy(args):
    """A device-to-device copy of --copy-mib MiB, 11 times (torch): the kernel of known read and
    write bytes whose rocprofv3 request counters calibrate the traffic summaries in profiles/."""
    import torch
    n = args.copy_mib << 17  # doubles
    a = torch.arange(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(11):
        b.copy_(a)
    torch.cuda.synchronize()
    return {"metric": "copy calibration", "value": args.copy_mib, "unit": "MiB", "n_gpus": 1, "steps": 11,
            "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
            "dtype": "f64", "data": "synthetic", "config": {"workload": f"device copy of {args.copy_mib} MiB"},
            "roofline": None, "cpu_baseline": None}


def bench_stream(args):
    """The streaming rates torch's own kernels reach on a 1 GiB fp64 buffer — a copy, a read (a
    sum) and a write (a fill), 20 launches each timed with HIP events: the practical ceiling of
    this GPU's HBM beside its 8 TB/s peak."""
    import torch
    n = (1 << 30) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    rates = {}
    for name, fn, nbytes in (("copy", lambda: y.copy_(x), 16 * n), ("read", lambda: x.sum(), 8 * n),
                             ("write", lambda: y.fill_(1.0), 8 * n)):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(20):
            fn()
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / 20
        rates[name] = {"us": round(us, 1), "GB/s": round(nbytes / us / 1e3, 1)}
    return {"metric": "HBM streaming rate", "value": rates["copy"]["GB/s"], "unit": "GB/s", "n_gpus": 1, "steps": 20,
            "warmup": 3, "ms_per_step": rates["copy"]["us"] / 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "1 GiB device copy / sum / fill (torch)"}, "rates": rates,
            "roofline": {"bound": "hbm", "achieved": rates["copy"]["GB/s"], "peak": 8000.0, "unit": "GB/s",
                         "frac": rates["copy"]["GB/s"] / 8000.0, "traffic": 16 * n},
            "cpu_baseline": None}


# ------------------------------------------------------------------ refinery
def bench_refinery(args, ps26119, lpgen):
    """Refinery planning LP, mixed precision in increment form vs plain fp64.
    (a) at T = --T periods (8760: 429,240 × 525,600): the iteration rate of both (loop wall time of
        --steps iterations, setup excluded) and the HBM rate of the increment-form primal kernel;
    (b) end to end at T = --T-small periods, both to 1e-8: seconds, iterations, and the objective
        difference between the two answers. `value` = fp64 seconds ÷ mixed seconds of (b)."""
    t = time.time()
    lp = lpgen.refinery_lp(T=args.T, seed=3)
    log(f"refinery T={args.T}: m={lp.m} n={lp.n} nnz={lp.nnz} in {time.time() - t:.1f}s")
    iters = max(64, args.steps // 64 * 64)
    rate = {}
    for prec in ("fp64", "mixed"):
        r = ps26119.solve_model(lp, precision=prec, stats=True, iteration_limit=iters, verbosity=1)
        rate[prec] = r.iterations / r.stats["iterate_seconds"]
        log(f"T={args.T} {prec}: {rate[prec]:.0f} iterations/s, gap {r.gap:.2e} after {r.iterations}")
    timed = ps26119.solve_model(lp, precision="mixed", params={"hip_timing": 1}, stats=True,
                                iteration_limit=min(iters, 2048))
    pb, db = delta_kernel_bytes(lp.m, lp.n, lp.nnz)
    roof, avg_us = roofline_of(timed.stats, pb, "k_delta_primal (fp32 increments: Ãᵀδy + projection/reflection/Halpern)")
    traffic_file = os.path.join(ROOT, "profiles", "k_delta_refinery8760_pmc_traffic.json")
    if args.T == 8760 and os.path.exists(traffic_file):
        with open(traffic_file) as f:  # HBM bytes per launch from the two rocprofv3 --pmc passes
            roof["traffic"] = json.load(f)["traffic"].get("k_delta_primal<float>")
        roof["traffic_source"] = "profiles/k_delta_refinery8760_pmc_traffic.md"
    if args.refinery_rate_only:
        line = base_line(args, rate["mixed"], "iterations/s", True, "f32+f64",
                         {"workload": f"refinery planning LP at T={args.T}: iteration rates only"})
        line.update({"roofline": roof, "iterations_per_s": rate, "cpu_baseline": None})
        return line
    small = lpgen.refinery_lp(T=args.T_small, seed=3)
    e2e = {}
    for prec in ("mixed", "fp64"):
        r = ps26119.solve_model(small, precision=prec, time_limit=args.time_limit, verbosity=1)
        e2e[prec] = r
        log(f"T={args.T_small} {prec}: {r.status_name} it={r.iterations} {r.seconds:.2f}s obj={r.objective!r} "
            f"check={r.check} {r.message[:120]}")
    mx, fp = e2e["mixed"], e2e["fp64"]
    both = mx.status == fp.status == ps26119.OPTIMAL
    value = fp.seconds / mx.seconds if both and mx.seconds > 0 else None
    line = base_line(args, value, "x (fp64 s / mixed s to 1e-8)", True, "f32+f64",
                     {"workload": f"refinery planning LP (49 rows x 60 columns per period): rates at T={args.T}, end "
                                  f"to end to 1e-8 at T={args.T_small}; mixed = fp32 increments on fp64 anchors",
                      "m": lp.m, "n": lp.n, "nnz": lp.nnz, "m_small": small.m})
    line["steps"] = iters
    line["ms_per_step"] = 1e3 / rate["mixed"]
    line.update({
        "roofline": roof, "iterations_per_s": rate, "rate_ratio_mixed_vs_fp64": rate["mixed"] / rate["fp64"],
        "end_to_end": {p: {"status": r.status_name, "seconds": r.seconds, "iterations": r.iterations, "check": r.check}
                       for p, r in e2e.items()},
        "objective_rel_diff_mixed_vs_fp64": abs(mx.objective - fp.objective) / (1 + abs(fp.objective)) if both else None,
    })
    line["cpu_baseline"] = None if args.no_cpu_baseline else cpu_baseline(lp, iters, iters / rate["fp64"],
                                                                           f"refinery T={args.T}")
    return line


# ------------------------------------------------------------------ scenarios
def bench_scenarios(args, ps26119, lpgen):
    """Three measurements, all bounded:
    (a) at the full size (100,000 × 200,000, 64 scenarios) the batch's rate in scenario-iterations
        per second, fp64 and mixed, against one LP alone (its loop time) — `value`. The rate runs
        use cost scenarios (c scaled by lognormal(0, 0.05) factors), which all stay live: scaling b,
        c and u together (below) makes most scenarios infeasible, and those leave the batch after a
        few checks with a certificate, which would change the width being timed;
    (b) end to end at --e2e-m rows (10,000 by default; 100,000 is the full size) with b, c and u
        scaled by lognormal(0, 0.05) factors, warm-started from the base solution, to 1e-6 and to
        1e-8: fp64 batch and mixed batch (each capped at 3M iterations or 150 s), and sequential
        single solves to 1e-6 (--seq-sample timed, scaled to 64), with the verdicts of each
        (Optimal, or Infeasible with a checked dual-ray certificate);
    (c) at 10,000 × 20,000 the 64 cost scenarios warm-started to 1e-6 and to 1e-8, fp64 batch
        against mixed batch (capped at 6M iterations or 150 s), both with the --scenario-epoch-cap
        knobs."""
    m = args.m if args.m != 1_000_000 else 100_000
    base = lpgen.random_lp(m, seed=args.seed)
    scen = [{"obj": sc["obj"]} for sc in lpgen.scenario_set(base, args.scenarios, sigma=0.05, seed=7)]
    log(f"scenarios: base m={base.m} n={base.n} nnz={base.nnz}, {len(scen)} cost scenarios")
    lo, hi = 64, args.scenario_iterations

    def timed(fn):
        t = time.time()
        fn()
        return time.time() - t

    ps26119.solve_model_batch(base, scen[:2], iteration_limit=lo)  # one-time costs (code objects, first contexts)
    # each timed solve includes its setup (≈ 2 s, with some jitter): the best of two per length
    tb = [min(timed(lambda k=k: ps26119.solve_model_batch(base, scen, iteration_limit=k, verbosity=1))
              for _ in range(2)) for k in (lo, hi)]
    tm = [min(timed(lambda k=k: ps26119.solve_model_batch(base, scen, iteration_limit=k, verbosity=1,
                                                          precision="mixed")) for _ in range(2)) for k in (lo, hi)]
    one = ps26119.solve_model(base, iteration_limit=hi, presolve=False, stats=True)
    batch_rate = len(scen) * (hi - lo) / (tb[1] - tb[0])
    mixed_rate = len(scen) * (hi - lo) / (tm[1] - tm[0])
    single_rate = one.iterations / one.stats["iterate_seconds"]
    log(f"batch: {batch_rate:.0f} scenario-iterations/s ({(tb[1] - tb[0]) / (hi - lo) * 1e3:.3f} ms/iteration for "
        f"{len(scen)}), mixed {mixed_rate:.0f}; single LP: {single_rate:.0f} iterations/s")
    W, nz = len(scen), base.nnz
    # algorithmic bytes of one batched iteration (all slots live): the matrix once, per (row, slot) the
    # vectors read and written, and one gathered value per (nonzero, slot)
    b64 = (12 * nz + 4 * (base.n + 1) + base.n * W * 56 + nz * W * 8) + (12 * nz + 4 * (base.m + 1) + base.m * W * 40 +
                                                                         nz * W * 8)
    b32 = (8 * nz + 4 * (base.n + 1) + base.n * W * 24 + nz * W * 4) + (8 * nz + 4 * (base.m + 1) + base.m * W * 20 +
                                                                        nz * W * 4)
    loop_gbs = {"fp64": b64 * batch_rate / W / 1e9, "mixed": b32 * mixed_rate / W / 1e9}
    if args.scenario_rate_only:
        line = base_line(args, batch_rate, "scenario-iterations/s", True, "f64",
                         {"workload": f"{W} cost scenarios sharing a random {base.m}x{base.n} matrix: batch rates only"})
        line.update({"mixed_scenario_iterations_per_s": mixed_rate, "loop_GBs": loop_gbs, "roofline": None,
                     "cpu_baseline": None})
        return line
    # (b) end to end at --e2e-m rows, b, c and u scaled, to 1e-6 and to 1e-8
    small = lpgen.random_lp(10_000, seed=args.seed)
    e2e_lp = small if args.e2e_m == small.m else lpgen.random_lp(args.e2e_m, seed=args.seed)
    sscen = lpgen.scenario_set(e2e_lp, args.scenarios, sigma=0.05, seed=7)
    cap = {"iteration_limit": 3_000_000, "time_limit": 150.0}
    runs, warm = {}, {}
    for tol in (1e-6, 1e-8):
        base_sol = ps26119.solve_model(e2e_lp, tolerance=tol)
        warm[tol] = base_sol
        for prec in ("fp64", "mixed"):
            t = time.time()
            res = ps26119.solve_model_batch(e2e_lp, sscen, tolerance=tol, warm_x=base_sol.x, warm_y=base_sol.y,
                                            verbosity=1, precision=prec, **cap)
            runs[(prec, tol)] = (time.time() - t, res)
            log(f"{e2e_lp.m} rows, b/c/u scenarios to {tol:.0e}, {prec} batch: {runs[(prec, tol)][0]:.2f}s")
    b = warm[1e-6]
    k_seq = min(args.seq_sample, len(sscen))
    t = time.time()
    seq = [ps26119.solve_model(lpgen.with_scenario(e2e_lp, sc), tolerance=1e-6, warm_x=b.x, warm_y=b.y,
                               presolve=False, **cap) for sc in sscen[:k_seq]]
    seq_s = (time.time() - t) / k_seq * len(sscen)

    def verdicts(res):
        out = {}
        for r in res:
            out[r.status_name] = out.get(r.status_name, 0) + 1
        return out

    agree = sum(r.status == q.status for r, q in zip(runs[("fp64", 1e-6)][1], seq))
    # (c) feasible scenarios at 10k rows (costs scaled only), fp64 batch vs mixed batch to 1e-6
    cost = {}
    if not args.skip_cost:
        b = ps26119.solve_model(small, tolerance=1e-6)
        cscen = [{"obj": sc["obj"]} for sc in lpgen.scenario_set(small, args.scenarios, sigma=0.05, seed=7)]
    for tol in ((1e-6, 1e-8) if not args.skip_cost else ()):
        for prec in ("fp64", "mixed"):
            t = time.time()
            res = ps26119.solve_model_batch(small, cscen, tolerance=tol, warm_x=b.x, warm_y=b.y, verbosity=1,
                                            precision=prec, iteration_limit=6_000_000, time_limit=150.0,
                                            params=epoch_knobs(args, args.scenario_epoch_cap))
            cost[f"{prec}_{tol:.0e}"] = {"seconds": time.time() - t, "verdicts": verdicts(res),
                                         "iterations_max": max(r.iterations for r in res)}
            log(f"10k cost scenarios to {tol:.0e}, {prec} batch: {cost[f'{prec}_{tol:.0e}']}")
        fp, mx = cost[f"fp64_{tol:.0e}"], cost[f"mixed_{tol:.0e}"]
        if fp["verdicts"].get("Optimal") == mx["verdicts"].get("Optimal") == len(cscen):
            cost[f"mixed_speedup_{tol:.0e}"] = fp["seconds"] / mx["seconds"]
    f6, m6 = runs[("fp64", 1e-6)], runs[("mixed", 1e-6)]
    log(f"{e2e_lp.m} rows end to end to 1e-6: fp64 batch {f6[0]:.2f}s {verdicts(f6[1])}, mixed batch "
        f"{m6[0]:.2f}s {verdicts(m6[1])}, sequential estimate {seq_s:.2f}s "
        f"(sample {verdicts(seq)}, {agree}/{k_seq} verdicts equal to the batch's)")
    line = base_line(args, batch_rate, "scenario-iterations/s", True, "f64",
                     {"workload": f"{W} scenarios sharing a random {base.m}x{base.n} matrix (6 per row), c scaled by "
                                  f"lognormal(0, 0.05); one batched r2HPDHG (SpMM over 64-wide block vectors), fp64; "
                                  f"rate over iterations {lo}..{hi}", "m": base.m, "n": base.n, "nnz": base.nnz,
                      "scenarios": W})
    line["steps"] = hi - lo
    line["ms_per_step"] = (tb[1] - tb[0]) / (hi - lo) * 1e3
    line.update({
        "roofline": {"bound": "hbm", "achieved": loop_gbs["fp64"], "peak": 8000.0, "unit": "GB/s",
                     "frac": loop_gbs["fp64"] / 8000.0, "traffic": None,
                     "kernel": "batched iteration loop (k_bprimal2 + k_bdual2 + the check every 64), from the loop rate",
                     "bytes_per_iteration": b64},
        "loop_GBs": loop_gbs, "single_lp_iterations_per_s": single_rate,
        "speedup_vs_single_lp_iterations": batch_rate / single_rate,
        "mixed_scenario_iterations_per_s": mixed_rate,
        "end_to_end": {"m": e2e_lp.m, "n": e2e_lp.n, "scenarios": len(sscen), "warm_start": "base solution",
                       "perturbation": "b, c and u scaled by lognormal(0, 0.05) factors", "caps": "3M iterations, 150 s",
                       **{f"{prec}_batch_{tol:.0e}": {"seconds": runs[(prec, tol)][0],
                                                       "verdicts": verdicts(runs[(prec, tol)][1]),
                                                       "iterations_max": max(r.iterations for r in runs[(prec, tol)][1])}
                          for prec in ("fp64", "mixed") for tol in (1e-6, 1e-8)},
                       "sequential_seconds_estimate_1e-06": seq_s, "sequential_sample": k_seq,
                       "sequential_sample_verdicts": verdicts(seq),
                       "speedup_fp64_batch_1e-06": seq_s / f6[0] if f6[0] > 0 else None},
        "cost_scenarios_10k": None if args.skip_cost else {"perturbation": "c scaled by lognormal(0, 0.05) factors", "tolerances": [1e-6, 1e-8],
                               "warm_start": "base solution", "caps": "6M iterations, 150 s",
                               "epoch_cap": args.scenario_epoch_cap, **cost},
    })
    line["cpu_baseline"] = None
    return line


# ------------------------------------------------------------------ serving
def serving_stream(ps26119, lpgen, N, m_lo, m_hi, rate, budget_gb, precision, seed, time_limit, eager=False,
                   sequential=True, params=None):
    """A Poisson stream of N random LPs (6 per row, n = 2m, m log-uniform in [m_lo, m_hi]; half at
    tolerance 1e-6, half at 1e-8) served with continuous batching under an HBM byte budget, then
    (sequential=True) the same requests as back-to-back plain fp64 solves, with the same engine
    knobs (`params`). HIP events around every served chunk give the device time of the served
    iterations (the roofline)."""
    rng = np.random.default_rng(seed)
    ms = np.exp(rng.uniform(np.log(m_lo), np.log(m_hi), N)).astype(int)
    tols = np.where(rng.random(N) < 0.5, 1e-6, 1e-8)
    gaps = rng.exponential(1.0 / rate, N)
    arrivals = np.cumsum(gaps) - gaps[0]
    lps = [lpgen.random_lp(int(m), seed=1000 + k) for k, m in enumerate(ms)]
    log(f"serving: {N} requests, m in [{ms.min()}, {ms.max()}], rate {rate}/s, budget {budget_gb} GB")
    srv = ps26119.Server(budget_bytes=int(budget_gb * (1 << 30)), max_instances=256, setup_threads=8, eager=eager,
                         timing=True)
    ids = [None] * N

    def feed():
        t0 = time.time()
        for k in range(N):
            d = arrivals[k] - (time.time() - t0)
            if d > 0:
                time.sleep(d)
            ids[k] = srv.submit(lps[k], tolerance=float(tols[k]), time_limit=time_limit, precision=precision,
                                params=params)

    t = time.time()
    th = threading.Thread(target=feed)
    th.start()
    th.join()
    res = []
    for k, i in enumerate(ids):  # one line per finished request: progress for a long stream
        res.append(srv.wait(i))
        log(f"request {k}: m={ms[k]} tol={tols[k]:.0e} {res[-1].status_name} it={res[-1].iterations} "
            f"latency {res[-1].stats['latency']:.3f}s")
    served = time.time() - t
    kt = srv.kernel_time()
    srv.close()
    roof = None
    if kt and kt["iterations"]:
        avg_us = kt["ms"] / kt["iterations"] * 1e3
        gbs = kt["bytes"] / (kt["ms"] * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS,
                "traffic": None, "kernel": "k_srv_primal + k_srv_dual (one served iteration of every live instance)",
                "bytes_per_launch": kt["bytes"] / kt["iterations"], "avg_launch_us": avg_us,
                "launches": kt["iterations"], "timing": "HIP events around every chunk on the serving stream"}
        log(f"served iterations: {kt['iterations']}, avg {avg_us:.1f} us, {gbs:.0f} GB/s algorithmic")
    lat = np.array([r.stats["latency"] for r in res])
    ok = sum(r.status == ps26119.OPTIMAL for r in res)
    log(f"served: {ok}/{N} optimal in {served:.2f}s, p50 {np.percentile(lat, 50):.3f}s p99 {np.percentile(lat, 99):.3f}s")
    out = {"requests": N, "m_range": [int(ms.min()), int(ms.max())], "precision": precision, "engine_knobs": params or {},
           "solves_per_s": N / served, "optimal": ok, "p50_latency_s": float(np.percentile(lat, 50)),
           "p99_latency_s": float(np.percentile(lat, 99)), "served_seconds": served, "roofline": roof}
    if sequential:  # the same requests as plain fp64 solves, back to back
        t = time.time()
        for k, (lp, tol) in enumerate(zip(lps, tols)):
            r = ps26119.solve_model(lp, tolerance=float(tol), time_limit=time_limit, params=params)
            log(f"sequential {k}: {r.status_name} it={r.iterations} {r.seconds:.3f}s")
        seq = time.time() - t
        log(f"sequential: {seq:.2f}s")
        out["sequential"] = {"seconds": seq, "solves_per_s": N / seq}
        out["speedup_vs_sequential"] = seq / served
    return out


def bench_serving(args, ps26119, lpgen):
    r = serving_stream(ps26119, lpgen, args.requests, args.m_lo, args.m_hi, args.rate, args.budget_gb,
                       args.serving_precision, args.seed, args.time_limit, eager=args.serving_eager,
                       sequential=not args.serving_only, params=serving_knobs(args))
    N = args.requests
    dtype = "f32+f64" if args.serving_precision == "mixed" else "f64"
    if args.serving_only:
        line = base_line(args, r["solves_per_s"], "solves/s", True, dtype,
                         {"workload": f"serving: {N} requests, served part only", "requests": N})
    else:
        line = base_line(args, r["solves_per_s"], "solves/s", True, dtype,
                         {"workload": f"serving: {N} Poisson arrivals at {args.rate}/s of random LPs (6 per row, n=2m), m "
                                      f"log-uniform in [{args.m_lo}, {args.m_hi}], half at tolerance 1e-6 and half at 1e-8, "
                                      f"HBM budget {args.budget_gb} GB, continuous batching, "
                                      + ("mixed precision (fp32 increments on fp64 anchors)"
                                         if args.serving_precision == "mixed" else "fp64")
                                      + "; baseline: the same requests as sequential plain fp64 solves",
                          "requests": N})
    line.update({k: v for k, v in r.items() if k not in ("requests", "solves_per_s")})
    line["cpu_baseline"] = None
    return line


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16384,
                    help="refinery / scenarios: iterations of the rate measurements; lp1m runs to 1e-8 and "
                         "reports the iterations it ran as steps")
    ap.add_argument("--warmup", type=int, default=1, help="untimed warm-up iterations (whole 64-iteration blocks)")
    ap.add_argument("--workload", default="lp1m", choices=["lp1m", "refinery", "scenarios", "serving", "copy",
                                                          "stream"])
    ap.add_argument("--m", type=int, default=1_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--time-limit", type=float, default=450.0)
    ap.add_argument("--compare-seconds", type=float, default=0.0,
                    help="lp1m: both restart schedules also run this long and report the gap they reach (0 = skip)")
    ap.add_argument("--copy-mib", type=int, default=64, help="copy: buffer size of the calibration copy")
    ap.add_argument("--iteration-limit", type=int, default=0,
                    help="stop each timed solve after this many iterations (profiling runs; 0 = to 1e-8)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-reference-instances", action="store_true",
                    help="lp1m: leave out the reference's own published instances (rand-1e6, refinery-T8760)")
    ap.add_argument("--epoch-cap", type=int, default=4096,
                    help="lp1m: hip_max_epoch with hip_max_epoch_gap_only (0 = the reference's restart schedule alone)")
    ap.add_argument("--scenario-epoch-cap", type=int, default=8192,
                    help="scenarios: the same for the cost-scenario batches (0 = the reference's schedule)")
    ap.add_argument("--precision", default="fp64", choices=["fp64", "mixed"],
                    help="lp1m: mixed = fp32 increments on fp64 anchors (the line's dtype says which)")
    ap.add_argument("--T", type=int, default=8760, help="refinery: planning periods of the rate measurement")
    ap.add_argument("--T-small", type=int, default=24, help="refinery: planning periods of the end-to-end comparison")
    ap.add_argument("--refinery-rate-only", action="store_true", help="refinery: only the rates at --T (profiling runs)")
    ap.add_argument("--scenarios", type=int, default=64, help="scenarios")
    ap.add_argument("--seq-sample", type=int, default=4, help="scenarios: sequential solves timed for the comparison")
    ap.add_argument("--e2e-m", type=int, default=10_000, help="scenarios: rows of the end-to-end b/c/u scenario set")
    ap.add_argument("--skip-cost", action="store_true", help="scenarios: leave out the 10k cost-scenario comparison")
    ap.add_argument("--scenario-iterations", type=int, default=16384, help="scenarios: iterations of the rate measurement")
    ap.add_argument("--scenario-rate-only", action="store_true", help="scenarios: only the batch rates (profiling runs)")
    ap.add_argument("--requests", type=int, default=48, help="serving")
    ap.add_argument("--line-requests", type=int, default=24,
                    help="lp1m: requests of the short serving stream reported in the line's `serving` (0 = skip)")
    ap.add_argument("--rate", type=float, default=8.0, help="serving: arrivals per second")
    ap.add_argument("--m-lo", type=int, default=1000, help="serving")
    ap.add_argument("--m-hi", type=int, default=20000,
                    help="serving (at 2e5 rows a single request needs minutes: the default keeps the stream short)")
    ap.add_argument("--budget-gb", type=float, default=16.0, help="serving: HBM byte budget of the server")
    ap.add_argument("--serving-precision", default="mixed", choices=["fp64", "mixed"], help="serving: precision of the served requests")
    ap.add_argument("--serving-only", action="store_true", help="serving: skip the sequential comparison (profiling runs)")
    ap.add_argument("--serving-epoch-cap", type=int, default=0,
                    help="serving: hip_max_epoch with hip_max_epoch_gap_only for the served requests (0 = off)")
    ap.add_argument("--serving-eager", action="store_true",
                    help="serving: launch the kernels directly and time every primal launch (the line's roofline)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print(f"bench.py measures one GPU: the r2HPDHG path here runs on a single MI355X; --gpus {args.gpus} "
              "is not supported (multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "copy":
        print(json.dumps(bench_copy(args)), flush=True)
        return
    if args.workload == "stream":
        print(json.dumps(bench_stream(args)), flush=True)
        return
    import ps26119
    from ps26119_amd import lpgen
    if ps26119.device_count() <= 0:
        log("no GPU visible to the HIP backend")
        sys.exit(1)
    fn = {"lp1m": bench_lp1m, "refinery": bench_refinery, "scenarios": bench_scenarios, "serving": bench_serving}[args.workload]
    print(json.dumps(fn(args, ps26119, lpgen)), flush=True)


if __name__ == "__main__":
    main()
