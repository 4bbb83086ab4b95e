"""Benchmark of the Yukawa pair-force + velocity-Verlet hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-batch]
                    [--no-serve] [--forms fp64,mixed]

Headline workload: the 27,648-ion crust mixture (0.6 Z=8 A=16, 0.3 Z=26 A=56,
0.1 Z=34 A=80; n = 7.18e-5 fm^-3, kappa = 0.7, Gamma(Z=8) = 100, cutoff
min(L/2, 12 lambda), dt = 0.05/omega_p), positions uniform random from seed
1001, state resident in HBM before timing.  A step is one full velocity-Verlet
step (drift + pair forces from the device-built pair lists + corrector), with
energies/pressure measured every 10th step.  ``value`` is the FP64 drop-in
form; the mixed-precision form is reported beside it.

Further sections of the same JSON line:
  forms.mixed_batched  4 replicas of the same mixture (seeds 2001..) stepped
                       together by the batched engine in mixed precision;
  batched_ensemble     256 replicas (seed 1003: N log-uniform [1024, 16384],
                       pure or mixture, Gamma U[150,220], kappa U[0.5,2]),
                       500 mixed steps in one persistent launch per step;
  serving              the first requests of the seeded serving stream served
                       with continuous batching: p50/p99 completion latency and
                       aggregate pair interactions/s;
  roofline             the FP64 pair-list kernel against the FP64 vector peak,
                       its average duration from HIP events recorded on the
                       stream it runs on (in a second timed run of the same
                       steps: the events' dispatch gaps stay out of the
                       headline run), and its profiled HBM bytes;
  cpu_baseline         the reference's own Fortran OpenMP build (oracle/_ref)
                       on this host's cores, on a bounded sample.

``value`` = pair interactions/s = N(N-1)/2 x steps / wall seconds.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 vector
PEAK_FP32_TFLOPS = 157.3     # MI355X FP32 vector (packed)
FLOPS_PER_PAIR_FP64 = 90     # directed pair, force-only step, special functions costed
FLOPS_PER_PAIR_MEAS = 109    # measurement step
FLOPS_PER_PAIR_PURE = 89     # pure ions (no Z_j factor)
FLOPS_PER_PAIR_PURE_MEAS = 108
# mixed far pair, FP32 (FMA = 2): displacement 3, r^2 6, cutoff 1, r 1, exponent
# 3, (1 + mu r) e 3, Z_j 1, r^-3 2, force factor 1, accumulate 6; the hardware
# rsq and exp2 counted as one each
FLOPS_PER_PAIR_MIXED_FP32 = 29
FM_C_SECONDS = 3.33564e-24
# HBM bytes per launch of the FP64 force-only cluster-list kernel at N = 27,648,
# from rocprofv3 FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0 (scale factors from a plain
# 256 MiB copy); counters cannot be read inside a timed run, so the figure is
# the profiled one, with its commands, in the file named here.
TRAFFIC_BYTES_PER_LAUNCH = 1.026e8
TRAFFIC_SOURCE = "profiles/hbm_traffic_k_pairs_clist_ion_mixture_27648.md"
# VALU issue of the pair kernels at N = 27,648 from rocprofv3 counters:
# SQ_INSTS_VALU x 4 cycles (a wave64 VALU instruction on a 16-lane SIMD, FP64 or
# unpacked FP32) / (GRBM_GUI_ACTIVE x 1,024 SIMDs); read from the profile named
# here, since counters cannot be collected inside a timed run.  The mixed
# kernel's figure leaves out its transcendental instructions, which run beside
# the other VALU work (with them the count exceeds the budget: 1.04)
VALU_ISSUE_FRAC_FP64 = 0.82          # of which FP64 arithmetic 0.62
VALU_ISSUE_FRAC_FP64_ARITH = 0.62
VALU_ISSUE_FRAC_MIXED = 0.97
COUNTER_SOURCE = "profiles/valu_counters_k_pairs_clist_ion_mixture_27648.md"
# bytes per particle and average durations of the O(N) step kernels at
# N = 27,648 (rocprofv3 kernel trace of this benchmark), and their HBM rate
INTEGRATORS = {
    "kernel": "k_finalize<MIX,force-only,vel_verlet_2,cluster partials,+next step's drift>",
    "bytes_per_particle_per_step": 997, "avg_us": 12.2, "gb_per_s": 2260,
    "hbm_frac": 2260 / 8000,
    "drift_alone": {"bytes_per_particle": 180, "avg_us": 9.1, "gb_per_s": 549},
    "source": TRAFFIC_SOURCE,
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nind", type=int, default=10, help="measure every nind steps")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-calls", type=int, default=25)
    ap.add_argument("--cpu-calls-1", type=int, default=2, help="calls timed on one core")
    ap.add_argument("--forms", default="fp64,mixed")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-replicas", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=500)
    ap.add_argument("--batch-crust", type=int, default=4,
                    help="replicas of the 27,648-ion mixture in the batched mixed form")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the pure-ion (N=3,456) and nucleon (N=51,200) lines")
    ap.add_argument("--config-steps", type=int, default=100)
    ap.add_argument("--config-only", default=None,
                    choices=["pure_ion_3456", "nucleon_51200", "ion_mixture_262144"],
                    help="only run that config's line (both forms) and print it as the JSON "
                         "line (for kernel traces of one workload)")
    ap.add_argument("--no-drift", action="store_true",
                    help="skip the energy drift over 1,000 steps (mixture and nucleons)")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=40)
    ap.add_argument("--serve-rate", type=float, default=3.0, help="Poisson arrivals per s")
    ap.add_argument("--serve-round", type=int, default=50, help="steps per serving round")
    ap.add_argument("--serve-budget-gib", type=float, default=64.0,
                    help="HBM the serving loop admits requests against")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (reference "
                         "kernel for scaling HBM byte counters) and exit")
    return ap.parse_args()


def run_form(w, precision, steps, warmup, nind, timing):
    """timing=False: the product path as a caller runs it (the headline);
    timing=True: HIP events around every force launch (the kernel's average
    duration for the roofline; the events add a few microseconds of dispatch
    gap per step, so this run does not give the headline)."""
    from yukawamd_amd.system import from_workload
    s = from_workload(w, precision=precision, timing=timing)
    try:
        s.accel(True)
        if warmup:
            s.run(warmup, nind)
        s.reset_timing()
        t0 = time.perf_counter()
        recs = s.run(steps, nind)      # returns after the device has finished
        t = time.perf_counter() - t0
        tim = s.timing()
        return {"seconds": t, "timing": tim, "records": recs}
    finally:
        s.close()


def form_figures(w, prec, steps, warmup, nind, flops_pair=None, flops_pair_meas=None):
    """Headline-style figures of one workload in one precision: the product
    run (no events) for the rates, a second event-timed run of the same steps
    for the force kernel's average duration and the pairs it evaluated."""
    n = w.n
    pairs = n * (n - 1) // 2
    r = run_form(w, prec, steps, warmup, nind, timing=False)
    t = r["seconds"]
    tim = run_form(w, prec, steps, warmup, nind, timing=True)["timing"]
    launches = max(tim["force_launches"], 1)
    avg_ms = tim["force_ms"] / launches
    swept = tim["pairs_swept"] / launches     # directed pairs evaluated per launch
    out = {
        "pair_interactions_per_s": pairs * steps / t,
        "ms_per_step": 1e3 * t / steps,
        "steps_per_s": steps / t,
        "ns_per_day": w.dt * FM_C_SECONDS * 1e9 * (steps / t) * 86400,
        "omega_p_t_per_day": 0.05 * (steps / t) * 86400,
        "force_kernel_avg_ms": avg_ms,
        "directed_pairs_evaluated_per_launch": swept,
        "evaluated_fraction_of_all_pairs": swept / (n * (n - 1)),
        "last_record": r["records"][-1] if r["records"] else None,
    }
    if flops_pair:
        nmeas = steps // nind if nind else 0
        flops = (steps - nmeas) * flops_pair + nmeas * flops_pair_meas
        out["force_kernel_tflops"] = swept * flops / steps / (avg_ms * 1e-3) / 1e12
    return out


def in_cutoff_fraction(p):
    """Fraction of all pairs inside the cutoff in a uniform fluid (rc <= L/2):
    the cutoff sphere's share of the box."""
    import math
    return 4.0 / 3.0 * math.pi * p.rccut ** 3 / (p.xl[0] * p.xl[1] * p.xl[2])


def energy_drift(w, warm_groups=20, warm_steps=25, steps=1000, nind=10, forms=("fp64", "mixed")):
    """Relative energy drift over `steps` NVE steps: a thermostatted warm-up
    from the random start in FP64 (warm_groups groups of warm_steps steps, each
    followed by cancel_vcom + tnorm, as src/md.F:88-101), then from that state
    `steps` steps in each precision, E = ev + ek per particle recorded every
    nind steps; drift = max |E(t) - E(0)| / |E(0)|."""
    from yukawamd_amd.system import from_workload
    s = from_workload(w, precision="fp64")
    try:
        s.accel(True)
        for _ in range(warm_groups):
            s.run(warm_steps)
            s.cancel_vcom()
        x, v, _ = s.get_state()
    finally:
        s.close()
    out = {"warmup_steps": warm_groups * warm_steps, "steps": steps, "nind": nind}
    for prec in forms:
        s = from_workload(w, precision=prec)
        try:
            s.put_state(x, v)
            e0 = s.accel(True)["ev"] + s.ttot()
            recs = s.run(steps, nind)
        finally:
            s.close()
        e = [e0] + [r["ev"] + r["ek"] for r in recs]
        out[prec] = {"relative_drift": max(abs(ei - e0) for ei in e) / abs(e0),
                     "e0_mev": e0, "e_end_mev": e[-1]}
    return out


def run_batch(ws, steps, warmup, nind):
    """Step the replicas ws together in one ensemble (budgets warmup + steps);
    the timed region is the `steps` steps with every state resident."""
    from yukawamd_amd.batch import Ensemble
    cap = sum(((w.n + 255) // 256) * 256 for w in ws)
    if nind:   # warm up over whole measuring blocks so the timed run replays the same graph
        warmup = max(1, -(-warmup // nind)) * nind
    nrec = ((warmup + steps) // nind + 1) if nind else 1
    with Ensemble(capacity=cap, max_replicas=len(ws), max_records=nrec) as e:
        ids = [e.add_workload(w, steps=warmup + steps) for w in ws]
        e.run(warmup, nind)     # accelerations of the initial state + warm-up
        t0 = time.perf_counter()
        e.run(steps, nind)
        t = time.perf_counter() - t0
        lr = e.last_run()
        left = [e.status(r)[0] for r in ids]
        last = e.get(ids[0])["records"]
    pairs = sum(w.n * (w.n - 1) // 2 for w in ws)
    return {
        "replicas": len(ws),
        "particles": int(sum(w.n for w in ws)),
        "pairs_per_step": pairs,
        "steps": steps,
        "seconds": t,
        "device_ms": lr["ms"],
        "ms_per_step": 1e3 * t / steps,
        "pair_interactions_per_s": pairs * steps / t,
        "work_items_per_step": lr["items"],
        "all_budgets_done": all(v == 0 for v in left),
        "replica0_last_record": last[-1] if last else None,
    }


def run_serving(count, rate, round_steps, budget_gib=64.0):
    """The first `count` requests of the seeded serving stream (seed 1004),
    arriving in real time at `rate` per second, served to completion."""
    from yukawamd_amd.serving import Server, stream, summarize
    reqs = stream(count, rate)
    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    with Server(hbm_budget_bytes=budget_gib * 2 ** 30, round_steps=round_steps, log=log) as srv:
        t0 = time.perf_counter()
        res = srv.serve(reqs)
        wall = time.perf_counter() - t0
        out = summarize(res, wall, srv.device_ms)
        out.update({"rounds": srv.rounds, "peak_concurrent_runs": srv.peak_live,
                    "peak_footprint_gib": srv.peak_bytes / 2 ** 30,
                    "device_utilization": srv.device_ms * 1e-3 / wall})
    out["window"] = f"requests 0-{count - 1} of the 2,000-request stream"
    out["workload"] = (f"first {count} requests of the serving stream (seed 1004): crust "
                       "mixture, N log-uniform [2048, 262144], steps U[100, 2000], Gamma "
                       f"U[150,220], kappa U[0.5,2], 80% mixed / 20% FP64; Poisson arrivals "
                       f"at {rate}/s in real time; rounds of {round_steps} steps; HBM budget "
                       f"{budget_gib:g} GiB")
    return out


def cpu_baseline(calls, threads):
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    env.setdefault("OMP_PROC_BIND", "close")
    cmd = [sys.executable, "-m", "oracle.cpu_bench", "--workload", "ion_mixture_27648",
           "--calls", str(calls)]
    out = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    if out.returncode != 0:
        return {"error": out.stderr[-400:]}
    return json.loads(out.stdout.strip().splitlines()[-1])


def copy_calibration(mib):
    import torch
    n = mib * 2 ** 20 // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_mib": mib, "bytes_read": n * 8, "bytes_written": n * 8}))
    return 0


def config_line(wc, args, steps):
    fl = FLOPS_PER_PAIR_PURE if wc.params.sim_type == "pure-ion" else None
    flm = FLOPS_PER_PAIR_PURE_MEAS if fl else None
    out = {"n": wc.n, "steps": steps, "dt_fm_per_c": wc.dt,
           "in_cutoff_fraction": in_cutoff_fraction(wc.params)}
    for prec in args.forms.split(","):
        out[prec] = form_figures(wc, prec, steps, args.warmup, args.nind, fl, flm)
        if fl:
            out[prec]["roofline_frac_fp64"] = out[prec]["force_kernel_tflops"] / PEAK_FP64_TFLOPS
    return out


def config_only(args):
    from yukawamd_amd import workloads as W
    if args.config_only == "pure_ion_3456":
        wc, steps = W.pure_ion(), args.config_steps
    elif args.config_only == "ion_mixture_262144":   # the largest system on cluster lists
        wc, steps = W.ion_mixture(n=262144, seed=1006), max(args.config_steps // 5, 10)
    else:
        wc, steps = W.nucleon(), max(args.config_steps // 2, 10)
    print(json.dumps({"config_only": wc.name, wc.name: config_line(wc, args, steps)}))
    return 0


def main():
    args = parse()
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU; --gpus {args.gpus} is not supported "
              "(a single system is not sharded across GPUs; independent replicas would run "
              "one ensemble per GPU)",
              file=sys.stderr)
        return 2
    from yukawamd_amd import workloads as W
    if args.config_only:
        return config_only(args)
    w = W.ion_mixture()
    n = w.n
    pairs = n * (n - 1) // 2
    forms = {}
    for prec in args.forms.split(","):
        forms[prec] = form_figures(w, prec, args.steps, args.warmup, args.nind,
                                   FLOPS_PER_PAIR_FP64, FLOPS_PER_PAIR_MEAS)
    f64 = forms.get("fp64") or next(iter(forms.values()))
    line = {
        "metric": "pair interactions/s at N=27,648 (FP64 drop-in; mixed beside it)",
        "value": f64["pair_interactions_per_s"],
        "unit": "pair interactions/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": f64["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform-random positions, Maxwell-Boltzmann velocities)",
        "config": {"workload": "ion_mixture_27648", "n": n, "species": "0.6 Z8A16/0.3 Z26A56/"
                   "0.1 Z34A80", "kappa": 0.7, "gamma_z8": 100, "rccut_fm": w.params.rccut,
                   "box_fm": w.params.xl[0], "dt_fm_per_c": w.dt, "nind": args.nind,
                   "integrator": "vel_verlet_2", "parallelism": "single-gpu"},
        "ns_per_day": f64["ns_per_day"],
        "forms": forms,
    }
    if "mixed" in forms and "fp64" in forms:
        line["mixed_over_fp64"] = (forms["mixed"]["pair_interactions_per_s"]
                                   / forms["fp64"]["pair_interactions_per_s"])
    if not args.no_configs:
        cfg = {}
        for wc, steps in ((W.pure_ion(), args.config_steps),
                          (W.nucleon(), max(args.config_steps // 2, 10))):
            cfg[wc.name] = config_line(wc, args, steps)
        line["configs"] = cfg
    if not args.no_drift:
        line["energy_drift_1000_steps"] = {
            "ion_mixture_27648": energy_drift(w),
            "nucleon_51200": energy_drift(W.nucleon()),
        }
    if not args.no_batch:
        from yukawamd_amd.batch import TUNING_FILE
        crust = [W.ion_mixture(seed=2001 + k, precision="mixed") for k in range(args.batch_crust)]
        if crust:
            b = run_batch(crust, args.steps, args.warmup, args.nind)
            b["workload"] = f"{len(crust)} x ion_mixture_27648 (seeds 2001..), mixed"
            forms["mixed_batched"] = b
            if "fp64" in forms:
                line["mixed_batched_over_fp64"] = (b["pair_interactions_per_s"]
                                                   / forms["fp64"]["pair_interactions_per_s"])
        ens = W.ensemble(args.batch_replicas)
        b = run_batch(ens, args.batch_steps, args.warmup, args.nind)
        b["workload"] = (f"batched ensemble: {len(ens)} replicas (seed 1003), N log-uniform "
                         "[1024, 16384], pure/mixture p=0.5, Gamma U[150,220], kappa U[0.5,2], "
                         "mixed precision")
        b["tuning"] = str(TUNING_FILE.relative_to(ROOT)) if TUNING_FILE.exists() else "default"
        line["batched_ensemble"] = b
    if not args.no_serve:
        line["serving"] = run_serving(args.serve_requests, args.serve_rate, args.serve_round,
                                      args.serve_budget_gib)
    incut = in_cutoff_fraction(w.params) * n * (n - 1)    # directed pairs inside rc
    swept = f64["directed_pairs_evaluated_per_launch"]
    line["roofline"] = {
        "bound": "valu",
        "peak_kind": "FP64 vector",
        "kernel": "k_pairs_clist<MIX,FP64>",
        "achieved": f64["force_kernel_tflops"],
        "peak": PEAK_FP64_TFLOPS,
        "unit": "TFLOP/s",
        "frac": f64["force_kernel_tflops"] / PEAK_FP64_TFLOPS,
        "traffic": TRAFFIC_BYTES_PER_LAUNCH,
        "traffic_source": TRAFFIC_SOURCE,
        "flops_per_directed_pair": FLOPS_PER_PAIR_FP64,
        "directed_pairs_per_launch": swept,
        # the pairs inside the cutoff (a uniform fluid's share) over those the
        # lists make the kernel evaluate, and the rate counted on them alone
        "in_cutoff_fraction_of_evaluated": incut / swept,
        "useful_frac": f64["force_kernel_tflops"] / PEAK_FP64_TFLOPS * incut / swept,
        # counter cross-check (rocprofv3 --pmc, file named below): VALU
        # instructions x 4 cycles / (GPU-active cycles x SIMDs)
        "valu_issue_frac": VALU_ISSUE_FRAC_FP64,
        "valu_issue_frac_fp64_arith": VALU_ISSUE_FRAC_FP64_ARITH,
        "valu_issue_frac_mixed": VALU_ISSUE_FRAC_MIXED,
        "counter_source": COUNTER_SOURCE,
    }
    if "mixed" in forms:
        mx = forms["mixed"]
        line["roofline"]["mixed_fp32_tflops"] = (
            mx["directed_pairs_evaluated_per_launch"] * FLOPS_PER_PAIR_MIXED_FP32
            / (mx["force_kernel_avg_ms"] * 1e-3) / 1e12)
        line["roofline"]["mixed_fp32_frac"] = line["roofline"]["mixed_fp32_tflops"] / PEAK_FP32_TFLOPS
        line["roofline"]["flops_per_directed_pair_mixed_fp32"] = FLOPS_PER_PAIR_MIXED_FP32
    line["integrators"] = INTEGRATORS
    if not args.no_cpu:
        share = min(len(os.sched_getaffinity(0)), 16)
        cb = cpu_baseline(args.cpu_calls, share)
        c1 = cpu_baseline(args.cpu_calls_1, 1)
        if "error" in cb:
            line["cpu_baseline"] = {"value": None, "error": cb["error"]}
        else:
            line["cpu_baseline"] = {
                "value": cb["pair_interactions_per_s"], "unit": "pair interactions/s",
                "cores": cb["threads"], "kind": cb["kind"],
                "sample": f"{cb['calls']} force-only accel calls of ion_mixture_27648 "
                          f"({cb['seconds']:.1f} s, {cb['threads']} threads of a {cb['cpu']} "
                          f"host with {os.cpu_count()} logical CPUs, this job's share "
                          f"{len(os.sched_getaffinity(0))})",
                "reference_pair_loop_s_per_call": cb["pair_loop_seconds"] / cb["calls"],
                "one_core": None if "error" in c1 else {
                    "value": c1["pair_interactions_per_s"], "calls": c1["calls"],
                    "seconds": c1["seconds"],
                    "reference_pair_loop_s_per_call": c1["pair_loop_seconds"] / c1["calls"]},
            }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
