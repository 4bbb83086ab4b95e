"""CPU baseline timing of the force step (TEST/BENCH INFRASTRUCTURE ONLY).

Times force-only accel calls of a workload on the host cores and prints one
JSON line.  Kind "reference": the reference's own Fortran OpenMP build
(oracle/_ref/libykref.so, the same loops the reference runs); kind "port":
the OpenMP restatement in oracle/yk_oracle.c.  Run as a separate process so
OMP_NUM_THREADS takes effect before the OpenMP runtime starts:

    OMP_NUM_THREADS=16 python -m oracle.cpu_bench --workload ion_mixture_27648 --calls 2
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent.parent))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="ion_mixture_27648")
    ap.add_argument("--calls", type=int, default=2)
    ap.add_argument("--kind", default="auto", choices=["auto", "reference", "port"])
    args = ap.parse_args()
    import numpy as np
    import oracle
    from yukawamd_amd import workloads as W
    name = args.workload
    if name.startswith("ion_mixture_"):
        w = W.ion_mixture(n=int(name.rsplit("_", 1)[1]))
    elif name.startswith("pure_ion_"):
        w = W.pure_ion(n=int(name.rsplit("_", 1)[1]))
    else:
        w = W.nucleon(n=int(name.rsplit("_", 1)[1]))
    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    kind = args.kind
    if kind == "auto":
        kind = "reference" if oracle.ref_available() else "port"
    if kind == "reference":
        oracle.ref_setup(w.params, w.x, w.v, w.Z, w.A)
        lib = oracle.load_ref()
        a = np.zeros((w.n + 3, 3))
        m = np.zeros(oracle.NMEAS)
        aptr = a.ctypes.data_as(oracle._D)
        mptr = m.ctypes.data_as(oracle._D)
        lib.ykref_accel(0, aptr, mptr)              # warm-up
        oracle.ref_reset_timers()
        t0 = time.perf_counter()
        for _ in range(args.calls):
            lib.ykref_accel(0, aptr, mptr)
        t = time.perf_counter() - t0
        tim = oracle.ref_timers()
        t_pair = tim["t_calc_a"]
    else:
        oracle.accel_omp(w.params, w.x, w.Z, w.A, threads)
        t0 = time.perf_counter()
        for _ in range(args.calls):
            _, used = oracle.accel_omp(w.params, w.x, w.Z, w.A, threads)
        t = time.perf_counter() - t0
        t_pair = t
    pairs = w.n * (w.n - 1) // 2
    print(json.dumps({
        "kind": kind, "workload": name, "calls": args.calls, "threads": threads,
        "seconds": t, "pair_loop_seconds": t_pair,
        "pair_interactions_per_s": pairs * args.calls / t,
        "cpu": _cpu_model(),
    }))


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


if __name__ == "__main__":
    main()
