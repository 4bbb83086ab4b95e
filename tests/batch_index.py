"""Driver of the CPU check of the batched engine's device indexing
(tests/native/batch_index_check.cpp): builds it with g++ into
tests/native/_build/ and runs engine lives (admissions, removals, runs) under a
tuning table, reporting every index that would fall outside its buffer.

    python -m tests.batch_index [--slots N] [--exhaustive]

runs the tuning sweep's own ensembles (every bucket, every candidate shape of
the sweep that faulted, at the sweep's ensemble size by default) and prints a
line per case."""
from __future__ import annotations

import argparse
import ctypes as C
import math
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "native" / "batch_index_check.cpp"
OUT = ROOT / "tests" / "native" / "_build" / "libbatch_index_check.so"
CSRC = ROOT / "yukawamd_amd" / "csrc"
DEPS = [SRC, CSRC / "yk_batch_plan.hpp", CSRC / "yk_layout.hpp"]

ADD, REMOVE, RUN = 1, 2, 3

# The candidate grid of the sweep that ended in a GPU memory fault (the list
# build's segment offsets sign-extended past 2^31 list entries, a device-code
# conversion outside what this check restates): waves 1, 2, 4; tile items of
# 2-128 source tiles (at most 64 items per target tile); 2, 4, 8 or 16 list
# items per tile up to 65,536 particles.
FAULT_SWEEP_WAVES = [1, 2, 4]
FAULT_SWEEP_CHUNKS = [2, 4, 8, 16, 32, 64, 128]
FAULT_SWEEP_LIST_ITEMS = [2, 4, 8, 16]


def fault_sweep_candidates(n_max: int):
    nt = (n_max + 63) // 64
    out = []
    for c in FAULT_SWEEP_CHUNKS:
        if c > nt and c != FAULT_SWEEP_CHUNKS[0]:
            continue
        if (nt + c - 1) // c > 64:
            continue
        for w in FAULT_SWEEP_WAVES:
            out.append({"list_items": 0, "chunk_tiles": min(c, nt), "waves_per_block": w})
    if n_max <= 65536:
        for mm in FAULT_SWEEP_LIST_ITEMS:
            for w in FAULT_SWEEP_WAVES:
                out.append({"list_items": mm, "chunk_tiles": 8, "waves_per_block": w})
    return out


class Replica(C.Structure):
    _fields_ = [("n", C.c_int32), ("steps", C.c_int32), ("L", C.c_double * 3),
                ("rccut", C.c_double), ("near_r", C.c_double), ("x", C.POINTER(C.c_double))]


class EngineCfg(C.Structure):
    _fields_ = [("capacity", C.c_int64), ("max_replicas", C.c_int32),
                ("max_records", C.c_int32), ("cus", C.c_int32), ("resort_every", C.c_int32),
                ("nbuckets", C.c_int32), ("n_upper", C.POINTER(C.c_int32)),
                ("chunk_tiles", C.POINTER(C.c_int32)), ("waves", C.POINTER(C.c_int32)),
                ("list_items", C.POINTER(C.c_int32)), ("plant", C.c_int32)]


class Report(C.Structure):
    _fields_ = [("accesses", C.c_int64), ("violations", C.c_int64), ("launches", C.c_int64),
                ("items_max", C.c_int64), ("list_entries_read", C.c_int64),
                ("clist_cap", C.c_int64), ("part_items_cap", C.c_int64), ("span", C.c_int32),
                ("waves", C.c_int32), ("blocks", C.c_int32), ("status", C.c_int32),
                ("refused", C.c_int32), ("first", C.c_char * 512)]


_lib = None


def load():
    """Builds the checker if a source is newer than the library, then loads it."""
    global _lib
    if _lib is not None:
        return _lib
    if not OUT.exists() or OUT.stat().st_mtime < max(p.stat().st_mtime for p in DEPS):
        OUT.parent.mkdir(parents=True, exist_ok=True)
        tmp = OUT.with_suffix(".so.tmp")
        subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-shared", "-fPIC",
                        "-I", str(CSRC), str(SRC), "-o", str(tmp)], check=True)
        tmp.replace(OUT)
    lib = C.CDLL(str(OUT))
    lib.ykc_check.restype = C.c_int
    lib.ykc_check.argtypes = [C.POINTER(EngineCfg), C.POINTER(Replica), C.c_int32,
                              C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.POINTER(Report)]
    lib.ykc_clear_cache.restype = None
    lib.ykc_cursor_walk.restype = C.c_int64
    lib.ykc_cursor_walk.argtypes = [C.POINTER(C.c_int32), C.c_int32, C.c_int32, C.c_int32,
                                    C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
    _lib = lib
    return lib


class Case:
    """A set of replicas (workloads) whose position arrays stay alive while
    the case runs lives over them."""

    def __init__(self, workloads, steps=None):
        self.lib = load()
        self.lib.ykc_clear_cache()
        self.w = list(workloads)
        self.x = [np.ascontiguousarray(w.x, dtype=np.float64).reshape(-1) for w in self.w]
        self.reps = (Replica * len(self.w))()
        for k, (w, x) in enumerate(zip(self.w, self.x)):
            p = w.params
            r = self.reps[k]
            r.n = x.size // 3
            r.steps = int(steps[k]) if steps is not None else 1 << 30
            for d in range(3):
                r.L[d] = p.xl[d]
            r.rccut = p.rccut
            a = (3.0 / (4.0 * math.pi * p.rho)) ** (1.0 / 3.0)
            r.near_r = 4.0 * a if p.precision == "mixed" else -1.0
            r.x = x.ctypes.data_as(C.POINTER(C.c_double))

    def run(self, tuning, ops, capacity=None, max_replicas=None, max_records=1, cus=256,
            resort_every=10, exhaustive=False, plant=0):
        nb = len(tuning)
        arr = lambda key, dflt: (C.c_int32 * max(nb, 1))(*[int(b.get(key, dflt)) for b in tuning])
        cfg = EngineCfg()
        cfg.capacity = capacity or sum((w.n + 255) // 256 * 256 for w in self.w)
        cfg.max_replicas = max_replicas or len(self.w)
        cfg.max_records = max_records
        cfg.cus = cus
        cfg.resort_every = resort_every
        cfg.nbuckets = nb
        keep = [arr("n_upper", 0), arr("chunk_tiles", 8), arr("waves_per_block", 4),
                arr("list_items", 0)]
        cfg.n_upper, cfg.chunk_tiles, cfg.waves, cfg.list_items = keep
        cfg.plant = plant
        o = np.ascontiguousarray(np.asarray(ops, dtype=np.int32).reshape(-1))
        rep = Report()
        self.lib.ykc_check(C.byref(cfg), self.reps, len(self.w),
                           o.ctypes.data_as(C.POINTER(C.c_int32)), len(o) // 2,
                           1 if exhaustive else 0, C.byref(rep))
        out = {k: getattr(rep, k) for k, _ in Report._fields_ if k != "first"}
        out["first"] = rep.first.decode()
        return out


def sweep_ops(nrep, warmup=3, steps=8):
    """The sweep's life of a candidate ensemble: admit every replica, run the
    warm-up, run the timed steps."""
    return [[ADD, k] for k in range(nrep)] + [[RUN, warmup], [RUN, steps]]


def cursor_walk(counts, segcap, lo, hi, iters, stride=16, plant=0):
    """Replays the pair kernels' cursor walk over one cluster's list
    (ykc_cursor_walk): returns (mismatches, visits per entry of [0, hi))."""
    lib = load()
    c = (C.c_int32 * 8)(*[int(v) for v in np.asarray(counts).reshape(-1)])
    visits = np.zeros(max(hi, 1), dtype=np.int32)
    bad = lib.ykc_cursor_walk(c, segcap, lo, hi, iters, stride, plant,
                              visits.ctypes.data_as(C.POINTER(C.c_int32)))
    return int(bad), visits[:hi]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--slots", type=int, default=1 << 20,
                    help="particles per bucket ensemble (the sweep's: 1,048,576)")
    ap.add_argument("--buckets", default="2048,4096,8192,16384,65536,262144")
    ap.add_argument("--seed", type=int, default=1005)
    ap.add_argument("--exhaustive", action="store_true")
    args = ap.parse_args(argv)
    sys.path.insert(0, str(ROOT))
    from yukawamd_amd.autotune import bucket_replicas
    lo, bad = 1023, 0
    for k, hi in enumerate(int(v) for v in args.buckets.split(",")):
        t0 = time.perf_counter()
        reps = bucket_replicas(lo, hi, args.seed + k, slots=args.slots)
        case = Case(reps)
        print(f"bucket ({lo}, {hi}]: {len(reps)} replicas, {sum(w.n for w in reps)} particles "
              f"({time.perf_counter() - t0:.1f} s to draw)", flush=True)
        for cand in fault_sweep_candidates(hi):
            t1 = time.perf_counter()
            r = case.run([dict(cand, n_upper=hi)], sweep_ops(len(reps)),
                         exhaustive=args.exhaustive)
            bad += r["violations"] != 0 or r["status"] != 0
            print(f"  {cand}: {r['violations']} violations of {r['accesses']:.3e} indices, "
                  f"{r['launches']} launches, items {r['items_max']}, list arena "
                  f"{r['clist_cap']} entries, {r['blocks']} x {r['waves']} waves, status "
                  f"{r['status']} ({time.perf_counter() - t1:.1f} s) {r['first']}", flush=True)
        lo = hi
    return 1 if bad else 0


if __name__ == "__main__":
    raise SystemExit(main())

