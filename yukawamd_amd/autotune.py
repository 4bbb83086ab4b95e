"""Autotuner of the batched engine's work shape, per N bucket.

    python -m yukawamd_amd.autotune [--out PATH] [--steps K] [--nind M] [--buckets 2048,...]

For every N bucket an ensemble of replicas drawn from that bucket alone
(N log-uniform inside the bucket, pure Z=8 or the three-species mixture with
probability 1/2, Gamma U[150, 220], kappa U[0.5, 2], mixed precision: the
batched-ensemble parameter ranges) is admitted afresh under every candidate
shape and stepped (K = 10 timed steps measuring every M = 10th, as records
are taken in production runs: the measuring step's kernels weigh in), and the
shortest device time per step wins.  The knobs:

  list_items       > 0: the replicas get 16-slot cluster pair lists and this
                   many work items per 64-slot target tile (item = an equal
                   share of each of the tile's four cluster lists); 0: tile
                   items (replicas above 262,144 particles always use them);
  chunk_tiles      tile items: source tiles per item, the item grain (load
                   balance against per-item overhead: partial-sum traffic and
                   finalize work grow with the items per tile);
  waves_per_block  waves per workgroup of the persistent launch (LDS per
                   workgroup = waves x 64 x 32 B; occupancy and queue
                   contention).

Tile-item candidates keep at most 64 items per target tile.  The table is
written as JSON ({"device": ..., "buckets": [{"n_upper", "list_items",
"chunk_tiles", "waves_per_block", "ms_per_step", "pairs_per_s"}, ...],
"candidates": ...}); yukawamd_amd/tuning/batch_gfx950.json is the committed
result for MI355X, which ``Ensemble`` loads by default.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

DEFAULT_BUCKETS = [2048, 4096, 8192, 16384, 65536, 262144]
WAVES = [2, 4, 8]
CHUNKS = [4, 8, 16, 32, 64, 128]
LIST_ITEMS = [4, 8, 16, 32]
CLIST_MAX_N = 262144
MAX_ITEMS_PER_TILE = 64
BUCKET_SLOTS = 1 << 20          # particles per tuning ensemble (at least one replica)


def bucket_replicas(lo: int, hi: int, seed: int, slots: int = BUCKET_SLOTS):
    """The tuning ensemble of the bucket (lo, hi]: replicas until `slots`
    particles (at most 256 replicas)."""
    from . import workloads as W
    rng = np.random.Generator(np.random.PCG64(seed))
    reps, total = [], 0
    while total < slots or not reps:
        n = int(round(math.exp(rng.uniform(math.log(lo + 1), math.log(hi)))))
        n = max(lo + 1, min(hi, n))
        w = W.replica(int(rng.integers(0, 2 ** 31 - 1)), n, bool(rng.random() < 0.5),
                      float(rng.uniform(150.0, 220.0)), float(rng.uniform(0.5, 2.0)), "mixed")
        reps.append(w)
        total += n
        if len(reps) >= 256:
            break
    return reps


def candidates(n_max: int):
    nt = (n_max + 63) // 64
    out = []
    for c in CHUNKS:
        if c > nt and c != CHUNKS[0]:
            continue
        if (nt + c - 1) // c > MAX_ITEMS_PER_TILE:
            continue
        for w in WAVES:
            out.append({"list_items": 0, "chunk_tiles": min(c, nt), "waves_per_block": w})
    if n_max <= CLIST_MAX_N:
        for mm in LIST_ITEMS:
            for w in WAVES:
                out.append({"list_items": mm, "chunk_tiles": 8, "waves_per_block": w})
    return out


def tune_bucket(lo, hi, steps, warmup, seed, log, nind=0):
    from .batch import Ensemble
    reps = bucket_replicas(lo, hi, seed)
    cap = sum(((w.n + 255) // 256) * 256 for w in reps)
    pairs = sum(w.n * (w.n - 1) // 2 for w in reps)
    results = []
    for cand in candidates(hi):
        tuning = [dict(cand, n_upper=hi)]
        # the list / tile choice is made at admission: a fresh ensemble each
        nrec = max(1, (steps + warmup) // nind) if nind else 1
        with Ensemble(capacity=cap, max_replicas=len(reps), max_records=nrec,
                      tuning=tuning) as e:
            for w in reps:
                e.add_workload(w, steps=steps + warmup)
            e.run(warmup)
            e.run(steps, nind)
            ms = e.last_run()["ms"] / steps
        r = dict(cand, ms_per_step=ms, pairs_per_s=pairs / (ms * 1e-3))
        results.append(r)
        log(f"  n<={hi} {cand}: {ms:.4f} ms/step {pairs / (ms * 1e-3):.3e} pairs/s")
    best = min(results, key=lambda r: r["ms_per_step"])
    return best, results, len(reps), pairs


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--nind", type=int, default=10, help="measure every nind-th step")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--buckets", default=",".join(map(str, DEFAULT_BUCKETS)))
    ap.add_argument("--seed", type=int, default=1005)
    args = ap.parse_args(argv)
    import torch  # noqa: F401  (device name only)
    dev = torch.cuda.get_device_name(0) if torch.cuda.is_available() else "unknown"

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    uppers = [int(v) for v in args.buckets.split(",")]
    table, detail = [], {}
    lo = 1023          # the batched ensemble's smallest replica has 1,024 particles
    t0 = time.perf_counter()
    for k, hi in enumerate(uppers):
        log(f"bucket ({lo}, {hi}]")
        best, results, nrep, pairs = tune_bucket(lo, hi, args.steps, args.warmup,
                                                  args.seed + k, log, args.nind)
        table.append({"n_upper": hi, "list_items": best["list_items"],
                      "chunk_tiles": best["chunk_tiles"],
                      "waves_per_block": best["waves_per_block"],
                      "ms_per_step": best["ms_per_step"], "pairs_per_s": best["pairs_per_s"]})
        detail[str(hi)] = {"replicas": nrep, "pairs_per_step": pairs, "results": results}
        lo = hi
    out = {"device": dev, "precision": "mixed", "steps": args.steps, "nind": args.nind,
           "seconds": time.perf_counter() - t0, "buckets": table, "candidates": detail}
    text = json.dumps(out, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(json.dumps({"buckets": table}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
