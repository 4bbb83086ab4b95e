"""Autotuner of the batched engine's work shape, per N bucket.

    python -m yukawamd_amd.autotune [--out PATH] [--steps K] [--buckets 2048,4096,...]

For every N bucket an ensemble of replicas drawn from that bucket alone
(N log-uniform inside the bucket, pure Z=8 or the three-species mixture with
probability 1/2, Gamma U[150, 220], kappa U[0.5, 2], mixed precision: the
batched-ensemble parameter ranges) is stepped under every candidate shape,
and the shortest device time per step wins.  The two knobs are

  chunk_tiles      source tiles per work item: one wave sweeps one 64-particle
                   target tile against chunk_tiles source tiles, so this sets
                   the item grain (load balance against per-item overhead:
                   partial-sum traffic and finalize work grow with the number
                   of items per tile);
  waves_per_block  waves per workgroup of the persistent launch (LDS per
                   workgroup = waves x 64 x 32 B; occupancy and queue
                   contention).

Candidates keep at most 64 items per target tile.  The table is written as
JSON ({"device": ..., "buckets": [{"n_upper", "chunk_tiles",
"waves_per_block", "ms_per_step", "pairs_per_s"}, ...], "candidates": ...});
yukawamd_amd/tuning/batch_gfx950.json is the committed result for MI355X,
which ``Ensemble`` loads by default.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

DEFAULT_BUCKETS = [2048, 4096, 8192, 16384, 65536, 262144]
WAVES = [1, 2, 4, 8]
CHUNKS = [1, 2, 4, 8, 16, 32, 64, 128, 256]
MAX_ITEMS_PER_TILE = 64
BUCKET_SLOTS = 1 << 20          # particles per tuning ensemble (at least one replica)


def bucket_replicas(lo: int, hi: int, seed: int):
    from . import workloads as W
    rng = np.random.Generator(np.random.PCG64(seed))
    reps, total = [], 0
    while total < BUCKET_SLOTS or not reps:
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
            out.append((min(c, nt), w))
    return sorted(set(out))


def tune_bucket(lo, hi, steps, warmup, seed, log):
    from .batch import Ensemble
    reps = bucket_replicas(lo, hi, seed)
    cap = sum(((w.n + 255) // 256) * 256 for w in reps)
    pairs = sum(w.n * (w.n - 1) // 2 for w in reps)
    results = []
    with Ensemble(capacity=cap, max_replicas=len(reps), max_records=1, tuning=[]) as e:
        big = steps * 1000 + warmup * 1000
        for w in reps:
            e.add_workload(w, steps=big)
        for c, wv in candidates(hi):
            e.set_tuning([{"n_upper": hi, "chunk_tiles": c, "waves_per_block": wv}])
            e.run(warmup)
            e.run(steps)
            ms = e.last_run()["ms"] / steps
            results.append({"chunk_tiles": c, "waves_per_block": wv, "ms_per_step": ms,
                            "pairs_per_s": pairs / (ms * 1e-3)})
            log(f"  n<={hi} chunk={c} waves={wv}: {ms:.4f} ms/step "
                f"{pairs / (ms * 1e-3):.3e} pairs/s")
    best = min(results, key=lambda r: r["ms_per_step"])
    return best, results, len(reps), pairs


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None)
    ap.add_argument("--steps", type=int, default=10)
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
                                                  args.seed + k, log)
        table.append({"n_upper": hi, "chunk_tiles": best["chunk_tiles"],
                      "waves_per_block": best["waves_per_block"],
                      "ms_per_step": best["ms_per_step"], "pairs_per_s": best["pairs_per_s"]})
        detail[str(hi)] = {"replicas": nrep, "pairs_per_step": pairs, "results": results}
        lo = hi
    out = {"device": dev, "precision": "mixed", "steps": args.steps,
           "seconds": time.perf_counter() - t0, "buckets": table, "candidates": detail}
    text = json.dumps(out, indent=1)
    if args.out:
        with open(args.out, "w") as f:
            f.write(text + "\n")
    print(json.dumps({"buckets": table}))
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
