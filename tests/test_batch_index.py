"""CPU check of the batched engine's device indexing (no GPU): the engine's
own planner and list walker drive a thread-by-thread restatement of the
indices every accel-pass and step kernel forms (tests/native/
batch_index_check.cpp), and every index is checked against the buffer the
engine allocates for it.

Covered: the tuning sweep that ended in a GPU memory fault (its candidate
shapes: 1, 2, 4 waves per workgroup; tile items of 2-128 source tiles; 2, 4,
8, 16 list items per tile) on its own kind of ensembles at reduced size, the
committed tuning table on the ensemble draw, a serving-like churn of
admissions and removals, and planted mistakes (each must be caught).  The
sweep's full ensemble size (1,048,576 particles per bucket) runs with
``python -m tests.batch_index``."""
import json

import numpy as np
import pytest

from tests import batch_index as B
from yukawamd_amd import workloads as W
from yukawamd_amd.autotune import DEFAULT_BUCKETS, bucket_replicas, candidates
from yukawamd_amd.batch import TUNING_FILE


def _ok(r):
    return r["violations"] == 0 and r["status"] == 0


@pytest.mark.parametrize("k", range(len(DEFAULT_BUCKETS)))
def test_fault_sweep_shapes_in_bounds(k):
    """Every candidate shape of the sweep that faulted, on that bucket's
    sweep ensemble (reduced to 8,192 particles): every list entry read is
    checked."""
    lo = 1023 if k == 0 else DEFAULT_BUCKETS[k - 1]
    hi = DEFAULT_BUCKETS[k]
    reps = bucket_replicas(lo, hi, 1005 + k, slots=8192)
    case = B.Case(reps)
    # one replica of at least 65,537 particles: two steps (its tile items
    # alone are 1e8 checked indices per step)
    ops = B.sweep_ops(len(reps)) if hi <= 65536 else B.sweep_ops(len(reps), 1, 1)
    for cand in B.fault_sweep_candidates(hi):
        r = case.run([dict(cand, n_upper=hi)], ops, exhaustive=True)
        assert _ok(r), (cand, r["first"])
        assert r["launches"] > 0 and r["accesses"] > 0


def test_current_candidates_are_a_subset_of_the_checked_grid():
    for hi in DEFAULT_BUCKETS:
        checked = {tuple(sorted(c.items())) for c in B.fault_sweep_candidates(hi)}
        for c in candidates(hi):
            if c["waves_per_block"] == 8 or c["list_items"] == 32 or \
                    (c["list_items"] > 0 and hi > 65536):
                continue   # shapes the current grid adds: covered by the next test
            assert tuple(sorted(c.items())) in checked, (hi, c)


def test_current_candidates_in_bounds():
    """The current sweep grid (8-wave workgroups, 32 list items and lists up
    to 262,144 particles added) on every bucket, reduced ensembles."""
    lo = 1023
    for k, hi in enumerate(DEFAULT_BUCKETS):
        reps = bucket_replicas(lo, hi, 2005 + k, slots=4096)
        case = B.Case(reps)
        ops = B.sweep_ops(len(reps)) if hi <= 65536 else B.sweep_ops(len(reps), 1, 1)
        for cand in candidates(hi):
            r = case.run([dict(cand, n_upper=hi)], ops)
            assert _ok(r), (hi, cand, r["first"])
        lo = hi


def test_committed_table_on_the_ensemble_draw():
    """The committed tuning table on 24 replicas of the ensemble draw plus a
    replica above 65,536 particles (tile items), 25 steps
    with re-sorts every 10, measuring every 10th."""
    table = json.loads(TUNING_FILE.read_text())["buckets"]
    reps = W.ensemble(24) + [W.replica(77, 70000, True, 180.0, 1.0, "mixed")]
    case = B.Case(reps)
    ops = [[B.ADD, k] for k in range(len(reps))] + [[B.RUN, 25]]
    r = case.run(table, ops, exhaustive=True)
    assert _ok(r), r["first"]
    assert r["list_entries_read"] > 0


def test_serving_churn():
    """Admissions and removals between runs (continuous batching) in an arena
    with room for a fraction of the requests: first-fit slots and list
    regions reused, the list arena grown by copy, items re-planned."""
    rng = np.random.Generator(np.random.PCG64(11))
    reqs = []
    for k in range(16):
        n = int(np.exp(rng.uniform(np.log(1100), np.log(20000))))
        reqs.append(W.replica(500 + k, n, bool(k % 2), 180.0, float(rng.uniform(0.5, 2)),
                              "mixed" if k % 5 else "fp64"))
    table = json.loads(TUNING_FILE.read_text())["buckets"]
    case = B.Case(reqs)
    ops, live = [], []
    for k in range(len(reqs)):
        ops.append([B.ADD, k])
        live.append(k)
        if len(live) > 5:
            gone = live.pop(int(rng.integers(0, len(live))))
            ops.append([B.REMOVE, gone])
        ops.append([B.RUN, int(rng.integers(1, 12))])
    cap = sum((w.n + 255) // 256 * 256 for w in reqs) // 2
    r = case.run(table, ops, capacity=cap, max_replicas=10, max_records=4, exhaustive=True)
    assert _ok(r), r["first"]
    assert r["refused"] < len(reqs) // 2


@pytest.mark.parametrize("plant", [1, 2, 3, 4, 5, 6])
def test_planted_mistakes_are_caught(plant):
    """A mistake planted in the restatement (partial sums one item short, a
    list share one entry long, a segment count above its segment, a tile
    item one source tile long, a list region one cluster off, an admitted
    replica's lists never built) is reported."""
    reps = [W.replica(900 + k, n, True, 180.0, 1.0, "mixed") for k, n in
            enumerate([3000, 1500, 5000])]
    case = B.Case(reps)
    tuning = [{"n_upper": 2048, "list_items": 0, "chunk_tiles": 4, "waves_per_block": 4},
              {"n_upper": 65536, "list_items": 8, "chunk_tiles": 8, "waves_per_block": 4}]
    ops = [[B.ADD, k] for k in range(len(reps))] + [[B.RUN, 3]]
    clean = case.run(tuning, ops, exhaustive=True)
    assert _ok(clean), clean["first"]
    r = case.run(tuning, ops, exhaustive=True, plant=plant)
    assert r["violations"] > 0, plant
    print(plant, r["first"])


def _walk_cases():
    """Segment layouts of one cluster's list: (fast, back) counts per segment,
    empty segments and runs among them, shares cut anywhere."""
    rng = np.random.default_rng(2027)
    cases = [
        ([[5, 3], [0, 0], [40, 17], [2, 0]], 64),
        ([[0, 0], [0, 0], [0, 0], [0, 0]], 16),
        ([[0, 33], [0, 0], [0, 1], [0, 0]], 40),
        ([[64, 0], [1, 0], [0, 0], [31, 0]], 80),
    ]
    for _ in range(40):
        segcap = int(rng.integers(16, 400))
        counts = []
        for _ in range(4):
            f = int(rng.integers(0, segcap + 1)) if rng.random() > 0.2 else 0
            b = int(rng.integers(0, segcap - f + 1)) if rng.random() > 0.3 else 0
            counts.append([f, b])
        cases.append((counts, segcap))
    return cases


@pytest.mark.parametrize("plant", [0, 1, 2])
def test_list_cursor_walk(plant):
    """The pair kernels walk a cluster's segmented list through a cursor that
    steps +-16 within a segment's run and recomputes at a run's end.  Over
    every share of every layout, each entry the walk reads equals
    SegList::offset's (the direct formula), lies in its segment's written
    part, and every entry of the share is read exactly once.  The two planted
    mistakes (no recompute at a run's end; fast runs stepped backwards) must
    be caught."""
    rng = np.random.default_rng(7)
    caught = 0
    for counts, segcap in _walk_cases():
        total = int(np.sum(counts))
        for M in (1, 3, 8):
            for m in range(M):
                per = (total + M - 1) // M
                lo = min(m * per, total)
                hi = min(lo + per, total)
                iters = hi - lo + int(rng.integers(0, 40))   # the wave's longest share
                bad, visits = B.cursor_walk(counts, segcap, lo, hi, iters, plant=plant)
                if plant == 0:
                    assert bad == 0, (counts, segcap, lo, hi)
                    assert np.all(visits[lo:hi] == 1) and np.all(visits[:lo] == 0)
                caught += bad > 0
    if plant:
        assert caught > 0
