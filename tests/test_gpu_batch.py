"""Batched ensembles (yk_batch_*) on the GPU, through the C ABI.

* FP64 replicas of different kinds and sizes in one arena reproduce the
  oracle's accelerations (condition-aware 1e-12) and energies (1e-13);
* a replica gives bitwise the same trajectory and records alone and inside a
  mix of other replicas;
* the batch matches the single-system path to summation-order rounding;
* step budgets, removal/re-admission and the capacity limit behave;
* mixed replicas meet the mixed bars against the oracle;
* work shapes outside the committed table reproduce the oracle;
* a list arena past 2^31 entries gives the replica above that line the same
  trajectory as alone.
"""
import numpy as np
import pytest

import oracle
from _forces import masses
from yukawamd_amd import workloads as W
from yukawamd_amd._lib import YkError
from yukawamd_amd.batch import Ensemble
from yukawamd_amd.system import YukawaSystem

pytestmark = pytest.mark.gpu

# the two pair paths of the engine: cluster pair lists, and tile items with
# AABB culling
MODES = {
    "lists": [{"n_upper": 1 << 20, "chunk_tiles": 8, "waves_per_block": 4, "list_items": 16}],
    "tiles": [{"n_upper": 1 << 20, "chunk_tiles": 4, "waves_per_block": 4, "list_items": 0}],
}


def small_set():
    return [W.replica(11, 300, False, 175.0, 1.0, "fp64"),
            W.replica(12, 500, True, 160.0, 0.7, "fp64"),
            W.replica(13, 777, True, 210.0, 1.8, "fp64")]


@pytest.mark.parametrize("mode", list(MODES))
def test_batch_fp64_accel_matches_oracle(mode):
    reps = small_set()
    with Ensemble(capacity=4096, max_replicas=8, tuning=MODES[mode]) as e:
        ids = [e.add_workload(w, steps=0) for w in reps]
        e.run(0)
        for rid, w in zip(ids, reps):
            g = e.get(rid)
            a_o, m_o = oracle.accel(w.params, w.x, w.Z, w.A, True, ref_f1=0)
            scale = np.abs(a_o).max()
            err = np.abs(g["a"] - a_o).max() / scale
            print(w.name, w.n, "max |da|/max|a|", err)
            assert err <= 1e-12
            assert g["init"]["ev"] == pytest.approx(m_o["ev"], rel=1e-13)
            assert g["init"]["px"] == pytest.approx(m_o["px"], rel=1e-13)


@pytest.mark.parametrize("mode", list(MODES))
def test_mixed_replicas_against_oracle(mode):
    """Mixed-precision replicas against the FP64 oracle: on cluster lists (the
    FP64 near field) every particle's force within 1e-6 relative, on tile
    items (every pair in FP32) within 1e-4; energies within 1e-6."""
    reps = [W.replica(15, 2500, True, 175.0, 0.7, "mixed"),
            W.replica(16, 1800, False, 200.0, 1.6, "mixed")]
    with Ensemble(capacity=4608, max_replicas=2, tuning=MODES[mode]) as e:
        ids = [e.add_workload(w, steps=0) for w in reps]
        e.run(0)
        for rid, w in zip(ids, reps):
            g = e.get(rid)
            p = w.params.replace(precision="fp64")
            a_o, m_o = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=0)
            m = masses(p, w.A)[:, None]
            rel = (np.linalg.norm((g["a"] - a_o) * m, axis=1)
                   / np.linalg.norm(a_o * m, axis=1)).max()
            print(mode, w.name, w.n, "mixed replica: max |dF|/|F|", rel)
            assert rel <= (1e-6 if mode == "lists" else 1e-4)
            assert g["init"]["ev"] == pytest.approx(m_o["ev"], rel=1e-6)


@pytest.mark.parametrize("mode", list(MODES))
def test_unscreened_mixed_replica(mode):
    """A mixed-precision replica with mu = 0 (plain Coulomb): the list items'
    far pairs work in lengths scaled by mu log2 e, so a list replica runs in
    FP64 (FP64 bar); tile items keep the mixed pair (mixed bar)."""
    w = W.replica(14, 600, True, 170.0, 1.0, "mixed")
    with Ensemble(capacity=4096, max_replicas=2, tuning=MODES[mode]) as e:
        rid = e.add_workload(w, steps=0, xmuc=0.0)
        e.run(0)
        g = e.get(rid)
    p = w.params.replace(xmuc=0.0, precision="fp64")
    a_o, m_o = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=0)
    err = np.abs(g["a"] - a_o).max() / np.abs(a_o).max()
    print(mode, "mu = 0 mixed replica: max |da|/max|a|", err)
    if mode == "lists":
        assert err <= 1e-12
        assert g["init"]["ev"] == pytest.approx(m_o["ev"], rel=1e-13)
    else:
        assert err <= 1e-5
        assert g["init"]["ev"] == pytest.approx(m_o["ev"], rel=1e-6)


@pytest.mark.parametrize("mode", list(MODES))
def test_replica_alone_equals_in_mix(mode):
    target = W.replica(21, 1500, True, 180.0, 0.9, "mixed")
    others = [W.replica(22 + k, n, k % 2 == 0, 170.0, 1.3, "mixed")
              for k, n in enumerate([3000, 700, 2100, 1100])]
    results = []
    for company in ([], others):
        with Ensemble(capacity=16384, max_replicas=8, max_records=8, tuning=MODES[mode]) as e:
            ids = [e.add_workload(w, steps=20) for w in company[:2]]
            tid = e.add_workload(target, steps=20)
            ids += [e.add_workload(w, steps=20) for w in company[2:]]
            e.run(20, nind=10)
            results.append(e.get(tid))
    a, b = results
    assert np.array_equal(a["x"], b["x"]) and np.array_equal(a["v"], b["v"])
    assert a["records"] == b["records"] and len(a["records"]) == 2


@pytest.mark.parametrize("mode", list(MODES))
def test_batch_matches_single_system(mode):
    w = W.replica(31, 2000, True, 190.0, 1.2, "fp64")
    with Ensemble(capacity=4096, max_replicas=2, max_records=4, tuning=MODES[mode]) as e:
        rid = e.add_workload(w, steps=10)
        e.run(10, nind=5)
        g = e.get(rid)
    with YukawaSystem(w.params, Z=w.Z, A=w.A, dt=w.dt) as s:
        s.put_state(w.x, w.v)
        s.accel(True)
        recs = s.run(10, nind=5)
        x, v, a = s.get_state()
    L = w.params.xl[0]
    dx = np.abs(x - g["x"])
    dx = np.minimum(dx, L - dx)
    assert dx.max() <= 1e-10 * L
    np.testing.assert_allclose(g["v"], v, rtol=1e-8, atol=1e-12 * np.abs(v).max())
    for r1, r2 in zip(g["records"], recs):
        assert r1["ev"] == pytest.approx(r2["ev"], rel=1e-11)
        assert r1["ek"] == pytest.approx(r2["ek"], rel=1e-10)


def test_step_budgets_and_readmission():
    ws = [W.replica(41 + k, 400 + 100 * k, k == 1, 175.0, 1.0, "fp64") for k in range(3)]
    budgets = [5, 12, 20]
    with Ensemble(capacity=2048, max_replicas=4, max_records=4) as e:
        ids = [e.add_workload(w, steps=s) for w, s in zip(ws, budgets)]
        e.run(20)
        for rid in ids:
            assert e.status(rid)[0] == 0
        first = e.get(ids[0])
        # the same replica alone for exactly 5 steps
        with Ensemble(capacity=1024, max_replicas=1) as e2:
            r2 = e2.add_workload(ws[0], steps=5)
            e2.run(5)
            alone = e2.get(r2)
        assert np.array_equal(first["x"], alone["x"])
        # remove and admit again into the freed slots
        e.remove(ids[1])
        nid = e.add_workload(ws[1], steps=3)
        e.run(3)
        assert e.status(nid)[0] == 0


def test_capacity_limit():
    w = W.replica(51, 900, False, 175.0, 1.0, "mixed")
    with Ensemble(capacity=1024, max_replicas=4) as e:
        rid = e.add_workload(w, steps=1)
        with pytest.raises(YkError) as err:
            e.add_workload(w, steps=1)
        assert err.value.code == 7
        e.remove(rid)
        e.add_workload(w, steps=1)


def test_resort_changes_only_rounding():
    w = W.replica(61, 3000, True, 180.0, 0.8, "fp64")
    outs = []
    for every in (0, 7):
        with Ensemble(capacity=4096, max_replicas=1, max_records=4) as e:
            e.set_resort(every)
            rid = e.add_workload(w, steps=30)
            e.run(30, nind=10)
            outs.append(e.get(rid))
    a, b = outs
    L = w.params.xl[0]
    dx = np.abs(a["x"] - b["x"])
    dx = np.minimum(dx, L - dx)
    assert dx.max() <= 1e-9 * L
    assert len(a["records"]) == len(b["records"]) == 3
    for r1, r2 in zip(a["records"], b["records"]):
        assert r1["ev"] == pytest.approx(r2["ev"], rel=1e-10)
        assert r1["ek"] == pytest.approx(r2["ek"], rel=1e-10)


@pytest.mark.parametrize("mode", list(MODES))
def test_staggered_admission_with_resort_is_bitwise(mode):
    target = W.replica(71, 1800, True, 170.0, 0.9, "mixed")
    other = W.replica(72, 2500, False, 200.0, 1.5, "mixed")
    with Ensemble(capacity=8192, max_replicas=4, max_records=8, tuning=MODES[mode]) as e:
        e.set_resort(10)
        e.add_workload(other, steps=100)
        e.run(13, nind=0)                 # the other replica is 13 steps ahead
        tid = e.add_workload(target, steps=45)
        e.run(17, nind=0)
        e.run(40, nind=0)
        assert e.status(tid)[0] == 0
        mixed_in = e.get(tid)
    with Ensemble(capacity=4096, max_replicas=1, max_records=8, tuning=MODES[mode]) as e:
        e.set_resort(10)
        rid = e.add_workload(target, steps=45)
        e.run(45, nind=0)
        alone = e.get(rid)
    for key in ("x", "v", "a"):
        assert np.array_equal(mixed_in[key], alone[key]), key


def test_list_replica_above_65536():
    """A 100,000-particle FP64 replica on cluster lists (16 items per tile)
    beside a small one: its initial accelerations match the single-system
    path's cluster lists to summation-order rounding, and after 4 steps its
    state still does."""
    big = W.replica(17, 100_000, True, 180.0, 1.0, "fp64")
    small = W.replica(18, 3000, False, 170.0, 1.5, "fp64")
    tuning = [{"n_upper": 1 << 20, "chunk_tiles": 8, "waves_per_block": 4, "list_items": 16}]
    with Ensemble(capacity=104_192, max_replicas=2, max_records=2, tuning=tuning) as e:
        e.add_workload(small, steps=4)
        rid = e.add_workload(big, steps=4)
        e.run(0)
        g0 = e.get(rid)
        e.run(4, nind=2)
        g = e.get(rid)
    with YukawaSystem(big.params, Z=big.Z, A=big.A, dt=big.dt) as s:
        s.put_state(big.x, big.v)
        m0 = s.accel(True)
        a0 = s.a
        recs = s.run(4, nind=2)
        x, v, a = s.get_state()
    m = masses(big.params, big.A)[:, None]
    rel = (np.linalg.norm((g0["a"] - a0) * m, axis=1) / np.linalg.norm(a0 * m, axis=1)).max()
    print("n=100000 batched list replica vs single system: max |dF|/|F|", rel)
    assert rel <= 1e-10
    assert g0["init"]["ev"] == pytest.approx(m0["ev"], rel=1e-12)
    L = big.params.xl[0]
    dx = np.abs(x - g["x"])
    dx = np.minimum(dx, L - dx)
    assert dx.max() <= 1e-10 * L
    for r1, r2 in zip(g["records"], recs):
        assert r1["ev"] == pytest.approx(r2["ev"], rel=1e-11)
        assert r1["ek"] == pytest.approx(r2["ek"], rel=1e-10)


def test_out_of_memory_at_admission():
    """Replicas of 262,144 particles on cluster lists take 16 GiB of list
    arena each; the arena doubles as it grows (16, 32, 64, 128 GiB), so the
    ninth admission needs a 256 GiB arena beside the 128 GiB one and is
    refused with YK_ERR_NOMEM.  The ensemble keeps running (the failed
    allocation leaves no error behind for the next launch check), and after a
    removal the same replica is admitted."""
    tuning = [{"n_upper": 1 << 20, "chunk_tiles": 8, "waves_per_block": 4, "list_items": 4}]
    reps = [W.replica(131 + k, 262144, False, 175.0, 1.0, "mixed") for k in range(9)]
    with Ensemble(capacity=9 * 262144, max_replicas=9, max_records=1, tuning=tuning) as e:
        ids = [e.add_workload(w, steps=2) for w in reps[:8]]
        with pytest.raises(YkError) as err:
            e.add_workload(reps[8], steps=2)
        assert err.value.code == 3
        e.run(1)
        e.remove(ids[0])
        ids[0] = e.add_workload(reps[8], steps=1)
        e.run(1)
        assert all(e.status(r)[0] == 0 for r in ids)


def test_list_arena_beyond_2_31_entries():
    """Nine replicas of 65,536 particles on cluster lists hold 9 x 2^28 list
    entries: the ninth replica's lists start at entry 2^31, where the build
    kernel's segment offsets need their full 64-bit widening (a sign-extended
    low half addressed 16 EiB away and faulted).  The ninth replica matches
    itself alone, whose lists start at entry 0, bit for bit."""
    reps = [W.replica(91 + k, 65536, k % 2 == 0, 175.0, 1.0, "mixed") for k in range(9)]
    tuning = [{"n_upper": 1 << 20, "chunk_tiles": 8, "waves_per_block": 4, "list_items": 8}]
    with Ensemble(capacity=9 * 65536, max_replicas=9, max_records=2, tuning=tuning) as e:
        ids = [e.add_workload(w, steps=4) for w in reps]
        e.run(4, nind=2)
        last = e.get(ids[-1])
        assert all(e.status(r)[0] == 0 for r in ids)
    with Ensemble(capacity=65536, max_replicas=1, max_records=2, tuning=tuning) as e:
        rid = e.add_workload(reps[-1], steps=4)
        e.run(4, nind=2)
        alone = e.get(rid)
    for key in ("x", "v", "a"):
        assert np.array_equal(last[key], alone[key]), key
    assert last["records"] == alone["records"] and len(last["records"]) == 2


@pytest.mark.parametrize("shape", [
    {"list_items": 8, "chunk_tiles": 8, "waves_per_block": 1},
    {"list_items": 2, "chunk_tiles": 8, "waves_per_block": 4},
    {"list_items": 4, "chunk_tiles": 8, "waves_per_block": 2},
    {"list_items": 0, "chunk_tiles": 2, "waves_per_block": 4},
    {"list_items": 0, "chunk_tiles": 16, "waves_per_block": 4},
    {"list_items": 0, "chunk_tiles": 32, "waves_per_block": 2},
    {"list_items": 0, "chunk_tiles": 64, "waves_per_block": 4},
    {"list_items": 32, "chunk_tiles": 8, "waves_per_block": 8},
])
def test_tuning_shapes_against_oracle(shape):
    """Shapes outside the committed table: 1- and 8-wave workgroups, 2, 4 and
    32 list items per tile, tile items of 2 to 64 source tiles.  FP64
    replicas against the oracle (condition-free 1e-12 of the largest
    acceleration) and a step budget run out."""
    reps = small_set()
    with Ensemble(capacity=4096, max_replicas=4, tuning=[dict(shape, n_upper=1 << 20)]) as e:
        ids = [e.add_workload(w, steps=3) for w in reps]
        e.run(0)
        for rid, w in zip(ids, reps):
            g = e.get(rid)
            a_o, m_o = oracle.accel(w.params, w.x, w.Z, w.A, True, ref_f1=0)
            err = np.abs(g["a"] - a_o).max() / np.abs(a_o).max()
            print(shape, w.name, w.n, "max |da|/max|a|", err)
            assert err <= 1e-12
            assert g["init"]["ev"] == pytest.approx(m_o["ev"], rel=1e-13)
        e.run(3)
        assert all(e.status(r)[0] == 0 for r in ids)
