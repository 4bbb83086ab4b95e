"""The named workloads at their real sizes on the GPU, through the C ABI.

* ion mixture, N = 27,648 (the headline workload): FP64 forces, energy,
  pressure and tensor against the oracle (src/int_ion_mix.f:105-295);
  mixed-precision per-particle force errors against the same oracle result;
* nucleon matter, N = 51,200: FP64 against the oracle (src/int_nn.f:130-345),
  Newton's third law, bitwise determinism, a run loop equal to repeated
  newton steps, and the mixed form's errors;
* batched ensemble: 32 replicas of the 256-replica draw (seed 1003; every N
  bucket of the tuning table) under the committed tuning table, each replica
  bitwise the same alone and in the ensemble, FP64 copies against the oracle;
  one replica per N bucket over 1,000 steps of the mixed batched form: energy
  drift within the single systems' bar;
* serving stream: the first 10 requests of the seed-1004 stream (two above
  65,536 particles, four in FP64), steps capped at 20: every served result
  bitwise the same as the request run alone, and an FP64 request's
  trajectory against the oracle's velocity-Verlet steps.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from _forces import check_meas, force_error_arrays, gpu_accel, masses
from yukawamd_amd import workloads as W
from yukawamd_amd.batch import Ensemble
from yukawamd_amd.serving import Server, stream
from yukawamd_amd.system import YukawaSystem

pytestmark = pytest.mark.gpu


def report(tag, rel, rc, ratio):
    worst = np.argsort(rel)[::-1][:5]
    print(f"{tag}: max rel {rel.max():.3e} (99.9% {np.percentile(rel, 99.9):.2e}, "
          f"median {np.median(rel):.2e}) cond {rc.max():.3e}; worst particles "
          + ", ".join(f"#{i} rel {rel[i]:.2e} |F|/sum|F_ij| {ratio[i]:.2e}" for i in worst))


@pytest.fixture(scope="module")
def mixture():
    w = W.ion_mixture()
    a_o, m_o = oracle.accel(w.params, w.x, w.Z, w.A, True, ref_f1=0)
    cond, _ = oracle.pair_abs_sum(w.params, w.x, w.Z)
    return w, a_o, m_o, cond


@pytest.fixture(scope="module")
def nucleons():
    w = W.nucleon()
    a_o, m_o = oracle.accel(w.params, w.x, w.Z, w.A, True, ref_f1=0)
    cond, _ = oracle.pair_abs_sum(w.params, w.x, w.Z)
    return w, a_o, m_o, cond


def test_ion_mixture_27648_fp64_against_oracle(mixture):
    w, a_o, m_o, cond = mixture
    a_g, m_g = gpu_accel(w.params, w.x, w.Z, w.A, True)
    rel, rc, ratio = force_error_arrays(w.params, a_g, a_o, w.x, w.Z, w.A, cond)
    report("ion_mixture_27648 fp64", rel, rc, ratio)
    assert rc.max() <= 1e-12 and rel.max() <= 1e-10
    check_meas(m_g, m_o)


def test_ion_mixture_27648_mixed_against_oracle(mixture):
    """The mixed form's force error on the headline workload, all 27,648
    particles: maximum |dF_i|/|F_i| <= 1e-6 (the stated tolerance), energy
    within 1e-6."""
    w, a_o, m_o, cond = mixture
    a_g, m_g = gpu_accel(w.params.replace(precision="mixed"), w.x, w.Z, w.A, True)
    rel, rc, ratio = force_error_arrays(w.params, a_g, a_o, w.x, w.Z, w.A, cond)
    report("ion_mixture_27648 mixed", rel, rc, ratio)
    print(f"ev rel {abs(m_g['ev'] / m_o['ev'] - 1):.2e} px rel {abs(m_g['px'] / m_o['px'] - 1):.2e}")
    assert rc.max() <= 1e-7
    assert rel.max() <= 1e-6
    assert m_g["ev"] == pytest.approx(m_o["ev"], rel=1e-6)


def test_nucleon_51200_fp64_against_oracle(nucleons):
    w, a_o, m_o, cond = nucleons
    a_g, m_g = gpu_accel(w.params, w.x, w.Z, w.A, True)
    rel, rc, ratio = force_error_arrays(w.params, a_g, a_o, w.x, w.Z, w.A, cond)
    report("nucleon_51200 fp64", rel, rc, ratio)
    assert rc.max() <= 1e-12 and rel.max() <= 1e-10
    check_meas(m_g, m_o)


def test_nucleon_51200_mixed_against_oracle(nucleons):
    w, a_o, m_o, cond = nucleons
    a_g, m_g = gpu_accel(w.params.replace(precision="mixed"), w.x, w.Z, w.A, True)
    rel, rc, ratio = force_error_arrays(w.params, a_g, a_o, w.x, w.Z, w.A, cond)
    report("nucleon_51200 mixed", rel, rc, ratio)
    print(f"ev rel {abs(m_g['ev'] / m_o['ev'] - 1):.2e}")
    assert rc.max() <= 3e-6
    assert m_g["ev"] == pytest.approx(m_o["ev"], rel=1e-5)


def test_nucleon_51200_properties():
    """Full-size nucleons: total force zero, bitwise determinism, run loop ==
    repeated newton (src/md.F:124-133 cadence, measuring every 2nd step)."""
    w = W.nucleon()
    p = w.params
    with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
        s.put_state(w.x, w.v)
        m1 = s.accel(True)
        a1 = s.a
        m2 = s.accel(True)
        a2 = s.a
        assert np.array_equal(a1, a2) and m1 == m2
        F = a1 * masses(p, w.A)[:, None]
        assert np.abs(F.sum(0)).max() <= 1e-12 * np.abs(F).sum()
        for st in range(1, 5):
            s.newton(st % 2 == 0)
        xa, va, aa = s.get_state()
    with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
        s.put_state(w.x, w.v)
        s.accel(True)
        recs = s.run(4, nind=2)
        xb, vb, ab = s.get_state()
    assert len(recs) == 2
    assert np.array_equal(xa, xb) and np.array_equal(va, vb) and np.array_equal(aa, ab)


# --------------------------------------------------------------------------
def ensemble_sample():
    """32 replicas of the seed-1003 draw, covering every bucket up to 16,384."""
    reps = W.ensemble(32)
    assert {(w.n > 2048) + (w.n > 4096) + (w.n > 8192) for w in reps} == {0, 1, 2, 3}
    return reps


def test_ensemble_sample_alone_equals_in_ensemble():
    """Under the committed tuning table (tile items up to 4,096 particles,
    cluster lists above): each of 32 replicas of the ensemble draw, plus FP64
    copies of one replica per bucket, gives bitwise the same state and records
    alone and in the ensemble over 30 steps (re-sorts at 10 and 20)."""
    reps = ensemble_sample()
    by_bucket = {}
    for w in reps:
        by_bucket.setdefault((w.n > 2048) + (w.n > 4096) + (w.n > 8192), w)
    fp64 = [W.replica(7000 + k, w.n, w.params.sim_type == "ion-mixture", 175.0, 1.0, "fp64")
            for k, w in sorted(by_bucket.items())]
    members = reps + fp64
    steps, nind = 30, 10
    cap = sum((w.n + 255) // 256 * 256 for w in members)
    with Ensemble(capacity=cap, max_replicas=len(members), max_records=4) as e:
        ids = [e.add_workload(w, steps=steps) for w in members]
        e.run(0)
        for rid, w in zip(ids[len(reps):], fp64):
            g = e.get(rid)
            a_o, m_o = oracle.accel(w.params, w.x, w.Z, w.A, True, ref_f1=0)
            rel, rc, ratio = force_error_arrays(w.params, g["a"], a_o, w.x, w.Z, w.A)
            print(f"fp64 copy n={w.n}: cond {rc.max():.2e} rel {rel.max():.2e}")
            assert rc.max() <= 1e-12
            assert g["init"]["ev"] == pytest.approx(m_o["ev"], rel=1e-13)
            assert g["init"]["px"] == pytest.approx(m_o["px"], rel=1e-13)
        e.run(steps, nind)
        together = [e.get(rid) for rid in ids]
    for w, t in zip(members, together):
        with Ensemble(capacity=(w.n + 255) // 256 * 256, max_replicas=1, max_records=4) as e:
            rid = e.add_workload(w, steps=steps)
            e.run(steps, nind)
            alone = e.get(rid)
        for key in ("x", "v", "a"):
            assert np.array_equal(t[key], alone[key]), (w.name, w.n, key)
        assert t["init"] == alone["init"] and t["records"] == alone["records"]
        assert len(t["records"]) == steps // nind


def test_serving_window_served_equals_alone():
    """Requests 0-9 of the seed-1004 stream (N from 3,282 to 243,423, on the
    committed table's cluster lists and tile items; FP64 and mixed), steps
    capped at 20, arriving at 10 per second: each result is bitwise the same as
    the request alone; the FP64 request 0 (N = 3,851) follows the oracle."""
    reqs = stream(10, rate_per_s=10.0)
    for r in reqs:
        r.steps = min(r.steps, 20)
    assert any(r.n > 65536 for r in reqs) and any(r.precision == "fp64" for r in reqs)
    with Server(hbm_budget_bytes=16 * 2 ** 30, capacity_slots=1 << 20, max_runs=16,
                round_steps=10, nind=10, max_records=4) as srv:
        results = srv.serve(reqs)
        assert srv.peak_live >= 2
    assert sorted(r.request.id for r in results) == list(range(10))
    for res in results:
        req = res.request
        w = req.workload()
        with Ensemble(capacity=(req.n + 255) // 256 * 256, max_replicas=1, max_records=4) as e:
            rid = e.add_workload(w, steps=req.steps)
            e.run(req.steps, 10)
            ref = e.get(rid)
        for key in ("x", "v", "a"):
            assert np.array_equal(getattr(res, key), ref[key]), (req.id, req.n, key)
        assert res.init == ref["init"] and res.records == ref["records"]
        if req.id == 0:
            assert req.precision == "fp64"
            p = w.params
            a, m0 = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=0)
            assert res.init["ev"] == pytest.approx(m0["ev"], rel=1e-13)
            x, v, vold = w.x, w.v, np.zeros_like(w.v)
            for st in range(1, req.steps + 1):
                x, v, vold, a, m = oracle.newton(p, w.dt, x, v, vold, a, w.Z, w.A, st % 10 == 0)
            L = p.xl[0]
            dx = np.abs(res.x - x)
            dx = np.minimum(dx, L - dx)
            assert dx.max() <= 1e-10 * L
            np.testing.assert_allclose(res.v, v, rtol=1e-8, atol=1e-11 * np.abs(v).max())
            assert res.records[-1]["ev"] == pytest.approx(m["ev"], rel=1e-11)
            assert res.records[-1]["ek"] == pytest.approx(m["ek"], rel=1e-10)


# --------------------------------------------------------------------------
def energy_drift(w, steps=1000, nind=10, warm_groups=20, warm_steps=25):
    """Relative drift max |E(t) - E(0)| / |E(0)| of E = ev + ek per particle
    over `steps` NVE steps in each precision, from the state of a thermostatted
    FP64 warm-up (warm_groups groups of warm_steps steps, each followed by
    cancel_vcom + tnorm, as src/md.F:88-101)."""
    from yukawamd_amd.system import from_workload
    with from_workload(w, precision="fp64") as s:
        s.accel(True)
        for _ in range(warm_groups):
            s.run(warm_steps)
            s.cancel_vcom()
        x, v, _ = s.get_state()
    out = {}
    for prec in ("fp64", "mixed"):
        with from_workload(w, precision=prec) as s:
            s.put_state(x, v)
            e0 = s.accel(True)["ev"] + s.ttot()
            recs = s.run(steps, nind)
        e = np.array([r["ev"] + r["ek"] for r in recs])
        out[prec] = float(np.abs(e - e0).max() / abs(e0))
    return out


def test_energy_drift_ion_mixture_27648():
    """The headline workload over 1,000 steps: the mixed form's relative drift
    is at most 1e-5 (the stated tolerance) and no larger than the FP64 form's
    beyond 1e-7."""
    d = energy_drift(W.ion_mixture())
    print(f"ion_mixture_27648 drift per 1,000 steps: fp64 {d['fp64']:.3e} mixed {d['mixed']:.3e}")
    assert d["fp64"] <= 1e-5
    assert d["mixed"] <= 1e-5 and d["mixed"] <= d["fp64"] + 1e-7


def test_energy_drift_batched_replicas():
    """Replicas of the 256-replica ensemble draw, one per N bucket up to 16,384,
    in the batched engine's mixed form: from an FP64-warmed state (20 groups of
    25 steps, each followed by cancel_vcom), 1,000 steps measuring every 10th;
    each replica's relative drift is at most 1e-5, the single systems' bar."""
    ens = W.ensemble(256)
    picks = [next(w for w in ens if lo < w.n <= hi)
             for lo, hi in ((1023, 2048), (2048, 4096), (4096, 8192), (8192, 16384))]
    warmed = []
    for w in picks:
        with YukawaSystem(w.params.replace(precision="fp64"), Z=w.Z, A=w.A, dt=w.dt) as s:
            s.put_state(w.x, w.v)
            s.accel(True)
            for _ in range(20):
                s.run(25)
                s.cancel_vcom()
            x, v, _ = s.get_state()
            s.put_state(x, v)
            e0 = s.accel(True)["ev"] + s.ttot()
        warmed.append((dataclasses.replace(w, x=x, v=v), e0))
    cap = sum(((w.n + 255) // 256) * 256 for w, _ in warmed)
    with Ensemble(capacity=cap, max_replicas=4, max_records=100) as e:
        ids = [e.add_workload(w, steps=1000) for w, _ in warmed]
        e.run(1000, nind=10)
        for rid, (w, e0) in zip(ids, warmed):
            recs = e.get(rid)["records"]
            assert len(recs) == 100
            drift = max(abs(r["ev"] + r["ek"] - e0) for r in recs) / abs(e0)
            print(f"{w.name} (N = {w.n}, mixed, batched) drift per 1,000 steps: {drift:.3e}")
            assert drift <= 1e-5


def test_energy_drift_nucleon_51200():
    """Nucleons (dt = 2 fm/c against the HPP core's ~15 fm/c period): the
    drift is the integrator's, the same in both precisions within 20 %."""
    d = energy_drift(W.nucleon())
    print(f"nucleon_51200 drift per 1,000 steps: fp64 {d['fp64']:.3e} mixed {d['mixed']:.3e}")
    assert d["fp64"] <= 1e-3
    assert d["mixed"] <= 1.2 * d["fp64"]
