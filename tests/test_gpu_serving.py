"""Serving loop on the GPU: continuous batching of MD requests.

* every request's result (positions, velocities, accelerations, initial and
  periodic measurement records) is bitwise the same served among others,
  admitted at different times, as run alone in an empty engine;
* admission never exceeds the HBM budget, and requests that do not fit wait
  and complete later.
"""
import numpy as np
import pytest

from yukawamd_amd.batch import Ensemble
from yukawamd_amd.serving import Request, Server, summarize

pytestmark = pytest.mark.gpu


def requests():
    spec = [  # n, steps, precision, mixture, arrival_s
        (700, 40, "mixed", True, 0.0),
        (1500, 30, "fp64", True, 0.0),
        (2300, 60, "mixed", False, 0.0),
        (400, 20, "fp64", False, 0.0),
        (3100, 50, "mixed", True, 0.05),
        (1000, 70, "mixed", True, 0.2),
    ]
    return [Request(id=k, arrival_s=a, n=n, steps=st, gamma=150.0 + 10 * k, kappa=0.6 + 0.2 * k,
                    precision=p, seed=7000 + k, mixture=m)
            for k, (n, st, p, m, a) in enumerate(spec)]


def alone(req, nind):
    w = req.workload()
    with Ensemble(capacity=4096, max_replicas=1, max_records=16) as e:
        rid = e.add_workload(w, steps=req.steps)
        e.run(req.steps, nind)
        assert e.status(rid)[0] == 0
        return e.get(rid)


def test_served_equals_alone():
    reqs = requests()
    with Server(hbm_budget_bytes=2 ** 30, capacity_slots=16384, max_runs=8, round_steps=10,
                nind=10, max_records=16) as srv:
        results = srv.serve(reqs)
        assert srv.peak_live >= 4          # the runs did overlap on the device
    assert sorted(r.request.id for r in results) == [r.id for r in reqs]
    for res in results:
        ref = alone(res.request, 10)
        for key in ("x", "v", "a"):
            assert np.array_equal(getattr(res, key), ref[key]), (res.request.id, key)
        assert res.init == ref["init"]
        assert res.records == ref["records"]
        assert len(res.records) == res.request.steps // 10


def test_admission_by_footprint():
    reqs = requests()
    # the engine's default work shape (footprints then do not depend on the
    # tuning table)
    probe = Server(hbm_budget_bytes=2 ** 30, capacity_slots=16384, max_runs=8, tuning=[])
    sizes = {r.id: probe.footprint(r) for r in reqs}
    probe.close()
    budget = max(sizes.values())      # the largest run alone, or a few small ones
    first = [r for r in reqs if r.arrival_s == 0.0]
    assert sum(sizes[r.id] for r in first) > budget
    with Server(hbm_budget_bytes=budget, capacity_slots=16384, max_runs=8,
                round_steps=10, tuning=[]) as srv:
        results = srv.serve(reqs)
        assert srv.peak_bytes <= budget
        s = summarize(results, 1.0, srv.device_ms)
    assert s["requests"] == len(reqs)
    first_done = min(r.done_s for r in results)
    waited = [r for r in results if r.request.arrival_s == 0.0 and r.admitted_s >= first_done]
    assert waited                      # some request was held back until a run finished
    with Server(hbm_budget_bytes=budget, capacity_slots=16384, max_runs=8, tuning=[]) as srv:
        big = Request(id=99, arrival_s=0.0, n=16000, steps=1, gamma=170.0, kappa=1.0,
                      precision="mixed", seed=1)
        with pytest.raises(ValueError):
            srv.submit(big)
