"""Host logic of the serving loop (no GPU): admission by footprint, backfill,
the bypass limit, retirement, latency bookkeeping.  A stand-in engine counts
steps down the way the device does (a run stops when its budget is used up)
and refuses admission when its slot table is full."""
import pytest

from yukawamd_amd.serving import Request, Server, summarize


class FakeEngine:
    def __init__(self, slots):
        self.slots = slots
        self.used = 0
        self.runs = {}
        self.next_id = 0
        self.steps_run = 0

    def footprint(self, n):
        return n, 100 * n

    def add_workload(self, w, steps):
        if self.used + w.n > self.slots:
            e = RuntimeError("capacity")
            e.code = 7
            raise e
        rid = self.next_id
        self.next_id += 1
        self.runs[rid] = [w.n, steps]
        self.used += w.n
        return rid

    def run(self, nsteps, nind):
        for r in self.runs.values():
            r[1] = max(0, r[1] - nsteps)
        self.steps_run += nsteps

    def last_run(self):
        return {"ms": 1.0}

    def status(self, rid):
        return self.runs[rid][1], 0

    def get(self, rid):
        return {"x": None, "v": None, "a": None, "init": {}, "records": []}

    def remove(self, rid):
        self.used -= self.runs.pop(rid)[0]

    def close(self):
        pass


class FakeWorkload:
    def __init__(self, n):
        self.n = n


class Clock:
    def __init__(self):
        self.t = 0.0

    def __call__(self):
        self.t += 0.001
        return self.t

    def sleep(self, dt):
        self.t += dt


def req(k, n, steps, arrival):
    r = Request(id=k, arrival_s=arrival, n=n, steps=steps, gamma=170.0, kappa=1.0,
                precision="mixed", seed=k)
    r.workload = lambda n=n: FakeWorkload(n)
    return r


def test_all_served_within_budget():
    eng = FakeEngine(slots=10_000)
    reqs = [req(k, 1000 + 500 * (k % 4), 30 + 10 * k, 0.01 * k) for k in range(12)]
    clk = Clock()
    srv = Server(hbm_budget_bytes=400_000, round_steps=10, ensemble=eng)
    res = srv.serve(reqs, clock=clk, sleep=clk.sleep)
    assert sorted(r.request.id for r in res) == list(range(12))
    assert srv.peak_bytes <= 400_000
    assert srv.used_bytes == 0 and not srv.live and not srv.pending
    for r in res:
        assert r.done_s >= r.admitted_s >= r.request.arrival_s
    s = summarize(res, clk.t, srv.device_ms)
    assert s["p50_latency_s"] <= s["p99_latency_s"] <= s["max_latency_s"]


def test_backfill_and_bypass_limit():
    # head needs the whole budget; a small request arriving later may pass it
    # while the bypass limit allows, never after
    eng = FakeEngine(slots=100_000)
    clk = Clock()
    srv = Server(hbm_budget_bytes=300_000, round_steps=10, ensemble=eng, max_bypass_s=1e9)
    reqs = [req(0, 2000, 100, 0.0), req(1, 3000, 10, 0.0), req(2, 500, 10, 0.0)]
    res = srv.serve(reqs, clock=clk, sleep=clk.sleep)
    order = {r.request.id: r.admitted_s for r in res}
    assert order[2] < order[1]          # backfilled past the blocked request 1
    eng = FakeEngine(slots=100_000)
    clk = Clock()
    srv = Server(hbm_budget_bytes=300_000, round_steps=10, ensemble=eng, max_bypass_s=0.0)
    res = srv.serve(reqs, clock=clk, sleep=clk.sleep)
    order = {r.request.id: r.admitted_s for r in res}
    assert order[2] >= order[1]         # no bypass allowed


def test_oversized_request_rejected():
    srv = Server(hbm_budget_bytes=1000, ensemble=FakeEngine(slots=100))
    with pytest.raises(ValueError):
        srv.submit(req(0, 20, 5, 0.0))


def test_round_steps_multiple_of_nind():
    with pytest.raises(ValueError):
        Server(round_steps=15, nind=10, ensemble=FakeEngine(slots=10))
