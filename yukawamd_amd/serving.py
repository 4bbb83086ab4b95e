"""Serving loop: MD runs accepted as they arrive, continuously batched on one GPU.

A request names an ion system (the three-species crust mixture or pure Z=8,
N, Gamma, kappa, a seed for its initial state), a step count and a precision
flag.  ``Server`` keeps every admitted run's state resident in one device
arena (``Ensemble``) and advances all of them together in rounds of
``round_steps`` velocity-Verlet steps (each round replays a captured hipGraph;
a run whose budget ends inside a round idles for the rest of it; with
``nind`` > 0 every nind-th step of a run is measured, counted from its own
first step).  Between rounds it

  * retires finished runs (downloads x, v, a and the measurement records,
    frees their slots),
  * admits arrived requests in arrival order while their HBM footprint
    (arena slots + work-item partial sums, as the engine reports it) fits in
    the free budget and the arena holds a contiguous run of slots for them;
    a request that does not fit waits, and later, smaller arrivals may pass
    it (backfill) unless it has waited longer than ``max_bypass_s``.  The
    budget counts the engine's per-request footprints; the engine's own growth
    headroom comes on top (its list arena at least doubles when it grows and
    holds both copies while it copies; partial sums are allocated 25 % ahead),
    and an admission that finds the device short of memory waits for a
    retirement like one that finds the arena full.  The next pending requests'
    initial states are generated on host threads while the device runs.

A run's arithmetic depends only on its own state: its result is the same
whether it ran alone or among others, admitted early or late.

Latency is measured from a request's arrival to the end of the round in which
its last step ran (its result downloaded).
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import workloads as W
from .batch import Ensemble


@dataclass
class Request:
    id: int
    arrival_s: float
    n: int
    steps: int
    gamma: float
    kappa: float
    precision: str
    seed: int
    mixture: bool = True

    @classmethod
    def from_dict(cls, d):
        return cls(**{k: d[k] for k in ("id", "arrival_s", "n", "steps", "gamma", "kappa",
                                         "precision", "seed") if k in d},
                   mixture=d.get("mixture", True))

    def workload(self):
        return W.replica(self.seed, self.n, self.mixture, self.gamma, self.kappa,
                         self.precision)

    @property
    def pair_steps(self) -> float:
        return self.n * (self.n - 1) / 2 * self.steps


@dataclass
class Result:
    request: Request
    admitted_s: float
    done_s: float
    x: np.ndarray
    v: np.ndarray
    a: np.ndarray
    init: dict
    records: list

    @property
    def latency_s(self) -> float:
        return self.done_s - self.request.arrival_s


@dataclass
class _Live:
    request: Request
    rid: int
    bytes: int
    admitted_s: float


class Server:
    def __init__(self, hbm_budget_bytes: float = 64 * 2 ** 30, capacity_slots: int | None = None,
                 max_runs: int = 1024, round_steps: int = 50, nind: int = 0,
                 max_records: int = 64, max_bypass_s: float = 30.0, tuning=None,
                 ensemble=None, log=None):
        if nind and round_steps % nind:
            raise ValueError("round_steps must be a multiple of nind")
        self.budget = int(hbm_budget_bytes)
        if ensemble is not None:
            self.ens = ensemble
        else:
            if capacity_slots is None:
                # slot arrays are ~130 B per particle, the work-item partial sums
                # of a large run many times more: 1 KiB of budget per slot
                capacity_slots = max(1 << 16, min(self.budget // 1024, 1 << 25))
            self.ens = Ensemble(capacity=capacity_slots, max_replicas=max_runs,
                                max_records=max_records, tuning=tuning)
        self.log = log            # callable(str) told of every completion, or None
        self.round_steps = int(round_steps)
        self.nind = int(nind)
        self.max_bypass_s = float(max_bypass_s)
        # initial states of the next pending requests, generated on host threads
        # while the device runs a round (the engine call releases the GIL)
        self._gen = ThreadPoolExecutor(max_workers=2)
        self._states: dict = {}
        self.prefetch = 8
        self.pending: list[Request] = []
        self.live: dict[int, _Live] = {}
        self.used_bytes = 0
        self.results: list[Result] = []
        self.rounds = 0
        self.device_ms = 0.0
        self.peak_bytes = 0
        self.peak_live = 0

    def close(self):
        self._gen.shutdown(wait=True)
        self.ens.close()

    def _prefetch(self):
        for req in self.pending[:self.prefetch]:
            if req.id not in self._states:
                self._states[req.id] = self._gen.submit(req.workload)

    def _workload(self, req):
        fut = self._states.pop(req.id, None)
        return fut.result() if fut is not None else req.workload()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- admission -----------------------------------------------------------
    def footprint(self, req: Request) -> int:
        return self.ens.footprint(req.n)[1]

    def submit(self, req: Request):
        if self.footprint(req) > self.budget:
            raise ValueError(f"request {req.id}: N={req.n} needs {self.footprint(req)} bytes, "
                             f"more than the whole budget of {self.budget}")
        self.pending.append(req)

    def _admit(self, now: float) -> int:
        admitted = 0
        blocked_head = None
        keep = []
        for req in self.pending:
            if blocked_head is not None and now - blocked_head.arrival_s > self.max_bypass_s:
                keep.append(req)          # the head has waited long enough: no backfill
                continue
            need = self.footprint(req)
            ok = self.used_bytes + need <= self.budget
            rid = None
            if ok:
                w = self._workload(req)
                try:
                    rid = self.ens.add_workload(w, steps=req.steps)
                except Exception as e:
                    # arena fragmentation / replica table full (YK_ERR_CAPACITY),
                    # or device memory short of the engine's own growth headroom
                    # (YK_ERR_NOMEM): the request waits for a retirement
                    if getattr(e, "code", None) not in (3, 7):
                        raise
                    rid = None
                    self._states[req.id] = self._gen.submit(lambda w=w: w)
            if rid is None:
                if blocked_head is None:
                    blocked_head = req
                keep.append(req)
                continue
            self.live[rid] = _Live(req, rid, need, now)
            self.used_bytes += need
            self.peak_bytes = max(self.peak_bytes, self.used_bytes)
            self.peak_live = max(self.peak_live, len(self.live))
            admitted += 1
        self.pending = keep
        return admitted

    # -- rounds --------------------------------------------------------------
    def _round(self):
        self.ens.run(self.round_steps, self.nind)
        self.device_ms += self.ens.last_run()["ms"]
        self.rounds += 1

    def _retire(self, now: float):
        for rid in list(self.live):
            left, _ = self.ens.status(rid)
            if left == 0:
                lv = self.live.pop(rid)
                g = self.ens.get(rid)
                self.ens.remove(rid)
                self.used_bytes -= lv.bytes
                self.results.append(Result(lv.request, lv.admitted_s, now, g["x"], g["v"],
                                           g["a"], g["init"], g["records"]))
                if self.log:
                    r = lv.request
                    self.log(f"served request {r.id} (N={r.n}, {r.steps} steps, {r.precision}) "
                             f"at {now:.1f} s, latency {now - r.arrival_s:.2f} s; "
                             f"{len(self.results)} done, {len(self.live)} running, "
                             f"{len(self.pending)} waiting")

    def serve(self, requests, clock=time.perf_counter, sleep=time.sleep):
        """Serve `requests` (arrival_s relative to the call) to completion."""
        reqs = sorted((r if isinstance(r, Request) else Request.from_dict(r) for r in requests),
                      key=lambda r: r.arrival_s)
        t0 = clock()
        k = 0
        while k < len(reqs) or self.pending or self.live:
            now = clock() - t0
            while k < len(reqs) and reqs[k].arrival_s <= now:
                self.submit(reqs[k])
                k += 1
            self._admit(now)
            self._prefetch()
            if self.live:
                self._round()
                self._retire(clock() - t0)
            elif k < len(reqs):
                sleep(max(0.0, reqs[k].arrival_s - (clock() - t0)))
            elif self.pending:
                raise RuntimeError("pending requests cannot be admitted into an empty server")
        return self.results


def summarize(results, wall_s: float, device_ms: float) -> dict:
    lat = np.array([r.latency_s for r in results])
    pair_steps = sum(r.request.pair_steps for r in results)
    return {
        "requests": len(results),
        "p50_latency_s": float(np.percentile(lat, 50)) if len(lat) else None,
        "p99_latency_s": float(np.percentile(lat, 99)) if len(lat) else None,
        "max_latency_s": float(lat.max()) if len(lat) else None,
        "mean_queue_s": float(np.mean([r.admitted_s - r.request.arrival_s for r in results]))
        if results else None,
        "wall_s": wall_s,
        "device_busy_s": device_ms * 1e-3,
        "pair_interactions_per_s": pair_steps / wall_s if wall_s > 0 else None,
        "pair_interactions_per_device_s": pair_steps / (device_ms * 1e-3) if device_ms else None,
        "particle_steps": float(sum(r.request.n * r.request.steps for r in results)),
    }


def stream(count: int, rate_per_s: float, seed: int = 1004, **kw):
    """The first `count` requests of the seeded serving stream (Poisson arrivals
    at rate_per_s; N log-uniform [2048, 262144], steps U[100, 2000], Gamma
    U[150, 220], kappa U[0.5, 2], 80% mixed precision)."""
    return [Request.from_dict(d) for d in W.serving_requests(count, seed, rate_per_s=rate_per_s,
                                                             **kw)]


def expected_pair_steps(count: int = 2000, seed: int = 1004) -> float:
    return sum(r["n"] * (r["n"] - 1) / 2 * r["steps"] for r in W.serving_requests(count, seed))


__all__ = ["Request", "Result", "Server", "summarize", "stream", "expected_pair_steps"]
