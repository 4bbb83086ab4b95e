"""Batched ensembles of independent ion systems over the C ABI (yk_batch_*).

``Ensemble`` owns one device arena.  Replicas (pure or mixture, FP64 or
mixed, each with its own N, box, screening, cutoff, timestep and step budget)
are admitted with ``add`` -- refused with ``YK_ERR_CAPACITY`` when the arena
or the replica table is full -- and advance together with ``run``: one
persistent pair launch per step pulling (replica, tile, source range) items
from a device queue, the step replayed from a captured hipGraph.  ``remove``
frees a replica's slots for the next admission (continuous batching).
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

from . import _lib
from .system import _as_aos, _dptr

TUNING_FILE = Path(__file__).resolve().parent / "tuning" / "batch_gfx950.json"


class Ensemble:
    def __init__(self, capacity: int, max_replicas: int = 256, max_records: int = 64,
                 tuning: dict | None = None):
        self.lib = _lib.load()
        self._b = C.c_void_p()
        _lib.check_batch(self.lib.yk_batch_create(int(capacity), int(max_replicas),
                                                  int(max_records), C.byref(self._b)))
        self.max_records = int(max_records)
        self.replicas = {}
        self.tuning = []
        if tuning is None and TUNING_FILE.exists():
            tuning = json.loads(TUNING_FILE.read_text())["buckets"]
        if tuning:
            self.set_tuning(tuning)

    def close(self):
        if self._b:
            self.lib.yk_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def footprint(self, n: int):
        """(arena slots, device bytes) a replica of n particles would hold."""
        slots, nbytes = C.c_int64(), C.c_int64()
        _lib.check_batch(self.lib.yk_batch_footprint(self._b, int(n), C.byref(slots),
                                                     C.byref(nbytes)))
        return slots.value, nbytes.value

    def set_tuning(self, buckets):
        """buckets: [{"n_upper": N, "chunk_tiles": C, "waves_per_block": W,
        "list_items": M}, ...] in increasing n_upper; list_items > 0 selects
        cluster pair lists with M work items per tile, 0 tile items of C source
        tiles."""
        nb = len(buckets)
        a = (C.c_int32 * nb)(*[int(b["n_upper"]) for b in buckets])
        c = (C.c_int32 * nb)(*[int(b.get("chunk_tiles", 8)) for b in buckets])
        w = (C.c_int32 * nb)(*[int(b["waves_per_block"]) for b in buckets])
        m = (C.c_int32 * nb)(*[int(b.get("list_items", 0)) for b in buckets])
        _lib.check_batch(self.lib.yk_batch_set_tuning(self._b, nb, a, c, w, m))
        self.tuning = list(buckets)

    def set_resort(self, every: int):
        """Re-sort each replica into Hilbert order every `every` of its own steps
        (0: never)."""
        _lib.check_batch(self.lib.yk_batch_set_resort(self._b, int(every)))

    def add(self, params, x, v, Z, A, dt: float, steps: int) -> int:
        n = x.shape[0]
        xa = _as_aos(x, n)
        va = _as_aos(v, n) if v is not None else None
        Za = np.ascontiguousarray(Z, dtype=np.float64)
        Aa = np.ascontiguousarray(A, dtype=np.float64)
        ps = _lib.params_struct(params)
        rid = C.c_int32()
        _lib.check_batch(self.lib.yk_batch_add(self._b, C.byref(ps), n, _dptr(Za), _dptr(Aa),
                                               _dptr(xa), _dptr(va), float(dt), int(steps),
                                               C.byref(rid)))
        self.replicas[rid.value] = n
        return rid.value

    def add_workload(self, w, steps: int, **over) -> int:
        p = w.params.replace(**over) if over else w.params
        return self.add(p, w.x, w.v, w.Z, w.A, w.dt, steps)

    def remove(self, rid: int):
        _lib.check_batch(self.lib.yk_batch_remove(self._b, int(rid)))
        self.replicas.pop(rid, None)

    def run(self, nsteps: int, nind: int = 0):
        _lib.check_batch(self.lib.yk_batch_run(self._b, int(nsteps), int(nind)))

    def status(self, rid: int):
        left, nrec = C.c_int32(), C.c_int32()
        _lib.check_batch(self.lib.yk_batch_status(self._b, int(rid), C.byref(left),
                                                  C.byref(nrec)))
        return left.value, nrec.value

    def get(self, rid: int):
        n = self.replicas[rid]
        x, v, a = np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3))
        init = _lib.YkMeas()
        recs = (_lib.YkMeas * self.max_records)()
        nrec = C.c_int32()
        _lib.check_batch(self.lib.yk_batch_get(self._b, int(rid), _dptr(x), _dptr(v), _dptr(a),
                                               C.byref(init), recs, self.max_records,
                                               C.byref(nrec)))
        return {"x": x, "v": v, "a": a, "init": init.as_dict(),
                "records": [recs[k].as_dict() for k in range(nrec.value)]}

    def last_run(self):
        ms, steps, items = C.c_double(), C.c_int64(), C.c_int64()
        _lib.check_batch(self.lib.yk_batch_last_run(self._b, C.byref(ms), C.byref(steps),
                                                    C.byref(items)))
        return {"ms": ms.value, "steps": steps.value, "items": items.value}
