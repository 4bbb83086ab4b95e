"""Host-side mirror of the reference's hot-path interface over the C ABI.

``YukawaSystem`` holds one device-resident particle system and exposes the
reference's operator names with the same argument meaning:

    accel(do_measurements)        src/int_control.f:52-81
    newton(do_measurements)       src/vel_verlet_2.f:18-120 (dt from self.dt)
    ttot()                        src/gen_utils.f:40-55

After each call the scalar outputs the reference leaves in ``md_globals``
(ev, ek, px, pp) are available as attributes, and ``time`` advances by dt per
newton step as in src/vel_verlet_2.f:114.  Errors raise ``YkError`` instead
of the reference's md_exit (src/md_utils.f:158-183).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from .params import SimParams


def _dptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _as_aos(a, n) -> np.ndarray:
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64))
    if arr.shape != (n, 3):
        raise ValueError(f"expected an array of shape ({n}, 3), got {arr.shape}")
    return arr


class YukawaSystem:
    def __init__(self, params: SimParams, Z=None, A=None, n: int | None = None,
                 dt: float = 0.0):
        self.lib = _lib.load()
        self.params = params
        if Z is not None:
            Z = np.ascontiguousarray(Z, dtype=np.float64)
            n = Z.shape[0] if n is None else n
        if A is not None:
            A = np.ascontiguousarray(A, dtype=np.float64)
        if n is None:
            raise ValueError("n or Z is required")
        if Z is not None and Z.shape != (n,):
            raise ValueError(f"Z must have shape ({n},)")
        if A is not None and A.shape != (n,):
            raise ValueError(f"A must have shape ({n},)")
        self.n = int(n)
        self.Z, self.A = Z, A
        self.dt = float(dt)
        self.time = 0.0
        self.ev = self.ek = self.px = self.vir = 0.0
        self.pp = np.zeros(6)
        self._h = C.c_void_p()
        ps = _lib.params_struct(params)
        _lib.check(self.lib.yk_create(C.byref(ps), self.n, _dptr(Z), _dptr(A),
                                      C.byref(self._h)))

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if self._h:
            self.lib.yk_destroy(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------
    def set_params(self, params: SimParams):
        ps = _lib.params_struct(params)
        _lib.check(self.lib.yk_set_params(self._h, C.byref(ps)))
        self.params = params

    def put_state(self, x, v=None):
        xa = _as_aos(x, self.n)
        va = _as_aos(v, self.n) if v is not None else None
        _lib.check(self.lib.yk_put_state(self._h, _dptr(xa), _dptr(va)))

    def put_accel(self, a):
        aa = _as_aos(a, self.n)
        _lib.check(self.lib.yk_put_accel(self._h, _dptr(aa)))

    def get_state(self):
        x = np.empty((self.n, 3))
        v = np.empty((self.n, 3))
        a = np.empty((self.n, 3))
        _lib.check(self.lib.yk_get_state(self._h, _dptr(x), _dptr(v), _dptr(a)))
        return x, v, a

    @property
    def a(self) -> np.ndarray:
        a = np.empty((self.n, 3))
        _lib.check(self.lib.yk_get_state(self._h, None, None, _dptr(a)))
        return a

    # -- hot path ----------------------------------------------------------
    def _take(self, m: _lib.YkMeas) -> dict:
        d = m.as_dict()
        self.ev, self.ek, self.px, self.vir = d["ev"], d["ek"], d["px"], d["vir"]
        self.pp = np.array(d["pp"])
        return d

    def accel(self, do_measurements: bool = True) -> dict:
        m = _lib.YkMeas()
        _lib.check(self.lib.yk_accel(self._h, 1 if do_measurements else 0, C.byref(m)))
        return self._take(m)

    def newton(self, do_measurements: bool = False, dt: float | None = None) -> dict:
        dt = self.dt if dt is None else float(dt)
        m = _lib.YkMeas()
        _lib.check(self.lib.yk_newton(self._h, dt, 1 if do_measurements else 0, C.byref(m)))
        self.time += dt
        return self._take(m)

    def run(self, nsteps: int, nind: int = 0, dt: float | None = None) -> list:
        """nsteps device-resident steps; a record every nind steps (md.F:129)."""
        dt = self.dt if dt is None else float(dt)
        nrec = nsteps // nind if nind > 0 else 0
        recs = (_lib.YkMeas * max(nrec, 1))()
        _lib.check(self.lib.yk_run(self._h, dt, int(nsteps), int(nind), recs))
        self.time += dt * nsteps
        out = [recs[r].as_dict() for r in range(nrec)]
        if out:
            self._take(recs[nrec - 1])
        return out

    def ttot(self) -> float:
        e = C.c_double()
        _lib.check(self.lib.yk_ttot(self._h, C.byref(e)))
        return e.value

    # -- step-loop utilities (src/gen_utils.f:61-107, src/strain.f) ---------
    def tnorm(self):
        _lib.check(self.lib.yk_tnorm(self._h))

    def cancel_vcom(self):
        _lib.check(self.lib.yk_cancel_vcom(self._h))

    def strain(self, deps, xl0, dt: float | None = None) -> tuple:
        """One step of strain rates deps (None or XNOSTRAIN: unstrained axis)
        from the gage dimensions xl0; returns the new box."""
        dt = self.dt if dt is None else float(dt)
        d = np.array([_lib.XNOSTRAIN if e is None else e for e in deps], dtype=np.float64)
        x0 = np.ascontiguousarray(xl0, dtype=np.float64)
        box = np.zeros(3)
        _lib.check(self.lib.yk_strain(self._h, dt, _dptr(d), _dptr(x0), _dptr(box)))
        self.params = self.params.replace(xl=tuple(float(b) for b in box))
        return self.params.xl

    def run_md(self, nsteps: int, istep0: int = 0, nind: int = 0, nptensor: int = 0,
               ncom: int = 0, ntnorm: int = 0, deps=None, xl0=None,
               dt: float | None = None) -> list:
        """The measurement-phase step loop of src/md.F:124-133 on the device
        (cancel_vcom / tnorm / strain / newton cadence); one record per
        measuring step, with its 'istep'."""
        dt = self.dt if dt is None else float(dt)
        lp = _lib.YkLoop()
        lp.dt, lp.nsteps, lp.istep0 = dt, int(nsteps), int(istep0)
        lp.nind, lp.nptensor, lp.ncom, lp.ntnorm = int(nind), int(nptensor), int(ncom), int(ntnorm)
        for k in range(3):
            e = None if deps is None else deps[k]
            lp.deps[k] = _lib.XNOSTRAIN if e is None else float(e)
            lp.xl0[k] = float((self.params.xl if xl0 is None else xl0)[k])
        cap = max(int(nsteps), 1)
        recs = (_lib.YkMeas * cap)()
        steps = (C.c_int32 * cap)()
        nrec = C.c_int32()
        box = np.zeros(3)
        _lib.check(self.lib.yk_run_md(self._h, C.byref(lp), recs, steps, cap, C.byref(nrec),
                                      _dptr(box)))
        self.time += dt * nsteps
        self.params = self.params.replace(xl=tuple(float(b) for b in box))
        out = [dict(recs[r].as_dict(), istep=int(steps[r])) for r in range(nrec.value)]
        if out:
            self._take(recs[nrec.value - 1])
        return out

    # -- timing --------------------------------------------------------------
    def timing(self) -> dict:
        t = _lib.YkTiming()
        _lib.check(self.lib.yk_get_timing(self._h, C.byref(t)))
        return {"force_ms": t.force_ms, "force_launches": t.force_launches,
                "step_ms": t.step_ms, "steps": t.steps, "pairs_swept": t.pairs_swept}

    def reset_timing(self):
        _lib.check(self.lib.yk_reset_timing(self._h))


def from_workload(w, **over) -> YukawaSystem:
    """Create a system from a workloads.Workload and upload its state."""
    p = w.params.replace(**over) if over else w.params
    Z = w.Z
    A = w.A
    s = YukawaSystem(p, Z=Z, A=A, dt=w.dt)
    s.put_state(w.x, w.v)
    return s
