"""yk_accel_noalloc: accel on device-resident arrays with no allocation,
against yk_accel (the handle-based entry point) through the C ABI.

Two calls in a row share the library's static work memory: a 3,000-ion
mixture measuring, then a smaller 1,100-ion pure system (zii/aii NULL), each
checked against yk_accel on the same positions: FP64 condition-aware
|dF_i| / sum_j |F_ij| <= 1e-12 and energy / pressure / tensor within 1e-13;
mixed precision condition-aware <= 1e-6 and energy within 1e-6.  The test
allocates its device buffers with hipMalloc itself (the entry point does not).
"""
import ctypes as C
import os

import numpy as np
import pytest

from _forces import check_meas, force_errors
from yukawamd_amd import _lib
from yukawamd_amd import workloads as W
from yukawamd_amd._lib import YkError, YkMeas
from yukawamd_amd.system import YukawaSystem

pytestmark = pytest.mark.gpu

MAX_N = 16384   # YK_NOALLOC_MAX_N


def hip():
    for name in ("libamdhip64.so", os.path.join("/opt/rocm", "lib", "libamdhip64.so")):
        try:
            h = C.CDLL(name)
            break
        except OSError:
            continue
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipFree.argtypes = [C.c_void_p]
    for f in (h.hipMalloc, h.hipMemcpy, h.hipFree, h.hipDeviceSynchronize):
        f.restype = C.c_int
    return h


class Dev:
    """A device buffer holding a float64 numpy array (test plumbing)."""

    def __init__(self, h, arr):
        self.h = h
        self.arr = np.ascontiguousarray(arr, dtype=np.float64)
        self.nbytes = self.arr.nbytes
        self.ptr = C.c_void_p()
        assert h.hipMalloc(C.byref(self.ptr), self.nbytes) == 0
        assert h.hipMemcpy(self.ptr, self.arr.ctypes.data, self.nbytes, 1) == 0   # H2D

    def get(self):
        out = np.empty_like(self.arr)
        assert self.h.hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2) == 0   # D2H
        return out

    def free(self):
        self.h.hipFree(self.ptr)


def noalloc_accel(lib, h, p, x, Z, A, meas, with_za=True):
    n = x.shape[0]
    dx = Dev(h, x)
    dz = Dev(h, Z) if with_za else None
    da = Dev(h, A) if with_za else None
    dacc = Dev(h, np.full((n, 3), np.nan))
    dm = Dev(h, np.full(10, np.nan))
    ps = _lib.params_struct(p)
    try:
        _lib.check(lib.yk_accel_noalloc(C.byref(ps), n, dx.ptr, dz.ptr if dz else None,
                                        da.ptr if da else None, dacc.ptr, 1 if meas else 0,
                                        dm.ptr, None))
        assert h.hipDeviceSynchronize() == 0
        a = dacc.get()
        mv = dm.get()
    finally:
        for d in (dx, dz, da, dacc, dm):
            if d is not None:
                d.free()
    m = YkMeas()
    C.memmove(C.byref(m), mv.ctypes.data, C.sizeof(YkMeas))
    return a, m.as_dict()


def handle_accel(p, x, Z, A, meas):
    with YukawaSystem(p, Z=Z, A=A) as s:
        s.put_state(x)
        m = s.accel(meas)
        return s.a, m


def as_ref_meas(m):
    d = {"ev": m["ev"], "px": m["px"]}
    for k, key in enumerate(("pp11", "pp12", "pp13", "pp22", "pp23", "pp33")):
        d[key] = m["pp"][k]
    return d


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_accel_on_two_calls(precision):
    lib = _lib.load()
    h = hip()
    cases = [(W.ion_mixture(n=3000, seed=141, precision=precision), True, True),
             (W.pure_ion(n=1100, seed=142, precision=precision), True, False)]
    for w, meas, with_za in cases:
        p = w.params
        a_n, m_n = noalloc_accel(lib, h, p, w.x, w.Z, w.A, meas, with_za)
        a_r, m_r = handle_accel(p, w.x, w.Z, w.A, meas)
        rel, cond = force_errors(p, a_n, a_r, w.x, w.Z, w.A)
        print(f"{w.name} {precision}: max rel {rel:.2e} cond {cond:.2e} "
              f"ev {m_n['ev']:.12e} vs {m_r['ev']:.12e}")
        assert np.isfinite(a_n).all()
        if precision == "fp64":
            assert cond <= 1e-12 and rel <= 1e-10
            check_meas(m_n, as_ref_meas(m_r))
        else:
            assert cond <= 1e-6
            assert m_n["ev"] == pytest.approx(m_r["ev"], rel=1e-6)


def test_noalloc_not_measuring_zeroes_meas():
    lib = _lib.load()
    h = hip()
    w = W.ion_mixture(n=900, seed=143)
    a_n, m_n = noalloc_accel(lib, h, w.params, w.x, w.Z, w.A, False)
    a_r, _ = handle_accel(w.params, w.x, w.Z, w.A, False)
    rel, cond = force_errors(w.params, a_n, a_r, w.x, w.Z, w.A)
    assert cond <= 1e-12
    assert m_n["ev"] == 0.0 and m_n["px"] == 0.0 and all(v == 0.0 for v in m_n["pp"])


def test_noalloc_refuses_without_launching():
    """Sizes beyond the limit and host (non-device) arrays are refused with
    YK_ERR_ARG before any work is enqueued."""
    lib = _lib.load()
    h = hip()
    w = W.ion_mixture(n=64, seed=144)
    ps = _lib.params_struct(w.params)
    dx = Dev(h, w.x)
    dacc = Dev(h, np.zeros((64, 3)))
    dz, da = Dev(h, w.Z), Dev(h, w.A)
    try:
        rc = lib.yk_accel_noalloc(C.byref(ps), MAX_N + 1, dx.ptr, dz.ptr, da.ptr, dacc.ptr, 0,
                                  None, None)
        assert rc == 1 and b"no-allocation limit" in lib.yk_last_error()
        host = np.ascontiguousarray(w.x)
        rc = lib.yk_accel_noalloc(C.byref(ps), 64, C.c_void_p(host.ctypes.data), dz.ptr,
                                  da.ptr, dacc.ptr, 0, None, None)
        assert rc == 1 and b"not device memory" in lib.yk_last_error()
        rc = lib.yk_accel_noalloc(C.byref(ps), 200, dx.ptr, dz.ptr, da.ptr, dacc.ptr, 0,
                                  None, None)     # x holds 64 particles only
        assert rc == 1 and b"fewer than" in lib.yk_last_error()
        with pytest.raises(YkError):
            _lib.check(rc)
    finally:
        for d in (dx, dacc, dz, da):
            d.free()
