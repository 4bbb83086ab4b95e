"""Oracle for the yukawamd hot path -- TEST INFRASTRUCTURE ONLY.

Two checkers, both CPU:
  * ``liboracle_yk.so`` (oracle/yk_oracle.c): a C restatement of accel_ion_pure,
    accel_ion_mix, accel_nn and newton (vel_verlet_2) with the reference's loop
    nest and evaluation order, plus a long-double "truth" evaluation;
  * ``_ref/libykref.so``: the reference's own Fortran compiled from its sources
    with oracle/ref_harness.f90 (built only where the reference is present).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use this
package.  The product (yukawamd_amd) never imports it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "_build" / "liboracle_yk.so"
REF_SO = HERE / "_ref" / "libykref.so"

KINDS = {"pure-ion": 0, "ion-mixture": 1, "nucleon": 2, "B-field": 3}
NMEAS = 10
MEAS_KEYS = ("ev", "vir", "px", "pp11", "pp12", "pp13", "pp22", "pp23", "pp33", "ek")


class YkoParams(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("coulomb_on", C.c_int), ("nuclear_on", C.c_int),
        ("ref_f1", C.c_int), ("xl", C.c_double * 3),
        ("xmuc", C.c_double), ("rccut", C.c_double), ("frp", C.c_double),
        ("vc", C.c_double), ("xmass", C.c_double), ("rho", C.c_double), ("kT", C.c_double),
        ("zi", C.c_double), ("ai", C.c_double), ("aa", C.c_double), ("bb", C.c_double),
        ("cc", C.c_double), ("xpacket", C.c_double), ("rncut", C.c_double),
        ("bfield", C.c_double),
    ]


_D = C.POINTER(C.c_double)
_lib = None
_ref = None


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_D)


def load():
    global _lib
    if _lib is None:
        if not ORACLE_SO.exists():
            raise FileNotFoundError(f"{ORACLE_SO} not built (make -C oracle)")
        lib = C.CDLL(str(ORACLE_SO))
        P = C.POINTER(YkoParams)
        lib.yko_accel.argtypes = [P, C.c_int, _D, _D, _D, C.c_int, _D, _D]
        lib.yko_newton.argtypes = [P, C.c_int, C.c_double, _D, _D, _D, _D, _D, _D, C.c_int, _D]
        lib.yko_newton_kdk.argtypes = [P, C.c_int, C.c_double, _D, _D, _D, _D, _D, C.c_int, _D]
        lib.yko_newton_sw.argtypes = [P, C.c_int, C.c_double, _D, _D, _D, _D, _D, C.c_int, _D]
        lib.yko_ttot.argtypes = [P, C.c_int, _D, _D]
        lib.yko_ttot.restype = C.c_double
        lib.yko_tnorm.argtypes = [P, C.c_int, _D, _D]
        lib.yko_cancel_vcom.argtypes = [P, C.c_int, _D, _D]
        lib.yko_strain.argtypes = [P, C.c_int, _D, _D, _D, C.c_double]
        lib.yko_accel_truth.argtypes = [P, C.c_int, _D, _D, _D, _D]
        lib.yko_accel_omp.argtypes = [P, C.c_int, _D, _D, _D, _D, C.c_int]
        lib.yko_pair_abs_sum.argtypes = [P, C.c_int, _D, _D, _D, _D, C.c_int]
        _lib = lib
    return _lib


def ref_available() -> bool:
    return REF_SO.exists()


def load_ref():
    global _ref
    if _ref is None:
        if not REF_SO.exists():
            raise FileNotFoundError(f"{REF_SO} not built (make -C oracle ref)")
        lib = C.CDLL(str(REF_SO))
        lib.ykref_setup.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _D, _D, _D, _D, _D]
        lib.ykref_setup.restype = C.c_int
        lib.ykref_accel.argtypes = [C.c_int, _D, _D]
        lib.ykref_newton.argtypes = [C.c_int, C.c_double, C.c_int, _D]
        lib.ykref_newton_alt.argtypes = [C.c_int, C.c_int, C.c_double, _D]
        lib.ykref_set_accel.argtypes = [_D]
        lib.ykref_get.argtypes = [_D, _D, _D]
        lib.ykref_ttot.restype = C.c_double
        lib.ykref_strain.argtypes = [_D, _D, C.c_double, _D]
        lib.ykref_md_loop.argtypes = [C.c_int] * 6 + [C.c_double, _D, _D, C.c_int, _D,
                                                        C.POINTER(C.c_int), _D]
        lib.ykref_timers.argtypes = [_D]
        _ref = lib
    return _ref


def params_struct(p, ref_f1=None) -> YkoParams:
    s = YkoParams()
    s.kind = KINDS[p.sim_type]
    s.coulomb_on = 1 if p.coulomb_on else 0
    s.nuclear_on = 1 if p.nuclear_on else 0
    s.ref_f1 = int(p.ref_compat_f1 if ref_f1 is None else ref_f1)
    for k in range(3):
        s.xl[k] = p.xl[k]
    for name in ("xmuc", "rccut", "frp", "vc", "xmass", "rho", "kT", "zi", "ai", "aa", "bb",
                 "cc", "xpacket", "rncut", "bfield"):
        setattr(s, name, float(getattr(p, name)))
    return s


def _arr(a, shape):
    out = np.ascontiguousarray(a, dtype=np.float64)
    assert out.shape == shape, (out.shape, shape)
    return out


def _meas(m):
    return dict(zip(MEAS_KEYS, (float(v) for v in m)))


def accel(p, x, Z, A, do_meas=True, ref_f1=None):
    """accel(do_measurements) of the restatement: (a (n,3), meas dict)."""
    n = x.shape[0]
    x = _arr(x, (n, 3))
    Z = _arr(Z, (n,))
    A = _arr(A, (n,))
    a = np.zeros((n + 3, 3))
    m = np.zeros(NMEAS)
    ps = params_struct(p, ref_f1)
    rc = load().yko_accel(C.byref(ps), n, _ptr(x), _ptr(Z), _ptr(A), int(do_meas), _ptr(a),
                          _ptr(m))
    if rc != 0:
        raise RuntimeError(f"yko_accel failed ({rc})")
    return a[:n].copy(), _meas(m)


def newton(p, dt, x, v, vold, a, Z, A, do_meas=False, ref_f1=None, kdk=False, sw=False):
    """One step on copies of the state; returns (x, v, vold, a, meas).
    Integrator: vel_verlet_2 (default), vel_verlet_1 (kdk) or vel_verlet_3 (sw)."""
    n = x.shape[0]
    x = _arr(x, (n, 3)).copy()
    v = _arr(v, (n, 3)).copy()
    vold = _arr(vold, (n, 3)).copy()
    ab = np.zeros((n + 3, 3))
    ab[:n] = a
    Z = _arr(Z, (n,))
    A = _arr(A, (n,))
    m = np.zeros(NMEAS)
    ps = params_struct(p, ref_f1)
    lib = load()
    if kdk or sw:
        fn = lib.yko_newton_sw if sw else lib.yko_newton_kdk
        rc = fn(C.byref(ps), n, dt, _ptr(x), _ptr(v), _ptr(ab), _ptr(Z), _ptr(A),
                int(do_meas), _ptr(m))
    else:
        rc = lib.yko_newton(C.byref(ps), n, dt, _ptr(x), _ptr(v), _ptr(vold), _ptr(ab), _ptr(Z),
                            _ptr(A), int(do_meas), _ptr(m))
    if rc != 0:
        raise RuntimeError(f"yko_newton failed ({rc})")
    return x, v, vold, ab[:n].copy(), _meas(m)


def ttot(p, v, A):
    n = v.shape[0]
    ps = params_struct(p)
    return load().yko_ttot(C.byref(ps), n, _ptr(_arr(v, (n, 3))), _ptr(_arr(A, (n,))))


XNOSTRAIN = 1000.0   # src/md_globals.f:22


def tnorm(p, v, A):
    """tnorm (src/gen_utils.f:61-75) on a copy of v."""
    n = v.shape[0]
    v = _arr(v, (n, 3)).copy()
    load().yko_tnorm(C.byref(params_struct(p)), n, _ptr(v), _ptr(_arr(A, (n,))))
    return v


def cancel_vcom(p, v, A):
    """cancel_vcom (src/gen_utils.f:82-107, tnorm included) on a copy of v."""
    n = v.shape[0]
    v = _arr(v, (n, 3)).copy()
    load().yko_cancel_vcom(C.byref(params_struct(p)), n, _ptr(v), _ptr(_arr(A, (n,))))
    return v


def strain(p, x, xl0, deps, dt):
    """strain (src/strain.f:19-82): (positions, new box)."""
    n = x.shape[0]
    x = _arr(x, (n, 3)).copy()
    ps = params_struct(p)
    load().yko_strain(C.byref(ps), n, _ptr(x), _ptr(_arr(xl0, (3,))), _ptr(_arr(deps, (3,))),
                      float(dt))
    return x, tuple(ps.xl[k] for k in range(3))


def _mdmod(p, q):
    return -1 if q == 0 else p % q   # src/md.F:248-256


def md_loop(p, dt, x, v, a, Z, A, nsteps, istep0=0, nind=0, nptensor=0, ncom=0, ntnorm=0,
            deps=(XNOSTRAIN,) * 3, xl0=None, ref_f1=None):
    """The measurement-phase step loop of src/md.F:124-133 over the restated
    routines: per step cancel_vcom (every ncom), tnorm (every ntnorm),
    strain, newton measuring when mdmod(istep+1,nind) or
    mdmod(istep+1,nptensor) is 0.  Returns (x, v, a, records, box) with one
    record per measuring step (its 'istep' included)."""
    xl0 = p.xl if xl0 is None else xl0
    x, v, a = _arr(x, x.shape).copy(), _arr(v, v.shape).copy(), _arr(a, a.shape).copy()
    vold = np.zeros_like(v)
    istep = istep0
    recs = []
    for _ in range(nsteps):
        if _mdmod(istep, ncom) == 0:
            v = cancel_vcom(p, v, A)
        if _mdmod(istep, ntnorm) == 0:
            v = tnorm(p, v, A)
        x, box = strain(p, x, xl0, deps, dt)
        p = p.replace(xl=box)
        dm = _mdmod(istep + 1, nind) == 0 or _mdmod(istep + 1, nptensor) == 0
        x, v, vold, a, m = newton(p, dt, x, v, vold, a, Z, A, dm, ref_f1=ref_f1)
        istep += 1
        if dm:
            recs.append(dict(m, istep=istep))
    return x, v, a, recs, p.xl


def truth(p, x, Z):
    """Long-double full-matrix evaluation: (force (n,3) MeV/fm, ev_total, vir, pp[6])."""
    n = x.shape[0]
    x = _arr(x, (n, 3))
    Z = _arr(Z, (n,))
    f = np.zeros((n, 3))
    out = np.zeros(8)
    ps = params_struct(p, 0)
    load().yko_accel_truth(C.byref(ps), n, _ptr(x), _ptr(Z), _ptr(f), _ptr(out))
    return f, float(out[0]), float(out[1]), out[2:].copy()


def accel_omp(p, x, Z, A, nthreads=0):
    """OpenMP restatement of the ions' force-only branch with the reference's
    last-particle behaviour (F1), for CPU timing; returns (a, threads)."""
    if p.sim_type not in ("pure-ion", "ion-mixture"):
        raise ValueError(f"accel_omp covers the ion kernels only, not {p.sim_type!r}")
    n = x.shape[0]
    a = np.zeros((n, 3))
    ps = params_struct(p, 0)
    used = load().yko_accel_omp(C.byref(ps), n, _ptr(_arr(x, (n, 3))), _ptr(_arr(Z, (n,))),
                                _ptr(_arr(A, (n,))), _ptr(a), int(nthreads))
    return a, used


def pair_abs_sum(p, x, Z, nthreads=0):
    """Conditioning of each particle's force (test metric): (sum_j |F_ij|,
    sqrt(sum_j |F_ij|^2)) in MeV/fm, over the pairs inside the cutoffs."""
    n = x.shape[0]
    sa, rms = np.zeros(n), np.zeros(n)
    ps = params_struct(p, 0)
    load().yko_pair_abs_sum(C.byref(ps), n, _ptr(_arr(x, (n, 3))), _ptr(_arr(Z, (n,))),
                            _ptr(sa), _ptr(rms), int(nthreads))
    return sa, rms


# ---------------------------------------------------------------------------
# The reference's own Fortran (oracle/_ref/libykref.so)
def _ref_prm(p):
    return np.array([*p.xl, p.xmuc, p.rccut, p.frp, p.xmass, p.rho, p.kT, p.zi, p.ai,
                     p.aa, p.bb, p.cc, p.xpacket, p.rncut, p.bfield], dtype=np.float64)


def ref_setup(p, x, v, Z, A):
    n = x.shape[0]
    prm = _ref_prm(p)
    x = _arr(x, (n, 3))
    v = _arr(v, (n, 3)) if v is not None else np.zeros((n, 3))
    rc = load_ref().ykref_setup(KINDS[p.sim_type], n, int(p.coulomb_on), int(p.nuclear_on),
                                _ptr(prm), _ptr(x), _ptr(v), _ptr(_arr(Z, (n,))),
                                _ptr(_arr(A, (n,))))
    if rc != 0:
        raise RuntimeError("ykref_setup failed")


def ref_accel(p, x, Z, A, do_meas=True):
    """The reference's accel(do_measurements) on these positions."""
    n = x.shape[0]
    ref_setup(p, x, None, Z, A)
    a = np.zeros((n + 3, 3))
    m = np.zeros(NMEAS)
    load_ref().ykref_accel(int(do_meas), _ptr(a), _ptr(m))
    return a[:n].copy(), _meas(m)


def ref_newton(p, dt, x, v, a, Z, A, nsteps=1, meas_last=True):
    """nsteps of the reference's newton (vel_verlet_2) from (x, v, a)."""
    n = x.shape[0]
    ref_setup(p, x, v, Z, A)
    lib = load_ref()
    a = _arr(a, (n, 3))
    lib.ykref_set_accel(_ptr(a))
    m = np.zeros(NMEAS)
    lib.ykref_newton(int(nsteps), float(dt), int(meas_last), _ptr(m))
    xo, vo, ao = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    lib.ykref_get(_ptr(xo), _ptr(vo), _ptr(ao))
    return xo, vo, ao, _meas(m)


def ref_newton_alt(p, dt, x, v, a, Z, A, which, nsteps=1):
    """nsteps of the reference's vel_verlet_1 (which='kdk') or vel_verlet_3
    (which='sw') newton from (x, v, a).  Both call accel without an argument
    in the reference; the harness gives them accel(.true.) (measuring)."""
    n = x.shape[0]
    ref_setup(p, x, v, Z, A)
    lib = load_ref()
    lib.ykref_set_accel(_ptr(_arr(a, (n, 3))))
    m = np.zeros(NMEAS)
    lib.ykref_newton_alt({"kdk": 1, "sw": 3}[which], int(nsteps), float(dt), _ptr(m))
    xo, vo, ao = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    lib.ykref_get(_ptr(xo), _ptr(vo), _ptr(ao))
    return xo, vo, ao, _meas(m)


def ref_tnorm(p, v, A):
    n = v.shape[0]
    ref_setup(p, np.zeros((n, 3)), v, np.ones(n), A)
    load_ref().ykref_tnorm()
    xo, vo, ao = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    load_ref().ykref_get(_ptr(xo), _ptr(vo), _ptr(ao))
    return vo


def ref_cancel_vcom(p, v, A):
    n = v.shape[0]
    ref_setup(p, np.zeros((n, 3)), v, np.ones(n), A)
    load_ref().ykref_cancel_vcom()
    xo, vo, ao = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    load_ref().ykref_get(_ptr(xo), _ptr(vo), _ptr(ao))
    return vo


def ref_strain(p, x, xl0, deps, dt):
    n = x.shape[0]
    ref_setup(p, x, None, np.ones(n), np.ones(n))
    xl = np.zeros(3)
    load_ref().ykref_strain(_ptr(_arr(deps, (3,))), _ptr(_arr(xl0, (3,))), float(dt), _ptr(xl))
    xo, vo, ao = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    load_ref().ykref_get(_ptr(xo), _ptr(vo), _ptr(ao))
    return xo, tuple(xl)


def ref_md_loop(p, dt, x, v, a, Z, A, nsteps, istep0=0, nind=0, nptensor=0, ncom=0, ntnorm=0,
                deps=(XNOSTRAIN,) * 3, xl0=None):
    """The same loop on the reference's own routines (ref_harness.f90)."""
    n = x.shape[0]
    xl0 = p.xl if xl0 is None else xl0
    ref_setup(p, x, v, Z, A)
    lib = load_ref()
    lib.ykref_set_accel(_ptr(_arr(a, (n, 3))))
    maxrec = max(nsteps, 1)
    rec = np.zeros((maxrec, 10))
    nrec = C.c_int(0)
    xl = np.zeros(3)
    lib.ykref_md_loop(int(nsteps), int(istep0), int(nind), int(nptensor), int(ncom), int(ntnorm),
                      float(dt), _ptr(_arr(deps, (3,))), _ptr(_arr(xl0, (3,))), maxrec,
                      _ptr(rec), C.byref(nrec), _ptr(xl))
    xo, vo, ao = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
    lib.ykref_get(_ptr(xo), _ptr(vo), _ptr(ao))
    keys = ("ev", "ek", "px", "pp11", "pp12", "pp13", "pp22", "pp23", "pp33")
    recs = [dict(zip(keys, r[1:]), istep=int(r[0])) for r in rec[:nrec.value]]
    return xo, vo, ao, recs, tuple(xl)


def ref_timers():
    t = np.zeros(6)
    load_ref().ykref_timers(_ptr(t))
    return {"t_calc_a": t[0], "n_calc_a": int(t[1]), "t_accel": t[2], "n_accel": int(t[3]),
            "t_newton": t[4], "n_newton": int(t[5])}


def ref_reset_timers():
    load_ref().ykref_reset_timers()
