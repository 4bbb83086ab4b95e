"""ctypes binding of libyukawa_hip.so (the C ABI in include/yk.h).

The library is loaded from the package directory only; there is no other
implementation to fall back to, so a missing or unloadable library raises
``LibraryMissing`` at the first call.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libyukawa_hip.so"


class LibraryMissing(RuntimeError):
    pass


class YkParams(C.Structure):
    _fields_ = [
        ("sim_type", C.c_int32), ("precision", C.c_int32), ("flags", C.c_uint32),
        ("coulomb_on", C.c_int32), ("nuclear_on", C.c_int32), ("resort_every", C.c_int32),
        ("xl", C.c_double * 3),
        ("xmuc", C.c_double), ("rccut", C.c_double), ("frp", C.c_double), ("vc", C.c_double),
        ("xmass", C.c_double), ("rho", C.c_double), ("kT", C.c_double),
        ("zi", C.c_double), ("ai", C.c_double),
        ("aa", C.c_double), ("bb", C.c_double), ("cc", C.c_double),
        ("xpacket", C.c_double), ("rncut", C.c_double), ("bfield", C.c_double),
    ]


class YkMeas(C.Structure):
    _fields_ = [("ev", C.c_double), ("ek", C.c_double), ("px", C.c_double),
                ("vir", C.c_double), ("pp", C.c_double * 6)]

    def as_dict(self) -> dict:
        return {"ev": self.ev, "ek": self.ek, "px": self.px, "vir": self.vir,
                "pp": list(self.pp)}


XNOSTRAIN = 1000.0   # YK_XNOSTRAIN: no strain rate on an axis (src/md_globals.f:22)


class YkLoop(C.Structure):
    _fields_ = [("dt", C.c_double), ("nsteps", C.c_int32), ("istep0", C.c_int32),
                ("nind", C.c_int32), ("nptensor", C.c_int32), ("ncom", C.c_int32),
                ("ntnorm", C.c_int32), ("deps", C.c_double * 3), ("xl0", C.c_double * 3)]


class YkTiming(C.Structure):
    _fields_ = [("force_ms", C.c_double), ("force_launches", C.c_int64),
                ("step_ms", C.c_double), ("steps", C.c_int64), ("pairs_swept", C.c_int64)]


STATUS = {0: "YK_OK", 1: "YK_ERR_ARG", 2: "YK_ERR_HIP", 3: "YK_ERR_NOMEM", 4: "YK_ERR_STATE",
          5: "YK_ERR_NONFINITE", 6: "YK_ERR_UNSUPPORTED", 7: "YK_ERR_CAPACITY"}


class YkError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


_lib = None

P = C.c_void_p
D = C.POINTER(C.c_double)

# name -> (restype, argtypes); every function declared in include/yk.h
SIGNATURES = {
    "yk_abi_version": (C.c_int32, []),
    "yk_last_error": (C.c_char_p, []),
    "yk_create": (C.c_int, [C.POINTER(YkParams), C.c_int32, D, D, C.POINTER(P)]),
    "yk_destroy": (None, [P]),
    "yk_set_params": (C.c_int, [P, C.POINTER(YkParams)]),
    "yk_put_state": (C.c_int, [P, D, D]),
    "yk_put_accel": (C.c_int, [P, D]),
    "yk_get_state": (C.c_int, [P, D, D, D]),
    "yk_accel": (C.c_int, [P, C.c_int32, C.POINTER(YkMeas)]),
    "yk_newton": (C.c_int, [P, C.c_double, C.c_int32, C.POINTER(YkMeas)]),
    "yk_run": (C.c_int, [P, C.c_double, C.c_int32, C.c_int32, C.POINTER(YkMeas)]),
    "yk_ttot": (C.c_int, [P, D]),
    "yk_tnorm": (C.c_int, [P]),
    "yk_cancel_vcom": (C.c_int, [P]),
    "yk_strain": (C.c_int, [P, C.c_double, D, D, D]),
    "yk_run_md": (C.c_int, [P, C.POINTER(YkLoop), C.POINTER(YkMeas), C.POINTER(C.c_int32),
                            C.c_int32, C.POINTER(C.c_int32), D]),
    "yk_get_timing": (C.c_int, [P, C.POINTER(YkTiming)]),
    "yk_reset_timing": (C.c_int, [P]),
    "yk_accel_noalloc": (C.c_int, [C.POINTER(YkParams), C.c_int32, P, P, P, P, C.c_int32, P,
                                   P]),
    # batched ensembles / continuous batching
    "yk_batch_last_error": (C.c_char_p, []),
    "yk_batch_create": (C.c_int, [C.c_int64, C.c_int32, C.c_int32, C.POINTER(P)]),
    "yk_batch_destroy": (None, [P]),
    "yk_batch_footprint": (C.c_int, [P, C.c_int32, C.POINTER(C.c_int64),
                                     C.POINTER(C.c_int64)]),
    "yk_batch_add": (C.c_int, [P, C.POINTER(YkParams), C.c_int32, D, D, D, D, C.c_double,
                               C.c_int32, C.POINTER(C.c_int32)]),
    "yk_batch_remove": (C.c_int, [P, C.c_int32]),
    "yk_batch_set_tuning": (C.c_int, [P, C.c_int32, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32)]),
    "yk_batch_set_resort": (C.c_int, [P, C.c_int32]),
    "yk_batch_run": (C.c_int, [P, C.c_int32, C.c_int32]),
    "yk_batch_status": (C.c_int, [P, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "yk_batch_get": (C.c_int, [P, C.c_int32, D, D, D, C.POINTER(YkMeas), C.POINTER(YkMeas),
                               C.c_int32, C.POINTER(C.c_int32)]),
    "yk_batch_last_run": (C.c_int, [P, C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                    C.POINTER(C.c_int64)]),
}


def load(path: os.PathLike | None = None):
    """Load (once) and return the ctypes library."""
    global _lib
    if _lib is not None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise LibraryMissing(f"{p} is not built: run `python -m yukawamd_amd.build` "
                             "(the HIP library is the only implementation)")
    try:
        lib = C.CDLL(str(p))
    except OSError as e:
        raise LibraryMissing(f"cannot load {p}: {e}") from e
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(code: int):
    if code != 0:
        msg = load().yk_last_error()
        raise YkError(code, msg.decode() if msg else "")


def check_batch(code: int):
    if code != 0:
        msg = load().yk_batch_last_error()
        raise YkError(code, msg.decode() if msg else "")


def params_struct(p) -> YkParams:
    from .params import PRECISIONS, SIM_TYPES
    s = YkParams()
    s.sim_type = SIM_TYPES[p.sim_type]
    s.precision = PRECISIONS[p.precision]
    s.flags = p.flags
    s.coulomb_on = 1 if p.coulomb_on else 0
    s.nuclear_on = 1 if p.nuclear_on else 0
    s.resort_every = int(p.resort_every)
    for k in range(3):
        s.xl[k] = p.xl[k]
    for name in ("xmuc", "rccut", "frp", "vc", "xmass", "rho", "kT", "zi", "ai", "aa", "bb",
                 "cc", "xpacket", "rncut", "bfield"):
        setattr(s, name, float(getattr(p, name)))
    return s
