"""Build recipe for the native parts of yukawamd_amd (in-tree, gfx950).

    python -m yukawamd_amd.build          # library only
    python -m yukawamd_amd.build --all    # library, Fortran binding, oracle

The HIP library is compiled with hipcc for gfx950 only.  Packed FP32 VALU ops
are disabled (-packed-fp32-ops) and FP contraction is off: the kernels write
every fused multiply-add they want explicitly, and the integrator keeps the
reference's separate multiply and add roundings.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "libyukawa_hip.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_library(verbose: bool = True) -> Path:
    srcs = [CSRC / "yk_lib.hip", CSRC / "yk_batch.hip"]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, *HIP_FLAGS, "-o", tmp, *srcs])
    os.replace(tmp, LIB)
    return LIB


FLANG = os.environ.get("FLANG", "/opt/rocm/lib/llvm/bin/flang")
FORTRAN = PKG / "fortran"


YKMD = FORTRAN / "ykmd"


def build_fortran() -> Path:
    """Compile the iso_c_binding module (yk_bind.mod / yk_bind.o) in-tree, and
    the run driver yukawamd_amd/fortran/ykmd (yk_host.f90, yk_config.f90 and
    ykmd.f90, linked
    against libyukawa_hip.so; unformatted files big-endian like the
    reference's build)."""
    out = FORTRAN / "_build"
    out.mkdir(exist_ok=True)
    flags = ["-O2", "-fPIC", "-fconvert=big-endian", "-module-dir", out]
    _run([FLANG, *flags, "-c", FORTRAN / "yk_bind.f90", "-o", out / "yk_bind.o"])
    _run([FLANG, *flags, "-c", FORTRAN / "yk_host.f90", "-o", out / "yk_host.o"])
    _run([FLANG, *flags, "-c", FORTRAN / "yk_config.f90", "-o", out / "yk_config.o"])
    _run([FLANG, *flags, "-c", FORTRAN / "ykmd.f90", "-o", out / "ykmd.o"])
    tmp = YKMD.with_suffix(".tmp")
    _run([FLANG, "-o", tmp, out / "ykmd.o", out / "yk_config.o", out / "yk_host.o",
          out / "yk_bind.o",
          "-L", PKG, "-lyukawa_hip", "-Wl,-rpath,$ORIGIN/.."])
    os.replace(tmp, YKMD)
    return out / "yk_bind.o"


def build_oracle() -> None:
    """The checker: C restatement always; the reference build and the drop-in
    executables where the reference sources are present."""
    _run(["make", "-s", "-C", ROOT / "oracle"])
    if Path("/root/reference/src").is_dir():
        _run(["make", "-s", "-C", ROOT / "oracle", "ref", "dropin"])


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    build_library()
    if "--all" in argv:
        build_fortran()
        build_oracle()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
