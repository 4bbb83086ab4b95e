"""Run parameters of the hot path, mirroring the reference's module globals.

The reference keeps every parameter in ``module md_globals``
(src/md_globals.f:36-140); accel/newton read them directly.  ``SimParams``
carries the subset the pair kernels and the integrator use, with the same
names, units and defaults, and converts to the C-ABI struct ``yk_params``
(include/yk.h).
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field

# src/md_globals.f:24-28, :122
HBARC = 197.327
ALPHA = 1.0 / 137.036
VC = ALPHA * HBARC          # Coulomb coupling constant, MeV fm (= 1.439964680814...)
PI = 3.14159265358979324
XMASS_DEFAULT = 931.0       # src/md_globals.f:64
FM_PER_C_SECONDS = 3.33564e-24   # one fm/c in seconds

SIM_TYPES = {"pure-ion": 0, "ion-mixture": 1, "nucleon": 2, "B-field": 3}
PRECISIONS = {"fp64": 0, "mixed": 1}

FLAG_REF_F1 = 1 << 0
FLAG_KDK = 1 << 1
FLAG_TIMING = 1 << 2
FLAG_ALLPAIRS = 1 << 3
FLAG_SW = 1 << 4
FLAG_TILE_LISTS = 1 << 5


@dataclass
class SimParams:
    """Parameters of one system (names follow src/md_globals.f)."""

    sim_type: str                       # 'pure-ion' | 'ion-mixture' | 'nucleon' | 'B-field'
    xl: tuple = (0.0, 0.0, 0.0)         # box edges (fm)
    xmuc: float = 0.0                   # inverse screening length (1/fm)
    rccut: float = 1.0e50               # Coulomb cutoff (fm)
    frp: float = 1.0                    # ion form factor
    vc: float = VC
    xmass: float = XMASS_DEFAULT        # MeV
    rho: float = 0.0                    # 1/fm^3
    kT: float = 0.0                     # MeV
    zi: float = 0.0                     # pure-ion Z
    ai: float = 0.0                     # pure-ion A
    aa: float = 110.0                   # HPP (MeV), src/md_globals.f:127-131
    bb: float = -26.0
    cc: float = 24.0
    xpacket: float = 1.25               # fm^2
    rncut: float = 0.0                  # fm
    coulomb: str = "screened-coulomb"   # src/md_globals.f:118
    nuclear: str = "HPP"                # src/md_globals.f:126
    bfield: float = 0.0                 # Gauss
    precision: str = "fp64"             # 'fp64' | 'mixed'
    ref_compat_f1: bool = False         # keep the mixture kernel's last-particle defect
    integrator: str = "vv2"             # 'vv2' (vel_verlet_2.f) | 'kdk' (vel_verlet_1.f)
                                        # | 'sw' (vel_verlet_3.f, Spreiter-Walter B field)
    timing: bool = False
    resort_every: int = 0               # device re-sort cadence (0 = library default)
    allpairs: bool = False              # AABB-culled tile pairs instead of pair lists
    tile_lists: bool = False            # 64-particle tile lists only (no cluster lists)

    def __post_init__(self):
        if self.sim_type not in SIM_TYPES:
            raise ValueError(f"sim_type={self.sim_type!r}: program not compiled to do this "
                             "simulation type")
        if self.precision not in PRECISIONS:
            raise ValueError(f"precision={self.precision!r} is not 'fp64' or 'mixed'")
        if self.integrator not in ("vv2", "kdk", "sw"):
            raise ValueError(f"integrator={self.integrator!r} is not 'vv2', 'kdk' or 'sw'")
        self.xl = tuple(float(v) for v in self.xl)

    @property
    def coulomb_on(self) -> bool:
        return self.coulomb == "screened-coulomb"

    @property
    def nuclear_on(self) -> bool:
        return self.nuclear == "HPP"

    @property
    def flags(self) -> int:
        f = 0
        if self.ref_compat_f1:
            f |= FLAG_REF_F1
        if self.integrator == "kdk":
            f |= FLAG_KDK
        if self.integrator == "sw":
            f |= FLAG_SW
        if self.timing:
            f |= FLAG_TIMING
        if self.allpairs:
            f |= FLAG_ALLPAIRS
        if self.tile_lists:
            f |= FLAG_TILE_LISTS
        return f

    def replace(self, **kw) -> "SimParams":
        return dataclasses.replace(self, **kw)

    @property
    def volume(self) -> float:
        return self.xl[0] * self.xl[1] * self.xl[2]


def auto_rncut(aa: float = 110.0, bb: float = -26.0, cc: float = 24.0,
               xpacket: float = 1.25) -> float:
    """Nuclear cutoff chosen when rncut=0 (src/input1.F:296-311)."""
    xepsilon = (math.sqrt(2.0 * xpacket) / (2.0 * (-bb + cc))) * 0.0001 * (2.0 ** (-53))
    rncut = math.sqrt(-math.log((-bb - cc) / (2.0 * aa)))
    xdeltar = 1.0
    rncut0 = rncut
    while True:
        while rncut * math.exp(-rncut ** 2) > xepsilon:
            rncut0 = rncut
            rncut = rncut + xdeltar
        if xdeltar < 0.11:
            break
        xdeltar = xdeltar / 10.0
        rncut = rncut0
    return rncut * math.sqrt(2.0 * xpacket)
