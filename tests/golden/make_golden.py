"""Writes the golden vectors in tests/golden/ from the reference's own Fortran.

Runs oracle/_ref/libykref.so (the reference compiled from its sources by
oracle/Makefile) in THIS container, one OpenMP thread, on small seeded
systems of each sim_type, and stores inputs and outputs as .npz (no pickles):
  accel with do_measurements true and false (accelerations, ev, px, pp),
  three vel_verlet_2 steps (positions, velocities, accelerations, ev, ek, px,
  pp of the last step).

    OMP_NUM_THREADS=1 python tests/golden/make_golden.py
"""
import json
import os
import sys
from pathlib import Path

os.environ.setdefault("OMP_NUM_THREADS", "1")
HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import numpy as np  # noqa: E402

import oracle  # noqa: E402
from yukawamd_amd import workloads as W  # noqa: E402

CASES = {
    "pure_ion_96": lambda: W.pure_ion(n=96, seed=101),
    "ion_mixture_160": lambda: W.ion_mixture(n=160, seed=102),
    "nucleon_256": lambda: W.nucleon(n=256, seed=103, rho=0.05),
    "ion_mixture_bcc_128": lambda: W.ion_mixture(n=128, seed=104, start="bcc", kappa=1.3),
}

PARAM_KEYS = ("sim_type", "xl", "xmuc", "rccut", "frp", "vc", "xmass", "rho", "kT", "zi",
              "ai", "aa", "bb", "cc", "xpacket", "rncut", "coulomb", "nuclear", "bfield")


def main():
    for name, make in CASES.items():
        w = make()
        p = w.params
        out = {"x": w.x, "v": w.v, "Z": w.Z, "A": w.A, "dt": np.array(w.dt)}
        for dm in (1, 0):
            a, m = oracle.ref_accel(p, w.x, w.Z, w.A, bool(dm))
            out[f"a_meas{dm}"] = a
            out[f"m_meas{dm}"] = np.array([m[k] for k in oracle.MEAS_KEYS])
        a0, _ = oracle.ref_accel(p, w.x, w.Z, w.A, True)
        x3, v3, a3, m3 = oracle.ref_newton(p, w.dt, w.x, w.v, a0, w.Z, w.A, nsteps=3,
                                           meas_last=True)
        out.update(x3=x3, v3=v3, a3=a3, m3=np.array([m3[k] for k in oracle.MEAS_KEYS]))
        np.savez(HERE / f"{name}.npz", **out)
        meta = {k: getattr(p, k) for k in PARAM_KEYS}
        meta["xl"] = list(p.xl)
        meta["n"] = w.n
        meta["source"] = "reference Fortran (oracle/_ref), OMP_NUM_THREADS=1, ref F1 behaviour"
        (HERE / f"{name}.json").write_text(json.dumps(meta, indent=1) + "\n")
        print("wrote", name, w.n)


if __name__ == "__main__":
    main()
