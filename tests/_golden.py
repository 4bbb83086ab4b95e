"""Loads the golden vectors written by tests/golden/make_golden.py."""
import json
from pathlib import Path

import numpy as np

from yukawamd_amd.params import SimParams

GOLDEN = Path(__file__).resolve().parent / "golden"
NAMES = sorted(p.stem for p in GOLDEN.glob("*.npz"))


def load(name):
    meta = json.loads((GOLDEN / f"{name}.json").read_text())
    d = dict(np.load(GOLDEN / f"{name}.npz", allow_pickle=False))
    kw = {k: meta[k] for k in meta if k not in ("n", "source")}
    kw["xl"] = tuple(kw["xl"])
    return SimParams(**kw), d
