"""The oracle against the reference's golden vectors (tests/golden/*.npz).

The vectors were produced by the reference's own Fortran (one thread); the C
restatement must reproduce them bit for bit, with the reference's mixture
last-particle behaviour switched on (ref_f1=1).
"""
import numpy as np
import pytest

import oracle
from _golden import NAMES, load

KEYS = ("ev", "px", "pp11", "pp12", "pp13", "pp22", "pp23", "pp33")


@pytest.mark.parametrize("name", NAMES)
def test_golden_accel(name):
    p, g = load(name)
    for dm in (1, 0):
        a, m = oracle.accel(p, g["x"], g["Z"], g["A"], bool(dm), ref_f1=1)
        assert np.array_equal(a, g[f"a_meas{dm}"])
        ref = dict(zip(oracle.MEAS_KEYS, g[f"m_meas{dm}"]))
        for k in KEYS:
            assert m[k] == ref[k], k


@pytest.mark.parametrize("name", NAMES)
def test_golden_three_steps(name):
    p, g = load(name)
    a, _ = oracle.accel(p, g["x"], g["Z"], g["A"], True, ref_f1=1)
    x, v, vold = g["x"], g["v"], np.zeros_like(g["v"])
    for s in range(3):
        x, v, vold, a, m = oracle.newton(p, float(g["dt"]), x, v, vold, a, g["Z"], g["A"],
                                         s == 2, ref_f1=1)
    assert np.array_equal(x, g["x3"]) and np.array_equal(v, g["v3"])
    assert np.array_equal(a, g["a3"])
    ref = dict(zip(oracle.MEAS_KEYS, g["m3"]))
    for k in KEYS + ("ek",):
        assert m[k] == ref[k], k
