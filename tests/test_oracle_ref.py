"""The C restatement against the reference's own Fortran, bit for bit.

oracle/_ref/libykref.so is the reference compiled from its sources
(src/int_ion_pure.f, src/int_ion_mix.f, src/int_nn.f, src/vel_verlet_2.f, ...)
with oracle/ref_harness.f90; with one OpenMP thread both must agree exactly on
accelerations, energies, pressure and pressure tensor, for both branches of
every kernel and for whole velocity-Verlet steps.
"""
import numpy as np
import pytest

import oracle
from yukawamd_amd import workloads as W

pytestmark = pytest.mark.skipif(not oracle.ref_available(),
                                reason="reference build oracle/_ref not present")

CASES = [
    ("pure", lambda: W.pure_ion(n=250, seed=11)),
    ("mix", lambda: W.ion_mixture(n=300, seed=12)),
    ("nn", lambda: W.nucleon(n=700, seed=13)),
    ("mix-kappa2", lambda: W.ion_mixture(n=200, seed=14, kappa=2.0)),
]


@pytest.mark.parametrize("name,make", CASES)
@pytest.mark.parametrize("do_meas", [True, False])
def test_accel_bitexact(name, make, do_meas):
    w = make()
    a1, m1 = oracle.accel(w.params, w.x, w.Z, w.A, do_meas, ref_f1=1)
    a2, m2 = oracle.ref_accel(w.params, w.x, w.Z, w.A, do_meas)
    assert np.array_equal(a1, a2)
    for k in ("ev", "px", "pp11", "pp12", "pp13", "pp22", "pp23", "pp33"):
        assert m1[k] == m2[k], k


@pytest.mark.parametrize("name,make", CASES)
def test_newton_steps_bitexact(name, make):
    w = make()
    p = w.params
    a, _ = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=1)
    x, v, vold = w.x, w.v, np.zeros_like(w.v)
    for step in range(3):
        x, v, vold, a, m = oracle.newton(p, w.dt, x, v, vold, a, w.Z, w.A, step == 2, ref_f1=1)
    a0, _ = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=1)
    xr, vr, ar, mr = oracle.ref_newton(p, w.dt, w.x, w.v, a0, w.Z, w.A, nsteps=3, meas_last=True)
    assert np.array_equal(x, xr) and np.array_equal(v, vr) and np.array_equal(a, ar)
    for k in ("ev", "ek", "px", "pp11", "pp12", "pp13", "pp22", "pp23", "pp33"):
        assert m[k] == mr[k], k


def test_bfield_step_bitexact():
    w = W.ion_mixture(n=120, seed=15)
    p = w.params.replace(bfield=3.0e15)
    a, _ = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=1)
    x, v, _, a1, m = oracle.newton(p, w.dt, w.x, w.v, np.zeros_like(w.v), a, w.Z, w.A, True,
                                   ref_f1=1)
    xr, vr, ar, mr = oracle.ref_newton(p, w.dt, w.x, w.v, a, w.Z, w.A, 1, True)
    assert np.array_equal(v, vr) and np.array_equal(a1, ar) and m["ek"] == mr["ek"]


def test_ttot_bitexact():
    w = W.ion_mixture(n=300, seed=16)
    oracle.ref_setup(w.params, w.x, w.v, w.Z, w.A)
    assert oracle.ttot(w.params, w.v, w.A) == oracle.load_ref().ykref_ttot()


def test_reference_f1_last_particle():
    """The reference's mixture kernel leaves out zii(n-1) for the last particle
    (src/int_ion_mix.f:158, :201, :250): its acceleration is a_true / Z(n-1)."""
    w = W.ion_mixture(n=150, seed=17)
    ar, _ = oracle.ref_accel(w.params, w.x, w.Z, w.A, False)
    af, _ = oracle.accel(w.params, w.x, w.Z, w.A, False, ref_f1=0)
    assert np.array_equal(ar[:-1], af[:-1])
    np.testing.assert_allclose(ar[-1] * w.Z[-1], af[-1], rtol=1e-15, atol=0)
    assert not np.allclose(ar[-1], af[-1])


@pytest.mark.parametrize("which", ["kdk", "sw"])
@pytest.mark.parametrize("name,make", CASES[:2])
def test_alternate_integrators_bitexact(which, name, make):
    """vel_verlet_1 (kick-drift-kick) and vel_verlet_3 (Spreiter-Walter, with a
    field strong enough that the rotation terms matter) against the
    reference's own sources, three measured steps."""
    w = make()
    p = w.params.replace(bfield=2.0e17) if which == "sw" else w.params
    a0, _ = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=1)
    x, v, vold, a = w.x, w.v, np.zeros_like(w.v), a0
    for _ in range(3):
        x, v, vold, a, m = oracle.newton(p, w.dt, x, v, vold, a, w.Z, w.A, True, ref_f1=1,
                                         kdk=which == "kdk", sw=which == "sw")
    xr, vr, ar, mr = oracle.ref_newton_alt(p, w.dt, w.x, w.v, a0, w.Z, w.A, which, nsteps=3)
    assert np.array_equal(x, xr) and np.array_equal(v, vr) and np.array_equal(a, ar)
    for k in ("ev", "ek", "px", "pp11", "pp12", "pp13", "pp22", "pp23", "pp33"):
        assert m[k] == mr[k], k


# --- velocity utilities, strain and the md.F step loop -----------------------
X = oracle.XNOSTRAIN


@pytest.mark.parametrize("name,make", CASES[:2])
def test_tnorm_cancel_vcom_bitexact(name, make):
    """tnorm and cancel_vcom (src/gen_utils.f:61-107) on velocities with a
    drift and the wrong temperature."""
    w = make()
    v = w.v * 1.7 + np.array([0.002, -0.001, 0.0005])
    assert np.array_equal(oracle.tnorm(w.params, v, w.A), oracle.ref_tnorm(w.params, v, w.A))
    assert np.array_equal(oracle.cancel_vcom(w.params, v, w.A),
                          oracle.ref_cancel_vcom(w.params, v, w.A))


STRAINS = [(X, X, X), (2e-6, X, X), (X, -3e-6, X), (2e-6, -1e-6, X), (X, X, 3e-6),
           (-2e-6, X, 1e-6), (X, 1e-6, -2e-6), (1e-6, 2e-6, -3e-6)]


@pytest.mark.parametrize("deps", STRAINS)
def test_strain_bitexact(deps):
    """Every case of src/strain.f:31-72 (one, two or three strained axes, or
    none), from a box that differs from the gage, with particles at the upper
    faces so that the shrinking edges wrap them."""
    w = W.ion_mixture(n=300, seed=16)
    L = np.array(w.params.xl)
    xl0 = tuple(0.97 * L)
    x = w.x.copy()
    x[:6] = L * 0.99999
    dt = w.dt * 200
    x1, b1 = oracle.strain(w.params, x, xl0, deps, dt)
    x2, b2 = oracle.ref_strain(w.params, x, xl0, deps, dt)
    assert b1 == b2
    assert np.array_equal(x1, x2)
    if deps != (X, X, X):
        assert b1 != tuple(L) and not np.array_equal(x1, x)


def test_md_loop_bitexact():
    """The measurement-phase loop of src/md.F:124-133 -- cancel_vcom every 5,
    tnorm every 3, an x strain each step, newton measuring at the nind=4 and
    nptensor=6 steps -- restated against the reference's own routines."""
    w = W.ion_mixture(n=200, seed=17)
    p = w.params
    a, _ = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=1)
    kw = dict(nsteps=13, istep0=0, nind=4, nptensor=6, ncom=5, ntnorm=3, deps=(2e-6, X, X),
              xl0=p.xl)
    x1, v1, a1, r1, b1 = oracle.md_loop(p, w.dt, w.x, w.v, a, w.Z, w.A, ref_f1=1, **kw)
    x2, v2, a2, r2, b2 = oracle.ref_md_loop(p, w.dt, w.x, w.v, a, w.Z, w.A, **kw)
    assert b1 == b2 and b1 != p.xl
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2) and np.array_equal(a1, a2)
    assert [r["istep"] for r in r1] == [r["istep"] for r in r2] == [4, 6, 8, 12]
    for q1, q2 in zip(r1, r2):
        for k in q2:
            assert q1[k] == q2[k], (q2["istep"], k)
