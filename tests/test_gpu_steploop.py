"""Device-resident step-loop utilities against the oracle (needs a GPU).

tnorm / cancel_vcom (src/gen_utils.f:61-107), strain (src/strain.f:19-82) and
the measurement-phase loop of src/md.F:124-133 (yk_run_md), all through the
C ABI.  The oracle's restatements are bit-exact against the reference built
from its sources (tests/test_oracle_ref.py); the device sums in another
order, so velocities and energies agree to rounding (bars in each test),
while box edges, which follow from host arithmetic, agree exactly.
"""
import numpy as np
import pytest

import oracle
from yukawamd_amd import workloads as W
from yukawamd_amd.system import YukawaSystem

pytestmark = pytest.mark.gpu
X = oracle.XNOSTRAIN


def test_tnorm_and_cancel_vcom_match_oracle():
    w = W.ion_mixture(n=3000, seed=41)
    v0 = w.v * 1.7 + np.array([0.002, -0.001, 0.0005])
    for op in ("tnorm", "cancel_vcom"):
        with YukawaSystem(w.params, Z=w.Z, A=w.A, dt=w.dt) as s:
            s.put_state(w.x, v0)
            getattr(s, op)()
            v = s.get_state()[1]
            eka = s.ttot()
        ref = getattr(oracle, op)(w.params, v0, w.A)
        np.testing.assert_allclose(v, ref, rtol=1e-13, atol=1e-15 * np.abs(ref).max(),
                                   err_msg=op)
        assert eka == pytest.approx(1.5 * w.params.kT, rel=1e-13)
        if op == "cancel_vcom":
            p = (w.A[:, None] * v).sum(axis=0)
            assert np.abs(p).max() <= 1e-13 * (w.A[:, None] * np.abs(v)).sum()


@pytest.mark.parametrize("deps", [(X, X, X), (2e-6, X, X), (X, -3e-6, X), (2e-6, -1e-6, X),
                                  (X, X, 3e-6), (-2e-6, X, 1e-6), (X, 1e-6, -2e-6),
                                  (1e-6, 2e-6, -3e-6)])
def test_strain_matches_oracle(deps):
    """Every case of src/strain.f: the box exactly, the wrapped positions
    exactly, the accelerations untouched."""
    w = W.ion_mixture(n=2000, seed=42)
    L = np.array(w.params.xl)
    xl0 = tuple(0.97 * L)
    x = w.x.copy()
    x[:6] = L * 0.99999
    dt = w.dt * 200
    with YukawaSystem(w.params, Z=w.Z, A=w.A, dt=w.dt) as s:
        s.put_state(x, w.v)
        s.accel(False)
        a0 = s.a
        box = s.strain(deps, xl0, dt)
        xs, _, a1 = s.get_state()
    xr, boxr = oracle.strain(w.params, x, xl0, deps, dt)
    assert box == boxr
    assert np.array_equal(xs, xr)
    assert np.array_equal(a1, a0)


@pytest.mark.parametrize("deps", [(2e-6, X, X), (X, X, X)], ids=["strain", "no_strain"])
def test_run_md_matches_oracle_loop(deps):
    """13 steps with cancel_vcom every 5, tnorm every 3, with and without an x
    strain each step, measuring at the nind=4 and nptensor=6 steps
    (src/md.F:124-133).  Without strain the steps between the utilities run
    with the next step's drift done by the finalize kernel."""
    w = W.ion_mixture(n=1500, seed=43)
    p = w.params
    kw = dict(nsteps=13, istep0=0, nind=4, nptensor=6, ncom=5, ntnorm=3, deps=deps,
              xl0=p.xl)
    with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
        s.put_state(w.x, w.v)
        s.accel(True)
        recs = s.run_md(**kw)
        x, v, a = s.get_state()
        box = s.params.xl
    a0, _ = oracle.accel(p, w.x, w.Z, w.A, True)
    xr, vr, ar, rr, boxr = oracle.md_loop(p, w.dt, w.x, w.v, a0, w.Z, w.A, **kw)
    assert box == boxr
    assert [r["istep"] for r in recs] == [r["istep"] for r in rr] == [4, 6, 8, 12]
    Lr = np.array(boxr)
    dx = np.abs(x - xr)
    dx = np.minimum(dx, Lr - dx)
    assert dx.max() <= 1e-10 * Lr.max()
    np.testing.assert_allclose(v, vr, rtol=1e-9, atol=1e-12 * np.abs(vr).max())
    np.testing.assert_allclose(a, ar, rtol=1e-7, atol=1e-10 * np.abs(ar).max())
    for g, r in zip(recs, rr):
        for k in ("ev", "ek", "px"):
            assert g[k] == pytest.approx(r[k], rel=1e-10), (r["istep"], k)
        np.testing.assert_allclose(g["pp"], [r[k] for k in ("pp11", "pp12", "pp13", "pp22",
                                                            "pp23", "pp33")],
                                   rtol=1e-9, atol=1e-12 * abs(r["pp11"]))


def test_run_md_without_utilities_is_run():
    """With no cancel_vcom, tnorm or strain the loop is yk_run, bit for bit."""
    w = W.ion_mixture(n=2500, seed=44)
    out = []
    for md in (False, True):
        with YukawaSystem(w.params, Z=w.Z, A=w.A, dt=w.dt) as s:
            s.put_state(w.x, w.v)
            s.accel(True)
            recs = s.run_md(nsteps=30, nind=10) if md else s.run(30, nind=10)
            out.append((s.get_state(), recs))
    (sa, ra), (sb, rb) = out
    for u, v in zip(sa, sb):
        assert np.array_equal(u, v)
    assert [r["istep"] for r in rb] == [10, 20, 30]
    for r1, r2 in zip(ra, rb):
        for k in ("ev", "ek", "px", "vir", "pp"):
            assert r1[k] == r2[k], k


def test_new_box_from_set_params_wraps_device_positions():
    """Positions stay continuous on the device between re-sorts; a new box from
    yk_set_params first wraps them into the old box (as the reference's
    per-step wrap had), so the forces are those of the reference's positions
    in the new box."""
    w = W.ion_mixture(n=3000, seed=45)
    p = w.params
    with YukawaSystem(p.replace(resort_every=1000), Z=w.Z, A=w.A, dt=w.dt * 4) as s:
        s.put_state(w.x, w.v)
        s.accel(False)
        s.run(12)
        x_ref = s.get_state()[0]          # in the old box (host wrap)
        p2 = p.replace(xl=tuple(1.002 * np.array(p.xl)), resort_every=1000)
        s.set_params(p2)
        s.accel(False)
        a = s.a
    ar, _ = oracle.accel(p2, x_ref, w.Z, w.A, False)
    np.testing.assert_allclose(a, ar, rtol=1e-9, atol=1e-11 * np.abs(ar).max())
