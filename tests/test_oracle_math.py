"""Pins the oracle to the mathematics, independently of the reference build.

* two- and three-body closed forms evaluated in mpmath (Yukawa with Z_i Z_j,
  HPP with both isospin signs, p-p screened Coulomb), including the minimum
  image across the box edge, the inclusive cutoff and the energy shift;
* F = -grad V by central differences of the oracle's own total energy;
* Newton's third law, virial = sum r.F, tr(pressure tensor) = virial;
* the long-double full-matrix "truth" against the oracle at a few hundred
  particles;
* velocity-Verlet algebra: free flight is exact, energy is conserved to
  O(dt^2) for a bound pair.
"""
import math

import mpmath as mp
import numpy as np
import pytest

import oracle
from yukawamd_amd import workloads as W
from yukawamd_amd.params import VC, SimParams

mp.mp.dps = 40


def ion_params(kind="ion-mixture", L=100.0, mu=0.1, rc=40.0, **kw):
    d = dict(sim_type=kind, xl=(L, L, L), xmuc=mu, rccut=rc, rho=1e-4, kT=0.05)
    d.update(kw)
    return SimParams(**d)


def yukawa_pair_mp(zi, zj, r, mu, rc, frp=1.0, L=None):
    """Pair force magnitude (along d) and energy from the closed form."""
    r = mp.mpf(r)
    mu = mp.mpf(mu)
    s = mp.mpf(frp) * mp.mpf(VC) * zi * zj * (1 + mu * r) * mp.e ** (-mu * r) / r ** 3
    diag = mp.sqrt(3) * L / 2 if L is not None else mp.inf
    evcut = mp.e ** (-mu * rc) / rc if rc <= diag else 0
    phi = mp.mpf(frp) * mp.mpf(VC) * zi * zj * (mp.e ** (-mu * r) / r - evcut)
    return s, phi


def test_two_ions_closed_form_across_boundary():
    L = 100.0
    p = ion_params(L=L, mu=0.07, rc=45.0)
    # separation 7 fm across the x boundary: minimum image
    x = np.array([[1.5, 20.0, 30.0], [L - 3.0, 22.0, 33.0]])
    Z = np.array([8.0, 26.0])
    A = np.array([16.0, 56.0])
    d = np.array([1.5 - (L - 3.0) + L, -2.0, -3.0])
    r = float(np.sqrt((d * d).sum()))
    s, phi = yukawa_pair_mp(8, 26, r, 0.07, 45.0, L=L)
    for dm in (True, False):
        a, m = oracle.accel(p, x, Z, A, dm, ref_f1=0)
        f0 = a[0] * A[0] * p.xmass
        f1 = a[1] * A[1] * p.xmass
        for k in range(3):
            assert f0[k] == pytest.approx(float(s * d[k]), rel=1e-14)
            assert f1[k] == pytest.approx(-float(s * d[k]), rel=1e-14)
        if dm:
            assert m["ev"] * 2 == pytest.approx(float(phi), rel=1e-13)
            vir = float(s) * r * r
            assert m["pp11"] + m["pp22"] + m["pp33"] == pytest.approx(vir, rel=1e-13)
            assert m["pp12"] == pytest.approx(float(s) * d[0] * d[1], rel=1e-13)
            assert m["px"] == pytest.approx(p.rho * (p.kT + vir / 6.0), rel=1e-13)


def test_cutoff_inclusive_and_energy_shift_switch():
    # r exactly rc is included (r2 .le. rccut2); just beyond it is not
    L = 64.0
    x = np.array([[10.0, 10.0, 10.0], [10.0 + 16.0, 10.0, 10.0]])
    Z = np.array([8.0, 8.0])
    A = np.array([16.0, 16.0])
    p = ion_params("pure-ion", L=L, mu=0.1, rc=16.0, zi=8.0, ai=16.0)
    a, m = oracle.accel(p, x, Z, A, True)
    assert a[0, 0] != 0.0 and m["ev"] == pytest.approx(0.0, abs=1e-15)   # phi(rc) - evcut
    p2 = p.replace(rccut=15.999999)
    a2, m2 = oracle.accel(p2, x, Z, A, True)
    assert a2[0, 0] == 0.0 and m2["ev"] == 0.0
    # rccut beyond half the box diagonal: no shift (evcut = 0)
    p3 = p.replace(rccut=0.51 * math.sqrt(3) * L)
    _, m3 = oracle.accel(p3, x, Z, A, True)
    _, phi = yukawa_pair_mp(8, 8, 16.0, 0.1, p3.rccut, L=L)
    assert m3["ev"] * 2 == pytest.approx(float(phi), rel=1e-13)
    assert float(phi) == pytest.approx(64 * VC * math.exp(-1.6) / 16.0, rel=1e-13)


@pytest.mark.parametrize("zi,zj,sign", [(1.0, 1.0, +1), (0.0, 0.0, +1), (1.0, 0.0, -1)])
def test_nucleon_pair_closed_form(zi, zj, sign):
    L = 60.0
    p = SimParams(sim_type="nucleon", xl=(L, L, L), xmuc=0.1, rccut=20.0, rncut=20.0,
                  rho=0.05, kT=1.0, xmass=939.0)
    x = np.array([[5.0, 5.0, 5.0], [5.7, 6.1, 4.2]])
    Z = np.array([zi, zj])
    A = np.ones(2)
    d = x[0] - x[1]
    r2 = mp.mpf(float((d * d).sum()))
    lam = mp.mpf(1.25)
    g = mp.e ** (-r2 / (2 * lam))
    bc = mp.mpf(-26 + sign * 24)
    V = 110 * g * g + bc * g
    s = (2 * 110 * g + bc) * g / lam          # -dV/dr / r
    if zi == 1.0 and zj == 1.0:
        r = mp.sqrt(r2)
        evcut = mp.mpf(VC) * mp.e ** (-mp.mpf(0.1) * 20) / 20
        V += mp.mpf(VC) * mp.e ** (-mp.mpf(0.1) * r) / r - evcut
        s += mp.mpf(VC) * (1 + mp.mpf(0.1) * r) * mp.e ** (-mp.mpf(0.1) * r) / r ** 3
    for dm in (True, False):
        a, m = oracle.accel(p, x, Z, A, dm)
        for k in range(3):
            assert a[0, k] * 939.0 == pytest.approx(float(s * d[k]), rel=1e-13)
            assert a[1, k] == -a[0, k]
        if dm:
            assert m["ev"] * 2 == pytest.approx(float(V), rel=1e-13)


def _total_energy(p, x, Z, A):
    _, m = oracle.accel(p, x, Z, A, True, ref_f1=0)
    return m["ev"] * x.shape[0]


@pytest.mark.parametrize("make", [lambda: W.ion_mixture(n=40, seed=21, kappa=1.5),
                                  lambda: W.nucleon(n=60, seed=22, rho=0.01)])
def test_force_is_minus_gradient(make):
    w = make()
    p = w.params.replace(rccut=1e50) if w.params.sim_type != "nucleon" else w.params
    if p.sim_type == "nucleon":
        p = p.replace(rccut=1e50, rncut=1e50)
    a, _ = oracle.accel(p, w.x, w.Z, w.A, False, ref_f1=0)
    mass = (w.A if p.sim_type != "nucleon" else np.ones(w.n)) * p.xmass
    F = a * mass[:, None]
    rng = np.random.default_rng(3)
    for i in rng.choice(w.n, 4, replace=False):
        for k in range(3):
            h = 1e-4
            xp = w.x.copy(); xp[i, k] += h
            xm = w.x.copy(); xm[i, k] -= h
            grad = (_total_energy(p, xp, w.Z, w.A) - _total_energy(p, xm, w.Z, w.A)) / (2 * h)
            assert -grad == pytest.approx(F[i, k], rel=2e-6, abs=1e-9 * np.abs(F).max())


@pytest.mark.parametrize("make", [lambda: W.pure_ion(n=128, seed=31),
                                  lambda: W.ion_mixture(n=160, seed=32),
                                  lambda: W.nucleon(n=400, seed=33)])
def test_invariants(make):
    w = make()
    p = w.params
    a, m = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=0)
    mass = (w.A if p.sim_type != "nucleon" else np.ones(w.n)) * p.xmass
    F = a * mass[:, None]
    scale = np.abs(F).sum()
    assert np.abs(F.sum(axis=0)).max() < 1e-13 * scale          # Newton's third law
    # virial from the pressure: vir = 3n(px/rho - kT); trace of the tensor = vir
    vir = 3 * w.n * (m["px"] / p.rho - p.kT)
    assert m["pp11"] + m["pp22"] + m["pp33"] == pytest.approx(vir, rel=1e-9)
    # translation invariance (positions shifted by a whole box are identical)
    shift = np.array([0.37, -1.1, 2.9])
    xs = np.mod(w.x + shift, p.xl[0])
    a2, m2 = oracle.accel(p, xs, w.Z, w.A, True, ref_f1=0)
    np.testing.assert_allclose(a2, a, rtol=1e-9, atol=1e-12 * np.abs(a).max())
    assert m2["ev"] == pytest.approx(m["ev"], rel=1e-12)
    # permutation equivariance
    perm = np.random.default_rng(5).permutation(w.n)
    a3, m3 = oracle.accel(p, w.x[perm], w.Z[perm], w.A[perm], True, ref_f1=0)
    np.testing.assert_allclose(a3, a[perm], rtol=1e-11, atol=1e-13 * np.abs(a).max())


@pytest.mark.parametrize("make", [lambda: W.ion_mixture(n=300, seed=41),
                                  lambda: W.nucleon(n=500, seed=42)])
def test_against_long_double_truth(make):
    w = make()
    p = w.params
    a, m = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=0)
    f, ev, vir, pp = oracle.truth(p, w.x, w.Z)
    mass = (w.A if p.sim_type != "nucleon" else np.ones(w.n)) * p.xmass
    F = a * mass[:, None]
    rowscale = np.abs(F).max()
    np.testing.assert_allclose(F, f, rtol=1e-11, atol=1e-13 * rowscale)
    assert m["ev"] * w.n == pytest.approx(ev, rel=1e-13)
    assert m["pp11"] == pytest.approx(pp[0], rel=1e-12)
    assert m["pp23"] == pytest.approx(pp[4], rel=1e-9, abs=1e-12 * abs(pp[0]))


def test_free_flight_is_exact_and_wraps():
    # 'B-field' sim type sets a = 0 (src/int_control.f:71-72): x(t) = x0 + v t
    L = 50.0
    p = SimParams(sim_type="B-field", xl=(L, L, L), rccut=10.0)
    x = np.array([[1.0, 2.0, 3.0], [49.0, 25.0, 0.5]])
    v = np.array([[-0.25, 0.0, 0.125], [0.5, -0.5, -0.25]])
    Z = np.ones(2)
    A = np.ones(2)
    a = np.zeros((2, 3))
    xo, vo, _, ao, _ = oracle.newton(p, 8.0, x, v, np.zeros_like(v), a, Z, A)
    np.testing.assert_array_equal(vo, v)
    np.testing.assert_array_equal(xo, np.mod(x + 8.0 * v, L))


def test_vv_energy_conservation_second_order():
    # a bound-ish pair of ions in a large box; energy error scales as dt^2
    L = 400.0
    p = SimParams(sim_type="ion-mixture", xl=(L, L, L), xmuc=0.02, rccut=150.0, rho=1e-6,
                  kT=0.01)
    Z = np.array([8.0, 8.0])
    A = np.array([16.0, 16.0])
    x0 = np.array([[100.0, 100.0, 100.0], [108.0, 100.0, 100.0]])
    v0 = np.array([[0.0, 0.002, 0.0], [0.0, -0.002, 0.0]])

    def drift(dt, T=2000.0):
        a, m = oracle.accel(p, x0, Z, A, True, ref_f1=0)
        e0 = m["ev"] + oracle.ttot(p, v0, A)
        x, v, vold = x0, v0, np.zeros_like(v0)
        worst = 0.0
        for s in range(int(round(T / dt))):
            x, v, vold, a, m = oracle.newton(p, dt, x, v, vold, a, Z, A, True, ref_f1=0)
            worst = max(worst, abs(m["ev"] + m["ek"] - e0))
        return worst

    e1, e2 = drift(20.0), drift(10.0)
    assert e2 < e1
    assert 2.5 < e1 / e2 < 6.0     # second order: ratio ~4


def test_auto_rncut_matches_closed_form_loop():
    from yukawamd_amd.params import auto_rncut
    r = auto_rncut()
    # the nuclear force at rncut is below 2^-53 * 1e-4 of the packet scale
    assert 10.0 < r < 13.0


def test_spreiter_walter_reduces_to_kdk_without_field():
    # vel_verlet_3 with B = 0: both stages become plain half kicks, so the step
    # is vel_verlet_1's kick-drift-kick, bit for bit
    w = W.ion_mixture(n=150, seed=17)
    p = w.params
    a0, _ = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=0)
    z = np.zeros_like(w.v)
    r1 = oracle.newton(p, w.dt, w.x, w.v, z, a0, w.Z, w.A, True, ref_f1=0, kdk=True)
    r3 = oracle.newton(p, w.dt, w.x, w.v, z, a0, w.Z, w.A, True, ref_f1=0, sw=True)
    for u, v in zip(r1[:4], r3[:4]):
        np.testing.assert_array_equal(u, v)
    assert r1[4] == r3[4]


def test_spreiter_walter_free_particle_gyrates():
    # no pair force ('B-field' sim type): a charge in B along z keeps v_z, drifts
    # exactly, and its transverse velocity turns by the cyclotron angle
    # -omega*dt (to first order in omega*dt) in the sense of F = q v x B
    L = 1.0e6
    dt = 2.0
    p = SimParams(sim_type="B-field", xl=(L, L, L), rccut=10.0, bfield=1.0e16)
    Z = np.array([8.0, 26.0])
    A = np.array([16.0, 56.0])
    x = np.array([[100.0, 200.0, 300.0], [400.0, 500.0, 600.0]])
    v = np.array([[0.01, -0.02, 0.005], [-0.03, 0.01, -0.004]])
    a = np.zeros((2, 3))
    xo, vo, _, ao, _ = oracle.newton(p, dt, x, v, np.zeros_like(v), a, Z, A, sw=True)
    np.testing.assert_array_equal(xo, x + v * dt)
    np.testing.assert_array_equal(vo[:, 2], v[:, 2])
    omega = np.float64(np.float32(3.218404e-20)) * Z * p.bfield / A
    th = omega * dt
    assert np.all(np.abs(th) < 1e-2) and np.all(np.abs(th) > 1e-4)
    exp_vx = v[:, 0] + th * v[:, 1]
    exp_vy = v[:, 1] - th * v[:, 0]
    scale = np.abs(v[:, :2]).max(axis=1)
    assert np.all(np.abs(vo[:, 0] - exp_vx) <= 2 * th ** 2 * scale)
    assert np.all(np.abs(vo[:, 1] - exp_vy) <= 2 * th ** 2 * scale)
    # transverse speed is kept to second order
    s0 = np.hypot(v[:, 0], v[:, 1])
    s1 = np.hypot(vo[:, 0], vo[:, 1])
    assert np.all(np.abs(s1 - s0) <= 2 * th ** 2 * s0)


# --- velocity utilities and strain: what the mathematics fixes ---------------
def test_tnorm_sets_the_temperature():
    """After tnorm the kinetic energy per particle is 1.5 kT (equipartition,
    src/gen_utils.f:58-60), whatever the starting scale."""
    w = W.ion_mixture(n=500, seed=31)
    for scale in (0.3, 1.0, 4.0):
        v = oracle.tnorm(w.params, w.v * scale, w.A)
        assert oracle.ttot(w.params, v, w.A) == pytest.approx(1.5 * w.params.kT, rel=1e-14)
        # a uniform rescale: directions unchanged
        np.testing.assert_allclose(v / np.linalg.norm(v), w.v / np.linalg.norm(w.v), rtol=1e-13)


def test_cancel_vcom_removes_momentum():
    """Mass-weighted momentum vanishes after cancel_vcom, and the temperature is
    restored to 1.5 kT by its closing tnorm."""
    w = W.ion_mixture(n=500, seed=32)
    v0 = w.v + np.array([0.003, -0.002, 0.001])
    v = oracle.cancel_vcom(w.params, v0, w.A)
    ptot = (w.A[:, None] * v).sum(axis=0)
    assert np.abs(ptot).max() <= 1e-13 * (w.A[:, None] * np.abs(v)).sum()
    assert oracle.ttot(w.params, v, w.A) == pytest.approx(1.5 * w.params.kT, rel=1e-14)


@pytest.mark.parametrize("deps", [(2e-6, None, None), (None, -3e-6, None), (None, None, 1e-6),
                                  (2e-6, -1e-6, None), (-2e-6, None, 1e-6),
                                  (None, 1e-6, -2e-6)])
def test_strain_keeps_the_volume(deps):
    """One or two strained axes: the others follow so that the volume stays the
    gage volume (src/strain.f:15-17); a strained edge grows by deps*dt*xl0."""
    X = oracle.XNOSTRAIN
    d = tuple(X if e is None else e for e in deps)
    w = W.ion_mixture(n=50, seed=33)
    xl0 = w.params.xl
    p = w.params
    dt = 1000.0
    box = p.xl
    for _ in range(5):
        _, box = oracle.strain(p.replace(xl=box), w.x, xl0, d, dt)
    assert math.prod(box) == pytest.approx(math.prod(xl0), rel=1e-14)
    for k in range(3):
        if deps[k] is not None:
            assert box[k] == pytest.approx(xl0[k] * (1 + 5 * deps[k] * dt), rel=1e-14)


def test_strain_wraps_only_beyond_the_new_box():
    """Positions above a shrunk edge move down by one edge length; all others
    stay (src/strain.f:76-82)."""
    X = oracle.XNOSTRAIN
    w = W.ion_mixture(n=400, seed=34)
    p = w.params
    x1, box = oracle.strain(p, w.x, p.xl, (-1e-3, X, X), 10.0)
    assert box[0] < p.xl[0] and box[1] > p.xl[1]
    moved = w.x[:, 0] > box[0]
    assert moved.any()
    np.testing.assert_array_equal(x1[moved, 0], w.x[moved, 0] - box[0])
    np.testing.assert_array_equal(x1[~moved], w.x[~moved])
    np.testing.assert_array_equal(x1[:, 1:], w.x[:, 1:])


@pytest.mark.parametrize("make", [lambda: W.ion_mixture(n=900, seed=131),
                                  lambda: W.pure_ion(n=700, seed=132)])
@pytest.mark.parametrize("threads", [1, 3])
def test_openmp_restatement_matches_the_force_branch(make, threads):
    """yko_accel_omp (the CPU-baseline timing restatement: the force-only
    branch with the reference's j > i loop split over threads) gives the
    force-only branch's accelerations, with the reference's last-particle
    behaviour (F1), to rounding: the per-thread reductions only reorder sums."""
    w = make()
    a_ref, _ = oracle.accel(w.params, w.x, w.Z, w.A, False, ref_f1=1)
    a_omp, used = oracle.accel_omp(w.params, w.x, w.Z, w.A, threads)
    assert used == threads
    assert np.abs(a_omp - a_ref).max() <= 1e-13 * np.abs(a_ref).max()


def test_openmp_restatement_is_ions_only():
    w = W.nucleon(n=200, seed=133)
    with pytest.raises(ValueError):
        oracle.accel_omp(w.params, w.x, w.Z, w.A, 1)


@pytest.mark.parametrize("make", [lambda: W.ion_mixture(n=600, seed=134),
                                  lambda: W.pure_ion(n=500, seed=135),
                                  lambda: W.nucleon(n=900, seed=136)])
def test_pair_abs_sum_conditioning(make):
    """oracle.pair_abs_sum (the parity tests' conditioning metric) against a
    numpy restatement, and the bounds it must satisfy: |F_i| <= sum_j |F_ij|
    (triangle inequality, F_i from the long-double truth) and
    sqrt(sum_j |F_ij|^2) <= sum_j |F_ij|."""
    import sys
    from pathlib import Path
    sys.path.insert(0, str(Path(__file__).resolve().parent))
    from _forces import abs_pair_force_sum_np
    w = make()
    sa, rms = oracle.pair_abs_sum(w.params, w.x, w.Z, 2)
    ref = abs_pair_force_sum_np(w.params, w.x, w.Z)
    np.testing.assert_allclose(sa, ref, rtol=1e-12, atol=0)
    f, _, _, _ = oracle.truth(w.params, w.x, w.Z)
    F = np.linalg.norm(f, axis=1)
    assert np.all(F <= sa * (1 + 1e-12))
    assert np.all(rms <= sa * (1 + 1e-12)) and np.all(rms > 0)
