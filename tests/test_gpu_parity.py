"""HIP path against the oracle and the reference's golden vectors (needs a GPU).

Bars (stated per test):
  FP64 forces    condition-aware |dF_i| / sum_j |F_ij| <= 1e-12, and the plain
                 per-particle |dF_i|/|F_i| is reported and bounded by 1e-10
                 (summation order differs from the reference, so a particle
                 whose force nearly cancels loses relative digits in both);
  FP64 energy, pressure, pressure-tensor diagonal: 1e-13 relative;
  mixed form     max |dF_i|/|F_i| and drift reported; bounded in the test.
All calls go through the C ABI (libyukawa_hip.so).
"""
import numpy as np
import pytest

import oracle
from _forces import check_meas, force_errors, gpu_accel, masses
from _golden import NAMES, load
from yukawamd_amd import workloads as W
from yukawamd_amd.system import YukawaSystem

pytestmark = pytest.mark.gpu


# --------------------------------------------------------------------------
@pytest.mark.parametrize("name", NAMES)
def test_golden_fp64_accel(name):
    p, g = load(name)
    p = p.replace(ref_compat_f1=True)
    for dm in (1, 0):
        a, m = gpu_accel(p, g["x"], g["Z"], g["A"], bool(dm))
        rel, cond = force_errors(p, a, g[f"a_meas{dm}"], g["x"], g["Z"], g["A"])
        print(f"{name} meas={dm}: max rel {rel:.2e} cond {cond:.2e}")
        assert cond <= 1e-12 and rel <= 1e-10
        if dm:
            check_meas(m, dict(zip(oracle.MEAS_KEYS, g["m_meas1"])))


@pytest.mark.parametrize("name", NAMES)
def test_golden_fp64_three_steps(name):
    p, g = load(name)
    p = p.replace(ref_compat_f1=True)
    with YukawaSystem(p, Z=g["Z"], A=g["A"], dt=float(g["dt"])) as s:
        s.put_state(g["x"], g["v"])
        s.accel(True)
        s.newton(False)
        s.newton(False)
        m = s.newton(True)
        x, v, a = s.get_state()
    L = p.xl[0]
    dx = np.abs(x - g["x3"])
    dx = np.minimum(dx, L - dx)
    assert dx.max() <= 1e-12 * L
    assert np.abs(v - g["v3"]).max() <= 1e-11 * np.abs(g["v3"]).max()
    mref = dict(zip(oracle.MEAS_KEYS, g["m3"]))
    assert m["ek"] == pytest.approx(mref["ek"], rel=1e-12)
    assert m["ev"] == pytest.approx(mref["ev"], rel=1e-12)


@pytest.mark.parametrize("make", [
    lambda: W.pure_ion(n=3456, seed=1000),          # the pure-ion workload at full size
    lambda: W.ion_mixture(n=2000, seed=51),
    lambda: W.nucleon(n=4000, seed=52),
    lambda: W.ion_mixture(n=1024, seed=53, kappa=2.0),
])
@pytest.mark.parametrize("path", ["clist", "tilelist", "allpairs"])
def test_fp64_against_oracle(make, path):
    w = make()
    p = w.params
    for dm in (True, False):
        a_o, m_o = oracle.accel(p, w.x, w.Z, w.A, dm, ref_f1=0)
        a_g, m_g = gpu_accel(p, w.x, w.Z, w.A, dm, path)
        rel, cond = force_errors(p, a_g, a_o, w.x, w.Z, w.A)
        print(f"{w.name} {path} meas={dm}: max rel {rel:.2e} cond {cond:.2e}")
        assert cond <= 1e-12 and rel <= 1e-10
        if dm:
            check_meas(m_g, m_o)


@pytest.mark.parametrize("make", [lambda: W.ion_mixture(n=3000, seed=61),
                                  lambda: W.pure_ion(n=3456, seed=1000),
                                  lambda: W.nucleon(n=6000, seed=62),
                                  # the ensemble's screening range: the FP64 near
                                  # field's radius follows the force's decay
                                  lambda: W.ion_mixture(n=4000, seed=63, kappa=0.5),
                                  lambda: W.ion_mixture(n=4000, seed=64, kappa=2.0),
                                  lambda: W.pure_ion(n=4000, seed=65, kappa=2.0)])
def test_mixed_precision_error(make):
    w = make()
    p = w.params
    a_o, m_o = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=0)
    a_g, m_g = gpu_accel(p.replace(precision="mixed"), w.x, w.Z, w.A, True)
    rel, cond = force_errors(p, a_g, a_o, w.x, w.Z, w.A)
    print(f"{w.name} mixed: max rel {rel:.2e} cond {cond:.2e} "
          f"ev rel {abs(m_g['ev'] / m_o['ev'] - 1):.2e}")
    if p.sim_type == "nucleon":     # no near field: every pair in FP32
        assert cond <= 1e-6 and rel <= 1e-4
    else:                           # FP64 near field (cluster lists)
        assert cond <= 1e-7 and rel <= 1e-6
    assert m_g["ev"] == pytest.approx(m_o["ev"], rel=1e-6)


def test_ion_mixture_27648_properties():
    """Full-size mixture: Newton's 3rd law, determinism, run == repeated newton."""
    w = W.ion_mixture()
    p = w.params
    with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
        s.put_state(w.x, w.v)
        m1 = s.accel(True)
        a1 = s.a
        m2 = s.accel(True)
        a2 = s.a
        assert np.array_equal(a1, a2) and m1 == m2            # bitwise deterministic
        F = a1 * masses(p, w.A)[:, None]
        assert np.abs(F.sum(0)).max() <= 1e-12 * np.abs(F).sum()
        for st in range(1, 5):
            s.newton(st % 2 == 0)     # the cadence run(4, nind=2) uses
        xa, va, aa = s.get_state()
    with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
        s.put_state(w.x, w.v)
        s.accel(True)
        recs = s.run(4, nind=2)
        xb, vb, ab = s.get_state()
    assert len(recs) == 2
    assert np.array_equal(xa, xb) and np.array_equal(va, vb) and np.array_equal(aa, ab)


@pytest.mark.parametrize("kind", ["pure", "mixture"])
def test_run_is_repeated_newton_across_resorts(kind):
    """A run loop does each step's drift inside the previous step's finalize
    except before a re-sort and at its end; 23 steps with a re-sort every 5
    and measurements every 3rd cross all of these and must equal 23 single
    newton calls bit for bit, records included."""
    w = W.ion_mixture(n=3000, seed=77) if kind == "mixture" else W.pure_ion(n=3000, seed=77)
    p = w.params.replace(resort_every=5)
    out = []
    for loop in (False, True):
        with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
            s.put_state(w.x, w.v)
            s.accel(True)
            if loop:
                recs = s.run(23, nind=3)
            else:
                recs = [s.newton(st % 3 == 0) for st in range(1, 24)]
                recs = [r for st, r in enumerate(recs, 1) if st % 3 == 0]
            out.append((s.get_state(), recs))
    (sa, ra), (sb, rb) = out
    for u, v in zip(sa, sb):
        assert np.array_equal(u, v)
    assert len(ra) == len(rb) == 7
    for r1, r2 in zip(ra, rb):
        for k in ("ev", "ek", "px", "vir", "pp"):
            assert np.array_equal(r1[k], r2[k]), k


@pytest.mark.parametrize("n", [6000, 10000, 17000])
def test_item_partial_sums_against_oracle(n):
    """The cluster-list launch takes 128 items per cluster set below 8,192
    particles, 64 up to 16,384 and 32 above; from 64 items the partials are
    summed by k_part_reduce (16 and 8 threads per slot) before the finalize,
    below that by the finalize itself.  Each against the oracle, measuring
    and not, and a run loop (fused drift) equal to single newton steps."""
    w = W.ion_mixture(n=n, seed=95)
    p = w.params
    for dm in (True, False):
        a_o, m_o = oracle.accel(p, w.x, w.Z, w.A, dm, ref_f1=0)
        a_g, m_g = gpu_accel(p, w.x, w.Z, w.A, dm)
        rel, cond = force_errors(p, a_g, a_o, w.x, w.Z, w.A)
        print(f"n={n} meas={dm}: max rel {rel:.2e} cond {cond:.2e}")
        assert cond <= 1e-12 and rel <= 1e-10
        if dm:
            check_meas(m_g, m_o)
    out = []
    for loop in (False, True):
        with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
            s.put_state(w.x, w.v)
            s.accel(True)
            if loop:
                recs = s.run(6, nind=3)
            else:
                recs = [s.newton(st % 3 == 0) for st in range(1, 7)]
                recs = [r for st, r in enumerate(recs, 1) if st % 3 == 0]
            out.append((s.get_state(), recs))
    (sa, ra), (sb, rb) = out
    for u, v in zip(sa, sb):
        assert np.array_equal(u, v)
    assert len(ra) == len(rb) == 2
    for r1, r2 in zip(ra, rb):
        assert r1 == r2


def test_cluster_lists_above_65536():
    """A single system of 100,000 particles runs on the cluster lists (6,250
    clusters: after a re-sort the cluster order comes from a device radix sort
    instead of k_corder's one-workgroup bitonic sort).  A run with re-sorts
    every 5 steps equals single newton steps bit for bit, and the forces at
    its end match the oracle (OpenMP restatement) and the tile lists."""
    w = W.ion_mixture(n=100_000, seed=96)
    p = w.params.replace(resort_every=5)
    out = []
    for loop in (False, True):
        with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
            s.put_state(w.x, w.v)
            s.accel(True)
            if loop:
                recs = s.run(12, nind=6)
            else:
                recs = [s.newton(st % 6 == 0) for st in range(1, 13)]
                recs = [r for st, r in enumerate(recs, 1) if st % 6 == 0]
            out.append((s.get_state(), recs))
    (sa, ra), (sb, rb) = out
    for u, v in zip(sa, sb):
        assert np.array_equal(u, v)
    assert len(ra) == len(rb) == 2 and ra == rb
    x, _, a = sa
    a_o, _ = oracle.accel_omp(p, x, w.Z, w.A)
    # accel_omp keeps the reference's last-particle behaviour (F1): the last
    # particle's force lacks its own factor Z
    a_o[-1] *= w.Z[-1]
    cond_sum = oracle.pair_abs_sum(p, x, w.Z)[0]
    a_t, _ = gpu_accel(p, x, w.Z, w.A, False, "tilelist")
    for name, a_ref in (("oracle", a_o), ("tile lists", a_t)):
        rel, cond = force_errors(p, a, a_ref, x, w.Z, w.A, cond_sum)
        print(f"n=100000 cluster lists vs {name}: max rel {rel:.2e} cond {cond:.2e}")
        assert cond <= 1e-12 and rel <= 1e-10


def test_large_systems_mixed_and_nucleon():
    """Above 65,536 particles on cluster lists (radix-sorted cluster order):
    the mixed ion form at N = 100,000 against the FP64 tile lists (the mixed
    bar: max |dF|/|F| <= 1e-6), and nucleons at N = 100,000 (charge-sorted
    Hilbert runs, neutral sources listed to the nuclear radius only) against
    the tile lists in FP64 (|dF_i| <= 1e-12 of the largest force: the two
    differ only in summation order), with the total force vanishing."""
    w = W.ion_mixture(n=100_000, seed=98)
    p = w.params
    a_t, m_t = gpu_accel(p, w.x, w.Z, w.A, True, "tilelist")
    a_m, m_m = gpu_accel(p.replace(precision="mixed"), w.x, w.Z, w.A, True, "clist")
    m = masses(p, w.A)[:, None]
    rel = (np.linalg.norm((a_m - a_t) * m, axis=1) / np.linalg.norm(a_t * m, axis=1)).max()
    print(f"n=100000 mixed cluster lists vs FP64 tile lists: max |dF|/|F| {rel:.2e}")
    assert rel <= 1e-6
    assert m_m["ev"] == pytest.approx(m_t["ev"], rel=1e-6)
    w = W.nucleon(n=100_000, seed=99)
    p = w.params
    a_t, m_t = gpu_accel(p, w.x, w.Z, w.A, True, "tilelist")
    a_c, m_c = gpu_accel(p, w.x, w.Z, w.A, True, "clist")
    m = masses(p, w.A)[:, None]
    dF = np.linalg.norm((a_c - a_t) * m, axis=1)
    F = np.linalg.norm(a_t * m, axis=1)
    print(f"n=100000 nucleons cluster lists vs tile lists: max |dF|/|F| "
          f"{(dF / F).max():.2e}, max |dF| / max |F| {dF.max() / F.max():.2e}")
    assert dF.max() <= 1e-12 * F.max()
    assert m_c["ev"] == pytest.approx(m_t["ev"], rel=1e-12)
    Fc = a_c * m
    assert np.abs(Fc.sum(0)).max() <= 1e-12 * np.abs(Fc).sum()


def test_cluster_lists_at_the_size_limit():
    """262,144 particles, the largest single system on cluster lists: 2^32
    list entries (16 GiB), so segment offsets pass 2^31 entries (a regression
    test for their widening in the build kernel).  Forces and measurements
    against the tile lists; the total force vanishes (Newton's third law)."""
    w = W.ion_mixture(n=262_144, seed=97)
    p = w.params
    a_c, m_c = gpu_accel(p, w.x, w.Z, w.A, True, "clist")
    a_t, m_t = gpu_accel(p, w.x, w.Z, w.A, True, "tilelist")
    m = masses(p, w.A)[:, None]
    dF = np.linalg.norm((a_c - a_t) * m, axis=1)
    F = np.linalg.norm(a_t * m, axis=1)
    rel = (dF / F).max()
    print(f"n=262144 cluster lists vs tile lists: max |dF|/|F| {rel:.2e}")
    assert rel <= 1e-10
    assert m_c["ev"] == pytest.approx(m_t["ev"], rel=1e-12)
    assert m_c["px"] == pytest.approx(m_t["px"], rel=1e-12)
    diag = max(abs(m_t["pp"][k]) for k in (0, 3, 5))
    assert np.abs(np.array(m_c["pp"]) - np.array(m_t["pp"])).max() <= 1e-12 * diag
    Fc = a_c * m
    assert np.abs(Fc.sum(0)).max() <= 1e-12 * np.abs(Fc).sum()


def test_force_independent_of_measurement():
    """Unlike the reference's two branches (F2), the device force does not depend
    on do_measurements: both templates accumulate it with the same FMAs."""
    w = W.ion_mixture(n=5000, seed=91)
    with YukawaSystem(w.params, Z=w.Z, A=w.A) as s:
        s.put_state(w.x)
        s.accel(True)
        a1 = s.a
        s.accel(False)
        a2 = s.a
    assert np.array_equal(a1, a2)


def test_kdk_against_oracle():
    w = W.ion_mixture(n=700, seed=71)
    p = w.params.replace(integrator="kdk")
    a, _ = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=0)
    x, v, vold = w.x, w.v, np.zeros_like(w.v)
    for st in range(3):
        x, v, vold, a, m = oracle.newton(p, w.dt, x, v, vold, a, w.Z, w.A, st == 2, kdk=True)
    with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
        s.put_state(w.x, w.v)
        s.accel(True)
        s.newton(False)
        s.newton(False)
        mg = s.newton(True)
        xg, vg, ag = s.get_state()
    assert np.abs(vg - v).max() <= 1e-11 * np.abs(v).max()
    assert mg["ek"] == pytest.approx(m["ek"], rel=1e-12)


@pytest.mark.parametrize("integrator", ["sw", "vv2"])
@pytest.mark.parametrize("make", [lambda: W.ion_mixture(n=900, seed=72),
                                  lambda: W.pure_ion(n=500, seed=73)])
def test_bfield_integrators_against_oracle(make, integrator):
    """Uniform B field along z: the Spreiter-Walter integrator
    (src/vel_verlet_3.f) and vel_verlet_2's Lorentz term
    (src/vel_verlet_2.f:61-69), three steps, against the oracle."""
    w = make()
    p = w.params.replace(integrator=integrator, bfield=2.0e17)
    a, _ = oracle.accel(p, w.x, w.Z, w.A, True, ref_f1=0)
    x, v, vold = w.x, w.v, np.zeros_like(w.v)
    for st in range(3):
        x, v, vold, a, m = oracle.newton(p, w.dt, x, v, vold, a, w.Z, w.A, st == 2,
                                         sw=integrator == "sw")
    with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
        s.put_state(w.x, w.v)
        s.accel(True)
        s.newton(False)
        s.newton(False)
        mg = s.newton(True)
        xg, vg, ag = s.get_state()
    L = p.xl[0]
    dx = np.abs(xg - x)
    dx = np.minimum(dx, L - dx)
    assert dx.max() <= 1e-12 * L
    assert np.abs(vg - v).max() <= 1e-11 * np.abs(v).max()
    assert mg["ek"] == pytest.approx(m["ek"], rel=1e-12)
    assert mg["pp"][0] == pytest.approx(m["pp11"], rel=1e-11)


def test_error_paths():
    from yukawamd_amd._lib import YkError
    w = W.ion_mixture(n=64, seed=81)
    with YukawaSystem(w.params, Z=w.Z, A=w.A, dt=w.dt) as s:
        bad = w.x.copy()
        bad[3, 1] = np.nan
        with pytest.raises(YkError) as e:
            s.put_state(bad)
        assert e.value.code == 5
        s.put_state(w.x, w.v)
        with pytest.raises(YkError) as e:
            s.newton(False)
        assert e.value.code == 4
        s.accel(False)
        s.newton(False)


def test_positions_outside_the_box():
    """Positions handed over outside [0, L) give the forces of their wrapped
    images (minimum image is translation-invariant by whole boxes)."""
    w = W.ion_mixture(n=1500, seed=101)
    L = w.params.xl[0]
    shift = np.random.default_rng(3).integers(-1, 2, size=w.x.shape) * L
    a0, m0 = gpu_accel(w.params, w.x, w.Z, w.A, True)
    a1, m1 = gpu_accel(w.params, w.x + shift, w.Z, w.A, True)
    np.testing.assert_allclose(a1, a0, rtol=1e-10, atol=1e-12 * np.abs(a0).max())
    assert m1["ev"] == pytest.approx(m0["ev"], rel=1e-12)


def test_pairlist_matches_allpairs_over_a_trajectory():
    """60 steps from a bcc start: the pair lists (rebuilt on the device when the
    two largest displacements since the build add up to the skin) and the
    all-pairs kernel give the same trajectory to summation-order rounding."""
    w = W.ion_mixture(n=3456, seed=102, start="bcc")
    out = {}
    for ap in (False, True):
        with YukawaSystem(w.params.replace(allpairs=ap), Z=w.Z, A=w.A, dt=w.dt * 4) as s:
            s.put_state(w.x, w.v)
            s.accel(True)
            recs = s.run(60, nind=20)
            out[ap] = (s.get_state(), recs)
    (xa, va, aa), ra = out[False]
    (xb, vb, ab), rb = out[True]
    L = w.params.xl[0]
    dx = np.abs(xa - xb)
    dx = np.minimum(dx, L - dx)
    assert dx.max() <= 1e-9 * L
    np.testing.assert_allclose(aa, ab, rtol=1e-6, atol=1e-9 * np.abs(ab).max())
    for r1, r2 in zip(ra, rb):
        assert r1["ev"] == pytest.approx(r2["ev"], rel=1e-11)
        assert r1["ek"] == pytest.approx(r2["ek"], rel=1e-10)


def test_nucleon_lists_valid_along_a_trajectory():
    """Nucleon cluster lists (neutrons and protons in separate Hilbert runs,
    neutral sources and all sources of a proton-free cluster listed only out
    to the nuclear radius) stay complete while they are reused: every five
    steps the run's forces match the all-pairs kernel on the same positions,
    and at the end the oracle, which evaluates the Gaussian out to rncut."""
    w = W.nucleon(n=8000, seed=104)
    p = w.params
    with YukawaSystem(p, Z=w.Z, A=w.A, dt=w.dt) as s:
        s.put_state(w.x, w.v)
        s.accel(True)
        for k in range(4):
            s.run(5, nind=5)
            x, v, a = s.get_state()
            a_ap, _ = gpu_accel(p, x, w.Z, w.A, False, "allpairs")
            rel, cond = force_errors(p, a, a_ap, x, w.Z, w.A)
            print(f"after {5 * (k + 1)} steps: lists vs all-pairs max rel {rel:.2e} cond {cond:.2e}")
            assert cond <= 1e-12
    a_o, _ = oracle.accel(p, x, w.Z, w.A, False, ref_f1=0)
    rel, cond = force_errors(p, a, a_o, x, w.Z, w.A)
    print(f"against the oracle: max rel {rel:.2e} cond {cond:.2e}")
    assert cond <= 1e-12


@pytest.mark.parametrize("path", ["clist", "tilelist"])
def test_pairlist_valid_with_one_fast_ion(path):
    """One fast ion (0.45 skin per step) among nearly cold ones: the lists are
    kept while the two largest displacements since the build add up to at
    most the skin, so the fast ion uses lists built up to 0.9 skin away.  After
    every step the forces from the lists match the all-pairs kernel on the
    same positions to summation-order rounding (a pair missed near the cutoff
    would change forces far beyond that)."""
    w = W.ion_mixture(n=3456, seed=103, start="bcc")
    p = w.params
    skin = 0.05 * p.rccut
    v = w.v * 1e-3
    u = np.array([0.6, 0.5, 0.62])
    v[17] = u / np.linalg.norm(u) * 0.45 * skin / w.dt
    lp = p.replace(tile_lists=path == "tilelist")
    with YukawaSystem(lp, Z=w.Z, A=w.A, dt=w.dt) as s, \
            YukawaSystem(p.replace(allpairs=True), Z=w.Z, A=w.A) as ref:
        s.put_state(w.x, v)
        s.accel(False)
        x0 = s.get_state()[0]
        for k in range(12):
            s.run(1)
            x, vv, a = s.get_state()
            ref.put_state(x, vv)
            ref.accel(False)
            ar = ref.a
            np.testing.assert_allclose(a, ar, rtol=1e-10, atol=1e-12 * np.abs(ar).max(),
                                       err_msg=f"{path} step {k + 1}")
    L = p.xl[0]
    d = np.abs(x - x0)
    d = np.minimum(d, L - d)
    assert np.sqrt((d[17] ** 2).sum()) > 3 * skin     # the ion did travel
