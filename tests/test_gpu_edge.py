"""Edge cases of the force path against the oracle (needs a GPU).

* Tiny and ragged systems: 1 ion (no pair at all: zero force, ev = 0,
  px = rho kT), 2 ions, 17 (one full 16-slot cluster plus one slot) and 65
  (one full 64-slot tile plus one slot), so partial clusters and tiles, idle
  lanes and a single-entry list are all exercised.
* Cutoffs beyond half the box (pairs between L/2 and rc exist only through
  the minimum image along some axes) and beyond half the box diagonal (every
  pair within the cutoff and, as in src/int_ion_mix.f:140-145, no energy
  shift).
* Boxes with three different edges.
* A cutoff of 1e8 fm, past the kernels' 1e6/mu limit, and screening so strong
  that mu r exceeds that limit inside the cutoff (exact zeros).
* The 'B-field' sim type: accelerations exactly zero (src/int_control.f:71-72).
* Unscreened Coulomb (mu = 0), FP64 and mixed.

Every case runs the cluster lists, the 64-slot tile lists and the culled
all-pairs kernel, FP64, measuring and force-only, with the bars of the
parity tests: condition-aware force error <= 1e-12, energies and pressure
<= 1e-13 relative; and in mixed precision with the mixed-precision bars.
"""
import numpy as np
import pytest

import oracle
from _forces import check_meas, force_errors, gpu_accel
from yukawamd_amd import workloads as W
from yukawamd_amd._lib import YkError
from yukawamd_amd.params import SimParams
from yukawamd_amd.system import YukawaSystem

pytestmark = pytest.mark.gpu

SPECIES = [(8.0, 16.0), (26.0, 56.0), (34.0, 80.0)]


def mixture(n, xl, rc, seed, kappa=1.0):
    """n ions of the three-species mixture, uniform in the box xl."""
    rng = np.random.Generator(np.random.PCG64(seed))
    k = rng.integers(0, 3, size=n)
    Z = np.array([SPECIES[i][0] for i in k])
    A = np.array([SPECIES[i][1] for i in k])
    x = rng.uniform(0.0, 1.0, size=(n, 3)) * np.array(xl)
    rho = 7.18e-5
    lam = (3.0 / (4.0 * np.pi * rho)) ** (1.0 / 3.0) / kappa
    p = SimParams(sim_type="ion-mixture", xl=tuple(xl), xmuc=1.0 / lam, rccut=rc, rho=rho,
                  kT=0.06)
    return p, x, Z, A


CASES = {
    # name: (n, box edges (fm), cutoff (fm))
    "n1": (1, (100.0, 100.0, 100.0), 40.0),
    "n2": (2, (100.0, 100.0, 100.0), 80.0),
    "n17": (17, (120.0, 120.0, 120.0), 50.0),
    "n65": (65, (150.0, 150.0, 150.0), 60.0),
    "rc_beyond_half_box": (700, (220.0, 220.0, 220.0), 0.6 * 220.0),
    "rc_beyond_half_diagonal": (500, (200.0, 200.0, 200.0), 0.9 * 200.0),
    "three_edges": (900, (180.0, 260.0, 130.0), 60.0),
    "three_edges_rc_beyond_half": (600, (150.0, 240.0, 200.0), 90.0),
    # beyond 1e6 / mu the kernels' cutoff test stops at 1e6 / mu (the FP64
    # exp's range): here that is ~1.5e7 fm, still past every pair of the box
    "rc_huge": (500, (200.0, 200.0, 200.0), 1.0e8),
}


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("path", ["clist", "tilelist", "allpairs"])
def test_edge_case_against_oracle(case, path):
    n, xl, rc = CASES[case]
    p, x, Z, A = mixture(n, xl, rc, seed=900 + len(case) * 7 + n)
    for dm in (True, False):
        a_o, m_o = oracle.accel(p, x, Z, A, dm, ref_f1=0)
        a_g, m_g = gpu_accel(p, x, Z, A, dm, path)
        if n == 1:
            assert np.array_equal(a_g, np.zeros((1, 3))) and np.array_equal(a_o, a_g)
            if dm:
                assert m_g["ev"] == 0.0 and m_o["ev"] == 0.0
                assert m_g["px"] == pytest.approx(p.rho * p.kT, rel=1e-15)
            continue
        rel, cond = force_errors(p, a_g, a_o, x, Z, A)
        print(f"{case} {path} meas={dm}: max rel {rel:.2e} cond {cond:.2e}")
        assert cond <= 1e-12 and rel <= 1e-10
        if dm:
            check_meas(m_g, m_o)


@pytest.mark.parametrize("case", sorted(c for c in CASES if c != "n1"))
@pytest.mark.parametrize("path", ["clist", "tilelist", "allpairs"])
def test_edge_case_mixed_precision(case, path):
    """The mixed form on the same cases, with the mixed-precision bars:
    condition-aware force error <= 1e-6, ev <= 1e-6, and per particle <= 1e-6
    on the cluster lists (FP64 near field) and <= 1e-4 on the tile lists and
    the all-pairs kernel (every pair in FP32)."""
    n, xl, rc = CASES[case]
    p, x, Z, A = mixture(n, xl, rc, seed=900 + len(case) * 7 + n)
    a_o, m_o = oracle.accel(p, x, Z, A, True, ref_f1=0)
    a_g, m_g = gpu_accel(p.replace(precision="mixed"), x, Z, A, True, path)
    rel, cond = force_errors(p, a_g, a_o, x, Z, A)
    print(f"{case} {path} mixed: max rel {rel:.2e} cond {cond:.2e}")
    assert cond <= 1e-6 and rel <= (1e-6 if path == "clist" else 1e-4)
    assert m_g["ev"] == pytest.approx(m_o["ev"], rel=1e-6)


@pytest.mark.parametrize("path", ["clist", "tilelist", "allpairs"])
def test_exponent_beyond_range_contributes_zero(path):
    """Screening so strong that mu r reaches 2e6 inside the cutoff: e^{-mu r}
    is exactly 0 for every pair (so is the energy shift), and the kernels,
    which test pairs against min(rc, 1e6/mu), must return exact zeros too."""
    p, x, Z, A = mixture(400, (200.0, 200.0, 200.0), 150.0, seed=5, kappa=2.0e5)
    assert p.xmuc * p.rccut > 1.5e6
    a_o, m_o = oracle.accel(p, x, Z, A, True, ref_f1=0)
    a_g, m_g = gpu_accel(p, x, Z, A, True, path)
    assert np.array_equal(a_o, np.zeros_like(a_o)) and np.array_equal(a_g, a_o)
    assert m_o["ev"] == 0.0 and m_g["ev"] == 0.0
    assert m_g["px"] == pytest.approx(p.rho * p.kT, rel=1e-15)


@pytest.mark.parametrize("path", ["clist", "tilelist", "allpairs"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_unscreened_coulomb(path, precision):
    """mu = 0 (plain Coulomb with the energy shift 1/rc): the cluster-list ion
    forms work in lengths scaled by mu, so they take the unscaled FP64 form
    there, also when mixed precision is asked for (FP64 bars); the mixed form
    of the tile lists and the all-pairs kernel keeps the mixed bars."""
    p, x, Z, A = mixture(900, (180.0, 180.0, 180.0), 70.0, seed=12)
    p = p.replace(xmuc=0.0, precision=precision)
    a_o, m_o = oracle.accel(p.replace(precision="fp64"), x, Z, A, True, ref_f1=0)
    a_g, m_g = gpu_accel(p, x, Z, A, True, path)
    rel, cond = force_errors(p, a_g, a_o, x, Z, A)
    print(f"mu = 0 {path} {precision}: max rel {rel:.2e} cond {cond:.2e}")
    if precision == "fp64" or path == "clist":
        assert cond <= 1e-12 and rel <= 1e-10
        check_meas(m_g, m_o)
    else:
        assert cond <= 1e-6 and rel <= 1e-4
        assert m_g["ev"] == pytest.approx(m_o["ev"], rel=1e-6)


def test_bfield_sim_type_has_no_pair_force():
    rng = np.random.Generator(np.random.PCG64(7))
    n, L = 300, 100.0
    x = rng.uniform(0.0, L, size=(n, 3))
    p = SimParams(sim_type="B-field", xl=(L, L, L), rccut=30.0, rho=n / L ** 3, kT=1.0)
    Z, A = np.ones(n), np.ones(n)
    a_o, _ = oracle.accel(p, x, Z, A, True, ref_f1=0)
    with YukawaSystem(p, Z=Z, A=A) as s:
        s.put_state(x)
        s.accel(True)
        a_g = s.a
    assert np.array_equal(a_o, np.zeros((n, 3))) and np.array_equal(a_g, a_o)


def test_out_of_memory_at_create():
    """A system whose tile lists would take 25 TB is refused with
    YK_ERR_NOMEM, and the failed allocation leaves no error behind: a system
    created next computes its forces."""
    p = W.pure_ion(n=1000).params.replace(tile_lists=True)
    with pytest.raises(YkError) as err:
        YukawaSystem(p, n=20_000_000)
    assert err.value.code == 3
    w = W.pure_ion(n=1000, seed=5)
    with YukawaSystem(w.params, Z=w.Z, A=w.A) as s:
        s.put_state(w.x)
        s.accel(False)
        assert np.isfinite(s.a).all() and np.abs(s.a).max() > 0
