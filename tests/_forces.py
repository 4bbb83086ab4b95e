"""Force-comparison helpers shared by the GPU parity tests.

The conditioning of particle i's force is sum_j |F_ij| (oracle.pair_abs_sum):
summation-order rounding changes F_i by a few ulps of that sum, so FP64 results
are compared by the condition-aware error |dF_i| / sum_j |F_ij| as well as the
plain |dF_i| / |F_i|.
"""
import numpy as np
import pytest

import oracle
from yukawamd_amd.system import YukawaSystem


def masses(p, A):
    return (A if p.sim_type != "nucleon" else np.ones_like(A)) * p.xmass


def abs_pair_force_sum_np(p, x, Z):
    """sum_j |F_ij| per particle in numpy (a second restatement of
    oracle.pair_abs_sum, for small systems)."""
    n = x.shape[0]
    L = np.array(p.xl)
    out = np.zeros(n)
    for i0 in range(0, n, 256):
        xi = x[i0:i0 + 256, None, :]
        d = xi - x[None, :, :]
        d = np.where(d > 0.5 * L, d - L, d)
        d = np.where(d < -0.5 * L, d + L, d)
        r2 = (d * d).sum(-1)
        idx = np.arange(i0, min(i0 + 256, n))
        r2[np.arange(len(idx)), idx] = 1e300     # self pair: beyond any cutoff, no NaN
        r = np.sqrt(r2)
        with np.errstate(over="ignore", invalid="ignore"):
            if p.sim_type == "nucleon":
                g = np.exp(-r2 / (2 * p.xpacket))
                same = Z[i0:i0 + 256, None] == Z[None, :]
                bc = np.where(same, p.bb + p.cc, p.bb - p.cc)
                s = np.where(r2 <= p.rncut ** 2, np.abs((2 * p.aa * g + bc) * g / p.xpacket), 0)
                pp = (Z[i0:i0 + 256, None] == 1) & (Z[None, :] == 1) & (r2 <= p.rccut ** 2)
                s += np.where(pp, p.vc * (1 + p.xmuc * r) * np.exp(-p.xmuc * r) / (r * r2), 0)
            else:
                zz = Z[i0:i0 + 256, None] * Z[None, :] if p.sim_type == "ion-mixture" \
                    else p.zi ** 2
                s = np.where(r2 <= p.rccut ** 2,
                             np.abs(p.frp * p.vc * zz) * (1 + p.xmuc * r) * np.exp(-p.xmuc * r)
                             / (r * r2), 0)
        out[i0:i0 + 256] = (s * r).sum(-1)
    return out


def gpu_accel(p, x, Z, A, do_meas=True, path="clist"):
    """path: 'clist' (cluster lists, the default for N <= 262,144), 'tilelist'
    (64-particle tile lists only) or 'allpairs' (culled tile pairs)."""
    p = p.replace(allpairs=path == "allpairs", tile_lists=path == "tilelist")
    with YukawaSystem(p, Z=Z, A=A) as s:
        s.put_state(x)
        m = s.accel(do_meas)
        return s.a, m


def force_error_arrays(p, a_gpu, a_ref, x, Z, A, cond=None):
    """Per particle: (|dF_i| / |F_i|, |dF_i| / sum_j |F_ij|, |F_i| / sum_j |F_ij|)."""
    m = masses(p, A)[:, None]
    dF = np.linalg.norm((a_gpu - a_ref) * m, axis=1)
    F = np.linalg.norm(a_ref * m, axis=1)
    if cond is None:
        cond = oracle.pair_abs_sum(p, x, Z)[0]
    ok = cond > 0
    rel = np.where(ok, dF / np.where(F > 0, F, 1.0), 0.0)
    rc = np.where(ok, dF / np.where(ok, cond, 1.0), 0.0)
    ratio = np.where(ok, F / np.where(ok, cond, 1.0), 1.0)
    return rel, rc, ratio


def force_errors(p, a_gpu, a_ref, x, Z, A, cond=None):
    rel, rc, _ = force_error_arrays(p, a_gpu, a_ref, x, Z, A, cond)
    return rel.max(), rc.max()


def check_meas(m_gpu, m_ref, rtol=1e-13):
    assert m_gpu["ev"] == pytest.approx(m_ref["ev"], rel=rtol)
    assert m_gpu["px"] == pytest.approx(m_ref["px"], rel=rtol)
    diag = max(abs(m_ref["pp11"]), abs(m_ref["pp22"]), abs(m_ref["pp33"]))
    for k_gpu, k_ref in zip(range(6), ("pp11", "pp12", "pp13", "pp22", "pp23", "pp33")):
        tol = rtol * abs(m_ref[k_ref]) if k_ref in ("pp11", "pp22", "pp33") else rtol * diag
        assert abs(m_gpu["pp"][k_gpu] - m_ref[k_ref]) <= tol, k_ref
