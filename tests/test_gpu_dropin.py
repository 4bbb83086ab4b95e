"""Drop-in test through the reference's own Fortran host (needs a GPU).

oracle/_ref/md_omp is the reference program built from its sources;
oracle/_ref/md_gpu is the same program (md.F, input1, initconfig, gen_utils,
md_utils, ...) linked with yukawamd_amd/fortran/md_yk_shim.f90 in place of
int_control.o and vel_verlet_2.o, so accel/newton run on the GPU through the
C ABI.  Both read the same runmd.in (random start from the reference's own
RNG), and write the reference's .log and binary .meas files.

Checks: the initial potential energy, kinetic energy, total energy and
pressure agree to the 10 digits the log prints; the energy series of the
first steps agrees to 1e-9 relative; after the reference's own thermostatted
warm-up, 1,000 NVE steps follow the reference's energy series record by
record (1e-10) and the relative energy drift matches the reference's; the
mixed-precision kernel's drift over the same run stays below 1e-5.
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = ROOT / "oracle" / "_ref"
MD_OMP = REF / "md_omp"
MD_GPU = REF / "md_gpu"

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not (MD_OMP.exists() and MD_GPU.exists()),
                                 reason="drop-in executables not built (make -C oracle dropin)")]

RUNMD = """&run_parameters
 sim_type='ion', start='random', tstart=0.0, dt={dt},
 nwgroup=0, nwsteps=0, ngroup=1, ntot={ntot}, nind={nind},
 {species}
 rho=7.18e-5, kT={kT},
 coulomb='screened-coulomb', xlambda={lam}, rccut={rc},
 irnd=1, iseed={seed},
/
"""


def pure_runmd(n=3456, ntot=10, nind=1, seed=1234):
    # the pure-ion workload: Z=8 A=16, kappa=1, Gamma=175, rc = 12 lambda
    return RUNMD.format(dt=21.1627, ntot=ntot, nind=nind, kT=0.0352829, lam=14.9255,
                        rc=179.106, seed=seed,
                        species=f"spec_list(1)%n={n}, spec_list(1)%z=8.0, spec_list(1)%a=16.0,")


def mix_runmd(ntot=10, nind=1, seed=777):
    sp = ("spec_list(1)%n=1200, spec_list(1)%z=8.0, spec_list(1)%a=16.0,\n"
          " spec_list(2)%n=600, spec_list(2)%z=26.0, spec_list(2)%a=56.0,\n"
          " spec_list(3)%n=200, spec_list(3)%z=34.0, spec_list(3)%a=80.0,")
    return RUNMD.format(dt=13.1422, ntot=ntot, nind=nind, kT=0.0617451, lam=21.3221,
                        rc=200.0, seed=seed, species=sp)


def run(exe, runmd, workdir, threads=8, env_extra=None):
    workdir.mkdir(parents=True, exist_ok=True)
    (workdir / "runmd.in").write_text(runmd)
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    env.update(env_extra or {})
    out = subprocess.run([str(exe)], cwd=workdir, env=env, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]
    log = next(workdir.glob("md.*.log")).read_text()
    meas = read_meas(next(workdir.glob("md.*.meas")))
    return log, meas


def read_meas(path):
    """Binary .meas records (src/md.F:151): istep, time, ev, ek, px, pp x6."""
    rec = np.dtype([("m0", ">i4"), ("istep", ">i4"), ("v", ">f8", (10,)), ("m1", ">i4")])
    data = np.fromfile(path, dtype=rec)
    assert (data["m0"] == 84).all() and (data["m1"] == 84).all()
    return data["istep"], data["v"]


def initial_line(log):
    line = next(l for l in log.splitlines() if "Initial Vtot" in l)
    return np.array([float(t) for t in line.split("=")[1].split()])


@pytest.mark.parametrize("name,runmd", [("pure_ion_3456", pure_runmd()),
                                        ("ion_mixture_2000", mix_runmd())],
                         ids=["pure_ion_3456", "ion_mixture_2000"])
def test_dropin_matches_reference_host(tmp_path, name, runmd):
    log_c, (st_c, m_c) = run(MD_OMP, runmd, tmp_path / "cpu", threads=8)
    log_g, (st_g, m_g) = run(MD_GPU, runmd, tmp_path / "gpu", threads=1)
    i_c, i_g = initial_line(log_c), initial_line(log_g)
    print(name, "initial", i_c, i_g)
    # the log prints 10 significant digits
    np.testing.assert_allclose(i_g, i_c, rtol=2e-9)
    assert (st_c == st_g).all()
    # time, ev, ek, px to 1e-9 over the first 10 steps
    np.testing.assert_allclose(m_g[:, :4], m_c[:, :4], rtol=1e-9)


def warm_runmd(seed=99, nwgroup=10, nwsteps=50, ntot=100, nind=10):
    """Pure-ion workload with the reference's own warm-up (nwgroup groups of
    nwsteps steps, cancel_vcom + tnorm after each, src/md.F:88-101) before
    ntot*nind measured NVE steps."""
    r = pure_runmd(n=3456, ntot=ntot, nind=nind, seed=seed)
    return r.replace("nwgroup=0, nwsteps=0", f"nwgroup={nwgroup}, nwsteps={nwsteps}")


def drift(m):
    e = m[:, 1] + m[:, 2]
    return e, np.abs(e - e[0]).max() / abs(e[0])


def test_dropin_energy_drift_1000_steps(tmp_path):
    """1,000 NVE steps after a thermostatted warm-up: the GPU run follows the
    reference's energy series record by record and its relative drift matches."""
    runmd = warm_runmd()
    _, (_, m_c) = run(MD_OMP, runmd, tmp_path / "cpu", threads=16)
    _, (_, m_g) = run(MD_GPU, runmd, tmp_path / "gpu", threads=1)
    e_c, d_c = drift(m_c)
    e_g, d_g = drift(m_g)
    print(f"relative drift over 1000 steps: reference {d_c:.3e}  gpu {d_g:.3e}")
    np.testing.assert_allclose(e_g, e_c, rtol=1e-10)
    assert d_g == pytest.approx(d_c, rel=0.05)


def test_dropin_mixed_precision_drift(tmp_path):
    """The mixed-precision kernel through the same host: relative energy drift
    per 1,000 steps at most 1e-5, and within a few times the FP64 drift."""
    runmd = warm_runmd()
    _, (_, m_d) = run(MD_GPU, runmd, tmp_path / "fp64", threads=1)
    _, (_, m_m) = run(MD_GPU, runmd, tmp_path / "mixed", threads=1,
                      env_extra={"YK_PRECISION": "mixed"})
    _, d_d = drift(m_d)
    e_m, d_m = drift(m_m)
    print(f"relative drift over 1000 steps: fp64 {d_d:.3e}  mixed {d_m:.3e}")
    assert d_m <= 1e-5
