"""The run driver's initial state against the reference program (CPU).

yukawamd_amd/fortran/ykmd (our Fortran driver, yk_host.f90) and
oracle/_ref/md_omp (the reference built from its sources, one OpenMP
thread) read the same runmd.in with start='random' and the default
generator (MT19937 seeded as src/randnums.f:239-256).  The initial
configuration each writes (md.<tstart>.xv8b, src/initconfig.f:385-390) must
agree bit for bit: the box from n, rho and the aspect ratios, the
positions, and the Maxwell velocities after cancel_vcom and tnorm.
`ykmd --init-only` stops there, so no GPU is needed.
"""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

from yukawamd_amd.xvb import read_xvb

ROOT = Path(__file__).resolve().parent.parent
MD_OMP = ROOT / "oracle" / "_ref" / "md_omp"
YKMD = ROOT / "yukawamd_amd" / "fortran" / "ykmd"

pytestmark = pytest.mark.skipif(not (MD_OMP.exists() and YKMD.exists()),
                                reason="reference program or driver not built")

MIX = ("spec_list(1)%n=600, spec_list(1)%z=8.0, spec_list(1)%a=16.0,\n"
       " spec_list(2)%n=300, spec_list(2)%z=26.0, spec_list(2)%a=56.0,\n"
       " spec_list(3)%n=100, spec_list(3)%z=34.0, spec_list(3)%a=80.0,")
PURE = "spec_list(1)%n=1024, spec_list(1)%z=8.0, spec_list(1)%a=16.0,"


def runmd(species, seed, extra=""):
    return f"""&run_parameters
 sim_type='ion', start='random', tstart=0.0, dt=13.1422,
 nwgroup=0, nwsteps=0, ngroup=0, ntot=0, nind=0,
 {species}
 rho=7.18e-5, kT=0.0617451,
 coulomb='screened-coulomb', rccut=150.0,
 iseed={seed}, {extra}
/
"""


def run_only(cmd, text, workdir):
    workdir.mkdir(parents=True, exist_ok=True)
    (workdir / "runmd.in").write_text(text)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    out = subprocess.run(cmd, cwd=workdir, env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]


def run(cmd, text, workdir):
    run_only(cmd, text, workdir)
    return read_xvb(next(workdir.glob("md.*.xv8b")))


@pytest.mark.parametrize("species,seed,extra", [(MIX, 4321, ""), (PURE, 17, "aspect=1.0,1.0,2.0,"),
                                                (MIX, 99991, "irnd=4,")])
def test_random_start_bitexact(tmp_path, species, seed, extra):
    text = runmd(species, seed, extra)
    ref = run([str(MD_OMP)], text, tmp_path / "ref")
    ours = run([str(YKMD), "--init-only"], text, tmp_path / "ours")
    assert ours["n"] == ref["n"] and ours["xl"] == ref["xl"]
    assert np.array_equal(ours["x"], ref["x"])
    assert np.array_equal(ours["v"], ref["v"])
    # what the mathematics fixes: positions in the box, zero momentum, 1.5 kT
    assert (ref["x"] >= 0).all() and (ref["x"] <= np.array(ref["xl"])).all()


@pytest.mark.parametrize("seed", [5489, 4357, 123456789])
def test_generator_against_numpy_mt19937(tmp_path, seed):
    """An independent MT19937: numpy's bit generator with its state set to the
    reference's seeding (word i = seed * 69069^i mod 2^32, src/randnums.f:
    250-253) yields the raw 32-bit outputs; the driver's first 3n uniforms
    are the positions x(k,i) = xl(k) * y / (2^32 - 1), so rounding
    x / xl * (2^32 - 1) recovers each y exactly (the error is ~1e-6)."""
    n = 1024
    ours = run([str(YKMD), "--init-only"], runmd(PURE, seed), tmp_path / "ours")
    key = np.zeros(624, dtype=np.uint32)
    key[0] = seed & 0xFFFFFFFF
    for i in range(1, 624):
        key[i] = (69069 * int(key[i - 1])) & 0xFFFFFFFF
    bg = np.random.MT19937()
    bg.state = {"bit_generator": "MT19937", "state": {"key": key, "pos": 624}}
    raw = bg.random_raw(3 * n).astype(np.float64)
    y = ours["x"] / np.array(ours["xl"]) * 4294967295.0
    assert np.abs(y.ravel() - np.round(y.ravel())).max() < 1e-3
    assert np.array_equal(np.round(y.ravel()), raw)


def echo(workdir):
    """The log's parameter echo, 'SIMULATION TYPE' to the 'nckpt' line."""
    lines = next(workdir.glob("md.*.log")).read_text().splitlines()
    i = next(k for k, ln in enumerate(lines) if "SIMULATION TYPE" in ln)
    j = next(k for k, ln in enumerate(lines) if ln.startswith("  nckpt:"))
    return lines[i:j + 1]


NUCLEON = """&run_parameters
 sim_type='nucleon', start='random', tstart=0.0, dt=2.0,
 nwgroup=0, nwsteps=0, ngroup=0, ntot=0, nind=0, nptensor=0, ngofr=1,
 spec_list(1)%n=200, spec_list(1)%z=0.0, spec_list(1)%a=1.0,
 spec_list(2)%n=100, spec_list(2)%z=1.0, spec_list(2)%a=1.0,
 rho=0.05, kT=1.0, xmass=939.0, rccut=20.0,
 coulomb='screened-coulomb', xlambda=10.0, nuclear='HPP',
 iseed=77, irnd=4, nout=7, ftype='xv4b', append=.true., nckpt=11,
/
"""


@pytest.mark.parametrize("case", ["mixture_thomas_fermi_strain", "nucleon_hpp"])
def test_log_parameter_echo_matches_reference(tmp_path, case):
    """The parameter echo of src/input1.F / src/initconfig.f, line for line:
    Thomas-Fermi screening (no xlambda) with strain rates on two axes and a
    B field, or the nucleon model with the automatic HPP cutoff, g(r)
    requested (disabled), the MT19937 line and the output options."""
    if case == "nucleon_hpp":
        text = NUCLEON
    else:
        text = runmd(MIX, 2468, "deps(1)=1.0e-6, deps(3)=-2.0e-6, bfield=1.0e13, "
                                "nout=3, ftype='x8b', nckpt=9,")
    run([str(MD_OMP)], text, tmp_path / "ref")
    run([str(YKMD), "--init-only"], text, tmp_path / "ours")
    assert echo(tmp_path / "ours") == echo(tmp_path / "ref")


def test_xv8b_start_echo_and_box_match_reference(tmp_path):
    """A start from an xv8b file: the density is recomputed from the file's
    box and n, and with no aspect given the gage box is the file's box
    (src/initconfig.f:295-308); the echo names the file."""
    run([str(MD_OMP)], runmd(MIX, 1357), tmp_path / "gen")
    start = next((tmp_path / "gen").glob("md.*.xv8b"))
    text = runmd(MIX, 1357).replace("start='random', tstart=0.0,", "start='start.xv8b',") \
                           .replace("rho=7.18e-5,", "rho=5.0e-5,")
    for d in ("ref", "ours"):
        (tmp_path / d).mkdir()
        (tmp_path / d / "start.xv8b").write_bytes(start.read_bytes())
    run_only([str(MD_OMP)], text, tmp_path / "ref")
    run_only([str(YKMD), "--init-only"], text, tmp_path / "ours")
    assert echo(tmp_path / "ours") == echo(tmp_path / "ref")


NUCLEUS = """&run_parameters
 sim_type='nucleon', start='nuclear', tstart=0.0, dt=2.0,
 nwgroup=0, nwsteps=0, ngroup=0, ntot=0, nind=0,
 spec_list(1)%n=126, spec_list(1)%z=0.0, spec_list(1)%a=1.0,
 spec_list(2)%n=82, spec_list(2)%z=1.0, spec_list(2)%a=1.0,
 rho=0.01, kT=1.0, xmass=939.0, rccut=20.0, rmax={rmax},
 coulomb='screened-coulomb', xlambda=10.0, nuclear='HPP',
 iseed={seed},
/
"""


@pytest.mark.parametrize("rmax,seed", [(7.5, 208), (1.0e3, 31)])
def test_nuclear_start_bitexact(tmp_path, rmax, seed):
    """start='nuclear' (src/initconfig.f:227-253): positions uniform in a
    sphere of radius min(rmax, half the shortest edge) about the box centre,
    three draws per particle; the initial configuration bit for bit and the
    echo (with the radius line) line for line.  rmax = 1000 fm is clamped to
    half the box."""
    text = NUCLEUS.format(rmax=rmax, seed=seed)
    ref = run([str(MD_OMP)], text, tmp_path / "ref")
    ours = run([str(YKMD), "--init-only"], text, tmp_path / "ours")
    assert ours["n"] == ref["n"] == 208 and ours["xl"] == ref["xl"]
    assert np.array_equal(ours["x"], ref["x"])
    assert np.array_equal(ours["v"], ref["v"])
    assert echo(tmp_path / "ours") == echo(tmp_path / "ref")
    # the mathematics: every particle within the radius of the box centre
    c = 0.5 * np.array(ref["xl"])
    r = np.linalg.norm(ref["x"] - c, axis=1)
    assert r.max() <= min(rmax, 0.5 * min(ref["xl"])) * (1 + 1e-12)
