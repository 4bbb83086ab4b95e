"""The run driver's species files and start-file types against the reference
program (CPU).

Each case writes a start (and species) file from a state the reference
generated, gives both programs -- yukawamd_amd/fortran/ykmd --init-only and
oracle/_ref/md_omp (the reference built from its sources, one OpenMP
thread) -- the same runmd.in, and compares the initial configuration each
writes (md.<tstart>.xv8b, src/initconfig.f:385-390) bit for bit, header
included, and the log's parameter echo line for line.  Covered: species
files 'ZAfa' and 'ZAfb' (src/rw_xvfa.f:84-98, src/rw_xvfb.f:11-62); the start
types of src/initconfig.f:254-345 -- formatted revision a 'onp', 'md1',
'xfa', 'xvfa', 'xvZAfa', binary revision a 'md0', 'x4a', 'xv4a', 'x8a',
'xv8a', binary revision b 'x4b', 'x8b', 'xv4b', formatted revision b 'xfb',
'xvfb', 'xZAfb', 'xvZAfb'; Maxwell velocities drawn when the file has none.
The files are written here in the layouts the reference's readers take.
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

SPECIES = [(60, 8.0, 16.0), (30, 26.0, 56.0), (10, 34.0, 80.0)]
N = sum(k for k, _, _ in SPECIES)
SPEC_LIST = ", ".join(f"spec_list({i + 1})%n={k}, spec_list({i + 1})%z={z}, "
                      f"spec_list({i + 1})%a={a}" for i, (k, z, a) in enumerate(SPECIES))


def runmd(start, species=SPEC_LIST, extra=""):
    return f"""&run_parameters
 sim_type='ion', start='{start}', dt=13.1422,
 nwgroup=0, nwsteps=0, ngroup=0, ntot=0, nind=0,
 {species}
 rho=7.18e-5, kT=0.0617451,
 coulomb='screened-coulomb', xlambda=40.0, rccut=150.0,
 iseed=2718, {extra}
/
"""


def run_in(cmd, text, workdir, files):
    workdir.mkdir(parents=True, exist_ok=True)
    for name, data in files.items():
        (workdir / name).write_bytes(data)
    (workdir / "runmd.in").write_text(text)
    env = dict(os.environ, OMP_NUM_THREADS="1")
    out = subprocess.run(cmd, cwd=workdir, env=env, capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout[-2000:] + out.stderr[-2000:]


def echo(workdir):
    lines = next(workdir.glob("md.*.log")).read_text().splitlines()
    i = next(k for k, ln in enumerate(lines) if "SIMULATION TYPE" in ln)
    j = next(k for k, ln in enumerate(lines) if ln.startswith("  nckpt:"))
    return lines[i:j + 1]


def both(tmp_path, text, files):
    run_in([str(MD_OMP)], text, tmp_path / "ref", files)
    run_in([str(YKMD), "--init-only"], text, tmp_path / "ours", files)
    ref = read_xvb(next((tmp_path / "ref").glob("md.*.xv8b")))
    ours = read_xvb(next((tmp_path / "ours").glob("md.*.xv8b")))
    assert echo(tmp_path / "ours") == echo(tmp_path / "ref")
    for key in ("n", "xl", "time", "ev", "ek", "px", "sim_type"):
        assert ours[key] == ref[key], key
    assert np.array_equal(ours["pp"], ref["pp"])
    assert np.array_equal(ours["x"], ref["x"])
    assert np.array_equal(ours["v"], ref["v"])
    return ref, ours


@pytest.fixture(scope="module")
def state(tmp_path_factory):
    """A state of the reference's own random start: positions, velocities,
    box and time (tstart = 250.5 fm/c)."""
    d = tmp_path_factory.mktemp("gen")
    run_in([str(MD_OMP)], runmd("random", extra="tstart=250.5,"), d, {})
    s = read_xvb(next(d.glob("md.*.xv8b")))
    z = np.concatenate([np.full(k, zz) for k, zz, _ in SPECIES])
    a = np.concatenate([np.full(k, aa) for k, _, aa in SPECIES])
    return dict(s, z=z, a=a)


# ---- writers of the layouts the reference's readers take --------------------
def num(v):
    return repr(float(v))


def rows(*cols):
    return "".join(" ".join(num(c) for c in row) + "\n" for row in zip(*cols))


def record(b: bytes) -> bytes:
    ln = len(b).to_bytes(4, "big")
    return ln + b + ln


def rev_a_binary(ftype, s):
    x, v, t = s["x"], s["v"], s["time"]
    if ftype == "md0":
        return record(np.hstack([x, v]).astype(">f8").tobytes())
    f4 = ftype.endswith("4a")
    tb = np.array([t], dtype=">f4" if f4 else ">f8").tobytes()
    data = np.hstack([x, v]) if ftype.startswith("xv") else x
    return record(tb) + record(data.astype(">f4" if f4 else ">f8").tobytes())


def rev_b_binary(ftype, s):
    head = np.concatenate([[s["time"]], s["xl"], [s["ev"], s["ek"], s["px"]],
                           s["pp"].T.ravel()]).astype(">f8").tobytes()
    head += np.array([s["n"]], dtype=">i4").tobytes()
    f4 = "4" in ftype
    data = np.hstack([s["x"], s["v"]]) if ftype.startswith("xv") else s["x"]
    return (record(ftype.ljust(6).encode()) + record(b"MD".ljust(10) + b"6.2.0".ljust(8))
            + record(b"20261204" + b"120000.000" + b"+0000")
            + record(b"ion-mixture".ljust(20)) + record(head)
            + record(data.astype(">f4" if f4 else ">f8").tobytes()))


def rev_a_formatted(ftype, s):
    x, v, t = s["x"], s["v"], s["time"]
    c = [x[:, 0], x[:, 1], x[:, 2]]
    cv = c + [v[:, 0], v[:, 1], v[:, 2]]
    if ftype == "onp":
        return rows(*c, np.zeros(len(x))).encode()
    if ftype == "md1":
        return rows(*cv).encode()
    if ftype == "xfa":
        return (num(t) + "\n" + rows(*c)).encode()
    if ftype == "xvfa":
        return (num(t) + "\n" + rows(*cv)).encode()
    # xvZAfa: comments, the count (i10), the time, then x v Z A
    return ("# a comment\n# another\n" + f"{len(x):10d}\n" + num(t) + "\n"
            + rows(*cv, s["z"], s["a"])).encode()


def rev_b_formatted(ftype, s):
    x, v = s["x"], s["v"]
    head = (f"{ftype}\nMD 6.2.0\n20261204 120000.000 +0000\nion-mixture\n"
            + " ".join(num(q) for q in s["xl"]) + "\n"
            + " ".join(num(q) for q in [s["time"], *s["xl"], s["ev"], s["ek"], s["px"],
                                        *s["pp"].T.ravel()]) + f" {s['n']}\n")
    c = [x[:, 0], x[:, 1], x[:, 2]]
    cv = c + [v[:, 0], v[:, 1], v[:, 2]]
    body = {"xfb": rows(*c), "xvfb": rows(*cv), "xZAfb": rows(*c, s["z"], s["a"]),
            "xvZAfb": rows(*cv, s["z"], s["a"])}[ftype]
    return (head + body).encode()


# ---- species files -----------------------------------------------------------
def zafb():
    return ("# species\n" + f"  {N}\n"
            + "".join(f"{k} {z} {a}\n" for k, z, a in SPECIES)).encode()


def zafa():
    lines = ["# species, one line per particle", f"{N:10d}"]
    i = 1
    for k, z, a in SPECIES:
        for _ in range(k):
            lines.append(f"{i} {z} {a}")
            i += 1
    return ("\n".join(lines) + "\n").encode()


@pytest.mark.parametrize("kind", ["ZAfb", "ZAfa"])
def test_species_file_random_start(tmp_path, kind):
    """Charges and masses from a species file (no spec_list): the species
    echo lines, n from the file, the random start bit for bit (the initial
    file is written before 'ion' resolves, src/input1.F:186-198)."""
    data = zafb() if kind == "ZAfb" else zafa()
    text = runmd("random", species=f"spec_file='species.{kind}',", extra="tstart=0.0,")
    ref, _ = both(tmp_path, text, {f"species.{kind}": data})
    assert ref["n"] == N and ref["sim_type"] == "ion"


def test_species_file_thomas_fermi(tmp_path):
    """'ZAfb' sums the total charge, so Thomas-Fermi screening (no xlambda)
    follows the file's species."""
    text = runmd("random", species="spec_file='species.ZAfb',", extra="tstart=0.0,") \
        .replace("xlambda=40.0, ", "")
    both(tmp_path, text, {"species.ZAfb": zafb()})


def test_spec_list_wins_over_spec_file(tmp_path):
    """With both given, the spec_list is used (src/input1.F:136-137)."""
    text = runmd("random", species=SPEC_LIST + ", spec_file='missing.ZAfb',",
                 extra="tstart=0.0,")
    both(tmp_path, text, {})


STARTS = ["onp", "md1", "xfa", "xvfa", "md0", "x4a", "xv4a", "x8a", "xv8a",
          "x4b", "x8b", "xv4b", "xfb", "xvfb"]


@pytest.mark.parametrize("ftype", STARTS)
def test_start_file_type(tmp_path, state, ftype):
    """Start from a file of type ftype with the species of the spec_list:
    the time from the file where it has one (tstart not given), the box from
    n and rho (revision a) or from the file (revision b), velocities drawn
    from the Maxwell distribution where the file has none."""
    if ftype in ("onp", "md1", "xfa", "xvfa"):
        data = rev_a_formatted(ftype, state)
    elif ftype in ("md0", "x4a", "xv4a", "x8a", "xv8a"):
        data = rev_a_binary(ftype, state)
    elif ftype in ("x4b", "x8b", "xv4b"):
        data = rev_b_binary(ftype, state)
    else:
        data = rev_b_formatted(ftype, state)
    extra = "tstart=75.0," if ftype in ("onp", "md1", "md0") else ""
    ref, ours = both(tmp_path, runmd(f"start.{ftype}", extra=extra), {f"start.{ftype}": data})
    assert ref["n"] == N
    if ftype.startswith("xv") or ftype in ("md1", "md0"):
        # velocities taken from the file (rounded to float for the 4-byte types)
        tol = 1e-6 if "4" in ftype else 0.0
        np.testing.assert_allclose(ref["v"], state["v"], rtol=tol, atol=0)


@pytest.mark.parametrize("ftype", ["xvZAfa", "xZAfb", "xvZAfb"])
def test_start_file_with_charges(tmp_path, state, ftype):
    """Start files that carry the charges and masses (no spec_list): n from
    the file, the 'This file also includes charges and masses' echo."""
    data = rev_a_formatted(ftype, state) if ftype == "xvZAfa" else rev_b_formatted(ftype, state)
    ref, _ = both(tmp_path, runmd(f"start.{ftype}", species=""), {f"start.{ftype}": data})
    assert ref["n"] == N
    log = next((tmp_path / "ours").glob("md.*.log")).read_text()
    assert "This file also includes charges and masses." in log


def test_refused_start_types(tmp_path, state):
    """'xva8a' (whose reference reader takes the particles from the time
    record) and an unknown extension stop the driver with a message."""
    for name in ("start.xva8a", "start.bogus"):
        d = tmp_path / name
        d.mkdir()
        (d / name).write_bytes(b"\0" * 64)
        (d / "runmd.in").write_text(runmd(name))
        out = subprocess.run([str(YKMD), "--init-only"], cwd=d, capture_output=True,
                             text=True, timeout=60)
        assert out.returncode != 0 and "not supported" in out.stderr
