"""Reader of the reference's revision-b configuration files ('x4b', 'xv4b',
'x8b', 'xv8b', 'xva8b'; src/rw_xvb.f:148-203): Fortran unformatted sequential records,
big-endian, each framed by 4-byte lengths -- file type; code name and
version; date, time, zone; sim_type; (time, box, ev, ek, px, pp(3,3), n);
the particle record."""
from __future__ import annotations

import numpy as np


def _records(raw: bytes):
    out, o = [], 0
    while o < len(raw):
        ln = int.from_bytes(raw[o:o + 4], "big")
        out.append(raw[o + 4:o + 4 + ln])
        o += 8 + ln
    return out


_COLS = {"x4b": (3, ">f4"), "xv4b": (6, ">f4"), "x8b": (3, ">f8"), "xv8b": (6, ">f8"),
         "xva8b": (9, ">f8")}


def _config(recs) -> dict:
    ftype = recs[0].decode().strip()
    head = recs[4]
    vals = np.frombuffer(head[:16 * 8], dtype=">f8")
    n = int(np.frombuffer(head[16 * 8:16 * 8 + 4], dtype=">i4")[0])
    cols, dt = _COLS[ftype]
    data = np.frombuffer(recs[5], dtype=dt).reshape(n, cols).astype(np.float64)
    out = {"ftype": ftype, "code": recs[1].decode().strip(), "sim_type": recs[3].decode().strip(),
           "time": float(vals[0]), "xl": tuple(float(v) for v in vals[1:4]), "ev": float(vals[4]),
           "ek": float(vals[5]), "px": float(vals[6]), "pp": vals[7:16].reshape(3, 3).T.copy(),
           "n": n, "x": data[:, 0:3].copy()}
    if cols >= 6:
        out["v"] = data[:, 3:6].copy()
    if cols == 9:
        out["a"] = data[:, 6:9].copy()
    return out


def read_xvb(path) -> dict:
    """The first (usually only) configuration of a revision-b file."""
    return _config(_records(open(path, "rb").read())[:6])


def read_xvb_all(path) -> list:
    """Every configuration of a revision-b trajectory (six records each)."""
    recs = _records(open(path, "rb").read())
    return [_config(recs[k:k + 6]) for k in range(0, len(recs), 6)]


__all__ = ["read_xvb", "read_xvb_all"]
