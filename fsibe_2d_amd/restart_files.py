"""Host-side reading of the reference's restart files: Fortran unformatted
sequential records (a 4-byte byte count before and after each record) as
the reference's build (ifort -r8) writes them, and the choice between the
two alternating backup files.

* runtime_backup<n>.in: MakeRuntimeBackup (save_fields.f90:109-152), read
  back by init_fields with inicond = 2 (init_fields.f90:44-97);
* the final vorticity file (time_step.f90:350-355), read by init_fields
  with inicond = 111 (init_fields.f90:109-128).

The library writes and reads these files itself (fsibe_backup_write/read,
fsibe_inicond_write/read); this module parses them into arrays for checks
and post-processing, and picks the file init_fields would resume from.
"""
from __future__ import annotations

import os
import struct

import numpy as np


class FortranRecordError(ValueError):
    pass


def read_records(path) -> list:
    """every record of a Fortran unformatted sequential file, as bytes"""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    o = 0
    while o < len(data):
        if o + 4 > len(data):
            raise FortranRecordError(f"{path}: truncated record marker at byte {o}")
        (n,) = struct.unpack_from("<i", data, o)
        if n < 0 or o + 8 + n > len(data):
            raise FortranRecordError(f"{path}: bad record length {n} at byte {o}")
        rec = data[o + 4:o + 4 + n]
        (n2,) = struct.unpack_from("<i", data, o + 4 + n)
        if n2 != n:
            raise FortranRecordError(f"{path}: record markers {n} and {n2} differ at byte {o}")
        out.append(rec)
        o += 8 + n
    return out


def write_records(path, records) -> None:
    """records: a sequence of bytes-like objects, one per record"""
    with open(path, "wb") as f:
        for r in records:
            b = bytes(r)
            f.write(struct.pack("<i", len(b)))
            f.write(b)
            f.write(struct.pack("<i", len(b)))


class _Cursor:
    def __init__(self, rec: bytes, what: str):
        self.b, self.o, self.what = rec, 0, what

    def take(self, fmt_or_count, dtype=None):
        if dtype is None:
            n = struct.calcsize("<" + fmt_or_count)
            if self.o + n > len(self.b):
                raise FortranRecordError(f"{self.what}: record too short")
            v = struct.unpack_from("<" + fmt_or_count, self.b, self.o)
            self.o += n
            return v if len(v) > 1 else v[0]
        cnt = int(fmt_or_count)
        n = cnt * np.dtype(dtype).itemsize
        if self.o + n > len(self.b):
            raise FortranRecordError(f"{self.what}: record too short")
        a = np.frombuffer(self.b, dtype=dtype, count=cnt, offset=self.o).copy()
        self.o += n
        return a

    def end(self):
        if self.o != len(self.b):
            raise FortranRecordError(f"{self.what}: {len(self.b) - self.o} bytes left in the record")


def read_backup(path, nx: int, ny: int, ns: int, nbeams: int = 1) -> dict:
    """runtime_backup<n>.in as arrays: fields shaped (level, ny, nx) or
    (ny, nx) in the reference's memory order; logicals as bools (odd or
    nonzero is .true.)"""
    recs = read_records(path)
    N = nx * ny
    r = iter(recs)

    def nxt(what):
        try:
            return _Cursor(next(r), what)
        except StopIteration:
            raise FortranRecordError(f"{path}: missing record {what}") from None

    d = {}
    c = nxt("time"); d["time"] = c.take("d"); c.end()
    c = nxt("checksum"); d["checksum"] = c.take("d"); c.end()
    c = nxt("fields")
    d["n1"] = c.take("i")
    d["dt1"] = c.take("d")
    for k in ("vortk", "nlk", "workvis"):
        d[k] = c.take(2 * N, "<f8").reshape(2, ny, nx)
    for k in ("mask", "maskvx", "maskvy"):
        d[k] = c.take(N, "<f8").reshape(ny, nx)
    d["ivideo"] = c.take("i")
    d["u"] = c.take(2 * N, "<f8").reshape(2, ny, nx)
    c.end()
    c = nxt("ns"); d["ns"] = c.take("i"); c.end()
    if d["ns"] != ns:
        raise FortranRecordError(f"{path}: ns {d['ns']} in the file, {ns} expected")
    c = nxt("colorscale"); d["colorscale"] = c.take("d"); d["colorscale_done"] = c.take("i") != 0; c.end()
    beams = []
    for i in range(nbeams):
        b = {}
        c = nxt(f"beam {i} positions")
        for k in ("x", "y", "vx", "vy", "theta", "theta_dot"):
            b[k] = c.take(ns, "<f8")
        c.end()
        c = nxt(f"beam {i} loads")
        for k in ("pressure_new", "pressure_old", "tau_new", "tau_old"):
            b[k] = c.take(ns, "<f8")
        c.end()
        c = nxt(f"beam {i} forces")
        for k in ("Force", "Force_unst", "Force_press"):
            b[k] = c.take(2, "<f8")
        for k in ("E_kinetic", "E_pot", "E_elastic", "x0", "y0"):
            b[k] = c.take("d")
        c.end()
        c = nxt(f"beam {i} motion")
        b["AngleBeam"] = c.take("d")
        b["iMouvement"] = c.take("i")
        for k in ("drag_unst_new", "drag_unst_old", "lift_unst_new", "lift_unst_old"):
            b[k] = c.take("d")
        c.end()
        c = nxt(f"beam {i} solver")
        b["UnsteadyCorrectionsReady"] = c.take("i") != 0
        b["dt_old"] = c.take("d")
        b["beam_oldold"] = c.take(6 * ns, "<f8").reshape(6, ns)
        c.end()
        c = nxt(f"beam {i} accelerations")
        b["ax"] = c.take(ns, "<f8")
        b["ay"] = c.take(ns, "<f8")
        b["Inertial_Force"] = c.take(2, "<f8")
        c.end()
        beams.append(b)
    d["beams"] = beams
    c = nxt("vor_mean"); d["vor_mean"] = c.take(N, "<f8").reshape(ny, nx); c.end()
    if next(r, None) is not None:
        raise FortranRecordError(f"{path}: records after vor_mean")
    return d


def read_inicond(path) -> tuple:
    """(nx, ny, vorticity (ny, nx)) of the final vorticity file"""
    recs = read_records(path)
    if len(recs) != 1:
        raise FortranRecordError(f"{path}: {len(recs)} records, 1 expected")
    c = _Cursor(recs[0], path)
    nx, ny = c.take("i"), c.take("i")
    v = c.take(nx * ny, "<f8").reshape(ny, nx)
    c.end()
    return nx, ny, v


def latest_backup(directory) -> str | None:
    """the backup init_fields resumes from (init_fields.f90:45-52): of
    runtime_backup0.in and runtime_backup1.in, the one whose first record
    holds the later time (the first on a tie); None if neither is readable"""
    best, best_t = None, 0.0
    for i in (0, 1):
        p = os.path.join(directory, f"runtime_backup{i}.in")
        try:
            with open(p, "rb") as f:
                head = f.read(12)
            (n,) = struct.unpack_from("<i", head, 0)
            if n != 8:
                continue
            (t,) = struct.unpack_from("<d", head, 4)
        except (OSError, struct.error):
            continue
        if t > best_t:
            best, best_t = p, t
    return best
