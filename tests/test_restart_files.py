"""The restart-file layout on the host (no GPU): Fortran unformatted
records, the runtime_backup record list of MakeRuntimeBackup
(save_fields.f90:118-146) as parsed by fsibe_2d_amd.restart_files, the same
file read by a Fortran program with init_fields' read list
(init_fields.f90:55-94), and the choice of the newer backup
(init_fields.f90:45-52)."""
import os
import struct
import subprocess

import numpy as np
import pytest

from fsibe_2d_amd import restart_files as rf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
READER = os.path.join(ROOT, "tests", "fortran", "build", "read_backup")


def synthetic_backup(path, nx=8, ny=6, ns=5, time=0.375, n1=1, seed=0):
    rng = np.random.default_rng(seed)
    N = nx * ny
    f = {k: rng.standard_normal(2 * N) for k in ("vortk", "nlk", "workvis", "u")}
    f.update({k: rng.standard_normal(N) for k in ("mask", "maskvx", "maskvy", "vor_mean")})
    b = {k: rng.standard_normal(ns) for k in ("x", "y", "vx", "vy", "theta", "theta_dot", "pn", "po", "tn", "to",
                                             "ax", "ay")}
    oo = rng.standard_normal(6 * ns)
    d = lambda *v: struct.pack("<%dd" % len(v), *v)  # noqa: E731
    i4 = lambda v: struct.pack("<i", v)  # noqa: E731
    recs = [
        d(time), d(123.0),
        i4(n1) + d(2.5e-4) + f["vortk"].tobytes() + f["nlk"].tobytes() + f["workvis"].tobytes()
        + f["mask"].tobytes() + f["maskvx"].tobytes() + f["maskvy"].tobytes() + i4(17) + f["u"].tobytes(),
        i4(ns), d(3.5) + i4(-1),
        b"".join(b[k].tobytes() for k in ("x", "y", "vx", "vy", "theta", "theta_dot")),
        b"".join(b[k].tobytes() for k in ("pn", "po", "tn", "to")),
        d(1.0, 2.0, 3.0, 4.0, 5.0, 6.0, 7.0, 8.0, 9.0, 1.5, 2.5),
        d(10.0) + i4(601) + d(0.1, 0.2, 0.3, 0.4),
        i4(0) + d(1e-4) + oo.tobytes(),
        b["ax"].tobytes() + b["ay"].tobytes() + d(0.5, -0.5),
        f["vor_mean"].tobytes(),
    ]
    rf.write_records(path, recs)
    return f, b, oo


def test_records_round_trip_and_markers(tmp_path):
    p = tmp_path / "r.bin"
    rf.write_records(p, [b"abc", b"", bytes(range(256))])
    assert rf.read_records(p) == [b"abc", b"", bytes(range(256))]
    raw = bytearray(open(p, "rb").read())
    assert raw[:4] == struct.pack("<i", 3) and raw[7:11] == struct.pack("<i", 3)
    raw[7] = 4  # trailing marker differs from the leading one
    open(p, "wb").write(raw)
    with pytest.raises(rf.FortranRecordError):
        rf.read_records(p)


def test_backup_parser(tmp_path):
    p = tmp_path / "runtime_backup0.in"
    f, b, oo = synthetic_backup(p)
    d = rf.read_backup(p, 8, 6, 5)
    assert d["time"] == 0.375 and d["checksum"] == 123.0 and d["n1"] == 1 and d["dt1"] == 2.5e-4
    assert d["ivideo"] == 17 and d["colorscale"] == 3.5 and d["colorscale_done"]
    assert np.array_equal(d["vortk"].ravel(), f["vortk"]) and np.array_equal(d["u"].ravel(), f["u"])
    assert np.array_equal(d["maskvy"].ravel(), f["maskvy"]) and np.array_equal(d["vor_mean"].ravel(), f["vor_mean"])
    bb = d["beams"][0]
    assert np.array_equal(bb["theta"], b["theta"]) and np.array_equal(bb["tau_old"], b["to"])
    assert list(bb["Force"]) == [1.0, 2.0] and list(bb["Force_press"]) == [5.0, 6.0] and bb["y0"] == 2.5
    assert bb["iMouvement"] == 601 and bb["lift_unst_old"] == 0.4 and not bb["UnsteadyCorrectionsReady"]
    assert np.array_equal(bb["beam_oldold"].ravel(), oo) and list(bb["Inertial_Force"]) == [0.5, -0.5]
    with pytest.raises(rf.FortranRecordError):
        rf.read_backup(p, 8, 6, 6)  # ns differs
    with pytest.raises(rf.FortranRecordError):
        rf.read_backup(p, 8, 8, 5)  # field record of another size


@pytest.mark.skipif(not os.path.exists(READER), reason="tests/fortran/build/read_backup is built by build()")
def test_fortran_reader_agrees(tmp_path):
    """the Fortran reader (amdflang, default REAL 8 bytes) reads the same
    values from the same bytes"""
    p = tmp_path / "runtime_backup1.in"
    f, b, oo = synthetic_backup(p, seed=3)
    out = subprocess.run([READER, str(p), "8", "6", "5"], capture_output=True, text=True, check=True).stdout
    got = {ln.split()[0]: ln.split()[1:] for ln in out.splitlines() if ln.strip()}
    assert float(got["checksum"][0]) == 123.0 and int(got["n1"][0]) == 1 and int(got["ivideo"][0]) == 17
    assert got["colorscale_done"][0] == "T" and got["UnsteadyCorrectionsReady"][0] == "F"
    N = 48
    for key, arr in (("vortk", f["vortk"]), ("u", f["u"])):
        for lev in (0, 1):
            assert float(got[key][lev]) == pytest.approx(arr[lev * N:(lev + 1) * N].sum(), rel=1e-13, abs=1e-13)
    assert float(got["vor_mean"][0]) == pytest.approx(f["vor_mean"].sum(), rel=1e-13, abs=1e-13)
    assert float(got["oldold"][0]) == pytest.approx(oo.sum(), rel=1e-13, abs=1e-13)
    assert [float(v) for v in got["forces"]] == [1.0, 2.0, 3.0, 4.0, 5.0, 6.0]


def test_latest_backup(tmp_path):
    assert rf.latest_backup(tmp_path) is None
    synthetic_backup(tmp_path / "runtime_backup0.in", time=0.5)
    synthetic_backup(tmp_path / "runtime_backup1.in", time=0.75)
    assert rf.latest_backup(tmp_path).endswith("runtime_backup1.in")
    synthetic_backup(tmp_path / "runtime_backup1.in", time=0.5)  # a tie keeps the first
    assert rf.latest_backup(tmp_path).endswith("runtime_backup0.in")


def test_inicond_parser(tmp_path):
    p = tmp_path / "inicond"
    v = np.arange(12.0)
    rf.write_records(p, [struct.pack("<2i", 4, 3) + v.tobytes()])
    nx, ny, w = rf.read_inicond(p)
    assert (nx, ny) == (4, 3) and np.array_equal(w, v.reshape(3, 4))
