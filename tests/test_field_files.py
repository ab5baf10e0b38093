"""Field outputs on the host (no GPU): SaveField's text file
(FieldExport.f90:612-648) written by the library (fsibe_save_field) against
the Fortran runtime's own ES/I formatting of the same values and records
(tests/fortran/write_field.f90), and the oracle's ComputePressureSnapshot
(NavierStokes.f90:596-665) pinned by the Taylor-Green vortex."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
WRITER = os.path.join(ROOT, "tests", "fortran", "build", "write_field")


def _values(nx, ny, seed=0):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((ny, nx)) * 10.0 ** rng.integers(-6, 6, (ny, nx))
    v[0, 0], v[0, 1], v[1, 0] = 0.0, 1e-120, -3.5e150  # zero, three-digit exponents
    v[1, 1], v[2, 2] = 9.99995, -0.000123455  # values near a rounding edge
    return v


@pytest.mark.skipif(not os.path.exists(WRITER), reason="tests/fortran/build/write_field is built by build()")
@pytest.mark.parametrize("ftype,nsave", [("vorticity", 1), ("precision", 1), ("pressure", 3)])
def test_save_field_matches_fortran_formatting(tmp_path, ftype, nsave):
    from fsibe_2d_amd.solver import save_field

    nx, ny = 13, 7
    v = _values(nx, ny, seed=nsave)
    raw = tmp_path / "v.bin"
    v.astype("<f8").tofile(raw)  # [ny][nx] = Fortran field(nx, ny)
    ref = tmp_path / "ref.txt"
    subprocess.run([WRITER, str(raw), str(nx), str(ny), str(nsave), ftype, str(ref)], check=True)
    out = tmp_path / "lib.txt"
    save_field(out, v, nsave=nsave, xl=4.0, yl=2.5, fieldtype=ftype, header="% nx=64 ny=64\n% Re=1000")
    a, b = open(out, "rb").read(), open(ref, "rb").read()
    assert a == b
    lines = a.decode().splitlines()
    assert len(lines[0]) == 256 and lines[2] == "%format1" and lines[3] == "00013   % nx"
    assert len(lines) == 2 + 7 + len(range(0, ny, nsave)) * len(range(0, nx, nsave))


def test_pressure_snapshot_of_taylor_green(oracle_mod):
    """u = (sin x cos y, -cos x sin y) on (0, 2 pi)^2, no mask, no mean flow:
    the static pressure is (cos 2x + cos 2y)/4 (u.grad u = -grad p); the snapshot's Poisson
    solve drops the mean of p + |u|^2/2"""
    n = 32
    o = oracle_mod.Oracle(nx=n, ny=n, iBeam=0, iSponge=0, iMeanVelocity=0, xl=2 * np.pi, yl=2 * np.pi)
    x = np.arange(n) * 2 * np.pi / n
    X, Y = np.meshgrid(x, x)
    o.set_vortk(o.coftxy(2 * np.sin(X) * np.sin(Y)))
    p = o.pressure_snapshot(0.0, 0)
    u2 = (np.sin(X) * np.cos(Y)) ** 2 + (np.cos(X) * np.sin(Y)) ** 2
    pe = (np.cos(2 * X) + np.cos(2 * Y)) / 4
    q = pe + 0.5 * u2
    np.testing.assert_allclose(p, pe - q.mean(), rtol=0, atol=1e-13)
