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


# ---- SaveGIF -------------------------------------------------------------------
GIF_REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "save_gif_ref")


def _gif_fields():
    rng = np.random.default_rng(7)
    y, x = np.mgrid[0:64, 0:96]
    smooth = np.sin(x / 9.0) * np.cos(y / 7.0) + 0.3 * x / 96.0
    noise = rng.standard_normal((150, 200))  # enough distinct runs to fill and clear the 4096-code table
    return [("smooth", smooth, 0, None), ("noise", noise, 0, None), ("band14", smooth, 14, (-0.5, 0.9)),
            ("spring11", noise[:40, :70], 11, None), ("clamped13", 3.0 * smooth, 13, (-1.0, 1.0)),
            ("constant", np.full((16, 24), 0.25), 0, None)]  # 0/0: every pixel gets colour 0


def _save(path, f, pal, rng):
    from fsibe_2d_amd.solver import save_gif

    if rng is None:
        save_gif(path, f, palette=pal or 13)
    else:
        save_gif(path, f, palette=pal, color_min=rng[0], color_max=rng[1])


@pytest.mark.skipif(not os.path.exists(GIF_REF), reason="oracle/_ref/save_gif_ref not built (needs the reference sources)")
@pytest.mark.parametrize("case", range(6))
def test_save_gif_matches_reference_bytes(tmp_path, case):
    """the GIF of fsibe_save_gif against the one the reference's SaveGIF
    (FieldExport.f90 with gif_util.f90, built from the reference's sources)
    writes for the same field: identical bytes (palette, index mapping,
    orientation, LZW codes, sub-blocks)"""
    name, f, pal, rng = _gif_fields()[case]
    ny, nx = f.shape
    raw = tmp_path / "in.bin"
    with open(raw, "wb") as h:
        h.write(np.array([nx, ny, pal, 0 if rng is None else 1], dtype="<i4").tobytes())
        h.write(np.array(rng if rng is not None else (0.0, 0.0), dtype="<f8").tobytes())
        h.write(np.ascontiguousarray(f, dtype="<f8").tobytes())
    subprocess.run([GIF_REF, str(raw), str(tmp_path / "ref")], check=True, capture_output=True)
    _save(str(tmp_path / "ours"), f, pal, rng)
    a = (tmp_path / "ours.gif").read_bytes()
    b = (tmp_path / "ref.gif").read_bytes()
    print(f"[gif {name}] {len(a)} bytes", flush=True)
    assert a == b


def test_save_gif_decodes_to_the_index_raster(tmp_path):
    """an independent GIF decoder (Pillow) reads back the raster: size, the
    field's last row on top, and the palette's ends (palette 13 runs from
    (0, 0, 143) to (128, 0, 0))"""
    from PIL import Image

    y, x = np.mgrid[0:50, 0:300]
    f = x + 0.001 * y  # every column its own colour band, rows distinguishable
    _save(str(tmp_path / "g"), f, 13, None)
    im = Image.open(tmp_path / "g.gif")
    assert im.size == (300, 50) and im.mode == "P"
    idx = np.array(im)
    lo, hi = f.min(), f.max()
    want = np.clip(1 + np.floor(255.0 * (f - lo) / (hi - lo) + 0.5).astype(int), 1, 256) - 1
    assert np.array_equal(idx, want[::-1])
    pal = np.array(im.getpalette()[:768]).reshape(256, 3)
    assert tuple(pal[0]) == (0, 0, 143) and tuple(pal[255]) == (128, 0, 0)


def test_save_gif_refuses_tabulated_palettes(tmp_path):
    from fsibe_2d_amd import FsibeError

    with pytest.raises(FsibeError):
        _save(str(tmp_path / "g"), np.zeros((8, 8)), 5, None)
