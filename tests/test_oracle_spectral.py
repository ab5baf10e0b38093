"""The oracle's FFT plugin and spectral operators, pinned without the oracle
itself: a brute-force DFT, numpy's FFT, single-mode closed forms of the
packed layout (coftxy, cof_fftw33.f90:142-193; cofitxy :197-243), analytic
derivatives (cofdx.f90, cofdy.f90, cofdxdx.f90, cofdydy.f90, cofdxdy.f90),
the Poisson solve (poisson.f90), the viscous factor (cal_vis.f90) and the
dealiasing ellipse (dealiase_mask.f90)."""
import numpy as np
import pytest

from modes import field_of, mode, packed_entry, packed_of, random_modes

SIZES = [(8, 8), (16, 12), (12, 16), (64, 48), (96, 128)]


def _oracle(oracle_mod, nx, ny, xl=4.0, yl=3.0, nu=1e-2):
    return oracle_mod.Oracle(nx=nx, ny=ny, xl=xl, yl=yl, nu=nu, iBeam=0, iSponge=0)


@pytest.mark.parametrize("nx,ny", [(8, 8), (16, 12), (12, 16), (20, 10)])
def test_rfft2_brute_force_dft(oracle_mod, nx, ny):
    """r2c half spectrum against the O(N^3) matrix DFT."""
    o = _oracle(oracle_mod, nx, ny)
    f = np.random.default_rng(nx * 100 + ny).normal(size=(ny, nx))
    Wx = np.exp(-2j * np.pi * np.outer(np.arange(nx), np.arange(nx // 2 + 1)) / nx)
    Wy = np.exp(-2j * np.pi * np.outer(np.arange(ny), np.arange(ny)) / ny)
    ref = Wy @ f @ Wx
    got = o.rfft2(f)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    back = o.irfft2(got)  # unnormalised c2r
    assert np.abs(back - nx * ny * f).max() <= 1e-12 * nx * ny * np.abs(f).max()


@pytest.mark.parametrize("nx,ny", [(64, 64), (96, 128), (256, 256), (384, 192), (250, 100)])
def test_rfft2_numpy(oracle_mod, nx, ny):
    o = _oracle(oracle_mod, nx, ny)
    f = np.random.default_rng(7).normal(size=(ny, nx))
    ref = np.fft.rfft2(f)
    got = o.rfft2(f)
    assert np.abs(got - ref).max() <= 2e-13 * np.abs(ref).max()


@pytest.mark.parametrize("nx,ny", SIZES)
def test_coftxy_single_modes(oracle_mod, nx, ny):
    """every Nyquist-free mode lands on exactly its packed entry"""
    xl, yl = 4.0, 3.0
    o = _oracle(oracle_mod, nx, ny, xl, yl)
    rng = np.random.default_rng(1)
    for kx, ky, sx, sy, _ in random_modes(rng, nx, ny, 24):
        fk = o.coftxy(mode(nx, ny, xl, yl, kx, ky, sx, sy))
        iy, ix, v = packed_entry(kx, ky, sx, sy)
        assert abs(fk[iy, ix] - v) < 1e-14, (kx, ky, sx, sy)
        fk[iy, ix] = 0.0
        assert np.abs(fk).max() < 1e-14, (kx, ky, sx, sy)


@pytest.mark.parametrize("nx,ny", SIZES)
def test_coftxy_drops_nyquist(oracle_mod, nx, ny):
    """the kx = nx/2 and ky = ny/2 modes are dropped (cof_fftw33.f90:166-190)"""
    xl, yl = 4.0, 3.0
    o = _oracle(oracle_mod, nx, ny, xl, yl)
    for f in (mode(nx, ny, xl, yl, nx // 2, 0, "c", "c"), mode(nx, ny, xl, yl, 0, ny // 2, "c", "c"),
              mode(nx, ny, xl, yl, nx // 2, 3, "c", "s"), mode(nx, ny, xl, yl, 1, ny // 2, "s", "c")):
        assert np.abs(o.coftxy(f)).max() < 1e-14


@pytest.mark.parametrize("nx,ny", SIZES)
def test_cofitxy_is_inverse_on_nyquist_free_fields(oracle_mod, nx, ny):
    xl, yl = 4.0, 3.0
    o = _oracle(oracle_mod, nx, ny, xl, yl)
    ms = random_modes(np.random.default_rng(2), nx, ny, 16)
    f = field_of(nx, ny, xl, yl, ms)
    fk = o.coftxy(f)
    assert np.abs(fk - packed_of(nx, ny, ms)).max() < 1e-13
    assert np.abs(o.cofitxy(fk) - f).max() < 1e-12 * max(1.0, np.abs(f).max())
    # cofitxy of a packed single entry is the mode itself
    kx, ky = 2, 1
    fk = packed_of(nx, ny, [(kx, ky, "s", "c", 1.0)])
    assert np.abs(o.cofitxy(fk) - mode(nx, ny, xl, yl, kx, ky, "s", "c")).max() < 1e-13


@pytest.mark.parametrize("nx,ny", [(16, 12), (64, 48), (96, 128)])
def test_derivatives_closed_form(oracle_mod, nx, ny):
    """cofdx/cofdy/cofdxdx/cofdydy/cofdxdy of band-limited fields equal the
    packed transform of the analytic derivative"""
    xl, yl = 4.0, 3.0
    o = _oracle(oracle_mod, nx, ny, xl, yl)
    ms = random_modes(np.random.default_rng(3), nx, ny, 20)
    fk = packed_of(nx, ny, ms)
    # cofdxdy scales every block by fac = -kx ky (2pi/xl)(2pi/yl)
    # (cofdxdy.f90:21-27), which is MINUS the mixed derivative
    cases = [("cofdx", 1, 0, 1), ("cofdy", 0, 1, 1), ("cofdxdx", 2, 0, 1), ("cofdydy", 0, 2, 1),
             ("cofdxdy", 1, 1, -1)]
    for name, ox, oy, sign in cases:
        ref = sign * o.coftxy(field_of(nx, ny, xl, yl, ms, ox, oy))
        got = getattr(o, name)(fk)
        scale = max(1.0, np.abs(ref).max())
        assert np.abs(got - ref).max() < 1e-12 * scale, name


@pytest.mark.parametrize("nx,ny", [(16, 12), (64, 48)])
def test_poisson_closed_form(oracle_mod, nx, ny):
    """poisson divides each mode by |k|^2 (phys), the mean mode goes to 0"""
    xl, yl = 4.0, 3.0
    o = _oracle(oracle_mod, nx, ny, xl, yl)
    ms = random_modes(np.random.default_rng(4), nx, ny, 20)
    k2 = lambda kx, ky: (2 * np.pi * kx / xl) ** 2 + (2 * np.pi * ky / yl) ** 2
    ref = packed_of(nx, ny, ms, scale=lambda kx, ky: 0.0 if kx == ky == 0 else 1.0 / k2(kx, ky))
    got = o.poisson(packed_of(nx, ny, ms))
    assert np.abs(got - ref).max() < 1e-14 * max(1.0, np.abs(ref).max())
    # -laplacian(poisson(f)) = f minus its mean
    lap = o.cofdxdx(got) + o.cofdydy(got)
    f = packed_of(nx, ny, ms)
    f[0, 0] = 0.0
    assert np.abs(-lap - f).max() < 1e-12


@pytest.mark.parametrize("nx,ny", [(16, 12), (64, 48)])
def test_vis_closed_form(oracle_mod, nx, ny):
    """vis(dt) = exp(-nu dt |k|^2) on every 2x2 block (cal_vis.f90:20-31)"""
    xl, yl, nu, dt = 4.0, 3.0, 1.3e-2, 0.37
    o = _oracle(oracle_mod, nx, ny, xl, yl, nu)
    w = o.vis(dt)
    kx = np.repeat(np.arange(nx // 2), 2)[None, :]
    ky = np.repeat(np.arange(ny // 2), 2)[:, None]
    ref = np.exp(-nu * dt * ((2 * np.pi * kx / xl) ** 2 + (2 * np.pi * ky / yl) ** 2))
    assert np.abs(w - ref).max() <= 4e-16


@pytest.mark.parametrize("nx,ny", [(16, 12), (64, 48), (256, 256)])
def test_dealiase_ellipse(oracle_mod, nx, ny):
    """retained set = ellipse kx^2/kt_x^2 + ky^2/kt_y^2 < 1, kt = 2/3 (n/2 - 1)"""
    o = _oracle(oracle_mod, nx, ny)
    d = o.dealiase_mask()
    assert set(np.unique(d)) <= {0.0, 1.0}
    ktx, kty = 2.0 / 3.0 * (nx / 2 - 1), 2.0 / 3.0 * (ny / 2 - 1)
    n_in = sum(1 for a in range(nx // 2) for b in range(ny // 2) if a * a / ktx ** 2 + b * b / kty ** 2 < 1.0)
    assert int(d.sum()) == 4 * n_in
    # 2x2 blocks are uniform and (0,0) is kept, the corner mode is cut
    blocks = d.reshape(ny // 2, 2, nx // 2, 2)
    assert np.all(blocks == blocks[:, :1, :, :1])
    assert d[0, 0] == 1.0 and d[-1, -1] == 0.0
