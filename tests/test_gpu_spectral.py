"""Tier A parity on the GPU: the FFT/operator plugin against the oracle.

Tolerances: the packed transforms agree with the oracle to a few ulps of the
field's largest coefficient times log2(n) (the two FFTs round differently);
the elementwise operators (cofdx .. poisson) are the same IEEE operations in
the same order and must agree bit for bit; vis uses exp, which libm and the
GPU may round 1 ulp apart.
"""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

SIZES = [(8, 8), (16, 12), (64, 48), (96, 128), (256, 256), (384, 384), (768, 512), (1024, 1024)]


@pytest.fixture(scope="module")
def spectral_mod():
    from fsibe_2d_amd import spectral

    return spectral


@pytest.mark.parametrize("nx,ny", SIZES)
def test_coftxy_cofitxy(nx, ny, spectral_mod, oracle_mod):
    xl, yl = 4.0, 3.0
    o = oracle_mod.Oracle(nx=nx, ny=ny, xl=xl, yl=yl, iBeam=0, iSponge=0)
    s = spectral_mod.Spectral(nx, ny, xl, yl, 1e-3)
    rng = np.random.default_rng(nx * 1000 + ny)
    f = rng.standard_normal((ny, nx))
    fk_o = o.coftxy(f)
    fk_g = s.coftxy(f)
    err = np.abs(fk_g - fk_o).max() / np.abs(fk_o).max()
    print(f"coftxy {nx}x{ny} max rel err {err:.2e}", flush=True)
    assert err < 2e-14
    # cofitxy of a packed field
    g_o = o.cofitxy(fk_o)
    g_g = s.cofitxy(fk_o)
    err2 = np.abs(g_g - g_o).max() / np.abs(g_o).max()
    print(f"cofitxy {nx}x{ny} max rel err {err2:.2e}", flush=True)
    assert err2 < 2e-14
    # round trip is the identity on Nyquist-free data
    assert rel_l2(s.coftxy(g_g), fk_o) < 1e-14


@pytest.mark.parametrize("nx,ny", [(16, 12), (96, 128), (256, 256)])
def test_operators_bitexact(nx, ny, spectral_mod, oracle_mod):
    xl, yl = 6.283185307179586, 2.0
    o = oracle_mod.Oracle(nx=nx, ny=ny, xl=xl, yl=yl, iBeam=0, iSponge=0, nu=1e-3)
    s = spectral_mod.Spectral(nx, ny, xl, yl, 1e-3)
    fk = np.random.default_rng(7).standard_normal((ny, nx))
    for name in ("cofdx", "cofdy", "cofdxdx", "cofdydy", "cofdxdy", "poisson"):
        a, b = getattr(s, name)(fk), getattr(o, name)(fk)
        assert np.array_equal(a, b), (name, np.abs(a - b).max())
    v_g, v_o = s.vis(3.7e-4), o.vis(3.7e-4)
    assert np.abs(v_g - v_o).max() <= 2.3e-16 * np.abs(v_o).max()
    assert np.array_equal(s.dealiase_mask(), o.dealiase_mask())


def test_unsupported_size_raises(spectral_mod):
    from fsibe_2d_amd import FsibeError

    with pytest.raises(FsibeError):
        spectral_mod.Spectral(100, 100, 1.0, 1.0)  # 50 = 2 * 5^2
