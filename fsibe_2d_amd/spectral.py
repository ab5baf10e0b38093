"""Tier A: the reference's FFT / spectral-operator plugin, same routine names.

    fft_initialize / fft_free      cof_fftw33.f90:18-66, 119-137
    coftxy(f) -> fk                 cof_fftw33.f90:142-193
    cofitxy(fk) -> f                cof_fftw33.f90:197-243
    cofdx, cofdy, cofdxdx, cofdydy, cofdxdy   cofd*.f90
    poisson(f) -> ans               poisson.f90:1-30
    vis(dt) -> work                 cal_vis.f90:1-36
    dealiase_mask() -> dealiase     dealiase_mask.f90:1-33

Arrays are float64 of shape (ny, nx): a[iy, ix] is the reference's f(ix, iy).
The reference reads nx, ny, xl, yl, nu from module share_vars; here they are
given to fft_initialize, which returns the plugin context.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib


class Spectral:
    def __init__(self, nx: int, ny: int, xl: float, yl: float, nu: float = 0.0):
        self.nx, self.ny = int(nx), int(ny)
        self.xl, self.yl, self.nu = float(xl), float(yl), float(nu)
        h = C.c_void_p()
        check(lib().fsibe_spectral_create(self.nx, self.ny, self.xl, self.yl, self.nu, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fsibe_spectral_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.ny, self.nx):
            raise ValueError(f"expected shape {(self.ny, self.nx)}, got {a.shape}")
        return a

    def _call(self, name, a):
        out = np.empty((self.ny, self.nx), dtype=np.float64)
        check(getattr(lib(), name)(self.h, self._arr(a), out))
        return out

    def coftxy(self, f):
        return self._call("fsibe_coftxy", f)

    def cofitxy(self, fk):
        return self._call("fsibe_cofitxy", fk)

    def cofdx(self, fk):
        return self._call("fsibe_cofdx", fk)

    def cofdy(self, fk):
        return self._call("fsibe_cofdy", fk)

    def cofdxdx(self, fk):
        return self._call("fsibe_cofdxdx", fk)

    def cofdydy(self, fk):
        return self._call("fsibe_cofdydy", fk)

    def cofdxdy(self, fk):
        return self._call("fsibe_cofdxdy", fk)

    def poisson(self, f):
        return self._call("fsibe_poisson", f)

    def vis(self, dt):
        out = np.empty((self.ny, self.nx), dtype=np.float64)
        check(lib().fsibe_vis(self.h, float(dt), out))
        return out

    def dealiase_mask(self):
        out = np.empty((self.ny, self.nx), dtype=np.float64)
        check(lib().fsibe_dealiase_mask(self.h, out))
        return out


def fft_initialize(nx, ny, xl, yl, nu=0.0) -> Spectral:
    return Spectral(nx, ny, xl, yl, nu)


def fft_free(ctx: Spectral):
    ctx.close()


def cofts(f, L: int, n: int = 1):
    """n real sequences of length L (columns of f, shape (n, L)) -> the
    truncated r2c coefficients / L (cof_fftw33.f90:248-296), host-side."""
    f = np.ascontiguousarray(f, dtype=np.float64).reshape(n, L)
    fk = np.empty_like(f)
    check(lib().fsibe_cofts(f, fk, L, n))
    return fk


def cofits(fk, L: int, n: int = 1):
    """inverse of cofts without the 1/L (cof_fftw33.f90:300-340), host-side."""
    fk = np.ascontiguousarray(fk, dtype=np.float64).reshape(n, L)
    f = np.empty_like(fk)
    check(lib().fsibe_cofits(fk, f, L, n))
    return f
