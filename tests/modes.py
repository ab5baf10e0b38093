"""Analytic test fields on the reference grid x_i = i*dx, y_j = j*dy.

Arrays are (ny, nx) with a[j, i] = f(x_i, y_j), the reference's f(i, j).
A trigonometric mode cos/sin(a x) * cos/sin(b y) with a = 2 pi kx / xl,
b = 2 pi ky / yl lands on exactly one entry of the packed spectral layout
(coftxy, cof_fftw33.f90:142-193), with the value and sign worked out from the
DFT of the mode:

    cos cos -> +1/4 at (2kx,   2ky)      sin sin -> +1/4 at (2kx+1, 2ky+1)
    sin cos -> -1/4 at (2kx+1, 2ky)      cos sin -> -1/4 at (2kx,   2ky+1)

for kx, ky >= 1; for ky = 0 the row 0 holds the 1-D r2c coefficients
(cos -> +1/2 at 2kx, sin -> -1/2 at 2kx+1), for kx = 0 the column 0 the
folded ky pair (cos -> +1/2 at 2ky, sin -> -1/2 at 2ky+1), and the
constant 1 -> 1 at (0, 0).
"""
from __future__ import annotations

import numpy as np


def grid(nx, ny, xl, yl):
    x = np.arange(nx) * (xl / nx)
    y = np.arange(ny) * (yl / ny)
    return np.meshgrid(x, y)  # X[j, i], Y[j, i]


def trig(kind, k, x, L):
    a = 2.0 * np.pi * k / L
    return np.cos(a * x) if kind == "c" else np.sin(a * x)


def mode(nx, ny, xl, yl, kx, ky, sx, sy, dx_order=0, dy_order=0):
    X, Y = grid(nx, ny, xl, yl)
    fx = trig(sx, kx, X, xl) if dx_order == 0 else _deriv(sx, kx, X, xl, dx_order)
    fy = trig(sy, ky, Y, yl) if dy_order == 0 else _deriv(sy, ky, Y, yl, dy_order)
    return fx * fy


def _deriv(kind, k, x, L, n):
    a = 2.0 * np.pi * k / L
    # d^n/dx^n cos(ax) = a^n cos(ax + n pi/2); sin(ax) -> a^n sin(ax + n pi/2)
    ph = n * np.pi / 2.0
    return a ** n * (np.cos(a * x + ph) if kind == "c" else np.sin(a * x + ph))


def packed_entry(kx, ky, sx, sy):
    """(row index iy, column index ix, value) of the packed entry of a mode"""
    if kx == 0 and ky == 0:
        assert sx == "c" and sy == "c"
        return 0, 0, 1.0
    if ky == 0:
        assert sy == "c"
        return 0, 2 * kx + (sx == "s"), (0.5 if sx == "c" else -0.5)
    if kx == 0:
        assert sx == "c"
        return 2 * ky + (sy == "s"), 0, (0.5 if sy == "c" else -0.5)
    v = {("c", "c"): 0.25, ("s", "s"): 0.25, ("s", "c"): -0.25, ("c", "s"): -0.25}[(sx, sy)]
    return 2 * ky + (sy == "s"), 2 * kx + (sx == "s"), v


def random_modes(rng, nx, ny, n, kmax=None):
    """n distinct (kx, ky, sx, sy, amplitude) tuples, Nyquist-free"""
    kxm = nx // 2 - 1 if kmax is None else min(kmax, nx // 2 - 1)
    kym = ny // 2 - 1 if kmax is None else min(kmax, ny // 2 - 1)
    out, seen = [], set()
    while len(out) < n:
        kx, ky = int(rng.integers(0, kxm + 1)), int(rng.integers(0, kym + 1))
        sx = "c" if kx == 0 else ("c" if rng.random() < 0.5 else "s")
        sy = "c" if ky == 0 else ("c" if rng.random() < 0.5 else "s")
        key = (kx, ky, sx, sy)
        if key in seen:
            continue
        seen.add(key)
        out.append((kx, ky, sx, sy, float(rng.normal())))
    return out


def field_of(nx, ny, xl, yl, modes, dx_order=0, dy_order=0):
    f = np.zeros((ny, nx))
    for kx, ky, sx, sy, amp in modes:
        f += amp * mode(nx, ny, xl, yl, kx, ky, sx, sy, dx_order, dy_order)
    return f


def packed_of(nx, ny, modes, scale=None):
    """packed layout of sum amp * mode; scale(kx, ky) multiplies each mode"""
    fk = np.zeros((ny, nx))
    for kx, ky, sx, sy, amp in modes:
        iy, ix, v = packed_entry(kx, ky, sx, sy)
        fk[iy, ix] += amp * v * (1.0 if scale is None else scale(kx, ky))
    return fk
