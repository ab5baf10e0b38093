#!/usr/bin/env python3
"""Writes the golden vectors under tests/golden/ (run from the repository root
after `make -f oracle/ref.mk`):

* ref_operators_32x24.npz -- the reference's own cofdx, cofdy, cofdxdx,
  cofdydy, cofdxdy, poisson (cofd*.f90, poisson.f90), vis (cal_vis.f90) and
  dealiase_mask (dealiase_mask.f90), compiled unmodified into
  oracle/_ref/libfsibe_ref.so, on a seeded random packed field;
* ref_kinematics.npz -- the reference's mouvement (mouvement.f90) for every
  active motion case at several times, and integrate_position
  (integrate_position.f90) for seeded random beam angles;
* packed_fft_64x48.npz -- the packed spectrum of coftxy (cof_fftw33.f90:142-193)
  built from numpy.fft.rfft2 and the packing rule, independent of the oracle
  and of the GPU library.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from refbuild import Ref  # noqa: E402

MOTIONS = [0, 2, 3, 4, 201, 206, 213, 300, 350, 401, 405, 413]
TIMES = [0.0, 0.013, 0.5, 0.9999, 1.0, 1.37, 4.2]


def pack(f):
    """coftxy of a real (ny, nx) field from numpy's FFT (Nyquist modes dropped)."""
    ny, nx = f.shape
    c = np.fft.rfft2(f) / (nx * ny)
    fk = np.zeros((ny, nx))
    H, hy = nx // 2, ny // 2
    for kx in range(H):
        fk[0, 2 * kx] = c[0, kx].real
        fk[0, 2 * kx + 1] = c[0, kx].imag
        for ky in range(1, hy):
            cp, cm = c[ky, kx], c[ny - ky, kx]
            fk[2 * ky, 2 * kx] = 0.5 * (cp.real + cm.real)
            fk[2 * ky + 1, 2 * kx] = 0.5 * (cp.imag - cm.imag)
            fk[2 * ky, 2 * kx + 1] = 0.5 * (cp.imag + cm.imag)
            fk[2 * ky + 1, 2 * kx + 1] = 0.5 * (-cp.real + cm.real)
    return fk


def main():
    nx, ny, xl, yl, nu = 32, 24, 4.0, 3.0, 1.3e-2
    r = Ref(nx, ny, xl=xl, yl=yl, nu=nu)
    fk = np.random.default_rng(11).normal(size=(ny, nx))
    ops = {name: r.op(i, fk) for i, name in enumerate(
        ["cofdx", "cofdy", "cofdxdx", "cofdydy", "cofdxdy", "poisson"])}
    np.savez_compressed(os.path.join(HERE, "ref_operators_32x24.npz"), nx=nx, ny=ny, xl=xl, yl=yl, nu=nu,
                        fk=fk, vis_dt=0.37, vis=r.vis(0.37), dealiase=r.dealiase(), **ops)

    ns, length = 32, 1.25
    r = Ref(16, 16, ns=ns, length=length)
    mv = np.zeros((len(MOTIONS), len(TIMES), 9))
    for a, im in enumerate(MOTIONS):
        for b, t in enumerate(TIMES):
            al, at, att, le = r.mouvement(im, 0.8, 1.9, 7.5, 0.25, t)
            mv[a, b] = [al, at, att, *le]
    rng = np.random.default_rng(12)
    ip_cases = [(0, 0.0), (3, 0.7), (201, 0.4), (4, 1.1)]
    theta = 0.3 * rng.normal(size=(len(ip_cases), ns))
    theta_dot = rng.normal(size=(len(ip_cases), ns))
    pos = np.zeros((len(ip_cases), 4, ns))
    for k, (im, t) in enumerate(ip_cases):
        pos[k] = r.integrate_position(im, 0.8, 1.9, 7.5, 0.0, t, theta[k], theta_dot[k])
    np.savez_compressed(os.path.join(HERE, "ref_kinematics.npz"), motions=np.array(MOTIONS),
                        times=np.array(TIMES), x0=0.8, y0=1.9, angle=7.5, phase=0.25, mouvement=mv,
                        ns=ns, length=length, ip_cases=np.array(ip_cases), theta=theta,
                        theta_dot=theta_dot, positions=pos)

    f = np.random.default_rng(13).normal(size=(48, 64))
    np.savez_compressed(os.path.join(HERE, "packed_fft_64x48.npz"), f=f, fk=pack(f))


if __name__ == "__main__":
    main()
