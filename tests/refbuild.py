"""ctypes access to oracle/_ref/libfsibe_ref.so: the reference's own Fortran
operators, kinematics and mean flow, compiled unmodified by oracle/ref.mk.
Tests that use it skip when it has not been built (it needs the reference
sources, which only the build container has)."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PATH = os.path.join(ROOT, "oracle", "_ref", "libfsibe_ref.so")


class Ref:
    def __init__(self, nx, ny, ns=3, xl=4.0, yl=4.0, nu=1e-2, length=1.0):
        L = C.CDLL(PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        L.ref_setup.argtypes = [C.c_int, C.c_int, C.c_int] + [C.c_double] * 4
        L.ref_set_flow.argtypes = [C.c_int, C.c_double, C.c_double]
        L.ref_op.argtypes = [C.c_int, dp, dp]
        L.ref_vis.argtypes = [C.c_double, dp]
        L.ref_dealiase.argtypes = [dp]
        L.ref_mean_ux.restype = C.c_double
        L.ref_mean_ux.argtypes = [C.c_double]
        L.ref_mean_uy.restype = C.c_double
        L.ref_mean_uy.argtypes = [C.c_double]
        dref = C.POINTER(C.c_double)
        L.ref_mouvement.argtypes = [C.c_int] + [C.c_double] * 5 + [dref, dref, dref, dp]
        L.ref_integrate_position.argtypes = [C.c_int] + [C.c_double] * 5 + [dp] * 6
        self.L = L
        self.nx, self.ny, self.ns = nx, ny, ns
        L.ref_setup(nx, ny, ns, xl, yl, nu, length)

    def op(self, which, fk):
        out = np.zeros((self.ny, self.nx))
        self.L.ref_op(which, np.ascontiguousarray(fk, dtype=np.float64), out)
        return out

    def vis(self, dt):
        out = np.zeros((self.ny, self.nx))
        self.L.ref_vis(dt, out)
        return out

    def dealiase(self):
        out = np.zeros((self.ny, self.nx))
        self.L.ref_dealiase(out)
        return out

    def set_flow(self, iMean, theta_inf=0.0, T_fullspeed=0.0):
        self.L.ref_set_flow(iMean, theta_inf, T_fullspeed)

    def mean_u(self, t):
        return self.L.ref_mean_ux(t), self.L.ref_mean_uy(t)

    def mouvement(self, im, x0, y0, angle, phase, t):
        a, at, att = C.c_double(), C.c_double(), C.c_double()
        le = np.zeros(6)
        self.L.ref_mouvement(im, x0, y0, angle, phase, t, C.byref(a), C.byref(at), C.byref(att), le)
        return a.value, at.value, att.value, le

    def integrate_position(self, im, x0, y0, angle, phase, t, theta, theta_dot):
        x, y, vx, vy = (np.zeros(self.ns) for _ in range(4))
        self.L.ref_integrate_position(im, x0, y0, angle, phase, t, np.ascontiguousarray(theta, dtype=np.float64),
                                      np.ascontiguousarray(theta_dot, dtype=np.float64), x, y, vx, vy)
        return x, y, vx, vy


def available() -> bool:
    return os.path.exists(PATH)


needs_ref = pytest.mark.skipif(not available(), reason="oracle/_ref not built (needs the reference sources)")
