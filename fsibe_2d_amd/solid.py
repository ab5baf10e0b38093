"""The library's host solid solver for one beam (SolidSolver.f90), through
the C ABI fsibe_solid_* of include/fsibe_hip.h.

    b = Solid(params)                         init_beams + InitializeSolidSolver
    b.set("pressure_new", p); b.set("tau_new", tau)   GetForces' loads
    b.step(time, dt)                          GravityImpulse + SolidSolverWrapper
                                              (time_step.f90:209-228)
    b.get("theta"), b.info()                  type(solid) arrays and scalars

guess / residual / jacobian / rhs expose the Newton system of the next step
(predictor, F_nonlinear, Jacobi) and RHS_beameqn.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import SOLID_ARRAYS, FsibeSolidInfo, check, lib
from .params import Params


class Solid:
    def __init__(self, params: Params | None = None, handle=None, owner=None, beam: int = 0):
        """params: a new solver for beam `beam` at its placement
        (init_beam.f90:44-76); handle: a solver owned by `owner`"""
        if handle is None:
            self.p = params
            self._pc = params.to_c()
            h = C.c_void_p()
            check(lib().fsibe_solid_create_beam(C.byref(self._pc), int(beam), C.byref(h)))
            self.h, self._owned = h, True
        else:
            self.p = params
            self.h, self._owned = C.c_void_p(handle), False
        self._owner = owner  # keeps the simulation that owns the handle alive
        self.ns = self.p.ns
        self.n = 2 * self.ns + 4

    def close(self):
        if getattr(self, "h", None) and self._owned:
            lib().fsibe_solid_destroy(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _size(self, name):
        return 6 * self.ns if name == "oldold" else self.ns

    def get(self, name: str) -> np.ndarray:
        out = np.zeros(self._size(name))
        check(lib().fsibe_solid_get(self.h, SOLID_ARRAYS.index(name), out))
        return out.reshape(6, self.ns) if name == "oldold" else out

    def set(self, name: str, values):
        a = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        assert a.size == self._size(name), (name, a.size)
        check(lib().fsibe_solid_set(self.h, SOLID_ARRAYS.index(name), a))

    def info(self) -> dict:
        o = FsibeSolidInfo()
        check(lib().fsibe_solid_get_info(self.h, C.byref(o)))
        return dict(dt_old=o.dt_old, grav=o.grav, E_kinetic=o.E_kinetic, E_pot=o.E_pot, E_elastic=o.E_elastic,
                    Inertial_Force=tuple(o.Inertial_Force), StartupStep=bool(o.StartupStep), iter=o.iter,
                    brake=bool(o.brake), method=o.method)

    def set_info(self, dt_old: float, grav: float, StartupStep: bool):
        o = FsibeSolidInfo()
        o.dt_old, o.grav, o.StartupStep = float(dt_old), float(grav), int(bool(StartupStep))
        check(lib().fsibe_solid_set_info(self.h, C.byref(o)))

    def step(self, time: float, dt: float):
        check(lib().fsibe_solid_step(self.h, float(time), float(dt)))

    def guess(self, time: float, dt: float) -> np.ndarray:
        x = np.zeros(self.n)
        check(lib().fsibe_solid_guess(self.h, float(time), float(dt), x))
        return x

    def residual(self, time: float, dt: float, x) -> np.ndarray:
        F = np.zeros(self.n)
        check(lib().fsibe_solid_residual(self.h, float(time), float(dt), np.ascontiguousarray(x, np.float64), F))
        return F

    def jacobian(self, time: float, dt: float, x) -> np.ndarray:
        J = np.zeros((self.n, self.n))
        check(lib().fsibe_solid_jacobian(self.h, float(time), float(dt), np.ascontiguousarray(x, np.float64), J))
        return J

    def rhs(self, time: float, theta, theta_dot, pressure, tau):
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64)
        th_t, th_tt, T = np.zeros(self.ns), np.zeros(self.ns), np.zeros(self.ns)
        check(lib().fsibe_solid_rhs(self.h, float(time), f(theta), f(theta_dot), f(pressure), f(tau),
                                    th_t, th_tt, T))
        return th_t, th_tt, T
