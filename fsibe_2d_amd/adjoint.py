"""The discrete adjoint of the GPU fluid step (fsibe_adjoint_* of
include/fsibe_hip.h): J(c) = (1/T) sum_n (-Force(1)^n) dt over nsteps steps of
a prescribed-motion simulation (motion case 601, c = motion_coef[8]) and its
gradient dJ/dc by a reverse sweep over units of `chk` steps.  `snapshots`
bounds the checkpoint slots (None: one per unit start); with fewer slots the
sweep follows the binomial (revolve) schedule, with identical results."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, lib
from .params import Params


class Adjoint:
    def __init__(self, params: Params, vortk=None, nsteps: int = 1000, chk: int = 50, snapshots=None):
        self.p = params
        self._pc = params.to_c()
        vk = None
        if vortk is not None:
            vk = np.ascontiguousarray(vortk, dtype=np.float64)
            assert vk.shape == (params.ny, params.nx)
        h = C.c_void_p()
        check(lib().fsibe_adjoint_create_binomial(C.byref(self._pc), vk.ctypes.data if vk is not None else None,
                                                  int(nsteps), int(chk), -1 if snapshots is None else int(snapshots),
                                                  C.byref(h)))
        self.h = h
        self.nsteps, self.chk = nsteps, chk

    def close(self):
        if getattr(self, "h", None):
            lib().fsibe_adjoint_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def objective(self, coef) -> float:
        c = np.ascontiguousarray(coef, dtype=np.float64)
        assert c.size == 8
        J = C.c_double()
        check(lib().fsibe_adjoint_objective(self.h, c, C.byref(J)))
        return J.value

    def gradient(self, coef):
        c = np.ascontiguousarray(coef, dtype=np.float64)
        assert c.size == 8
        J, g = C.c_double(), np.zeros(8)
        check(lib().fsibe_adjoint_gradient(self.h, c, C.byref(J), g))
        return J.value, g

    def timings(self) -> dict:
        f, g, r = C.c_double(), C.c_double(), C.c_double()
        check(lib().fsibe_adjoint_timings(self.h, C.byref(f), C.byref(g), C.byref(r)))
        return {"forward_ms": f.value, "gradient_ms": g.value, "reverse_ms": r.value}

    def stats(self) -> dict:
        """snapshot slots; steps the last call ran without and with a tape"""
        q, adv, rec = C.c_int(), C.c_longlong(), C.c_longlong()
        check(lib().fsibe_adjoint_stats(self.h, C.byref(q), C.byref(adv), C.byref(rec)))
        return {"slots": q.value, "steps_advanced": adv.value, "steps_recorded": rec.value}


def plan(nunits: int, nslots: int):
    """the reverse schedule as a list of (op, unit, arg) and the units it
    advances (fsibe_adjoint_plan; host only)"""
    n, adv = C.c_int(), C.c_longlong()
    check(lib().fsibe_adjoint_plan(int(nunits), int(nslots), None, 0, C.byref(n), C.byref(adv)))
    buf = (C.c_int * (3 * max(n.value, 1)))()
    check(lib().fsibe_adjoint_plan(int(nunits), int(nslots), buf, n.value, C.byref(n), C.byref(adv)))
    acts = [(buf[3 * i], buf[3 * i + 1], buf[3 * i + 2]) for i in range(n.value)]
    return acts, adv.value


def plan_cost(nunits: int, nslots: int) -> int:
    """the least number of units any schedule advances (fsibe_adjoint_plan_cost)"""
    return int(lib().fsibe_adjoint_plan_cost(int(nunits), int(nslots)))
