"""Tier B: the device-resident fluid step with the reference's call sequence.

    sim.create_mask(time, beam, alpha, LeadingEdge)  create_mask.f90:2-218
    dt1 = sim.evolve_fluid_explicit(time)            NavierStokes.f90:422-459
    dt1 = sim.evolve_fluid_ab2(time, dt0)            NavierStokes.f90:463-508
    p, tau, Fp = sim.get_forces(time)                BeamForces.f90:6-288
    sim.swap_levels()                                time_step.f90:152, 222
    sim.run(nsteps)                                  time_step.f90:120-341
                                                     (prescribed kinematics, or
                                                     iFLUSI=1 with sim.solid)

State lives in GPU memory between calls.  Every method raises FsibeError on a
nonzero status; there is no CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from ._lib import FsibeParams, FsibeStepOut, check, lib
from .params import Params

FIELDS = {
    "vortk": 0, "nlk": 1, "u1": 3, "u2": 4, "press": 5, "mask": 6, "maskvx": 7,
    "maskvy": 8, "mask_sponge": 9, "vor_mean": 10, "vort": 14,
}
BOX_FIELDS = {"press": 0, "stress_a": 1, "stress_b": 2, "mask": 3, "maskvx": 4, "maskvy": 5}


def step_out_dict(o: FsibeStepOut) -> dict:
    return dict(dt1=o.dt1, u_max=o.u_max, Force=tuple(o.Force), Force_unst=tuple(o.Force_unst),
                Force_press=tuple(o.Force_press), energies=tuple(o.energies), bsi=tuple(o.bsi),
                fx_int=o.fx_int, fy_int=o.fy_int, status=o.status, ibes_iter=o.ibes_iter)


def mouvement(p: Params, time: float, iMouvement: int | None = None, phase: float = 0.0):
    """(alpha, alpha_t, alpha_tt, LeadingEdge[6]) -- mouvement.f90:5-241."""
    a, at, att = C.c_double(), C.c_double(), C.c_double()
    le = np.zeros(6)
    pc = p.to_c()
    im = p.iMotion if iMouvement is None else iMouvement
    check(lib().fsibe_mouvement(C.byref(pc), im, p.x0, p.y0, p.AngleBeam, phase, float(time),
                                C.byref(a), C.byref(at), C.byref(att), le))
    return a.value, at.value, att.value, le


def integrate_position(p: Params, time: float, theta, theta_dot, beam: int = 0):
    """Beam positions and velocities from theta (integrate_position.f90) for
    beam `beam` at its placement (beam_placement)."""
    ns = p.ns
    th = np.ascontiguousarray(theta, dtype=np.float64)
    thd = np.ascontiguousarray(theta_dot, dtype=np.float64)
    x, y, vx, vy = (np.zeros(ns) for _ in range(4))
    pc = p.to_c()
    check(lib().fsibe_integrate_position_beam(C.byref(pc), int(beam), float(time), th, thd, x, y, vx, vy))
    return x, y, vx, vy


def beam_placement(p: Params, beam: int):
    """(x0, y0, phase) of beam `beam` (init_beam.f90:44-76): the second beam
    of the dragonfly motions 300..307 at x0 + 1.25 with phase (iMotion-300) pi/4"""
    x0, y0, ph = C.c_double(), C.c_double(), C.c_double()
    pc = p.to_c()
    check(lib().fsibe_beam_placement(C.byref(pc), int(beam), C.byref(x0), C.byref(y0), C.byref(ph)))
    return x0.value, y0.value, ph.value


class FluidSim:
    def __init__(self, params: Params):
        self.p = params
        self.nx, self.ny, self.ns = params.nx, params.ny, params.ns
        self.nb = params.iBeam
        self._pc = params.to_c()
        h = C.c_void_p()
        check(lib().fsibe_sim_create(C.byref(self._pc), C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().fsibe_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- state -----------------------------------------------------------------
    def set_vortk(self, vortk_packed):
        a = np.ascontiguousarray(vortk_packed, dtype=np.float64)
        assert a.shape == (self.ny, self.nx)
        check(lib().fsibe_set_vortk(self.h, a))

    def field(self, name: str) -> np.ndarray:
        out = np.empty((self.ny, self.nx), dtype=np.float64)
        check(lib().fsibe_get_field(self.h, FIELDS[name], out))
        return out

    def box(self, name: str):
        """(array, (x0, x1, y0, y1)) of a field kept on the beam's box."""
        b = (C.c_int * 4)()
        check(lib().fsibe_get_box(self.h, BOX_FIELDS[name], None, b))
        x0, x1, y0, y1 = b
        if x1 < x0 or y1 < y0:
            return np.zeros((0, 0)), (x0, x1, y0, y1)
        out = np.empty((y1 - y0 + 1, x1 - x0 + 1), dtype=np.float64)
        check(lib().fsibe_get_box(self.h, BOX_FIELDS[name], out.ctypes.data, b))
        return out, (x0, x1, y0, y1)

    # ---- the reference's per-step calls ----------------------------------------
    def create_mask(self, time, x, y, vx, vy, theta, alpha, LeadingEdge):
        """one beam: arrays of ns values, alpha a number, LeadingEdge 6 values;
        iBeam beams: arrays (iBeam, ns), alpha (iBeam,), LeadingEdge (iBeam, 6)"""
        f = lambda a: np.ascontiguousarray(a, dtype=np.float64).ravel()
        if self.nb > 1:
            check(lib().fsibe_create_mask_beams(self.h, float(time), f(x), f(y), f(vx), f(vy), f(theta),
                                                f(alpha), f(LeadingEdge)))
        else:
            check(lib().fsibe_create_mask(self.h, float(time), f(x), f(y), f(vx), f(vy), f(theta),
                                          float(alpha), f(LeadingEdge)))

    def evolve_fluid_explicit(self, time) -> float:
        d = C.c_double()
        check(lib().fsibe_evolve_explicit(self.h, float(time), C.byref(d)))
        return d.value

    def evolve_fluid_ab2(self, time, dt0) -> float:
        d = C.c_double()
        check(lib().fsibe_evolve_ab2(self.h, float(time), float(dt0), C.byref(d)))
        return d.value

    def get_forces(self, time):
        """(pressure, tau, Force_press): ns, ns, 2 values for one beam; (iBeam,
        ns), (iBeam, ns), (iBeam, 2) for several"""
        nb = max(self.nb, 1)
        pb, tb, fp = np.zeros(max(nb * self.ns, 1)), np.zeros(max(nb * self.ns, 1)), np.zeros(2 * nb)
        check(lib().fsibe_get_forces(self.h, float(time), pb, tb, fp))
        if self.nb > 1:
            return pb.reshape(nb, self.ns), tb.reshape(nb, self.ns), fp.reshape(nb, 2)
        return pb[: self.ns], tb[: self.ns], fp

    def step_out(self, beam: int | None = None) -> dict:
        """the last step's results; beam: the per-beam fields of that beam
        (default: the first)"""
        o = FsibeStepOut()
        if beam is None:
            check(lib().fsibe_get_step_out(self.h, C.byref(o)))
        else:
            check(lib().fsibe_get_step_out_beam(self.h, int(beam), C.byref(o)))
        return step_out_dict(o)

    def swap_levels(self):
        check(lib().fsibe_swap_levels(self.h))

    # ---- the time loop --------------------------------------------------------
    def run(self, nsteps: int) -> dict:
        o = FsibeStepOut()
        check(lib().fsibe_run(self.h, int(nsteps), C.byref(o)))
        return step_out_dict(o)

    def time_state(self):
        t, d, it = C.c_double(), C.c_double(), C.c_int()
        check(lib().fsibe_get_time(self.h, C.byref(t), C.byref(d), C.byref(it)))
        return t.value, d.value, it.value

    def beam(self, beam: int = 0) -> dict:
        """state of beam `beam` held by run()"""
        nb = max(self.nb, 1)
        if not 0 <= beam < nb:
            raise IndexError(f"beam {beam} of {self.nb}")
        n = max(nb * self.ns, 1)
        arrs = [np.zeros(n) for _ in range(7)]
        check(lib().fsibe_get_beam(self.h, *arrs))
        keys = ("x", "y", "vx", "vy", "theta", "pressure", "tau")
        o = beam * self.ns
        return {k: a[o: o + self.ns] for k, a in zip(keys, arrs)}

    def backup_write(self, path, ivideo: int = 1, colorscale: float = 0.0, colorscale_done: bool = False):
        """MakeRuntimeBackup (save_fields.f90:109-152) into `path`"""
        check(lib().fsibe_backup_write(self.h, os.fsencode(path), int(ivideo), float(colorscale),
                                       1 if colorscale_done else 0))

    def backup_read(self, path) -> dict:
        """init_fields with inicond = 2 from one backup file; returns the
        host's counters stored with it"""
        iv, cs, cd = C.c_int(), C.c_double(), C.c_int()
        check(lib().fsibe_backup_read(self.h, os.fsencode(path), C.byref(iv), C.byref(cs), C.byref(cd)))
        return {"ivideo": iv.value, "colorscale": cs.value, "colorscale_done": bool(cd.value)}

    def inicond_write(self, path):
        """the final vorticity file of time_step.f90:350-355"""
        check(lib().fsibe_inicond_write(self.h, os.fsencode(path)))

    def inicond_read(self, path):
        """init_fields with inicond = 111 (before the first step)"""
        check(lib().fsibe_inicond_read(self.h, os.fsencode(path)))

    def pressure_snapshot(self, time: float, level: int) -> np.ndarray:
        """ComputePressureSnapshot(time, vortk(:,:,level)) (NavierStokes.f90:598-665)"""
        out = np.zeros((self.p.ny, self.p.nx))
        check(lib().fsibe_pressure_snapshot(self.h, float(time), int(level), out))
        return out

    @property
    def stream(self):
        return lib().fsibe_sim_stream(self.h)

    @property
    def solid(self):
        """the solid solver fsibe_run uses for iFLUSI=1 (fsibe_sim_solid)"""
        return self.solid_of(0)

    def solid_of(self, beam: int):
        """the solid solver of beam `beam` (fsibe_sim_solid_beam)"""
        from .solid import Solid

        h = C.c_void_p()
        check(lib().fsibe_sim_solid_beam(self.h, int(beam), C.byref(h)))
        return Solid(self.p, handle=h.value, owner=self)


def save_field(path, field, nsave: int = 1, xl: float = 4.0, yl: float = 4.0, fieldtype: str = "vorticity",
               header: str | None = None):
    """SaveField (FieldExport.f90:612-648): the reference's text field file of
    a host field [ny][nx]; header: the PARAMS lines written before it"""
    f = np.ascontiguousarray(field, dtype=np.float64)
    ny, nx = f.shape
    check(lib().fsibe_save_field(os.fsencode(path), f, nx, ny, int(nsave), float(xl), float(yl),
                                 fieldtype.encode(), header.encode() if header is not None else None))


def save_gif(path, field, palette: int = 13, color_min: float | None = None, color_max: float | None = None):
    """SaveGIF (FieldExport.f90:305-384): path + ".gif" from a host field
    [ny][nx], coloured by FargePalette's palette 11, 13 or 14 between
    color_min and color_max (default: the field's own range)"""
    f = np.ascontiguousarray(field, dtype=np.float64)
    ny, nx = f.shape
    rng = color_min is not None and color_max is not None
    check(lib().fsibe_save_gif(os.fsencode(path), f, nx, ny, int(palette), int(rng),
                               float(color_min) if rng else 0.0, float(color_max) if rng else 0.0))


class _Member(FluidSim):
    """a member of a Batch (owned by the batch, not destroyed on its own)"""

    def __init__(self, params: Params, handle):
        self.p = params
        self.nx, self.ny, self.ns = params.nx, params.ny, params.ns
        self.nb = params.iBeam
        self.h = C.c_void_p(handle)

    def close(self):
        self.h = None


class Batch:
    """Independent simulations advanced together: every AB2 step is one
    launch per stage and grid size (fsibe_batch_*).  Member i behaves like a
    FluidSim with params[i] run alone.

    graph: replay captured HIP graphs (True, the library's default) or
    launch every stage (False); autotune: measure the column stages' launch
    choices (True, done at creation by default) or use the single path's
    (False).  Neither changes a result."""

    def __init__(self, params, vortk=None, graph=None, autotune=None):
        n = len(params)
        self.params = list(params)
        arr = (FsibeParams * n)(*[p.to_c() for p in params])
        self._keep = []
        ptrs = (C.c_void_p * n)()
        for i in range(n):
            if vortk is not None and vortk[i] is not None:
                a = np.ascontiguousarray(vortk[i], dtype=np.float64)
                assert a.shape == (params[i].ny, params[i].nx)
                self._keep.append(a)
                ptrs[i] = a.ctypes.data
        h = C.c_void_p()
        check(lib().fsibe_batch_create(arr, n, ptrs, C.byref(h)))
        self.h = h
        self._keep = []
        self.members = [_Member(params[i], lib().fsibe_batch_member(h, i)) for i in range(n)]
        if graph is not None or autotune is not None:
            self.configure(graph, autotune)

    def run(self, nsteps: int):
        check(lib().fsibe_batch_run(self.h, int(nsteps)))

    def configure(self, graph=None, autotune=None):
        """graph / autotune: True, False or None (unchanged); autotune=True
        measures again"""
        g = -1 if graph is None else int(bool(graph))
        a = -1 if autotune is None else int(bool(autotune))
        check(lib().fsibe_batch_configure(self.h, g, a))

    def set_tiles(self, group: int, stage: int, ct: bool, tx: int):
        """pin column stage `stage` (1, 4, 5, 7) of size group `group` to the
        compile-time (ct) or runtime FFT plan with tx columns per tile"""
        check(lib().fsibe_batch_set_tiles(self.h, int(group), int(stage), int(bool(ct)), int(tx)))

    def describe(self) -> dict:
        """groups, launch choices (with the autotuner's times in us) and the
        step mode"""
        import json

        buf = C.create_string_buffer(1 << 16)
        check(lib().fsibe_batch_describe(self.h, buf, len(buf)))
        return json.loads(buf.value.decode())

    def close(self):
        if getattr(self, "h", None):
            lib().fsibe_batch_destroy(self.h)
            self.h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
