"""ctypes wrapper of the CPU oracle (oracle/fsibe_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check in
__graft_entry__.py and by the cpu_baseline leg of bench.py.  The product
package fsibe_2d_amd never imports this module.

Arrays are numpy float64 of shape (ny, nx), C order, so that a[iy, ix] is the
reference's f(ix, iy) (column-major (0:nx-1, 0:ny-1), x fastest).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle_fsibe.so")


class OParams(C.Structure):
    """Mirror of o_params (fsibe_oracle.h)."""

    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("ns", C.c_int32),
        ("iMeanVelocity", C.c_int32), ("iSponge", C.c_int32), ("iBeam", C.c_int32),
        ("iWalls", C.c_int32), ("iCylinder", C.c_int32), ("iSharpTrailing", C.c_int32),
        ("iFLUSI", C.c_int32), ("iMotion", C.c_int32), ("iViscous", C.c_int32),
        ("iSaveBeam", C.c_int32), ("idealis", C.c_int32),
        ("xl", C.c_double), ("yl", C.c_double), ("nu", C.c_double), ("eps", C.c_double),
        ("eps_sponge", C.c_double), ("cfl", C.c_double), ("dt_fixed", C.c_double),
        ("Time_end", C.c_double),
        ("N_smooth", C.c_double), ("t_beam", C.c_double), ("length", C.c_double),
        ("SpongeSize", C.c_double), ("R_cylinder", C.c_double), ("h_channel", C.c_double),
        ("x0", C.c_double), ("y0", C.c_double), ("AngleBeam", C.c_double),
        ("theta_inf", C.c_double), ("T_fullspeed", C.c_double), ("T_release", C.c_double),
        ("tau", C.c_double), ("U_mean_true", C.c_double),
        ("heave_amp", C.c_double), ("heave_freq", C.c_double),
        ("motion_coef", C.c_double * 8),
        ("mue", C.c_double), ("eta", C.c_double), ("sigma", C.c_double), ("grav", C.c_double),
        ("TimeMethodSolid", C.c_int32), ("iImpulse", C.c_int32),
        ("iSmoothCutoff", C.c_int32),
    ]


class OBeam(C.Structure):
    _fields_ = [
        ("x", C.POINTER(C.c_double)), ("y", C.POINTER(C.c_double)),
        ("vx", C.POINTER(C.c_double)), ("vy", C.POINTER(C.c_double)),
        ("theta", C.POINTER(C.c_double)), ("theta_dot", C.POINTER(C.c_double)),
        ("pressure_new", C.POINTER(C.c_double)), ("tau_new", C.POINTER(C.c_double)),
        ("Force", C.c_double * 2), ("Force_unst", C.c_double * 2),
        ("Force_press", C.c_double * 2),
        ("x0", C.c_double), ("y0", C.c_double), ("AngleBeam", C.c_double),
        ("phase", C.c_double),
        ("drag_unst_new", C.c_double), ("drag_unst_old", C.c_double),
        ("lift_unst_new", C.c_double), ("lift_unst_old", C.c_double),
        ("UnsteadyCorrectionsReady", C.c_int32), ("iMouvement", C.c_int32),
    ]


class ODiag(C.Structure):
    _fields_ = [
        ("time", C.c_double), ("dt0", C.c_double), ("dt1", C.c_double),
        ("u_max", C.c_double), ("energies", C.c_double * 6), ("bsi", C.c_double * 11),
        ("fx_int", C.c_double), ("fy_int", C.c_double),
        ("n0", C.c_int32), ("it", C.c_int32), ("started", C.c_int32), ("pad", C.c_int32),
        ("u_mean", C.c_double * 2),
    ]


F = {
    "vortk": 0, "nlk": 1, "workvis": 2, "u1": 3, "u2": 4, "press": 5, "mask": 6,
    "maskvx": 7, "maskvy": 8, "mask_sponge": 9, "vor_mean": 10, "dealiase": 11,
    "stress_a": 12, "stress_b": 13, "vort": 14,
}

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
        vp = C.c_void_p
        L.o_sim_create.restype = vp
        L.o_sim_create.argtypes = [C.POINTER(OParams)]
        L.o_sim_destroy.argtypes = [vp]
        L.o_set_threads.argtypes = [C.c_int]
        for name in ("o_coftxy", "o_cofitxy", "o_rfft2", "o_irfft2", "o_cofdx", "o_cofdy",
                     "o_cofdxdx", "o_cofdydy", "o_cofdxdy", "o_poisson"):
            getattr(L, name).argtypes = [vp, dp, dp]
        L.o_cfft1.argtypes = [C.c_int, dp, C.c_int]
        L.o_vis.argtypes = [vp, C.c_double, dp]
        L.o_dealiase_mask.argtypes = [vp, dp]
        L.o_create_mask.argtypes = [vp, C.c_double, C.POINTER(OBeam)]
        L.o_create_sponge_mask.argtypes = [vp]
        L.o_pressure_snapshot.argtypes = [vp, C.c_double, C.c_int, dp]
        L.o_cal_velocity.argtypes = [vp, C.c_double, dp, dp, dp]
        L.o_cal_nlk.argtypes = [vp, C.c_double, C.POINTER(C.c_double), dp, dp, dp]
        L.o_get_beam.restype = C.POINTER(OBeam)
        L.o_get_beam.argtypes = [vp]
        L.o_get_beam_k.restype = C.POINTER(OBeam)
        L.o_get_beam_k.argtypes = [vp, C.c_int]
        L.o_nbeams.restype = C.c_int
        L.o_nbeams.argtypes = [vp]
        L.o_get_beam_bsi.argtypes = [vp, C.c_int, dp]
        L.o_mouvement.argtypes = [vp, C.c_double, C.POINTER(OBeam), C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), dp]
        L.o_integrate_position.argtypes = [vp, C.c_double, C.POINTER(OBeam)]
        L.o_mean_ux.restype = C.c_double
        L.o_mean_ux.argtypes = [vp, C.c_double]
        L.o_mean_uy.restype = C.c_double
        L.o_mean_uy.argtypes = [vp, C.c_double]
        L.o_smoothstep.argtypes = [vp, C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double]
        L.o_delta_phi4star.restype = C.c_double
        L.o_delta_phi4star.argtypes = [C.c_double, C.c_double, C.c_double]
        L.o_delta_interp.restype = C.c_double
        L.o_delta_interp.argtypes = [vp, C.c_double, C.c_double, dp, C.POINTER(C.c_int)]
        L.o_get_forces.argtypes = [vp, C.c_double, C.POINTER(OBeam), dp, dp]
        L.o_surface_integral.argtypes = [vp, C.c_double, dp, dp, dp, dp]
        L.o_sponge10_term.argtypes = [vp, dp, dp]
        L.o_cal_drag.argtypes = [vp, C.c_double, C.c_double, dp, C.POINTER(OBeam)]
        L.o_set_vortk.argtypes = [vp, dp]
        L.o_host_create_mask.argtypes = [vp, C.c_double]
        L.o_host_evolve.restype = C.c_double
        L.o_host_evolve.argtypes = [vp, C.c_double, C.c_double, C.c_int]
        L.o_host_get_forces.argtypes = [vp, C.c_double]
        L.o_host_swap.argtypes = [vp]
        L.o_run.argtypes = [vp, C.c_int]
        L.o_startup_step.argtypes = [vp]
        L.o_ab2_step.argtypes = [vp]
        L.o_get_field.argtypes = [vp, C.c_int, dp]
        L.o_set_field.argtypes = [vp, C.c_int, dp]
        L.o_get_diag.argtypes = [vp, C.POINTER(ODiag)]
        L.o_normals.argtypes = [C.c_uint64, C.c_int, dp]
        L.o_init_noise.argtypes = [vp, C.c_uint64, C.c_double, C.c_int, dp]
        L.o_pi.restype = C.c_double
        L.o_pi.argtypes = [vp]
        _lib = L
    return _lib


def make_params(**kw) -> OParams:
    """Defaults: the 256^2 workload (domain 4x4, Re=200,
    sponge 3, beam of 64 nodes at (1,2), heave 0.1 sin(2 pi t))."""
    d = dict(
        nx=256, ny=256, ns=64, iMeanVelocity=3, iSponge=3, iBeam=1, iWalls=0,
        iCylinder=0, iSharpTrailing=0, iFLUSI=0, iMotion=600, iViscous=1, iSaveBeam=1,
        idealis=1, xl=4.0, yl=4.0, nu=1.0 / 200.0, eps=1e-3, eps_sponge=1e-3, cfl=1.0,
        dt_fixed=2e-4, Time_end=1e6, N_smooth=2.0, t_beam=0.0312, length=1.0,
        SpongeSize=0.2, R_cylinder=0.0, h_channel=0.0, x0=1.0, y0=2.0, AngleBeam=0.0,
        theta_inf=0.0, T_fullspeed=0.0, T_release=0.0, tau=0.0, U_mean_true=1.0,
        heave_amp=0.1, heave_freq=1.0,
    )
    coef = kw.pop("motion_coef", None)
    d.update(kw)
    p = OParams()
    for k, v in d.items():
        setattr(p, k, v)
    if coef is not None:
        for i, c in enumerate(coef):
            p.motion_coef[i] = c
    return p


class Oracle:
    """One oracle simulation (share_vars globals + time_step locals)."""

    def __init__(self, params: OParams | None = None, **kw):
        self.L = lib()
        self.p = params if params is not None else make_params(**kw)
        self.nx, self.ny = self.p.nx, self.p.ny
        self.h = self.L.o_sim_create(C.byref(self.p))

    def close(self):
        if self.h:
            self.L.o_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _new(self):
        return np.zeros((self.ny, self.nx), dtype=np.float64)

    def _as(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (self.ny, self.nx), a.shape
        return a

    # ---- FFT plugin and operators ------------------------------------------
    def _op(self, name, a):
        out = self._new()
        getattr(self.L, name)(self.h, self._as(a), out)
        return out

    def coftxy(self, f):
        return self._op("o_coftxy", f)

    def cofitxy(self, fk):
        return self._op("o_cofitxy", fk)

    def cofdx(self, fk):
        return self._op("o_cofdx", fk)

    def cofdy(self, fk):
        return self._op("o_cofdy", fk)

    def cofdxdx(self, fk):
        return self._op("o_cofdxdx", fk)

    def cofdydy(self, fk):
        return self._op("o_cofdydy", fk)

    def cofdxdy(self, fk):
        return self._op("o_cofdxdy", fk)

    def poisson(self, fk):
        return self._op("o_poisson", fk)

    def vis(self, dt):
        out = self._new()
        self.L.o_vis(self.h, dt, out)
        return out

    def dealiase_mask(self):
        out = self._new()
        self.L.o_dealiase_mask(self.h, out)
        return out

    def rfft2(self, f):
        """Unnormalised r2c half spectrum, complex array (ny, nx/2+1)."""
        out = np.zeros((self.ny, 2 * (self.nx // 2 + 1)), dtype=np.float64)
        self.L.o_rfft2(self.h, self._as(f), out)
        return out[:, 0::2] + 1j * out[:, 1::2]

    def irfft2(self, ft):
        buf = np.zeros((self.ny, 2 * (self.nx // 2 + 1)), dtype=np.float64)
        buf[:, 0::2] = ft.real
        buf[:, 1::2] = ft.imag
        out = self._new()
        self.L.o_irfft2(self.h, buf, out)
        return out

    # ---- geometry / kinematics ---------------------------------------------
    @property
    def beam(self):
        return self.L.o_get_beam(self.h).contents

    @property
    def nbeams(self):
        return self.L.o_nbeams(self.h)

    def _beam_ptr(self, ib):
        if not 0 <= ib < max(self.nbeams, 1):
            raise IndexError(f"beam {ib} of {self.nbeams}")
        return self.L.o_get_beam_k(self.h, ib) if self.nbeams else self.L.o_get_beam(self.h)

    def beam_arrays(self, ib=0):
        b = self._beam_ptr(ib).contents
        ns = self.p.ns
        g = lambda ptr: np.ctypeslib.as_array(ptr, shape=(ns,)).copy()
        return dict(x=g(b.x), y=g(b.y), vx=g(b.vx), vy=g(b.vy), theta=g(b.theta),
                    theta_dot=g(b.theta_dot), pressure=g(b.pressure_new), tau=g(b.tau_new),
                    Force=tuple(b.Force), Force_unst=tuple(b.Force_unst),
                    Force_press=tuple(b.Force_press), x0=b.x0, y0=b.y0, phase=b.phase)

    def set_beam_arrays(self, ib=0, **arrs):
        b = self._beam_ptr(ib).contents
        ns = self.p.ns
        for k, v in arrs.items():
            dst = np.ctypeslib.as_array(getattr(b, k), shape=(ns,))
            dst[:] = np.asarray(v, dtype=np.float64)

    def beam_bsi(self, ib=0):
        """BeamSurfaceIntegral's 10 values of beam ib from the last
        get_forces, then GetForcesInterp's fx_int, fy_int"""
        out = np.zeros(12)
        self.L.o_get_beam_bsi(self.h, ib, out)
        return out

    def mouvement(self, time, ib=0):
        a, at, att = C.c_double(), C.c_double(), C.c_double()
        le = np.zeros(6)
        self.L.o_mouvement(self.h, time, self._beam_ptr(ib), C.byref(a), C.byref(at), C.byref(att), le)
        return a.value, at.value, att.value, le

    def integrate_position(self, time, ib=None):
        """integrate_position of beam ib, or of every beam (ib None)"""
        for k in (range(max(self.nbeams, 1)) if ib is None else (ib,)):
            self.L.o_integrate_position(self.h, time, self._beam_ptr(k))

    def beam_view(self, ib):
        """the kinematics interface of one beam (for SolidOracle)"""
        return BeamView(self, ib)

    def create_mask(self, time):
        self.L.o_create_mask(self.h, time, self.L.o_get_beam(self.h))

    def create_sponge_mask(self):
        self.L.o_create_sponge_mask(self.h)

    def smoothstep(self, x, t, h):
        f = C.c_double()
        self.L.o_smoothstep(self.h, C.byref(f), x, t, h)
        return f.value

    def delta_interp(self, x, y, field):
        err = C.c_int(0)
        v = self.L.o_delta_interp(self.h, x, y, self._as(field), C.byref(err))
        if err.value:
            raise ValueError("DeltaInterpolation: target outside field")
        return v

    def get_forces(self, time, press, u):
        """GetForces (BeamForces.f90:6-51) on the given pressure and velocity
        (u: (2, ny, nx)); results land in the beam (pressure, tau, Force_press)."""
        self.L.o_get_forces(self.h, time, self.L.o_get_beam(self.h), self._as(press),
                            np.ascontiguousarray(u, dtype=np.float64))

    def surface_integral(self, time, u, stress_a, stress_b, press):
        """BeamSurfaceIntegral (BeamForces.f90:306-509) of the held beam on
        the given fields; returns the 10 values of diag().bsi"""
        self.L.o_surface_integral(self.h, time, np.ascontiguousarray(u, dtype=np.float64), self._as(stress_a),
                                  self._as(stress_b), self._as(press))
        return np.array(list(self.diag().bsi)[:10])

    def sponge10_term(self, u1):
        """the iSponge = 10 increment of the x-penalisation
        (NavierStokes.f90:103-121) for the velocity u1 and the held
        mask_sponge"""
        out = np.zeros((self.ny, self.nx))
        self.L.o_sponge10_term(self.h, self._as(u1), out)
        return out

    def cal_drag(self, time, dt1, penal):
        """cal_drag (NavierStokes.f90:512-593) of the held beam and masks on
        penal (2, ny, nx); Force / Force_unst land in the beam"""
        self.L.o_cal_drag(self.h, time, dt1, np.ascontiguousarray(penal, dtype=np.float64),
                          self.L.o_get_beam(self.h))

    # per-step calls for a host that moves the beam itself (iFLUSI=1)
    def host_create_mask(self, time):
        self.L.o_host_create_mask(self.h, time)

    def host_evolve(self, time, dt0=0.0, explicit=False):
        return self.L.o_host_evolve(self.h, time, dt0, int(explicit))

    def host_get_forces(self, time):
        self.L.o_host_get_forces(self.h, time)

    def host_swap(self):
        self.L.o_host_swap(self.h)

    # ---- solver ---------------------------------------------------------------
    def set_vortk(self, vk):
        self.L.o_set_vortk(self.h, self._as(vk))

    def run(self, nsteps):
        self.L.o_run(self.h, int(nsteps))

    def field(self, name):
        out = self._new()
        self.L.o_get_field(self.h, F[name], out)
        return out

    def set_field(self, name, a):
        self.L.o_set_field(self.h, F[name], self._as(a))

    def diag(self):
        d = ODiag()
        self.L.o_get_diag(self.h, C.byref(d))
        return d

    def init_noise(self, seed=0, amp=1e-3, kmax2=64):
        out = self._new()
        self.L.o_init_noise(self.h, seed, amp, kmax2, out)
        return out

    def cal_velocity(self, time, vk):
        u = np.zeros((2, self.ny, self.nx))
        st = self._new()
        self.L.o_cal_velocity(self.h, time, self._as(vk), u, st)
        return u, st

    def pressure_snapshot(self, time, level):
        out = self._new()
        self.L.o_pressure_snapshot(self.h, C.c_double(time), int(level), out)
        return out

    def cal_nlk(self, time, vk, u):
        nlk = self._new()
        dt1 = C.c_double()
        self.L.o_cal_nlk(self.h, time, C.byref(dt1), nlk, self._as(vk),
                         np.ascontiguousarray(u, dtype=np.float64))
        return nlk, dt1.value


class BeamView:
    """One beam of an Oracle, with the kinematics interface SolidOracle
    reads (p, beam_arrays, set_beam_arrays, mouvement, integrate_position)."""

    def __init__(self, o, ib):
        self.o, self.ib, self.p = o, ib, o.p

    def beam_arrays(self):
        return self.o.beam_arrays(self.ib)

    def set_beam_arrays(self, **arrs):
        self.o.set_beam_arrays(self.ib, **arrs)

    def mouvement(self, time):
        return self.o.mouvement(time, self.ib)

    def integrate_position(self, time):
        self.o.integrate_position(time, self.ib)


def normals(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.float64)
    lib().o_normals(seed, n, out)
    return out


def set_threads(n: int):
    lib().o_set_threads(int(n))
