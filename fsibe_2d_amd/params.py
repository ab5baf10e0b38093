"""Run parameters (the share_vars values the fluid step reads) and the
workloads the benchmarks and tests use.

Field names follow share_vars.f90:8-25 and the order PARAMS.m lists them
(params.f90:9-81).  nu is the kinematic viscosity (params.f90:50-51 reads the
Reynolds number and stores nu = 1/Re).
"""
from __future__ import annotations

import dataclasses

from ._lib import FsibeParams

MOTION_HEAVE = 600         # y = heave_amp sin(2 pi heave_freq t)  (no reference case)
MOTION_HEAVE_PITCH = 601   # Fourier heave + pitch                   (no reference case)


@dataclasses.dataclass
class Params:
    nx: int = 256
    ny: int = 256
    ns: int = 64
    iMeanVelocity: int = 3
    iSponge: int = 3
    iBeam: int = 1
    iWalls: int = 0
    iCylinder: int = 0
    iSharpTrailing: int = 0
    iFLUSI: int = 0
    iMotion: int = MOTION_HEAVE
    iViscous: int = 1
    iSaveBeam: int = 1
    idealis: int = 1
    xl: float = 4.0
    yl: float = 4.0
    nu: float = 1.0 / 200.0
    eps: float = 1e-3
    eps_sponge: float = 1e-3
    cfl: float = 1.0
    dt_fixed: float = 2e-4
    Time_end: float = 1e6
    N_smooth: float = 2.0
    t_beam: float = 0.0312
    length: float = 1.0
    SpongeSize: float = 0.2
    R_cylinder: float = 0.0
    h_channel: float = 0.0
    x0: float = 1.0
    y0: float = 2.0
    AngleBeam: float = 0.0
    theta_inf: float = 0.0
    T_fullspeed: float = 0.0
    T_release: float = 0.0
    tau: float = 0.0
    U_mean_true: float = 1.0
    heave_amp: float = 0.1
    heave_freq: float = 1.0
    motion_coef: tuple = (0.0,) * 8
    # beam material for the host solid solver (iFLUSI=1): mass ratio,
    # stiffness, damping, gravity, time scheme (5 = BDF2), gravity impulse
    mue: float = 1.0
    eta: float = 0.1
    sigma: float = 0.0
    grav: float = 0.0
    TimeMethodSolid: int = 5
    iImpulse: int = 0
    # 1: the mask's cut-offs and cal_drag's window moved to where the smooth
    # step is below rounding (continuous Force; no reference counterpart)
    iSmoothCutoff: int = 0

    def to_c(self) -> FsibeParams:
        p = FsibeParams()
        for f in dataclasses.fields(self):
            if f.name == "motion_coef":
                for i, c in enumerate(self.motion_coef):
                    p.motion_coef[i] = float(c)
            else:
                setattr(p, f.name, getattr(self, f.name))
        return p

    def replace(self, **kw) -> "Params":
        return dataclasses.replace(self, **kw)

    @property
    def dx(self) -> float:
        return self.xl / self.nx

    @property
    def dy(self) -> float:
        return self.yl / self.ny


def fsi_256(coupled: bool = True) -> Params:
    """256^2, domain 4x4, Re=200, vorticity sponge 0.2 wide around the domain
    (iSponge=3) with inflow U=1 (iMeanVelocity=3), beam of length 1 with 64
    nodes clamped at (1,2), heave 0.1 sin(2 pi t), eta=1e-3, 2 smoothing cells,
    2/3 dealiasing, dt=2e-4.  coupled: the flexible beam (mass ratio 1,
    stiffness 0.1) solved by the solid solver (iFLUSI=1); otherwise the beam
    follows the heave rigidly (iFLUSI=0).  Choices left open by the workload
    description: eps_sponge = eps = 1e-3, t_beam = 0.0312 (PARAMS.m default),
    tau = 0, no structural damping (sigma = 0), BDF2 for the beam."""
    return Params(iFLUSI=1 if coupled else 0, mue=1.0, eta=0.1)


def fsi_large(n: int = 1024, coupled: bool = True) -> Params:
    """The same physics at n^2 (1024 or 2048), Re=1000, 128 beam nodes,
    dt = 5e-5, beam mass ratio 0.5 and stiffness 0.05."""
    return Params(nx=n, ny=n, ns=128, nu=1.0 / 1000.0, dt_fixed=5e-5, iFLUSI=1 if coupled else 0,
                  mue=0.5, eta=0.05)


def turbulence_4096(n: int = 4096) -> Params:
    """Solver-only decaying 2-D turbulence: no mask, no sponge, fluid at rest
    (iMeanVelocity=4), nu=1e-5, dt=1e-4, domain 2 pi x 2 pi."""
    import math

    return Params(nx=n, ny=n, ns=64, iBeam=0, iSponge=0, iSaveBeam=0, iMeanVelocity=4,
                  nu=1e-5, dt_fixed=1e-4, cfl=1e6, xl=2 * math.pi, yl=2 * math.pi, iMotion=0)


def adjoint_512(n: int = 512, smooth_cutoff: bool = True) -> Params:
    """The differentiation workload: n^2, Re = 500, dt = 1e-4 fixed, the beam
    following prescribed rigid heave + pitch (motion 601: two harmonics each,
    base frequency 1), no pressure or stresses; physics otherwise as
    fsi_256().  smooth_cutoff: the mask's cut-offs and cal_drag's window
    moved to where the smooth step is below rounding (iSmoothCutoff = 1), so
    that J is a continuous function of the coefficients; with the reference's
    cut-offs J jumps where a pixel crosses t* or the window's edge moves."""
    return Params(nx=n, ny=n, ns=64, nu=1.0 / 500.0, dt_fixed=1e-4, cfl=100.0, iMotion=MOTION_HEAVE_PITCH,
                  heave_freq=1.0, iSaveBeam=0, iFLUSI=0, iSmoothCutoff=1 if smooth_cutoff else 0)


def adjoint_coefficients(seed: int = 3) -> list:
    """the 8 motion coefficients of the differentiation workload: N(0, 0.05)
    draws of the SplitMix64 / Box-Muller generator"""
    from .inputs import normals

    return [0.05 * float(v) for v in normals(seed, 8)]


def batch_configs(n: int = 64, seed: int = 2, coupled: bool = True) -> list:
    """The batched workload: n independent FSI simulations with uneven grids
    and parameters, drawn per member in this order from SplitMix64(seed)
    uniforms u = (x >> 11) 2^-53: grid in {256, 384, 512, 768}, beam
    stiffness log-uniform in [0.01, 1], mass ratio uniform in [0.2, 2], heave
    frequency uniform in [0.5, 2], Re uniform in [100, 1000], nodes uniform
    in {64, ..., 128}.  Physics otherwise as fsi_256(); dt = 2e-4 (256/N).
    coupled: each member's flexible beam is solved by the solid solver
    (iFLUSI=1); otherwise it follows the heave rigidly."""
    from .inputs import splitmix64

    z = splitmix64(seed, 6 * n)
    u = [float(int(v) >> 11) * (1.0 / 9007199254740992.0) for v in z]
    out = []
    for i in range(n):
        g, eta, mu, f, re, nsu = u[6 * i: 6 * i + 6]
        N = (256, 384, 512, 768)[min(int(g * 4), 3)]
        ns = 64 + min(int(nsu * 65), 64)
        out.append(Params(nx=N, ny=N, ns=ns, nu=1.0 / (100.0 + 900.0 * re), heave_freq=0.5 + 1.5 * f,
                          dt_fixed=2e-4 * 256.0 / N, iFLUSI=1 if coupled else 0,
                          eta=0.01 * 100.0 ** eta, mue=0.2 + 1.8 * mu))
    return out
