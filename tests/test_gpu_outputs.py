"""ComputePressureSnapshot on the GPU (fsibe_pressure_snapshot) against the
oracle's restatement (NavierStokes.f90:596-665), for both stored vorticity
levels, and the simulation state left alone by it."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

PKEYS = ("nx", "ny", "ns", "iMeanVelocity", "iSponge", "iBeam", "iWalls", "iCylinder", "iSharpTrailing",
         "iFLUSI", "iMotion", "iViscous", "iSaveBeam", "idealis", "xl", "yl", "nu", "eps", "eps_sponge", "cfl",
         "dt_fixed", "Time_end", "N_smooth", "t_beam", "length", "SpongeSize", "R_cylinder", "h_channel", "x0",
         "y0", "AngleBeam", "theta_inf", "T_fullspeed", "T_release", "tau", "U_mean_true", "heave_amp",
         "heave_freq", "motion_coef", "iSmoothCutoff")


@pytest.mark.parametrize("kw", [dict(), dict(iSaveBeam=0, iMeanVelocity=11), dict(iWalls=1, h_channel=0.4,
                                                                                   U_mean_true=0.8)])
def test_pressure_snapshot(oracle_mod, kw):
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solver import FluidSim

    p = Params(nx=96, ny=64, ns=32, dt_fixed=2e-4, **kw)
    o = oracle_mod.Oracle(oracle_mod.make_params(**{k: getattr(p, k) for k in PKEYS}))
    vk = o.init_noise(seed=2, amp=1e-3, kmax2=64)
    o.set_vortk(vk)
    g = FluidSim(p)
    g.set_vortk(vk)
    o.run(9)
    g.run(9)
    t = g.time_state()[0]
    for level in (0, 1):
        a, b = g.pressure_snapshot(t, level), o.pressure_snapshot(t, level)
        e = rel_l2(a, b)
        print(f"[snapshot {kw} level {level}] rel L2 {e:.1e}", flush=True)
        assert e <= 1e-10
    o.run(3)
    g.run(3)
    assert rel_l2(g.field("vortk"), o.field("vortk")) <= 1e-10
    g.close()
