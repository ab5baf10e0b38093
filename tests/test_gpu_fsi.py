"""Fluid-structure interaction through fsibe_run with iFLUSI = 1: the fluid
step on the GPU and the flexible beam solved by the library's host solid
solver, against the oracle's fluid step and the oracle's solid solver in the
same loop (time_step.f90:120-230: create_mask -> evolve -> GetForces ->
GravityImpulse + SolidSolverWrapper -> swap -> time += dt1)."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

PKEYS = ("nx", "ny", "ns", "iMeanVelocity", "iSponge", "iBeam", "iWalls", "iCylinder", "iSharpTrailing",
         "iFLUSI", "iMotion", "iViscous", "iSaveBeam", "idealis", "xl", "yl", "nu", "eps", "eps_sponge", "cfl",
         "dt_fixed", "Time_end", "N_smooth", "t_beam", "length", "SpongeSize", "R_cylinder", "h_channel", "x0",
         "y0", "AngleBeam", "theta_inf", "T_fullspeed", "T_release", "tau", "U_mean_true", "heave_amp",
         "heave_freq", "motion_coef", "iSmoothCutoff")


def oracle_fsi(oracle_mod, p, vk, nsteps, history=None):
    """the coupled loop on the oracle; history (a list) receives per step
    (time, tip y, Newton iterations, Force)"""
    from oracle.solid import SolidOracle

    o = oracle_mod.Oracle(oracle_mod.make_params(**{k: getattr(p, k) for k in PKEYS}))
    so = SolidOracle(o, mue=p.mue, eta=p.eta, sigma=p.sigma, grav=p.grav, TimeMethodSolid=p.TimeMethodSolid,
                     iImpulse=p.iImpulse)
    o.set_vortk(vk)
    time, dt0 = 0.0, 0.0
    for step in range(nsteps):
        o.set_beam_arrays(x=so.x, y=so.y, vx=so.vx, vy=so.vy, theta=so.theta, theta_dot=so.theta_dot)
        o.host_create_mask(time)
        dt1 = o.host_evolve(time, explicit=True) if step == 0 else o.host_evolve(time, dt0)
        o.host_get_forces(time)
        b = o.beam_arrays()
        so.pressure_new, so.tau_new = b["pressure"].copy(), b["tau"].copy()
        so.step(time, dt1)
        o.host_swap()
        time, dt0 = time + dt1, dt1
        if history is not None:
            history.append((time, so.y[-1], so.iter, tuple(o.beam_arrays()["Force"])))
    return o, so, time


def _floor_close(a, b, rel):
    a, b = np.asarray(a, float), np.asarray(b, float)
    floor = 1e-6 * max(np.abs(b).max(), 1e-300)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)) <= rel


def _compare(g, o, so, t_o, tag, unst_rel=1e-9):
    out = g.step_out()
    t_g = g.time_state()[0]
    gb, ob = g.beam(), o.beam_arrays()
    e_v = rel_l2(g.field("vortk"), o.field("vortk"))
    e_th = np.abs(gb["theta"] - so.theta).max() / max(np.abs(so.theta).max(), 1e-300)
    e_tip = abs(gb["y"][-1] - so.y[-1]) / abs(so.y[-1] - g.p.y0)
    print(f"[{tag}] vortk {e_v:.1e} theta {e_th:.1e} tip deflection {e_tip:.1e} "
          f"tip {so.y[-1] - g.p.y0:.3e} iters {out['ibes_iter']}/{so.iter}", flush=True)
    assert t_g == pytest.approx(t_o, rel=1e-14)
    assert e_v <= 1e-10
    assert e_th <= 1e-9 and e_tip <= 1e-9
    assert np.abs(gb["x"] - so.x).max() <= 1e-12 * np.abs(so.x).max()
    assert out["ibes_iter"] == so.iter
    for k in ("Force", "Force_press"):
        assert _floor_close(out[k], ob[k], 1e-9), (k, out[k], ob[k])
    assert _floor_close(out["Force_unst"], ob["Force_unst"], unst_rel)
    s = g.solid
    assert np.array_equal(s.get("theta"), gb["theta"]) and np.array_equal(s.get("y"), gb["y"])
    assert np.abs(s.get("theta_dot") - so.theta_dot).max() <= 1e-8 * max(np.abs(so.theta_dot).max(), 1e-300)


CASES = {
    "bdf2": dict(),
    "crank_nicolson_damped": dict(TimeMethodSolid=2, sigma=0.5, heave_freq=1.5),
    "implicit_euler_gravity": dict(TimeMethodSolid=1, grav=0.5, mue=0.6, eta=0.05, AngleBeam=10.0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_coupled_loop(name, oracle_mod):
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solver import FluidSim

    kw = dict(nx=128, ny=128, ns=48, iFLUSI=1, mue=1.0, eta=0.1, dt_fixed=2e-4)
    kw.update(CASES[name])
    p = Params(**kw)
    o0 = oracle_mod.Oracle(nx=p.nx, ny=p.ny, iBeam=0, iSponge=0)
    vk = o0.init_noise(seed=3, amp=1e-3, kmax2=64)
    nsteps = 25
    o, so, t_o = oracle_fsi(oracle_mod, p, vk, nsteps)
    g = FluidSim(p)
    g.set_vortk(vk)
    g.run(nsteps)
    _compare(g, o, so, t_o, name)
    assert np.abs(so.theta).max() > 1e-6  # the beam bends
    g.close()


def test_coupled_loop_long_horizon_cfl(oracle_mod):
    """800 coupled steps at 128^2 with the time step chosen by the CFL bound
    each step (dt_fixed = 0: dt1 from the device's u_max), so that the
    rounding of every stage and of the step sizes accumulates over a long
    horizon: the time, vortk, theta and the tip stay within the bars of
    the short runs"""
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solver import FluidSim

    p = Params(nx=128, ny=128, ns=48, iFLUSI=1, mue=1.0, eta=0.1, dt_fixed=0.0, cfl=0.4)
    o0 = oracle_mod.Oracle(nx=p.nx, ny=p.ny, iBeam=0, iSponge=0)
    vk = o0.init_noise(seed=5, amp=1e-3, kmax2=64)
    nsteps = 800
    o, so, t_o = oracle_fsi(oracle_mod, p, vk, nsteps)
    g = FluidSim(p)
    g.set_vortk(vk)
    g.run(nsteps)
    _compare(g, o, so, t_o, "long horizon, CFL dt")
    assert t_o > 0.1 and np.abs(so.theta).max() > 1e-4  # a long run, and the beam bends
    g.close()


def test_geometry_gate_timeout_is_an_error(monkeypatch):
    """the coupled loop queues the next step's geometry copy before the
    beam solve, behind a bounded wait on a gate in host memory that the
    host opens once the record is written.  With the bound set to 0
    wall-clock ticks (FSIBE_GATE_TICKS, read when a simulation is created)
    the copy stops waiting before the host opens the gate (the 128-node
    beam solve takes longer than the 64^2 kernels ahead of the copy), and
    the loop reports the stale record as an error instead of stepping on
    with it; every kernel still ends normally"""
    from fsibe_2d_amd import FsibeError, Params
    from fsibe_2d_amd.solver import FluidSim

    monkeypatch.setenv("FSIBE_GATE_TICKS", "0")
    p = Params(nx=64, ny=64, ns=128, iFLUSI=1, dt_fixed=2e-4)
    g = FluidSim(p)
    with pytest.raises(FsibeError, match="stale"):
        g.run(10)
    g.close()


def test_coupled_loop_in_two_calls(oracle_mod):
    """fsibe_run(10) then fsibe_run(15) is the same run as fsibe_run(25)"""
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solver import FluidSim

    p = Params(nx=96, ny=64, ns=40, iFLUSI=1, dt_fixed=2e-4)
    o0 = oracle_mod.Oracle(nx=p.nx, ny=p.ny, iBeam=0, iSponge=0)
    vk = o0.init_noise(seed=4, amp=1e-3, kmax2=64)
    a, b = FluidSim(p), FluidSim(p)
    a.set_vortk(vk)
    b.set_vortk(vk)
    a.run(25)
    b.run(10)
    b.run(15)
    assert a.time_state() == b.time_state()
    assert np.array_equal(a.field("vortk"), b.field("vortk"))
    assert np.array_equal(a.beam()["y"], b.beam()["y"])


@pytest.mark.parametrize("n,steps", [(1024, 3), (2048, 2)])
def test_full_size_coupled(n, steps, oracle_mod):
    """the benchmark workload (1024^2, Re 1000, 128 nodes, mass ratio 0.5,
    stiffness 0.05, dt 5e-5) coupled, and the same at 2048^2; Force_unst
    keeps the 1e-5 bar of its conditioning at these sizes (a one-step
    difference of a sum over ~4000 pixels divided by dt)"""
    from fsibe_2d_amd import fsi_large
    from fsibe_2d_amd.solver import FluidSim

    p = fsi_large(n)
    o0 = oracle_mod.Oracle(nx=p.nx, ny=p.ny, iBeam=0, iSponge=0)
    vk = o0.init_noise(seed=0, amp=1e-3, kmax2=64)
    o, so, t_o = oracle_fsi(oracle_mod, p, vk, steps)
    g = FluidSim(p)
    g.set_vortk(vk)
    g.run(steps)
    _compare(g, o, so, t_o, f"fsi coupled {n}", unst_rel=1e-5)
    g.close()


def test_autotuned_tiles_do_not_change_results(oracle_mod, monkeypatch):
    """fsibe_run measures the column stages' launch choices at the first run
    of a grid; FSIBE_AUTOTUNE=0 keeps the static ones.  A coupled run with
    each gives bit-identical fields, loads and beam (the choices differ only
    in tiles per block, never in the arithmetic)"""
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solver import FluidSim

    p = Params(nx=384, ny=384, ns=48, iFLUSI=1, mue=0.7, eta=0.08, dt_fixed=1e-4, nu=2e-3)
    o = oracle_mod.Oracle(nx=p.nx, ny=p.ny, iBeam=0, iSponge=0, xl=p.xl, yl=p.yl)
    vk = o.init_noise(seed=3, amp=1e-3, kmax2=64)
    res = []
    for tune in ("0", "1"):
        monkeypatch.setenv("FSIBE_AUTOTUNE", tune)
        g = FluidSim(p)
        g.set_vortk(vk)
        out = g.run(6)
        res.append((g.field("vortk"), out["Force"], out["Force_press"], g.beam()["theta"]))
        g.close()
    a, b = res
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1] and a[2] == b[2]
    assert np.array_equal(a[3], b[3])


def test_workload_256_coupled_100_steps(oracle_mod):
    """fsi_256(): 256^2, Re 200, the flexible beam of 64 nodes (mass ratio 1,
    stiffness 0.1, BDF2) clamped at (1, 2) and heaving 0.1 sin(2 pi t),
    band-limited noise start (seed 0), 100 coupled steps.  Every step: the
    tip's y and cal_drag's Force within 1e-9 (relative, floor 1e-6 of the
    run's maximum) and equal Newton iteration counts; after the last step
    the vorticity within 1e-10 and the full comparison."""
    from fsibe_2d_amd import fsi_256
    from fsibe_2d_amd.solver import FluidSim

    p = fsi_256()
    o0 = oracle_mod.Oracle(nx=p.nx, ny=p.ny, iBeam=0, iSponge=0)
    vk = o0.init_noise(seed=0, amp=1e-3, kmax2=64)
    nsteps = 100
    hist = []
    o, so, t_o = oracle_fsi(oracle_mod, p, vk, nsteps, history=hist)
    g = FluidSim(p)
    g.set_vortk(vk)
    tips, forces = [], []
    for step in range(nsteps):
        out = g.run(1)
        t_g = g.time_state()[0]
        assert t_g == pytest.approx(hist[step][0], rel=1e-14)
        assert out["ibes_iter"] == hist[step][2], (step, out["ibes_iter"], hist[step][2])
        tips.append(g.beam()["y"][-1])
        forces.append(out["Force"])
    tip_o = np.array([h[1] for h in hist]) - p.y0
    tip_g = np.array(tips) - p.y0
    assert _floor_close(tip_g, tip_o, 1e-9), np.abs(tip_g - tip_o).max()
    f_o = np.array([h[3] for h in hist])
    f_g = np.array(forces)
    for c in range(2):
        assert _floor_close(f_g[:, c], f_o[:, c], 1e-9), (c, np.abs(f_g[:, c] - f_o[:, c]).max())
    print(f"[fsi_256 x100] tip deflection {tip_o[-1]:.4e} max tip err {np.abs(tip_g - tip_o).max():.1e}",
          flush=True)
    assert abs(so.y[-1] - so.y[0]) > 1e-5  # the beam bends: its tip lags the clamped end
    _compare(g, o, so, t_o, "fsi_256 coupled x100")
    g.close()
