"""The oracle's fluid step pinned by closed forms (cal_velocity / cal_nlk /
EvolveFluidExplicit / EvolveFluidAB2, NavierStokes.f90:9-508):

* a two-mode vorticity field whose -u.grad(omega) is a single mode;
* the penalisation term -curl(chi (u - u_s)) for an analytic mask;
* the CFL time step and the Time_end clamp (:59-78);
* Taylor-Green decay (u.grad(omega) = 0) for 50 steps of startup + AB2;
* second-order convergence of AB2 on the advected mode exp(-nu t) cos(x - t).
"""
import numpy as np
import pytest

TWO_PI = 2.0 * np.pi


def _flow(oracle_mod, n=32, **kw):
    d = dict(nx=n, ny=n, xl=TWO_PI, yl=TWO_PI, iBeam=0, iSponge=0, iMeanVelocity=0, nu=0.01,
             dt_fixed=1e-3, cfl=100.0, Time_end=1e6, iSaveBeam=0)
    d.update(kw)
    return oracle_mod.Oracle(**d)


def _packed(n, entries):
    fk = np.zeros((n, n))
    for (iy, ix), v in entries.items():
        fk[iy, ix] += v
    return fk


def test_nonlinear_term_closed_form(oracle_mod):
    """omega = cos x + cos 2y: psi = cos x + cos(2y)/4, u = -sin(2y)/2 + U,
    v = sin x, so -u.grad(omega) = 1.5 sin x sin 2y + U sin x"""
    n = 32
    vk = _packed(n, {(0, 2): 0.5, (4, 0): 0.5})
    for imean, U in ((0, 0.0), (3, 1.0)):
        o = _flow(oracle_mod, n, iMeanVelocity=imean)
        u, _ = o.cal_velocity(0.0, vk)
        X, Y = np.meshgrid(np.arange(n) * TWO_PI / n, np.arange(n) * TWO_PI / n)
        assert np.abs(u[0] - (-np.sin(2 * Y) / 2 + U)).max() < 1e-14
        assert np.abs(u[1] - np.sin(X)).max() < 1e-14
        nlk, _ = o.cal_nlk(0.0, vk, u)
        ref = _packed(n, {(5, 3): 1.5 * 0.25, (0, 3): -0.5 * U})
        assert np.abs(nlk - ref).max() < 1e-14, imean


def test_penalisation_term_closed_form(oracle_mod):
    """chi = c0 + c1 cos x, u_s = 0: curl(chi u) = chi omega + chi_x v
    = c0 cos x + c0 cos 2y + c1 cos 2x + c1 cos x cos 2y (same omega as above);
    cal_nlk adds minus that (NavierStokes.f90:88-132)"""
    n, c0, c1 = 32, 3.0, 1.25
    vk = _packed(n, {(0, 2): 0.5, (4, 0): 0.5})
    o = _flow(oracle_mod, n, iCylinder=1, R_cylinder=0.1)
    X = np.meshgrid(np.arange(n) * TWO_PI / n, np.arange(n) * TWO_PI / n)[0]
    o.set_field("mask", c0 + c1 * np.cos(X))
    o.set_field("maskvx", np.zeros((n, n)))
    o.set_field("maskvy", np.zeros((n, n)))
    u, _ = o.cal_velocity(0.0, vk)
    nlk, _ = o.cal_nlk(0.0, vk, u)
    ref = _packed(n, {(5, 3): 0.375, (0, 2): -0.5 * c0, (4, 0): -0.5 * c0, (0, 4): -0.5 * c1,
                      (4, 2): -0.25 * c1})
    assert np.abs(nlk - ref).max() < 1e-13
    # a uniform solid velocity u_s = (a, b) with constant chi adds nothing: curl(chi u_s) = 0
    o.set_field("mask", np.full((n, n), c0))
    o.set_field("maskvx", np.full((n, n), 0.7))
    o.set_field("maskvy", np.full((n, n), -0.4))
    nlk2, _ = o.cal_nlk(0.0, vk, u)
    ref2 = _packed(n, {(5, 3): 0.375, (0, 2): -0.5 * c0, (4, 0): -0.5 * c0})
    assert np.abs(nlk2 - ref2).max() < 1e-13


def test_cfl_time_step_and_time_end_clamp(oracle_mod):
    n = 64
    o = _flow(oracle_mod, n, dt_fixed=0.0, cfl=0.3, Time_end=2.0)
    vk = o.init_noise(seed=5, amp=1.0, kmax2=36)
    u, _ = o.cal_velocity(0.0, vk)
    umax = np.sqrt(u[0] ** 2 + u[1] ** 2).max()
    _, dt1 = o.cal_nlk(0.0, vk, u)
    assert abs(dt1 - 0.3 * (TWO_PI / n) / umax) <= 1e-15 * dt1
    _, dt_end = o.cal_nlk(2.0 - 1e-5, vk, u)
    assert abs(dt_end - 1e-5) < 1e-16
    # with a fixed dt and penalisation the clamps are min(dt_fixed, 0.9 eps, CFL)
    o2 = _flow(oracle_mod, n, dt_fixed=5e-3, cfl=0.3, iCylinder=1, R_cylinder=0.1, eps=2e-3)
    assert o2.cal_nlk(0.0, vk, u)[1] == min(5e-3, 0.9 * 2e-3)


def test_taylor_green_decay(oracle_mod):
    """modes with one |k|^2 = 25 have psi = omega/25 and u.grad(omega) = 0:
    omega_hat(t) = omega_hat(0) exp(-25 nu t) exactly through the integrating
    factor (cal_vis.f90), startup Euler and AB2 alike"""
    n = 64
    vk = _packed(n, {(8, 6): 0.25 * 1.0, (7, 9): 0.25 * -0.6, (0, 10): 0.5 * 0.8, (11, 0): -0.5 * 0.3})
    o = _flow(oracle_mod, n, nu=0.02, dt_fixed=2e-3)
    o.set_vortk(vk)
    o.run(50)
    d = o.diag()
    assert d.it == 49 and abs(d.time - 0.1) < 1e-14
    ref = vk * np.exp(-0.02 * 25.0 * d.time)
    assert np.abs(o.field("vortk") - ref).max() < 1e-14


def _advected_error(oracle_mod, dt, T=0.5):
    n, nu = 32, 0.01
    o = _flow(oracle_mod, n, nu=nu, iMeanVelocity=3, dt_fixed=dt)
    vk = _packed(n, {(0, 2): 0.5})  # cos x
    o.set_vortk(vk)
    o.run(int(round(T / dt)))
    t = o.diag().time
    exact = _packed(n, {(0, 2): 0.5 * np.cos(t), (0, 3): -0.5 * np.sin(t)}) * np.exp(-nu * t)
    return np.abs(o.field("vortk") - exact).max()


def test_ab2_second_order(oracle_mod):
    """uniform advection by U = 1 of cos x: exact exp(-nu t) cos(x - t); the
    error of startup Euler + AB2 falls by 4 when dt halves"""
    e1, e2, e3 = (_advected_error(oracle_mod, dt) for dt in (0.02, 0.01, 0.005))
    assert 3.6 < e1 / e2 < 4.4 and 3.6 < e2 / e3 < 4.4, (e1, e2, e3)


def test_energy_diagnostics_and_running_mean(oracle_mod):
    """cal_energy (NavierStokes.f90:222-263) and time_avg_vort (:670-683) for
    omega = a cos x after one step: omega = a g cos x, u = (0, a g sin x),
    g = exp(-nu dt); no mask, so the three penalisation terms are 0, and the
    running mean after the first step is the field itself"""
    n, a, nu, dt = 32, 0.7, 0.01, 1e-3
    o = _flow(oracle_mod, n, nu=nu, dt_fixed=dt)
    o.set_vortk(_packed(n, {(0, 2): 0.5 * a}))
    o.run(1)
    e = list(o.diag().energies)
    dxdy, g = (TWO_PI / n) ** 2, np.exp(-nu * dt)
    s2 = a * a * g * g * n * n / 2  # sum of omega^2 = sum of |u|^2 over the grid
    ref = [dxdy * np.sqrt(s2), 0.5 * dxdy * s2, nu * dxdy * s2, 0.0, 0.0, 0.0]
    assert np.allclose(e, ref, rtol=1e-14, atol=0.0)
    X = np.arange(n) * TWO_PI / n
    assert np.abs(o.field("vor_mean") - a * g * np.cos(X)[None, :]).max() < 1e-15


def test_sponge_term_closed_form(oracle_mod):
    """the vorticity sponge adds -F[chi_s omega] (cal_nlk NavierStokes.f90:
    139-164); chi_s = c + d cos x with omega = cos x + cos 2y gives
    chi_s omega = c cos x + c cos 2y + d/2 + d/2 cos 2x + d cos x cos 2y"""
    n, c, d = 32, 2.0, 0.6
    vk = _packed(n, {(0, 2): 0.5, (4, 0): 0.5})
    o = _flow(oracle_mod, n, iSponge=3, eps_sponge=1.0)
    X = np.meshgrid(np.arange(n) * TWO_PI / n, np.arange(n) * TWO_PI / n)[0]
    o.set_field("mask_sponge", c + d * np.cos(X))
    u, _ = o.cal_velocity(0.0, vk)
    nlk, _ = o.cal_nlk(0.0, vk, u)
    ref = _packed(n, {(5, 3): 0.375, (0, 2): -0.5 * c, (4, 0): -0.5 * c, (0, 0): -0.5 * d,
                      (0, 4): -0.25 * d, (4, 2): -0.25 * d})
    assert np.abs(nlk - ref).max() < 1e-13
