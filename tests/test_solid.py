"""The library's host solid solver (SolidSolver.f90; C++ in
fsibe_2d_amd/csrc/solid.hpp, no GPU involved) against the oracle's numpy
restatement (oracle/solid.py) and against what the mechanics fixes.

The reference's solver cannot be built here (it needs MKL's PARDISO and
LAPACK95 modules), so the oracle is pinned by: the complex-step derivative of
its F_nonlinear against its analytic Jacobi (the two routines restated
independently); the static cantilever under uniform load (Euler-Bernoulli:
tip angle q L^3 / (6 EI), tip deflection q L^4 / (8 EI)); the first bending
frequency of the clamped-free beam, 1.8751^2 sqrt(EI / (mu L^4)); a
static equilibrium has zero acceleration in RHS_beameqn; the centrifugal
tension of a rotating straight beam, mu w^2 (L^2 - s^2) / 2; an unloaded
straight beam stays straight."""
import numpy as np
import pytest

NS = 40


def _kin(oracle_mod, **kw):
    d = dict(nx=32, ny=32, ns=NS, iFLUSI=1, iMotion=600, heave_amp=0.1, heave_freq=1.0)
    d.update(kw)
    return oracle_mod.Oracle(oracle_mod.make_params(**d)), d


def _pair(oracle_mod, mat, **kw):
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solid import Solid
    from oracle.solid import SolidOracle

    o, d = _kin(oracle_mod, **kw)
    return Solid(Params(**d, **mat)), SolidOracle(o, **mat)


def _loads(ns, t):
    s = np.linspace(0.0, 1.0, ns)
    return 0.5 * np.sin(2 * np.pi * (s - t)) * (1 + s), 0.02 * np.cos(3 * s + t)


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300)


def _state(b, so, ns, seed):
    """a bent, moving beam with loads, set on both solvers"""
    rng = np.random.default_rng(seed)
    s = np.linspace(0.0, 1.0, ns)
    th = 0.2 * np.sin(2.3 * s) * s + 0.01 * rng.standard_normal(ns) * s
    th[0] = 0.0
    thd = 0.3 * s ** 2 + 0.05 * rng.standard_normal(ns)
    pr, tau = _loads(ns, 0.37)
    oo = np.stack([s, s, s, s, 0.9 * th, 0.8 * thd])
    for name, v in (("theta", th), ("theta_dot", thd), ("pressure_new", pr), ("tau_new", tau),
                    ("pressure_old", 0.9 * pr), ("tau_old", 0.9 * tau), ("oldold", oo)):
        b.set(name, v)
        setattr(so, name, np.array(v))


MATS = {
    "bdf2": dict(mue=0.7, eta=0.08, sigma=0.05, TimeMethodSolid=5),
    "crank_nicolson": dict(mue=1.3, eta=0.2, sigma=0.0, TimeMethodSolid=2, grav=0.3),
    "implicit_euler": dict(mue=0.4, eta=0.03, sigma=0.2, TimeMethodSolid=1),
}


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("motion", [600, 3])
def test_jacobi_is_the_derivative_of_f_nonlinear(name, motion, oracle_mod):
    """the oracle's analytic Jacobi against the complex-step Jacobian of its
    F_nonlinear, and the library's residual / Jacobian / predictor / RHS
    against the oracle's"""
    b, so = _pair(oracle_mod, MATS[name], iMotion=motion, AngleBeam=5.0)
    _state(b, so, NS, 1)
    if MATS[name]["TimeMethodSolid"] == 5:  # the BDF2 system proper, not its Crank-Nicolson start
        b.set_info(dt_old=1.3e-3, grav=0.0, StartupStep=False)
        so.dt_old, so.StartupStep = 1.3e-3, False
    t, dt = 0.41, 1e-3
    x = so.guess(t, dt)
    assert _rel(b.guess(t, dt), x) < 1e-13
    x = x + 1e-3 * np.random.default_rng(2).standard_normal(x.size)
    c = so.context(t, dt)
    J, Jc = so.jacobian(x, c), so.complex_step_jacobian(x, c)
    floor = 1e-9 * np.abs(Jc).max()
    assert np.max(np.abs(J - Jc) / np.maximum(np.abs(Jc), floor)) < 1e-10
    assert np.array_equal(J != 0, Jc != 0) or np.abs(J[(J != 0) != (Jc != 0)]).max() < floor
    F = so.residual(x, c)
    assert _rel(b.residual(t, dt, x), F) < 1e-13
    assert _rel(b.jacobian(t, dt, x), J) < 1e-13
    th_t, th_tt, T = b.rhs(t, so.theta, so.theta_dot, so.pressure_new, so.tau_new)
    o_t, o_tt, oT = so.rhs_beameqn(t, so.theta, so.theta_dot, so.pressure_new, so.tau_new)
    assert _rel(th_t, o_t) == 0.0 and _rel(th_tt, o_tt) < 1e-12 and _rel(T, oT) < 1e-12


@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("motion", [600, 3])
def test_steps_match_the_oracle(name, motion, oracle_mod):
    """40 steps of variable length under travelling loads: every array of
    type(solid) and the Newton iteration count"""
    b, so = _pair(oracle_mod, MATS[name], iMotion=motion, AngleBeam=5.0)
    t, worst = 0.0, 0.0
    for k in range(40):
        dt = 2e-3 * (1.0 + 0.3 * np.sin(k))  # variable steps exercise BDF2's ratio
        pr, tau = _loads(NS, t)
        b.set("pressure_new", pr)
        b.set("tau_new", tau)
        so.pressure_new, so.tau_new = pr.copy(), tau.copy()
        b.step(t, dt)
        so.step(t, dt)
        t += dt
        for k2 in ("theta", "theta_dot", "x", "y", "vx", "vy", "ax", "ay", "pressure_old", "tau_old"):
            worst = max(worst, _rel(b.get(k2), getattr(so, k2)))
        assert _rel(b.get("oldold"), so.oldold) < 1e-11
        assert b.info()["iter"] == so.iter
        e, eo = b.info(), so.energies()
        for k2 in ("E_kinetic", "E_pot", "E_elastic"):
            assert abs(e[k2] - eo[k2]) <= 1e-11 * max(abs(eo[k2]), 1e-12)
    print(f"[solid {name} motion {motion}] worst rel {worst:.1e}")
    assert worst < 1e-11
    assert b.info()["dt_old"] == so.dt_old and not b.info()["StartupStep"]


def test_gravity_impulse_ramp(oracle_mod):
    """iImpulse = 3: grav ramps 0 -> 0.7 over t in [2, 3] (SolidSolver.f90:893-906)"""
    b, so = _pair(oracle_mod, dict(mue=1.0, eta=0.1, TimeMethodSolid=5, iImpulse=3), iMotion=0)
    so.iImpulse = 3
    t = 2.0
    for k in range(30):
        b.step(t, 0.05)
        so.step(t, 0.05)
        assert b.info()["grav"] == pytest.approx(so.grav, rel=1e-14, abs=1e-300)
        t += 0.05
    assert b.info()["grav"] == pytest.approx(0.7)
    assert _rel(b.get("y"), so.y) < 1e-12


def _solid(**kw):
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solid import Solid

    d = dict(nx=32, ny=32, ns=32, iFLUSI=1, iMotion=0, mue=1.0, eta=0.1)
    d.update(kw)
    return Solid(Params(**d))


def test_straight_unloaded_beam_stays_straight():
    b = _solid(TimeMethodSolid=5)
    y0 = b.get("y").copy()
    t = 0.0
    for k in range(20):
        b.step(t, 1e-2)
        t += 1e-2
    assert np.all(b.get("theta") == 0.0) and np.array_equal(b.get("y"), y0)
    assert b.info()["iter"] == 3  # the reference iterates at least three times


def _static_cantilever(ns, q=0.008):
    b = _solid(ns=ns, grav=q, TimeMethodSolid=1)
    t = 0.0
    for k in range(60):  # implicit Euler with long steps relaxes to the static state
        b.step(t, 0.5)
        t += 0.5
    return b


def test_static_cantilever_under_gravity():
    """uniform load q = mue g on a clamped-free beam of length 1, EI = eta"""
    q, EI = 0.008, 0.1
    b = _static_cantilever(32, q)
    th, y = b.get("theta"), b.get("y")
    assert th[-1] == pytest.approx(-q / (6 * EI), rel=1e-3)
    # integrate_position sums theta with the left rectangle rule: first order in ds
    assert y[-1] - 2.0 == pytest.approx(-q / (8 * EI), rel=3.0 / 31)
    b64 = _static_cantilever(64, q)
    assert abs(b64.get("y")[-1] - 2.0 + q / (8 * EI)) < 0.6 * abs(y[-1] - 2.0 + q / (8 * EI))


def test_static_equilibrium_has_zero_acceleration():
    q = 0.008
    b = _static_cantilever(32, q)
    th = b.get("theta")
    z = np.zeros(32)
    _, acc, _ = b.rhs(100.0, th, z, z, z)
    ds = 1.0 / 31
    scale = 0.1 * np.abs(np.diff(th, 2)).max() / ds ** 2 / 1.0  # eta theta_ss / mue
    assert np.abs(acc).max() < 1e-3 * scale


@pytest.mark.parametrize("method,dt", [(5, 2e-3), (2, 2e-3), (1, 2e-4)])
def test_first_bending_frequency(method, dt):
    """free vibration released from the static shape: the period between
    upward zero crossings of the tip against 2 pi / (1.8751^2 sqrt(EI / mu))"""
    th0 = _static_cantilever(32).get("theta")
    b = _solid(TimeMethodSolid=method)
    b.set("theta", th0)
    w = 1.875104068711961 ** 2 * np.sqrt(0.1 / 1.0)
    n = int(2.2 * 2 * np.pi / w / dt)
    t, ys, ts = 0.0, np.empty(n), np.empty(n)
    for k in range(n):
        b.step(t, dt)
        t += dt
        ys[k], ts[k] = b.get("y")[-1] - 2.0, t
    i = np.where((ys[:-1] < 0) & (ys[1:] >= 0))[0]
    tc = ts[i] - ys[i] * (ts[i + 1] - ts[i]) / (ys[i + 1] - ys[i])
    assert len(tc) >= 2
    assert np.diff(tc)[0] == pytest.approx(2 * np.pi / w, rel=3e-3)


def test_centrifugal_tension_of_a_rotating_beam():
    """iMotion = 3 at t = 0: the leading edge turns at w = 80 deg * pi rad/s
    with no angular acceleration; a straight beam carries T = mue w^2 (1 - s^2) / 2"""
    b = _solid(iMotion=3, mue=0.7, ns=33)
    s = np.linspace(0.0, 1.0, 33)
    z = np.zeros(33)
    _, _, T = b.rhs(0.0, z, z, z, z)
    w = 80.0 * np.pi / 180.0 * np.pi
    assert np.abs(T - 0.7 * w ** 2 * (1 - s ** 2) / 2).max() < 1e-9 * 0.7 * w ** 2


def test_held_until_release():
    """time <= T_release: the beam only follows its leading edge
    (SolidSolverWrapper, SolidSolver.f90:20-35)"""
    b = _solid(iMotion=600, T_release=0.05, TimeMethodSolid=5)
    b.set("pressure_new", np.linspace(0, 1, 32))
    t = 0.0
    while t + 1e-2 <= 0.05:
        b.step(t, 1e-2)
        t += 1e-2
    assert np.all(b.get("theta") == 0.0) and b.info()["StartupStep"]
    b.step(0.06, 1e-2)
    assert np.abs(b.get("theta")).max() > 0 and not b.info()["StartupStep"]


def test_emergency_brake_flag():
    """theta_dot changing by more than 100 in one step ends the reference's
    run (SolidSolver.f90:370-380): a beam spinning at up to 120 rad/s under
    strong damping (sigma dt / mue = 10) nearly stops in one implicit step"""
    b = _solid(sigma=1e4, TimeMethodSolid=1)
    s = np.linspace(0.0, 1.0, 32)
    b.set("theta_dot", 120.0 * np.sin(np.pi * s) ** 2)
    b.step(0.0, 1e-3)  # time = T_release: still held
    assert not b.info()["brake"]
    b.step(1e-3, 1e-3)
    assert b.info()["brake"] and np.abs(b.get("theta_dot")).max() < 20.0


def test_newton_failure_is_an_error():
    """no convergence in 49 Newton iterations: the reference stops
    (SolidSolver.f90:311-314), the library returns FSIBE_EDOMAIN"""
    from fsibe_2d_amd import FsibeError

    b = _solid(mue=1e-3, eta=1e-3, TimeMethodSolid=1)
    b.set("pressure_new", np.linspace(0.0, 50.0, 32))
    with pytest.raises(FsibeError, match="did not converge") as e:
        b.step(0.01, 1e-2)
    assert e.value.code == 5


def test_argument_errors():
    from fsibe_2d_amd import FsibeError

    with pytest.raises(FsibeError, match="ns >= 8"):
        _solid(ns=6)
    with pytest.raises(FsibeError, match="TimeMethodSolid"):
        _solid(TimeMethodSolid=3)
    b = _solid()
    with pytest.raises(FsibeError, match="dt must be"):
        b.step(0.0, 0.0)
