"""Beam loads of the oracle pinned by closed forms: the moment conditions of
the smoothed 4-point kernel phi* (Interpolation.f90:82-91, Yang et al. JCP
2009), exact reproduction of linear fields by DeltaInterpolation
(Interpolation.f90:6-44), the pressure jump of a linear pressure field across
a straight beam and the shear stress of a sinusoidal flow (GetForcesInterp,
BeamForces.f90:63-288)."""
import numpy as np
import pytest


def test_phi4star_moments(oracle_mod):
    L = oracle_mod.lib()
    pi = np.pi
    rng = np.random.default_rng(0)
    for x in rng.uniform(0.0, 1.0, 50):
        w = np.array([L.o_delta_phi4star(x - i, 1.0, pi) for i in range(-4, 5)])
        r = x - np.arange(-4, 5)
        assert abs(w.sum() - 1.0) < 1e-14
        assert abs((r * w).sum()) < 1e-14
        assert np.all(w >= -1e-15)
    # support |r| <= 2.5, symmetric
    assert L.o_delta_phi4star(2.5001, 1.0, pi) == 0.0
    assert L.o_delta_phi4star(0.7, 1.0, pi) == L.o_delta_phi4star(-0.7, 1.0, pi)


def _oracle(oracle_mod, **kw):
    d = dict(nx=64, ny=64, ns=16, iBeam=1, iSponge=0, iMotion=0, x0=1.0, y0=2.0, length=1.5,
             T_release=0.0, tau=0.0, iViscous=1, nu=1e-2)
    d.update(kw)
    return oracle_mod.Oracle(**d)


def test_delta_interp_reproduces_linear_fields(oracle_mod):
    o = _oracle(oracle_mod)
    X, Y = np.meshgrid(np.arange(64) * (4.0 / 64), np.arange(64) * (4.0 / 64))
    f = 0.3 - 1.7 * X + 2.2 * Y
    rng = np.random.default_rng(1)
    for x, y in rng.uniform(0.5, 3.4, size=(40, 2)):
        assert abs(o.delta_interp(x, y, f) - (0.3 - 1.7 * x + 2.2 * y)) < 1e-13
    with pytest.raises(ValueError):
        o.delta_interp(0.05, 2.0, f)  # stencil leaves the field (Interpolation.f90:19-24)


@pytest.mark.parametrize("angle", [0.0, 30.0, -55.0])
def test_pressure_jump_of_linear_pressure(oracle_mod, angle):
    """p = a x + b y: every node carries 2 t_beam (grad p . n), n = (-sin, cos)"""
    ns, t = 16, 0.0312
    o = _oracle(oracle_mod, ns=ns, AngleBeam=angle, t_beam=t)
    th = np.deg2rad(angle)
    s = np.arange(ns) * (1.5 / (ns - 1))
    o.set_beam_arrays(x=1.2 + s * np.cos(th), y=1.7 + s * np.sin(th), theta=np.zeros(ns))
    X, Y = np.meshgrid(np.arange(64) * (4.0 / 64), np.arange(64) * (4.0 / 64))
    a, b = 0.8, -1.9
    o.get_forces(1.0, a * X + b * Y, np.zeros((2, 64, 64)))
    r = o.beam_arrays()
    dp = 2.0 * t * (-a * np.sin(th) + b * np.cos(th))
    assert np.abs(r["pressure"] - dp).max() < 1e-12
    assert np.abs(r["tau"]).max() == 0.0
    ds = 1.5 / (ns - 1)
    fx, fy = r["Force_press"]
    assert abs(fx - (-np.sin(th) * dp * ns * ds)) < 1e-11
    assert abs(fy - (np.cos(th) * dp * ns * ds)) < 1e-11


def test_soft_startup_scales_the_loads(oracle_mod):
    """soft = 0 before T_release, cubic ramp over tau (BeamForces.f90:246-252)"""
    ns = 8
    o = _oracle(oracle_mod, ns=ns, T_release=0.2, tau=0.5, t_beam=0.05)
    s = np.arange(ns) * (1.5 / (ns - 1))
    o.set_beam_arrays(x=1.2 + s, y=np.full(ns, 1.7), theta=np.zeros(ns))
    X, Y = np.meshgrid(np.arange(64) * (4.0 / 64), np.arange(64) * (4.0 / 64))
    for time, soft in ((0.1, 0.0), (0.45, None), (0.9, 1.0)):
        o.get_forces(time, Y, np.zeros((2, 64, 64)))
        if soft is None:
            a = time - 0.2
            soft = a ** 3 / (-0.5 * 0.5 ** 3) + 3 * a ** 2 / 0.5 ** 2
        assert np.allclose(o.beam_arrays()["pressure"], 2 * 0.05 * soft, atol=1e-13)


def test_shear_stress_of_sinusoidal_flow(oracle_mod):
    """u1 = sin(k y): stress_b = nu k cos(k y), the jump of sigma_t across a
    horizontal beam is tau = -2 nu k sin(k y0) sin(k r), r = t + N dx; the
    Delta kernel reproduces cos only to O(dy^2)"""
    ns, t, N, nu = 8, 0.05, 2.0, 1e-2
    o = _oracle(oracle_mod, ns=ns, t_beam=t, N_smooth=N, nu=nu)
    s = np.arange(ns) * (1.5 / (ns - 1))
    y0 = 1.7
    o.set_beam_arrays(x=1.2 + s, y=np.full(ns, y0), theta=np.zeros(ns))
    k = 2 * np.pi / 4.0
    Y = np.meshgrid(np.arange(64) * (4.0 / 64), np.arange(64) * (4.0 / 64))[1]
    u = np.stack([np.sin(k * Y), np.zeros_like(Y)])
    o.get_forces(1.0, np.zeros((64, 64)), u)
    r = o.beam_arrays()
    rr = t + N * 4.0 / 64
    tau = -2 * nu * k * np.sin(k * y0) * np.sin(k * rr)
    assert np.abs(r["tau"] - tau).max() < 1e-2 * abs(tau)
    # exact for the discrete kernel: the weights phi*(y_j - y) applied to the
    # analytic stress nu k cos(k y_j) at the grid rows
    L, dy = oracle_mod.lib(), 4.0 / 64

    def interp_cos(yt):
        j0 = int(yt / dy)
        js = np.arange(j0 - 4, j0 + 5)
        w = np.array([L.o_delta_phi4star(abs(j * dy - yt), dy, np.pi) for j in js])
        return (w * nu * k * np.cos(k * js * dy)).sum()  # the x weights sum to 1

    tau_d = interp_cos(y0 + rr) - interp_cos(y0 - rr)
    assert np.abs(r["tau"] - tau_d).max() < 1e-12 * abs(tau_d)
    # sigma_n = -A = 0 on a horizontal beam, so the pressure load is 0
    assert np.abs(r["pressure"]).max() < 1e-15


def test_surface_integral_pressure_of_a_linear_field(oracle_mod):
    """the pressure part of BeamSurfaceIntegral (BeamForces.f90:306-489) on a
    straight horizontal beam and p = a (y - y_beam), which the Delta
    interpolation reproduces exactly: the two lines at +-t_beam give
    -2 a t_beam ds per node, and the two end semicircles of radius
    rad = t_beam + N dx (128 points each over pi, both ends included) give
    -a rad^2 pi 129/127, the trapezoid sum of the circle's area; no
    x-component"""
    p = oracle_mod.make_params(nx=128, ny=128, ns=24)
    o = oracle_mod.Oracle(p)
    o.create_mask(0.0)
    b = o.beam_arrays()
    assert np.all(b["y"] == 2.0) and np.all(b["theta"] == 0.0)
    n, dx, a = 128, 4.0 / 128, 0.7
    Y = np.repeat((np.arange(n) * dx)[:, None], n, axis=1)
    o.get_forces(0.0, a * (Y - 2.0), np.zeros((2, n, n)))
    bsi = np.array(list(o.diag().bsi))
    rad = p.t_beam + p.N_smooth * dx
    ds = p.length / (p.ns - 1)
    want = -a * (2.0 * p.t_beam * ds * p.ns + np.pi * rad ** 2 * 129.0 / 127.0)
    assert abs(bsi[6]) < 1e-15 and bsi[9] == pytest.approx(want, rel=1e-13)
    assert np.all(bsi[[4, 5, 7, 8]] == 0.0)  # no velocity: no stresses
