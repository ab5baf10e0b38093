"""The oracle against the reference's own Fortran routines (oracle/_ref,
built from the unmodified sources by oracle/ref.mk): bit-for-bit agreement
of the spectral operators (cofd*.f90, poisson.f90), vis (cal_vis.f90),
dealiase_mask.f90, Mean_ux/Mean_uy (mean_velocity.f90), mouvement
(mouvement.f90) and integrate_position (integrate_position.f90)."""
import numpy as np
import pytest

from refbuild import Ref, needs_ref

pytestmark = needs_ref

OPS = ["cofdx", "cofdy", "cofdxdx", "cofdydy", "cofdxdy", "poisson"]
MOTIONS = [0, 2, 3, 4, 201, 205, 212, 213, 300, 333, 350, 401, 407, 413]


@pytest.mark.parametrize("nx,ny,xl,yl", [(64, 48, 4.0, 3.0), (96, 128, 2.0, 5.0), (256, 256, 4.0, 4.0)])
def test_operators_bit_exact(oracle_mod, nx, ny, xl, yl):
    o = oracle_mod.Oracle(nx=nx, ny=ny, xl=xl, yl=yl, nu=1e-2, iBeam=0, iSponge=0)
    r = Ref(nx, ny, xl=xl, yl=yl, nu=1e-2)
    fk = np.random.default_rng(nx + ny).normal(size=(ny, nx))
    for which, name in enumerate(OPS):
        assert np.array_equal(getattr(o, name)(fk), r.op(which, fk)), name
    for dt in (1e-4, 0.37):
        assert np.array_equal(o.vis(dt), r.vis(dt))
    assert np.array_equal(o.dealiase_mask(), r.dealiase())


@pytest.mark.parametrize("iMean", [0, 1, 2, 3, 6, 22])
def test_mean_velocity_bit_exact(oracle_mod, iMean):
    th, tf = 0.3, 0.7
    o = oracle_mod.Oracle(nx=16, ny=16, iBeam=0, iSponge=0, iMeanVelocity=iMean, theta_inf=th, T_fullspeed=tf)
    r = Ref(16, 16)
    r.set_flow(iMean, th, tf)
    for t in (0.0, 0.1, 0.35, 0.69999, 0.7, 2.5):
        assert o.L.o_mean_ux(o.h, t) == r.mean_u(t)[0]
        assert o.L.o_mean_uy(o.h, t) == r.mean_u(t)[1]


@pytest.mark.parametrize("im", MOTIONS)
def test_mouvement_bit_exact(oracle_mod, im):
    ns = 16
    o = oracle_mod.Oracle(nx=16, ny=16, ns=ns, iMotion=im, AngleBeam=12.5, x0=1.25, y0=2.5)
    r = Ref(16, 16, ns=ns)
    o.beam.phase = 0.4
    for t in (0.0, 0.05, 0.5, 0.999, 1.0, 1.7, 3.3):
        a, at, att, le = o.mouvement(t)
        ra, rat, ratt, rle = r.mouvement(im, 1.25, 2.5, 12.5, 0.4, t)
        assert (a, at, att) == (ra, rat, ratt), t
        assert np.array_equal(le, rle), t


@pytest.mark.parametrize("im", [0, 3, 201, 4, 407])
def test_integrate_position_bit_exact(oracle_mod, im):
    ns = 24
    o = oracle_mod.Oracle(nx=16, ny=16, ns=ns, iMotion=im, AngleBeam=-7.0, x0=0.75, y0=1.5, length=1.3)
    r = Ref(16, 16, ns=ns, length=1.3)
    rng = np.random.default_rng(im)
    for t in (0.0, 0.3, 1.1):
        th = 0.2 * rng.normal(size=ns)
        thd = rng.normal(size=ns)
        o.set_beam_arrays(theta=th, theta_dot=thd)
        o.integrate_position(t)
        b = o.beam_arrays()
        x, y, vx, vy = r.integrate_position(im, 0.75, 1.5, -7.0, 0.0, t, th, thd)
        for k, v in (("x", x), ("y", y), ("vx", vx), ("vy", vy)):
            assert np.array_equal(b[k], v), (k, t)
