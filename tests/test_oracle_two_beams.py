"""Two beams in the oracle (iBeam = 2, the reference's loops over beams:
init_beam.f90:48-76, create_mask.f90:74-119, NavierStokes.f90:527-591
(cal_drag), BeamForces.f90:24-47 (GetForces), time_step.f90:215-218).
Each per-beam quantity equals what the single-beam routines (pinned
elsewhere) give for that beam alone on the same fields, and the mask is
drawn beam after beam with the strict > of every primitive."""
import numpy as np
import pytest

PH = {300: 0.0, 301: 0.25, 302: 0.5, 303: 0.75, 304: 1.0, 305: 1.25, 306: 1.5, 307: 1.75}


def _two(oracle_mod, im=302, n=128, **kw):
    p = dict(nx=n, ny=n, ns=32, iBeam=2, iMotion=im, iSponge=0, x0=0.9, y0=2.0)
    p.update(kw)
    return oracle_mod.Oracle(**p)


def _one(oracle_mod, o2, ib, n=128, **kw):
    """a single-beam oracle holding beam ib of o2 (placement, phase, arrays)"""
    p = dict(nx=n, ny=n, ns=32, iBeam=1, iMotion=o2.p.iMotion, iSponge=0, x0=0.9, y0=2.0)
    p.update(kw)
    o = oracle_mod.Oracle(**p)
    b = o2.beam_arrays(ib)
    bb = o.beam
    bb.x0, bb.y0, bb.phase = b["x0"], b["y0"], b["phase"]
    o.set_beam_arrays(**{k: b[k] for k in ("x", "y", "vx", "vy", "theta", "theta_dot")})
    return o


@pytest.mark.parametrize("im", [300, 302, 307])
def test_second_beam_placement_and_phase(oracle_mod, im):
    o = _two(oracle_mod, im)
    assert o.nbeams == 2
    b0, b1 = o.beam_arrays(0), o.beam_arrays(1)
    assert b1["x0"] == pytest.approx(0.9 + 1.25, abs=1e-15) and b1["y0"] == 2.0
    assert b1["phase"] == pytest.approx(PH[im] * np.pi, abs=1e-15) and b0["phase"] == 0.0
    for t in (0.0, 0.13, 0.41):
        a0, at0, att0, le0 = o.mouvement(t, 0)
        a1, at1, att1, le1 = o.mouvement(t, 1)
        if im == 300:  # in phase: the second beam is the first shifted by 1.25
            assert a1 == a0 and le1[0] == pytest.approx(le0[0] + 1.25, abs=1e-14)
            np.testing.assert_array_equal(le1[1:], le0[1:])
        elif t > 0.0:  # the soft start k(t) (mouvement.f90:49-58) is 0 at t = 0
            assert le1[1] != le0[1] and le1[3] != le0[3]
    o.integrate_position(0.0)
    if im == 300:
        np.testing.assert_allclose(o.beam_arrays(1)["x"], o.beam_arrays(0)["x"] + 1.25, rtol=0, atol=1e-14)


def test_mask_is_drawn_beam_after_beam(oracle_mod):
    o2 = _two(oracle_mod, 305)
    t = 0.2
    o2.integrate_position(t)
    o2.create_mask(t)
    m, vx, vy = (o2.field(k) for k in ("mask", "maskvx", "maskvy"))
    singles = []
    for ib in range(2):
        o = _one(oracle_mod, o2, ib)
        o.create_mask(t)
        singles.append([o.field(k) for k in ("mask", "maskvx", "maskvy")])
    (m0, vx0, vy0), (m1, vx1, vy1) = singles
    assert np.count_nonzero(m0) > 100 and np.count_nonzero(m1) > 100
    want = np.maximum(m0, m1)
    np.testing.assert_array_equal(m, want)
    second = m1 > m0  # strict: where the second beam's value is larger it wins
    np.testing.assert_array_equal(vx, np.where(second, vx1, vx0))
    np.testing.assert_array_equal(vy, np.where(second, vy1, vy0))


def test_drag_and_loads_per_beam(oracle_mod):
    n = 128
    o2 = _two(oracle_mod, 303, n=n)
    t = 0.27
    o2.integrate_position(t)
    o2.create_mask(t)
    rng = np.random.default_rng(4)
    X, Y = np.meshgrid(np.arange(n) * (4.0 / n), np.arange(n) * (4.0 / n))
    pen = rng.normal(size=(2, n, n))
    press = np.sin(X) * np.cos(1.3 * Y) + 0.2 * X
    u = np.stack([np.cos(X) * np.sin(Y), -np.sin(X) * np.cos(Y)])
    o2.cal_drag(t, 2e-4, pen)
    o2.get_forces(t, press, u)
    for ib in range(2):
        o = _one(oracle_mod, o2, ib, n=n)
        for k in ("mask", "maskvx", "maskvy"):
            o.set_field(k, o2.field(k))
        o.cal_drag(t, 2e-4, pen)
        o.get_forces(t, press, u)
        b, b2 = o.beam_arrays(), o2.beam_arrays(ib)
        assert b2["Force"] == b["Force"], ib
        np.testing.assert_array_equal(b2["pressure"], b["pressure"])
        np.testing.assert_array_equal(b2["tau"], b["tau"])
        assert b2["Force_press"] == b["Force_press"]
        np.testing.assert_array_equal(o2.beam_bsi(ib), o.beam_bsi(0))
    assert o2.beam_arrays(0)["Force"] != o2.beam_arrays(1)["Force"]


def test_two_beam_run(oracle_mod):
    """prescribed motion of both beams through the oracle's loop: each beam
    is where integrate_position puts it at the last step's start time"""
    o = _two(oracle_mod, 304, n=96, nx=96, ny=96, dt_fixed=2e-4)
    o.set_vortk(o.init_noise(seed=1, amp=1e-3, kmax2=64))
    o.run(6)
    t_last = o.diag().time - o.diag().dt1
    ref = _two(oracle_mod, 304, n=96, nx=96, ny=96)
    ref.integrate_position(t_last)
    for ib in range(2):
        np.testing.assert_array_equal(o.beam_arrays(ib)["y"], ref.beam_arrays(ib)["y"])
    # the beams heave out of phase (still inside the soft start k ~ 35 t^4)
    assert np.abs(o.beam_arrays(1)["y"] - o.beam_arrays(0)["y"]).max() > 0.0
