"""create_mask / create_sponge_mask of the oracle pinned by closed forms
(create_mask.f90): the erf step (SmoothStep :475-520), a cylinder drawn as
one dot (:67-69, DrawADot :660-701), up-down symmetry and rigid velocity of a
straight heaving beam, and the type-3 sponge (:315-330)."""
import numpy as np
from scipy.special import erf

GERF = abs((np.exp(-4.0) - 1.0) / np.sqrt(np.pi))


def smooth(x, t, h):
    d = h * GERF
    return 0.5 * (erf((t - x) / d) + erf((x + t) / d))


def test_smoothstep_closed_form(oracle_mod):
    o = oracle_mod.Oracle(nx=16, ny=16, iBeam=0, iSponge=0, N_smooth=2.0)
    for x in np.linspace(0.0, 0.3, 31):
        assert abs(o.smoothstep(x, 0.1, 0.04) - smooth(x, 0.1, 0.04)) < 2e-16
    # both erf terms saturate once |t -+ x| > 6 delta
    assert o.smoothstep(0.0, 0.2, 0.04) == 1.0 and o.smoothstep(0.5, 0.1, 0.04) == 0.0
    sharp = oracle_mod.Oracle(nx=16, ny=16, iBeam=0, iSponge=0, N_smooth=0.0)
    assert sharp.smoothstep(0.1, 0.1, 0.0) == 1.0 and sharp.smoothstep(0.1000001, 0.1, 0.0) == 0.0


def test_cylinder_is_a_radial_erf_profile(oracle_mod):
    n, R, N, eps = 64, 0.5, 2.0, 1e-3
    o = oracle_mod.Oracle(nx=n, ny=n, iBeam=0, iCylinder=1, R_cylinder=R, x0=2.0, y0=2.0, N_smooth=N,
                          eps=eps, iSponge=0)
    o.create_mask(0.0)
    dx = 4.0 / n
    X, Y = np.meshgrid(np.arange(n) * dx, np.arange(n) * dx)
    r = np.sqrt((X - (2.0 - R)) ** 2 + (Y - 2.0) ** 2)
    assert np.abs(o.field("mask") * eps - smooth(r, R, N * dx)).max() < 1e-15
    assert not o.field("maskvx").any() and not o.field("maskvy").any()


def test_straight_heaving_beam(oracle_mod):
    """heave y = A sin(2 pi t) at t = 0: a horizontal beam on grid row 64,
    moving rigidly with v = 2 pi A"""
    n, ns, t, N, eps, A = 128, 32, 0.05, 2.0, 1e-3, 0.1
    o = oracle_mod.Oracle(nx=n, ny=n, ns=ns, iBeam=1, iMotion=600, heave_amp=A, heave_freq=1.0, x0=1.0,
                          y0=2.0, t_beam=t, N_smooth=N, eps=eps, iSponge=0)
    o.integrate_position(0.0)
    b = o.beam_arrays()
    assert np.all(b["y"] == 2.0) and np.allclose(b["vy"], 2 * np.pi * A, rtol=0, atol=1e-15)
    o.create_mask(0.0)
    m, vx, vy = o.field("mask"), o.field("maskvx"), o.field("maskvy")
    j0 = 64  # y0 = 64 dy exactly
    for d in range(1, 30):
        assert np.array_equal(m[j0 + d], m[j0 - d]), d
    on = m > 0
    assert on.any() and np.all(vx[on] == 0.0) and np.all(np.abs(vy[on] - 2 * np.pi * A) < 1e-15)
    # beyond t + 7 delta the erf step is exactly 0; on the centre line inside
    # the beam it is erf(t / delta) / eps
    dx = 4.0 / n
    far = int(np.ceil((t + 7 * N * dx * GERF) / dx)) + 1
    assert not m[: j0 - far].any() and not m[j0 + far:].any()
    i_mid = int(round(1.5 / dx))
    assert abs(m[j0, i_mid] * eps - smooth(0.0, t, N * dx)) < 1e-15


def test_sponge_type3(oracle_mod):
    """chi_s = min(1, sum of four erf steps of the distance to each border)
    / eps_sponge with thickness SpongeSize - N dx"""
    n, S, N, es = 64, 0.4, 2.0, 5e-3
    o = oracle_mod.Oracle(nx=n, ny=n, iBeam=0, iSponge=3, SpongeSize=S, N_smooth=N, eps_sponge=es)
    o.create_sponge_mask()
    ms = o.field("mask_sponge")
    dx = 4.0 / n
    X, Y = np.meshgrid(np.arange(n) * dx, np.arange(n) * dx)
    th, h = S - N * dx, N * dx
    ref = np.minimum(1.0, smooth(X, th, h) + smooth(4.0 - X, th, h) + smooth(Y, th, h)
                     + smooth(4.0 - Y, th, h)) / es
    assert np.abs(ms - ref).max() < 1e-15 / es
    assert ms[0, 0] == 1.0 / es and ms[n // 2, n // 2] < 1e-12


def test_smooth_cutoff_extends_the_reference_mask(oracle_mod):
    """iSmoothCutoff = 1 draws segments and hinges out to t + 7 delta: on a
    straight beam each column through the beam's interior is then the erf
    profile of the distance to the centre line on every row, where the
    reference's t* = t + 2 N dx cuts it at SS(t*) ~ 2e-7, and the mask never
    decreases; Force's window then holds every pixel where the mask is not 0"""
    n, ns, t, N, eps, A = 128, 32, 0.05, 2.0, 1e-3, 0.1
    kw = dict(nx=n, ny=n, ns=ns, iBeam=1, iMotion=600, heave_amp=A, heave_freq=1.0, x0=1.0, y0=2.0, t_beam=t,
              N_smooth=N, eps=eps, iSponge=0)
    dx = 4.0 / n
    ms = {}
    for cut in (0, 1):
        o = oracle_mod.Oracle(iSmoothCutoff=cut, **kw)
        o.integrate_position(0.0)
        o.create_mask(0.0)
        ms[cut] = o.field("mask") * eps
    ref, sm = ms[0], ms[1]
    j0, i_mid = 64, int(round(1.5 / dx))
    d = np.abs(np.arange(n) - j0) * dx
    assert np.all(np.abs(sm[:, i_mid] - np.maximum(smooth(d, t, N * dx), 0.0)) < 1e-16)
    t_star = t + 2 * N * dx
    inner = d <= t_star
    assert np.array_equal(ref[inner, i_mid], sm[inner, i_mid])
    assert not ref[~inner, i_mid].any()
    step = smooth(t_star, t, N * dx)
    assert 1e-7 < step < 1e-6 and sm[~inner, i_mid].max() > 1e-9
    assert np.all(sm >= ref) and (sm - ref).max() <= step
