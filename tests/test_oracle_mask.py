"""create_mask / create_sponge_mask of the oracle pinned by closed forms
(create_mask.f90): the erf step (SmoothStep :475-520), a cylinder drawn as
one dot (:67-69, DrawADot :660-701), up-down symmetry and rigid velocity of a
straight heaving beam, the channel walls (:34-53), the sponge masks of
types 1, 2, 3 and 10 (:237-330), and cal_drag's window sums against the
integral of the penalisation (NavierStokes.f90:512-593)."""
import numpy as np
import pytest
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


def _ss2(x, t, h):
    """SmoothStep2, create_mask.f90:523-544 (cosine step)"""
    x = np.asarray(x, float)
    return np.where(x <= t - h, 1.0, np.where(x < t + h, 0.5 * (1.0 + np.cos((x - t + h) * np.pi / (2.0 * h))), 0.0))


def test_frame_around_the_domain(oracle_mod):
    """iSponge = 22: the frame is SmoothStep2 of the distance to the nearest
    edge over 4 t_beam (away from the rounded corners), 1 on the edge rows
    and in the corners, added to the mask at weight 1 (eps = 1e-3) with the
    mean velocity; the vorticity sponge is 1/eps_sponge in the outer band
    and 0 where the mask is positive or inside"""
    n, tb, S = 64, 0.05, 0.3
    o = oracle_mod.Oracle(nx=n, ny=n, iBeam=0, iSponge=22, SpongeSize=S, t_beam=tb, eps=1e-3, eps_sponge=1e-3,
                          iMeanVelocity=3, theta_inf=0.0)
    o.create_mask(0.0)
    o.create_sponge_mask()
    dx = 4.0 / n
    T = 4 * tb
    fr = o.field("mask") * 1e-3
    mid = n // 2
    x = np.arange(n) * dx
    near = x < 2 * T - 1e-12  # columns that only the edge distance reaches on the middle row
    np.testing.assert_allclose(fr[mid, near], _ss2(x[near], T / 2, T / 2), rtol=0, atol=1e-15)
    assert fr[mid, mid] == 0.0 and fr[0, mid] == 1.0 and fr[0, 0] == 1.0 and fr[n - 1, n - 1] == 1.0
    assert np.array_equal(fr, fr.T)  # square domain
    vx = o.field("maskvx")
    assert np.all(vx[fr > 0] == 1.0) and np.all(vx[fr == 0] == 0.0)
    ms = o.field("mask_sponge") * 1e-3
    k = int(round(0.25 / dx))  # inside the sponge band, beyond the frame's reach
    assert fr[mid, k] == 0.0 and ms[mid, k] == pytest.approx(1.0, rel=1e-15)
    assert ms[mid, mid] == 0.0 and ms[0, mid] == 0.0


def test_velocity_sponge_of_a_channel(oracle_mod):
    """iSponge = 4: over the first S of the channel the mask is at least the
    erf step of the sponge (left half: x = 0.75 S - x_i over a quarter of
    the core size; right half: x_i - 0.5 S over the smoothing layer) and
    maskvx is the parabola 1.5 Mean_ux y (H - y) / (H/2)^2 inside the
    channel, 0 in the walls"""
    nx, ny, h, S, N = 128, 64, 0.4, 0.5, 2.0
    o = oracle_mod.Oracle(nx=nx, ny=ny, iBeam=0, iWalls=1, h_channel=h, iSponge=4, SpongeSize=S, N_smooth=N,
                          iMeanVelocity=3, theta_inf=0.0, eps=1e-3)
    o.create_mask(0.0)
    m, vx = o.field("mask") * 1e-3, o.field("maskvx")
    dx, dy = 4.0 / nx, 4.0 / ny
    H = 4.0 - 2 * h
    ssz = S - 2 * N * dx
    yc = np.arange(ny) * dy - h
    prof = np.where((yc > 0) & (yc < H), 1.5 * yc * (H - yc) / (0.5 * H) ** 2, 0.0)
    imid, imax = int(round(0.5 * S / dx)), int(round(S / dx))
    for ix in (0, 3, imid - 1):
        step = float(smooth(0.75 * S - ix * dx, ssz * 0.5, 0.25 * ssz))
        assert np.all(m[:, ix] >= step - 1e-16)
        np.testing.assert_allclose(vx[:, ix], prof, rtol=1e-15, atol=0)
        inside = (yc > 2 * h) & (yc < H - 2 * h)  # rows the walls do not reach
        np.testing.assert_allclose(m[inside, ix], step, rtol=1e-15)
    for ix in (imid + 1, imax):
        step = float(smooth(ix * dx - 0.5 * S, ssz * 0.5, N * dx))
        inside = (yc > 2 * h) & (yc < H - 2 * h)
        np.testing.assert_allclose(m[inside, ix], step, rtol=1e-15, atol=1e-300)
        np.testing.assert_allclose(vx[:, ix], prof, rtol=1e-15, atol=0)
    assert np.all(vx[:, imax + 1:] == 0.0)


def test_zero_mean_pressure_forcing(oracle_mod):
    """iMeanVelocity = 11: u_mean advances with the mean penalisation force
    f = mean(-mask (u - u_s)) of each step's cal_nlk: u_mean = f dt1 after
    the startup step, u_mean + b10 f_n + b11 f_(n-1) after an AB2 step"""
    n = 64
    o = oracle_mod.Oracle(nx=n, ny=n, iBeam=0, iCylinder=1, R_cylinder=0.3, x0=2.0, y0=2.0, iSponge=0,
                          iMeanVelocity=11, iMotion=0, dt_fixed=2e-4)
    vk = o.init_noise(seed=5, amp=1e-2, kmax2=64)
    o.set_vortk(vk)
    u0, _ = o.cal_velocity(0.0, vk)
    o.run(1)
    mask, mvx, mvy = o.field("mask"), o.field("maskvx"), o.field("maskvy")
    f1 = np.array([np.mean(-(mask * (u0[0] - mvx))), np.mean(-(mask * (u0[1] - mvy)))])
    d = o.diag()
    um1 = np.array(d.u_mean)
    np.testing.assert_allclose(um1, f1 * d.dt1, rtol=1e-12, atol=1e-18)
    assert np.abs(um1).max() > 0
    u1 = np.stack([o.field("u1"), o.field("u2")])
    o.run(1)
    f2 = np.array([np.mean(-(mask * (u1[0] - mvx))), np.mean(-(mask * (u1[1] - mvy)))])
    d2 = o.diag()
    dt1, dt0 = d2.dt1, d2.dt0
    b10 = dt1 / dt0 * (0.5 * dt1 + dt0)
    b11 = -(0.5 * dt1 * dt1) / dt0
    np.testing.assert_allclose(np.array(d2.u_mean), um1 + b10 * f2 + b11 * f1, rtol=1e-12, atol=1e-18)


def test_channel_walls(oracle_mod):
    """iWalls = 1 (create_mask.f90:34-53): the bottom wall is the erf step
    of the height y about h_channel (width N dy), the top wall the same step
    of the distance yl - y to the top; 1 deep in the walls, 1/2 at y =
    h_channel, 0 in the channel; every column alike"""
    n, h, N, eps = 64, 0.5, 2.0, 1e-3
    o = oracle_mod.Oracle(nx=n, ny=n, iBeam=0, iSponge=0, iWalls=1, h_channel=h, N_smooth=N, eps=eps)
    o.create_mask(0.0)
    m = o.field("mask") * eps
    dy = 4.0 / n
    y = np.arange(n) * dy
    assert np.array_equal(m, np.repeat(m[:, :1], n, axis=1))
    jb = int(np.floor((h + N * dy) / dy + 0.5))  # rows the bottom wall is drawn on
    np.testing.assert_allclose(m[: jb + 1, 0], smooth(y[: jb + 1], h, N * dy), rtol=0, atol=1e-15)
    np.testing.assert_allclose(m[n - jb:, 0], smooth(4.0 - y[n - jb:], h, N * dy), rtol=0, atol=1e-15)
    assert m[0, 0] == pytest.approx(1.0, abs=1e-15) and m[n // 2, 0] == 0.0
    assert m[int(round(h / dy)), 0] == pytest.approx(0.5, abs=1e-15)
    np.testing.assert_allclose(m[1: jb + 1, 0], m[n - 1: n - jb - 1: -1, 0], rtol=0, atol=1e-15)


def test_outflow_inflow_and_sharp_sponges(oracle_mod):
    """create_sponge_mask types 10, 1 and 2 (create_mask.f90:237-306):
    10 is 1/eps_sponge on the columns from xl - S on; 1 (outflow) rises
    from ~0 at xl - S towards 1 at the sponge's middle xl - S/2 as the erf
    step of the distance to the middle about S/2 - N dx, and from the middle
    on follows the step of its right half; 2 (inflow) is the same
    step on both sides of its middle S/2, symmetric about it, ~0 at x = 0
    and x = S and 0 beyond; all uniform in y"""
    n, S, N, es = 64, 1.0, 2.0, 5e-3
    dx = 4.0 / n
    x = np.arange(n) * dx
    ms = {}
    for t in (10, 1, 2):
        o = oracle_mod.Oracle(nx=n, ny=n, iBeam=0, iSponge=t, SpongeSize=S, N_smooth=N, eps_sponge=es)
        o.create_mask(0.0)
        o.create_sponge_mask()
        ms[t] = o.field("mask_sponge") * es
        assert np.array_equal(ms[t], np.repeat(ms[t][:1], n, axis=0)), t
    r10 = ms[10][0]
    i0 = int(np.floor((4.0 - S) / dx + 0.5))
    assert np.all(r10[i0:] == 1.0) and np.all(r10[:i0] == 0.0)
    r1 = ms[1][0]
    imid = int(np.floor((4.0 - 0.5 * S) / dx + 0.5))
    th = 0.5 * (S - 2.0 * N * dx)
    np.testing.assert_allclose(r1[i0:imid], smooth(4.0 - 0.5 * S - x[i0:imid], th, N * dx), rtol=0, atol=1e-15)
    # from the middle column on, the right half (drawn second, so it owns the
    # middle column): the step of x - (xl - S/2) - dx about 0.4 (S - 2 N dx)
    np.testing.assert_allclose(r1[imid:], smooth(x[imid:] - 4.0 + 0.5 * S - dx, 0.8 * th, N * dx), rtol=0,
                               atol=1e-15)
    assert r1[i0] < 0.01 and r1[imid - 1] == pytest.approx(1.0, abs=1e-6)  # the smoothing layer's tail
    assert np.all(np.diff(r1[i0:imid]) >= 0.0) and np.all(r1[:i0] == 0.0)
    r2 = ms[2][0]
    m2 = int(np.floor(0.5 * S / dx + 0.5))
    np.testing.assert_allclose(r2[: m2 + 1], smooth(0.5 * S - x[: m2 + 1], th, N * dx), rtol=0, atol=1e-15)
    np.testing.assert_allclose(r2[: m2 + 1], r2[2 * m2: m2 - 1: -1], rtol=0, atol=1e-15)
    assert r2[0] < 0.01 and r2[m2] == pytest.approx(1.0, abs=1e-6)
    assert np.all(r2[2 * m2 + 1:] == 0.0)


def test_drag_window_sums(oracle_mod):
    """cal_drag's Force (NavierStokes.f90:512-593) is dx dy times the sum of
    the penalisation chi (u - u_s) over its window; BeamSurfaceIntegral's
    first two entries are the same sum over the whole grid.  With the
    window grown to the whole mask (iSmoothCutoff = 1) the two are equal,
    and both equal the integral of chi (u - u_s) computed here from the
    fields; with the reference's window (1.5 t_beam around the beam) the
    smooth tails outside it are left out, so the window sum is smaller"""
    res = {}
    for smooth in (1, 0):
        o = oracle_mod.Oracle(oracle_mod.make_params(nx=64, ny=64, ns=24, iSmoothCutoff=smooth, dt_fixed=2e-4))
        o.set_vortk(o.init_noise(seed=0, amp=1e-3, kmax2=64))
        o.run(5)
        F = np.array(o.beam_arrays()["Force"])
        bsi = np.array(list(o.diag().bsi)[:2])
        m, vx, vy = o.field("mask"), o.field("maskvx"), o.field("maskvy")
        u1, u2 = o.field("u1"), o.field("u2")
        dxdy = (4.0 / 64) ** 2
        full = np.array([np.sum(m * (u1 - vx)), np.sum(m * (u2 - vy))]) * dxdy
        np.testing.assert_allclose(bsi, full, rtol=1e-12)
        res[smooth] = (F, bsi)
    F1, b1 = res[1]
    assert np.array_equal(F1, b1)
    F0, b0 = res[0]
    assert np.all(np.abs(F0) < np.abs(b0))
