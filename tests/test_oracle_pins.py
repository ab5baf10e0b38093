"""Independent checks of the oracle's restatements that no golden vector or
reference build reaches (create_mask.f90 cannot be compiled here: Intel-only
logical ==):

* DrawHinge's angular sectors (create_mask.f90:563-656): a bent polyline
  drawn by segments, hinges and end dots is SmoothStep of the distance to
  the polyline inside t*, by a numpy brute force over every segment; the
  outer wedge of a bend carries the node's velocity.  The bends cover all
  nine branch pairs of beta1 and beta2, turning both ways.
* DrawSharpEnd (create_mask.f90:784-852): on a straight beam, the product of
  the two erf steps on its box, overwriting what the segments drew, and the
  end velocity beyond the end point.
* The iSponge = 10 term of cal_nlk (NavierStokes.f90:102-121): the
  trapezoid integral of -chi_s du1/dx of a single mode against the closed
  form of the discrete sum, its second-order convergence to the integral,
  and its place in nlk.
* cal_drag (NavierStokes.f90:512-593): the window holds the whole sharp mask;
  its bounds truncate (a shift of the beam by 0.4 dx adds one column); and
  Force_unst of a rigid heave at the top of its stroke is the body's volume
  times its acceleration.
* The stress half of BeamSurfaceIntegral (BeamForces.f90:306-509): constant
  stresses, which the Delta interpolation reproduces exactly, against the
  traction S.n of the stress tensor at the same surface points.
"""
import numpy as np
import pytest
from scipy.special import erf

GERF = abs((np.exp(-4.0) - 1.0) / np.sqrt(np.pi))


def smooth(x, t, h):
    d = h * GERF
    return 0.5 * (erf((t - x) / d) + erf((x + t) / d))


def _grid(n, L=4.0):
    d = L / n
    X, Y = np.meshgrid(np.arange(n) * d, np.arange(n) * d)
    return X, Y, d


def _seg_dist(X, Y, p, q):
    """distance of every pixel to the segment p-q, and the projection
    parameter along it (0 at p, |q-p| at q)"""
    ex, ey = q[0] - p[0], q[1] - p[1]
    L = np.hypot(ex, ey)
    ex, ey = ex / L, ey / L
    s = (X - p[0]) * ex + (Y - p[1]) * ey
    sc = np.clip(s, 0.0, L)
    return np.hypot(X - (p[0] + sc * ex), Y - (p[1] + sc * ey)), s, L


# ---------------------------------------------------------------------------
# DrawHinge
# ---------------------------------------------------------------------------
def _branch1(b):
    return "A" if b < 90.0 else ("B" if b < 270.0 else "C")


def _branch2(b):
    return "A" if b <= 90.0 else ("B" if b < 270.0 else "C")


BENDS = [(10, 70), (60, 130), (30, -50), (120, 50), (100, 170), (230, 290), (320, 30), (300, 230),
         (280, 340), (0, 85), (0, -85), (45, 52)]


def _polyline(theta_in, theta_out, seg=0.25):
    a, b = np.radians(theta_in), np.radians(theta_out)
    p2 = np.array([2.0, 2.0])
    pts = [p2 - 2 * seg * np.array([np.cos(a), np.sin(a)]), p2 - seg * np.array([np.cos(a), np.sin(a)]), p2,
           p2 + seg * np.array([np.cos(b), np.sin(b)]), p2 + 2 * seg * np.array([np.cos(b), np.sin(b)])]
    return np.array(pts)


@pytest.mark.parametrize("bend", BENDS)
def test_hinge_sectors_complete_the_polyline_distance(oracle_mod, bend):
    n, t, N, eps = 128, 0.0312, 2.0, 1e-3
    P = _polyline(*bend)
    ns = len(P)
    vx = 0.1 * (np.arange(ns) + 1)
    vy = -0.05 * (np.arange(ns) + 2)
    seg_ang = np.arctan2(np.diff(P[:, 1]), np.diff(P[:, 0]))
    o = oracle_mod.Oracle(nx=n, ny=n, ns=ns, iBeam=1, iMotion=0, iSponge=0, t_beam=t, N_smooth=N, eps=eps)
    o.set_beam_arrays(x=P[:, 0], y=P[:, 1], vx=vx, vy=vy, theta=np.r_[seg_ang, seg_ang[-1]],
                      theta_dot=np.zeros(ns))
    o.create_mask(0.0)
    m = o.field("mask") * eps
    mvx, mvy = o.field("maskvx"), o.field("maskvy")
    X, Y, dx = _grid(n)
    h = N * dx
    tstar = t + 2 * N * dx
    dists = [_seg_dist(X, Y, P[k], P[k + 1]) for k in range(ns - 1)]
    d = np.min([dd[0] for dd in dists], axis=0)
    r0 = np.hypot(X - P[0, 0], Y - P[0, 1])
    r4 = np.hypot(X - P[-1, 0], Y - P[-1, 1])
    want = np.maximum(np.where(d <= tstar, smooth(d, t, h), 0.0), np.maximum(smooth(r0, t, h), smooth(r4, t, h)))
    band = np.abs(d - tstar) < 1e-9
    assert np.abs(m - want)[~band].max() <= 1e-12
    assert np.abs(m - want).max() <= 2e-7  # the cut-off at t*, where the step is ~7e-8
    # the outer wedge of the bend at node 2 (beyond segment 1, before
    # segment 2, inside t*): only the hinge draws there, with the node's velocity
    marg = 1e-3 * dx
    s1, s2 = dists[1][1], dists[2][1]
    r2 = np.hypot(X - P[2, 0], Y - P[2, 1])
    wedge = (s1 > dists[1][2] + marg) & (s2 < -marg) & (r2 < tstar - marg)
    turn = (bend[1] - bend[0] + 180.0) % 360.0 - 180.0
    if abs(turn) >= 30.0:
        assert wedge.sum() >= 3, wedge.sum()
    assert np.all(mvx[wedge] == vx[2]) and np.all(mvy[wedge] == vy[2])
    # and the wedge lies on the outer side of the bend
    if wedge.any():
        side = np.sign(np.cos(seg_ang[1]) * (Y[wedge] - P[2, 1]) - np.sin(seg_ang[1]) * (X[wedge] - P[2, 0]))
        assert np.all(side == -np.sign(turn))


def test_hinge_bends_cover_every_branch_pair():
    """create_mask.f90:612-636 distinguishes beta1 < 90, [90, 270), >= 270
    and beta2 <= 90, [90, 270), >= 270; the bends above reach all nine pairs"""
    seen = set()
    for a, b in BENDS:
        b1, b2 = a % 360.0, b % 360.0
        seen.add((_branch1(b1), _branch2(b2)))
    assert len(seen) == 9, sorted(seen)


# ---------------------------------------------------------------------------
# DrawSharpEnd
# ---------------------------------------------------------------------------
def test_sharp_end_is_the_product_of_two_steps(oracle_mod):
    n, ns, t, N, eps, th0 = 128, 16, 0.0312, 2.0, 1e-3, np.radians(30.0)
    e = np.array([np.cos(th0), np.sin(th0)])
    p0 = np.array([1.4, 1.3])
    s_nodes = np.linspace(0.0, 1.0, ns)
    P = p0[None, :] + s_nodes[:, None] * e[None, :]
    vx = 0.2 + 0.01 * np.arange(ns)
    vy = -0.1 + 0.02 * np.arange(ns)
    o = oracle_mod.Oracle(nx=n, ny=n, ns=ns, iBeam=1, iMotion=0, iSponge=0, iSharpTrailing=1, t_beam=t,
                          N_smooth=N, eps=eps)
    o.set_beam_arrays(x=P[:, 0], y=P[:, 1], vx=vx, vy=vy, theta=np.full(ns, th0), theta_dot=np.zeros(ns))
    o.create_mask(0.0)
    m = o.field("mask") * eps
    mvx, mvy = o.field("maskvx"), o.field("maskvy")
    X, Y, dx = _grid(n)
    hh = dx  # max(dx, dy)
    tstar = t + 2 * N * hh
    # beam coordinates: xs along the beam from its leading edge, ys across
    xs = (X - p0[0]) * e[0] + (Y - p0[1]) * e[1]
    ys = -(X - p0[0]) * e[1] + (Y - p0[1]) * e[0]
    seg = np.where((xs >= 0) & (xs <= 1.0) & (np.abs(ys) <= tstar), smooth(np.abs(ys), t, N * hh), 0.0)
    dot0 = smooth(np.hypot(X - p0[0], Y - p0[1]), t, N * hh)
    a0, a1 = 1.0 - 2 * hh, 1.0 + 4 * hh   # the box: 2 cells inside the end, 4 beyond
    box = (xs >= a0) & (xs <= a1) & (np.abs(ys) <= tstar)
    prod = smooth(xs - a0, 2 * hh, N * hh) * smooth(np.abs(ys), t, N * hh)
    want = np.maximum(np.where(box, prod, seg), dot0)
    edge = (np.abs(xs - a0) < 1e-9) | (np.abs(xs - a1) < 1e-9) | (np.abs(np.abs(ys) - tstar) < 1e-9) | \
           (np.abs(xs - 1.0) < 1e-9) | (np.abs(xs) < 1e-9)
    assert box.sum() > 50
    assert np.abs(m - want)[~edge].max() <= 1e-12
    # inside the box the product overwrites smaller and larger values alike
    inner = box & ~edge & (xs < 1.0)
    assert np.any(prod[inner] < seg[inner] - 1e-3)
    # velocity: the end node's beyond the end point, the segments' before it
    beyond = box & ~edge & (xs - a0 >= 2 * hh + 1e-9)
    assert beyond.sum() > 10
    assert np.all(mvx[beyond] == vx[-1]) and np.all(mvy[beyond] == vy[-1])
    before = box & ~edge & (xs < 1.0 - 1e-9) & (seg > 0)
    assert np.all(np.abs(mvx[before] - (vx[-2] + (xs[before] - s_nodes[-2]) * (vx[-1] - vx[-2]) / (1.0 / 15)))
                  <= 1e-12)
    # no trailing dot with a sharp end: nothing beyond the box
    far = (xs > a1 + 1e-9) & (np.hypot(X - p0[0], Y - p0[1]) > 0.5)
    assert not m[far].any()


# ---------------------------------------------------------------------------
# iSponge = 10
# ---------------------------------------------------------------------------
def _trapz_cos_exact(a, b, h, th0):
    """sum_{m=a}^{b-1} (cos(th0 + m h) + cos(th0 + (m+1) h)) / 2 in closed form"""
    if b <= a:
        return 0.0
    cnt = b - a + 1
    full = np.sin(cnt * h / 2) / np.sin(h / 2) * np.cos(th0 + (a + b) * h / 2)
    return full - 0.5 * (np.cos(th0 + a * h) + np.cos(th0 + b * h))


def _sponge10(oracle_mod, n, c=3.0, S=1.5, ramp=False):
    o = oracle_mod.Oracle(nx=n, ny=n, iBeam=1, iSponge=10, SpongeSize=S, iMotion=0)
    X, Y, dx = _grid(n)
    A, k, phi, mm = 0.7, 2 * np.pi * 2 / 4.0, 0.3, 2 * np.pi / 4.0
    xa = round((4.0 - S) / dx) * dx
    chi = c * (X - xa) if ramp else np.full_like(X, c)
    o.set_field("mask_sponge", chi)
    u1 = A * np.sin(k * X + phi) * np.cos(mm * Y)
    return o, X, Y, dx, (A, k, phi, mm, c, xa), u1


@pytest.mark.parametrize("n", [64, 128])
def test_sponge10_is_the_trapezoid_sum(oracle_mod, n):
    o, X, Y, dx, (A, k, phi, mm, c, xa), u1 = _sponge10(oracle_mod, n)
    e = o.sponge10_term(u1)
    # the window runs from the sponge's inner edge to 16 cells before the
    # domain's end (empty unless the sponge is wider than 16 cells)
    ixmin, ixmax = int(round(2.5 / dx)), n - 16
    assert ixmax - ixmin >= 8
    want = np.zeros((n, n))
    for ix in range(ixmin, ixmax + 1):
        want[:, ix] = -c * A * k * dx * _trapz_cos_exact(ixmin, ix, k * dx, phi) * np.cos(mm * Y[:, 0])
    assert np.abs(e - want).max() <= 1e-13 * np.abs(want).max()
    assert not e[:, :ixmin + 1].any() and not e[:, ixmax + 1:].any()
    assert np.abs(e[:, ixmax]).max() > 0.1


def test_sponge10_converges_to_the_integral(oracle_mod):
    """with chi_s = c (x - x_a) the trapezoid sum of -chi_s du1/dx converges
    to its integral at second order"""
    errs = []
    for n in (64, 128, 256):
        o, X, Y, dx, (A, k, phi, mm, c, xa), u1 = _sponge10(oracle_mod, n, ramp=True)
        e = o.sponge10_term(u1)
        # the window's end (nx - 16) moves with n: compare on [2.5, 3], inside all three
        ix = np.arange(int(round(2.5 / dx)), int(round(3.0 / dx)) + 1)
        x = ix * dx
        # -c int_{xa}^{x} (s - xa) A k cos(k s + phi) ds, by parts
        F = lambda s: (s - xa) * np.sin(k * s + phi) + np.cos(k * s + phi) / k  # noqa: E731
        exact = -c * A * (F(x) - F(xa))[None, :] * np.cos(mm * Y[:, :1])
        errs.append(np.abs(e[:, ix] - exact).max() / np.abs(exact).max())
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert 3.6 < r1 < 4.4 and 3.6 < r2 < 4.4, errs


def test_sponge10_enters_nlk_as_minus_d_dy(oracle_mod):
    """nlk with iSponge = 10 minus nlk with the sponge off is
    -d/dy of the increment (nlk += pen2 - pen1, pen1 = cofdy(coftxy(.)))"""
    n = 64
    o10, X, Y, dx, _, u1 = _sponge10(oracle_mod, n)
    o0 = oracle_mod.Oracle(nx=n, ny=n, iBeam=1, iSponge=0, iMotion=0)
    u = np.stack([u1, 0.4 * np.cos(2 * np.pi * X / 4.0)])
    vk = np.zeros((n, n))
    nl10 = o10.cal_nlk(0.0, vk, u)[0]
    nl0 = o0.cal_nlk(0.0, vk, u)[0]
    e = o10.sponge10_term(u1)
    want = -o0.cofdy(o0.coftxy(e))
    assert np.abs((nl10 - nl0) - want).max() <= 1e-13 * np.abs(want).max()
    assert np.abs(want).max() > 1e-3


# ---------------------------------------------------------------------------
# cal_drag
# ---------------------------------------------------------------------------
def _heave_beam(oracle_mod, n=128, ns=24, time=0.0, **kw):
    p = dict(nx=n, ny=n, ns=ns, iBeam=1, iMotion=600, heave_amp=0.1, heave_freq=1.0, iSponge=0)
    p.update(kw)
    o = oracle_mod.Oracle(**p)
    o.integrate_position(time)
    o.create_mask(time)
    return o


def _window(b, t, dx, LE_x):
    """cal_drag's bounds (NavierStokes.f90:538-545) for a beam without a
    cylinder, in the positive quadrant: Fortran int() truncates"""
    ixmin = int(np.floor(min(LE_x - 1.5 * t, b["x"].min() - 1.5 * t) / dx))
    ixmax = int(np.floor((b["x"].max() + 1.5 * t) / dx))
    iymin = int(np.floor((b["y"].min() - 1.5 * t) / dx))
    iymax = int(np.floor((b["y"].max() + 1.5 * t) / dx))
    return ixmin, ixmax, iymin, iymax


def test_drag_window_holds_the_sharp_mask(oracle_mod):
    """with N_smooth = 0 the mask is 1/eps within t_beam of the beam and 0
    elsewhere, all inside the window: Force is dx dy times the sum of the
    penalisation over the whole grid"""
    n = 128
    o = _heave_beam(oracle_mod, n=n, N_smooth=0.0)
    m, mvx, mvy = o.field("mask"), o.field("maskvx"), o.field("maskvy")
    assert set(np.unique(m)) == {0.0, 1.0 / 1e-3}
    rng = np.random.default_rng(7)
    u = rng.normal(size=(2, n, n))
    pen = np.stack([-m * (u[0] - mvx), -m * (u[1] - mvy)])
    o.cal_drag(0.0, 2e-4, pen)
    F = np.array(o.beam_arrays()["Force"])
    dxdy = (4.0 / n) ** 2
    want = np.array([pen[0].sum(), pen[1].sum()]) * dxdy
    np.testing.assert_allclose(F, want, rtol=1e-12)
    # a field that is not zero off the body is summed over the window only
    o.cal_drag(0.0, 2e-4, np.ones((2, n, n)))
    F1 = np.array(o.beam_arrays()["Force"])
    assert F1[0] < n * n * dxdy / 4


@pytest.mark.parametrize("shift", [0.0, 0.4])
def test_drag_window_bounds_truncate(oracle_mod, shift):
    """the leading edge is placed so that (xmax + 1.5 t_beam)/dx = k + 0.7:
    int() gives column k (nint would give k + 1); moving the beam by 0.4 dx
    takes the bound to k + 1.1 and adds column k + 1.  With penal = 1 the
    Force is dx dy times the number of window pixels"""
    n, t, L = 128, 0.0312, 1.0
    dx = 4.0 / n
    k = 80
    x0 = (k + 0.7) * dx - 1.5 * t - L + shift * dx
    o = _heave_beam(oracle_mod, n=n, x0=x0, y0=2.0 + 0.3 * dx)
    b = o.beam_arrays()
    assert b["x"].max() == pytest.approx(x0 + L, abs=1e-12)
    ixmin, ixmax, iymin, iymax = _window(b, t, dx, x0)
    assert ixmax == (k if shift == 0.0 else k + 1)
    o.cal_drag(0.0, 2e-4, np.ones((2, n, n)))
    F = np.array(o.beam_arrays()["Force"])
    cnt = (ixmax - ixmin + 1) * (iymax - iymin + 1)
    assert F[0] == pytest.approx(cnt * dx * dx, rel=1e-14) and F[1] == F[0]


def test_force_unst_of_a_rigid_heave(oracle_mod):
    """Force_unst = d/dt of sum(eps mask u_s) dx dy over the window
    (NavierStokes.f90:577-589): for the rigid heave y = A sin(2 pi t) at the
    top of its stroke (t = 1/4, velocity 0, acceleration -A (2 pi)^2) it is
    the penalised body's volume times the acceleration, and 0 across the
    heave; the first call has no previous value and gives 0"""
    n, A, dt = 128, 0.1, 1e-6
    T = 0.25
    o = oracle_mod.Oracle(nx=n, ny=n, ns=24, iBeam=1, iMotion=600, heave_amp=A, heave_freq=1.0, iSponge=0)
    zero = np.zeros((2, n, n))
    out = []
    for tt in (T - dt, T):
        o.integrate_position(tt)
        o.create_mask(tt)
        o.cal_drag(tt, dt, zero)
        out.append(np.array(o.beam_arrays()["Force_unst"]))
    assert np.all(out[0] == 0.0)
    b = o.beam_arrays()
    dx = 4.0 / n
    ixmin, ixmax, iymin, iymax = _window(b, 0.0312, dx, 1.0)
    vol = 1e-3 * o.field("mask")[iymin:iymax + 1, ixmin:ixmax + 1].sum() * dx * dx
    acc = -A * (2 * np.pi) ** 2 * np.cos(2 * np.pi * dt / 2)  # mean acceleration over the step
    assert out[1][0] == 0.0
    assert out[1][1] == pytest.approx(vol * acc, rel=1e-6)


# ---------------------------------------------------------------------------
# BeamSurfaceIntegral, stresses
# ---------------------------------------------------------------------------
def _tractions(pts, A, B):
    """sum over surface points (x, y, nx, ny, tx, ty, w) of the normal and
    tangential tractions of the stress S = [[A, B], [B, -A]], with A and B
    functions of position"""
    Fn, Ft = np.zeros(2), np.zeros(2)
    for x, y, nx_, ny_, tx_, ty_, w in pts:
        a, b = A(x, y), B(x, y)
        S = np.array([[a, b], [b, -a]])
        n = np.array([nx_, ny_])
        tv = np.array([tx_, ty_])
        tr = S @ n
        Fn += w * (n @ tr) * n
        Ft += w * (tv @ tr) * tv
    return Fn, Ft


def _surface_points(b, p, dx, alpha=0.0, nc=128):
    """BeamSurfaceIntegral's quadrature: two half circles of radius
    t_beam + N dx around the end nodes (nc points each, both ends
    included), and the two sides at that distance along the normals"""
    rad = p["t_beam"] + p["N_smooth"] * dx
    ns = len(b["x"])
    ds = p["length"] / (ns - 1)
    dphi = np.pi / (nc - 1)
    pts = []
    beta = b["theta"][-1] + alpha
    for i in range(nc):  # trailing half circle
        ph = -(0.5 * np.pi - beta) + i * dphi
        pts.append((b["x"][-1] + rad * np.cos(ph), b["y"][-1] + rad * np.sin(ph), np.cos(ph), np.sin(ph),
                    -np.sin(ph), np.cos(ph), dphi * rad))
    beta = b["theta"][1] + alpha  # the leading circle takes node 1's angle (BeamForces.f90:390)
    for i in range(nc):
        ph = (0.5 * np.pi + beta) + i * dphi
        pts.append((b["x"][0] + rad * np.cos(ph), b["y"][0] + rad * np.sin(ph), np.cos(ph), np.sin(ph),
                    -np.sin(ph), np.cos(ph), dphi * rad))
    for x, y, th in zip(b["x"], b["y"], b["theta"] + alpha):  # upper and lower sides: outward normals
        pts.append((x - rad * np.sin(th), y + rad * np.cos(th), -np.sin(th), np.cos(th), np.cos(th), np.sin(th),
                    ds))
    for x, y, th in zip(b["x"], b["y"], b["theta"] + alpha):
        pts.append((x + rad * np.sin(th), y - rad * np.cos(th), np.sin(th), -np.cos(th), -np.cos(th), -np.sin(th),
                    ds))
    return pts


@pytest.mark.parametrize("bent", [False, True])
def test_surface_integral_stresses_of_a_constant_stress(oracle_mod, bent):
    n, ns = 128, 24
    kw = dict(nx=n, ny=n, ns=ns, iBeam=1, iMotion=0, iSponge=0, t_beam=0.0312, N_smooth=2.0, length=1.0,
              x0=1.2, y0=2.0)
    o = oracle_mod.Oracle(**kw)
    s = np.linspace(0.0, 1.0, ns)
    th = 0.4 * np.sin(2.0 * s) if bent else np.full(ns, 0.3)
    x = 1.2 + np.concatenate([[0.0], np.cumsum(np.cos(th[:-1]) * np.diff(s))])
    y = 2.0 + np.concatenate([[0.0], np.cumsum(np.sin(th[:-1]) * np.diff(s))])
    o.set_beam_arrays(x=x, y=y, vx=np.zeros(ns), vy=np.zeros(ns), theta=th, theta_dot=np.zeros(ns))
    # linear stresses, which the Delta interpolation reproduces exactly
    Af = lambda x, y: 0.37 + 0.5 * (x - 2.0) - 0.3 * (y - 2.0)  # noqa: E731
    Bf = lambda x, y: -0.81 + 0.2 * (x - 2.0) + 0.6 * (y - 2.0)  # noqa: E731
    X, Y, dx = _grid(n)
    z = np.zeros((n, n))
    bsi = o.surface_integral(0.0, np.zeros((2, n, n)), Af(X, Y), Bf(X, Y), z)
    Fn, Ft = _tractions(_surface_points(o.beam_arrays(), kw, dx), Af, Bf)
    # bsi: drag lift drag_surf lift_surf Ft1 Fn1 Fp1 Ft2 Fn2 Fp2
    np.testing.assert_allclose([bsi[4], bsi[7]], Ft, rtol=1e-11)
    np.testing.assert_allclose([bsi[5], bsi[8]], Fn, rtol=1e-11)
    assert bsi[6] == 0.0 and bsi[9] == 0.0
    np.testing.assert_allclose([bsi[2], bsi[3]], Fn + Ft, rtol=1e-11)
    assert np.abs(Ft).min() > 1e-2 and np.abs(Fn).min() > 1e-2
