"""The discrete adjoint (fsibe_adjoint_*): J(c) = (1/T) sum_n (-Force(1)^n) dt1^n
over a prescribed-motion run (motion 601, c = its 8 Fourier coefficients) and
dJ/dc by the reverse sweep.  The reference has no adjoint; the gradient is
pinned by central finite differences of the same GPU objective, whose
Force(1) terms are the cal_drag forces (NavierStokes.f90:567-575) that the
step's own parity tests hold to the oracle, and J itself is tied to the
step_out record of a plain FluidSim run."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(n, ns, seed=0, smooth=0):
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.inputs import band_limited_noise
    from fsibe_2d_amd.spectral import Spectral

    p = Params(nx=n, ny=n, ns=ns, iMotion=601, heave_freq=2.0, dt_fixed=2e-4, cfl=100.0, nu=1 / 500,
               iSaveBeam=0, iSmoothCutoff=smooth)
    sp = Spectral(n, n, p.xl, p.yl, p.nu)
    vk = band_limited_noise(sp, sp.dealiase_mask(), seed=seed, amp=1e-3, kmax2=64)
    sp.close()
    return p, vk


def _coef(seed=3):
    from fsibe_2d_amd.params import adjoint_coefficients

    return np.array(adjoint_coefficients(seed))


def _fd(a, c, h):
    fd = np.zeros(8)
    for k in range(8):
        e = np.zeros(8)
        e[k] = h
        fd[k] = (a.objective(c + e) - a.objective(c - e)) / (2 * h)
    return fd


@pytest.mark.parametrize("n,ns,steps,chk,smooth", [(64, 24, 30, 8, 0), (128, 48, 60, 16, 0),
                                                   (128, 48, 60, 16, 1)])
def test_gradient_against_finite_differences(n, ns, steps, chk, smooth):
    """each component of dJ/dc against (J(c + h e_k) - J(c - h e_k)) / 2h,
    h = 1e-6: relative error within 1e-6 (central differences of a smooth
    J carry O(h^2) truncation and O(eps J / h) rounding, ~1e-9 here).  Over
    these short runs the beam moves too little for a pixel to cross the
    reference's cut-offs, so both cut-off forms are smooth here."""
    from fsibe_2d_amd.adjoint import Adjoint

    p, vk = _case(n, ns, smooth=smooth)
    c = _coef()
    a = Adjoint(p, vk, steps, chk)
    J, g = a.gradient(c)
    assert a.objective(c) == J
    fd = _fd(a, c, 1e-6)
    rel = np.abs(fd - g) / np.abs(fd)
    print(f"[adjoint {n} cutoff {smooth}] J {J:.10e} grad {g} rel {rel}", flush=True)
    assert np.all(np.abs(g) > 1e-3)
    assert rel.max() <= 1e-6
    a.close()


def test_gradient_independent_of_checkpoint_spacing():
    """the checkpoint spacing changes only what is recomputed, not the
    arithmetic: J and dJ/dc are bit-identical for chk = 1, 7 and nsteps"""
    from fsibe_2d_amd.adjoint import Adjoint

    p, vk = _case(64, 24, seed=5)
    c = _coef(7)
    res = []
    for chk in (1, 7, 21):
        a = Adjoint(p, vk, 21, chk)
        res.append(a.gradient(c))
        a.close()
    for J, g in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(g, res[0][1])


def test_binomial_snapshot_bounds_give_identical_gradients():
    """fewer snapshot slots than unit starts (the revolve schedule) change
    only how often units are recomputed: J and dJ/dc are bit-identical for 0,
    1, 2 and 3 slots and for a slot per unit start (22 steps in 8 units of 3,
    the last unit 1 step), and the steps run forward are the schedule's
    least count of advanced units in steps, plus each step recorded once"""
    from fsibe_2d_amd.adjoint import Adjoint, plan

    p, vk = _case(64, 24, seed=6)
    c = _coef(9)
    nsteps, chk = 22, 3
    units = [min(chk, nsteps - k * chk) for k in range(8)]
    res = []
    for snaps in (None, 3, 2, 1, 0):
        a = Adjoint(p, vk, nsteps, chk, snapshots=snaps)
        res.append(a.gradient(c))
        st = a.stats()
        nslots = 7 if snaps is None else snaps
        acts, _ = plan(8, nslots)
        adv = sum(sum(units[u:u + n]) for op, u, n in acts if op == 1)
        print(f"[adjoint binomial] slots {st['slots']} advanced {st['steps_advanced']} "
              f"recorded {st['steps_recorded']}", flush=True)
        assert st["slots"] == nslots
        assert st["steps_recorded"] == nsteps
        assert st["steps_advanced"] == adv
        a.close()
    for J, g in res[1:]:
        assert J == res[0][0]
        assert np.array_equal(g, res[0][1])


def test_objective_is_the_force_record_of_a_plain_run():
    """J = (1/T) sum (-Force(1) dt1) from the step_out records of FluidSim
    stepping the same run one step at a time"""
    from fsibe_2d_amd.adjoint import Adjoint
    from fsibe_2d_amd.solver import FluidSim

    p, vk = _case(64, 24, seed=1)
    c = _coef(11)
    p.motion_coef = list(c)
    steps = 17
    a = Adjoint(p, vk, steps, 5)
    J = a.objective(c)
    a.close()
    g = FluidSim(p)
    g.set_vortk(vk)
    num = den = 0.0
    for _ in range(steps):
        out = g.run(1)
        num += -out["Force"][0] * out["dt1"]
        den += out["dt1"]
    g.close()
    assert J == pytest.approx(num / den, rel=1e-13)


def _workload(smooth):
    from fsibe_2d_amd.inputs import band_limited_noise
    from fsibe_2d_amd.params import adjoint_512
    from fsibe_2d_amd.spectral import Spectral

    p = adjoint_512(smooth_cutoff=smooth)
    sp = Spectral(p.nx, p.ny, p.xl, p.yl, p.nu)
    vk = band_limited_noise(sp, sp.dealiase_mask(), seed=0, amp=1e-3, kmax2=64)
    sp.close()
    return p, vk


def test_differentiation_workload():
    """the differentiation workload as specified (512^2, Re 500, dt 1e-4,
    1000 steps, checkpoints every 50, coefficients 0.05 N(0,1) of seed 3)
    with the continuous mask (iSmoothCutoff = 1): every component of the
    gradient against the central difference with h = 1e-6, within 1e-6"""
    from fsibe_2d_amd.adjoint import Adjoint

    p, vk = _workload(True)
    c = _coef()
    a = Adjoint(p, vk, 1000, 50)
    J, g = a.gradient(c)
    fd = _fd(a, c, 1e-6)
    rel = np.abs(fd - g) / np.abs(fd)
    print(f"[adjoint 512 x 1000] J {J:.10e} grad {g} rel {rel}", flush=True)
    assert rel.max() <= 1e-6
    a.close()


def test_full_size_directional_derivative():
    """the reference's cut-offs at the workload's grid and physics over 100
    steps (before the beam has moved far enough for J to jump): the
    gradient along a random direction d against the central difference of J
    along d"""
    from fsibe_2d_amd.adjoint import Adjoint
    from fsibe_2d_amd.inputs import normals

    p, vk = _workload(False)
    c = _coef()
    d = np.array(normals(4, 8))
    d /= np.linalg.norm(d)
    a = Adjoint(p, vk, 100, 50)
    J, g = a.gradient(c)
    h = 1e-6
    fd = (a.objective(c + h * d) - a.objective(c - h * d)) / (2 * h)
    print(f"[adjoint 512] J {J:.10e} g.d {g @ d:.10e} fd {fd:.10e}", flush=True)
    assert abs(g @ d - fd) <= 1e-6 * abs(fd)
    a.close()


def test_unsupported_configurations_are_refused():
    from fsibe_2d_amd import FsibeError, Params
    from fsibe_2d_amd.adjoint import Adjoint

    for kw in (dict(iFLUSI=1), dict(iMotion=600), dict(iSharpTrailing=1), dict(dt_fixed=0.0)):
        with pytest.raises(FsibeError):
            Adjoint(Params(nx=64, ny=64, ns=24, **({"iMotion": 601} | kw)), None, 10, 5)
