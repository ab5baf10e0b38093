"""Step-level parity on the GPU: the fused device step against the oracle's
restatement of the reference's call sequence (time_step.f90:120-341 with the
beam heaving as prescribed kinematics).

Bars: vorticity relative L2 <= 1e-10 after 100 steps; forces and
beam loads <= 1e-9 relative, with an absolute floor of 1e-6 x the largest
magnitude of that quantity (lift is ~0 by symmetry early in the run).
"""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _pair(oracle_mod, n, ns=64, **kw):
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solver import FluidSim

    p = Params(nx=n, ny=n, ns=ns, **kw)
    o = oracle_mod.Oracle(oracle_mod.make_params(**{k: getattr(p, k) for k in (
        "nx", "ny", "ns", "iMeanVelocity", "iSponge", "iBeam", "iWalls", "iCylinder",
        "iSharpTrailing", "iFLUSI", "iMotion", "iViscous", "iSaveBeam", "idealis", "xl", "yl",
        "nu", "eps", "eps_sponge", "cfl", "dt_fixed", "Time_end", "N_smooth", "t_beam",
        "length", "SpongeSize", "R_cylinder", "h_channel", "x0", "y0", "AngleBeam",
        "theta_inf", "T_fullspeed", "T_release", "tau", "U_mean_true", "heave_amp",
        "heave_freq")}))
    vk = o.init_noise(seed=0, amp=1e-3, kmax2=64)
    o.set_vortk(vk)
    g = FluidSim(p)
    g.set_vortk(vk)
    return o, g


def _close(a, b, rel, floor):
    a, b = np.asarray(a, float), np.asarray(b, float)
    return bool(np.all(np.abs(a - b) <= rel * np.maximum(np.abs(b), floor)))


def _compare(o, g, tag, tol_field=1e-10, tol_force=1e-9):
    d = o.diag()
    out = g.step_out()
    t, dt1, it = g.time_state()
    print(f"[{tag}] time {t:.6f} vs {d.time:.6f}, dt1 {dt1} vs {d.dt1}", flush=True)
    assert t == pytest.approx(d.time, rel=1e-14)
    assert dt1 == d.dt1
    errs = {}
    for name in ("vort", "u1", "u2", "vortk", "mask", "maskvx", "maskvy", "vor_mean"):
        errs[name] = rel_l2(g.field(name), o.field(name))
    print(f"[{tag}] field rel L2: " + " ".join(f"{k}={v:.2e}" for k, v in errs.items()), flush=True)
    for k in ("vort", "u1", "u2", "vortk", "vor_mean"):
        assert errs[k] <= tol_field, k
    for k in ("mask", "maskvx", "maskvy"):
        assert errs[k] <= 1e-12, k
    ob = o.beam_arrays()
    gb = g.beam()
    assert np.allclose(gb["y"], ob["y"], rtol=0, atol=1e-14)
    F = {"Force": (out["Force"], ob["Force"]), "Force_unst": (out["Force_unst"], ob["Force_unst"]),
         "Force_press": (out["Force_press"], ob["Force_press"]),
         "pressure": (gb["pressure"], ob["pressure"]), "tau": (gb["tau"], ob["tau"]),
         "energies": (out["energies"], list(d.energies)),
         "bsi": (out["bsi"], list(d.bsi)[:10])}
    for k, (a, b) in F.items():
        a, b = np.asarray(a, float), np.asarray(b, float)
        floor = 1e-6 * max(np.abs(b).max(), 1e-300)
        err = np.max(np.abs(a - b) / np.maximum(np.abs(b), floor))
        print(f"[{tag}] {k}: max rel err {err:.2e}", flush=True)
        assert _close(a, b, tol_force, floor), k


@pytest.mark.parametrize("n", [64, 128])
def test_startup_step(n, oracle_mod):
    o, g = _pair(oracle_mod, n)
    o.run(1)
    g.run(1)
    _compare(o, g, f"startup {n}")


@pytest.mark.parametrize("n", [64, 128])
def test_ab2_steps(n, oracle_mod):
    o, g = _pair(oracle_mod, n)
    o.run(5)
    g.run(5)
    _compare(o, g, f"ab2x4 {n}")


@pytest.mark.parametrize("n", [128, 256])
def test_100_steps(n, oracle_mod):
    o, g = _pair(oracle_mod, n)
    o.run(100)
    g.run(100)
    _compare(o, g, f"100 steps {n}")


def test_pressure_field(oracle_mod):
    o, g = _pair(oracle_mod, 64)
    o.run(3)
    g.run(3)
    pg, po = g.field("press"), o.field("press")
    err = rel_l2(pg, po)
    print(f"pressure rel L2 {err:.2e}", flush=True)
    assert err < 1e-10
