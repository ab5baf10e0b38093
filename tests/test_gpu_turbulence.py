"""The solver-only decaying-turbulence workload (no mask, no sponge, fluid at
rest: NavierStokes.f90:9-212 with iBeam = 0, time_step.f90:178-230) from the
random-phase start of fsibe_2d_amd.inputs.decaying_turbulence, on the GPU
against the oracle: 100 steps at 512^2 and 3 steps at the workload's 4096^2.
The nonlinear term u . grad(omega) of this field is of the order of the
vorticity itself per unit time, so the expensive half of the step is
exercised (a single-shell field has u . grad(omega) = 0)."""
import numpy as np
import pytest

from tests.conftest import rel_l2
from tests.test_gpu_variants import PKEYS

pytestmark = pytest.mark.gpu


def _pair(oracle_mod, n):
    from fsibe_2d_amd import turbulence_4096
    from fsibe_2d_amd.inputs import decaying_turbulence
    from fsibe_2d_amd.solver import FluidSim

    p = turbulence_4096(n)
    vk = decaying_turbulence(n, seed=1)
    o = oracle_mod.Oracle(oracle_mod.make_params(**{k: getattr(p, k) for k in PKEYS}))
    o.set_vortk(vk)
    g = FluidSim(p)
    g.set_vortk(vk)
    return p, vk, o, g


def _check(o, g, tag):
    errs = {k: rel_l2(g.field(k), o.field(k)) for k in ("vort", "u1", "u2", "vortk", "vor_mean")}
    d = o.diag()
    out = g.step_out()
    print(f"[{tag}] t={d.time:.4f} " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
    assert g.time_state()[0] == pytest.approx(d.time, rel=1e-14)
    assert all(v <= 1e-10 for v in errs.values()), errs
    e_o, e_g = np.array(d.energies), np.array(out["energies"])
    floor = 1e-6 * np.abs(e_o).max()
    assert np.max(np.abs(e_g - e_o) / np.maximum(np.abs(e_o), floor)) <= 1e-9, (e_g, e_o)


def test_turbulence_512_100_steps(oracle_mod):
    p, vk, o, g = _pair(oracle_mod, 512)
    o.run(100)
    g.run(100)
    _check(o, g, "turbulence 512 x100")
    # the kinetic energy starts at 1/2 <|u|^2> = 1
    u1, u2 = g.field("u1"), g.field("u2")
    assert 0.9 < 0.5 * np.mean(u1 * u1 + u2 * u2) <= 1.0
    g.close()


def test_turbulence_4096(oracle_mod):
    """the workload's own size, 3 steps (the oracle takes seconds per step)"""
    p, vk, o, g = _pair(oracle_mod, 4096)
    o.run(3)
    g.run(3)
    _check(o, g, "turbulence 4096 x3")
    g.close()
