"""Two beams (iBeam = 2, the dragonfly motions 300..307) on the GPU against the
oracle's loops over beams: init_beam.f90:44-76 (placement, phase),
create_mask.f90:74-119 (beam after beam), NavierStokes.f90:528-589 (cal_drag
per beam and window), BeamForces.f90:24-47 (GetForces per beam),
time_step.f90:199-218 (a solid solve or integrate_position per beam).

Bars as for one beam: vorticity relative L2 <= 1e-10; every per-beam force,
load and surface integral <= 1e-9 relative with an absolute floor of 1e-6 x
the largest magnitude of that quantity; masks <= 1e-12."""
import numpy as np
import pytest

from tests.conftest import rel_l2
from tests.test_gpu_fsi import PKEYS

pytestmark = pytest.mark.gpu


def _params(**kw):
    from fsibe_2d_amd import Params

    base = dict(nx=128, ny=128, ns=32, iBeam=2, iMotion=302, x0=0.9, y0=2.0, dt_fixed=2e-4)
    base.update(kw)
    return Params(**base)


def _oracle(oracle_mod, p):
    return oracle_mod.Oracle(oracle_mod.make_params(**{k: getattr(p, k) for k in PKEYS}))


def _noise(oracle_mod, p, seed=5):
    o0 = oracle_mod.Oracle(nx=p.nx, ny=p.ny, iBeam=0, iSponge=0)
    return o0.init_noise(seed=seed, amp=1e-3, kmax2=64)


def _err(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    floor = 1e-6 * max(np.abs(b).max(), 1e-300)
    return float(np.max(np.abs(a - b) / np.maximum(np.abs(b), floor)))


def _compare_beams(g, o, tag, tol=1e-9, beam_tol=1e-14):
    for ib in range(2):
        out, gb, ob = g.step_out(beam=ib), g.beam(ib), o.beam_arrays(ib)
        bsi = o.beam_bsi(ib)
        assert np.abs(gb["y"] - ob["y"]).max() <= beam_tol * np.abs(ob["y"]).max(), (tag, ib)
        assert np.abs(gb["x"] - ob["x"]).max() <= beam_tol * np.abs(ob["x"]).max(), (tag, ib)
        F = {"Force": (out["Force"], ob["Force"]), "Force_unst": (out["Force_unst"], ob["Force_unst"]),
             "Force_press": (out["Force_press"], ob["Force_press"]), "pressure": (gb["pressure"], ob["pressure"]),
             "tau": (gb["tau"], ob["tau"]), "bsi": (out["bsi"], bsi[:10]),
             "fxy_int": ((out["fx_int"], out["fy_int"]), bsi[10:])}
        errs = {k: _err(a, b) for k, (a, b) in F.items()}
        print(f"[{tag} beam {ib}] " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
        for k, e in errs.items():
            assert e <= tol, (tag, ib, k, F[k])
    # the two beams carry different loads (different phase and position)
    assert g.step_out(beam=0)["Force"] != g.step_out(beam=1)["Force"]


@pytest.mark.parametrize("im", [300, 302, 305])
def test_tandem_prescribed_motion(im, oracle_mod):
    """fsibe_run with both beams heaving as prescribed (iFLUSI = 0), forces
    every step (iSaveBeam = 1): 30 steps against the oracle's loop"""
    from fsibe_2d_amd.solver import FluidSim

    p = _params(iMotion=im)
    vk = _noise(oracle_mod, p)
    o = _oracle(oracle_mod, p)
    o.set_vortk(vk)
    g = FluidSim(p)
    g.set_vortk(vk)
    for chunk in (1, 9, 20):
        o.run(chunk)
        g.run(chunk)
        d = o.diag()
        t = g.time_state()[0]
        assert t == pytest.approx(d.time, rel=1e-14)
        errs = {k: rel_l2(g.field(k), o.field(k)) for k in ("vortk", "mask", "maskvx", "maskvy", "u1", "u2")}
        print(f"[im={im} t={t:.4f}] " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
        assert errs["vortk"] <= 1e-10 and errs["u1"] <= 1e-10 and errs["u2"] <= 1e-10
        for k in ("mask", "maskvx", "maskvy"):
            assert errs[k] <= 1e-12, k
        out = g.step_out()
        assert _err(out["energies"], list(d.energies)) <= 1e-9
        _compare_beams(g, o, f"im={im} t={t:.4f}")
    # the second beam is where integrate_position puts it with its own phase
    x0, y0, ph = o.beam_arrays(1)["x0"], o.beam_arrays(1)["y0"], o.beam_arrays(1)["phase"]
    assert x0 == pytest.approx(p.x0 + 1.25, abs=1e-15) and ph == pytest.approx((im - 300) * np.pi / 4, abs=1e-15)


def test_tandem_step_api(oracle_mod):
    """the per-step calls: fsibe_create_mask_beams with each beam's
    mouvement(time), evolve, GetForces of both beams"""
    from fsibe_2d_amd.solver import FluidSim

    p = _params(iMotion=306)
    vk = _noise(oracle_mod, p, seed=8)
    o = _oracle(oracle_mod, p)
    o.set_vortk(vk)
    g = FluidSim(p)
    g.set_vortk(vk)
    t = 0.35
    o.integrate_position(t)
    arr = [o.beam_arrays(ib) for ib in range(2)]
    mv = [o.mouvement(t, ib) for ib in range(2)]
    st = lambda k: np.stack([a[k] for a in arr])
    g.create_mask(t, st("x"), st("y"), st("vx"), st("vy"), st("theta"), np.array([m[0] for m in mv]),
                  np.stack([m[3] for m in mv]))
    o.host_create_mask(t)
    for k in ("mask", "maskvx", "maskvy"):
        assert rel_l2(g.field(k), o.field(k)) <= 1e-12, k
    dt_g = g.evolve_fluid_explicit(t)
    dt_o = o.host_evolve(t, explicit=True)
    assert dt_g == dt_o
    pb, tb, fp = g.get_forces(t)
    o.host_get_forces(t)
    assert pb.shape == (2, p.ns) and fp.shape == (2, 2)
    for ib in range(2):
        ob = o.beam_arrays(ib)
        assert _err(pb[ib], ob["pressure"]) <= 1e-9 and _err(tb[ib], ob["tau"]) <= 1e-9
        assert _err(fp[ib], ob["Force_press"]) <= 1e-9
        assert _err(g.step_out(beam=ib)["Force"], ob["Force"]) <= 1e-9
    with pytest.raises(Exception):  # one beam's call on a two-beam simulation
        from fsibe_2d_amd._lib import check, lib

        check(lib().fsibe_create_mask(g.h, t, arr[0]["x"], arr[0]["y"], arr[0]["vx"], arr[0]["vy"],
                                      arr[0]["theta"], mv[0][0], mv[0][3]))


def test_tandem_coupled(oracle_mod):
    """iFLUSI = 1: each beam solved by its own solid solver (SolidSolver.f90:25-27)
    from its own loads, 20 steps against two oracle solid solvers"""
    from oracle.solid import SolidOracle
    from fsibe_2d_amd.solver import FluidSim

    p = _params(iMotion=304, iFLUSI=1, mue=1.0, eta=0.1)
    vk = _noise(oracle_mod, p, seed=2)
    o = _oracle(oracle_mod, p)
    sos = [SolidOracle(o.beam_view(ib), mue=p.mue, eta=p.eta, sigma=p.sigma, grav=p.grav,
                       TimeMethodSolid=p.TimeMethodSolid, iImpulse=p.iImpulse) for ib in range(2)]
    o.set_vortk(vk)
    time, dt0 = 0.0, 0.0
    nsteps = 20
    for step in range(nsteps):
        for ib, so in enumerate(sos):
            o.set_beam_arrays(ib, x=so.x, y=so.y, vx=so.vx, vy=so.vy, theta=so.theta, theta_dot=so.theta_dot)
        o.host_create_mask(time)
        dt1 = o.host_evolve(time, explicit=True) if step == 0 else o.host_evolve(time, dt0)
        o.host_get_forces(time)
        for ib, so in enumerate(sos):
            b = o.beam_arrays(ib)
            so.pressure_new, so.tau_new = b["pressure"].copy(), b["tau"].copy()
            so.step(time, dt1)
        o.host_swap()
        time, dt0 = time + dt1, dt1
    g = FluidSim(p)
    g.set_vortk(vk)
    out = g.run(nsteps)
    assert g.time_state()[0] == pytest.approx(time, rel=1e-14)
    e_v = rel_l2(g.field("vortk"), o.field("vortk"))
    print(f"[coupled] vortk {e_v:.1e}", flush=True)
    assert e_v <= 1e-10
    for ib, so in enumerate(sos):
        gb, ob = g.beam(ib), o.beam_arrays(ib)
        gout = g.step_out(beam=ib)
        e_th = np.abs(gb["theta"] - so.theta).max() / max(np.abs(so.theta).max(), 1e-300)
        e_tip = abs(gb["y"][-1] - so.y[-1]) / abs(so.y[-1] - p.y0)
        print(f"[coupled beam {ib}] theta {e_th:.1e} tip {e_tip:.1e} iters {gout['ibes_iter']}/{so.iter}",
              flush=True)
        assert e_th <= 1e-9 and e_tip <= 1e-9
        assert gout["ibes_iter"] == so.iter
        for k in ("Force", "Force_press"):
            assert _err(gout[k], ob[k]) <= 1e-9, (ib, k)
        s = g.solid_of(ib)
        assert np.array_equal(s.get("theta"), gb["theta"])
    assert out["Force"] == g.step_out(beam=0)["Force"]
    # the beams bend differently (opposite heave phase, the second in the first's wake)
    assert np.abs(g.beam(0)["theta"] - g.beam(1)["theta"]).max() > 0.0


@pytest.mark.parametrize("coupled", [False, True])
def test_tandem_restart(tmp_path, coupled, oracle_mod):
    """six backup records per beam (save_fields.f90:132-139): the file holds
    each beam at its own placement, and a restored run continues the
    uninterrupted one (12 + 9 against 21 steps)"""
    from fsibe_2d_amd import restart_files as rf
    from fsibe_2d_amd.solver import FluidSim

    p = _params(nx=96, ny=64, iMotion=301, iFLUSI=1 if coupled else 0, mue=1.0, eta=0.1)
    vk = _noise(oracle_mod, p, seed=4)
    a = FluidSim(p)
    a.set_vortk(vk)
    a.run(12)
    path = tmp_path / "runtime_backup0.in"
    a.backup_write(path)
    d = rf.read_backup(path, p.nx, p.ny, p.ns, nbeams=2)
    for ib in range(2):
        bd, gb = d["beams"][ib], a.beam(ib)
        assert bd["x0"] == p.x0 + 1.25 * ib and bd["y0"] == p.y0
        assert np.array_equal(bd["y"], gb["y"]) and np.array_equal(bd["pressure_new"], gb["pressure"])
        assert tuple(bd["Force"]) == a.step_out(beam=ib)["Force"]
    b = FluidSim(p)
    b.backup_read(path)
    if coupled:  # the reference's restart begins each BDF2 solver with a Crank-Nicolson step
        for ib in range(2):
            s = a.solid_of(ib)
            i = s.info()
            s.set_info(i["dt_old"], i["grav"], True)
    a.run(9)
    b.run(9)
    errs = {k: rel_l2(a.field(k), b.field(k)) for k in ("vortk", "u1", "u2", "mask")}
    print(f"[tandem restart coupled={coupled}] " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
    assert all(v <= 1e-13 for v in errs.values()), errs
    for ib in range(2):
        oa, ob = a.step_out(beam=ib), b.step_out(beam=ib)
        for k in ("Force", "Force_press", "Force_unst"):
            x, y = np.asarray(oa[k]), np.asarray(ob[k])
            assert np.abs(x - y).max() <= 1e-11 * max(np.abs(y).max(), 1e-300), (ib, k)
        ba, bb = a.beam(ib), b.beam(ib)
        for k in ba:
            # the loads are differences of pressures interpolated on either
            # side of the beam, larger than the loads early in the soft start:
            # the restored state's rounding shows there relative to the load
            tol = 1e-9 if k in ("pressure", "tau") else 1e-11
            assert np.abs(ba[k] - bb[k]).max() <= tol * max(np.abs(bb[k]).max(), 1e-300), (ib, k)
    a.close()
    b.close()


def test_tandem_batch_member(oracle_mod):
    """a two-beam member of a batch beside one-beam members, prescribed and
    solved: bit-identical to its own run"""
    from fsibe_2d_amd.solver import Batch, FluidSim

    ps = [_params(iMotion=302), _params(iBeam=1, iMotion=2), _params(iMotion=306, iFLUSI=1, mue=1.0, eta=0.1),
          _params(nx=96, ny=96, iMotion=300)]
    vks = [_noise(oracle_mod, p, 30 + i) for i, p in enumerate(ps)]
    b = Batch(ps, vks)
    steps = 8
    b.run(steps)
    for i, p in enumerate(ps):
        g = FluidSim(p)
        g.set_vortk(vks[i])
        g.run(steps)
        m = b.members[i]
        assert m.time_state() == g.time_state(), i
        assert np.array_equal(m.field("vortk"), g.field("vortk")), i
        for ib in range(p.iBeam):
            for k in ("x", "y", "theta", "pressure", "tau"):
                assert np.array_equal(m.beam(ib)[k], g.beam(ib)[k]), (i, ib, k)
            so, go = m.step_out(beam=ib), g.step_out(beam=ib)
            for k in ("Force", "Force_unst", "Force_press", "bsi", "ibes_iter"):
                assert so[k] == go[k], (i, ib, k)
        g.close()
    b.close()
