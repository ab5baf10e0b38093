"""Restart files written and read by the library (fsibe_backup_*,
fsibe_inicond_*): the runtime_backup file holds the device state in the
reference's records (MakeRuntimeBackup, save_fields.f90:109-152), a Fortran
reader with init_fields' read list reads the same values, and a simulation
restored from it continues exactly as the one that wrote it
(init_fields.f90:44-97, time_step.f90:78-79: no startup step, create_mask
again from the stored beam)."""
import os
import subprocess

import numpy as np
import pytest

from fsibe_2d_amd import restart_files as rf

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
READER = os.path.join(ROOT, "tests", "fortran", "build", "read_backup")


def _sim(p, seed=0):
    from fsibe_2d_amd.inputs import band_limited_noise
    from fsibe_2d_amd.solver import FluidSim
    from fsibe_2d_amd.spectral import Spectral

    sp = Spectral(p.nx, p.ny, p.xl, p.yl, p.nu)
    vk = band_limited_noise(sp, sp.dealiase_mask(), seed=seed, amp=1e-3, kmax2=64)
    sp.close()
    g = FluidSim(p)
    g.set_vortk(vk)
    return g


def _vis_packed(p, dt):
    nx, ny = p.nx, p.ny
    sx, sy = (2 * np.pi / p.xl) ** 2, (2 * np.pi / p.yl) ** 2
    kx, ky = np.arange(nx // 2), np.arange(ny // 2)
    v = np.exp((-dt * p.nu * sx) * (kx[None, :] ** 2).astype(float) + (-dt * p.nu * sy) * (ky[:, None] ** 2).astype(float))
    return np.repeat(np.repeat(v, 2, axis=0), 2, axis=1)


def _same(a, b, tag):
    """The file's packed spectra hold one entry of each pair (kx = 0, +-ky),
    the device's half spectrum both, which the FFTs leave conjugate only to
    rounding; a restored state is therefore the written one symmetrised, and
    the runs agree to rounding rather than bit for bit."""
    from tests.conftest import rel_l2

    errs = {k: rel_l2(a.field(k), b.field(k)) for k in ("vortk", "u1", "u2", "vor_mean", "mask")}
    print(f"[restart {tag}] " + " ".join(f"{k}={v:.1e}" for k, v in errs.items()), flush=True)
    assert all(v <= 1e-13 for v in errs.values()), errs
    assert a.time_state()[:2] == b.time_state()[:2]
    oa, ob = a.step_out(), b.step_out()
    for k in ("Force", "Force_press", "energies"):
        x, y = np.asarray(oa[k]), np.asarray(ob[k])
        assert np.abs(x - y).max() <= 1e-11 * max(np.abs(y).max(), 1e-300), k
    ba, bb = a.beam(), b.beam()
    for k in ba:
        assert np.abs(ba[k] - bb[k]).max() <= 1e-11 * max(np.abs(bb[k]).max(), 1e-300), k


def test_backup_file_contents(tmp_path):
    from fsibe_2d_amd import Params

    p = Params(nx=96, ny=64, ns=32, dt_fixed=2e-4)
    g = _sim(p)
    g.run(12)
    path = tmp_path / "runtime_backup0.in"
    g.backup_write(path, ivideo=5, colorscale=2.5, colorscale_done=True)
    d = rf.read_backup(path, p.nx, p.ny, p.ns)
    t, dt1, _ = g.time_state()
    assert d["checksum"] == 123.0 and d["time"] == t and d["dt1"] == dt1 and d["ivideo"] == 5
    assert d["colorscale"] == 2.5 and d["colorscale_done"]
    n1 = d["n1"]
    n0 = 1 - n1
    assert np.array_equal(d["vortk"][n0], g.field("vortk").reshape(p.ny, p.nx))
    assert np.array_equal(d["u"][0], g.field("u1").reshape(p.ny, p.nx))
    assert np.array_equal(d["u"][1], g.field("u2").reshape(p.ny, p.nx))
    assert np.array_equal(d["vor_mean"], g.field("vor_mean").reshape(p.ny, p.nx))
    assert np.array_equal(d["mask"], g.field("mask").reshape(p.ny, p.nx))
    np.testing.assert_allclose(d["workvis"][n0], _vis_packed(p, dt1), rtol=4e-16, atol=0)
    np.testing.assert_allclose(d["workvis"][n1], _vis_packed(p, dt1), rtol=4e-16, atol=0)  # dt fixed
    b, gb = d["beams"][0], g.beam()
    for k in ("x", "y", "vx", "vy", "theta"):
        assert np.array_equal(b[k], gb[k]), k
    assert np.array_equal(b["pressure_new"], gb["pressure"]) and np.array_equal(b["tau_old"], gb["tau"])
    out = g.step_out()
    assert list(b["Force"]) == list(out["Force"]) and list(b["Force_unst"]) == list(out["Force_unst"])
    assert b["UnsteadyCorrectionsReady"] and b["iMouvement"] == p.iMotion
    if os.path.exists(READER):
        txt = subprocess.run([READER, str(path), str(p.nx), str(p.ny), str(p.ns)], capture_output=True, text=True,
                             check=True).stdout
        got = {ln.split()[0]: ln.split()[1:] for ln in txt.splitlines() if ln.strip()}
        assert int(got["n1"][0]) == n1 and float(got["time"][0]) == pytest.approx(t, rel=1e-16)
        # sums in another order: within the rounding of the sum of magnitudes
        for lev in (0, 1):
            a = d["vortk"][lev]
            assert abs(float(got["vortk"][lev]) - a.sum()) <= 1e-13 * np.abs(a).sum()
        a = d["vor_mean"]
        assert abs(float(got["vor_mean"][0]) - a.sum()) <= 1e-13 * np.abs(a).sum()
    g.close()


@pytest.mark.parametrize("coupled", [False, True])
def test_restart_continues_the_run(tmp_path, coupled):
    """run 12 steps, back up, restore into a fresh simulation, run 9 more:
    the state of 21 steps in one simulation, to rounding (for the coupled beam the
    reference's restart begins the BDF2 solver again with a Crank-Nicolson
    step; the continuous run is given the same StartupStep)"""
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solver import FluidSim

    p = Params(nx=96, ny=64, ns=32, dt_fixed=2e-4, iFLUSI=1 if coupled else 0, mue=1.0, eta=0.1)
    a = _sim(p)
    a.run(12)
    path = tmp_path / "runtime_backup1.in"
    a.backup_write(path, ivideo=3)
    b = FluidSim(p)
    info = b.backup_read(path)
    assert info == {"ivideo": 3, "colorscale": 0.0, "colorscale_done": False}
    if coupled:
        s = a.solid
        i = s.info()
        s.set_info(i["dt_old"], i["grav"], True)
    a.run(9)
    b.run(9)
    _same(a, b, "coupled" if coupled else "prescribed")
    if coupled:
        td_a, td_b = a.solid.get("theta_dot"), b.solid.get("theta_dot")
        assert np.abs(td_a - td_b).max() <= 1e-9 * np.abs(td_b).max()
        assert np.abs(a.solid.get("theta")).max() > 0
    a.close()
    b.close()


@pytest.mark.parametrize("coupled", [False, True])
def test_read_then_write_gives_the_same_file(tmp_path, coupled):
    """a backup written right after fsibe_backup_read, before any step, is
    the file that was read, byte for byte: the AB2 history nlk(n0) (the last
    nonlinear term times workvis of its step) and workvis(n1), which the
    device does not hold in that form, and the masks come from the file"""
    from fsibe_2d_amd import Params
    from fsibe_2d_amd.solver import FluidSim

    p = Params(nx=96, ny=64, ns=32, dt_fixed=2e-4, iFLUSI=1 if coupled else 0, mue=1.0, eta=0.1)
    a = _sim(p)
    a.run(7)
    pa, pb = tmp_path / "a.in", tmp_path / "b.in"
    a.backup_write(pa, ivideo=4, colorscale=1.5, colorscale_done=True)
    b = FluidSim(p)
    b.backup_read(pa)
    b.backup_write(pb, ivideo=4, colorscale=1.5, colorscale_done=True)
    assert pa.read_bytes() == pb.read_bytes()
    d = rf.read_backup(pb, p.nx, p.ny, p.ns)
    n0 = 1 - d["n1"]
    assert np.abs(d["workvis"][d["n1"]]).min() > 0 and np.abs(d["nlk"][n0]).max() > 0
    # after a step the device state is written again, as after any step
    if coupled:  # the restart begins the BDF2 beam solver with a Crank-Nicolson step
        i = a.solid.info()
        a.solid.set_info(i["dt_old"], i["grav"], True)
    a.run(1)
    b.run(1)
    a.backup_write(pa)
    b.backup_write(pb)
    da, db = rf.read_backup(pa, p.nx, p.ny, p.ns), rf.read_backup(pb, p.nx, p.ny, p.ns)
    for k in ("vortk", "nlk", "workvis"):
        x, y = np.asarray(da[k]), np.asarray(db[k])
        assert np.abs(x - y).max() <= 1e-13 * np.abs(x).max(), k
    a.close()
    b.close()


def test_inicond_file(tmp_path):
    """the final vorticity file holds level 0 (after an even number of steps
    the newest) and reads back as vortk(:,:,0) = coftxy(vorticity)"""
    from fsibe_2d_amd import FsibeError, Params
    from fsibe_2d_amd.solver import FluidSim

    p = Params(nx=96, ny=64, ns=32, dt_fixed=2e-4)
    a = _sim(p)
    a.run(4)
    path = tmp_path / "inicond"
    a.inicond_write(path)
    nx, ny, v = rf.read_inicond(path)
    assert (nx, ny) == (p.nx, p.ny)
    assert np.array_equal(v.ravel(), a.field("vort").ravel())
    b = FluidSim(p)
    b.inicond_read(path)
    va, vb = a.field("vortk"), b.field("vortk")
    assert np.linalg.norm(va - vb) <= 1e-14 * np.linalg.norm(va)
    b.run(2)
    with pytest.raises(FsibeError):
        b.inicond_read(path)  # only before the first step
    c = FluidSim(Params(nx=64, ny=64, ns=32))
    with pytest.raises(FsibeError):
        c.inicond_read(path)  # another resolution
    for s in (a, b, c):
        s.close()
