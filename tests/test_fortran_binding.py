"""The reference-side Fortran binding on the GPU: two Fortran programs built on
the reference's own modules (share_vars, motion, integrate_position) call the
library through fsibe_2d_amd/fortran (cof_hip.f90, ops_hip.f90,
fluid_hip.f90) exactly as the reference's callers would, and their outputs are
compared with the oracle on the same inputs.  The programs are built by
`make -C fsibe_2d_amd/fortran drivers` (part of __graft_entry__.build() where
the reference sources exist); the tests skip when they are absent."""
import os
import subprocess

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "fortran", "build")


def _run(name, tmp_path, *args):
    exe = os.path.join(BIN, name)
    if not os.path.exists(exe):
        pytest.skip(f"{name} not built (needs the reference's modules)")
    out = str(tmp_path / (name + ".bin"))
    print(f"running {name} {' '.join(args)}", flush=True)
    r = subprocess.run([exe, out, *args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and os.path.exists(out), r.stdout + r.stderr
    return np.fromfile(out, dtype=np.uint8)


class _Reader:
    def __init__(self, buf):
        self.b, self.o = buf, 0

    def ints(self, n):
        v = np.frombuffer(self.b, dtype=np.int32, count=n, offset=self.o)
        self.o += 4 * n
        return [int(x) for x in v]

    def reals(self, n, shape=None):
        v = np.frombuffer(self.b, dtype=np.float64, count=n, offset=self.o).copy()
        self.o += 8 * n
        return v.reshape(shape) if shape else v


def test_plugin_through_fortran_shim(tmp_path, oracle_mod):
    rd = _Reader(_run("drive_cof_hip", tmp_path))
    nx, ny = rd.ints(2)
    sh = (ny, nx)
    f, fk, g = rd.reals(nx * ny, sh), rd.reals(nx * ny, sh), rd.reals(nx * ny, sh)
    ops = {k: rd.reals(nx * ny, sh) for k in ("cofdx", "cofdy", "cofdxdx", "cofdydy", "cofdxdy", "poisson")}
    vis, dea = rd.reals(nx * ny, sh), rd.reals(nx * ny, sh)
    L, n = rd.ints(2)
    t, tk, ti = (rd.reals(L * n, (n, L)) for _ in range(3))
    o = oracle_mod.Oracle(nx=nx, ny=ny, xl=4.0, yl=3.0, nu=1.3e-2, iBeam=0, iSponge=0)
    fk_o = o.coftxy(f)
    assert np.abs(fk - fk_o).max() <= 2e-14 * np.abs(fk_o).max()
    assert np.abs(g - o.cofitxy(fk)).max() <= 2e-14 * np.abs(f).max()
    for name, v in ops.items():  # elementwise on the same input: bit for bit
        assert np.array_equal(v, getattr(o, name)(fk)), name
    assert np.abs(vis - o.vis(0.37)).max() <= 2.3e-16
    assert np.array_equal(dea, o.dealiase_mask())
    c = np.fft.rfft(t, axis=1) / L
    inter = np.empty((n, 2 * c.shape[1]))
    inter[:, 0::2], inter[:, 1::2] = c.real, c.imag
    assert np.abs(tk - inter[:, :L]).max() < 1e-15
    k = (L - 1) // 2  # odd L: cofits(cofts(t)) loses the last imaginary part
    lost = -2 * c[:, k].imag[:, None] * np.sin(2 * np.pi * k * np.arange(L) / L)[None, :]
    assert np.abs(ti - (t - lost)).max() < 1e-13


def test_time_loop_through_fortran_module(tmp_path, oracle_mod):
    """startup + 9 AB2 steps of time_step.f90's loop driven from Fortran
    (reference mouvement case 2 and integrate_position) against the oracle"""
    rd = _Reader(_run("drive_step_hip", tmp_path))
    nx, ny, ns = rd.ints(3)
    f = rd.reals(nx * ny, (ny, nx))
    time, dt1 = rd.reals(2)
    vortk = rd.reals(nx * ny, (ny, nx))
    force, force_unst, force_press = rd.reals(2), rd.reals(2), rd.reals(2)
    p_beam, tau_beam, bx, by = (rd.reals(ns) for _ in range(4))
    o = oracle_mod.Oracle(nx=nx, ny=ny, ns=ns, xl=4.0, yl=4.0, nu=5e-3, eps=1e-3, eps_sponge=1e-3, cfl=1.0,
                          dt_fixed=2e-4, Time_end=1e6, N_smooth=2.0, t_beam=0.05, length=1.0,
                          SpongeSize=0.4, x0=1.0, y0=2.0, AngleBeam=5.0, iMeanVelocity=3, iSponge=3,
                          iBeam=1, iMotion=2, iViscous=1, iSaveBeam=1, idealis=1, U_mean_true=1.0)
    o.set_vortk(o.coftxy(f))
    o.run(10)
    d = o.diag()
    assert abs(time - d.time) < 1e-15 and dt1 == d.dt1
    err = rel_l2(vortk, o.field("vortk"))
    print(f"fortran loop: vortk rel L2 {err:.2e}", flush=True)
    assert err < 1e-10
    b = o.beam_arrays()
    assert np.array_equal(bx, b["x"]) and np.array_equal(by, b["y"])

    def close(a, e):
        e = np.asarray(e)
        floor = 1e-6 * max(np.abs(e).max(), 1e-300)
        return np.all(np.abs(np.asarray(a) - e) <= 1e-9 * np.maximum(np.abs(e), floor))

    assert close(force, b["Force"]) and close(force_press, b["Force_press"])
    assert close(p_beam, b["pressure"]) and close(tau_beam, b["tau"])
    assert close(force_unst, b["Force_unst"])


def test_two_beam_time_loop_through_fortran_module(tmp_path, oracle_mod):
    """the same loop with two beams in tandem (iMotion 303): the Fortran
    module's loops over beams(1:2) (create_mask with each beam's mouvement,
    cal_drag and GetForces per beam) against the oracle's"""
    rd = _Reader(_run("drive_step_hip", tmp_path, "2"))
    nx, ny, ns = rd.ints(3)
    f = rd.reals(nx * ny, (ny, nx))
    time, dt1 = rd.reals(2)
    vortk = rd.reals(nx * ny, (ny, nx))
    beams = []
    for _ in range(2):
        force, force_unst, force_press = rd.reals(2), rd.reals(2), rd.reals(2)
        beams.append((force, force_unst, force_press, *(rd.reals(ns) for _ in range(4))))
    o = oracle_mod.Oracle(nx=nx, ny=ny, ns=ns, xl=4.0, yl=4.0, nu=5e-3, eps=1e-3, eps_sponge=1e-3, cfl=1.0,
                          dt_fixed=2e-4, Time_end=1e6, N_smooth=2.0, t_beam=0.05, length=1.0,
                          SpongeSize=0.4, x0=0.8, y0=2.0, AngleBeam=5.0, iMeanVelocity=3, iSponge=3,
                          iBeam=2, iMotion=303, iViscous=1, iSaveBeam=1, idealis=1, U_mean_true=1.0)
    o.set_vortk(o.coftxy(f))
    o.run(10)
    d = o.diag()
    assert abs(time - d.time) < 1e-15 and dt1 == d.dt1
    err = rel_l2(vortk, o.field("vortk"))
    print(f"fortran two-beam loop: vortk rel L2 {err:.2e}", flush=True)
    assert err < 1e-10

    def close(a, e):
        e = np.asarray(e)
        floor = 1e-6 * max(np.abs(e).max(), 1e-300)
        return np.all(np.abs(np.asarray(a) - e) <= 1e-9 * np.maximum(np.abs(e), floor))

    for ib, (force, force_unst, force_press, p_beam, tau_beam, bx, by) in enumerate(beams):
        b = o.beam_arrays(ib)
        assert np.array_equal(bx, b["x"]) and np.array_equal(by, b["y"]), ib
        assert close(force, b["Force"]) and close(force_press, b["Force_press"]), ib
        assert close(p_beam, b["pressure"]) and close(tau_beam, b["tau"]), ib
        assert close(force_unst, b["Force_unst"]), ib
    assert not np.array_equal(beams[0][0], beams[1][0])
