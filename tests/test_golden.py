"""The committed golden vectors (tests/golden/, written by make_golden.py from
the reference's own Fortran routines and from numpy's FFT) against the
oracle and the host kinematics on the CPU, and against the GPU plugin."""
import os

import numpy as np
import pytest

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OPS = ["cofdx", "cofdy", "cofdxdx", "cofdydy", "cofdxdy", "poisson"]


def _load(name):
    return np.load(os.path.join(G, name))


def test_oracle_operators_match_reference_vectors(oracle_mod):
    d = _load("ref_operators_32x24.npz")
    o = oracle_mod.Oracle(nx=int(d["nx"]), ny=int(d["ny"]), xl=float(d["xl"]), yl=float(d["yl"]),
                          nu=float(d["nu"]), iBeam=0, iSponge=0)
    for name in OPS:
        assert np.array_equal(getattr(o, name)(d["fk"]), d[name]), name
    assert np.array_equal(o.vis(float(d["vis_dt"])), d["vis"])
    assert np.array_equal(o.dealiase_mask(), d["dealiase"])


def test_oracle_coftxy_matches_numpy_packing(oracle_mod):
    d = _load("packed_fft_64x48.npz")
    o = oracle_mod.Oracle(nx=64, ny=48, iBeam=0, iSponge=0)
    assert np.abs(o.coftxy(d["f"]) - d["fk"]).max() < 1e-16 * 64 * 48
    assert np.abs(o.cofitxy(d["fk"]) - o.cofitxy(o.coftxy(d["f"]))).max() < 1e-13


def test_host_kinematics_match_reference_vectors():
    from fsibe_2d_amd.params import Params
    from fsibe_2d_amd.solver import integrate_position, mouvement

    d = _load("ref_kinematics.npz")
    x0, y0, ang, ph = (float(d[k]) for k in ("x0", "y0", "angle", "phase"))
    for a, im in enumerate(d["motions"]):
        p = Params(nx=64, ny=64, ns=8, iMotion=int(im), x0=x0, y0=y0, AngleBeam=ang)
        for b, t in enumerate(d["times"]):
            al, at, att, le = mouvement(p, float(t), phase=ph)
            assert np.array_equal(np.array([al, at, att, *le]), d["mouvement"][a, b]), (im, t)
    ns = int(d["ns"])
    for k, (im, t) in enumerate(d["ip_cases"]):
        p = Params(nx=64, ny=64, ns=ns, iMotion=int(im), x0=x0, y0=y0, AngleBeam=ang, length=float(d["length"]))
        got = integrate_position(p, float(t), d["theta"][k], d["theta_dot"][k])
        assert np.array_equal(np.array(got), d["positions"][k]), im


@pytest.mark.gpu
def test_gpu_plugin_matches_golden_vectors():
    from fsibe_2d_amd.spectral import Spectral

    d = _load("ref_operators_32x24.npz")
    s = Spectral(int(d["nx"]), int(d["ny"]), float(d["xl"]), float(d["yl"]), float(d["nu"]))
    for name in OPS:
        assert np.array_equal(getattr(s, name)(d["fk"]), d[name]), name
    assert np.abs(s.vis(float(d["vis_dt"])) - d["vis"]).max() <= 2.3e-16
    assert np.array_equal(s.dealiase_mask(), d["dealiase"])
    s.close()
    d = _load("packed_fft_64x48.npz")
    s = Spectral(64, 48, 4.0, 4.0, 0.0)
    assert np.abs(s.coftxy(d["f"]) - d["fk"]).max() < 1e-16 * 64 * 48
    assert np.abs(s.cofitxy(d["fk"]) - s.cofitxy(s.coftxy(d["f"]))).max() < 1e-13
    s.close()
