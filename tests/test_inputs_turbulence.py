"""The random-phase decaying-turbulence start of the solver-only workload
(fsibe_2d_amd.inputs.decaying_turbulence), pinned on the CPU by properties a
wrong generator fails:

* the mean kinetic energy (1/2)<|u|^2> of the velocity numpy derives from the
  vorticity in physical space is exactly the requested energy (a generator
  scaling psi instead of omega, or missing the conjugate half, fails);
* the shell-summed spectrum of that velocity is C k^4 exp(-2 (k/k0)^2) per
  shell (a per-mode instead of per-shell normalisation, or another exponent,
  fails);
* the field is real (the kx = 0 column is Hermitian);
* the packed layout equals the oracle's coftxy of the physical field (the
  oracle's packing is pinned by the golden vectors of numpy's FFT);
* the amplitudes do not depend on the seed, the phases do.
"""
import numpy as np
import pytest

from fsibe_2d_amd.inputs import decaying_turbulence, half_to_packed, turbulence_spectrum


def _full_spectrum(z):
    """numpy rfft2 layout (n, n/2+1) of the half spectrum z (n, n/2): the
    x-Nyquist column is zero"""
    n = z.shape[0]
    out = np.zeros((n, n // 2 + 1), dtype=complex)
    out[:, : n // 2] = z
    return out


def _velocity(z):
    """u = d psi/dy, v = -d psi/dx with omega = -lap psi, on the 2 pi grid,
    by numpy's inverse FFT (an evaluation independent of the generator)"""
    n = z.shape[0]
    zf = _full_spectrum(z)
    ky = np.fft.fftfreq(n, 1.0 / n)[:, None]
    kx = np.arange(n // 2 + 1)[None, :].astype(float)
    k2 = kx ** 2 + ky ** 2
    k2[0, 0] = 1.0
    psi = zf / k2
    psi[0, 0] = 0.0
    u = np.fft.irfft2(1j * ky * psi, s=(n, n)) * n * n
    v = np.fft.irfft2(-1j * kx * psi, s=(n, n)) * n * n
    w = np.fft.irfft2(zf, s=(n, n)) * n * n
    return u, v, w


@pytest.mark.parametrize("n", [64, 128, 256])
def test_mean_energy_is_exact(n):
    z = decaying_turbulence(n, packed=False)
    u, v, _ = _velocity(z)
    e = 0.5 * np.mean(u * u + v * v)
    assert abs(e - 1.0) <= 1e-13
    z2 = decaying_turbulence(n, packed=False, energy=0.25)
    u2, v2, _ = _velocity(z2)
    assert abs(0.5 * np.mean(u2 * u2 + v2 * v2) - 0.25) <= 1e-13


@pytest.mark.parametrize("n,k0", [(128, 12.0), (256, 12.0), (128, 6.0)])
def test_shell_spectrum_is_the_target(n, k0):
    z = decaying_turbulence(n, packed=False, k0=k0)
    u, v, _ = _velocity(z)
    # per-mode energy over the full plane from numpy's forward FFT
    U = np.fft.rfft2(u) / (n * n)
    V = np.fft.rfft2(v) / (n * n)
    em = 0.5 * (np.abs(U) ** 2 + np.abs(V) ** 2)
    kx = np.arange(n // 2 + 1)
    wt = np.where((kx > 0) & (kx < n // 2), 2.0, 1.0)[None, :]
    ky = np.fft.fftfreq(n, 1.0 / n)[:, None]
    shell = np.floor(np.sqrt(kx[None, :] ** 2 + ky ** 2) + 0.5).astype(int)
    Es = np.bincount(shell.ravel(), weights=(em * wt).ravel())
    s = np.arange(Es.size)
    # every shell, the partial ones beyond n/2 too, is normalised by its own
    # mode count, and C = 1 / sum_s E(s) over the shells that hold modes
    target = turbulence_spectrum(s, k0)
    C = 1.0 / target[Es > 0].sum()
    want = C * target
    err = np.abs(Es - want * (Es > 0))
    # relative per shell, absolute against the peak shell where the shell's
    # energy is below the rounding of the FFT of the whole field
    assert np.all(err <= 1e-10 * want + 1e-14 * want.max()), (err / np.maximum(want, 1e-300)).max()
    assert np.count_nonzero(want > 1e-6 * want.max()) >= 10
    assert abs(Es.sum() - 1.0) <= 1e-13


def test_field_is_real():
    n = 96
    z = decaying_turbulence(n, packed=False)
    full = np.zeros((n, n), dtype=complex)
    full[:, : n // 2] = z
    j = np.arange(n)
    for kx in range(1, n // 2):
        full[(-j) % n, n - kx] = np.conj(z[j, kx])
    w = np.fft.ifft2(full) * n * n
    assert np.abs(w.imag).max() <= 1e-13 * np.abs(w.real).max()
    # the ky = -n/2 row and the mean are absent
    assert np.all(z[n // 2] == 0) and z[0, 0] == 0


def test_packed_layout_matches_the_oracle_coftxy(oracle_mod):
    n = 64
    z = decaying_turbulence(n, packed=False)
    _, _, w = _velocity(z)
    o = oracle_mod.Oracle(nx=n, ny=n, xl=2 * np.pi, yl=2 * np.pi, iBeam=0, iSponge=0)
    P = decaying_turbulence(n)
    ref = o.coftxy(w)
    assert np.abs(P - ref).max() <= 1e-14 * np.abs(ref).max()
    # and back: the oracle's cofitxy of the packed field is the vorticity
    assert np.abs(o.cofitxy(P) - w).max() <= 1e-12 * np.abs(w).max()


def test_half_to_packed_single_modes():
    """each complex mode pair lands on the entries the packed layout gives
    a cos/sin product (tests/modes.py's table)"""
    from tests.modes import packed_of

    n = 16
    rng = np.random.default_rng(5)
    for _ in range(20):
        kx, ky = int(rng.integers(1, n // 2)), int(rng.integers(1, n // 2))
        a, b = rng.normal(size=2)
        # f = a cos(kx x + ky y) + b sin(kx x - ky y)
        #   = a (cc - ss) + b (sc - cs)
        z = np.zeros((n, n // 2), dtype=complex)
        z[ky, kx] += a / 2
        z[n - ky, kx] += b / 2j
        want = packed_of(n, n, [(kx, ky, "c", "c", a), (kx, ky, "s", "s", -a),
                                (kx, ky, "s", "c", b), (kx, ky, "c", "s", -b)])
        assert np.allclose(half_to_packed(z), want, rtol=0, atol=1e-15)


def test_seed_changes_phases_not_amplitudes():
    a = decaying_turbulence(64, seed=1, packed=False)
    b = decaying_turbulence(64, seed=2, packed=False)
    assert np.allclose(np.abs(a), np.abs(b), rtol=1e-15, atol=0)
    assert not np.allclose(a, b)
    assert np.array_equal(a, decaying_turbulence(64, seed=1, packed=False))


def test_nonlinear_term_is_not_trivial(oracle_mod):
    """the start's u . grad(omega) is of the order of omega (k0 = 12):
    |nlk| / |omega_k| > 1 in L2, so the GPU parity of this workload tests
    the nonlinear term (a single-shell field has u . grad(omega) = 0)"""
    from fsibe_2d_amd import turbulence_4096
    from tests.test_gpu_variants import PKEYS

    p = turbulence_4096(128)
    vk = decaying_turbulence(128)
    o = oracle_mod.Oracle(oracle_mod.make_params(**{k: getattr(p, k) for k in PKEYS}))
    o.set_vortk(vk)
    o.run(2)  # after the startup step and one AB2 step, nlk(n0) is N(omega_0)
    nl = np.linalg.norm(o.field("nlk"))
    assert nl > 1.0 * np.linalg.norm(vk), (nl, np.linalg.norm(vk))
    # kinetic energy (cal_energy, the integral of |u|^2/2 over the 2 pi square)
    assert o.diag().energies[1] == pytest.approx(4 * np.pi ** 2, rel=1e-6)
