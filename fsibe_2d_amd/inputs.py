"""Synthetic initial conditions for the benchmark workloads.

Random numbers: SplitMix64(seed) -> u = (x >> 11) 2^-53 -> Box-Muller, drawn in
column-major order (x fastest).  u1 = ((x >> 11) + 1) 2^-53 lies in (0, 1] so
the logarithm is finite.  The band-limited noise start replaces the
reference's inicond cases (init_fields.f90:38-290), which have no random
start: vort = amp * xi, vortk = coftxy(vort) with every (kx, ky) block with
kx^2 + ky^2 > kmax2 zeroed, then multiplied by the dealiasing mask.
"""
from __future__ import annotations

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int) -> np.ndarray:
    i = np.arange(1, n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + i * GAMMA
        z = (z ^ (z >> np.uint64(30))) * M1
        z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def normals(seed: int, n: int) -> np.ndarray:
    m = (n + 1) // 2
    z = splitmix64(seed, 2 * m)
    two53 = 1.0 / 9007199254740992.0
    u1 = ((z[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * two53
    u2 = (z[1::2] >> np.uint64(11)).astype(np.float64) * two53
    r = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = r * np.cos(6.283185307179586 * u2)
    out[1::2] = r * np.sin(6.283185307179586 * u2)
    return out[:n]


def band_limited_noise(spectral, dealiase, seed: int = 0, amp: float = 1e-3,
                       kmax2: int = 64) -> np.ndarray:
    """Packed vortk of band-limited Gaussian noise, computed with the GPU
    coftxy of `spectral` (a fsibe_2d_amd.spectral.Spectral)."""
    nx, ny = spectral.nx, spectral.ny
    f = (amp * normals(seed, nx * ny)).reshape(ny, nx)
    vk = spectral.coftxy(f)
    kx = np.arange(nx // 2)
    ky = np.arange(ny // 2)
    drop = (kx[None, :] ** 2 + ky[:, None] ** 2) > kmax2
    blk = np.repeat(np.repeat(drop, 2, axis=0), 2, axis=1)
    vk[blk] = 0.0
    return vk * dealiase
