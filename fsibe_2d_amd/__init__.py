"""fsibe_2d_amd -- MI355X (gfx950) implementation of the FSIBE_2D fluid step.

The compute lives in libfsibe_hip.so (HIP, C ABI in include/fsibe_hip.h).
This package is the host-side mirror of the reference's interface for that
path: `spectral` (the FFT/operator plugin) and `solver` (the device-resident
step and the time loop with prescribed beam kinematics).
"""
from ._lib import FsibeError, lib  # noqa: F401
from .params import Params, batch_configs, fsi_256, fsi_large, turbulence_4096  # noqa: F401

__all__ = ["FsibeError", "lib", "Params", "batch_configs", "fsi_256", "fsi_large", "turbulence_4096"]
