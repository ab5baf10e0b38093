"""ctypes binding of libfsibe_hip.so (C ABI declared in include/fsibe_hip.h).

The library is built in-tree by fsibe_2d_amd.build_ext.  There is no CPU
fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfsibe_hip.so")

OK, EARG, ESIZE, EHIP, EUNSUPPORTED, EDOMAIN, ESTATE = range(7)


class FsibeError(RuntimeError):
    """Nonzero status of a libfsibe_hip call (the reference STOPs instead)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"fsibe error {code}: {msg}")
        self.code = code


class FsibeParams(C.Structure):
    """fsibe_params (include/fsibe_hip.h): share_vars values of the step."""

    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("ns", C.c_int32),
        ("iMeanVelocity", C.c_int32), ("iSponge", C.c_int32), ("iBeam", C.c_int32),
        ("iWalls", C.c_int32), ("iCylinder", C.c_int32), ("iSharpTrailing", C.c_int32),
        ("iFLUSI", C.c_int32), ("iMotion", C.c_int32), ("iViscous", C.c_int32),
        ("iSaveBeam", C.c_int32), ("idealis", C.c_int32),
        ("xl", C.c_double), ("yl", C.c_double), ("nu", C.c_double), ("eps", C.c_double),
        ("eps_sponge", C.c_double), ("cfl", C.c_double), ("dt_fixed", C.c_double),
        ("Time_end", C.c_double),
        ("N_smooth", C.c_double), ("t_beam", C.c_double), ("length", C.c_double),
        ("SpongeSize", C.c_double), ("R_cylinder", C.c_double), ("h_channel", C.c_double),
        ("x0", C.c_double), ("y0", C.c_double), ("AngleBeam", C.c_double),
        ("theta_inf", C.c_double), ("T_fullspeed", C.c_double), ("T_release", C.c_double),
        ("tau", C.c_double), ("U_mean_true", C.c_double),
        ("heave_amp", C.c_double), ("heave_freq", C.c_double),
        ("motion_coef", C.c_double * 8),
        ("mue", C.c_double), ("eta", C.c_double), ("sigma", C.c_double), ("grav", C.c_double),
        ("TimeMethodSolid", C.c_int32), ("iImpulse", C.c_int32),
        ("iSmoothCutoff", C.c_int32),
    ]


class FsibeStepOut(C.Structure):
    _fields_ = [
        ("dt1", C.c_double), ("u_max", C.c_double),
        ("Force", C.c_double * 2), ("Force_unst", C.c_double * 2), ("Force_press", C.c_double * 2),
        ("energies", C.c_double * 6), ("bsi", C.c_double * 10),
        ("fx_int", C.c_double), ("fy_int", C.c_double),
        ("status", C.c_int32), ("ibes_iter", C.c_int32),
    ]


class FsibeSolidInfo(C.Structure):
    """fsibe_solid_info: scalars of type(solid) and of the solid solver."""

    _fields_ = [
        ("dt_old", C.c_double), ("grav", C.c_double), ("E_kinetic", C.c_double), ("E_pot", C.c_double),
        ("E_elastic", C.c_double), ("Inertial_Force", C.c_double * 2),
        ("StartupStep", C.c_int32), ("iter", C.c_int32), ("brake", C.c_int32), ("method", C.c_int32),
    ]


SOLID_ARRAYS = ("x", "y", "vx", "vy", "theta", "theta_dot", "ax", "ay", "pressure_old", "pressure_new",
                "tau_old", "tau_new", "oldold")

# every symbol include/fsibe_hip.h declares
EXPORTS = (
    "fsibe_last_error", "fsibe_spectral_create", "fsibe_spectral_destroy", "fsibe_coftxy",
    "fsibe_cofitxy", "fsibe_cofdx", "fsibe_cofdy", "fsibe_cofdxdx", "fsibe_cofdydy",
    "fsibe_cofdxdy", "fsibe_poisson", "fsibe_vis", "fsibe_dealiase_mask", "fsibe_dev_coftxy", "fsibe_dev_coftxy_noalloc",
    "fsibe_dev_cofitxy", "fsibe_dev_spectral_op", "fsibe_dev_vis", "fsibe_spectral_sync",
    "fsibe_sim_create", "fsibe_sim_destroy", "fsibe_set_vortk", "fsibe_create_mask",
    "fsibe_evolve_explicit", "fsibe_evolve_ab2", "fsibe_get_forces", "fsibe_get_step_out",
    "fsibe_swap_levels", "fsibe_get_field", "fsibe_get_box", "fsibe_mouvement",
    "fsibe_integrate_position", "fsibe_run", "fsibe_get_time", "fsibe_get_beam",
    "fsibe_sim_stream", "fsibe_profile_enable", "fsibe_profile_read", "fsibe_cofts", "fsibe_cofits",
    "fsibe_batch_create", "fsibe_batch_destroy", "fsibe_batch_run", "fsibe_batch_size", "fsibe_batch_member",
    "fsibe_batch_configure", "fsibe_batch_describe", "fsibe_batch_set_tiles",
    "fsibe_solid_create", "fsibe_solid_destroy", "fsibe_solid_get", "fsibe_solid_set", "fsibe_solid_get_info",
    "fsibe_solid_set_info", "fsibe_solid_step", "fsibe_solid_guess", "fsibe_solid_residual",
    "fsibe_solid_jacobian", "fsibe_solid_rhs", "fsibe_sim_solid", "fsibe_solid_wrapper", "fsibe_gravity_impulse",
    "fsibe_adjoint_create", "fsibe_adjoint_destroy", "fsibe_adjoint_objective", "fsibe_adjoint_gradient",
    "fsibe_adjoint_timings", "fsibe_backup_write", "fsibe_backup_read", "fsibe_inicond_write",
    "fsibe_inicond_read", "fsibe_pressure_snapshot", "fsibe_save_field", "fsibe_save_gif",
    "fsibe_create_mask_beams", "fsibe_get_step_out_beam", "fsibe_beam_placement", "fsibe_integrate_position_beam",
    "fsibe_solid_create_beam", "fsibe_sim_solid_beam", "fsibe_adjoint_create_binomial", "fsibe_adjoint_stats",
    "fsibe_adjoint_plan", "fsibe_adjoint_plan_cost",
)

_lib = None


def lib():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FsibeError(EHIP, f"{LIB_PATH} is not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    vp = C.c_void_p
    pvp = C.POINTER(C.c_void_p)
    L.fsibe_last_error.restype = C.c_char_p
    L.fsibe_spectral_create.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, pvp]
    L.fsibe_spectral_destroy.argtypes = [vp]
    for n in ("fsibe_coftxy", "fsibe_cofitxy", "fsibe_cofdx", "fsibe_cofdy", "fsibe_cofdxdx",
              "fsibe_cofdydy", "fsibe_cofdxdy", "fsibe_poisson"):
        getattr(L, n).argtypes = [vp, dp, dp]
    L.fsibe_vis.argtypes = [vp, C.c_double, dp]
    L.fsibe_dealiase_mask.argtypes = [vp, dp]
    L.fsibe_dev_coftxy.argtypes = [vp, vp, vp, vp]
    L.fsibe_dev_cofitxy.argtypes = [vp, vp, vp, vp]
    L.fsibe_dev_coftxy_noalloc.argtypes = [C.c_int, C.c_int, vp, vp, vp]
    L.fsibe_dev_spectral_op.argtypes = [vp, C.c_int, vp, vp, vp]
    L.fsibe_dev_vis.argtypes = [vp, C.c_double, vp, vp]
    L.fsibe_spectral_sync.argtypes = [vp]
    L.fsibe_sim_create.argtypes = [C.POINTER(FsibeParams), pvp]
    L.fsibe_sim_destroy.argtypes = [vp]
    L.fsibe_set_vortk.argtypes = [vp, dp]
    L.fsibe_create_mask.argtypes = [vp, C.c_double, dp, dp, dp, dp, dp, C.c_double, dp]
    L.fsibe_evolve_explicit.argtypes = [vp, C.c_double, C.POINTER(C.c_double)]
    L.fsibe_evolve_ab2.argtypes = [vp, C.c_double, C.c_double, C.POINTER(C.c_double)]
    L.fsibe_get_forces.argtypes = [vp, C.c_double, dp, dp, dp]
    L.fsibe_get_step_out.argtypes = [vp, C.POINTER(FsibeStepOut)]
    L.fsibe_swap_levels.argtypes = [vp]
    L.fsibe_get_field.argtypes = [vp, C.c_int, dp]
    L.fsibe_get_box.argtypes = [vp, C.c_int, vp, C.POINTER(C.c_int)]
    L.fsibe_mouvement.argtypes = [C.POINTER(FsibeParams), C.c_int, C.c_double, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.POINTER(C.c_double),
                                  C.POINTER(C.c_double), C.POINTER(C.c_double), dp]
    L.fsibe_integrate_position.argtypes = [C.POINTER(FsibeParams), C.c_int, C.c_double, dp, dp,
                                           dp, dp, dp, dp]
    L.fsibe_run.argtypes = [vp, C.c_int, C.POINTER(FsibeStepOut)]
    L.fsibe_get_time.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int)]
    L.fsibe_get_beam.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp]
    L.fsibe_sim_stream.restype = vp
    L.fsibe_sim_stream.argtypes = [vp]
    L.fsibe_profile_enable.argtypes = [vp, C.c_int]
    L.fsibe_batch_create.argtypes = [C.POINTER(FsibeParams), C.c_int, C.POINTER(C.c_void_p), C.POINTER(vp)]
    L.fsibe_batch_destroy.argtypes = [vp]
    L.fsibe_batch_run.argtypes = [vp, C.c_int]
    L.fsibe_batch_size.argtypes = [vp]
    L.fsibe_batch_member.restype = vp
    L.fsibe_batch_member.argtypes = [vp, C.c_int]
    L.fsibe_batch_configure.argtypes = [vp, C.c_int, C.c_int]
    L.fsibe_batch_describe.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.fsibe_batch_set_tiles.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.c_int]
    L.fsibe_cofts.argtypes = [dp, dp, C.c_int, C.c_int]
    L.fsibe_cofits.argtypes = [dp, dp, C.c_int, C.c_int]
    L.fsibe_profile_read.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong), C.c_int]
    L.fsibe_solid_create.argtypes = [C.POINTER(FsibeParams), pvp]
    L.fsibe_solid_destroy.argtypes = [vp]
    L.fsibe_solid_get.argtypes = [vp, C.c_int, dp]
    L.fsibe_solid_set.argtypes = [vp, C.c_int, dp]
    L.fsibe_solid_get_info.argtypes = [vp, C.POINTER(FsibeSolidInfo)]
    L.fsibe_solid_set_info.argtypes = [vp, C.POINTER(FsibeSolidInfo)]
    L.fsibe_solid_step.argtypes = [vp, C.c_double, C.c_double]
    L.fsibe_solid_guess.argtypes = [vp, C.c_double, C.c_double, dp]
    L.fsibe_solid_residual.argtypes = [vp, C.c_double, C.c_double, dp, dp]
    L.fsibe_solid_jacobian.argtypes = [vp, C.c_double, C.c_double, dp, dp]
    L.fsibe_solid_rhs.argtypes = [vp, C.c_double, dp, dp, dp, dp, dp, dp, dp]
    L.fsibe_sim_solid.argtypes = [vp, pvp]
    L.fsibe_solid_wrapper.argtypes = [vp, C.c_double, C.c_double]
    L.fsibe_gravity_impulse.argtypes = [C.POINTER(FsibeParams), C.c_double, C.POINTER(C.c_double)]
    L.fsibe_adjoint_create.argtypes = [C.POINTER(FsibeParams), vp, C.c_int, C.c_int, pvp]
    L.fsibe_adjoint_destroy.argtypes = [vp]
    L.fsibe_adjoint_objective.argtypes = [vp, dp, C.POINTER(C.c_double)]
    L.fsibe_adjoint_gradient.argtypes = [vp, dp, C.POINTER(C.c_double), dp]
    L.fsibe_adjoint_timings.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.fsibe_adjoint_create_binomial.argtypes = [C.POINTER(FsibeParams), vp, C.c_int, C.c_int, C.c_int, pvp]
    L.fsibe_adjoint_stats.argtypes = [vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]
    L.fsibe_adjoint_plan.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int),
                                     C.POINTER(C.c_longlong)]
    L.fsibe_adjoint_plan_cost.argtypes = [C.c_int, C.c_int]
    L.fsibe_adjoint_plan_cost.restype = C.c_longlong
    L.fsibe_backup_write.argtypes = [vp, C.c_char_p, C.c_int, C.c_double, C.c_int]
    L.fsibe_backup_read.argtypes = [vp, C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_double),
                                    C.POINTER(C.c_int)]
    L.fsibe_inicond_write.argtypes = [vp, C.c_char_p]
    L.fsibe_inicond_read.argtypes = [vp, C.c_char_p]
    L.fsibe_pressure_snapshot.argtypes = [vp, C.c_double, C.c_int, dp]
    L.fsibe_save_field.argtypes = [C.c_char_p, dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_char_p,
                                   C.c_char_p]
    L.fsibe_save_gif.argtypes = [C.c_char_p, dp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    L.fsibe_create_mask_beams.argtypes = [vp, C.c_double, dp, dp, dp, dp, dp, dp, dp]
    L.fsibe_get_step_out_beam.argtypes = [vp, C.c_int, C.POINTER(FsibeStepOut)]
    L.fsibe_beam_placement.argtypes = [C.POINTER(FsibeParams), C.c_int, C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]
    L.fsibe_integrate_position_beam.argtypes = [C.POINTER(FsibeParams), C.c_int, C.c_double, dp, dp,
                                                dp, dp, dp, dp]
    L.fsibe_solid_create_beam.argtypes = [C.POINTER(FsibeParams), C.c_int, pvp]
    L.fsibe_sim_solid_beam.argtypes = [vp, C.c_int, pvp]
    _lib = L
    return L


def check(rc: int):
    if rc != OK:
        raise FsibeError(rc, lib().fsibe_last_error().decode())
