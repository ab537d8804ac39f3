"""ctypes bindings of the C-ABI declared in include/hamers_convective.h."""
from __future__ import annotations

import ctypes
import os

from . import LIB_PATH

HM_MAX_EQN = 12
HM_MAX_SPECIES = 8
HM_NUM_STAGES = 5
HM_MAX_RK_COEFFS = 5

HM_OK = 0
HM_ERR_INVALID_ARGUMENT = 1
HM_DT_INFO_SLOTS = 8
HM_ERR_UNSUPPORTED = 2
HM_ERR_DEVICE = 3
HM_ERR_OUT_OF_MEMORY = 4

HM_MODEL_SINGLE_SPECIES = 0
HM_MODEL_FOUR_EQN_CONSERVATIVE = 1
HM_MODEL_FIVE_EQN_ALLAIRE = 2

HM_BC_NONE, HM_BC_FLOW, HM_BC_REFLECT, HM_BC_SYMMETRY, HM_BC_DIRICHLET = range(5)

HM_SCHEME_WCNS5_JS_HLLC_HLL = 0
HM_SCHEME_WCNS6_LD_HLLC_HLL = 1


class hm_config(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int),
        ("model", ctypes.c_int),
        ("num_species", ctypes.c_int),
        ("scheme", ctypes.c_int),
        ("constant_p", ctypes.c_int),
        ("constant_q", ctypes.c_int),
        ("constant_C", ctypes.c_double),
        ("constant_alpha_tau", ctypes.c_double),
        ("species_gamma", ctypes.c_double * HM_MAX_SPECIES),
        ("species_R", ctypes.c_double * HM_MAX_SPECIES),
        ("device", ctypes.c_int),
    ]


class hm_patch_view(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("ghost", ctypes.c_int * 3),
        ("dx", ctypes.c_double * 3),
        ("Q", ctypes.c_void_p * HM_MAX_EQN),
        ("F", ctypes.c_void_p * 3),
        ("S", ctypes.c_void_p),
    ]


class hm_vjp_patch_view(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("ghost", ctypes.c_int * 3),
        ("dx", ctypes.c_double * 3),
        ("Q", ctypes.c_void_p * HM_MAX_EQN),
        ("F_bar", ctypes.c_void_p * 3),
        ("S_bar", ctypes.c_void_p),
        ("Q_bar", ctypes.c_void_p * HM_MAX_EQN),
    ]


class hm_rk_patch_view(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int * 3),
        ("ghost", ctypes.c_int * 3),
        ("ghost_intermediate", ctypes.c_int * 3),
        ("dx", ctypes.c_double * 3),
        ("Q", ctypes.c_void_p * HM_MAX_EQN),
        ("Q_intermediate", (ctypes.c_void_p * HM_MAX_EQN) * HM_MAX_RK_COEFFS),
        ("F_intermediate", (ctypes.c_void_p * 3) * HM_MAX_RK_COEFFS),
        ("S_intermediate", ctypes.c_void_p * HM_MAX_RK_COEFFS),
        ("F_accumulated", ctypes.c_void_p * 3),
        ("S_accumulated", ctypes.c_void_p),
        ("F_diffusive_intermediate", (ctypes.c_void_p * 3) * HM_MAX_RK_COEFFS),
        ("F_diffusive_accumulated", ctypes.c_void_p * 3),
    ]


class hm_schedule(ctypes.Structure):
    _fields_ = [("tile_rows", ctypes.c_int), ("xcd_remap", ctypes.c_int), ("order", ctypes.c_int),
                ("persistent", ctypes.c_int)]

    def as_dict(self):
        return {"tile_rows": self.tile_rows, "xcd_remap": self.xcd_remap, "order": self.order,
                "persistent": self.persistent}


class hm_transport(ctypes.Structure):
    _fields_ = [
        ("species_mu", ctypes.c_double * HM_MAX_SPECIES),
        ("species_mu_v", ctypes.c_double * HM_MAX_SPECIES),
        ("species_kappa", ctypes.c_double * HM_MAX_SPECIES),
        ("species_D", ctypes.c_double * HM_MAX_SPECIES),
        ("species_M", ctypes.c_double * HM_MAX_SPECIES),
    ]


class LibraryMissing(RuntimeError):
    pass


def _load():
    # torch ships its own HIP runtime (soname libamdhip64.so.7) and links it by
    # the unversioned name.  Importing torch first lets the dynamic linker bind
    # this library's libamdhip64.so.7 dependency to that same runtime, so that
    # device pointers and streams handed over from torch belong to the runtime
    # the kernels are launched on (one HIP runtime per process).
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} is not built: run `python -c 'import __graft_entry__ as g; g.build()'` "
            "(the HIP library is the only compute path; there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    L.hm_ctx_create.argtypes = [ctypes.POINTER(hm_config), ctypes.POINTER(P)]
    L.hm_ctx_create.restype = ctypes.c_int
    L.hm_ctx_destroy.argtypes = [P]
    L.hm_ctx_destroy.restype = ctypes.c_int
    L.hm_get_num_ghosts.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    L.hm_get_num_ghosts.restype = ctypes.c_int
    L.hm_get_num_eqn.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    L.hm_get_num_eqn.restype = ctypes.c_int
    L.hm_last_error.argtypes = [P]
    L.hm_last_error.restype = ctypes.c_char_p
    L.hm_compute_convective_flux_and_source_on_patch.argtypes = [
        P, ctypes.POINTER(hm_patch_view), ctypes.c_double, ctypes.c_double, ctypes.c_int, P]
    L.hm_compute_convective_flux_and_source_on_patch.restype = ctypes.c_int
    L.hm_compute_convective_flux_and_source_on_patch_noalloc.argtypes = [
        P, ctypes.POINTER(hm_patch_view), ctypes.c_double, ctypes.c_double, ctypes.c_int, P]
    L.hm_compute_convective_flux_and_source_on_patch_noalloc.restype = ctypes.c_int
    L.hm_compute_convective_flux_and_source_on_patches.argtypes = [
        P, ctypes.POINTER(hm_patch_view), ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, P]
    L.hm_compute_convective_flux_and_source_on_patches.restype = ctypes.c_int
    L.hm_get_num_conservative_arrays.argtypes = [P, ctypes.POINTER(ctypes.c_int)]
    L.hm_get_num_conservative_arrays.restype = ctypes.c_int
    L.hm_profile_enable.argtypes = [P, ctypes.c_int]
    L.hm_profile_enable.restype = ctypes.c_int
    L.hm_profile_read.argtypes = [P, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
    L.hm_profile_read.restype = ctypes.c_int
    L.hm_synchronize.argtypes = [P, P]
    L.hm_synchronize.restype = ctypes.c_int
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.POINTER(ctypes.c_int)
    for name in ("hm_advance_single_step_on_patch",):
        getattr(L, name).argtypes = [P, ctypes.POINTER(hm_rk_patch_view), ctypes.c_int, D, D, D, P]
    L.hm_advance_single_step_on_patches.argtypes = [P, ctypes.POINTER(hm_rk_patch_view), ctypes.c_int,
                                                    ctypes.c_int, D, D, D, P]
    L.hm_level_create.argtypes = [P, ctypes.c_int, I, I, I, I, I, D, ctypes.POINTER(P)]
    L.hm_level_destroy.argtypes = [P]
    L.hm_level_set_rk_coefficients.argtypes = [P, ctypes.c_int, D, D, D]
    L.hm_level_write_patch.argtypes = [P, ctypes.c_int, D]
    L.hm_level_read_patch.argtypes = [P, ctypes.c_int, D]
    L.hm_level_patch_array.argtypes = [P, ctypes.c_int, ctypes.POINTER(P), I]
    L.hm_level_fill_ghosts.argtypes = [P, P]
    L.hm_level_set_boundary_conditions.argtypes = [P, I, D]
    L.hm_level_set_boundary_conditions.restype = ctypes.c_int
    S = ctypes.POINTER(hm_schedule)
    L.hm_autotune_on_patches.argtypes = [P, ctypes.POINTER(hm_patch_view), ctypes.c_int, ctypes.c_double,
                                         ctypes.c_int, P, S, D]
    L.hm_set_schedule_on_patches.argtypes = [P, ctypes.POINTER(hm_patch_view), ctypes.c_int, S]
    L.hm_get_schedule_on_patches.argtypes = [P, ctypes.POINTER(hm_patch_view), ctypes.c_int, S]
    L.hm_level_autotune.argtypes = [P, ctypes.c_double, ctypes.c_int, S, D]
    L.hm_level_set_schedule.argtypes = [P, S]
    L.hm_level_get_schedule.argtypes = [P, S]
    L.hm_hierarchy_create.argtypes = [P, P, ctypes.c_int, ctypes.POINTER(P)]
    L.hm_hierarchy_destroy.argtypes = [P]
    L.hm_hierarchy_advance.argtypes = [P, ctypes.c_double, ctypes.c_int, P]
    for name in ("hm_hierarchy_create", "hm_hierarchy_destroy", "hm_hierarchy_advance"):
        getattr(L, name).restype = ctypes.c_int
    for name in ("hm_autotune_on_patches", "hm_set_schedule_on_patches", "hm_get_schedule_on_patches",
                 "hm_level_autotune", "hm_level_set_schedule", "hm_level_get_schedule"):
        getattr(L, name).restype = ctypes.c_int
    L.hm_level_advance.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.c_int, P]
    L.hm_convective_flux_vjp_on_patch.argtypes = [P, ctypes.POINTER(hm_vjp_patch_view), ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_int, P]
    L.hm_convective_flux_vjp_on_patches.argtypes = [P, ctypes.POINTER(hm_vjp_patch_view), ctypes.c_int,
                                                    ctypes.c_double, ctypes.c_double, ctypes.c_int, P]
    L.hm_level_advance_recorded.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int, P]
    L.hm_level_adjoint.argtypes = [P, P]
    L.hm_level_set_face_state_budget.argtypes = [P, ctypes.c_int64]
    L.hm_level_face_states_kept.argtypes = [P]
    L.hm_level_write_adjoint.argtypes = [P, ctypes.c_int, D]
    L.hm_level_read_adjoint.argtypes = [P, ctypes.c_int, D]
    L.hm_set_diffusive_transport.argtypes = [P, ctypes.POINTER(hm_transport)]
    L.hm_get_diffusive_num_ghosts.argtypes = [P, I]
    L.hm_compute_diffusive_flux_on_patch.argtypes = [P, ctypes.POINTER(hm_patch_view), ctypes.c_double,
                                                     ctypes.c_double, ctypes.c_int, P]
    L.hm_compute_diffusive_flux_on_patches.argtypes = [P, ctypes.POINTER(hm_patch_view), ctypes.c_int,
                                                       ctypes.c_double, ctypes.c_double, ctypes.c_int, P]
    L.hm_compute_spectral_radiuses_and_stable_dt_on_patches.argtypes = [P, ctypes.POINTER(hm_patch_view), ctypes.c_int,
                                                                         D, P]
    L.hm_level_advance_cfl.argtypes = [P, ctypes.c_double, ctypes.c_int, ctypes.c_int, P]
    L.hm_level_dt_info.argtypes = [P, D]
    for name in ("hm_compute_spectral_radiuses_and_stable_dt_on_patches", "hm_level_advance_cfl",
                 "hm_level_dt_info"):
        getattr(L, name).restype = ctypes.c_int
    for name in ("hm_advance_single_step_on_patch", "hm_advance_single_step_on_patches", "hm_level_create",
                 "hm_level_destroy", "hm_level_set_rk_coefficients", "hm_level_write_patch",
                 "hm_level_read_patch", "hm_level_patch_array", "hm_level_fill_ghosts", "hm_level_advance",
                 "hm_convective_flux_vjp_on_patch", "hm_convective_flux_vjp_on_patches", "hm_level_advance_recorded",
                 "hm_level_adjoint", "hm_level_write_adjoint", "hm_level_read_adjoint",
                 "hm_level_set_face_state_budget", "hm_level_face_states_kept",
                 "hm_set_diffusive_transport", "hm_get_diffusive_num_ghosts", "hm_compute_diffusive_flux_on_patch",
                 "hm_compute_diffusive_flux_on_patches"):
        getattr(L, name).restype = ctypes.c_int
    return L


lib = _load()


class HamersError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"hamers_amd error {status}: {message}")
        self.status = status


def check(ctx, status: int):
    if status != HM_OK:
        msg = lib.hm_last_error(ctx)
        raise HamersError(status, msg.decode() if msg else "")
