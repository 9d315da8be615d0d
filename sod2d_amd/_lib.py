"""Loader and ctypes signatures of libsod2d_hip.so (the C-ABI in include/sod2d_hip.h).

The library is built in-tree by ``python -m sod2d_amd.build``.  There is no
fallback: if the library is missing, every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsod2d_hip.so")

P = C.c_void_p
I = C.c_int
D = C.c_double
F = C.c_float
SZ = C.c_size_t


class SodhipError(RuntimeError):
    pass


class ViscParams(C.Structure):
    """sodhip_visc_params: module globals read by smart_visc_spectral."""
    _fields_ = [("ce", D), ("cmax", D), ("cglob", D), ("Rgas", D),
                ("flag_normalise_entropy", I), ("flag_drop_c_in_envit", I)]


class MLP(C.Structure):
    """sodhip_mlp: device pointers of the surrogate MLP and its feature range."""
    _fields_ = [("w", C.c_void_p), ("wq", C.c_void_p), ("m", C.c_void_p), ("v", C.c_void_p),
                ("lo", C.c_float * 12), ("hi", C.c_float * 12)]


class StageParams(C.Structure):
    """sodhip_stage_params: physical constants of the fused RK stage."""
    _fields_ = [("gamma_gas", D), ("Rgas", D), ("Cp", D), ("Pr", D), ("mu_factor", D),
                ("c_rho", D), ("c_ener", D), ("fact", D), ("p_ref", D),
                ("mu_factor_e", C.c_void_p)]


class Instances(C.Structure):
    """sodhip_instances: the per-instance fields of joined meshes (device pointers)."""
    _fields_ = [("ninst", I), ("wseg", P), ("elem_inst", P), ("p_inf", P), ("rho_inf", P),
                ("mu_factor", P)]


# name -> (restype, argtypes)
_SIGS = {
    "sodhip_version": (I, []),
    "sodhip_last_error": (I, [C.c_char_p, I]),
    "sodhip_basis_tables": (I, [I, P, P, P, P, P, P, P, P, P]),
    "sodhip_order_tables": (I, [I] + [P] * 6),
    "sodhip_box_mesh_sizes": (I, [I, I, I, I, I, P, P, P]),
    "sodhip_box_mesh": (I, [I, I, I, I, I, P, D, C.c_ulonglong, P, P, P, P]),
    "sodhip_elem_jacobian": (I, [I, I, I, P, P, P, P]),
    "sodhip_lumped_mass": (I, [I, I, I, P, P, I, P, P, P]),
    "sodhip_char_length": (I, [I, I, I, P, P, P, P]),
    "sodhip_element_clusters": (I, [I, I, I, P, I, P] + [P] * 9),
    "sodhip_create": (I, [I, P]),
    "sodhip_destroy": (I, [P]),
    "sodhip_set_stream": (I, [P, P]),
    "sodhip_synchronize": (I, [P]),
    "sodhip_forget_mesh": (I, [P, P]),
    "sodhip_reserve_scratch": (I, [P, SZ, SZ]),
    "sodhip_pin_scratch": (I, [P, I]),
    "sodhip_enter_data": (I, [P, P, SZ, I]),
    "sodhip_update_device": (I, [P, P, SZ]),
    "sodhip_update_host": (I, [P, P, SZ]),
    "sodhip_exit_data": (I, [P, P, I]),
    "sodhip_deviceptr": (P, [P, P]),
    "sodhip_prof_enable": (I, [P, I]),
    "sodhip_prof_get": (I, [P, I, C.POINTER(D), C.POINTER(C.c_longlong)]),
    "sodhip_prof_reset": (I, [P]),
    "sodhip_copy_present": (I, [P, P, P, SZ]),
    "sodhip_set_stage_config": (I, [P, I, I, I]),
    "sodhip_get_stage_config": (I, [P, C.POINTER(I), C.POINTER(I), C.POINTER(I)]),
    "sodhip_smart_visc_spectral_ens_f64": (I, [P, I, I, I, I, P, P, C.POINTER(Instances), P, D]
                                           + [P] * 7 + [C.POINTER(ViscParams), P, P]),
    "sodhip_eos_entropy_step_ens_f64": (I, [P, I, I, P, D, D, C.POINTER(Instances)] + [P] * 11),
    "sodhip_sensor_features_ens_f64": (I, [P, I, I, I, I, P, P, C.POINTER(Instances), P, D]
                                       + [P] * 7 + [C.POINTER(ViscParams), P, P]),
    "sodhip_full_convec_ijk_f64": (I, [P, I, I, I] + [P] * 16 + [I, D]),
    "sodhip_full_convec_ijk_f32": (I, [P, I, I, I] + [P] * 16 + [I, F]),
    "sodhip_full_diffusion_ijk_f64": (I, [P, I, I, I] + [P] * 8 + [D, D] + [P] * 7 + [D, D]
                                      + [P] * 3 + [I, D]),
    "sodhip_full_diffusion_ijk_f32": (I, [P, I, I, I] + [P] * 8 + [F, F] + [P] * 7 + [F, F]
                                      + [P] * 3 + [I, F]),
    "sodhip_comp_tau_f64": (I, [P, I, I, I, P, P, P, P, D, I, P]),
    "sodhip_full_proj_ijk_f64": (I, [P, I, I, I, I] + [P] * 18),
    "sodhip_full_diff_stab_ijk_f64": (I, [P, I, I, I] + [P] * 8 + [D, D] + [P] * 13
                                      + [D, D, D] + [P] * 3 + [I, D]),
    "sodhip_generic_scalar_convec_ijk_f64": (I, [P, I, I, I] + [P] * 11 + [P, I, D]),
    "sodhip_smart_visc_spectral_f64": (I, [P, I, I, I, I, P, P, P, D] + [P] * 7
                                       + [C.POINTER(ViscParams), P, P]),
    "sodhip_lumped_solver_scal_f64": (I, [P, I, I, P, P, P]),
    "sodhip_lumped_solver_vect_f64": (I, [P, I, I, P, P, P]),
    "sodhip_rk_stage_update_f64": (I, [P, I, I, P, P, D, D, D, D] + [P] * 9),
    "sodhip_eos_stage_f64": (I, [P, I, I, P, D, D] + [P] * 8),
    "sodhip_eos_step_f64": (I, [P, I, I, P, D, D] + [P] * 9),
    "sodhip_sutherland_f64": (I, [P, I, P, P, D, P]),
    "sodhip_thermo_entropy_f64": (I, [P, I, I, P, D, D, D] + [P] * 5),
    "sodhip_eos_entropy_step_f64": (I, [P, I, I, P, D, D, D, D, D] + [P] * 11),
    "sodhip_entropy_convec_residual_f64": (I, [P, I, I, I, I] + [P] * 12 + [D, D, P, P]),
    "sodhip_entropy_residual_f64": (I, [P, I, I, P, D, D, P, P, P]),
    "sodhip_rk_stage_f64": (I, [P, I, I, I, I, P, P, I] + [P] * 9
                            + [C.POINTER(StageParams), D, D, D, D] + [P] * 9),
    "sodhip_rk_stage_next_f64": (I, [P, I, I, I, I, P, P, I] + [P] * 9
                                 + [C.POINTER(StageParams), D, D, D, D] + [P] * 9
                                 + [D, P, P, P]),
    "sodhip_mlp_sizes": (I, [P, P, P]),
    "sodhip_mlp_quantize": (I, [P, C.POINTER(MLP)]),
    "sodhip_sensor_features_f64": (I, [P, I, I, I, I, P, P, P, D] + [P] * 7
                                   + [C.POINTER(ViscParams), P, P]),
    "sodhip_mlp_forward": (I, [P, C.POINTER(MLP), I, P, P]),
    "sodhip_mlp_gradient": (I, [P, C.POINTER(MLP), I, P, P, P, P, P, I]),
    "sodhip_mlp_train_step": (I, [P, C.POINTER(MLP), I, P, P, I, F, P, P, P, I]),
    "sodhip_surrogate_visc_spectral_f64": (I, [P, I, I, I, I, P, P, P, D] + [P] * 7
                                           + [C.POINTER(ViscParams), C.POINTER(MLP)] + [P] * 7),
    "sodhip_rk4_stage_update_f64": (I, [P, I, I, P, P, D, D, D, I, I] + [P] * 12),
    "sodhip_rk4_stage_f64": (I, [P, I, I, I, I, P, P, I] + [P] * 9
                             + [C.POINTER(StageParams), D, D, D, I, I] + [P] * 15),
    "sodhip_rk_stage_noalloc_f64": (I, [P, I, I, I, I, P, P, I] + [P] * 9
                                    + [C.POINTER(StageParams), D, D, D, D] + [P] * 9),
    "sodhip_rk_stage_lps_f64": (I, [P, I, I, I, I, P, P] + [P] * 10 + [C.POINTER(StageParams)]
                                + [D] * 5 + [P] * 20 + [D, P, P, P]),
    "sodhip_adapt_dt_cfl_f64": (I, [P, I, I, I, P, P, P, P, D, D, P, P, P, P, C.POINTER(D)]),
    "sodhip_volavg_ek_f64": (I, [P, I, I, I, P, P, D, P, P, C.POINTER(D)]),
    "sodhip_visc_dissipation_f64": (I, [P, I, I, I, P, D, P, P, P, D] + [P] * 7
                                    + [C.POINTER(D)] * 3),
    "sodhip_max_mach_f64": (I, [P, I, I, P, P, C.POINTER(D)]),
}

_lib = None


def lib():
    """Load libsod2d_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise SodhipError(f"{LIB_PATH} is missing: build it with `python -m sod2d_amd.build`")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def declared_symbols():
    """Names declared in include/sod2d_hip.h (parsed from the header)."""
    import re
    hdr = os.path.join(os.path.dirname(HERE), "include", "sod2d_hip.h")
    txt = open(hdr).read()
    return sorted(set(re.findall(r"\b(sodhip_[a-z0-9_]+)\s*\(", txt)))


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().sodhip_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str = ""):
    if rc != 0:
        raise SodhipError(f"{what or 'sodhip call'} failed (code {rc}): {last_error()}")
