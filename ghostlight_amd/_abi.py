"""ctypes mirror of include/ghostlight_amd.h and the loader of the HIP library.

The product library is ``ghostlight_amd/_lib/libghostlight_amd.so`` (built by
``__graft_entry__.build()``).  Loading it never falls back to anything: a
missing or unloadable library raises ``ImportError`` naming the path.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

GL_OK = 0
GL_ERR_INVALID = -1
GL_ERR_DEVICE = -2
GL_ERR_NO_MEMORY = -3
GL_ERR_UNSUPPORTED = -4

GL_MAX_SURFACES = 64
GL_MAX_ASPHERE_TERMS = 8
GL_MAX_BLADES = 32

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libghostlight_amd.so")


class gl_surface(C.Structure):
    _fields_ = [
        ("radius", C.c_float), ("thickness", C.c_float), ("ior", C.c_float),
        ("abbe_v", C.c_float), ("semi_aperture", C.c_float), ("z", C.c_float),
        ("is_stop", C.c_int32), ("is_active", C.c_int32),
        ("aperture_shape", C.c_int32), ("aperture_blades", C.c_int32),
        ("aperture_rotation_rad", C.c_float), ("aperture_aspect", C.c_float),
        ("decenter_x", C.c_float), ("decenter_y", C.c_float),
        ("rot", C.c_float * 9),
        ("form", C.c_int32), ("conic_k", C.c_float),
        ("asphere_terms", C.c_float * GL_MAX_ASPHERE_TERMS),
        ("n_asphere_terms", C.c_int32), ("cyl_axis", C.c_int32),
        ("disp_model", C.c_int32),
        ("sellmeier_B", C.c_float * 3), ("sellmeier_C", C.c_float * 3),
        ("coating_model", C.c_int32), ("ar_layers", C.c_int32),
        ("out_of_range_discard", C.c_int32), ("angle_ref_ior", C.c_float),
        ("tint_r", C.c_float), ("tint_g", C.c_float), ("tint_b", C.c_float),
        ("tint_strength", C.c_float),
        ("gauss_sigma", C.c_float), ("gauss_background", C.c_float),
        ("gauss_peak", C.c_float), ("gauss_decenter_x", C.c_float),
        ("gauss_decenter_y", C.c_float),
        ("table_offset", C.c_int32), ("table_count", C.c_int32),
        ("sa_offset", C.c_int32), ("sa_n_wavelengths", C.c_int32),
        ("sa_n_angles", C.c_int32),
        ("aperture_curvature", C.c_float), ("aperture_twist", C.c_float),
        ("aperture_notch_rad", C.c_float), ("aperture_notch_angle_rad", C.c_float),
        ("aperture_semi_diameter", C.c_float), ("image_offset", C.c_int32), ("image_width", C.c_int32),
        ("image_height", C.c_int32), ("image_sdf", C.c_int32), ("image_sdf_offset", C.c_int32),
    ]


class gl_blade_profile(C.Structure):
    _fields_ = [("blades", C.c_int32), ("plain", C.c_int32), ("facets", C.c_int32)] + [
        (n, C.c_float) for n in ("rotation", "sigma", "half", "r_w", "phi_w", "u_w", "u_rise", "qR", "qL", "p",
                                 "beta", "eA", "eB", "area_frac")]


class gl_sample(C.Structure):
    _fields_ = [("u", C.c_float), ("v", C.c_float)]


class gl_spec(C.Structure):
    _fields_ = [("lambda_", C.c_float), ("cmf_r", C.c_float), ("cmf_g", C.c_float), ("cmf_b", C.c_float)]


class gl_pair(C.Structure):
    _fields_ = [("surf_a", C.c_int32), ("surf_b", C.c_int32), ("area_boost", C.c_float)]


class gl_source(C.Structure):
    _fields_ = [("angle_x", C.c_float), ("angle_y", C.c_float),
                ("r", C.c_float), ("g", C.c_float), ("b", C.c_float)]


GL_MAX_BAFFLES = 8


class gl_baffle(C.Structure):
    _fields_ = [("shape", C.c_int32)] + [(n, C.c_float) for n in ("dz", "top", "bottom", "left", "right", "cx", "cy",
                                                                   "radius", "aspect")]


class gl_flare_params(C.Structure):
    _fields_ = [
        ("ray_grid", C.c_int32), ("spectral_samples", C.c_int32),
        ("pupil_jitter", C.c_int32), ("jitter_seed", C.c_int32),
        ("sensor_model", C.c_int32), ("input_cs", C.c_int32), ("output_cs", C.c_int32),
        ("custom_input_to_xyz", C.c_float * 9), ("custom_xyz_to_output", C.c_float * 9),
        ("sensor_half_w", C.c_float), ("sensor_half_h", C.c_float),
        ("verbose", C.c_int32), ("collect_stats", C.c_int32),
        ("spectral_correction", C.c_int32),
        ("aperture_blades", C.c_int32), ("aperture_rotation", C.c_float),
        ("flare_gain", C.c_float), ("min_ghost_intensity", C.c_float),
        ("ghost_normalize", C.c_int32), ("max_area_boost", C.c_float),
        ("cull_dead_pairs", C.c_int32), ("concentrate_samples", C.c_int32),
        ("adaptive_sample_budgets", C.c_int32), ("adaptive_density_boost", C.c_float),
        ("adaptive_min_samples", C.c_int32),
        ("aov_mode", C.c_int32), ("aov_max_pairs", C.c_int32),
        ("ghost_filter_mode", C.c_int32), ("ghost_filter_count", C.c_int32),
        ("ghost_filter_pairs", C.POINTER(C.c_int32)),
        ("fp64_radius_min", C.c_float),
        ("batch_mode", C.c_int32),
        ("schedule", C.c_int32),
        ("frame_per_source", C.c_int32),
        ("hurb", C.c_int32), ("hurb_kick", C.c_int32), ("hurb_min_sigma_rad", C.c_float),
        ("hurb_max_kick_rad", C.c_float), ("n_baffles", C.c_int32), ("baffles", gl_baffle * GL_MAX_BAFFLES),
    ]


class gl_psf_source(C.Structure):
    _fields_ = [("angle_x", C.c_float), ("angle_y", C.c_float), ("r", C.c_float), ("g", C.c_float),
                ("b", C.c_float), ("chief_x_mm", C.c_float), ("chief_y_mm", C.c_float), ("tile_x0", C.c_int32),
                ("tile_y0", C.c_int32)]


class gl_psf_params(C.Structure):
    _fields_ = [("grid_nx", C.c_int32), ("grid_ny", C.c_int32), ("tile_w", C.c_int32), ("tile_h", C.c_int32),
                ("tile_extent_mm", C.c_float), ("monochromatic", C.c_int32), ("splat_sigma_um", C.c_float)]


class gl_ghost_stats(C.Structure):
    _fields_ = [
        ("ms_grid_build", C.c_float), ("ms_upload", C.c_float), ("ms_kernel", C.c_float),
        ("ms_download", C.c_float), ("ms_total", C.c_float),
        ("n_pairs", C.c_int32), ("n_sources", C.c_int32), ("n_grid", C.c_int32), ("n_spec", C.c_int32),
        ("traces_total", C.c_uint64), ("traces_survived", C.c_uint64), ("traces_on_sensor", C.c_uint64),
        ("segments", C.c_uint64),
        ("pair_cap", C.c_int32),
        ("pair_surf_a", C.POINTER(C.c_int32)), ("pair_surf_b", C.POINTER(C.c_int32)),
        ("pair_traces", C.POINTER(C.c_uint64)), ("pair_survived", C.POINTER(C.c_uint64)),
        ("pair_on_sensor", C.POINTER(C.c_uint64)),
        ("ps_cap", C.c_int32), ("n_ps", C.c_int32),
        ("ps_hits", C.POINTER(C.c_int32)), ("ps_rect", C.POINTER(C.c_float)),
        ("ps_budget", C.POINTER(C.c_int32)),
    ]


class gl_calibration(C.Structure):
    _fields_ = [(n, C.c_float) for n in (
        "sensor_half_w", "sensor_half_h", "max_half_angle_h", "max_half_angle_v", "image_circle_semi_w",
        "image_circle_semi_h", "focal_length_x", "focal_length_y", "entrance_pupil_semi_x", "entrance_pupil_semi_y",
        "f_number_x", "f_number_y", "pupil_area_frac")]


class gl_gate_params(C.Structure):
    _fields_ = [("thickness_mm", C.c_float), ("standoff_mm", C.c_float), ("offset_left_mm", C.c_float),
                ("offset_right_mm", C.c_float), ("offset_bottom_mm", C.c_float), ("offset_top_mm", C.c_float),
                ("lobe", C.c_int32), ("roughness_rad", C.c_float), ("groove_aniso", C.c_float),
                ("groove_angle_deg", C.c_float), ("max_kick_rad", C.c_float), ("reflectance_r0", C.c_float),
                ("gain", C.c_float), ("scatter_samples", C.c_int32), ("spectral_samples", C.c_int32)]


class gl_gate_resolved(C.Structure):
    _fields_ = [(n, C.c_float) for n in ("x_pos", "x_neg", "y_pos", "y_neg", "t_mm", "zs_mm", "sig_wide", "sig_tight",
                                         "cos_g", "sin_g", "r0", "max_kick", "inv_scatter", "gain")] + [
        ("n_scatter", C.c_int32), ("lobe", C.c_int32)]


class gl_gate_debug(C.Structure):
    _fields_ = [("gate", gl_gate_resolved), ("traces", C.c_uint64), ("scrapes", C.c_uint64), ("energy", C.c_double),
                ("reach_mm", C.c_float), ("direct_r", C.POINTER(C.c_float)), ("direct_g", C.POINTER(C.c_float)),
                ("direct_b", C.POINTER(C.c_float))]


class gl_trace_event(C.Structure):
    _fields_ = [("surface_index", C.c_int32), ("hit_point", C.c_float * 3), ("surface_normal", C.c_float * 3),
                ("ior_before", C.c_float), ("ior_after", C.c_float), ("fresnel_weight", C.c_float),
                ("status", C.c_int32), ("reflected", C.c_int32)]


class gl_source_map(C.Structure):
    _fields_ = [("angle_x", C.c_float), ("angle_y", C.c_float), ("status", C.c_int32), ("residual_mm", C.c_float),
                ("anchor_ax", C.c_float), ("anchor_ay", C.c_float), ("anchor_x", C.c_float), ("anchor_y", C.c_float),
                ("J", C.c_float * 4), ("eff_half_w", C.c_float), ("eff_half_h", C.c_float),
                ("inv_tan_w", C.c_float), ("inv_tan_h", C.c_float), ("probe_evals", C.c_int32)]


class gl_diffraction_params(C.Structure):
    _fields_ = [
        ("starburst_engine", C.c_int32), ("starburst_grid", C.c_int32), ("starburst_grid_cap", C.c_int32),
        ("pupil_fill", C.c_float), ("starburst_gain", C.c_float), ("scale_trim", C.c_float),
        ("spectral_samples", C.c_int32), ("use_survivor_mask", C.c_int32), ("survivor_grid", C.c_int32),
        ("veil_gain", C.c_float), ("veil_spread", C.c_float), ("veil_falloff", C.c_float),
        ("render_spectral_samples", C.c_int32), ("sensor_model", C.c_int32), ("output_cs", C.c_int32),
        ("custom_xyz_to_output", C.c_float * 9), ("aperture_blades", C.c_int32), ("aperture_rotation", C.c_float),
        ("mdft_precision", C.c_int32), ("n_baffles", C.c_int32), ("baffles", gl_baffle * GL_MAX_BAFFLES),
        ("use_surface_textures", C.c_int32),
    ]


class gl_starburst_debug(C.Structure):
    _fields_ = [
        ("grid", C.c_int32), ("grid_cap", C.c_int32),
        ("psf", C.POINTER(C.c_float)), ("sprite_rgb", C.POINTER(C.c_float)), ("pupil", C.POINTER(C.c_float)),
    ] + [(n, C.c_float) for n in (
        "dx_mm_x", "dx_mm_y", "f_number_x", "f_number_y", "focal_length_x", "focal_length_y",
        "entrance_pupil_semi_x", "entrance_pupil_semi_y", "lambda_ref_nm", "field_angle_x", "field_angle_y",
        "throughput", "chief_offset_x", "chief_offset_y")]


class gl_veil_debug(C.Structure):
    _fields_ = [("grid", C.c_int32), ("reference", C.POINTER(C.c_float)), ("sprite_rgb", C.POINTER(C.c_float))] + [
        (n, C.c_float) for n in ("dx_mm_x", "dx_mm_y", "core_mm", "falloff", "core_texels")]


F32P = C.POINTER(C.c_float)
I32P = C.POINTER(C.c_int32)
VP = C.c_void_p

# name -> (restype, argtypes); every symbol include/ghostlight_amd.h declares.
SIGNATURES = {
    "gl_device_available": (C.c_int, []),
    "gl_last_error": (C.c_char_p, []),
    "gl_ctx_create": (C.c_int, [C.POINTER(VP)]),
    "gl_ctx_destroy": (None, [VP]),
    "gl_ctx_synchronize": (C.c_int, [VP]),
    "gl_ctx_last_kernel_ms": (C.c_float, [VP]),
    "gl_ctx_last_splat_ms": (C.c_float, [VP]),
    "gl_ctx_last_splat_counts": (C.c_int, [VP, C.POINTER(C.c_uint64)]),
    "gl_device_memory": (C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "gl_ctx_memory_bytes": (C.c_uint64, [VP]),
    "gl_ctx_trim": (C.c_int, [VP]),
    "gl_ctx_last_schedule": (C.c_int, [VP, C.POINTER(C.c_int32)]),
    "gl_lens_create": (C.c_int, [C.POINTER(gl_surface), C.c_int32, F32P, C.c_int32, C.POINTER(VP)]),
    "gl_lens_destroy": (None, [VP]),
    "gl_flare_params_init": (None, [C.POINTER(gl_flare_params)]),
    "gl_build_spectral_table": (C.c_int, [C.c_int32, C.c_int32, F32P, C.POINTER(gl_spec)]),
    "gl_spectral_correction": (C.c_int, [C.POINTER(gl_spec), C.c_int32, C.c_int32, F32P, F32P]),
    "gl_color_matrices": (C.c_int, [C.c_int32, F32P, C.c_int32, F32P, F32P, F32P]),
    "gl_build_aperture_samples": (C.c_int, [C.POINTER(gl_surface), C.c_int32, C.c_int32, C.c_int32,
                                            C.c_int32, C.c_int32, C.c_float, C.POINTER(gl_sample),
                                            C.c_int32, I32P]),
    "gl_build_ior_tables": (C.c_int, [C.POINTER(gl_surface), C.c_int32, C.POINTER(gl_spec), C.c_int32,
                                      F32P, F32P]),
    "gl_bake_tmm_table": (C.c_int, [C.c_int32, F32P, F32P, F32P, F32P, F32P, F32P]),
    "gl_blade_profile_build": (C.c_int, [C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                                         C.POINTER(gl_blade_profile)]),
    "gl_blade_radius_at": (C.c_float, [C.POINTER(gl_blade_profile), C.c_float]),
    "gl_blade_dr_dtheta": (C.c_float, [C.POINTER(gl_blade_profile), C.c_float]),
    "gl_gate_resolve": (C.c_int, [C.POINTER(gl_gate_params), C.c_float, C.c_float, C.POINTER(gl_gate_resolved)]),
    "gl_gate_sample": (C.c_int, [C.POINTER(gl_gate_params), C.c_int32, C.c_uint32, C.POINTER(C.c_float),
                                 C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    "gl_hurb_sample": (C.c_int, [C.c_float, F32P, C.c_int32, C.c_int32, C.c_float, C.c_uint32, F32P, F32P]),
    "gl_aperture_edge_distance": (C.c_int, [C.POINTER(gl_surface), F32P, F32P, F32P]),
    "gl_bake_aperture_sdf": (C.c_int, [F32P, C.c_int32, C.c_int32, C.c_float, C.c_float, F32P]),
    "gl_enumerate_ghost_pairs": (C.c_int, [VP, C.POINTER(gl_pair), C.c_int32, I32P]),
    "gl_filter_ghost_pairs": (C.c_int, [VP, VP, C.c_float, C.c_float, C.POINTER(gl_flare_params),
                                        C.POINTER(gl_pair), C.c_int32, I32P]),
    "gl_launch_ghost_render": (C.c_int, [VP, VP, C.POINTER(gl_pair), C.c_int32, C.POINTER(gl_source),
                                         C.c_int32, C.c_float, C.c_float, C.c_int32, C.c_int32,
                                         C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                         C.POINTER(gl_flare_params), F32P, F32P, F32P,
                                         C.POINTER(F32P), C.POINTER(gl_ghost_stats)]),
    "gl_render_ghost_pipeline": (C.c_int, [VP, VP, C.POINTER(gl_source), C.c_int32, C.c_float, C.c_float,
                                           C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                           C.POINTER(gl_flare_params), F32P, F32P, F32P, F32P, I32P,
                                           C.c_int32, I32P, C.POINTER(gl_ghost_stats)]),
    "gl_trace_ghost_rays": (C.c_int, [VP, VP, C.c_int32, C.c_int32, F32P, C.c_int32, C.c_int32, C.c_float,
                                      F32P, F32P, I32P, I32P]),
    "gl_trace_ghost_rays_level": (C.c_int, [VP, VP, C.c_int32, C.c_int32, F32P, C.c_int32, C.c_int32, C.c_float,
                                            C.c_int32, F32P, F32P, I32P, I32P]),
    "gl_trace_rays_diagnostic": (C.c_int, [VP, VP, C.c_int32, C.c_int32, F32P, C.c_int32, C.c_int32, C.c_float,
                                           C.c_int32, C.POINTER(gl_trace_event), I32P, F32P, F32P, I32P]),
    "gl_selftest_sqrt_rn": (C.c_int, [VP, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint32)]),
    "gl_selftest_rn": (C.c_int, [VP, C.c_int32, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64),
                                 C.POINTER(C.c_uint32)]),
    "gl_fft2d_c2c_forward_inplace": (C.c_int, [VP, VP, C.c_int32]),
    "gl_fft2d_c2c_forward_host": (C.c_int, [VP, F32P, F32P, C.c_int32]),
    "gl_fft2d_time": (C.c_int, [VP, C.c_int32, C.c_int32, C.POINTER(C.c_float)]),
    "gl_render_starburst": (C.c_int, [VP, VP, C.POINTER(gl_calibration), C.POINTER(gl_source), F32P, F32P, C.c_int32,
                                      C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                      C.POINTER(gl_diffraction_params), F32P, F32P, F32P,
                                      C.POINTER(gl_starburst_debug)]),
    "gl_render_veil": (C.c_int, [VP, C.POINTER(gl_source), F32P, F32P, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                 C.c_int32, C.c_float, C.c_float, C.POINTER(gl_diffraction_params), F32P, F32P, F32P,
                                 C.POINTER(gl_veil_debug)]),
    "gl_calibrate_lens": (C.c_int, [VP, VP, C.c_float, C.c_float, C.POINTER(gl_calibration)]),
    "gl_solve_source_angle": (C.c_int, [VP, VP, C.POINTER(gl_calibration), C.c_float, C.c_float, C.c_float,
                                        C.c_float, C.POINTER(gl_source_map)]),
    "gl_psf_deep_pupil": (C.c_int, [VP, VP, C.c_int32, F32P, F32P, C.c_int32, F32P]),
    "gl_source_map_screen": (None, [C.POINTER(gl_source_map), C.c_float, C.c_float, F32P, F32P]),
    "gl_source_map_landing": (C.c_int, [VP, VP, C.POINTER(gl_calibration), C.c_float, C.c_float, F32P, F32P,
                                        I32P]),
    "gl_render_gate": (C.c_int, [VP, VP, C.POINTER(gl_source), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                 C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, C.POINTER(gl_flare_params),
                                 C.POINTER(gl_gate_params), F32P, F32P, F32P, C.POINTER(C.c_uint64)]),
    "gl_render_gate_debug": (C.c_int, [VP, VP, C.POINTER(gl_source), C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                       C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                       C.POINTER(gl_flare_params), C.POINTER(gl_gate_params), F32P, F32P, F32P,
                                       C.POINTER(gl_gate_debug)]),
    "gl_render_psf_tiles": (C.c_int, [VP, VP, C.POINTER(gl_psf_source), C.c_int32, C.POINTER(gl_flare_params),
                                      C.POINTER(gl_psf_params), F32P, F32P, F32P, F32P]),
    "gl_veil_convolve": (C.c_int, [VP, VP, VP, VP, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                   C.POINTER(gl_diffraction_params), C.c_int32, VP, VP, VP, F32P]),
    "gl_veil_convolve_noalloc": (C.c_int, [VP, VP, VP, VP, C.c_int32, C.c_int32, C.c_float, C.c_float,
                                           C.POINTER(gl_diffraction_params), VP, VP, VP]),
    "gl_cgemm_host": (C.c_int, [VP, C.c_int32, C.c_int32, C.c_int32, F32P, F32P, F32P, C.c_int32]),
    "gl_mft_psf": (C.c_int, [VP, F32P, C.c_int32, C.c_int32, C.c_float, C.c_float, F32P, C.c_int32, C.c_int32, F32P,
                             C.POINTER(F32P), F32P]),
    "gl_primary_survival": (C.c_int, [VP, VP, F32P, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float,
                                      C.c_float, C.POINTER(C.c_uint8)]),
}

_lib = None
_lock = threading.Lock()


def load_library(path: str = LIB_PATH):
    """Load the HIP product library once; raise ImportError when absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise ImportError(
                f"ghostlight_amd: HIP library not built ({path}); run __graft_entry__.build()")
        lib = C.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return lib


def fptr(a: np.ndarray):
    return a.ctypes.data_as(F32P)


def iptr(a: np.ndarray):
    return a.ctypes.data_as(I32P)


def check(status: int, what: str):
    """Map a C status to the reference binding's exception types."""
    if status == GL_OK:
        return
    msg = _lib.gl_last_error().decode(errors="replace") if _lib is not None else ""
    if status == GL_ERR_INVALID:
        raise ValueError(f"{what}: {msg}")
    if status == GL_ERR_UNSUPPORTED:
        raise NotImplementedError(f"{what}: {msg}")
    if status == GL_ERR_NO_MEMORY:
        raise MemoryError(f"{what}: {msg}")
    raise RuntimeError(f"{what}: HIP error: {msg}")
