"""Host preparation wrappers: spectral tables, colour matrices, pupil
samples, IOR tables and the TMM coating bake (all CPU code in the library)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .runtime import lib


def _custom9(m):
    a = (C.c_float * 9)()
    if m is not None:
        for k, v in enumerate(np.asarray(m, np.float32).reshape(-1)[:9]):
            a[k] = float(v)
    return a


def spectral_table(n: int, output_cs: int = 0, custom=None) -> np.ndarray:
    """(n, 4) float32 rows [lambda, cmf_r, cmf_g, cmf_b] (build_spectral_table,
    spectral.cpp:317; n == 3 gives the fixed RGB table)."""
    out = (_abi.gl_spec * n)()
    _abi.check(lib().gl_build_spectral_table(n, int(output_cs), _custom9(custom), out), "spectral_table")
    return np.frombuffer(bytes(out), np.float32).reshape(n, 4).copy()


def spectral_correction(table: np.ndarray, output_cs: int = 0, custom=None) -> np.ndarray:
    t = np.ascontiguousarray(table, np.float32)
    arr = (_abi.gl_spec * len(t)).from_buffer_copy(t.tobytes())
    C9 = (C.c_float * 9)()
    _abi.check(lib().gl_spectral_correction(arr, len(t), int(output_cs), _custom9(custom), C9), "spectral_correction")
    return np.array(C9[:], np.float32).reshape(3, 3)


def color_matrices(input_cs=0, output_cs=0, custom_in=None, custom_out=None):
    Mi, Mo = (C.c_float * 9)(), (C.c_float * 9)()
    _abi.check(lib().gl_color_matrices(int(input_cs), _custom9(custom_in), int(output_cs), _custom9(custom_out), Mi, Mo),
               "color_matrices")
    return np.array(Mi[:], np.float32).reshape(3, 3), np.array(Mo[:], np.float32).reshape(3, 3)


def aperture_samples(lens, ray_grid: int, pupil_jitter: int = 0, jitter_seed: int = 0, blades_override: int = 0,
                     rotation_deg: float = 0.0) -> np.ndarray:
    arr, _ = lens.pack()
    cap = max(1, ray_grid * ray_grid)
    out = (_abi.gl_sample * cap)()
    n = C.c_int32()
    _abi.check(lib().gl_build_aperture_samples(arr, lens.num_surfaces(), ray_grid, pupil_jitter, jitter_seed,
                                               blades_override, float(rotation_deg), out, cap, C.byref(n)),
               "aperture_samples")
    return np.frombuffer(bytes(out), np.float32).reshape(cap, 2)[: n.value].copy()


def ior_tables(lens, spec: np.ndarray):
    arr, _ = lens.pack()
    t = np.ascontiguousarray(spec, np.float32)
    sp = (_abi.gl_spec * len(t)).from_buffer_copy(t.tobytes())
    n = lens.num_surfaces()
    at = np.zeros((len(t), n), np.float32)
    bf = np.zeros((len(t), n), np.float32)
    _abi.check(lib().gl_build_ior_tables(arr, n, sp, len(t), _abi.fptr(at), _abi.fptr(bf)), "ior_tables")
    return at, bf


def bake_tmm_table(n_re, n_im, thickness_nm, n_ambient31, n_substrate31) -> np.ndarray:
    """SPECTRAL_ANGULAR table wl[31] | ang[18] | r[31*18] of a thin-film stack."""
    n_re = np.ascontiguousarray(n_re, np.float32)
    n_im = np.ascontiguousarray(n_im, np.float32)
    th = np.ascontiguousarray(thickness_nm, np.float32)
    amb = np.ascontiguousarray(n_ambient31, np.float32)
    sub = np.ascontiguousarray(n_substrate31, np.float32)
    out = np.zeros(31 + 18 + 31 * 18, np.float32)
    _abi.check(lib().gl_bake_tmm_table(len(th), _abi.fptr(n_re), _abi.fptr(n_im), _abi.fptr(th), _abi.fptr(amb),
                                       _abi.fptr(sub), _abi.fptr(out)), "bake_tmm_table")
    return out


def build_spectral_lambdas(n_samples: int, lambda_min: float = 400.0, lambda_max: float = 700.0) -> list:
    """Cell-centre wavelengths (lens_calibration.h:59-67), float32 arithmetic."""
    f = np.float32
    return [float(f(lambda_min) + f(f(lambda_max) - f(lambda_min)) * f(f(i) + f(0.5)) / f(n_samples))
            for i in range(n_samples)]
