"""Starburst and veil layers, the FFT seam and the reference's diffraction
debug entry points (render_starburst, diffraction.h:54; render_veil,
veil_render.h:54; _render_starburst_debug / _render_veil_debug,
bind_renderers.cpp:194-286)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _abi
from .config import FlareConfig, PointFlareConfig, fill_baffles
from .lens import LensCalibration, OpticalSystem
from .runtime import ctx_handle, lib


def calib_struct(calib: LensCalibration):
    c = _abi.gl_calibration()
    for f in LensCalibration.FIELDS:
        setattr(c, f, float(getattr(calib, f)))
    return c


def diffraction_params(cfg: FlareConfig, mdft_precision: int = 0):
    d = cfg.diffraction
    p = _abi.gl_diffraction_params()
    p.starburst_engine = int(d.starburst_engine)
    p.starburst_grid = int(d.starburst_grid)
    p.starburst_grid_cap = int(d.starburst_grid_cap)
    p.pupil_fill = float(d.pupil_fill)
    p.starburst_gain = float(d.starburst_gain)
    p.scale_trim = float(d.scale_trim)
    p.spectral_samples = int(d.spectral_samples)
    p.use_survivor_mask = int(bool(d.use_survivor_mask))
    p.survivor_grid = int(d.survivor_grid)
    p.veil_gain = float(d.veil_gain)
    p.veil_spread = float(d.veil_spread)
    p.veil_falloff = float(d.veil_falloff)
    p.render_spectral_samples = int(cfg.spectral_samples)
    p.sensor_model = int(cfg.sensor_model)
    p.output_cs = int(cfg.output_cs)
    for k, v in enumerate(np.asarray(cfg.custom_xyz_to_output, np.float32).reshape(-1)[:9]):
        p.custom_xyz_to_output[k] = float(v)
    p.aperture_blades = int(cfg.aperture_blades)
    p.aperture_rotation = float(cfg.aperture_rotation)
    p.mdft_precision = int(mdft_precision)
    p.n_baffles = fill_baffles(d, p.baffles)
    p.use_surface_textures = int(bool(d.use_surface_textures))
    return p


def source_pixels(lens: OpticalSystem, calib: LensCalibration, sources, width, height, half_w, half_h):
    """Buffer-pixel positions of field-angle sources through the first-order
    map y' = f tan(theta) with the calibrated focal lengths (used where no
    source-map solve is at hand)."""
    fx = calib.focal_length_x or lens.focal_length
    fy = calib.focal_length_y or lens.focal_length
    s = np.asarray(sources, np.float32).reshape(-1, 5)
    px = ((np.tan(s[:, 0].astype(np.float64)) * fx / (2.0 * half_w)) + 0.5) * width
    py = ((np.tan(s[:, 1].astype(np.float64)) * fy / (2.0 * half_h)) + 0.5) * height
    return px.astype(np.float32), py.astype(np.float32)


def _sources(sources):
    s = np.ascontiguousarray(sources, np.float32).reshape(-1, 5)
    arr = (_abi.gl_source * max(1, len(s)))()
    if len(s):
        C.memmove(arr, s.ctypes.data, s.nbytes)
    return arr, len(s)


def render_starburst(lens: OpticalSystem, calibration: LensCalibration, sources, src_px, src_py, width: int,
                     height: int, sensor_half_w: float, sensor_half_h: float, config: FlareConfig, fmt=None, ctx=None,
                     debug: bool = False, mdft_precision: int = 0):
    fw, fh = fmt if fmt is not None else (width, height)
    sarr, n = _sources(sources)
    px = np.ascontiguousarray(src_px, np.float32)
    py = np.ascontiguousarray(src_py, np.float32)
    planes = [np.zeros((height, width), np.float32) for _ in range(3)]
    p = diffraction_params(config, mdft_precision)
    cal = calib_struct(calibration)
    dbg = None
    keep = []
    if debug:
        cap = int(config.diffraction.starburst_grid_cap if config.diffraction.starburst_grid_cap >= 256 else 8192)
        cap = max(cap, int(config.diffraction.starburst_grid))
        cap = min(cap, 4096)
        psf = np.zeros(cap * cap, np.float32)
        spr = np.zeros(3 * cap * cap, np.float32)
        pup = np.zeros(cap * cap, np.float32)
        keep = [psf, spr, pup]
        dbg = _abi.gl_starburst_debug()
        dbg.grid_cap = cap
        dbg.psf, dbg.sprite_rgb, dbg.pupil = _abi.fptr(psf), _abi.fptr(spr), _abi.fptr(pup)
    dl = lens.device_lens()
    _abi.check(lib().gl_render_starburst(ctx_handle(ctx), dl.handle, C.byref(cal), sarr, _abi.fptr(px),
                                         _abi.fptr(py), n, width, height, fw, fh, float(sensor_half_w),
                                         float(sensor_half_h), C.byref(p), _abi.fptr(planes[0]), _abi.fptr(planes[1]),
                                         _abi.fptr(planes[2]), C.byref(dbg) if dbg is not None else None),
               "render_starburst")
    if not debug:
        return planes
    N = dbg.grid
    info = {k: getattr(dbg, k) for k, _t in _abi.gl_starburst_debug._fields_[5:]}
    info["grid"] = N
    if N and N <= dbg.grid_cap:
        info["psf"] = keep[0][: N * N].reshape(N, N).copy()
        info["sprite_rgb"] = keep[1][: 3 * N * N].reshape(N, N, 3).copy()
        info["pupil"] = keep[2][: N * N].reshape(N, N).copy()
    return planes, info


def render_veil(sources, src_px, src_py, width, height, sensor_half_w, sensor_half_h, config: FlareConfig, fmt=None,
                ctx=None, debug=False):
    fw, fh = fmt if fmt is not None else (width, height)
    sarr, n = _sources(sources)
    px = np.ascontiguousarray(src_px, np.float32)
    py = np.ascontiguousarray(src_py, np.float32)
    planes = [np.zeros((height, width), np.float32) for _ in range(3)]
    p = diffraction_params(config)
    dbg = None
    keep = []
    if debug:
        ref = np.zeros(1024 * 1024, np.float32)
        spr = np.zeros(3 * 1024 * 1024, np.float32)
        keep = [ref, spr]
        dbg = _abi.gl_veil_debug()
        dbg.reference, dbg.sprite_rgb = _abi.fptr(ref), _abi.fptr(spr)
    _abi.check(lib().gl_render_veil(ctx_handle(ctx), sarr, _abi.fptr(px), _abi.fptr(py), n, width, height, fw, fh,
                                    float(sensor_half_w), float(sensor_half_h), C.byref(p), _abi.fptr(planes[0]),
                                    _abi.fptr(planes[1]), _abi.fptr(planes[2]),
                                    C.byref(dbg) if dbg is not None else None), "render_veil")
    if not debug:
        return planes
    N = dbg.grid
    info = {k: getattr(dbg, k) for k, _t in _abi.gl_veil_debug._fields_[3:]}
    info.update(grid=N, reference=keep[0][: N * N].reshape(N, N).copy(),
                sprite_rgb=keep[1][: 3 * N * N].reshape(N, N, 3).copy())
    return planes, info


def fft2d(a: np.ndarray, ctx=None) -> np.ndarray:
    """Forward unnormalised 2-D FFT of a square complex array on the device
    (the fft_backend.h:25 seam)."""
    z = np.ascontiguousarray(a, np.complex64)
    N = z.shape[0]
    if z.shape != (N, N):
        raise ValueError("fft2d: square input required")
    out = np.empty_like(z)
    _abi.check(lib().gl_fft2d_c2c_forward_host(ctx_handle(ctx), z.view(np.float32).ctypes.data_as(_abi.F32P),
                                               out.view(np.float32).ctypes.data_as(_abi.F32P), N), "fft2d")
    return out


def fft2d_time(N: int, reps: int = 20, ctx=None) -> float:
    """Device ms of one in-place N x N forward transform (HIP events around
    `reps` back-to-back transforms on the context's stream)."""
    ms = C.c_float(0.0)
    _abi.check(lib().gl_fft2d_time(ctx_handle(ctx), int(N), int(reps), C.byref(ms)), "fft2d_time")
    return float(ms.value)


def render_starburst_layer(lens, calib, sources, width, height, hw, hh, config, px=None, py=None, ctx=None):
    if px is None or py is None:
        px, py = source_pixels(lens, calib, sources, width, height, hw, hh)
    r, g, b = render_starburst(lens, calib, sources, px, py, width, height, hw, hh, config, ctx=ctx)
    return {"starburst_r": r, "starburst_g": g, "starburst_b": b}


def render_veil_layer(lens, calib, sources, width, height, hw, hh, config, px=None, py=None, ctx=None):
    if px is None or py is None:
        px, py = source_pixels(lens, calib, sources, width, height, hw, hh)
    r, g, b = render_veil(sources, px, py, width, height, hw, hh, config, ctx=ctx)
    return {"veil_r": r, "veil_g": g, "veil_b": b}


def _debug_source(lens, calib, cfg: PointFlareConfig, w, h):
    from .render import source_rgb
    from .source_map import solve_source_map
    hw = cfg.sensor_half_w if cfg.sensor_half_w > 0 else calib.sensor_half_w
    hh = cfg.sensor_half_h if cfg.sensor_half_h > 0 else calib.sensor_half_h
    sm = solve_source_map(lens, calib, (cfg.source_x - 0.5) * 2.0, (cfg.source_y - 0.5) * 2.0, hw, hh)
    ax, ay = np.float32(sm.angle_x), np.float32(sm.angle_y)
    rgb = source_rgb(cfg)
    src = np.array([[ax, ay, rgb[0], rgb[1], rgb[2]]], np.float32)
    sx, sy = sm.screen(ax, ay)
    px = np.array([np.float32(sx) * np.float32(w)], np.float32)
    py = np.array([np.float32(sy) * np.float32(h)], np.float32)
    return src, px, py, hw, hh


def _render_starburst_debug(w: int, h: int, lens: OpticalSystem, calib: LensCalibration, cfg: PointFlareConfig,
                            ctx=None) -> dict:
    """bind_renderers.cpp:194-240: the starburst layer plus its intermediates."""
    src, px, py, hw, hh = _debug_source(lens, calib, cfg, w, h)
    planes, info = render_starburst(lens, calib, src, px, py, w, h, hw, hh, cfg, ctx=ctx, debug=True)
    out = {"starburst_r": planes[0], "starburst_g": planes[1], "starburst_b": planes[2],
           "source_px": float(px[0]), "source_py": float(py[0])}
    out.update(info)
    return out


def _render_veil_debug(w: int, h: int, lens: OpticalSystem, calib: LensCalibration, cfg: PointFlareConfig,
                       ctx=None) -> dict:
    """bind_renderers.cpp:244-286: the veil layer plus its intermediates."""
    src, px, py, hw, hh = _debug_source(lens, calib, cfg, w, h)
    planes, info = render_veil(src, px, py, w, h, hw, hh, cfg, ctx=ctx, debug=True)
    out = {"veil_r": planes[0], "veil_g": planes[1], "veil_b": planes[2],
           "source_px": float(px[0]), "source_py": float(py[0])}
    out.update(info)
    return out


MDFT_FP32 = 0
MDFT_BF16X3 = 1


def cgemm(a: np.ndarray, b: np.ndarray, precision: int = MDFT_FP32, ctx=None) -> np.ndarray:
    """Complex GEMM on the matrix cores (FP32 MFMA or the 3xBF16 split)."""
    A = np.ascontiguousarray(a, np.complex64)
    B = np.ascontiguousarray(b, np.complex64)
    M, K = A.shape
    K2, N = B.shape
    if K != K2:
        raise ValueError("cgemm: inner dimensions differ")
    out = np.empty((M, N), np.complex64)
    f = lambda x: x.view(np.float32).ctypes.data_as(_abi.F32P)  # noqa: E731
    _abi.check(lib().gl_cgemm_host(ctx_handle(ctx), M, N, K, f(A), f(B), f(out), int(precision)), "cgemm")
    return out


def mft_psf(pupil: np.ndarray, M: int, q_ref: float, lambdas, lambda_ref: float = 550.0,
            precision: int = MDFT_BF16X3, ctx=None, timing: bool = False):
    """Per-wavelength matrix-DFT PSFs |A_y P A_x^T|^2 (M x M each) of one
    complex pupil, lambda-scaled sampling q_l = q_ref lambda / lambda_ref."""
    P = np.ascontiguousarray(pupil, np.complex64)
    N = P.shape[0]
    lam = np.ascontiguousarray(lambdas, np.float32)
    out = np.empty((len(lam), M, M), np.float32)
    ms = C.c_float(0.0)
    _abi.check(lib().gl_mft_psf(ctx_handle(ctx), P.view(np.float32).ctypes.data_as(_abi.F32P), N, M, float(q_ref),
                                float(lambda_ref), _abi.fptr(lam), len(lam), int(precision), _abi.fptr(out), None,
                                C.byref(ms)), "mft_psf")
    return (out, ms.value) if timing else out


def primary_survival(lens: OpticalSystem, launch_xy, direction, z_spawn: float, lambda_nm: float = 550.0,
                     ctx=None) -> np.ndarray:
    """Which primary rays, launched from (x, y, z_spawn) along `direction`,
    reach the sensor (the survivor tracer behind the starburst's cat's-eye
    envelope).  Returns a bool array."""
    xy = np.ascontiguousarray(launch_xy, np.float32).reshape(-1, 2)
    out = np.zeros(len(xy), np.uint8)
    if len(xy) == 0:
        return out.astype(bool)
    d = np.asarray(direction, np.float32)
    dl = lens.device_lens()
    _abi.check(lib().gl_primary_survival(ctx_handle(ctx), dl.handle, _abi.fptr(xy), len(xy),
                                         float(d[0]), float(d[1]), float(d[2]), float(z_spawn), float(lambda_nm),
                                         out.ctypes.data_as(C.POINTER(C.c_ubyte))), "primary_survival")
    return out.astype(bool)


def veil_convolve(planes, sensor_half_w: float, sensor_half_h: float, config: FlareConfig, ctx=None,
                  timing: bool = False):
    """Veil applied to a whole image by FFT convolution with the unit-sum
    glare-spread kernel (the extended-source form of render_veil).  planes:
    three (H, W) float32 arrays; returns three arrays (and the device ms)."""
    ins = [np.ascontiguousarray(x, np.float32) for x in planes]
    H, W = ins[0].shape
    outs = [np.empty_like(ins[0]) for _ in range(3)]
    p = diffraction_params(config)
    ms = C.c_float(0.0)
    vp = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    _abi.check(lib().gl_veil_convolve(ctx_handle(ctx), vp(ins[0]), vp(ins[1]), vp(ins[2]), W, H,
                                      float(sensor_half_w), float(sensor_half_h), C.byref(p), 0, vp(outs[0]),
                                      vp(outs[1]), vp(outs[2]), C.byref(ms)), "veil_convolve")
    return (outs, ms.value) if timing else outs


def veil_convolve_device(d_in, d_out, width: int, height: int, sensor_half_w: float, sensor_half_h: float,
                         config: FlareConfig, ctx, noalloc: bool = False) -> None:
    """Device-plane form of veil_convolve: d_in / d_out are three device
    addresses (ints) of width*height float32 planes on ctx's device.  With
    noalloc the call goes to gl_veil_convolve_noalloc (no device allocation,
    images up to 1024 on the longer side, queued on ctx's stream without
    waiting); call ctx.synchronize() before reading d_out."""
    p = diffraction_params(config)
    h = ctx_handle(ctx)
    vp = C.c_void_p
    if noalloc:
        _abi.check(lib().gl_veil_convolve_noalloc(h, vp(d_in[0]), vp(d_in[1]), vp(d_in[2]), int(width), int(height),
                                                  float(sensor_half_w), float(sensor_half_h), C.byref(p),
                                                  vp(d_out[0]), vp(d_out[1]), vp(d_out[2])), "veil_convolve_noalloc")
    else:
        _abi.check(lib().gl_veil_convolve(h, vp(d_in[0]), vp(d_in[1]), vp(d_in[2]), int(width), int(height),
                                          float(sensor_half_w), float(sensor_half_h), C.byref(p), 1, vp(d_out[0]),
                                          vp(d_out[1]), vp(d_out[2]), None), "veil_convolve")
