"""Host checks of the edge-diffraction and film-gate pieces, under the
reference's binding names: `_hurb_sample_debug` (bind_renderers.cpp:457-481),
`_aperture_edge_distance_debug` (bind_trace.cpp:93), `_aperture_sdf_debug`
(bind_renderers.cpp:483-520), `_gate_params_debug` and `_gate_sample_debug`
(bind_renderers.cpp:416-455).  They call the same host functions the kernels'
math is built from (gl_hurb_sample, gl_aperture_edge_distance,
gl_bake_aperture_sdf, gl_gate_resolve, gl_gate_sample)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .lens import ApertureShape, OpticalSystem, Surface, Vec3f

HURB_GAUSS_K = np.float32(0.112539539)     # 1 / (2 sqrt 2 pi)  (hurb.h:40)
HURB_LORENTZ_K = np.float32(0.159154943)   # 1 / (2 pi)         (hurb.h:41)


def _hurb_sample_debug(lambda_nm: float, d_mm, dist: int, max_kick_rad: float, seed: int) -> dict:
    """Kick scale sigma = lambda K / d and one kick per distance, each from
    its own stream seeded wang(seed ^ (2654435761 i + 1))."""
    d = np.ascontiguousarray(d_mm, np.float32).reshape(-1)
    kicks = np.zeros(len(d), np.float32)
    sigma = np.zeros(len(d), np.float32)
    _abi.check(_abi.load_library().gl_hurb_sample(float(lambda_nm), _abi.fptr(d), len(d), int(dist),
                                                  float(max_kick_rad), int(seed) & 0xFFFFFFFF, _abi.fptr(kicks),
                                                  _abi.fptr(sigma)), "_hurb_sample_debug")
    return {"kicks": kicks, "sigma": sigma, "gauss_k": float(HURB_GAUSS_K), "lorentz_k": float(HURB_LORENTZ_K)}


def _aperture_edge_distance_debug(hit, surface: Surface):
    """(distance mm, edge normal as a Vec3f) from a world-space point on
    `surface` to its clear-aperture edge; an IMAGE aperture answers 1e30 and
    a zero normal (its edge distance lives in the stop's baked field)."""
    lens = OpticalSystem([surface])
    arr, _arena = lens.pack()
    h = (C.c_float * 3)(*[float(v) for v in hit])
    d = C.c_float()
    n = (C.c_float * 3)()
    _abi.check(_abi.load_library().gl_aperture_edge_distance(arr, h, C.byref(d), n), "_aperture_edge_distance_debug")
    return float(d.value), Vec3f(float(n[0]), float(n[1]), float(n[2]))


def _aperture_sdf_debug(lens: OpticalSystem) -> dict:
    """The image stop's baked distance field (find_sdf_target_surface +
    bake_aperture_sdf): stop_index (-1 when the stop is not a loaded IMAGE
    aperture), then sdf / nx / ny (H, W) and the texel pitches sx, sy."""
    out = {"stop_index": -1}
    k = next((i for i, s in enumerate(lens.surfaces) if s.is_stop), -1)
    if k < 0 or lens.surfaces[k].aperture_shape != ApertureShape.IMAGE:
        return out
    img = lens.aperture_images[k] if k < len(lens.aperture_images) else None
    if img is None or img.pixels is None or not len(img.pixels) or img.width <= 0 or img.height <= 0:
        return out
    s = lens.surfaces[k]
    semi = s.aperture_semi_diameter if s.aperture_semi_diameter > 0 else img.semi_diameter
    aspect = s.aperture_aspect if s.aperture_aspect > 0 else 1.0
    px = np.ascontiguousarray(img.pixels, np.float32).reshape(-1)
    f = np.zeros(4 * len(px), np.float32)
    if _abi.load_library().gl_bake_aperture_sdf(_abi.fptr(px), int(img.width), int(img.height), float(semi),
                                                 float(aspect), _abi.fptr(f)) != 0:
        return out
    f = f.reshape(img.height, img.width, 4)
    out.update(stop_index=k, width=img.width, height=img.height, sdf=f[..., 0].copy(), nx=f[..., 1].copy(),
               ny=f[..., 2].copy(), sx=2.0 * aspect * semi / img.width, sy=2.0 * semi / img.height)
    return out


def _gate_params_debug(config, sensor_half_w: float, sensor_half_h: float) -> dict:
    """The gate as the scatter pass resolves it from a GateConfig and the
    effective sensor half extents (every clamp applied)."""
    from .config import gate_params
    g = gate_params(config)
    r = _abi.gl_gate_resolved()
    _abi.check(_abi.load_library().gl_gate_resolve(C.byref(g), float(sensor_half_w), float(sensor_half_h),
                                                   C.byref(r)), "_gate_params_debug")
    return {name: getattr(r, name) for name, _t in _abi.gl_gate_resolved._fields_}


def _gate_sample_debug(config, n: int, seed: int) -> dict:
    """n draws of the wall's scatter lobe (axis, other, z) plus the lobe widths."""
    from .config import gate_params
    g = gate_params(config)
    n = max(int(n), 0)
    ax, ot, az = (np.zeros(n, np.float32) for _ in range(3))
    _abi.check(_abi.load_library().gl_gate_sample(C.byref(g), n, int(seed) & 0xFFFFFFFF, _abi.fptr(ax), _abi.fptr(ot),
                                                  _abi.fptr(az)), "_gate_sample_debug")
    p = _gate_params_debug(config, 1.0, 1.0)
    return {"axis": ax, "other": ot, "z": az, "sig_wide": p["sig_wide"], "sig_tight": p["sig_tight"]}
