"""Lens calibration and the screen <-> field-angle source map.

calibrate_lens follows lens_calibration.cpp:344-443 (Python binding
bind_calibration.cpp:40); solve_source_map, source_map_landing and
source_map_screen follow source_map.cpp:182-519 (bindings
bind_renderers.cpp:338-412).  The searches run in the library's host code and
every probe ray is traced on the device (gl_calibrate_lens,
gl_solve_source_angle, gl_source_map_landing).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Optional, Tuple

from . import _abi
from .lens import LensCalibration, OpticalSystem
from .runtime import ctx_handle, lib

D_LINE_NM = 587.56
STATUS_NAMES = {0: "solved", 1: "continued", 2: "fallback"}


def _calib_struct(calib: LensCalibration) -> _abi.gl_calibration:
    c = _abi.gl_calibration()
    for f in LensCalibration.FIELDS:
        setattr(c, f, float(getattr(calib, f)))
    return c


def calibrate_lens(lens: OpticalSystem, d_line_nm: float = D_LINE_NM, ctx=None) -> LensCalibration:
    """Covered field (90% of the axial spawn-disc throughput), image circle
    (5%), paraxial focal lengths, entrance pupil and f-numbers per axis."""
    out = _abi.gl_calibration()
    dl = lens.device_lens()
    _abi.check(lib().gl_calibrate_lens(ctx_handle(ctx), dl.handle, float(d_line_nm),
                                       float(getattr(lens, "focal_length", 0.0) or 0.0), C.byref(out)),
               "calibrate_lens")
    return LensCalibration(**{f: float(getattr(out, f)) for f in LensCalibration.FIELDS})


@dataclass
class SourceMapSolve:
    """A solved source position and its local forward map (source_map.h:20-52)."""
    angle_x: float = 0.0
    angle_y: float = 0.0
    status: str = "fallback"
    residual_mm: float = 0.0
    anchor_ax: float = 0.0
    anchor_ay: float = 0.0
    anchor_x_mm: float = 0.0
    anchor_y_mm: float = 0.0
    jacobian: Tuple[float, float, float, float] = (0.0, 0.0, 0.0, 0.0)
    eff_half_w: float = 0.0
    eff_half_h: float = 0.0
    probe_evals: int = 0
    _raw: Optional[_abi.gl_source_map] = field(default=None, repr=False, compare=False)

    def screen(self, angle_x: float, angle_y: float) -> Tuple[float, float]:
        """Where a source at (angle_x, angle_y) appears, as [0, 1] frame
        fractions (source_map_screen, source_map.cpp:456-484)."""
        sx, sy = C.c_float(), C.c_float()
        lib().gl_source_map_screen(C.byref(self._raw), float(angle_x), float(angle_y), C.byref(sx), C.byref(sy))
        return sx.value, sy.value


def solve_source_map(lens: OpticalSystem, calib: LensCalibration, ndc_x: float, ndc_y: float, eff_half_w: float,
                     eff_half_h: float, ctx=None) -> SourceMapSolve:
    """Screen NDC (-1..1, 0 = axis) of a frame with half-extents eff_half_w/h
    (mm) to the field angle whose spawn-disc survivor mean lands there."""
    out = _abi.gl_source_map()
    cs = _calib_struct(calib)
    dl = lens.device_lens()
    _abi.check(lib().gl_solve_source_angle(ctx_handle(ctx), dl.handle, C.byref(cs), float(ndc_x),
                                           float(ndc_y), float(eff_half_w), float(eff_half_h), C.byref(out)),
               "solve_source_angle")
    return SourceMapSolve(angle_x=out.angle_x, angle_y=out.angle_y, status=STATUS_NAMES.get(out.status, "fallback"),
                          residual_mm=out.residual_mm, anchor_ax=out.anchor_ax, anchor_ay=out.anchor_ay,
                          anchor_x_mm=out.anchor_x, anchor_y_mm=out.anchor_y, jacobian=tuple(out.J),
                          eff_half_w=out.eff_half_w, eff_half_h=out.eff_half_h, probe_evals=out.probe_evals,
                          _raw=out)


def _solve_source_map(lens: OpticalSystem, calibration: LensCalibration, source_x: float, source_y: float,
                      eff_half_w: float, eff_half_h: float, ctx=None) -> dict:
    """bind_renderers.cpp:338-375: the solve for a screen fraction source
    position ([0, 1], 0.5 = axis) as a dict with the reference's keys."""
    sm = solve_source_map(lens, calibration, (source_x - 0.5) * 2.0, (source_y - 0.5) * 2.0, eff_half_w, eff_half_h,
                          ctx=ctx)
    sx, sy = sm.screen(sm.angle_x, sm.angle_y)
    return {"angle_x": sm.angle_x, "angle_y": sm.angle_y, "status": sm.status, "residual_mm": sm.residual_mm,
            "anchor_ax": sm.anchor_ax, "anchor_ay": sm.anchor_ay, "anchor_x_mm": sm.anchor_x_mm,
            "anchor_y_mm": sm.anchor_y_mm, "jacobian": sm.jacobian, "probe_evals": sm.probe_evals,
            "screen_x": sx, "screen_y": sy}


def _source_map_landing(lens: OpticalSystem, calibration: LensCalibration, angle_x: float, angle_y: float,
                        ctx=None) -> Optional[Tuple[float, float]]:
    """bind_renderers.cpp:378-393: the survivor-mean landing (mm) at one
    angle, or None past the image circle."""
    x, y, ok = C.c_float(), C.c_float(), C.c_int32()
    cs = _calib_struct(calibration)
    dl = lens.device_lens()
    _abi.check(lib().gl_source_map_landing(ctx_handle(ctx), dl.handle, C.byref(cs), float(angle_x),
                                           float(angle_y), C.byref(x), C.byref(y), C.byref(ok)),
               "source_map_landing")
    return (x.value, y.value) if ok.value else None


def _source_map_screen(lens: OpticalSystem, calibration: LensCalibration, source_x: float, source_y: float,
                       eff_half_w: float, eff_half_h: float, angle_x: float, angle_y: float,
                       ctx=None) -> Tuple[float, float]:
    """bind_renderers.cpp:395-412: the forward map of the solve for a source
    position, evaluated at an arbitrary angle."""
    sm = solve_source_map(lens, calibration, (source_x - 0.5) * 2.0, (source_y - 0.5) * 2.0, eff_half_w, eff_half_h,
                          ctx=ctx)
    return sm.screen(angle_x, angle_y)
