"""Render entry points: the Python face of the drop-in boundary.

`render_point_flare`, `render_source_flare` mirror the reference binding
(ghostlight/bindings/python/src/bind_renderers.cpp:85-134): same argument
names and order, the same output dict keys / shapes / dtypes
(buffers.h:45-106) and the same exception types.  `render_ghosts` is the
lower seam of `launch_ghost_render` (ghost_render.h:75-94) with explicit
field-angle sources, used by the benchmark configurations.
"""
from __future__ import annotations

from dataclasses import dataclass

import ctypes as C
import math
from typing import Optional, Sequence

import numpy as np

from . import _abi, hostprep
from .config import FlareConfig, PointFlareConfig, to_params
from .lens import LensCalibration, OpticalSystem, TraceStatus, Vec3f
from .runtime import ctx_handle, lib


def _dims(fn: str, w: int, h: int):
    if w <= 0 or h <= 0:
        raise ValueError(f"{fn}: width and height must be positive")


def _sources_array(sources) -> "C.Array":
    s = np.ascontiguousarray(sources, np.float32).reshape(-1, 5)
    arr = (_abi.gl_source * max(1, len(s)))()
    if len(s):
        C.memmove(arr, s.ctypes.data, s.nbytes)
    return arr, len(s)


_PAIR_DTYPE = np.dtype([("a", np.int32), ("b", np.int32), ("boost", np.float32)])


def _pairs_array(pairs, boosts) -> "C.Array":
    """gl_pair array of (surf_a, surf_b) pairs and their area boosts."""
    n = len(pairs)
    arr = (_abi.gl_pair * max(1, n))()
    if n:
        ab = np.asarray(pairs, np.int64).reshape(n, 2)
        rec = np.empty(n, _PAIR_DTYPE)
        rec["a"], rec["b"] = ab[:, 0], ab[:, 1]
        bo = np.asarray(boosts, np.float64).ravel()
        if bo.size != n:
            raise ValueError(f"{bo.size} area boosts for {n} ghost pairs")
        rec["boost"] = bo
        C.memmove(arr, rec.ctypes.data, rec.nbytes)
    return arr


class _Stats:
    """Owner of the caller-side arrays of gl_ghost_stats."""

    def __init__(self, n_pairs_cap: int, n_ps_cap: int):
        self.pa = np.zeros(n_pairs_cap, np.int32)
        self.pb = np.zeros(n_pairs_cap, np.int32)
        self.pt = np.zeros(n_pairs_cap, np.uint64)
        self.ps = np.zeros(n_pairs_cap, np.uint64)
        self.po = np.zeros(n_pairs_cap, np.uint64)
        self.hits = np.zeros(max(1, n_ps_cap), np.int32)
        self.rect = np.zeros(max(1, n_ps_cap) * 4, np.float32)
        self.budget = np.zeros(max(1, n_ps_cap), np.int32)
        s = _abi.gl_ghost_stats()
        s.pair_cap = n_pairs_cap
        s.pair_surf_a = _abi.iptr(self.pa)
        s.pair_surf_b = _abi.iptr(self.pb)
        u64 = C.POINTER(C.c_uint64)
        s.pair_traces = self.pt.ctypes.data_as(u64)
        s.pair_survived = self.ps.ctypes.data_as(u64)
        s.pair_on_sensor = self.po.ctypes.data_as(u64)
        s.ps_cap = n_ps_cap
        s.ps_hits = _abi.iptr(self.hits)
        s.ps_rect = _abi.fptr(self.rect)
        s.ps_budget = _abi.iptr(self.budget)
        s.ms_grid_build = s.ms_upload = s.ms_kernel = s.ms_download = s.ms_total = -1.0
        self.s = s

    def to_dict(self) -> dict:
        s = self.s
        npair = s.n_pairs
        nps = s.n_ps
        return {
            "ms_grid_build": s.ms_grid_build, "ms_upload": s.ms_upload, "ms_kernel": s.ms_kernel,
            "ms_download": s.ms_download, "ms_total": s.ms_total,
            "n_pairs": s.n_pairs, "n_sources": s.n_sources, "n_grid": s.n_grid, "n_spec": s.n_spec,
            "traces_total": int(s.traces_total), "traces_survived": int(s.traces_survived),
            "traces_on_sensor": int(s.traces_on_sensor), "segments": int(s.segments),
            "pair_surf_a": self.pa[:npair].tolist(), "pair_surf_b": self.pb[:npair].tolist(),
            "pair_traces": self.pt[:npair].tolist(), "pair_survived": self.ps[:npair].tolist(),
            "pair_on_sensor": self.po[:npair].tolist(),
            "ps_hits": self.hits[:nps].tolist(), "ps_rect": self.rect[: 4 * nps].tolist(),
            "ps_budget": self.budget[:nps].tolist() if nps else [],
        }


def enumerate_ghost_pairs(lens: OpticalSystem):
    """All active (a < b) surface pairs (ghost.cpp:17)."""
    dl = lens.device_lens()
    n = lens.num_surfaces()
    cap = max(1, n * (n - 1) // 2)
    out = (_abi.gl_pair * cap)()
    k = C.c_int32()
    _abi.check(lib().gl_enumerate_ghost_pairs(dl.handle, out, cap, C.byref(k)), "enumerate_ghost_pairs")
    return [(out[i].surf_a, out[i].surf_b) for i in range(k.value)]


def filter_ghost_pairs(lens: OpticalSystem, sensor_half_w: float, sensor_half_h: float,
                       config: Optional[FlareConfig] = None, ctx=None):
    """(pairs, area_boosts) after the IOR-contrast and intensity gates
    (ghost.cpp:105)."""
    config = config or FlareConfig()
    dl = lens.device_lens()
    n = lens.num_surfaces()
    cap = max(1, n * (n - 1) // 2)
    out = (_abi.gl_pair * cap)()
    k = C.c_int32()
    p, _keep = to_params(config)
    _abi.check(lib().gl_filter_ghost_pairs(ctx_handle(ctx), dl.handle, float(sensor_half_w), float(sensor_half_h),
                                           C.byref(p), out, cap, C.byref(k)), "filter_ghost_pairs")
    return [(out[i].surf_a, out[i].surf_b) for i in range(k.value)], [out[i].area_boost for i in range(k.value)]


def render_ghosts(lens: OpticalSystem, sources, width: int, height: int, config: FlareConfig,
                  sensor_half_w: float, sensor_half_h: float, pairs=None, boosts=None,
                  fmt=None, ctx=None, device_out: bool = False, stats: bool = False):
    """launch_ghost_render (ghost_render.h:75-94) over field-angle sources.

    sources: (N, 5) [angle_x, angle_y, r, g, b] (radians, output colour
    space).  pairs/boosts: pre-filtered pairs; None runs filter_ghost_pairs.
    fmt: (fmt_w, fmt_h, x0, y0) format window, default the full buffer.
    Returns (r, g, b) host planes (H, W) float32, or three device pointers when
    device_out, plus the stats dict when requested."""
    _dims("render_ghosts", width, height)
    if pairs is None:
        pairs, boosts = filter_ghost_pairs(lens, sensor_half_w, sensor_half_h, config, ctx)
    if boosts is None:
        boosts = [1.0] * len(pairs)
    parr = _pairs_array(pairs, boosts)
    sarr, n_src = _sources_array(sources)
    fw, fh, x0, y0 = fmt if fmt is not None else (width, height, 0, 0)
    p, _keep = to_params(config)
    st = _Stats(max(1, len(pairs)), max(1, len(pairs) * n_src)) if stats else None
    dl = lens.device_lens()
    if device_out:
        ptrs = (C.POINTER(C.c_float) * 3)()
        _abi.check(lib().gl_launch_ghost_render(ctx_handle(ctx), dl.handle, parr, len(pairs), sarr, n_src,
                                                float(sensor_half_w), float(sensor_half_h), width, height, fw, fh, x0,
                                                y0, C.byref(p), None, None, None, ptrs,
                                                C.byref(st.s) if st else None), "render_ghosts")
        out = tuple(C.cast(ptrs[i], C.c_void_p).value for i in range(3))
    else:
        shape = (n_src, height, width) if getattr(config, "frame_per_source", False) else (height, width)
        planes = [np.zeros(shape, np.float32) for _ in range(3)]
        _abi.check(lib().gl_launch_ghost_render(ctx_handle(ctx), dl.handle, parr, len(pairs), sarr, n_src,
                                                float(sensor_half_w), float(sensor_half_h), width, height, fw, fh, x0,
                                                y0, C.byref(p), _abi.fptr(planes[0]), _abi.fptr(planes[1]),
                                                _abi.fptr(planes[2]), None, C.byref(st.s) if st else None),
                   "render_ghosts")
        out = tuple(planes)
    return (out, st.to_dict()) if stats else out


def trace_ghost_rays(lens: OpticalSystem, bounce_a: int, bounce_b: int, rays: np.ndarray, t_ge0: bool = True,
                     fp64_radius_min: float = 0.0, ctx=None, level: int = 0):
    """Device traces of (N, 7) rays [ox, oy, oz, dx, dy, dz, lambda]; bounce_b < 0
    traces the primary path.  Returns (pos (N,3), weight (N,), status (N,),
    segments (N,)).  level picks the device tracer's instantiation: 0 the
    generic surface step, 1 or 2 a predicated lean step (not below the level
    the lens needs), -1 the one a render of this lens runs."""
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 7)
    n = len(r)
    pos = np.zeros((n, 3), np.float32)
    w = np.zeros(n, np.float32)
    st = np.zeros(n, np.int32)
    sg = np.zeros(n, np.int32)
    dl = lens.device_lens()
    _abi.check(lib().gl_trace_ghost_rays_level(ctx_handle(ctx), dl.handle, int(bounce_a), int(bounce_b),
                                               _abi.fptr(r), n, int(bool(t_ge0)), float(fp64_radius_min), int(level),
                                               _abi.fptr(pos), _abi.fptr(w), _abi.iptr(st), _abi.iptr(sg)),
               "trace_ghost_rays")
    return pos, w, st, sg


@dataclass
class Ray:
    """A ray (vec3.h Ray): origin, unit direction, wavelength in nm."""
    origin: tuple = (0.0, 0.0, 0.0)
    direction: tuple = (0.0, 0.0, 1.0)
    wavelength: float = 550.0


@dataclass
class TraceResult:
    """Landing of one traced ray (TraceResult, trace_event.h:39-43)."""
    position: Vec3f
    weight: float
    status: TraceStatus


@dataclass
class TraceEvent:
    """One surface a diagnostic trace reached (TraceEvent, trace_event.h:48-58)."""
    surface_index: int
    hit_point: Vec3f
    surface_normal: Vec3f
    ior_before: float
    ior_after: float
    fresnel_weight: float
    status: TraceStatus
    reflected: bool


@dataclass
class RayPath:
    """A diagnostic trace: every surface event, then the landing (RayPath,
    trace_event.h:61-64)."""
    events: list
    result: TraceResult


@dataclass
class GhostPair:
    """Two-bounce ghost (ghost.h GhostPair): reflect at surf_b, then at surf_a."""
    surf_a: int
    surf_b: int


def _one_ray(ray: Ray) -> np.ndarray:
    r = np.zeros((1, 7), np.float32)
    r[0, :3] = ray.origin
    r[0, 3:6] = ray.direction
    r[0, 6] = ray.wavelength
    return r


def _result(pos, w, st) -> TraceResult:
    return TraceResult(Vec3f(float(pos[0]), float(pos[1]), float(pos[2])), float(w), TraceStatus(int(st)))


def trace_ghost_ray(ray: Ray, lens: OpticalSystem, bounce_a: int, bounce_b: int, ctx=None) -> TraceResult:
    """One ghost trace (bind_trace.cpp:57-63; trace.cpp:171): the device
    tracer in FP32 with the CPU tracer's t >= 0 landing rule."""
    _check_bounces(lens, bounce_a, bounce_b)
    pos, w, st, _ = trace_ghost_rays(lens, bounce_a, bounce_b, _one_ray(ray), t_ge0=True, ctx=ctx)
    return _result(pos[0], w[0], st[0])


def trace_primary_ray(ray: Ray, lens: OpticalSystem, ctx=None) -> TraceResult:
    """One primary trace (bind_trace.cpp:65-67; trace.cpp:243)."""
    pos, w, st, _ = trace_ghost_rays(lens, -1, -1, _one_ray(ray), t_ge0=True, ctx=ctx)
    return _result(pos[0], w[0], st[0])


def _check_bounces(lens: OpticalSystem, bounce_a: int, bounce_b: int) -> None:
    if bounce_a < 0 or bounce_b < 0 or bounce_a >= bounce_b or bounce_b >= lens.num_surfaces():
        raise ValueError("bounce indices must satisfy 0 <= bounce_a < bounce_b < num_surfaces")


TRACE_EVENT_DTYPE = np.dtype([("surface_index", np.int32), ("hit_point", np.float32, 3),
                              ("surface_normal", np.float32, 3), ("ior_before", np.float32),
                              ("ior_after", np.float32), ("fresnel_weight", np.float32), ("status", np.int32),
                              ("reflected", np.int32)])


def trace_rays_diagnostic(lens: OpticalSystem, bounce_a: int, bounce_b: int, rays: np.ndarray, t_ge0: bool = True,
                          fp64_radius_min: float = 0.0, ctx=None):
    """Device diagnostic traces of (N, 7) rays: (events (N, 3 n_surfaces)
    structured array of TRACE_EVENT_DTYPE, n_events (N,), pos (N, 3),
    weight (N,), status (N,)); bounce_b < 0 traces the primary path."""
    r = np.ascontiguousarray(rays, np.float32).reshape(-1, 7)
    n = len(r)
    m = max(1, 3 * lens.num_surfaces())
    ev = np.zeros((max(1, n), m), TRACE_EVENT_DTYPE)
    cnt = np.zeros(max(1, n), np.int32)
    pos = np.zeros((max(1, n), 3), np.float32)
    w = np.zeros(max(1, n), np.float32)
    st = np.zeros(max(1, n), np.int32)
    dl = lens.device_lens()
    _abi.check(lib().gl_trace_rays_diagnostic(ctx_handle(ctx), dl.handle, int(bounce_a), int(bounce_b), _abi.fptr(r),
                                              n, int(bool(t_ge0)), float(fp64_radius_min), m,
                                              ev.ctypes.data_as(C.POINTER(_abi.gl_trace_event)), _abi.iptr(cnt),
                                              _abi.fptr(pos), _abi.fptr(w), _abi.iptr(st)), "trace_rays_diagnostic")
    return ev[:n], cnt[:n], pos[:n], w[:n], st[:n]


def _ray_path(ev, cnt, pos, w, st) -> RayPath:
    events = [TraceEvent(int(e["surface_index"]), Vec3f(*(float(x) for x in e["hit_point"])),
                         Vec3f(*(float(x) for x in e["surface_normal"])), float(e["ior_before"]),
                         float(e["ior_after"]), float(e["fresnel_weight"]), TraceStatus(int(e["status"])),
                         bool(e["reflected"])) for e in ev[:int(cnt)]]
    return RayPath(events, _result(pos, w, st))


def trace_ghost_ray_diagnostic(ray: Ray, lens: OpticalSystem, bounce_a: int, bounce_b: int, ctx=None) -> RayPath:
    """One ghost trace with its per-surface events (bind_trace.cpp:71-79;
    trace.cpp:324-395), on the device."""
    _check_bounces(lens, bounce_a, bounce_b)
    ev, cnt, pos, w, st = trace_rays_diagnostic(lens, bounce_a, bounce_b, _one_ray(ray), ctx=ctx)
    return _ray_path(ev[0], cnt[0], pos[0], w[0], st[0])


def trace_primary_ray_diagnostic(ray: Ray, lens: OpticalSystem, ctx=None) -> RayPath:
    """One primary trace with its per-surface events (bind_trace.cpp:81-86;
    trace.cpp:264-320), on the device."""
    ev, cnt, pos, w, st = trace_rays_diagnostic(lens, -1, -1, _one_ray(ray), ctx=ctx)
    return _ray_path(ev[0], cnt[0], pos[0], w[0], st[0])


def solve_source_angle(lens: OpticalSystem, calib: LensCalibration, ndc_x: float, ndc_y: float,
                       half_w: float, half_h: float, ctx=None):
    """Screen NDC -> field angle by the traced source-map solve
    (source_map.cpp:182): the angle whose spawn-disc survivor mean lands at
    (ndc_x * half_w, ndc_y * half_h) mm."""
    from .source_map import solve_source_map
    sm = solve_source_map(lens, calib, ndc_x, ndc_y, half_w, half_h, ctx=ctx)
    return np.float32(sm.angle_x), np.float32(sm.angle_y)


def source_rgb(config: PointFlareConfig):
    """The source colour in the output space: Mout . Min . (r, g, b) in float32
    with the reference's summation order (SourceFlareRenderer.cpp:76-86)."""
    Min, Mout = hostprep.color_matrices(config.input_cs, config.output_cs, config.custom_input_to_xyz,
                                        config.custom_xyz_to_output)
    f = np.float32
    rgb = [f(config.source_r), f(config.source_g), f(config.source_b)]
    xyz = [f(f(f(Min[i, 0] * rgb[0]) + f(Min[i, 1] * rgb[1])) + f(Min[i, 2] * rgb[2])) for i in range(3)]
    return [f(f(f(Mout[i, 0] * xyz[0]) + f(Mout[i, 1] * xyz[1])) + f(Mout[i, 2] * xyz[2])) for i in range(3)]


def render_source_flare(offsets, width: int, height: int, lens: OpticalSystem, calibration: LensCalibration,
                        config: PointFlareConfig, ctx=None) -> dict:
    """Extended-source flare (SourceFlareRenderer.cpp:24-208): offsets is an
    (N, 3) float32 array of [d_angle_x, d_angle_y, weight] rows."""
    _dims("render_source_flare", width, height)
    off = np.asarray(offsets, np.float32)
    if off.ndim != 2 or off.shape[1] != 3 or off.shape[0] < 1:
        raise ValueError("render_source_flare: offsets must be a (N, 3) float32 array of "
                         "[d_angle_x, d_angle_y, weight] rows with N >= 1")
    if not np.all(np.isfinite(off)):
        raise RuntimeError("render_source_flare: HIP error: source offsets must be finite")
    if lens.num_surfaces() <= 0:
        raise RuntimeError("render_source_flare: HIP error: lens has no surfaces")
    hw = config.sensor_half_w if config.sensor_half_w > 0 else calibration.sensor_half_w
    hh = config.sensor_half_h if config.sensor_half_h > 0 else calibration.sensor_half_h
    from .source_map import solve_source_map
    smap = solve_source_map(lens, calibration, (config.source_x - 0.5) * 2.0, (config.source_y - 0.5) * 2.0, hw, hh,
                            ctx=ctx)
    ax0, ay0 = np.float32(smap.angle_x), np.float32(smap.angle_y)
    src_rgb = source_rgb(config)
    sources = np.zeros((len(off), 5), np.float32)
    sources[:, 0] = ax0 + off[:, 0]
    sources[:, 1] = ay0 + off[:, 1]
    for c in range(3):
        sources[:, 2 + c] = src_rgb[c] * off[:, 2]
    # Each source's buffer position from the solve's forward map
    # (SourceFlareRenderer.cpp:95-120), for the starburst and veil splats.
    src_px = np.zeros(len(off), np.float32)
    src_py = np.zeros(len(off), np.float32)
    for i in range(len(off)):
        sx, sy = smap.screen(sources[i, 0], sources[i, 1])
        src_px[i] = np.float32(sx) * np.float32(width)
        src_py[i] = np.float32(sy) * np.float32(height)
    return render_sources(sources, width, height, lens, calibration, config, ctx=ctx, src_px=src_px, src_py=src_py)


def render_sources(sources, width: int, height: int, lens: OpticalSystem, calibration: LensCalibration,
                   config: PointFlareConfig, ctx=None, src_px=None, src_py=None) -> dict:
    """The flare of (N, 5) [angle_x, angle_y, r, g, b] sources whose field
    angles are already solved and whose colour is already in the output space
    (the body of SourceFlareRenderer.cpp:24-208 after the source map and the
    colour conversion).  src_px/src_py: each source's buffer-pixel position
    for the starburst and veil splats (default: the first-order map through
    the calibrated focal lengths).  Same output dict as render_source_flare."""
    _dims("render_sources", width, height)
    sources = np.ascontiguousarray(sources, np.float32)
    if sources.ndim != 2 or sources.shape[1] != 5 or sources.shape[0] < 1:
        raise ValueError("render_sources: sources must be a (N, 5) float32 array with N >= 1")
    if not np.all(np.isfinite(sources)):
        raise RuntimeError("render_sources: HIP error: source rows must be finite")
    hw = config.sensor_half_w if config.sensor_half_w > 0 else calibration.sensor_half_w
    hh = config.sensor_half_h if config.sensor_half_h > 0 else calibration.sensor_half_h
    out = {"width": int(width), "height": int(height)}
    p, _keep = to_params(config)
    sarr, n_src = _sources_array(sources)
    dl = lens.device_lens()
    planes = [np.zeros((height, width), np.float32) for _ in range(3)]
    st = None
    n = lens.num_surfaces()
    max_pairs = max(1, n * (n - 1) // 2)
    if config.collect_stats:
        st = _Stats(max_pairs, max_pairs * n_src)
    aov_buf = aov_pairs = None
    aov_n = C.c_int32(0)
    aov_cap = 0
    if int(config.aov_mode) == 1:
        # One layer per pair that survives the pre-filter and the GhostFilter,
        # capped by aov_max_pairs: the pipeline's own selection, so the host
        # buffer holds exactly the layers it writes.
        fp, _ = filter_ghost_pairs(lens, hw, hh, config, ctx=ctx)
        gf = config.ghost_filter
        if int(gf.mode) != 0 and gf.pairs:
            want = {(int(a), int(b)) for a, b in gf.pairs}
            fp = [q for q in fp if (q in want) == (int(gf.mode) == 1)]
        aov_cap = len(fp) if config.aov_max_pairs < 0 else min(len(fp), int(config.aov_max_pairs))
        aov_buf = np.zeros((max(1, aov_cap), 3, height, width), np.float32)
        aov_pairs = np.zeros(2 * max(1, aov_cap), np.int32)
    _abi.check(lib().gl_render_ghost_pipeline(
        ctx_handle(ctx), dl.handle, sarr, n_src, float(calibration.sensor_half_w), float(calibration.sensor_half_h),
        width, height, width, height, 0, 0, C.byref(p), _abi.fptr(planes[0]), _abi.fptr(planes[1]),
        _abi.fptr(planes[2]), _abi.fptr(aov_buf) if aov_buf is not None else None,
        _abi.iptr(aov_pairs) if aov_pairs is not None else None, aov_cap,
        C.byref(aov_n), C.byref(st.s) if st else None), "render_source_flare")
    out["ghost_r"], out["ghost_g"], out["ghost_b"] = planes
    if config.diffraction.starburst:
        from . import diffraction
        out.update(diffraction.render_starburst_layer(lens, calibration, sources, width, height, hw, hh, config,
                                                      px=src_px, py=src_py,
                                                      ctx=ctx))
    if config.diffraction.veil:
        from . import diffraction
        out.update(diffraction.render_veil_layer(lens, calibration, sources, width, height, hw, hh, config,
                                                 px=src_px, py=src_py, ctx=ctx))
    if config.gate.enabled:
        out.update(render_gate_layer(sources, width, height, lens, config, hw, hh, ctx=ctx))
    if aov_buf is not None:
        for i in range(aov_n.value):
            a, b = int(aov_pairs[2 * i]), int(aov_pairs[2 * i + 1])
            for c, ch in enumerate("rgb"):
                out[f"ghost_s{a}_s{b}_{ch}"] = np.ascontiguousarray(aov_buf[i, c])
    if st is not None:
        out["stats"] = st.to_dict()
    return out


def render_gate_layer(sources, width: int, height: int, lens: OpticalSystem, config: PointFlareConfig,
                      sensor_half_w: float, sensor_half_h: float, ctx=None, counters: bool = False) -> dict:
    """Film-gate scatter layer (render_gate, gate_render.cu:280-514; called by
    SourceFlareRenderer.cpp:183-200 with the render's effective half
    extents): gate_r/g/b planes, plus [traces, scrapes] when counters."""
    from .config import gate_params
    src = np.ascontiguousarray(sources, np.float32).reshape(-1, 5)
    sarr, n_src = _sources_array(src)
    p, _keep = to_params(config)
    gp = gate_params(config.gate)
    planes = [np.zeros((height, width), np.float32) for _ in range(3)]
    ctr = (C.c_uint64 * 2)()
    dl = lens.device_lens()
    _abi.check(lib().gl_render_gate(ctx_handle(ctx), dl.handle, sarr, n_src, width, height, width,
                                    height, 0, 0, float(sensor_half_w), float(sensor_half_h), C.byref(p),
                                    C.byref(gp), _abi.fptr(planes[0]), _abi.fptr(planes[1]), _abi.fptr(planes[2]),
                                    ctr), "render_gate")
    out = {"gate_r": planes[0], "gate_g": planes[1], "gate_b": planes[2]}
    if counters:
        out["gate_counters"] = (int(ctr[0]), int(ctr[1]))
    return out


_GATE_KEYS = ("x_pos", "x_neg", "y_pos", "y_neg", "t_mm", "zs_mm", "sig_wide", "sig_tight")


def _render_gate_debug(w: int, h: int, lens: OpticalSystem, calibration: LensCalibration, config: PointFlareConfig,
                       ctx=None) -> dict:
    """The gate layer of one source with the pass's telemetry
    (bind_renderers.cpp:288-334): gate_r/g/b, the direct image of the same
    primary traces (direct_r/g/b), traces, scrapes, energy, the largest
    realised specular reach, the predicted capture bands t / (2 N) and the
    resolved gate.  The source is placed as the other debug entry points
    place it (source_x / source_y through the source map).  With the gate
    disabled the pass is a no-op: zero layer, no direct image, zero
    telemetry."""
    _dims("_render_gate_debug", w, h)
    from .config import gate_params
    from .diffraction import _debug_source
    src, _px, _py, hw, hh = _debug_source(lens, calibration, config, w, h)
    planes = [np.zeros((h, w), np.float32) for _ in range(3)]
    d = {"width": int(w), "height": int(h)}
    dbg = _abi.gl_gate_debug()
    if config.gate.enabled:
        direct = [np.zeros((h, w), np.float32) for _ in range(3)]
        dbg.direct_r, dbg.direct_g, dbg.direct_b = (_abi.fptr(x) for x in direct)
        sarr, n_src = _sources_array(src)
        p, _keep = to_params(config)
        gp = gate_params(config.gate)
        dl = lens.device_lens()
        _abi.check(lib().gl_render_gate_debug(ctx_handle(ctx), dl.handle, sarr, n_src, w, h, w, h, 0,
                                              0, float(hw), float(hh), C.byref(p), C.byref(gp),
                                              _abi.fptr(planes[0]), _abi.fptr(planes[1]), _abi.fptr(planes[2]),
                                              C.byref(dbg)), "render_gate")
    d["gate_r"], d["gate_g"], d["gate_b"] = planes
    if config.gate.enabled:
        d["direct_r"], d["direct_g"], d["direct_b"] = direct
    d["traces"] = int(dbg.traces)
    d["scrapes"] = int(dbg.scrapes)
    d["energy"] = float(dbg.energy)
    d["reach_mm"] = float(dbg.reach_mm)
    # gate_render.cu:507-510: the band t / (2 N) per axis, in float.
    f = np.float32
    t = f(dbg.gate.t_mm)
    nx = f(calibration.f_number_x) if calibration.f_number_x > 0 else f(0.0)
    ny = f(calibration.f_number_y) if calibration.f_number_y > 0 else f(0.0)
    d["band_x_mm"] = float(t / (f(2.0) * nx)) if nx > 0 else 0.0
    d["band_y_mm"] = float(t / (f(2.0) * ny)) if ny > 0 else 0.0
    for k in _GATE_KEYS:
        d[k] = float(getattr(dbg.gate, k))
    return d


def render_point_flare(width: int, height: int, lens: OpticalSystem, calibration: LensCalibration,
                       config: PointFlareConfig, ctx=None) -> dict:
    """Single point source (PointFlareRenderer.cpp:11-20): one zero offset."""
    _dims("render_point_flare", width, height)
    return render_source_flare(np.array([[0.0, 0.0, 1.0]], np.float32), width, height, lens, calibration, config,
                               ctx=ctx)
