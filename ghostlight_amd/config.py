"""Render configuration mirror.

Class names, field names, defaults and enum values follow the reference's
config hierarchy (ghostlight/src/render_config.h:14-383, spectral.h:16-38, bound
field by field in ghostlight/bindings/python/src/bind_config.cpp:70-377), so
code that sets fields by name keeps working.
"""
from __future__ import annotations

import copy
import enum
from dataclasses import dataclass, field
from typing import List, Tuple

import numpy as np

from . import _abi


class SensorModel(enum.IntEnum):
    SENSOR_CIE_1931 = 0


class InputColorSpace(enum.IntEnum):
    CS_INPUT_ACESCG = 0
    CS_INPUT_SRGB_LINEAR = 1
    CS_INPUT_CUSTOM = 2


class OutputColorSpace(enum.IntEnum):
    CS_ACESCG = 0
    CS_SRGB_LINEAR = 1
    CS_P3_D65 = 2
    CS_P3_D60 = 3
    CS_XYZ = 4
    CS_CUSTOM = 5


class GhostAovMode(enum.IntEnum):
    NONE = 0
    PER_PAIR = 1


class StarburstEngine(enum.IntEnum):
    SPRITE = 0
    MDFT = 1


class BaffleShape(enum.IntEnum):
    RECT = 0
    CIRCLE = 1


class HurbKickDistribution(enum.IntEnum):
    GAUSSIAN = 0
    LORENTZIAN = 1


class GateLobe(enum.IntEnum):
    CAUCHY_ISO = 0
    GROOVED = 1


class PSFCenterMode(enum.IntEnum):
    PSF_CENTER_CHIEF_CENTROID = 0
    PSF_CENTER_FIXED_TARGET = 1


MATTE_BOX_OPEN = 1.0e9


class GhostFilterMode(enum.IntEnum):
    ALL = 0
    INCLUDE = 1
    EXCLUDE = 2


@dataclass
class GhostFilter:
    """Pair selection after pre-filtering (render_config.h:52-63)."""
    Mode = GhostFilterMode
    mode: GhostFilterMode = GhostFilterMode.ALL
    pairs: List[Tuple[int, int]] = field(default_factory=list)


@dataclass
class MatteBox:
    enabled: bool = False
    z_front_mm: float = 60.0
    top: float = MATTE_BOX_OPEN
    bottom: float = MATTE_BOX_OPEN
    left: float = MATTE_BOX_OPEN
    right: float = MATTE_BOX_OPEN


@dataclass
class Baffle:
    shape: BaffleShape = BaffleShape.RECT
    z_front_mm: float = 60.0
    top: float = MATTE_BOX_OPEN
    bottom: float = MATTE_BOX_OPEN
    left: float = MATTE_BOX_OPEN
    right: float = MATTE_BOX_OPEN
    cx: float = 0.0
    cy: float = 0.0
    radius: float = MATTE_BOX_OPEN
    aspect: float = 1.0


@dataclass
class DiffractionConfig:
    """Starburst, veil and HURB settings (render_config.h:131-189)."""
    starburst: bool = False
    starburst_engine: StarburstEngine = StarburstEngine.SPRITE
    starburst_grid: int = 1024
    starburst_grid_cap: int = 8192
    pupil_fill: float = 0.30
    starburst_gain: float = 1.0
    scale_trim: float = 1.0
    spectral_samples: int = 0
    use_survivor_mask: bool = True
    use_surface_textures: bool = True
    survivor_grid: int = 192
    matte_box: MatteBox = field(default_factory=MatteBox)
    baffles: List[Baffle] = field(default_factory=list)
    veil: bool = False
    veil_gain: float = 0.03
    veil_spread: float = 0.12
    veil_falloff: float = 1.5
    hurb: bool = False
    hurb_kick: HurbKickDistribution = HurbKickDistribution.LORENTZIAN
    hurb_min_sigma_rad: float = 1.0e-5
    hurb_max_kick_rad: float = 0.35


@dataclass
class GateConfig:
    """Film-gate pass settings (render_config.h:212-269)."""
    enabled: bool = False
    thickness_mm: float = 0.8
    standoff_mm: float = 5.0
    offset_left_mm: float = 0.0
    offset_right_mm: float = 0.0
    offset_bottom_mm: float = 0.0
    offset_top_mm: float = 0.0
    lobe: GateLobe = GateLobe.GROOVED
    roughness_rad: float = 0.08
    groove_aniso: float = 0.12
    groove_angle_deg: float = 0.0
    max_kick_rad: float = 0.35
    reflectance_r0: float = 0.04
    gain: float = 1.0
    scatter_samples: int = 4
    spectral_samples: int = 0


def _zeros33():
    return [[0.0] * 3 for _ in range(3)]


@dataclass
class RenderConfig:
    """Base quality settings (render_config.h:14-46)."""
    ray_grid: int = 64
    spectral_samples: int = 16
    pupil_jitter: int = 0
    jitter_seed: int = 0
    sensor_model: SensorModel = SensorModel.SENSOR_CIE_1931
    input_cs: InputColorSpace = InputColorSpace.CS_INPUT_ACESCG
    output_cs: OutputColorSpace = OutputColorSpace.CS_ACESCG
    custom_input_to_xyz: list = field(default_factory=_zeros33)
    custom_xyz_to_output: list = field(default_factory=_zeros33)
    sensor_half_w: float = 0.0
    sensor_half_h: float = 0.0
    verbose: bool = False
    collect_stats: bool = False
    spectral_correction: bool = True

    def copy(self):
        return copy.deepcopy(self)


@dataclass
class FlareConfig(RenderConfig):
    """Ghost-path settings (render_config.h:273-337)."""
    aperture_blades: int = 0
    aperture_rotation: float = 0.0
    flare_gain: float = 1000.0
    min_ghost_intensity: float = 1e-7
    ghost_normalize: bool = True
    max_area_boost: float = 100.0
    cull_dead_pairs: bool = True
    concentrate_samples: bool = True
    adaptive_sample_budgets: bool = True
    adaptive_density_boost: float = 8.0
    adaptive_min_samples: int = 1024
    aov_mode: GhostAovMode = GhostAovMode.NONE
    aov_max_pairs: int = -1
    ghost_filter: GhostFilter = field(default_factory=GhostFilter)
    diffraction: DiffractionConfig = field(default_factory=DiffractionConfig)
    gate: GateConfig = field(default_factory=GateConfig)
    # Build-specific precision knob: spheres with |R| >= this (mm) use an FP64
    # ray-sphere solve; <= 0 disables it.
    fp64_radius_min: float = 1000.0
    # Build-specific launch form of the main trace: 0 grid, 1 persistent
    # (autotuned schedule), 2 persistent replayed from a hipGraph.
    batch_mode: int = 2
    schedule: int = -1          # -1 autotune, else a fixed candidate index
    # Build-specific: every source renders into a frame of its own (render_ghosts
    # then returns (n_sources, H, W) planes).
    frame_per_source: bool = False


@dataclass
class PointFlareConfig(FlareConfig):
    """Single screen-space point source (render_config.h:339-343)."""
    source_x: float = 0.5
    source_y: float = 0.5
    source_r: float = 1.0
    source_g: float = 1.0
    source_b: float = 1.0


@dataclass
class PSFConfig(RenderConfig):
    """PSF grid renderer settings (render_config.h:364-383)."""
    grid_nx: int = 8
    grid_ny: int = 8
    tile_w: int = 64
    tile_h: int = 64
    tile_extent_mm: float = 0.05
    monochromatic: bool = False
    center_mode: PSFCenterMode = PSFCenterMode.PSF_CENTER_CHIEF_CENTROID
    splat_sigma_um: float = 4.0
    aperture_blades: int = 0
    aperture_rotation: float = 0.0


def _flat9(m) -> list:
    a = np.asarray(m, np.float32).reshape(-1)
    return [float(x) for x in a[:9]] + [0.0] * (9 - min(9, a.size))


def render_params(cfg: RenderConfig):
    """The RenderConfig part of gl_flare_params (PSF renders), the ghost
    fields at their gl_flare_params_init defaults."""
    p = _abi.gl_flare_params()
    _abi.load_library().gl_flare_params_init(_abi.C.byref(p))
    p.ray_grid = int(cfg.ray_grid)
    p.spectral_samples = int(cfg.spectral_samples)
    p.pupil_jitter = int(cfg.pupil_jitter)
    p.jitter_seed = int(cfg.jitter_seed)
    p.sensor_model = int(cfg.sensor_model)
    p.input_cs = int(cfg.input_cs)
    p.output_cs = int(cfg.output_cs)
    for k, v in enumerate(_flat9(cfg.custom_input_to_xyz)):
        p.custom_input_to_xyz[k] = v
    for k, v in enumerate(_flat9(cfg.custom_xyz_to_output)):
        p.custom_xyz_to_output[k] = v
    p.verbose = int(bool(cfg.verbose))
    p.aperture_blades = int(getattr(cfg, "aperture_blades", 0))
    p.aperture_rotation = float(getattr(cfg, "aperture_rotation", 0.0))
    p.fp64_radius_min = float(getattr(cfg, "fp64_radius_min", 1000.0))
    return p


def to_params(cfg: FlareConfig):
    """Marshal a FlareConfig into the C ABI's gl_flare_params.  Returns the
    struct and the filter-pair buffer that must outlive the call."""
    p = _abi.gl_flare_params()
    p.ray_grid = int(cfg.ray_grid)
    p.spectral_samples = int(cfg.spectral_samples)
    p.pupil_jitter = int(cfg.pupil_jitter)
    p.jitter_seed = int(cfg.jitter_seed)
    p.sensor_model = int(cfg.sensor_model)
    p.input_cs = int(cfg.input_cs)
    p.output_cs = int(cfg.output_cs)
    for k, v in enumerate(_flat9(cfg.custom_input_to_xyz)):
        p.custom_input_to_xyz[k] = v
    for k, v in enumerate(_flat9(cfg.custom_xyz_to_output)):
        p.custom_xyz_to_output[k] = v
    p.sensor_half_w = float(cfg.sensor_half_w)
    p.sensor_half_h = float(cfg.sensor_half_h)
    p.verbose = int(bool(cfg.verbose))
    p.collect_stats = int(bool(cfg.collect_stats))
    p.spectral_correction = int(bool(cfg.spectral_correction))
    p.aperture_blades = int(cfg.aperture_blades)
    p.aperture_rotation = float(cfg.aperture_rotation)
    p.flare_gain = float(cfg.flare_gain)
    p.min_ghost_intensity = float(cfg.min_ghost_intensity)
    p.ghost_normalize = int(bool(cfg.ghost_normalize))
    p.max_area_boost = float(cfg.max_area_boost)
    p.cull_dead_pairs = int(bool(cfg.cull_dead_pairs))
    p.concentrate_samples = int(bool(cfg.concentrate_samples))
    p.adaptive_sample_budgets = int(bool(cfg.adaptive_sample_budgets))
    p.adaptive_density_boost = float(cfg.adaptive_density_boost)
    p.adaptive_min_samples = int(cfg.adaptive_min_samples)
    p.aov_mode = int(cfg.aov_mode)
    p.aov_max_pairs = int(cfg.aov_max_pairs)
    gf = cfg.ghost_filter
    buf = np.asarray([x for pr in gf.pairs for x in pr] or [0], np.int32)
    p.ghost_filter_mode = int(gf.mode)
    p.ghost_filter_count = len(gf.pairs)
    p.ghost_filter_pairs = _abi.iptr(buf)
    p.fp64_radius_min = float(cfg.fp64_radius_min)
    p.batch_mode = int(getattr(cfg, "batch_mode", 2))
    p.schedule = int(getattr(cfg, "schedule", -1))
    p.frame_per_source = int(bool(getattr(cfg, "frame_per_source", False)))
    d = cfg.diffraction
    p.hurb = int(bool(d.hurb))
    p.hurb_kick = int(d.hurb_kick)
    p.hurb_min_sigma_rad = float(d.hurb_min_sigma_rad)
    p.hurb_max_kick_rad = float(d.hurb_max_kick_rad)
    p.n_baffles = fill_baffles(d, p.baffles)
    return p, buf


SPAWN_OFFSET_MM = 20.0     # the spawn plane sits this far ahead of the front vertex (spawn_plane.h)


def fill_baffles(d: "DiffractionConfig", out) -> int:
    """build_gpu_baffles (baffle.h:124-153): the matte box as one RECT baffle,
    then the configured baffles, at most GL_MAX_BAFFLES; each plane's dz is
    measured from the spawn plane.  Fills the gl_baffle array `out`."""
    rows = []
    m = d.matte_box
    if m.enabled:
        rows.append((0, SPAWN_OFFSET_MM - m.z_front_mm, m.top, m.bottom, m.left, m.right, 0.0, 0.0,
                     MATTE_BOX_OPEN, 1.0))
    for b in d.baffles:
        rows.append((int(b.shape), SPAWN_OFFSET_MM - b.z_front_mm, b.top, b.bottom, b.left, b.right, b.cx, b.cy,
                     b.radius, b.aspect if b.aspect > 0.0 else 1.0))
    rows = rows[:_abi.GL_MAX_BAFFLES]
    for k, r in enumerate(rows):
        g = out[k]
        g.shape = r[0]
        g.dz, g.top, g.bottom, g.left, g.right, g.cx, g.cy, g.radius, g.aspect = (float(x) for x in r[1:])
    return len(rows)


def gate_params(gc: "GateConfig"):
    """GateConfig -> gl_gate_params (same fields and units)."""
    g = _abi.gl_gate_params()
    for name, _t in _abi.gl_gate_params._fields_:
        v = getattr(gc, name)
        setattr(g, name, int(v) if name in ("lobe", "scatter_samples", "spectral_samples") else float(v))
    return g
