"""ghostlight_amd — MI355X-native flare hot path of Ghostlight.

Ghost tracer, starburst and PSF/veil passes as hand-written gfx950 HIP kernels
behind a C ABI (include/ghostlight_amd.h), with the reference's Python API
(render_point_flare, render_source_flare, config classes, output dicts) on top.
"""
from .config import (BaffleShape, Baffle, DiffractionConfig, FlareConfig, GateConfig, GateLobe, GhostAovMode,
                     GhostFilter, HurbKickDistribution, InputColorSpace, MatteBox, OutputColorSpace,
                     PointFlareConfig, PSFCenterMode, PSFConfig, RenderConfig, SensorModel, StarburstEngine)
from .lens import (ApertureImage, ApertureProfile, ApertureShape, Coating, CoatingModel, CylinderAxis, DispersionModel,
                   LensCalibration, OpticalSystem, Surface, SurfaceForm, TraceStatus, Vec3f, cross, dot)
from .runtime import Context, device_available
from .render import (GhostPair, Ray, RayPath, TraceEvent, TraceResult, enumerate_ghost_pairs, filter_ghost_pairs,
                     render_ghosts, render_point_flare, render_source_flare, render_sources, solve_source_angle,
                     trace_ghost_ray, trace_ghost_ray_diagnostic, trace_ghost_rays, trace_primary_ray,
                     trace_primary_ray_diagnostic, trace_rays_diagnostic)
from .render import _render_gate_debug
from .hostprep import build_spectral_lambdas
from .diffraction import _render_starburst_debug, _render_veil_debug, fft2d
from .psf import PSFCellStatus, render_psf
from .config import MATTE_BOX_OPEN  # noqa: E402
from .lens_file import LensFileError, load_lens
from .source_map import (SourceMapSolve, _solve_source_map, _source_map_landing, _source_map_screen, calibrate_lens,
                         solve_source_map)


from .edge_debug import (_aperture_edge_distance_debug, _aperture_sdf_debug, _gate_params_debug,  # noqa: E402
                         _gate_sample_debug, _hurb_sample_debug)


def _cuda_available() -> bool:
    """Reference name for the device check (bind_renderers.cpp:518)."""
    return device_available()


__all__ = [n for n in dir() if not n.startswith("_")] + ["_cuda_available", "_render_starburst_debug",
                                                         "_render_veil_debug", "_solve_source_map",
                                                         "_source_map_landing", "_source_map_screen",
                                                         "_hurb_sample_debug", "_aperture_edge_distance_debug",
                                                         "_aperture_sdf_debug", "_gate_params_debug",
                                                         "_gate_sample_debug", "_render_gate_debug"]
