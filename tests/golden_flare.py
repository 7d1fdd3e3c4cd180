"""Shared set-up for the comparisons against the reference's point-flare
goldens (validation/goldens/point_flare_sampling, reduced to 8x8 box sums,
per-channel totals and peaks in tests/golden/point_flare_goldens.npz by
tests/golden/make_golden.py).

The goldens were rendered by the reference's CUDA build (approximate div/sqrt,
float atomics, validation/point_flare_sampling_baseline.py:28-48): 320x320,
Halton pupil (jitter 2, seed 12345), ray_grid 256, 16 lambdas, sensor half
extents 14 mm, gain 5000, source rgb 10, min_ghost_intensity 0, concentration
off, cull on/off, aperture blades 0 or 6.  The field angles of each source
position come from the reference's own solve_source_angle (source_map.cpp:182)
and are stored with the fixture."""
from __future__ import annotations

import os

import numpy as np

from ghostlight_amd import LensCalibration, OpticalSystem
from ghostlight_amd.config import PointFlareConfig

PATH = os.path.join(os.path.dirname(__file__), "golden", "point_flare_goldens.npz")
W = H = 320
SH = 14.0
DS = 8
LENSES = ("double_gauss", "sirui")
CONFIGS = [(0.5, 0.5, 0), (1.3, 0.5, 0), (1.6, 0.9, 0), (1.3, 0.5, 6)]
CASES = [(name, sx, sy, b, c) for name in LENSES for (sx, sy, b) in CONFIGS for c in (0, 1)]

# The reference renders with approximate division/square root and in a
# different summation order; against its images the oracle and the device
# path agree to ~1e-5 in per-channel totals and ~1e-4 in the 8x8-box rel-L2.
TOL_SUM = 5e-5
TOL_DS = 5e-4
TOL_PEAK = 5e-4

_Z = None


def fixture():
    global _Z
    if _Z is None:
        _Z = np.load(PATH)
    return _Z


def key(name, sx, sy, blades, cull):
    return f"{name}_{sx}_{sy}_b{blades}_c{cull}"


def lens(name):
    z = fixture()
    return OpticalSystem.from_packed(z[f"pf_{name}_surfaces"], z[f"pf_{name}_arena"], name=name)


def calibration(name):
    return LensCalibration.from_array(fixture()[f"pf_{name}_calib"])


def config(sx, sy, blades, cull):
    return PointFlareConfig(ray_grid=256, spectral_samples=16, pupil_jitter=2, jitter_seed=12345,
                            flare_gain=5000.0, min_ghost_intensity=0.0, cull_dead_pairs=bool(cull),
                            concentrate_samples=False, source_x=sx, source_y=sy, source_r=10.0, source_g=10.0,
                            source_b=10.0, sensor_half_w=SH, sensor_half_h=SH, aperture_blades=blades)


def angles(name, sx, sy, blades, cull):
    return fixture()[f"pf_{key(name, sx, sy, blades, cull)}_angles"]


def compare(img, name, sx, sy, blades, cull):
    """img (H, W, 3) -> (sum rel err per channel max, 8x8-box rel-L2, peak rel err)."""
    z = fixture()
    k = key(name, sx, sy, blades, cull)
    img = np.asarray(img, np.float64)
    s = img.sum(axis=(0, 1))
    gs = z[f"pf_{k}_sum"]
    ds = img.reshape(H // DS, DS, W // DS, DS, 3).sum(axis=(1, 3))
    g = z[f"pf_{k}_ds"].astype(np.float64)
    peak = float(img.sum(-1).max())
    gp = float(z[f"pf_{k}_peak"][0])
    return (float(np.max(np.abs(s - gs) / gs)), float(np.linalg.norm(ds - g) / np.linalg.norm(g)),
            abs(peak - gp) / gp)
