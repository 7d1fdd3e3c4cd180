"""The calibration / source-map fixture (tests/golden/source_map_goldens.npz,
written by tests/golden/make_golden.py from the reference's CPU code:
calibrate_lens, lens_calibration.cpp:344; solve_source_angle,
source_map.cpp:182; source_map_screen, :456; source_map_landing, :487) for
18 of the reference's lens files, plus the calibration manifest the reference
keeps beside its aperture goldens."""
from __future__ import annotations

import os

import numpy as np

from ghostlight_amd import LensCalibration, OpticalSystem

PATH = os.path.join(os.path.dirname(__file__), "golden", "source_map_goldens.npz")
LENSES = ("double_gauss", "helios44", "atlas_40_f3", "atlas_40_f35", "cooke_50_2x", "sirui_24", "sirui",
          "sirui_venus", "exp_helios_2x", "exp_panatar", "artisans_50", "canon_100", "canon_24", "cooke_pancro",
          "summicron_100", "hex_stop", "hex_rot30", "blade_shape")
# Polygonal stops: the product tests the pupil mask in half-plane form, the
# reference through atan2f/fmodf/cosf, so a sample on a blade edge can fall on
# the other side and the probe sums move by a sample's weight.
POLYGON_STOPS = ("sirui_venus", "hex_stop", "hex_rot30", "blade_shape")
SOLVE_FIELDS = ("angle_x", "angle_y", "status", "residual_mm", "anchor_ax", "anchor_ay", "anchor_x", "anchor_y",
                "J0", "J1", "J2", "J3", "eff_half_w", "eff_half_h", "inv_tan_w", "inv_tan_h", "probe_evals")

_Z = None


def fixture():
    global _Z
    if _Z is None:
        _Z = np.load(PATH)
    return _Z


def lens(name) -> OpticalSystem:
    z = fixture()
    return OpticalSystem.from_packed(z[f"sm_{name}_surfaces"], z[f"sm_{name}_arena"], name=name)


def calibration(name) -> LensCalibration:
    return LensCalibration.from_array(fixture()[f"sm_{name}_calib"])
