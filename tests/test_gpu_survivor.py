"""The starburst's off-axis effective pupil on the device (survivor centroid
and cat's-eye envelope, starburst_render.cu:393-582, on by default through
DiffractionConfig.use_survivor_mask) against the oracle restatement
(oracle/survivor_oracle.py, itself pinned by the two-disc closed form in
test_oracle_survivor.py), through _render_starburst_debug as
validate_starburst_pupil.py drives it: DoubleGauss and the 1.33x Sirui,
8-blade override rotated 22.5 deg, the source swept as in fig_cutoff: a 5x
sensor and field angles of 0, 0.5, 0.78, 0.97 and 1.1 times the lens's
calibrated field edge, so the cat's eye narrows until the rim vignetting
bites.

Bars: survivor centre bit-exact (same FP32 survival answers, FP64 sums in
the same raster order); pupil within 1e-6 per texel of the oracle's (the
same float32 envelope lookups, summed over the 4 x 4 supersamples in float32
on the device and in FP64 by the oracle); throughput within 1e-4; the
throughput never rises as the field grows and is below 1 at the field
edge."""
import math

import numpy as np
import pytest

import golden_flare as gf
from ghostlight_amd import _render_starburst_debug
from ghostlight_amd.config import PointFlareConfig
from oracle import diffraction_oracle as dor
from oracle import survivor_oracle as so

pytestmark = pytest.mark.gpu

GRID = 512


SCALE = 5.0
FRACS = (0.0, 0.5, 0.78, 0.97, 1.1)


def _sx_for(cal, frac):
    """Source x (0..1) whose field angle is frac x the calibrated edge on the
    SCALE x sensor (validate_starburst_pupil.py fig_cutoff)."""
    tw = SCALE * math.tan(cal.max_half_angle_h)
    return 0.5 + 0.5 * math.tan(frac * cal.max_half_angle_h) / tw


def _cfg(cal, sx, sy=0.5):
    c = PointFlareConfig(source_x=sx, source_y=sy, source_r=1.0, source_g=1.0, source_b=1.0, spectral_samples=12,
                         aperture_blades=8, aperture_rotation=22.5, sensor_half_w=cal.sensor_half_w * SCALE,
                         sensor_half_h=cal.sensor_half_h * SCALE)
    d = c.diffraction
    d.starburst = True
    d.spectral_samples = 12
    d.starburst_grid = GRID
    d.starburst_grid_cap = GRID          # keep the 5x-sensor sprite from growing the grid
    d.use_survivor_mask = True
    d.use_surface_textures = False
    return c


@pytest.mark.parametrize("name", gf.LENSES)
def test_off_axis_pupil_matches_oracle(name):
    lens, cal = gf.lens(name), gf.calibration(name)
    T_prev = 1.0
    for frac in FRACS:
        sx = _sx_for(cal, frac)
        cfg = _cfg(cal, sx)
        r = _render_starburst_debug(GRID, GRID, lens, cal, cfg)
        N = r["grid"]
        ax, ay = np.float32(r["field_angle_x"]), np.float32(r["field_angle_y"])
        T = float(r["throughput"])
        fill = cfg.diffraction.pupil_fill
        nrm = dor.mask_normals(8, float(np.float32(22.5) * np.float32(math.pi / 180.0)))
        plain = dor.build_pupil(N, fill, nrm)
        print(f"{name} {frac} x edge: field ({math.degrees(ax):.2f}, {math.degrees(ay):.2f}) deg, T {T:.5f}, "
              f"centre ({r['chief_offset_x']:.5f}, {r['chief_offset_y']:.5f})", flush=True)
        if frac == 0.0:
            # on axis: no envelope, the plain stop, T = 1
            assert abs(ax) < 1e-5 and abs(ay) < 1e-5
            assert T == 1.0
            assert np.array_equal(r["pupil"].astype(np.float64), plain)
            continue
        cx, cy, any_ = so.survivor_center(lens, ax, ay)
        assert any_
        assert np.float32(r["chief_offset_x"]) == cx and np.float32(r["chief_offset_y"]) == cy
        ep_x, ep_y = r["entrance_pupil_semi_x"], r["entrance_pupil_semi_y"]
        env = so.survivor_envelope(lens, None, ep_x, ep_y, cx, cy, ax, ay, cfg.diffraction.survivor_grid)
        pup = dor.build_pupil(N, fill, nrm, env=env)
        dev = r["pupil"].astype(np.float64)
        err = np.abs(dev - pup)
        T_ref = float((pup ** 2).sum() / (plain ** 2).sum())
        print(f"  pupil max |dev - oracle| {err.max():.2e}, T oracle {T_ref:.5f}, envelope mean "
              f"{float(env.mean()):.4f}", flush=True)
        assert err.max() <= 1e-6
        assert abs(T - T_ref) <= 1e-4
        assert 0.0 < T <= T_prev + 1e-7
        T_prev = T
    assert T_prev < 0.95, "no rim vignetting reached"
