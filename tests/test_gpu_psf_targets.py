"""PSF grid with FIXED_TARGET anchoring and the deep-stop fallback on real
lens files (test_render_psf.py:173-300, 458-525): the chief aimed onto every
sensor cell of a Super-35 frame, the visible peak centred in its tile,
vignetting flagged on an oversized sensor, and deep-stop lenses (stops far
smaller than the front element) resolved through the tracked axial window
instead of reported DARK."""
import math
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LENS_DIR = os.path.join(HERE, "golden", "lenses")

OK, VIGNETTED, DARK = 0, 1, 2


def _load(name):
    from ghostlight_amd import OpticalSystem
    return OpticalSystem.load(os.path.join(LENS_DIR, name))


@pytest.fixture(scope="module")
def psf_lens():
    """Double Gauss 58 mm f/2: contained coma off axis, image circle ~38 mm."""
    return _load("DoubleGauss.lens")


def _sensor_cells(grid_n, calib, half_w, half_h):
    """Raster-order cell centres (row 0 at -y) with tan-linear seed angles."""
    cal_hw = max(1e-6, float(calib.sensor_half_w))
    cal_hh = max(1e-6, float(calib.sensor_half_h))
    tan_h, tan_v = math.tan(float(calib.max_half_angle_h)), math.tan(float(calib.max_half_angle_v))
    seeds = np.empty((grid_n * grid_n, 2), np.float32)
    targets = np.empty((grid_n * grid_n, 2), np.float32)
    k = 0
    for row in range(grid_n):
        y = 0.0 if grid_n == 1 else -half_h + (row + 0.5) * (2 * half_h / grid_n)
        for col in range(grid_n):
            x = 0.0 if grid_n == 1 else -half_w + (col + 0.5) * (2 * half_w / grid_n)
            targets[k] = (x, y)
            rx, ry = max(-1.0, min(1.0, x / cal_hw)), max(-1.0, min(1.0, y / cal_hh))
            seeds[k] = (math.atan(rx * tan_h), math.atan(ry * tan_v))
            k += 1
    return seeds, targets


def _target_cfg(grid_n, tile=96, extent=0.5):
    from ghostlight_amd import PSFCenterMode, PSFConfig
    return PSFConfig(grid_nx=grid_n, grid_ny=grid_n, tile_w=tile, tile_h=tile, tile_extent_mm=extent, ray_grid=128,
                     spectral_samples=8, splat_sigma_um=4.0, pupil_jitter=2,
                     center_mode=PSFCenterMode.PSF_CENTER_FIXED_TARGET)


def test_target_mode_report_arrays(psf_lens):
    seeds, targets = _sensor_cells(3, psf_lens.calibration(), 12.4, 9.3)
    out = psf_lens.render_psf(seeds, _target_cfg(3), targets_mm=targets)
    for key in ("status", "pupil_fraction", "aim_residual_mm"):
        assert out[key].shape == (9,)
    assert out["status"].dtype == np.uint8


def test_target_mode_aims_chief_onto_cell(psf_lens):
    seeds, targets = _sensor_cells(7, psf_lens.calibration(), 24.89 / 2, 18.66 / 2)
    out = psf_lens.render_psf(seeds, _target_cfg(7), targets_mm=targets)
    err_um = np.hypot(out["chief_x_mm"] - targets[:, 0], out["chief_y_mm"] - targets[:, 1]) * 1e3
    print(f"DoubleGauss Super-35 aim: max chief error {err_um.max():.3f} um", flush=True)
    assert err_um.max() < 1.0
    assert out["aim_residual_mm"].max() * 1e3 < 1.0
    assert (out["status"] == OK).all()
    assert (out["pupil_fraction"] > 0).all() and (out["pupil_fraction"] <= 1).all()


def test_target_mode_peak_centred(psf_lens):
    grid_n, tile, extent = 7, 96, 0.5
    seeds, targets = _sensor_cells(grid_n, psf_lens.calibration(), 24.89 / 2, 18.66 / 2)
    out = psf_lens.render_psf(seeds, _target_cfg(grid_n, tile, extent), targets_mm=targets)
    lum = out["r"] + out["g"] + out["b"]
    mmpp = extent / tile
    offs = []
    for i in range(grid_n * grid_n):
        gx, gy = i % grid_n, i // grid_n
        t = lum[gy * tile:(gy + 1) * tile, gx * tile:(gx + 1) * tile]
        iy, ix = np.unravel_index(np.argmax(t), t.shape)
        offs.append(np.hypot(ix + 0.5 - tile / 2, iy + 0.5 - tile / 2) * mmpp * 1e3)
    offs = np.array(offs)
    print(f"peak offsets (um): max {offs.max():.2f}, corners {offs[[0, grid_n - 1, -grid_n, -1]]}", flush=True)
    assert offs.max() < 16.0
    assert offs[[0, grid_n - 1, -grid_n, -1]].max() < 13.0


def test_big_sensor_flags_vignetting(psf_lens):
    cfg = _target_cfg(7)
    seeds, targets = _sensor_cells(7, psf_lens.calibration(), 70.0 / 2, 48.5 / 2)
    out = psf_lens.render_psf(seeds, cfg, targets_mm=targets)
    status = out["status"]
    assert (status != OK).sum() > 0
    lum = out["r"] + out["g"] + out["b"]
    t = cfg.tile_w
    lit_vignetted = sum(1 for i in range(49) if status[i] == VIGNETTED
                        and lum[(i // 7) * t:(i // 7 + 1) * t, (i % 7) * t:(i % 7 + 1) * t].sum() > 0)
    assert lit_vignetted > 0, "vignetted cells rendered no light"


def test_smaller_sensor_fewer_vignetted(psf_lens):
    cfg = _target_cfg(7)

    def n_flagged(hw, hh):
        s, t = _sensor_cells(7, psf_lens.calibration(), hw, hh)
        return int((psf_lens.render_psf(s, cfg, targets_mm=t)["status"] != OK).sum())

    big, small = n_flagged(70.0 / 2, 48.5 / 2), n_flagged(24.89 / 2, 18.66 / 2)
    assert small < big and small == 0


def _deep_stop_cells(lens, grid_n, half_w, half_h):
    cal = lens.calibration()
    fx = cal.sensor_half_w / np.tan(cal.max_half_angle_h)
    fy = cal.sensor_half_h / np.tan(cal.max_half_angle_v)
    seeds, targets = [], []
    for j in range(grid_n):
        for i in range(grid_n):
            tx = (-1.0 + 2.0 * (i + 0.5) / grid_n) * half_w
            ty = (-1.0 + 2.0 * (j + 0.5) / grid_n) * half_h
            seeds.append([np.arctan(tx / fx), np.arctan(ty / fy), 1.0, 1.0, 1.0])
            targets.append([tx, ty])
    return np.asarray(seeds, np.float32), np.asarray(targets, np.float32)


@pytest.mark.parametrize("name", ["AtlasLensCo_40mm_F3.0_1.5x_US12379576B2_Table1.lens", "zodiac1.lens"])
def test_deep_stop_cells_are_not_reported_dark(name):
    """The Atlas anamorphic passes light through ~16% of its front radius and
    zodiac1 through ~4%: the whole-disc aim probes land under the minimum-hits
    bar, and only the tracked axial window resolves the lit field."""
    lens = _load(name)
    seeds, targets = _deep_stop_cells(lens, 7, 24.89 / 2, 18.66 / 2)
    out = lens.render_psf(seeds, _target_cfg(7, tile=32, extent=0.5), targets_mm=targets)
    status = out["status"]
    dark = int((status == DARK).sum())
    lit = status != DARK
    resid = out["aim_residual_mm"][lit]
    print(f"{name}: {dark}/49 DARK, {(status == VIGNETTED).sum()} vignetted, max residual "
          f"{resid.max() if resid.size else float('nan'):.4f} mm", flush=True)
    assert dark <= 3
    assert np.isfinite(resid).all() and resid.max() < 0.5
    # the lit cells' chiefs land near their targets and their tiles hold light
    lum = out["r"] + out["g"] + out["b"]
    lit_tiles = sum(1 for i in np.nonzero(lit)[0] if lum[(i // 7) * 32:(i // 7 + 1) * 32,
                                                         (i % 7) * 32:(i % 7 + 1) * 32].sum() > 0)
    assert lit_tiles >= int(lit.sum()) - 3


def test_deep_stop_centroid_mode_retries_dark_cells():
    """CHIEF_CENTROID mode: where the 7 x 7 disc probe finds no survivor on a
    deep-stop lens, the tracked-window retry reports the cell OK with a finite
    survivor-mean chief; the pupil fraction stays the disc measure."""
    from ghostlight_amd import PSFConfig
    from ghostlight_amd.psf import deep_pupil, pupil_probe
    lens = _load("zodiac1.lens")
    seeds, _t = _deep_stop_cells(lens, 5, 24.89 / 2, 18.66 / 2)
    cfg = PSFConfig(grid_nx=5, grid_ny=5, tile_w=32, tile_h=32, tile_extent_mm=0.5, ray_grid=64, spectral_samples=4)
    ax, ay = seeds[:, 0].astype(np.float64), seeds[:, 1].astype(np.float64)
    hits, frac, _mx, _my = pupil_probe(lens, ax, ay, 7, None, 1000.0)      # the pre-pass's FP64 threshold
    out = lens.render_psf(seeds, cfg)
    missed = hits == 0
    print(f"zodiac1 centroid: {missed.sum()}/25 cells missed by the disc probe, statuses {out['status']}",
          flush=True)
    assert missed.any(), "the disc probe should starve on a 4% stop"
    deep = deep_pupil(lens, seeds[missed, 0], seeds[missed, 1])
    found = deep[:, 5] > 0
    assert found.any()
    idx = np.nonzero(missed)[0]
    assert (out["status"][idx[found]] == OK).all()
    assert np.allclose(out["chief_x_mm"][idx[found]], deep[found, 2])
    assert np.allclose(out["chief_y_mm"][idx[found]], deep[found, 3])
    assert (out["pupil_fraction"][idx] == frac[idx]).all()


def test_deep_stop_fallback_leaves_a_well_behaved_lens_alone(psf_lens):
    seeds, targets = _deep_stop_cells(psf_lens, 7, 24.89 / 2, 18.66 / 2)
    out = psf_lens.render_psf(seeds, _target_cfg(7, tile=32, extent=0.5), targets_mm=targets)
    assert (out["status"] == OK).all()
    err_um = np.hypot(out["chief_x_mm"] - targets[:, 0], out["chief_y_mm"] - targets[:, 1]) * 1e3
    assert err_um.max() < 1.0
