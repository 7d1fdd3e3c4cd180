"""GPU: HURB edge diffraction and the front-of-lens baffle stack through the
public API against the CPU oracle (oracle/gl_oracle_trace.h restates hurb.h,
aperture_edge_distance and baffle.h).

* Baffles are geometry only: the device render matches the FP32 oracle to the
  FP32-path bar with the same executed segment count.
* HURB draws its kicks from integer-hashed streams identical on both sides;
  the kick angles go through tanf / logf / cosf, whose device and C-library
  results differ by an ulp, so a kicked ray lands a few ulp apart and the
  image is compared at 1e-3 rel-L2 (the build's bar for images against FP64).
* The starburst pupil clipped by a baffle is the numpy restatement's bit for
  bit, and its throughput is the clipped energy fraction."""
import math

import numpy as np
import pytest

from ghostlight_amd import FlareConfig, hostprep, render_ghosts, synthetic
from ghostlight_amd import diffraction as dfr
from ghostlight_amd.config import Baffle, BaffleShape, HurbKickDistribution, StarburstEngine, to_params
from oracle import diffraction_oracle as dor
from oracle import pyoracle

pytestmark = pytest.mark.gpu

W, H = 192, 128


@pytest.fixture(scope="module")
def setup():
    lens = synthetic.config1_lens()
    hw, hh = synthetic.sensor_for(lens, W, H, 14.0)
    pairs, boosts = pyoracle.filter_ghost_pairs(lens, hw, hh)
    return lens, hw, hh, pairs, boosts


def _pair(setup, cfg, src):
    lens, hw, hh, pairs, boosts = setup
    planes, stats = render_ghosts(lens, src, W, H, cfg, hw, hh, pairs=pairs, boosts=boosts, stats=True)
    spec = hostprep.spectral_table(cfg.spectral_samples)
    p, _keep = to_params(cfg)
    o32, info = pyoracle.render_ghosts(lens, pairs, boosts, src, spec, p, hw, hh, W, H,
                                       correction=hostprep.spectral_correction(spec))
    return np.stack(planes), stats, np.stack(o32), info


SRC = np.array([[math.radians(6.0), math.radians(-2.0), 1.0, 0.9, 0.8]], np.float32)


@pytest.mark.parametrize("conc", [False, True])
def test_baffled_render_matches_oracle(setup, conc):
    cfg = FlareConfig(ray_grid=48, spectral_samples=6, concentrate_samples=conc)
    d = cfg.diffraction
    d.matte_box.enabled = True
    d.matte_box.top, d.matte_box.right = 4.0, 6.5
    d.baffles = [Baffle(shape=BaffleShape.CIRCLE, z_front_mm=25.0, cx=0.5, cy=-0.3, radius=9.0, aspect=0.8)]
    img, stats, ref, info = _pair(setup, cfg, SRC)
    err = pyoracle.rel_l2(img, ref)
    print(f"baffles conc={conc}: rel-L2 {err:.2e}, segments {stats['segments']} / {info['segments']}", flush=True)
    assert err <= 1e-4
    assert stats["segments"] == info["segments"]
    plain, _s, _r, info0 = _pair(setup, FlareConfig(ray_grid=48, spectral_samples=6, concentrate_samples=conc), SRC)
    assert info["segments"] < 0.9 * info0["segments"]        # the stack clips real rays


@pytest.mark.parametrize("kick", [HurbKickDistribution.LORENTZIAN, HurbKickDistribution.GAUSSIAN])
def test_hurb_render_matches_oracle(setup, kick):
    cfg = FlareConfig(ray_grid=48, spectral_samples=6, jitter_seed=7)
    d = cfg.diffraction
    d.hurb = True
    d.hurb_kick = kick
    d.hurb_min_sigma_rad = 2.0e-6
    d.matte_box.enabled = True
    d.matte_box.left = 7.0                    # a baffle edge to kick at, too
    img, stats, ref, info = _pair(setup, cfg, SRC)
    err = pyoracle.rel_l2(img, ref)
    seg_rel = abs(stats["segments"] - info["segments"]) / info["segments"]
    print(f"HURB {kick.name}: rel-L2 vs oracle {err:.2e}, segments {stats['segments']} / {info['segments']}",
          flush=True)
    assert err <= 1e-3
    assert seg_rel <= 1e-4
    off = FlareConfig(ray_grid=48, spectral_samples=6, jitter_seed=7)
    off.diffraction.matte_box.enabled = True
    off.diffraction.matte_box.left = 7.0
    img_off, _s, _r, _i = _pair(setup, off, SRC)
    assert pyoracle.rel_l2(img, img_off) > 10 * max(err, 1e-6)   # the kicks move light
    again, _s2, _r2, _i2 = _pair(setup, cfg, SRC)
    assert pyoracle.rel_l2(again, img) <= 1e-6                    # deterministic streams


def test_starburst_pupil_clipped_by_baffles():
    lens = synthetic.config1_lens()
    cal = synthetic.calibration_for(lens, W, H)
    cfg = FlareConfig()
    cfg.aperture_blades = 6
    d = cfg.diffraction
    d.starburst = True
    d.starburst_engine = StarburstEngine.SPRITE
    d.starburst_grid = d.starburst_grid_cap = 256
    d.spectral_samples = 8
    d.use_survivor_mask = False
    d.matte_box.enabled = True
    d.matte_box.right = 0.3 * cal.entrance_pupil_semi_x      # a flag over the right of the pupil
    d.matte_box.z_front_mm = 20.0                             # at the spawn plane (dz = 0)
    src = np.array([[0.0, 0.0, 1.0, 1.0, 1.0]], np.float32)
    px, py = np.array([W / 2], np.float32), np.array([H / 2], np.float32)
    planes, dbg = dfr.render_starburst(lens, cal, src, px, py, W, H, cal.sensor_half_w, cal.sensor_half_h, cfg,
                                       debug=True)
    N = dbg["grid"]
    nrm = dor.mask_normals(6, 0.0)
    row = (0, 0.0, 1e9, 1e9, 1e9, d.matte_box.right, 0.0, 0.0, 1e9, 1.0)
    ep = (dbg["entrance_pupil_semi_x"], dbg["entrance_pupil_semi_y"])
    want = dor.build_pupil(N, d.pupil_fill, nrm, baffles=[row], ep=ep)
    assert np.array_equal(dbg["pupil"].astype(np.float64), want)
    full = dor.build_pupil(N, d.pupil_fill, nrm)
    T = float((want ** 2).sum() / (full ** 2).sum())
    print(f"baffled pupil throughput {dbg['throughput']:.5f} (numpy {T:.5f})", flush=True)
    assert dbg["throughput"] == pytest.approx(T, rel=1e-5)
    assert 0.5 < T < 0.9
