"""GPU: IMAGE apertures against the oracle.  The device samples the bitmap with
the same float operations as the oracle (bilinear between texel centres,
clamped edges, > 0.5 passes), so statuses and landings are bit-exact; the
ghost image through a bitmap stop matches at the FP32-path bar; the front
bitmap folded into the starburst pupil matches the numpy restatement."""
import math
import os

import numpy as np
import pytest

from ghostlight_amd import ApertureShape, FlareConfig, OpticalSystem, hostprep, render_ghosts, synthetic
from ghostlight_amd import diffraction as dfr
from ghostlight_amd import trace_ghost_rays
from ghostlight_amd.config import StarburstEngine, to_params
from oracle import diffraction_oracle as dor
from oracle import pyoracle

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
LENS = os.path.join(HERE, "golden", "lenses", "DoubleGauss.lens")


def _star(n=160, lobes=5, size=1.0):
    c = (np.arange(n) + 0.5) / n * 2 - 1
    X, Y = np.meshgrid(c, c)
    r, th = np.hypot(X, Y), np.arctan2(Y, X)
    edge = size * (0.62 + 0.22 * np.cos(lobes * th))
    return np.clip((edge - r) * 12.0 + 0.5, 0.0, 1.0).astype(np.float32)   # soft, sub-texel edge


def _image_lens(k_surface=None, size=1.0):
    lens = OpticalSystem.load(LENS)
    k = next(i for i, s in enumerate(lens.surfaces) if s.is_stop) if k_surface is None else k_surface
    s = lens.surfaces[k]
    s.aperture_shape = ApertureShape.IMAGE
    s.aperture_semi_diameter = s.semi_aperture
    lens.aperture_images[k].set_pixels(_star(size=size))
    lens.aperture_images[k].semi_diameter = s.semi_aperture
    return lens, k


def _rays(lens, n, seed):
    rng = np.random.default_rng(seed)
    fR = lens.surfaces[0].semi_aperture
    z0 = lens.surfaces[0].z - 20.0
    ang = np.radians(rng.uniform(-6, 6, (n, 2)))
    d = np.stack([np.tan(ang[:, 0]), np.tan(ang[:, 1]), np.ones(n)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 7), np.float32)
    r[:, :2] = rng.uniform(-1, 1, (n, 2)) * fR - 20.0 * d[:, :2] / d[:, 2:3]
    r[:, 2] = z0
    r[:, 3:6] = d
    r[:, 6] = rng.uniform(420, 680, n)
    return r


@pytest.mark.parametrize("pair", [(-1, -1), (1, 8), (3, 9)])
def test_image_stop_traces_bitexact(pair):
    lens, _k = _image_lens()
    rays = _rays(lens, 4000, 3)
    a, b = pair
    pos, _w, st, sg = trace_ghost_rays(lens, a, b, rays, t_ge0=True, fp64_radius_min=0.0)
    opos, _ow, ost, osg = pyoracle.trace_ghost(lens, a, b, rays, t_ge0=True)
    print(f"image stop pair {pair}: {int((st == 0).sum())} of {len(st)} pass", flush=True)
    assert np.array_equal(st, ost) and np.array_equal(sg, osg)
    ok = st == 0
    assert np.array_equal(pos[ok, :2], opos[ok, :2].astype(np.float32))
    plain = OpticalSystem.load(LENS)
    _p, _w2, st2, _g = trace_ghost_rays(plain, a, b, rays, t_ge0=True, fp64_radius_min=0.0)
    assert (st2 == 0).sum() > ok.sum()                  # the bitmap removes rays


def test_image_stop_ghost_render_matches_oracle():
    lens, _k = _image_lens()
    W = H = 160
    cal = lens.calibration()
    hw, hh = cal.sensor_half_w, cal.sensor_half_h
    pairs, boosts = pyoracle.filter_ghost_pairs(lens, hw, hh)
    cfg = FlareConfig(ray_grid=48, spectral_samples=6)
    src = np.array([[math.radians(3.0), math.radians(-1.5), 1.0, 1.0, 1.0]], np.float32)
    planes, stats = render_ghosts(lens, src, W, H, cfg, hw, hh, pairs=pairs, boosts=boosts, stats=True)
    spec = hostprep.spectral_table(6)
    p, _keep = to_params(cfg)
    ref, info = pyoracle.render_ghosts(lens, pairs, boosts, src, spec, p, hw, hh, W, H,
                                       correction=hostprep.spectral_correction(spec))
    err = pyoracle.rel_l2(np.stack(planes), np.stack(ref))
    print(f"image stop ghosts: rel-L2 {err:.2e}, segments {stats['segments']} / {info['segments']}", flush=True)
    assert err <= 1e-4 and stats["segments"] == info["segments"]


def test_front_bitmap_in_starburst_pupil():
    # a small star on the front element: the entrance pupil covers only the
    # middle of the front glass, so the bitmap is scaled into that disc
    lens, k = _image_lens(k_surface=0, size=0.35)
    cal = synthetic.calibration_for(lens, 192, 128)
    cfg = FlareConfig()
    d = cfg.diffraction
    d.starburst = True
    d.starburst_engine = StarburstEngine.SPRITE
    d.starburst_grid = d.starburst_grid_cap = 256
    d.spectral_samples = 8
    d.use_survivor_mask = False
    src = np.array([[0.0, 0.0, 1.0, 1.0, 1.0]], np.float32)
    px, py = np.array([96.0], np.float32), np.array([64.0], np.float32)
    _planes, dbg = dfr.render_starburst(lens, cal, src, px, py, 192, 128, cal.sensor_half_w, cal.sensor_half_h, cfg,
                                        debug=True)
    N = dbg["grid"]
    ep = (dbg["entrance_pupil_semi_x"], dbg["entrance_pupil_semi_y"])
    s0 = lens.surfaces[0]
    stop = lens.surfaces[next(i for i, s in enumerate(lens.surfaces) if s.is_stop)]
    nrm = dor.mask_normals(stop.aperture_blades, stop.aperture_rotation_rad) \
        if stop.aperture_shape == ApertureShape.POLYGON else None
    z0 = np.float32(s0.z)
    d0 = float(z0 - (z0 - np.float32(20.0)))           # the bitmap surface's distance from the spawn plane
    want = dor.build_pupil(N, d.pupil_fill, nrm, ep=ep, tex=(_star(size=0.35), 1.0, s0.semi_aperture, d0))
    err = np.max(np.abs(dbg["pupil"].astype(np.float64) - want))
    full = dor.build_pupil(N, d.pupil_fill, nrm)
    T = float((want ** 2).sum() / (full ** 2).sum())
    print(f"front bitmap pupil: max |diff| {err:.2e}, throughput {dbg['throughput']:.5f} (numpy {T:.5f})",
          flush=True)
    assert err <= 1e-6
    assert dbg["throughput"] == pytest.approx(T, rel=1e-4)
    assert T < 0.95


def test_hurb_kicks_at_the_image_stop_edge():
    """HURB on a bitmap stop: the kicks read the stop's baked distance field on
    both sides; the render matches the oracle and differs from HURB off."""
    from ghostlight_amd.config import HurbKickDistribution
    lens, _k = _image_lens()
    W = H = 128
    cal = lens.calibration()
    hw, hh = cal.sensor_half_w, cal.sensor_half_h
    pairs, boosts = pyoracle.filter_ghost_pairs(lens, hw, hh)
    src = np.array([[math.radians(2.0), math.radians(1.0), 1.0, 1.0, 1.0]], np.float32)
    spec = hostprep.spectral_table(6)

    def both(hurb):
        cfg = FlareConfig(ray_grid=48, spectral_samples=6, jitter_seed=3)
        cfg.diffraction.hurb = hurb
        cfg.diffraction.hurb_kick = HurbKickDistribution.GAUSSIAN
        cfg.diffraction.hurb_min_sigma_rad = 1.0e-6
        planes = np.stack(render_ghosts(lens, src, W, H, cfg, hw, hh, pairs=pairs, boosts=boosts))
        p, _keep = to_params(cfg)
        ref, _info = pyoracle.render_ghosts(lens, pairs, boosts, src, spec, p, hw, hh, W, H,
                                            correction=hostprep.spectral_correction(spec))
        return planes, np.stack(ref)

    on, on_ref = both(True)
    off, _ = both(False)
    err = pyoracle.rel_l2(on, on_ref)
    moved = pyoracle.rel_l2(on, off)
    print(f"HURB at the image stop: rel-L2 vs oracle {err:.2e}, vs HURB off {moved:.2e}", flush=True)
    assert err <= 1e-3 and moved > 10 * max(err, 1e-6)
