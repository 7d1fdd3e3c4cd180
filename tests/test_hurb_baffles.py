"""CPU: HURB edge diffraction (hurb.h) and the front-of-lens baffle stack
(baffle.h) on the host side of the product (gl_hurb_sample,
gl_aperture_edge_distance: the same __host__ __device__ functions the kernels
run) against the oracle and against closed forms, following the reference's
test_hurb.py and test_blade_shape_hurb.py:

  * sigma = lambda K / d with K = 1 / (2 sqrt 2 pi) (Gaussian) and 1 / (2 pi)
    (Lorentzian); the Gaussian's standard deviation is sigma; the Lorentzian is
    a Cauchy law with median |x| = IQR / 2 = gamma; widths scale with lambda;
    the clamp holds; streams are deterministic per seed;
  * over a slit the kick envelope sits at the diffraction scale lambda / w;
  * the edge distance: exact radial clearance of a circle, the apothem and the
    side line of a plain polygon, zero on a deformed blade's silhouette;
  * the baffle stack: an open stack changes nothing, a matte box flag that
    covers half the front element removes that half of the rays."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from ghostlight_amd import FlareConfig, _abi, synthetic
from ghostlight_amd.config import Baffle, BaffleShape, fill_baffles, to_params
from oracle import diffraction_oracle as dor
from oracle import pyoracle

G, L = 0, 1
BIG = 1.0e3


def _kicks(lam, d, dist, seed=1, max_kick=BIG, oracle=False):
    d = np.ascontiguousarray(d, np.float32)
    k = np.zeros(len(d), np.float32)
    s = np.zeros(len(d), np.float32)
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    if oracle:
        pyoracle.lib().orc_hurb_sample(C.c_float(lam), fp(d), len(d), dist, C.c_float(max_kick), C.c_uint32(seed),
                                       fp(k), fp(s))
    else:
        assert _abi.load_library().gl_hurb_sample(lam, fp(d), len(d), dist, max_kick, seed, fp(k), fp(s)) == 0
    return k, s


def test_sigma_formula_and_constants():
    d = np.array([0.05, 0.2, 1.0, 3.0], np.float32)
    _k, s = _kicks(550.0, d, G)
    assert np.allclose(s, 550e-6 / (2 * math.sqrt(2) * math.pi) / d, rtol=1e-5)
    _k, s = _kicks(550.0, d, L)
    assert np.allclose(s, 550e-6 / (2 * math.pi) / d, rtol=1e-5)


def test_no_kick_on_or_past_the_edge():
    k, s = _kicks(550.0, np.array([0.0, -1.0], np.float32), G)
    assert (s == 0).all() and (k == 0).all()


def test_product_streams_equal_the_oracle():
    d = np.random.default_rng(3).uniform(1e-3, 2.0, 20000).astype(np.float32)
    for dist in (G, L):
        kp, sp = _kicks(612.0, d, dist, seed=77, max_kick=0.35)
        ko, so = _kicks(612.0, d, dist, seed=77, max_kick=0.35, oracle=True)
        assert np.array_equal(sp, so)
        assert np.allclose(kp, ko, rtol=2e-6, atol=1e-12)


def test_gaussian_is_normal_sigma():
    sigma = 550e-6 / (2 * math.sqrt(2) * math.pi) / 0.5
    k, _ = _kicks(550.0, np.full(400_000, 0.5, np.float32), G, seed=11)
    assert np.std(k) == pytest.approx(sigma, rel=0.02)
    assert abs(np.mean(k)) < 0.02 * sigma


def test_lorentzian_is_cauchy_gamma():
    gamma = 550e-6 / (2 * math.pi) / 0.5
    k, _ = _kicks(550.0, np.full(400_000, 0.5, np.float32), L, seed=13)
    assert np.median(np.abs(k)) == pytest.approx(gamma, rel=0.03)
    assert (np.percentile(k, 75) - np.percentile(k, 25)) / 2 == pytest.approx(gamma, rel=0.03)


def test_width_scales_with_lambda_and_clamp_holds():
    d = np.full(300_000, 0.5, np.float32)
    base = np.std(_kicks(550.0, d, G, seed=5)[0])
    for lam in (450.0, 650.0):
        assert np.std(_kicks(lam, d, G, seed=5)[0]) / base == pytest.approx(lam / 550.0, rel=0.04)
    k, _ = _kicks(550.0, np.full(20000, 1e-4, np.float32), L, seed=7, max_kick=0.35)
    assert np.all(np.abs(k) <= 0.35 + 1e-6) and np.max(np.abs(k)) == pytest.approx(0.35, rel=0.05)


def test_streams_are_deterministic_per_seed():
    d = np.full(5000, 0.4, np.float32)
    a, b, c = (_kicks(550.0, d, L, seed=s)[0] for s in (99, 99, 100))
    assert np.array_equal(a, b) and not np.array_equal(a, c)


def test_slit_envelope_at_diffraction_scale():
    a = 0.5
    theta0 = 550e-6 / (2 * a)
    x = (np.random.default_rng(0).random(1_000_000).astype(np.float32) * 2 - 1) * a
    d = (a - np.abs(x)).astype(np.float32)
    kg, _ = _kicks(550.0, d, G, seed=3, max_kick=0.35)
    kl, _ = _kicks(550.0, d, L, seed=4, max_kick=0.35)
    for k in (kg, kl):
        assert 0.2 * theta0 <= np.median(np.abs(k)) <= 1.5 * theta0
    assert 0.70 <= float(np.mean(np.abs(kg) < theta0)) <= 0.98          # the sinc^2 core
    assert np.mean(np.abs(kl) > 3 * theta0) > 2 * np.mean(np.abs(kg) > 3 * theta0)   # heavier wings


def _surface(**kw):
    s = _abi.gl_surface()
    s.semi_aperture = 10.0
    s.aperture_aspect = 1.0
    s.is_active = 1
    s.rot[0] = s.rot[4] = s.rot[8] = 1.0
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def _edge(s, x, y, z=0.0, oracle=False):
    hit = (C.c_float * 3)(x, y, z)
    d = C.c_float()
    n = (C.c_float * 3)()
    if oracle:
        pyoracle.lib().orc_edge_distance(C.byref(s), hit, C.byref(d), n)
    else:
        assert _abi.load_library().gl_aperture_edge_distance(C.byref(s), hit, C.byref(d), n) == 0
    return d.value, np.array(n[:])


def test_edge_distance_circle_and_plain_polygon():
    s = _surface()
    for r in (0.0, 2.0, 7.5, 9.9):
        d, n = _edge(s, r * 0.6, r * 0.8)
        assert d == pytest.approx(10.0 - r if r > 0 else 10.0, abs=1e-5)
    hexs = _surface(aperture_shape=1, aperture_blades=6)
    apothem = 10.0 * math.cos(math.pi / 6)
    assert _edge(hexs, 0.0, 0.0)[0] == pytest.approx(apothem, abs=1e-5)
    assert _edge(hexs, 0.0, 4.0)[0] == pytest.approx(apothem - 4.0, abs=1e-5)
    d, n = _edge(hexs, 0.0, 4.0)
    assert abs(abs(n[1]) - 1.0) < 1e-5                 # the side's normal


@pytest.mark.parametrize("controls", [dict(aperture_curvature=-0.8), dict(aperture_curvature=0.6),
                                      dict(aperture_curvature=-0.5, aperture_twist=0.9),
                                      dict(aperture_curvature=-0.9, aperture_twist=-0.8,
                                           aperture_notch_rad=math.radians(35.8),
                                           aperture_notch_angle_rad=math.radians(45.0))])
def test_deformed_edge_distance_vanishes_on_the_silhouette(controls):
    s = _surface(aperture_shape=1, aperture_blades=6, **controls)
    pr = _abi.gl_blade_profile()
    _abi.load_library().gl_blade_profile_build(6, 0.0, s.aperture_curvature, s.aperture_twist, s.aperture_notch_rad,
                                               s.aperture_notch_angle_rad, C.byref(pr))
    lib = _abi.load_library()
    for th in np.linspace(0.0, 2 * math.pi, 97)[:-1]:
        rb = 10.0 * lib.gl_blade_radius_at(C.byref(pr), float(th))
        d_in, _ = _edge(s, 0.98 * rb * math.cos(th), 0.98 * rb * math.sin(th))
        d_on, _ = _edge(s, rb * math.cos(th), rb * math.sin(th))
        assert d_on <= 2e-3 and d_in > 0.0
        do, _ = _edge(s, 0.98 * rb * math.cos(th), 0.98 * rb * math.sin(th), oracle=True)
        assert do == pytest.approx(d_in, rel=1e-5, abs=1e-6)


def test_open_baffle_changes_nothing_and_a_flag_halves_the_rays():
    lens = synthetic.config1_lens()
    W = H = 128
    hw, hh = synthetic.sensor_for(lens, W, H, 14.0)
    pairs, boosts = pyoracle.filter_ghost_pairs(lens, hw, hh)
    pairs, boosts = pairs[:6], boosts[:6]
    src = np.array([[0.0, 0.0, 1.0, 1.0, 1.0]], np.float32)
    from ghostlight_amd import hostprep
    spec = hostprep.spectral_table(4)

    def render(cfg):
        p, _keep = to_params(cfg)
        img, info = pyoracle.render_ghosts(lens, pairs, boosts, src, spec, p, hw, hh, W, H, want_masks=True)
        return np.stack(img), info

    base = FlareConfig(ray_grid=48, spectral_samples=4, concentrate_samples=False, cull_dead_pairs=False)
    ref, info0 = render(base)
    far = FlareConfig(ray_grid=48, spectral_samples=4, concentrate_samples=False, cull_dead_pairs=False)
    far.diffraction.baffles = [Baffle(shape=BaffleShape.CIRCLE, z_front_mm=60.0, radius=1e4)]
    img, info1 = render(far)
    assert np.array_equal(img, ref) and info1["segments"] == info0["segments"]
    half = FlareConfig(ray_grid=48, spectral_samples=4, concentrate_samples=False, cull_dead_pairs=False)
    half.diffraction.matte_box.enabled = True
    half.diffraction.matte_box.left = 0.0              # the flag covers x < 0 at its plane
    img, info2 = render(half)
    # on axis the entering rays are parallel, so the flag takes every ray with
    # x < 0: very nearly half of the traced segments go
    assert 0.40 < info2["segments"] / info0["segments"] < 0.60


def test_baffle_stack_marshalling():
    cfg = FlareConfig()
    d = cfg.diffraction
    d.matte_box.enabled = True
    d.matte_box.top = 5.0
    d.baffles = [Baffle(shape=BaffleShape.CIRCLE, z_front_mm=30.0, cx=1.0, radius=12.0, aspect=0.0)] * 9
    out = (_abi.gl_baffle * _abi.GL_MAX_BAFFLES)()
    n = fill_baffles(d, out)
    assert n == _abi.GL_MAX_BAFFLES
    assert out[0].shape == 0 and out[0].dz == pytest.approx(20.0 - 60.0) and out[0].top == 5.0
    assert out[1].shape == 1 and out[1].dz == pytest.approx(-10.0) and out[1].aspect == 1.0


def test_pupil_baffle_clip_oracle_closed_form():
    # a flag at the pupil plane's centre line (dz = 0, on axis) clips exactly
    # the u < 0 half of a round pupil
    normals = None
    full = dor.build_pupil(64, 0.5, normals)
    half = dor.build_pupil(64, 0.5, normals, baffles=[(0, 0.0, 1e9, 1e9, 0.0, 1e9, 0, 0, 1e9, 1.0)],
                           ep=(10.0, 10.0))
    assert half.sum() == pytest.approx(full.sum() / 2, rel=0.02)
    assert np.all(half[:, :30] == 0)
