"""CPU: the reference's host debug entry points under `import ghostlight`
(bind_renderers.cpp:457-520, bind_trace.cpp:93), driven the way its
test_hurb.py, test_blade_shape_hurb.py and test_aperture_sdf.py drive them:

  * HURB constants K = 1 / (2 sqrt 2 pi), 1 / (2 pi), sigma = lambda K / d,
    no kick on or past the edge, one independent stream per entry;
  * the edge distance of a bladed stop: the circle at curvature +1, the
    waist at the centre, a unit normal in the surface plane;
  * the image-stop distance field against an analytic disc and straight
    edges (one-texel bar), the half-texel correction, stop-only scoping."""
import math

import numpy as np
import pytest

import ghostlight

_impl = ghostlight._ghostlight
G = int(ghostlight.HurbKickDistribution.GAUSSIAN)
L = int(ghostlight.HurbKickDistribution.LORENTZIAN)
SEMI = 10.0


def test_hurb_constants_and_sigma():
    r = _impl._hurb_sample_debug(550.0, np.array([0.5], np.float32), G, 1.0, 1)
    assert r["gauss_k"] == pytest.approx(1 / (2 * math.sqrt(2) * math.pi), rel=1e-6)
    assert r["lorentz_k"] == pytest.approx(1 / (2 * math.pi), rel=1e-6)
    for d in (0.05, 0.2, 1.0, 3.0):
        for dist, k in ((G, r["gauss_k"]), (L, r["lorentz_k"])):
            s = _impl._hurb_sample_debug(550.0, np.array([d], np.float32), dist, 1.0, 1)["sigma"][0]
            assert s == pytest.approx(550e-6 * k / d, rel=1e-5)
    z = _impl._hurb_sample_debug(550.0, np.array([0.0, -1.0], np.float32), G, 1.0, 1)
    assert np.all(z["sigma"] == 0) and np.all(z["kicks"] == 0)


def test_hurb_entries_are_independent_streams():
    d = np.full(200_000, 0.5, np.float32)
    k = _impl._hurb_sample_debug(550.0, d, G, 1e3, 11)["kicks"]
    sigma = 550e-6 / (2 * math.sqrt(2) * math.pi) / 0.5
    assert np.std(k) == pytest.approx(sigma, rel=0.02)
    # neighbouring entries share no stream: their lag-1 correlation is noise
    assert abs(np.corrcoef(k[:-1], k[1:])[0, 1]) < 0.01
    # the stream of entry i depends on the seed and i only
    head = _impl._hurb_sample_debug(550.0, d[:100], G, 1e3, 11)["kicks"]
    assert np.array_equal(head, k[:100])


def _stop(**controls):
    s = ghostlight.Surface()
    s.radius = 0.0
    s.thickness = 0.0
    s.ior = 1.0
    s.semi_aperture = SEMI
    s.is_stop = True
    s.disp_model = ghostlight.DispersionModel.AIR
    s.aperture_shape = ghostlight.ApertureShape.POLYGON
    s.aperture_blades = 6
    for k, v in controls.items():
        setattr(s, k, v)
    return s


def _distance(s, x, y):
    return _impl._aperture_edge_distance_debug(ghostlight.Vec3f(x, y, 0.0), s)[0]


def test_full_curvature_measures_the_circle():
    s = _stop(aperture_curvature=1.0)
    for i in range(32):
        th = 2 * math.pi * i / 32
        for r in (2.0, 6.0, 9.0):
            assert _distance(s, r * math.cos(th), r * math.sin(th)) == pytest.approx(SEMI - r, abs=2e-4)


def test_dead_centre_reports_the_waist():
    s = _stop(aperture_curvature=-0.8)
    assert s.aperture_profile.deformed()
    assert _distance(s, 0.0, 0.0) == pytest.approx(SEMI * s.aperture_profile.r_w, abs=1e-5)


@pytest.mark.parametrize("controls", [dict(aperture_curvature=-0.8), dict(aperture_curvature=-0.5, aperture_twist=0.9)])
def test_silhouette_and_unit_normal(controls):
    s = _stop(**controls)
    for th in np.linspace(0, 2 * math.pi, 41)[:-1]:
        rb = SEMI * s.aperture_profile.radius_at(th)
        assert _distance(s, rb * math.cos(th), rb * math.sin(th)) <= 2e-3
        d, n = _impl._aperture_edge_distance_debug(ghostlight.Vec3f(0.7 * rb * math.cos(th),
                                                                    0.7 * rb * math.sin(th), 0.0), s)
        assert d > 0 and math.hypot(n.x, n.y) == pytest.approx(1.0, abs=1e-5) and n.z == 0.0


def _image_stop_system(pixels, semi_diameter=5.0, aspect=1.0, semi_aperture=10.0):
    lens = ghostlight.OpticalSystem()
    stop = ghostlight.Surface()
    stop.radius, stop.thickness, stop.ior, stop.semi_aperture = 0.0, 0.0, 1.0, semi_aperture
    stop.is_stop = True
    stop.disp_model = ghostlight.DispersionModel.AIR
    stop.aperture_shape = ghostlight.ApertureShape.IMAGE
    stop.aperture_aspect = aspect
    stop.aperture_semi_diameter = semi_diameter
    lens.surfaces.append(stop)
    lens.finalize()
    img = lens.aperture_images[0]
    img.semi_diameter = semi_diameter
    img.pixels = pixels.astype(np.float32)
    return lens


def _texel_xy(W, H, sd, aspect):
    x = aspect * sd * ((2.0 * np.arange(W) + 1.0) / W - 1.0)
    y = sd * ((2.0 * np.arange(H) + 1.0) / H - 1.0)
    return np.meshgrid(x, y)


def test_sdf_disc_accuracy_and_radial_normal():
    W = H = 256
    sd, R = 8.0, 5.0
    xx, yy = _texel_xy(W, H, sd, 1.0)
    r = np.hypot(xx, yy)
    d = _impl._aperture_sdf_debug(_image_stop_system((r < R).astype(np.float32), semi_diameter=sd))
    assert d["stop_index"] == 0 and d["sdf"].shape == (H, W)
    band = np.abs(R - r) < 0.6 * R
    assert np.percentile(np.abs(d["sdf"][band] - (R - r)[band]), 95) < 1.2 * max(d["sx"], d["sy"])
    ring = (np.abs(R - r) < 1.0) & (r > 0.5)
    dot = np.abs(d["nx"][ring] * xx[ring] / r[ring] + d["ny"][ring] * yy[ring] / r[ring])
    assert np.median(dot) > 0.98


def test_sdf_anisotropic_edge_and_half_texel():
    W, H, sd, aspect, x_edge = 200, 120, 5.0, 2.0, 1.3
    xx, _yy = _texel_xy(W, H, sd, aspect)
    d = _impl._aperture_sdf_debug(_image_stop_system((xx < x_edge).astype(np.float32), sd, aspect))
    assert d["sx"] == pytest.approx(2 * aspect * sd / W, rel=1e-6) and d["sy"] == pytest.approx(2 * sd / H, rel=1e-6)
    band = np.abs(x_edge - xx) < 2.0
    assert np.percentile(np.abs(d["sdf"][band] - (x_edge - xx)[band]), 95) < 1.2 * d["sx"]
    m = np.zeros((64, 64), np.float32)
    m[:, :32] = 1.0
    h = _impl._aperture_sdf_debug(_image_stop_system(m, semi_diameter=4.0))
    assert h["sdf"][32, 31] == pytest.approx(0.5 * h["sx"], abs=0.35 * h["sx"])


def test_sdf_only_for_a_loaded_image_stop():
    lens = ghostlight.OpticalSystem()
    s0 = ghostlight.Surface(thickness=5.0, semi_aperture=10.0, disp_model=ghostlight.DispersionModel.AIR,
                            aperture_shape=ghostlight.ApertureShape.IMAGE, aperture_semi_diameter=5.0)
    stop = ghostlight.Surface(semi_aperture=8.0, is_stop=True, disp_model=ghostlight.DispersionModel.AIR)
    lens.surfaces += [s0, stop]
    lens.finalize()
    lens.aperture_images[0].semi_diameter = 5.0
    lens.aperture_images[0].pixels = np.ones((64, 64), np.float32)
    d = _impl._aperture_sdf_debug(lens)
    assert d["stop_index"] == -1 and "sdf" not in d
    empty = _image_stop_system(np.ones((8, 8), np.float32))
    empty.aperture_images[0].pixels = None
    assert _impl._aperture_sdf_debug(empty)["stop_index"] == -1


@pytest.mark.parametrize("controls", [dict(aperture_curvature=-0.8), dict(aperture_curvature=0.6),
                                      dict(aperture_curvature=-0.5, aperture_twist=0.9),
                                      dict(aperture_curvature=-0.9, aperture_twist=-0.8,
                                           aperture_notch_rad=math.radians(35.8),
                                           aperture_notch_angle_rad=math.radians(45.0))])
def test_profile_slope_matches_finite_differences(controls):
    p = _stop(**controls).aperture_profile
    for i in range(60):
        th = 2 * math.pi * (i + 0.31) / 60
        h = 1e-4
        fd = (p.radius_at(th + h) - p.radius_at(th - h)) / (2 * h)
        assert p.dr_dtheta(th) == pytest.approx(fd, abs=0.02, rel=0.05)
    assert _stop(aperture_shape=ghostlight.ApertureShape.CIRCLE).aperture_profile.dr_dtheta(0.3) == 0.0


def test_psf_status_enum_constants_and_vector_helpers():
    from ghostlight_amd import psf
    assert [int(v) for v in ghostlight.PSFCellStatus] == [psf.PSF_CELL_OK, psf.PSF_CELL_CHIEF_VIGNETTED,
                                                          psf.PSF_CELL_DARK] == [0, 1, 2]
    assert ghostlight.MATTE_BOX_OPEN == ghostlight.MatteBox().top == 1.0e9
    a, b = ghostlight.Vec3f(1.0, 2.0, 3.0), ghostlight.Vec3f(-2.0, 0.5, 4.0)
    c = ghostlight.cross(a, b)
    assert ghostlight.dot(a, b) == pytest.approx(11.0)
    assert abs(ghostlight.dot(c, a)) < 1e-6 and abs(ghostlight.dot(c, b)) < 1e-6
    assert (c.x, c.y, c.z) == pytest.approx((6.5, -10.0, 4.5))


def test_gate_params_opening_offsets_and_clamps():
    g = _impl._gate_params_debug(ghostlight.GateConfig(), 12.4, 7.0)
    assert (g["x_pos"], g["x_neg"], g["y_pos"], g["y_neg"]) == pytest.approx((12.4, -12.4, 7.0, -7.0))
    c = ghostlight.GateConfig()
    c.offset_right_mm, c.offset_left_mm, c.offset_top_mm, c.offset_bottom_mm = 2.0, -1.0, 0.5, -0.5
    g = _impl._gate_params_debug(c, 12.4, 7.0)
    assert (g["x_pos"], g["x_neg"], g["y_pos"], g["y_neg"]) == pytest.approx((14.4, -11.4, 7.5, -6.5))
    c = ghostlight.GateConfig()
    c.offset_left_mm = -50.0
    assert _impl._gate_params_debug(c, 12.4, 7.0)["x_neg"] == pytest.approx(-0.1)   # the opening cannot invert
    c = ghostlight.GateConfig()
    c.thickness_mm, c.standoff_mm, c.roughness_rad, c.groove_aniso = -1.0, -3.0, 99.0, 5.0
    c.reflectance_r0, c.scatter_samples, c.max_kick_rad = 7.0, 1000, 9.0
    g = _impl._gate_params_debug(c, 10.0, 10.0)
    assert g["t_mm"] > 0 and g["zs_mm"] == 0.0 and g["sig_wide"] == pytest.approx(0.5) and g["r0"] == 1.0
    assert g["n_scatter"] == 64 and g["inv_scatter"] == pytest.approx(1 / 64) and g["max_kick"] == pytest.approx(1.5)
    c = ghostlight.GateConfig()
    c.lobe = ghostlight.GateLobe.CAUCHY_ISO
    g = _impl._gate_params_debug(c, 10.0, 10.0)
    assert g["sig_tight"] == pytest.approx(g["sig_wide"])


def test_gate_lobe_anisotropy_and_streams():
    c = ghostlight.GateConfig()
    c.roughness_rad, c.groove_aniso, c.groove_angle_deg = 0.1, 0.1, 0.0
    c.lobe = ghostlight.GateLobe.GROOVED
    d = _impl._gate_sample_debug(c, 20000, 12345)
    across = float(np.percentile(np.abs(d["axis"]), 68))
    along = float(np.percentile(np.abs(d["other"]), 68))
    assert across > 3 * along                 # wide Cauchy across the marks, tight Gaussian along them
    assert np.allclose(d["z"], 0.5 * d["axis"], rtol=1e-6, atol=1e-9)   # groove angle 0: z is half the wide draw
    again = _impl._gate_sample_debug(c, 100, 12345)
    assert np.array_equal(again["axis"], d["axis"][:100])
    assert np.all(np.abs(d["axis"]) <= _impl._gate_params_debug(c, 1, 1)["max_kick"] + 1e-6)
