"""GPU parity of the diffraction passes against the numpy FP64 oracle (device
calls go through the C ABI).

Bars (float32 device arithmetic against complex128 / float64 references):
  FFT                rel-L2 <= 2e-6 (plus an impulse and a plane wave exact to 1e-6)
  complex GEMM       FP32 MFMA rel-L2 <= 2e-6; 3xBF16 split <= 6e-5; one bf16
                     product term alone would be ~3e-3, so the bar catches a
                     dropped correction term
  matrix-DFT PSF     FP32 <= 2e-5, 3xBF16 <= 2e-4 per wavelength vs FP64
  starburst pupil    bit-exact (same float32 sample positions / half-plane test)
  PSF / sprite / image  rel-L2 <= 1e-4 vs the oracle fed the device's own
                     upstream intermediate
  MDFT window        rel-L2 <= 1e-3 (FP32) / 2e-3 (3xBF16) vs the FP64 formula
  veil reference     rel-L2 <= 1e-5; image <= 1e-4
  survivor tracer    bit-exact status vs the FP32 oracle tracer
"""
import math

import numpy as np
import pytest

from ghostlight_amd import FlareConfig, hostprep, synthetic
from ghostlight_amd import diffraction as dfr
from ghostlight_amd.config import StarburstEngine
from oracle import diffraction_oracle as dor
from oracle import pyoracle

pytestmark = pytest.mark.gpu

W, H = 320, 200


def _crand(rng, *shape):
    return (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)


# ---------------------------------------------------------------- FFT seam


@pytest.mark.parametrize("N", [2, 8, 64, 256, 1024, 2048])
def test_fft2d_matches_numpy(N):
    rng = np.random.default_rng(N)
    a = _crand(rng, N, N)
    got = dfr.fft2d(a)
    ref = np.fft.fft2(a.astype(np.complex128))
    err = dor.rel_l2(got.view(np.float32), ref.astype(np.complex128).view(np.float64))
    print(f"fft N={N}: rel-L2 {err:.3e}", flush=True)
    assert err <= 2e-6


def test_fft2d_closed_forms():
    N = 512
    imp = np.zeros((N, N), np.complex64)
    imp[0, 0] = 1.0
    assert np.allclose(dfr.fft2d(imp), 1.0, atol=1e-6)
    kx, ky = 5, 17
    y, x = np.mgrid[0:N, 0:N]
    wave = np.exp(2j * np.pi * (kx * x + ky * y) / N).astype(np.complex64)
    F = dfr.fft2d(wave)
    peak = np.zeros((N, N))
    peak[ky, kx] = N * N
    assert np.abs(F - peak).max() <= 1e-6 * N * N


def test_fft2d_rejects_non_power_of_two():
    with pytest.raises(ValueError):
        dfr.fft2d(np.zeros((48, 48), np.complex64))


# ---------------------------------------------------------------- MFMA GEMM


@pytest.mark.parametrize("shape", [(64, 64, 64), (100, 70, 50), (257, 129, 33), (31, 300, 97), (1, 1, 1)])
def test_cgemm_fp32(shape):
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N)
    A, B = _crand(rng, M, K), _crand(rng, K, N)
    ref = A.astype(np.complex128) @ B.astype(np.complex128)
    got = dfr.cgemm(A, B, dfr.MDFT_FP32)
    err = dor.rel_l2(got.view(np.float32), ref.view(np.float64))
    print(f"cgemm fp32 {shape}: rel-L2 {err:.3e}", flush=True)
    assert err <= 2e-6


@pytest.mark.parametrize("shape", [(64, 64, 64), (100, 70, 50), (257, 129, 200), (128, 128, 32), (256, 384, 96),
                                   (384, 128, 1024)])
def test_cgemm_bf16x3(shape):
    """Unaligned shapes run the split-on-load kernel, tile-aligned ones the
    pre-split 128 x 128 kernel."""
    M, N, K = shape
    rng = np.random.default_rng(M * 11 + N)
    A, B = _crand(rng, M, K), _crand(rng, K, N)
    ref = A.astype(np.complex128) @ B.astype(np.complex128)
    got = dfr.cgemm(A, B, dfr.MDFT_BF16X3)
    err = dor.rel_l2(got.view(np.float32), ref.view(np.float64))
    print(f"cgemm 3xbf16 {shape}: rel-L2 {err:.3e}", flush=True)
    assert err <= 6e-5


def test_cgemm_structure():
    """Identity and permutation operands pin the output layout of the MFMA
    tiles (row/column placement, real/imaginary planes) exactly."""
    rng = np.random.default_rng(3)
    for M, K in [(96, 96), (130, 130)]:
        A = _crand(rng, M, K)
        perm = rng.permutation(K)
        P = np.zeros((K, K), np.complex64)
        P[np.arange(K), perm] = 1j
        got = dfr.cgemm(A, P, dfr.MDFT_FP32)
        ref = np.zeros_like(A)
        ref[:, perm] = A * 1j
        assert np.array_equal(got, ref)
    # The pre-split 3xBF16 kernel's placement: same product to within the
    # split's 2^-16 relative rounding.
    M = K = 256
    A = _crand(rng, M, K)
    perm = rng.permutation(K)
    P = np.zeros((K, K), np.complex64)
    P[np.arange(K), perm] = 1j
    got = dfr.cgemm(A, P, dfr.MDFT_BF16X3)
    ref = np.zeros_like(A)
    ref[:, perm] = A * 1j
    np.testing.assert_allclose(got, ref, rtol=0, atol=3e-5 * float(np.abs(A).max()))


# ---------------------------------------------------------------- matrix DFT


def _pupil(N, fill=0.5, blades=7):
    nrm = dor.mask_normals(blades, 0.2)
    return dor.build_pupil(N, fill, nrm).astype(np.complex64)


@pytest.mark.parametrize("prec,tol,N,M", [(0, 2e-5, 128, 96), (1, 2e-4, 128, 96), (1, 2e-4, 256, 384)])
def test_mft_psf_matches_fp64(prec, tol, N, M):
    q = 2.0
    P = _pupil(N)
    lams = np.array([430.0, 550.0, 680.0], np.float32)
    got = dfr.mft_psf(P, M, q, lams, precision=prec)
    for li, lam in enumerate(lams):
        E = dor.mft(P, M, q * float(lam) / 550.0)
        ref = np.abs(E) ** 2
        err = dor.rel_l2(got[li], ref)
        print(f"mft prec={prec} lambda={lam}: rel-L2 {err:.3e}", flush=True)
        assert err <= tol


def test_mft_psf_energy_and_fft_agreement():
    """q = 1 and M = N reduces the MFT to the centred FFT; Parseval fixes the total."""
    N = 64
    P = _pupil(N, fill=0.6)
    got = dfr.mft_psf(P, N, 1.0, np.array([550.0], np.float32), precision=0)[0]
    ref = np.abs(np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(P.astype(np.complex128))))) ** 2
    assert dor.rel_l2(got, ref) <= 2e-5
    assert abs(got.sum() / (N * N * (np.abs(P) ** 2).sum()) - 1.0) <= 1e-5


# ---------------------------------------------------------------- starburst


def _star_cfg(engine=StarburstEngine.SPRITE, grid=256, ns=8, blades=6, rot=0.0):
    cfg = FlareConfig()
    cfg.aperture_blades = blades
    cfg.aperture_rotation = rot
    d = cfg.diffraction
    d.starburst = True
    d.starburst_engine = engine
    d.starburst_grid = grid
    d.starburst_grid_cap = grid
    d.spectral_samples = ns
    d.use_survivor_mask = False
    return cfg


def _star_spec(ns, output_cs=0):
    t = hostprep.spectral_table(ns, output_cs).astype(np.float64)
    t[:, 1:] *= (700.0 - 400.0) / ns
    t[:, 1:] /= t[:, 1:].sum(0)
    return [tuple(r) for r in t]


@pytest.fixture(scope="module")
def c1():
    lens = synthetic.config1_lens()
    return lens, synthetic.calibration_for(lens, W, H)


# A 1.2 x 0.75 mm window keeps the diffraction pattern tens of pixels wide.
ZOOM_HW, ZOOM_HH = 0.6, 0.375


@pytest.mark.parametrize("rot,zoom", [(0.0, True), (15.0, True), (0.0, False)])
def test_starburst_sprite_chain(c1, rot, zoom):
    lens, cal = c1
    cfg = _star_cfg(rot=rot)
    src = np.array([[0.0, 0.0, 1.0, 0.8, 0.6], [0.0, 0.0, 0.3, 0.3, 0.9]], np.float32)
    px = np.array([W * 0.5, W * 0.25], np.float32)
    py = np.array([H * 0.5, H * 0.7], np.float32)
    hw, hh = (ZOOM_HW, ZOOM_HH) if zoom else (cal.sensor_half_w, cal.sensor_half_h)
    planes, dbg = dfr.render_starburst(lens, cal, src, px, py, W, H, hw, hh, cfg, debug=True)
    N = dbg["grid"]
    assert N == 256
    # pupil: same float32 sample positions and half-plane test -> bit-exact
    rot_rad = float(np.float32(rot) * np.float32(math.pi / 180.0))
    nrm = dor.mask_normals(6, rot_rad)
    fill = cfg.diffraction.pupil_fill
    pup = dor.build_pupil(N, fill, nrm)
    assert np.array_equal(dbg["pupil"].astype(np.float64), pup)
    # PSF from the device pupil (FP64 FFT, unit sum)
    psf_ref = dor.psf_from_pupil(dbg["pupil"])
    e_psf = dor.rel_l2(dbg["psf"], psf_ref)
    assert e_psf <= 1e-4, e_psf
    assert abs(float(dbg["psf"].astype(np.float64).sum()) - 1.0) <= 1e-5
    # lambda bake from the device PSF
    spr_ref = dor.bake_sprite(dbg["psf"].astype(np.float64), _star_spec(8))
    e_spr = dor.rel_l2(dbg["sprite_rgb"], spr_ref)
    assert e_spr <= 1e-4, e_spr
    # downsample + area resample of the device sprite
    px_mm_x, px_mm_y = 2 * hw / W, 2 * hh / H
    sx, sy = dbg["dx_mm_x"], dbg["dx_mm_y"]
    ds = dor.po2_floor(min(px_mm_x / (sx * 3), px_mm_y / (sy * 3)), N // 8)
    spr_ds = dor.downsample(dbg["sprite_rgb"].astype(np.float64), ds)
    fnx, fny = dbg["f_number_x"], dbg["f_number_y"]
    gain = cfg.diffraction.starburst_gain * dbg["throughput"] * (2.8 * 2.8) * cal.pupil_area_frac / max(fnx * fny, 1e-6)
    srcs = [(float(px[i]), float(py[i]), *map(float, src[i, 2:])) for i in range(len(src))]
    img_ref = dor.resample(spr_ds, sx * ds, sy * ds, px_mm_x, px_mm_y, srcs, gain, W, H)
    e_img = dor.rel_l2(np.stack(planes), img_ref)
    print(f"starburst rot={rot}: psf {e_psf:.2e} sprite {e_spr:.2e} image {e_img:.2e} ds={ds}", flush=True)
    assert e_img <= 1e-4, e_img


def test_starburst_zero_sources_and_empty(c1):
    lens, cal = c1
    cfg = _star_cfg()
    planes = dfr.render_starburst(lens, cal, np.zeros((0, 5), np.float32), [], [], W, H, cal.sensor_half_w,
                                  cal.sensor_half_h, cfg)
    assert all(not p.any() for p in planes)


@pytest.mark.parametrize("prec,tol", [(0, 1e-3), (1, 2e-3)])
def test_starburst_mdft_matches_formula(c1, prec, tol):
    lens, cal = c1
    cfg = _star_cfg(engine=StarburstEngine.MDFT)
    hw, hh = ZOOM_HW, ZOOM_HH
    src = np.array([[0.0, 0.0, 1.0, 0.9, 0.7]], np.float32)
    px = np.array([W * 0.4], np.float32)
    py = np.array([H * 0.55], np.float32)
    planes, dbg = dfr.render_starburst(lens, cal, src, px, py, W, H, hw, hh, cfg, debug=True, mdft_precision=prec)
    N = dbg["grid"]
    fill = cfg.diffraction.pupil_fill
    Cm, mc = dor.autocorr_C(dbg["pupil"], fill)
    dx, dy = dbg["dx_mm_x"], dbg["dx_mm_y"]
    px_mm_x, px_mm_y = 2 * hw / W, 2 * hh / H
    fnx, fny = dbg["f_number_x"], dbg["f_number_y"]
    gain = cfg.diffraction.starburst_gain * dbg["throughput"] * (2.8 * 2.8) * cal.pupil_area_frac / max(fnx * fny, 1e-6)
    kn = (px_mm_x * px_mm_y) / (dx * dy) * gain
    ref = dor.mdft_window(Cm, mc, _star_spec(8), N * dx, N * dy, px_mm_x, px_mm_y,
                          (float(px[0]), float(py[0]), 1.0, 0.9, 0.7), W, H, kn)
    assert (ref[0] > 0).sum() > 400          # the window spans many pixels
    err = dor.rel_l2(np.stack(planes), ref)
    print(f"mdft prec={prec}: N={N} mc={mc} rel-L2 {err:.3e}", flush=True)
    assert err <= tol


# ---------------------------------------------------------------- veil


@pytest.mark.parametrize("falloff", [1.0, 1.5, 2.5])
def test_veil_reference_and_image(falloff):
    cfg = FlareConfig()
    cfg.diffraction.veil = True
    cfg.diffraction.veil_falloff = falloff
    hw, hh = 18.0, 12.0
    src = np.array([[0.0, 0.0, 1.0, 0.5, 0.25], [0.0, 0.0, 0.2, 0.2, 0.2]], np.float32)
    px = np.array([W * 0.5, W * 0.9], np.float32)
    py = np.array([H * 0.5, H * 0.1], np.float32)
    planes, dbg = dfr.render_veil(src, px, py, W, H, hw, hh, cfg, debug=True)
    N = dbg["grid"]
    ref = dor.veil_reference(N, dbg["core_texels"], falloff)
    e_ref = dor.rel_l2(dbg["reference"], ref)
    assert e_ref <= 1e-5, e_ref
    px_mm_x, px_mm_y = 2 * hw / W, 2 * hh / H
    dxf = dbg["dx_mm_x"]
    ds = dor.po2_floor(min(px_mm_x / (dxf * 3), px_mm_y / (dxf * 3)), N // 8)
    sp = dor.downsample(dbg["reference"].astype(np.float64), ds)
    srcs = [(float(px[i]), float(py[i]), *map(float, src[i, 2:])) for i in range(len(src))]
    img = dor.resample(sp, dxf * ds, dxf * ds, px_mm_x, px_mm_y, srcs, cfg.diffraction.veil_gain, W, H)
    e_img = dor.rel_l2(np.stack(planes), img)
    print(f"veil p={falloff}: reference {e_ref:.2e} image {e_img:.2e}", flush=True)
    assert e_img <= 1e-4


def test_veil_linearity():
    cfg = FlareConfig()
    cfg.diffraction.veil = True
    src = np.array([[0.0, 0.0, 1.0, 1.0, 1.0]], np.float32)
    px, py = np.array([100.0], np.float32), np.array([80.0], np.float32)
    a = np.stack(dfr.render_veil(src, px, py, W, H, 18.0, 12.0, cfg))
    b = np.stack(dfr.render_veil(src * np.array([1, 1, 3, 3, 3], np.float32), px, py, W, H, 18.0, 12.0, cfg))
    assert np.allclose(b, 3 * a, rtol=1e-6, atol=0)


# ---------------------------------------------------------------- survivor tracer


@pytest.mark.parametrize("ang_deg", [0.0, 9.0, 18.0])
def test_primary_survival_bitexact(c1, ang_deg):
    lens, _ = c1
    rng = np.random.default_rng(int(ang_deg * 10) + 1)
    fR = lens.surfaces[0].semi_aperture
    xy = (rng.uniform(-1.2, 1.2, (4000, 2)) * fR).astype(np.float32)
    a = math.radians(ang_deg)
    d = np.array([math.tan(a), 0.3 * math.tan(a), 1.0])
    d = (d / np.linalg.norm(d)).astype(np.float32)
    z0 = np.float32(lens.surfaces[0].z - 20.0)
    got = dfr.primary_survival(lens, xy, d, float(z0), 550.0)
    rays = np.zeros((len(xy), 7), np.float32)
    rays[:, :2] = xy
    rays[:, 2] = z0
    rays[:, 3:6] = d
    rays[:, 6] = 550.0
    pos, _w, st, _sg = pyoracle.trace_ghost(lens, -1, -1, rays, dbl=False, t_ge0=True, fp64_radius_min=0.0)
    ok = (st == 0) & np.isfinite(pos[:, 0]) & np.isfinite(pos[:, 1])
    print(f"survival {ang_deg} deg: {int(ok.sum())} of {len(ok)}", flush=True)
    assert np.array_equal(got, ok)


def test_engine_gates(c1):
    """Engine gates of the reference's MDFT validation
    (validation/validate_starburst_mdft.py:99-108): SPRITE renders are
    bit-equal run to run, MDFT / SPRITE flux within 5%, MDFT flux invariant
    to the spectral sample count (spread < 0.5%), MDFT window non-negative."""
    lens, cal = c1
    src = np.array([[0.0, 0.0, 1.0, 1.0, 1.0]], np.float32)
    px, py = np.array([W * 0.5], np.float32), np.array([H * 0.5], np.float32)

    def run(engine, ns=12, prec=0):
        cfg = _star_cfg(engine=engine, ns=ns)
        return np.stack(dfr.render_starburst(lens, cal, src, px, py, W, H, ZOOM_HW, ZOOM_HH, cfg, mdft_precision=prec))

    a, b = run(StarburstEngine.SPRITE), run(StarburstEngine.SPRITE)
    assert np.array_equal(a, b)
    sp = a.astype(np.float64).sum()
    md = [run(StarburstEngine.MDFT, ns).astype(np.float64).sum() for ns in (3, 6, 12, 24)]
    ratio = md[2] / sp
    spread = max(md) / min(md)
    print(f"flux mdft/sprite {ratio:.4f}, ns spread {spread:.5f}", flush=True)
    assert abs(ratio - 1.0) < 0.05
    assert spread < 1.005
    m = run(StarburstEngine.MDFT, 12, 1)
    assert np.isfinite(m).all() and m.min() >= -1e-4 * m.max()


@pytest.mark.parametrize("shape,falloff", [((270, 480), 1.5), ((200, 200), 2.5)])
def test_veil_fft_convolution_matches_oracle(shape, falloff):
    H, W = shape
    rng = np.random.default_rng(H)
    planes = [np.zeros((H, W), np.float32) for _ in range(3)]
    for c in range(3):
        planes[c][rng.integers(0, H, 40), rng.integers(0, W, 40)] = rng.uniform(0.1, 5.0, 40)
        planes[c] += rng.uniform(0.0, 0.01, (H, W)).astype(np.float32)
    cfg = FlareConfig()
    cfg.diffraction.veil_falloff = falloff
    hw, hh = 18.0, 18.0 * H / W
    out, ms = dfr.veil_convolve(planes, hw, hh, cfg, timing=True)
    d = cfg.diffraction
    ref = dor.veil_convolve([p.astype(np.float64) for p in planes], 2 * hw / W, 2 * hh / H, d.veil_spread * hh,
                            falloff, d.veil_gain)
    err = dor.rel_l2(np.stack(out), np.stack(ref))
    print(f"veil FFT conv {W}x{H} p={falloff}: rel-L2 {err:.2e}, {ms:.3f} ms", flush=True)
    assert err <= 1e-5
    # energy: a unit-sum kernel keeps gain x flux up to what leaves the frame
    assert float(np.stack(out).sum()) <= d.veil_gain * float(np.stack(planes).sum()) * (1 + 1e-5)


@pytest.mark.parametrize("prec,tol", [(0, 2e-5), (1, 2e-4)])
def test_precision_pupil_mft(prec, tol):
    from ghostlight_amd import synthetic
    P = synthetic.precision_pupil(256, seed=3)
    M, q = 512, 2.0
    lams = np.array([400.0, 550.0, 700.0], np.float32)
    got = dfr.mft_psf(P, M, q, lams, precision=prec)
    for li, lam in enumerate(lams):
        ref = np.abs(dor.mft(P, M, q * float(lam) / 550.0)) ** 2
        err = dor.rel_l2(got[li], ref)
        print(f"precision pupil MFT prec={prec} lambda={lam}: rel-L2 vs FP64 {err:.2e}", flush=True)
        assert err <= tol


@pytest.mark.parametrize("sigma_um,mono", [(0.0, False), (4.0, False), (4.0, True)])
def test_psf_grid_matches_oracle(c1, sigma_um, mono):
    from ghostlight_amd import PSFConfig, hostprep
    from ghostlight_amd.psf import render_psf
    from oracle import psf_oracle
    lens, _cal = c1
    cfg = PSFConfig(grid_nx=3, grid_ny=2, tile_w=48, tile_h=40, tile_extent_mm=0.06, ray_grid=32, spectral_samples=6,
                    splat_sigma_um=sigma_um, monochromatic=mono)
    ang = np.radians([[0.0, 0.0], [5.0, 0.0], [10.0, 2.0], [-7.0, 4.0], [12.0, -6.0]])
    src = np.concatenate([ang, np.array([[1.0, 0.8, 0.6]] * len(ang))], 1).astype(np.float32)
    out = render_psf(src, lens, cfg)
    assert out["r"].shape == (80, 144) and out["status"].dtype == np.uint8
    # chief pre-pass (7 x 7 pupil mean, CPU t >= 0 rule)
    for i in range(len(src)):
        hits, mx, my = psf_oracle.pupil_probe_mean(lens, src[i, 0], src[i, 1])
        assert (out["status"][i] == 0) == (hits > 0)
        assert abs(out["chief_x_mm"][i] - mx) <= 1e-5 and abs(out["chief_y_mm"][i] - my) <= 1e-5
    grid = pyoracle.build_samples(lens, 32)
    if mono:
        spec = [(587.56, 0.0, 1.0, 0.0)]
    else:
        t = hostprep.spectral_table(6).astype(np.float64)
        t[:, 1:] *= 300.0 / 6
        spec = [tuple(r) for r in t]
    mmpp = 0.06 / 48
    sig = sigma_um / (mmpp * 1000.0)
    rad = min(16, max(1, math.ceil(2.5 * sig))) if sig > 0 else 0
    chief = np.stack([out["chief_x_mm"], out["chief_y_mm"]], 1).astype(np.float64)
    ref = psf_oracle.psf_splat(lens, src, chief, grid, spec, 48, 40, 3, 2, mmpp, sig, rad, mono=mono)
    got = np.stack([out["r"], out["g"], out["b"]])
    err = dor.rel_l2(got, ref)
    print(f"PSF grid sigma={sigma_um} mono={mono}: rel-L2 vs oracle {err:.2e}", flush=True)
    assert err <= 1e-4
    if mono:
        assert not out["g"].any() and not out["b"].any()


def test_psf_fixed_target_aims_the_chief(c1):
    from ghostlight_amd import PSFConfig
    from ghostlight_amd.psf import render_psf
    lens, cal = c1
    f = lens.focal_length
    cfg = PSFConfig(grid_nx=2, grid_ny=2, tile_w=32, tile_h=32, ray_grid=24, spectral_samples=4,
                    center_mode=1, splat_sigma_um=0.0)
    targets = np.array([[0.0, 0.0], [1.5, 0.0], [-2.0, 1.0], [3.0, -2.5]])
    seeds = np.arctan(targets / f)               # first-order seed
    src = np.concatenate([seeds, np.ones((4, 3)), targets], 1).astype(np.float32)
    out = render_psf(src, lens, cfg)
    ok = out["status"] == 0
    print(f"aim residuals (mm) {out['aim_residual_mm']}, status {out['status']}", flush=True)
    assert ok.all()
    assert np.all(out["aim_residual_mm"][ok] < 1e-3)
    assert np.allclose(out["chief_x_mm"], targets[:, 0], atol=1e-3)
    assert np.allclose(out["chief_y_mm"], targets[:, 1], atol=1e-3)
