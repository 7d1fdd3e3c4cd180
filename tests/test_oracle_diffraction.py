"""CPU checks that pin the numpy diffraction oracle (no device needed)."""
import math

import numpy as np

from oracle import diffraction_oracle as dor


def test_veil_fft_convolution_equals_direct_sum():
    rng = np.random.default_rng(4)
    img = np.zeros((12, 17))
    img[rng.integers(0, 12, 9), rng.integers(0, 17, 9)] = rng.uniform(0.1, 2.0, 9)
    a = dor.veil_convolve([img], 0.05, 0.04, 0.3, 1.5, 0.03)[0]
    b = dor.veil_convolve_direct(img, 0.05, 0.04, 0.3, 1.5, 0.03)
    assert np.allclose(a, b, rtol=1e-10, atol=1e-14)


def test_veil_kernel_unit_sum_and_point_response():
    g = dor.veil_kernel(64, 0.1, 0.1, 0.5, 2.0)
    assert abs(g.sum() - 1.0) < 1e-12
    # g(r) = (a^2/(r^2+a^2))^p up to the normalisation: ratio of two taps
    r = 3 * 0.1
    assert math.isclose(g[0, 3] / g[0, 0], (0.25 / (r * r + 0.25)) ** 2.0, rel_tol=1e-12)
    # a unit impulse far from the frame edge reproduces the kernel (x gain)
    img = np.zeros((20, 20))
    img[10, 10] = 1.0
    out = dor.veil_convolve([img], 0.1, 0.1, 0.5, 2.0, 1.0)[0]
    assert math.isclose(out[10, 13], g[0, 3], rel_tol=1e-9)


def test_mft_reduces_to_centred_fft():
    rng = np.random.default_rng(1)
    P = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    E = dor.mft(P, 16, 1.0)
    F = np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(P)))
    assert np.allclose(E, F, rtol=1e-10, atol=1e-10)


def test_autocorr_C_is_the_psf_transform():
    """C = conj(FFT(|FFT(A)|^2)) restricted to the lag window, normalised so
    C[0,0] * N^2 = 1 (mdft_render.cu:240-262); for a real pupil the
    autocorrelation is also the direct lag sum of A with itself."""
    N = 32
    A = np.zeros((N, N))
    A[10:20, 12:22] = 1.0
    C, mc = dor.autocorr_C(A, 0.4)
    direct00 = (A * A).sum()
    lag = (2, 1)          # (dy, dx)
    direct = (A[lag[0]:, lag[1]:] * A[:N - lag[0], :N - lag[1]]).sum()
    assert math.isclose(C[mc + lag[0], mc + lag[1]].real / C[mc, mc].real, direct / direct00, rel_tol=1e-9)
    assert math.isclose(C[mc, mc].real * N * N, 1.0, rel_tol=1e-12)


def test_rounded_polygon_area_closed_forms():
    """curvature 0 is the regular polygon (area n/2 sin(2 pi/n) R^2), 1 the
    circle (pi R^2); in between the area lies strictly between them."""
    from ghostlight_amd import synthetic
    fill, N = 0.9, 256
    scale = 4.0 / fill ** 2 / N ** 2          # coverage sum -> area in units of R^2
    for n in (5, 9, 11):
        a0 = synthetic.rounded_polygon_mask(N, n, 0.0).sum() * scale
        a1 = synthetic.rounded_polygon_mask(N, n, 1.0).sum() * scale
        ah = synthetic.rounded_polygon_mask(N, n, 0.5).sum() * scale
        assert math.isclose(a0, n / 2 * math.sin(2 * math.pi / n), rel_tol=1e-3)
        assert math.isclose(a1, math.pi, rel_tol=1e-3)
        assert a0 < ah < a1


def test_precision_pupil_ranges():
    from ghostlight_amd import synthetic
    P = synthetic.precision_pupil(256, seed=3)
    a = np.abs(P)
    assert P.dtype == np.complex64 and a.max() <= 1.0
    core = a[96:160, 96:160]
    assert core.max() <= 1.0 and np.median(core) >= 0.9       # dust transmission 0.9..1.0 inside the stop
    assert np.array_equal(P, synthetic.precision_pupil(256, seed=3))   # seeded


def test_psf_oracle_splat_conserves_weight():
    """Each landed ray deposits its full weight (bilinear weights and the
    normalised Gaussian both sum to one) when its footprint is inside the
    tile: the composite total equals the sum of landed weights x colour."""
    from ghostlight_amd import synthetic
    from oracle import psf_oracle, pyoracle
    lens = synthetic.config1_lens()
    grid = pyoracle.build_samples(lens, 16)
    src = np.array([[0.0, 0.0, 1.0, 1.0, 1.0]], np.float32)
    hits, mx, my = psf_oracle.pupil_probe_mean(lens, 0.0, 0.0)
    assert hits > 0
    chief = np.array([[mx, my]])
    spec = [(550.0, 0.0, 1.0, 0.0)]
    r = psf_oracle._rays(lens, 0.0, 0.0, grid[:, 0], grid[:, 1], 550.0)
    pos, w, st, _ = pyoracle.trace_ghost(lens, -1, -1, r, dbl=False, t_ge0=False, fp64_radius_min=1000.0)
    ok = (st == 0)
    expect = float(w[ok].sum()) / len(grid)
    for sigma, rad in ((0.0, 0), (2.0, 5)):
        out = psf_oracle.psf_splat(lens, src, chief, grid, spec, 512, 512, 1, 1, 0.01, sigma, rad)
        assert math.isclose(out[1].sum(), expect, rel_tol=1e-5)
