"""gl_veil_convolve_noalloc (the entry point that allocates nothing) against
gl_veil_convolve on the same device planes, through the C ABI: two calls in a
row, the second on a smaller image, each within 1e-6 rel-L2 of the allocating
entry point (the FP64 kernel-sum reduction may round its last bit
differently) and within 1e-5 of the numpy FP64 convolution; oversized images
are rejected."""
import numpy as np
import pytest

from ghostlight_amd import FlareConfig
from ghostlight_amd import diffraction as dfr
from ghostlight_amd.runtime import Context
from hip_mem import DeviceArray
from oracle import diffraction_oracle as dor

pytestmark = pytest.mark.gpu


def _planes(H, W, seed):
    rng = np.random.default_rng(seed)
    planes = [np.zeros((H, W), np.float32) for _ in range(3)]
    for c in range(3):
        planes[c][rng.integers(0, H, 30), rng.integers(0, W, 30)] = rng.uniform(0.1, 5.0, 30)
        planes[c] += rng.uniform(0.0, 0.01, (H, W)).astype(np.float32)
    return planes


def _run(ctx, planes, hw, hh, cfg, noalloc):
    H, W = planes[0].shape
    dev = [DeviceArray.of(p) for p in planes]
    out = [DeviceArray((H, W), fill=np.nan) for _ in range(3)]
    dfr.veil_convolve_device([d.address for d in dev], [d.address for d in out], W, H, hw, hh, cfg, ctx,
                             noalloc=noalloc)
    ctx.synchronize()
    return [d.download() for d in out]


def test_noalloc_veil_matches_allocating_entry_point():
    ctx = Context()
    cfg = FlareConfig()
    cfg.diffraction.veil_falloff = 1.8
    for (H, W), seed in (((300, 520), 1), ((120, 200), 2)):
        print(f"noalloc veil: {W}x{H}", flush=True)
        planes = _planes(H, W, seed)
        hw, hh = 18.0, 18.0 * H / W
        got = _run(ctx, planes, hw, hh, cfg, noalloc=True)
        want = _run(ctx, planes, hw, hh, cfg, noalloc=False)
        d = cfg.diffraction
        ref = dor.veil_convolve([p.astype(np.float64) for p in planes], 2 * hw / W, 2 * hh / H, d.veil_spread * hh,
                                d.veil_falloff, d.veil_gain)
        e_alloc = dor.rel_l2(np.stack(got), np.stack(want))
        e_ref = dor.rel_l2(np.stack(got), np.stack(ref))
        print(f"noalloc veil {W}x{H}: vs allocating {e_alloc:.2e}, vs FP64 {e_ref:.2e}", flush=True)
        assert np.isfinite(np.stack(got)).all()
        assert e_alloc <= 1e-6
        assert e_ref <= 1e-5


def test_noalloc_veil_rejects_oversized_image():
    ctx = Context()
    # The size check comes before any device work, so placeholder addresses
    # are never dereferenced.
    with pytest.raises(ValueError, match="2048"):
        dfr.veil_convolve_device([1, 1, 1], [1, 1, 1], 1025, 16, 18.0, 1.0, FlareConfig(), ctx, noalloc=True)
