"""The device ghost pipeline against the reference's own point-flare goldens
(all 16 images: two lens files x four source positions x cull on/off) through
render_sources, i.e. gl_render_ghost_pipeline with the pipeline's own pair
filter, probe cull and spectral chain.  Same bar as the CPU oracle's
comparison with the same goldens."""
import numpy as np
import pytest

import golden_flare as gf
from ghostlight_amd import render_sources
from ghostlight_amd.render import source_rgb

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", gf.CASES, ids=lambda c: gf.key(*c))
def test_device_matches_reference_golden(case):
    name, sx, sy, blades, cull = case
    lens = gf.lens(name)
    cfg = gf.config(sx, sy, blades, cull)
    ax, ay = gf.angles(*case)
    rgb = source_rgb(cfg)
    src = np.array([[ax, ay, rgb[0], rgb[1], rgb[2]]], np.float32)
    out = render_sources(src, gf.W, gf.H, lens, gf.calibration(name), cfg)
    img = np.stack([out["ghost_r"], out["ghost_g"], out["ghost_b"]], -1)
    e_sum, e_ds, e_peak = gf.compare(img, *case)
    print(f"{gf.key(*case)}: sum {e_sum:.2e}  8x8 rel-L2 {e_ds:.2e}  peak {e_peak:.2e}", flush=True)
    assert e_sum <= gf.TOL_SUM
    assert e_ds <= gf.TOL_DS
    assert e_peak <= gf.TOL_PEAK
