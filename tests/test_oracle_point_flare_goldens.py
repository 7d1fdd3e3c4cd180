"""The CPU oracle's ghost render against the reference's own point-flare
goldens: pins the whole render restatement (pupil sampler, filter, probe cull,
tracer, coatings, spectral chain, splat) at image level on two real lens
files, including the anamorphic cylinder lens and the 6-blade polygon stop."""
import numpy as np
import pytest

import golden_flare as gf
from ghostlight_amd import hostprep
from ghostlight_amd.config import to_params
from ghostlight_amd.render import source_rgb
from oracle import pyoracle

CPU_CASES = [("double_gauss", 0.5, 0.5, 0, 1), ("double_gauss", 1.6, 0.9, 0, 0),
             ("sirui", 1.3, 0.5, 6, 1), ("sirui", 1.6, 0.9, 0, 1)]


def oracle_render(name, sx, sy, blades, cull):
    lens = gf.lens(name)
    cfg = gf.config(sx, sy, blades, cull)
    ax, ay = gf.angles(name, sx, sy, blades, cull)
    rgb = source_rgb(cfg)
    src = np.array([[ax, ay, rgb[0], rgb[1], rgb[2]]], np.float32)
    pairs, boosts = pyoracle.filter_ghost_pairs(lens, gf.SH, gf.SH, min_intensity=0.0)
    spec = hostprep.spectral_table(16)
    p, _ = to_params(cfg)
    planes, _ = pyoracle.render_ghosts(lens, pairs, boosts, src, spec, p, gf.SH, gf.SH, gf.W, gf.H,
                                       correction=hostprep.spectral_correction(spec), threads=0)
    return np.stack(planes, -1)


def test_fixture_invariants():
    """Cull on and off render the same flux (validate_culling.py:100-114), and
    the 6-blade stop passes less light than the round one at the same field."""
    z = gf.fixture()
    for name in gf.LENSES:
        for sx, sy, b in gf.CONFIGS:
            s0 = z[f"pf_{gf.key(name, sx, sy, b, 0)}_sum"]
            s1 = z[f"pf_{gf.key(name, sx, sy, b, 1)}_sum"]
            np.testing.assert_allclose(s0, s1, rtol=5e-6)
        lens = gf.lens(name)
        assert lens.num_surfaces() == {"double_gauss": 11, "sirui": 23}[name]


@pytest.mark.parametrize("case", CPU_CASES, ids=lambda c: gf.key(*c))
def test_oracle_matches_reference_golden(case):
    img = oracle_render(*case)
    e_sum, e_ds, e_peak = gf.compare(img, *case)
    print(f"{gf.key(*case)}: sum {e_sum:.2e}  8x8 rel-L2 {e_ds:.2e}  peak {e_peak:.2e}")
    assert e_sum <= gf.TOL_SUM
    assert e_ds <= gf.TOL_DS
    assert e_peak <= gf.TOL_PEAK
