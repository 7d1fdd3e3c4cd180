"""Device calibration and source map against the reference's CPU code
(fixture from calibrate_lens / solve_source_angle / source_map_landing on 17
of the reference's lens files).  Every probe ray is traced on the device; the
searches and the survivor sums run in the library's host code in the
reference's order, so round-stop lenses match to the bit.  Polygonal stops
differ by the pupil-mask form (see source_map_fixture.POLYGON_STOPS) and are
held to a tolerance instead."""
import numpy as np
import pytest

import golden_flare as gf
import source_map_fixture as smf
from ghostlight_amd import LensCalibration, calibrate_lens, render_point_flare, solve_source_map
from ghostlight_amd.source_map import _source_map_landing

pytestmark = pytest.mark.gpu

# Polygon-stop tolerance: relative on calibration values, absolute on angles
# (rad) and landings (mm).
TOL_POLY_CAL = 2e-3
TOL_POLY_ANGLE = 2e-4
TOL_POLY_MM = 5e-3


@pytest.mark.parametrize("name", smf.LENSES)
def test_calibrate_lens(name):
    z = smf.fixture()
    want = z[f"sm_{name}_calib"]
    cal = calibrate_lens(smf.lens(name))
    got = np.array([getattr(cal, f) for f in LensCalibration.FIELDS], np.float32)
    rel = np.abs(got - want) / np.maximum(np.abs(want), 1e-6)
    print(f"{name}: calibration max rel diff {rel.max():.2e}", flush=True)
    if name in smf.POLYGON_STOPS:
        assert rel.max() <= TOL_POLY_CAL
    else:
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("name", smf.LENSES)
def test_solve_source_angle(name):
    z = smf.fixture()
    lens, cal = smf.lens(name), smf.calibration(name)
    for req, want in zip(z[f"sm_{name}_requests"], z[f"sm_{name}_solves"]):
        sm = solve_source_map(lens, cal, float(req[0]), float(req[1]), float(req[2]), float(req[3]))
        r = sm._raw
        got = np.array([r.angle_x, r.angle_y, r.status, r.residual_mm, r.anchor_ax, r.anchor_ay, r.anchor_x,
                        r.anchor_y, *r.J, r.eff_half_w, r.eff_half_h, r.inv_tan_w, r.inv_tan_h, r.probe_evals],
                       np.float32)
        d_ang = float(np.abs(got[:2] - want[:2]).max())
        print(f"{name} ndc ({req[0]:+.2f},{req[1]:+.2f}): status {int(got[2])}/{int(want[2])} "
              f"angle diff {d_ang:.2e} evals {int(got[16])}/{int(want[16])}", flush=True)
        if name in smf.POLYGON_STOPS:
            assert int(got[2]) == int(want[2])
            assert d_ang <= TOL_POLY_ANGLE
        else:
            np.testing.assert_array_equal(got, want, err_msg=f"{name} {req}")


@pytest.mark.parametrize("name", smf.LENSES)
def test_source_map_landing(name):
    z = smf.fixture()
    lens, cal = smf.lens(name), smf.calibration(name)
    for (ax, ay), (ok, x, y) in zip(z[f"sm_{name}_landing_angles"], z[f"sm_{name}_landings"]):
        got = _source_map_landing(lens, cal, float(ax), float(ay))
        assert (got is not None) == bool(ok), f"{name} ({ax}, {ay})"
        if got is None:
            continue
        if name in smf.POLYGON_STOPS:
            assert abs(got[0] - x) <= TOL_POLY_MM and abs(got[1] - y) <= TOL_POLY_MM
        else:
            assert (np.float32(got[0]), np.float32(got[1])) == (x, y)


@pytest.mark.parametrize("case", [c for c in gf.CASES if c[4] == 1], ids=lambda c: gf.key(*c))
def test_point_flare_end_to_end(case):
    """The reference's point-flare goldens from source_x/source_y: the
    device solves the source angle from the calibration, then renders."""
    name, sx, sy, blades, cull = case
    lens, cal = gf.lens(name), gf.calibration(name)
    cfg = gf.config(sx, sy, blades, cull)
    sm = solve_source_map(lens, cal, (sx - 0.5) * 2.0, (sy - 0.5) * 2.0, gf.SH, gf.SH)
    want = gf.angles(*case)
    assert (np.float32(sm.angle_x), np.float32(sm.angle_y)) == (want[0], want[1])
    out = render_point_flare(gf.W, gf.H, lens, cal, cfg)
    img = np.stack([out["ghost_r"], out["ghost_g"], out["ghost_b"]], -1)
    e_sum, e_ds, e_peak = gf.compare(img, *case)
    print(f"{gf.key(*case)} end to end: sum {e_sum:.2e}  8x8 rel-L2 {e_ds:.2e}  peak {e_peak:.2e}", flush=True)
    assert e_sum <= gf.TOL_SUM and e_ds <= gf.TOL_DS and e_peak <= gf.TOL_PEAK
