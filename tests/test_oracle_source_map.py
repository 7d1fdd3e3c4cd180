"""Host-side checks of the calibration / source-map fixture and of the
source map's forward arithmetic (gl_source_map_screen, no device work)."""
import ctypes as C
import math

import numpy as np
import pytest

import source_map_fixture as smf
from ghostlight_amd import _abi


def test_reference_calibration_matches_its_manifest():
    """The reference build's calibrate_lens reproduces the calibration the
    reference recorded with its aperture goldens (manifest.json)."""
    z = smf.fixture()
    for key, name in (("double_gauss", "double_gauss"), ("sirui", "sirui"), ("hex_stop", "hex_stop"),
                      ("hex_rot30", "hex_rot30"), ("blade_shape", "blade_shape")):
        np.testing.assert_array_equal(z[f"sm_{name}_calib"], z[f"manifest_{key}"], err_msg=name)


def test_fixture_invariants():
    """SOLVED entries anchor at their own angle and land within the accept
    bar; f-number = focal length / (2 entrance-pupil semi)."""
    z = smf.fixture()
    for name in smf.LENSES:
        cal = z[f"sm_{name}_calib"]
        for a in range(2):
            if cal[8 + a] > 0:
                assert cal[10 + a] == np.float32(cal[6 + a] / (np.float32(2.0) * cal[8 + a]))
        for req, r in zip(z[f"sm_{name}_requests"], z[f"sm_{name}_solves"]):
            if int(r[2]) == 0:
                assert r[4] == r[0] and r[5] == r[1]
                accept = max(1e-3, 0.01 * min(abs(req[2]), abs(req[3])))
                assert r[3] < accept


def _raw(r):
    s = _abi.gl_source_map()
    s.angle_x, s.angle_y, s.status, s.residual_mm = float(r[0]), float(r[1]), int(r[2]), float(r[3])
    s.anchor_ax, s.anchor_ay, s.anchor_x, s.anchor_y = (float(x) for x in r[4:8])
    for i in range(4):
        s.J[i] = float(r[8 + i])
    s.eff_half_w, s.eff_half_h, s.inv_tan_w, s.inv_tan_h = (float(x) for x in r[12:16])
    s.probe_evals = int(r[16])
    return s


@pytest.mark.parametrize("name", smf.LENSES)
def test_screen_map_bitexact(name):
    """source_map_screen of every reference solve at its angle and at two
    offsets, through the product's C ABI, equals the reference's."""
    lib = _abi.load_library()
    z = smf.fixture()
    offs = z[f"sm_{name}_screen_offsets"]
    for r, want in zip(z[f"sm_{name}_solves"], z[f"sm_{name}_screen"]):
        s = _raw(r)
        for (dx, dy), w in zip(offs, want):
            sx, sy = C.c_float(), C.c_float()
            lib.gl_source_map_screen(C.byref(s), float(np.float32(r[0] + dx)), float(np.float32(r[1] + dy)),
                                     C.byref(sx), C.byref(sy))
            assert (np.float32(sx.value), np.float32(sy.value)) == (w[0], w[1])


def test_screen_map_closed_forms():
    """FALLBACK is the constant-focal-length map; an anchored map is exact at
    its anchor and saturates its extension at 1.55 rad."""
    lib = _abi.load_library()
    s = _abi.gl_source_map()
    s.status = 2
    s.inv_tan_w, s.inv_tan_h = 2.0, 4.0
    sx, sy = C.c_float(), C.c_float()
    lib.gl_source_map_screen(C.byref(s), 0.1, -0.05, C.byref(sx), C.byref(sy))
    assert abs(sx.value - (0.5 + 0.5 * math.tan(0.1) * 2.0)) < 1e-6
    assert abs(sy.value - (0.5 + 0.5 * math.tan(-0.05) * 4.0)) < 1e-6
    s.status = 0
    s.anchor_ax, s.anchor_ay, s.anchor_x, s.anchor_y = 0.2, -0.1, 3.0, -1.5
    s.J[0], s.J[1], s.J[2], s.J[3] = 20.0, 1.0, -0.5, 18.0
    s.eff_half_w, s.eff_half_h = 12.0, 9.0
    lib.gl_source_map_screen(C.byref(s), 0.2, -0.1, C.byref(sx), C.byref(sy))
    assert sx.value == np.float32(0.5 + 0.5 * 3.0 / 12.0) and sy.value == np.float32(0.5 + 0.5 * -1.5 / 9.0)
    far = []
    for d in (2.0, 5.0):
        lib.gl_source_map_screen(C.byref(s), 0.2 + d, -0.1, C.byref(sx), C.byref(sy))
        far.append(sx.value)
    assert far[0] == far[1]
