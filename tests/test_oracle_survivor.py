"""The oracle's restatement of the starburst's off-axis pupil
(solve_survivor_center / build_survivor_envelope, starburst_render.cu:393-582)
pinned by a closed form: on a lens of two flat air-air apertures (front
radius R1 at z1, stop radius R2 at z2) a spawn-plane launch (x, y) heading
(tan a, 0, 1) survives iff it clears both discs, so the survivors are the
intersection of the discs centred at (-(z_k - z0) tan a, 0) with radius R_k
(the cat's eye of validate_starburst_pupil.py's two-circle model, exact
here).  Checked: the survivor centroid against an FP64 integral of that
region, envelope cells wholly inside / outside the region are exactly 1 / 0,
supersampled boundary cells against their exact coverage, and the dilation
band outside the stop."""
import math

import numpy as np
import pytest

from ghostlight_amd import OpticalSystem, Surface
from oracle import survivor_oracle as so

R1, R2 = 12.0, 8.0
T1, T2 = 40.0, 40.0          # front -> stop, stop -> sensor


def air_lens():
    lens = OpticalSystem([Surface(radius=0.0, thickness=T1, ior=1.0, semi_aperture=R1),
                          Surface(radius=0.0, thickness=T2, ior=1.0, semi_aperture=R2, is_stop=True)],
                         name="two-aperture air lens")
    lens.finalize()
    return lens


def discs(lens, ax):
    z0 = lens.surfaces[0].z - 20.0
    t = math.tan(ax)
    return [(-(s.z - z0) * t, 0.0, s.semi_aperture) for s in lens.surfaces]


def inside_both(x, y, dd):
    ok = np.ones(np.broadcast(x, y).shape, bool)
    for cx, cy, r in dd:
        ok &= (x - cx) ** 2 + (y - cy) ** 2 <= r * r
    return ok


def exact_centroid(dd, n=3000):
    lo = max(cx - r for cx, _, r in dd)
    hi = min(cx + r for cx, _, r in dd)
    rr = min(r for _, _, r in dd)
    xs = np.linspace(lo, hi, n)
    ys = np.linspace(-rr, rr, n)
    X, Y = np.meshgrid(xs, ys)
    m = inside_both(X, Y, dd)
    return float(X[m].mean()), float(Y[m].mean())


@pytest.mark.parametrize("deg", [5.0, 10.0, 14.0])
def test_survivor_center_is_the_two_disc_centroid(deg):
    lens = air_lens()
    ax = math.radians(deg)
    cx, cy, any_ = so.survivor_center(lens, ax, 0.0)
    assert any_
    ex, ey = exact_centroid(discs(lens, ax))
    front = lens.surfaces[0].semi_aperture
    cell = (1.1 * front + 20.0 * math.tan(ax)) / 80.0          # launch-grid spacing
    print(f"{deg} deg: centroid ({cx:.5f}, {cy:.5f}) exact ({ex:.5f}, {ey:.5f}), cell {cell:.4f} mm")
    assert abs(cx - ex) < 0.02 * cell
    assert abs(cy - ey) < 0.02 * cell


@pytest.mark.parametrize("deg", [10.0, 14.0])
def test_survivor_envelope_is_the_two_disc_coverage(deg):
    lens = air_lens()
    ax = math.radians(deg)
    cx, cy, _ = so.survivor_center(lens, ax, 0.0)
    ME, ep = 96, R2
    env = so.survivor_envelope(lens, None, ep, ep, cx, cy, ax, 0.0, ME)
    dd = discs(lens, ax)
    # exact coverage of every cell, 40 x 40 FP64 sub-grid per cell (cells
    # inside the unit stop disc only: the stop clip is the analytic stop's)
    c = (np.arange(ME) + 0.5) / ME * 2 - 1
    du = 2.0 / ME
    sub = (np.arange(40) + 0.5) / 40 - 0.5
    U, V = np.meshgrid(c, c)
    inside_stop = U * U + V * V <= 1.0
    cover = np.zeros((ME, ME))
    full_in = np.zeros((ME, ME), bool)
    full_out = np.zeros((ME, ME), bool)
    for j in range(ME):
        for i in range(ME):
            if not inside_stop[j, i]:
                continue
            uu = c[i] + du * sub
            vv = c[j] + du * sub
            GU, GV = np.meshgrid(uu, vv)
            st = GU * GU + GV * GV <= 1.0
            m = inside_both(cx + GU * ep, cy + GV * ep, dd)
            cover[j, i] = (m & st).sum() / max(1, st.sum())
            full_in[j, i] = m.all()
            full_out[j, i] = not m.any()
    e = env[inside_stop]
    assert np.array_equal(env[full_in & inside_stop], np.ones(int((full_in & inside_stop).sum()), np.float32))
    assert not env[full_out & inside_stop].any()
    # The reference supersamples a cell only when an in-stop 8-neighbour's
    # centre survives differently from its own; other cells keep the exact
    # survival of their centre.
    ctr = inside_both(cx + U * ep, cy + V * ep, dd)
    pad_s, pad_v = np.pad(ctr, 1), np.pad(inside_stop, 1)
    flagged = np.zeros((ME, ME), bool)
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            if di or dj:
                flagged |= pad_v[1 + dj:1 + dj + ME, 1 + di:1 + di + ME] & \
                    (pad_s[1 + dj:1 + dj + ME, 1 + di:1 + di + ME] != ctr)
    flagged &= inside_stop
    near = np.zeros((ME, ME), bool)       # centre within 1e-4 mm of a disc edge: float rounding decides
    for dcx, dcy, r in dd:
        near |= np.abs(np.hypot(cx + U * ep - dcx, cy + V * ep - dcy) - r) < 1e-4
    err = np.abs(env[flagged] - cover[flagged])
    plain = inside_stop & ~flagged & ~near
    print(f"{deg} deg: {int(flagged.sum())} supersampled cells, coverage error max {err.max():.3f} "
          f"mean {err.mean():.4f}; envelope mean {e.mean():.4f} vs exact {cover[inside_stop].mean():.4f}")
    assert flagged.any()
    assert err.max() <= 0.08
    assert np.array_equal(env[plain], ctr[plain].astype(np.float32))
    assert abs(float(e.mean()) - float(cover[inside_stop].mean())) < 5e-3
    # outside the stop: the largest inside value within 2 cells, else 0
    for j, i in np.argwhere(~inside_stop):
        win = env[max(0, j - 2):j + 3, max(0, i - 2):i + 3]
        msk = inside_stop[max(0, j - 2):j + 3, max(0, i - 2):i + 3]
        want = float(win[msk].max()) if msk.any() else 0.0
        assert env[j, i] == np.float32(max(0.0, want))


def test_envelope_on_axis_disc_inside_disc_is_full():
    """At 5 deg the stop's disc lies inside the front's: every cell survives."""
    lens = air_lens()
    ax = math.radians(5.0)
    (c1, _, r1), (c2, _, r2) = discs(lens, ax)
    assert abs(c1 - c2) + r2 < r1
    cx, cy, _ = so.survivor_center(lens, ax, 0.0)
    env = so.survivor_envelope(lens, None, R2 * 0.9, R2 * 0.9, c2, 0.0, ax, 0.0, 64)
    c = (np.arange(64) + 0.5) / 64 * 2 - 1
    U, V = np.meshgrid(c, c)
    assert np.all(env[U * U + V * V <= 1.0] == 1.0)
