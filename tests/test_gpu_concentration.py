"""Concentrated sampling and adaptive per-(pair, source) budgets on the
device, held to the reference's own statistical gates
(validate_concentration.py:148-155, validate_adaptive.py:166-187; restated in
concentration_gates.py) through the public render_point_flare, exactly as
those scripts call it: DoubleGauss and the 1.33x Sirui, 320x320, sensor half
extents 14 mm, Halton pupil (seed 12345), 16 lambda, gain 5000, source rgb
10, min_ghost_intensity 0, cull on; three off-axis source positions incl. the
far corner, round stop and the 6-blade override; a wang-lattice case per lens
for the lattice budget arithmetic.  Truth = classic sampling at ray_grid 768
(9x the samples of the 256 budget).

Both estimators are unbiased by construction (the weight scales by the rect
area over the reference area per sample drawn), so a wrong area weight, a
biased rect or an off-by-one budget moves the energy or the RMSE against the
truth; the CPU oracle runs the same gates with planted faults of each kind
(test_oracle_concentration.py) to show they are caught."""
import dataclasses

import numpy as np
import pytest

import concentration_gates as cg
import golden_flare as gf
from ghostlight_amd.config import PointFlareConfig

pytestmark = pytest.mark.gpu

W = H = 320
DENSE = 768
POSITIONS = [(0.9, 0.5), (1.3, 0.5), (1.6, 0.9)]


def _cfg(sx, sy, mode, grid=256, blades=0, jitter=2, seed=12345, stats=False):
    c = PointFlareConfig(source_x=sx, source_y=sy, source_r=10.0, source_g=10.0, source_b=10.0, flare_gain=5000.0,
                         ray_grid=grid, spectral_samples=16, pupil_jitter=jitter, jitter_seed=seed,
                         sensor_half_w=gf.SH, sensor_half_h=gf.SH, min_ghost_intensity=0.0, cull_dead_pairs=True,
                         concentrate_samples=(mode != "classic"), adaptive_sample_budgets=(mode == "adaptive"),
                         collect_stats=stats)
    if blades:
        c = dataclasses.replace(c, aperture_blades=blades)
    return c


_LENS = {}


def _render(name, sx, sy, mode, **kw):
    if name not in _LENS:
        _LENS[name] = (gf.lens(name), gf.calibration(name))
    lens, calib = _LENS[name]
    o = lens.render_point_flare(W, H, _cfg(sx, sy, mode, **kw), calib=calib)
    img = np.stack([np.asarray(o["ghost_r"]), np.asarray(o["ghost_g"]), np.asarray(o["ghost_b"])], -1)
    return (img, o.get("stats")) if kw.get("stats") else img


def _floor(name, sx, sy, blades):
    fa = _render(name, sx, sy, "classic", blades=blades, jitter=1, seed=311)
    fb = _render(name, sx, sy, "classic", blades=blades, jitter=1, seed=622)
    return cg.floor_energy(fa, fb)


CASES = [(name, sx, sy, b) for name in gf.LENSES for b in (0, 6) for (sx, sy) in POSITIONS]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}_{c[1]}_{c[2]}_b{c[3]}")
def test_concentration_and_adaptive_gates(case):
    name, sx, sy, blades = case
    off = _render(name, sx, sy, "classic", blades=blades)
    on = _render(name, sx, sy, "conc", blades=blades)
    ada, st = _render(name, sx, sy, "adaptive", blades=blades, stats=True)
    dense = _render(name, sx, sy, "classic", grid=DENSE, blades=blades)
    fe = _floor(name, sx, sy, blades)
    okc, mc = cg.concentration_gate(off, on, dense, fe)
    oka, ma = cg.adaptive_gate(off, on, ada, dense, fe)
    bud = np.asarray(st["ps_budget"], np.float64)
    rect = np.asarray(st["ps_rect"], np.float64).reshape(-1, 4)
    valid = rect[:, 0] <= rect[:, 2]
    bfrac = float(np.mean(bud[valid] / float(st["n_grid"]))) if valid.any() else float("nan")
    print(f"{name} ({sx},{sy}) b{blades}: conc {okc} {mc}; adaptive {oka} {ma}; "
          f"ada/conc RMSE {ma['rmse_ada'] / max(ma['rmse_con'], 1e-30):.3f}, mean budget {100 * bfrac:.1f}% "
          f"of the grid over {int(valid.sum())} concentrated (pair, source)", flush=True)
    assert okc, f"concentration gate: {mc}"
    assert oka, f"adaptive gate: {ma}"
    assert valid.any(), "no (pair, source) was concentrated"
    assert bfrac < 1.0, "adaptive budgets never shrank"


@pytest.mark.parametrize("name", gf.LENSES)
def test_adaptive_wang_lattice_gate(name):
    """validate_adaptive.py:149: the only case that exercises the lattice
    (pupil_jitter = 1) budget arithmetic."""
    sx, sy = 1.3, 0.5
    cla = _render(name, sx, sy, "classic", jitter=1)
    con = _render(name, sx, sy, "conc", jitter=1)
    ada = _render(name, sx, sy, "adaptive", jitter=1)
    dense = _render(name, sx, sy, "classic", grid=DENSE, jitter=1)
    fe = _floor(name, sx, sy, 0)
    oka, ma = cg.adaptive_gate(cla, con, ada, dense, fe)
    print(f"{name} wang: adaptive {oka} {ma}", flush=True)
    assert oka, f"adaptive gate (wang lattice): {ma}"
