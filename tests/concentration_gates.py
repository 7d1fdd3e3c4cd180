"""The statistical gates the reference applies to concentrated sampling and
adaptive per-(pair, source) budgets, restated for the tests:

  validate_concentration.py:81-101, 148-155
      classic (concentration off) and concentrated renders at the same pupil
      budget, each compared with a dense classic render at 3x the grid (9x the
      samples, the converged truth):
        centroid error (on) <= max(1.25 x centroid error (off), 0.05 px), < 0.5 px
        spread error (on)   <= max(1.25 x spread error (off), 0.005)
        |energy(on) / energy(off) - 1| <= max(3 x resample floor, 0.01)
        RMSE(on vs truth)   <= 1.05 x RMSE(off vs truth)
  validate_adaptive.py:166-187
      adaptive (concentration + budgets) against the same truth:
        RMSE(adaptive)      <= 1.05 x RMSE(classic)
        centroid error      < 0.8 px
        spread error        <= max(2 x spread error (classic), 0.01)
        |energy(adaptive) / energy(concentrated) - 1| <= max(3 x floor, 0.01)

The resample floor is |E(A) / E(B) - 1| for two classic renders on wang
lattices of different seeds (311, 622).  Images are (H, W, 3)."""
from __future__ import annotations

import math

import numpy as np


def moments(img):
    d = img.sum(-1).astype(np.float64)
    tot = d.sum()
    if tot < 1e-20:
        return None
    ys, xs = np.mgrid[0:d.shape[0], 0:d.shape[1]]
    cx = (d * xs).sum() / tot
    cy = (d * ys).sum() / tot
    mxx = (d * (xs - cx) ** 2).sum() / tot
    myy = (d * (ys - cy) ** 2).sum() / tot
    mxy = (d * (xs - cx) * (ys - cy)).sum() / tot
    ev = np.linalg.eigvalsh(np.array([[mxx, mxy], [mxy, myy]]))
    return cx, cy, math.sqrt(max(ev[0], 0.0)) + math.sqrt(max(ev[1], 0.0))


def shape_delta(truth, img):
    """(centroid distance in px, spread ratio img / truth)."""
    ma, mb = moments(truth), moments(img)
    if ma is None or mb is None:
        return 0.0, 1.0
    return math.hypot(ma[0] - mb[0], ma[1] - mb[1]), (mb[2] / ma[2] if ma[2] > 0 else 1.0)


def rmse(a, ref):
    return float(np.sqrt(np.mean((np.asarray(a, np.float64) - ref) ** 2)))


def floor_energy(fa, fb):
    return abs(float(fa.sum() / (fb.sum() + 1e-20)) - 1.0)


def concentration_gate(off, on, dense, floor_e):
    """-> (ok, metrics) for one case of validate_concentration.py."""
    cen_off, spr_off = shape_delta(dense, off)
    cen_on, spr_on = shape_delta(dense, on)
    spr_off_e, spr_on_e = abs(spr_off - 1.0), abs(spr_on - 1.0)
    energy = float(on.sum() / (off.sum() + 1e-20))
    r_off, r_on = rmse(off, dense), rmse(on, dense)
    m = {"cen_off": cen_off, "cen_on": cen_on, "spr_off": spr_off_e, "spr_on": spr_on_e, "energy": energy,
         "floor": floor_e, "rmse_off": r_off, "rmse_on": r_on}
    ok = (cen_on <= max(1.25 * cen_off, 0.05) and cen_on < 0.5
          and spr_on_e <= max(1.25 * spr_off_e, 0.005)
          and abs(energy - 1.0) <= max(3.0 * floor_e, 0.01)
          and r_on <= r_off * 1.05)
    return ok, m


def adaptive_gate(cla, con, ada, dense, floor_e):
    """-> (ok, metrics) for one case of validate_adaptive.py."""
    r_cla, r_con, r_ada = rmse(cla, dense), rmse(con, dense), rmse(ada, dense)
    cen_cla, spr_cla = shape_delta(dense, cla)
    cen_ada, spr_ada = shape_delta(dense, ada)
    spr_cla_e, spr_ada_e = abs(spr_cla - 1.0), abs(spr_ada - 1.0)
    energy = float(ada.sum() / (con.sum() + 1e-20))
    m = {"rmse_cla": r_cla, "rmse_con": r_con, "rmse_ada": r_ada, "cen_ada": cen_ada, "spr_cla": spr_cla_e,
         "spr_ada": spr_ada_e, "energy": energy, "floor": floor_e}
    ok = (r_ada <= r_cla * 1.05
          and cen_ada < 0.8
          and spr_ada_e <= max(2.0 * spr_cla_e, 0.01)
          and abs(energy - 1.0) <= max(3.0 * floor_e, 0.01))
    return ok, m
