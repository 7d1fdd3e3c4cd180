"""Test infrastructure (never shipped or called by the product): CPU
restatement of the starburst's off-axis effective pupil on the FP32 oracle
tracer.

  survivor_center    solve_survivor_center (starburst_render.cu:503-582):
                     81 x 81 launch grid over the spawn plane, binary survival,
                     then a coverage-weighted centroid in raster order where
                     boundary cells contribute 8 x 8 supersamples
  survivor_envelope  build_survivor_envelope (starburst_render.cu:393-487):
                     survival at ME x ME pupil cells inside the real stop,
                     10 x 10 supersampled coverage on cells whose 8-neighbour-
                     hood changes state, outward dilation of radius 2

Positions are formed in float32 in the reference's operation order; sums are
float64.  Survival comes from orc_primary_survival (t >= 0 rule, landing
finite).  Pinned by the two-disc closed form of a two-aperture lens
(tests/test_oracle_survivor.py)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import pyoracle
from .diffraction_oracle import mask_contains

SPAWN_OFFSET = 20.0
F = np.float32


def primary_survival(lens, launch, ax, ay, lam=550.0, threads=0):
    """launch (n, 2) spawn-plane positions -> bool (n,)."""
    surfs, arena = lens.pack()
    n = lens.num_surfaces()
    launch = np.ascontiguousarray(launch, np.float32).reshape(-1, 2)
    out = np.zeros(len(launch), np.uint8)
    if len(launch) == 0:
        return out.astype(bool)
    spec = np.array([[lam, 0, 0, 0]], np.float32)
    at, _bf = pyoracle.ior_tables(surfs, n, spec)
    z0 = F(lens.surfaces[0].z) - F(SPAWN_OFFSET)
    pyoracle.lib().orc_primary_survival(surfs, n, pyoracle._f(arena), pyoracle._f(launch), len(launch),
                                        C.c_float(ax), C.c_float(ay), C.c_float(z0), C.c_float(lam),
                                        pyoracle._f(np.ascontiguousarray(at, np.float32)),
                                        out.ctypes.data_as(C.POINTER(C.c_uint8)), int(threads))
    return out.astype(bool)


_libm = C.CDLL("libm.so.6")
_libm.tanf.restype = C.c_float
_libm.tanf.argtypes = [C.c_float]


def tanf(x):
    """The C library's float tangent (std::tan(float) in the reference)."""
    return F(_libm.tanf(F(x)))


def _boundary(state, valid=None):
    """True where an in-bounds 8-neighbour (counted only where `valid`) has a
    different state."""
    S0, S1 = state.shape
    sp = np.pad(state, 1)
    vp = np.pad(np.ones(state.shape, bool) if valid is None else valid, 1)
    b = np.zeros(state.shape, bool)
    for dj in (-1, 0, 1):
        for di in (-1, 0, 1):
            if di or dj:
                nb_s = sp[1 + dj:1 + dj + S0, 1 + di:1 + di + S1]
                nb_v = vp[1 + dj:1 + dj + S0, 1 + di:1 + di + S1]
                b |= nb_v & (nb_s != state)
    return b


def survivor_center(lens, ax, ay, lam=550.0):
    """-> (cx, cy, any) as float32 / bool."""
    front_R = F(lens.surfaces[0].semi_aperture)
    if not front_R > 0:
        return F(0), F(0), False
    bx, by = tanf(ax), tanf(ay)
    span_x = F(F(1.1) * front_R) + F(F(SPAWN_OFFSET) * abs(bx))
    span_y = F(F(1.1) * front_R) + F(F(SPAWN_OFFSET) * abs(by))
    S = 81
    t = (F(2.0) * np.arange(S, dtype=np.float32) / F(S - 1) - F(1.0)).astype(np.float32)
    lx = (span_x * t).astype(np.float32)
    ly = (span_y * t).astype(np.float32)
    LX, LY = np.meshgrid(lx, ly)
    surv = primary_survival(lens, np.stack([LX.ravel(), LY.ravel()], 1), ax, ay, lam).reshape(S, S)
    bnd = _boundary(surv)
    hx, hy = F(span_x / F(S - 1)), F(span_y / F(S - 1))
    K = 8
    inv_k2 = float(F(1.0) / F(K * K))
    off = (F(2.0) * (np.arange(K, dtype=np.float32) + F(0.5)) / F(K) - F(1.0)).astype(np.float32)
    # supersamples of every boundary cell, raster order, sy outer / sx inner
    cells = np.argwhere(bnd)
    ss = []
    for j, i in cells:
        lly = (ly[j] + hy * off).astype(np.float32)
        llx = (lx[i] + hx * off).astype(np.float32)
        GX, GY = np.meshgrid(llx, lly)
        ss.append(np.stack([GX.ravel(), GY.ravel()], 1))
    sv2 = primary_survival(lens, np.concatenate(ss), ax, ay, lam) if ss else np.zeros(0, bool)
    sx = sy = w = 0.0
    k = 0
    cell_of = {(int(j), int(i)): q for q, (j, i) in enumerate(cells)}
    for j in range(S):
        for i in range(S):
            q = cell_of.get((j, i))
            if q is None:
                if surv[j, i]:
                    sx += float(lx[i]); sy += float(ly[j]); w += 1.0
                continue
            pts = ss[q]
            ok = sv2[k:k + len(pts)]
            k += len(pts)
            for (px, py), o in zip(pts, ok):
                if o:
                    sx += float(px) * inv_k2; sy += float(py) * inv_k2; w += inv_k2
    if w > 0.0:
        return F(sx / w), F(sy / w), True
    return F(0), F(0), False


def survivor_envelope(lens, stop_normals, ep_x, ep_y, cx, cy, ax, ay, ME, lam=550.0):
    """-> (ME, ME) float32 envelope (coverage inside the stop, dilated out)."""
    f = F
    c = (((np.arange(ME, dtype=np.float32) + f(0.5)) / f(ME)) * f(2.0) - f(1.0)).astype(np.float32)
    U, V = np.meshgrid(c, c)
    inside = mask_contains(U, V, 1.0, stop_normals)
    env = np.full((ME, ME), -1.0, np.float32)
    lx = (f(cx) + U * f(ep_x)).astype(np.float32)
    ly = (f(cy) + V * f(ep_y)).astype(np.float32)
    sv = primary_survival(lens, np.stack([lx[inside], ly[inside]], 1), ax, ay, lam)
    env[inside] = np.where(sv, 1.0, 0.0)
    base = env.copy()
    state = base > 0.5
    bnd = _boundary(state, base >= 0.0) & (base >= 0.0)
    K = 10
    du = f(f(2.0) / f(ME))
    off = (((np.arange(K, dtype=np.float32) + f(0.5)) / f(K)) - f(0.5)).astype(np.float32)
    pts, owner, n_in = [], [], []
    cells = np.argwhere(bnd)
    for q, (j, i) in enumerate(cells):
        uu = (c[i] + du * off).astype(np.float32)
        vv = (c[j] + du * off).astype(np.float32)
        GU, GV = np.meshgrid(uu, vv)
        ins = mask_contains(GU, GV, 1.0, stop_normals)
        n_in.append(int(ins.sum()))
        pts.append(np.stack([(f(cx) + GU[ins] * f(ep_x)).astype(np.float32),
                             (f(cy) + GV[ins] * f(ep_y)).astype(np.float32)], 1))
        owner.append(np.full(int(ins.sum()), q))
    if pts:
        s2 = primary_survival(lens, np.concatenate(pts), ax, ay, lam)
        hits = np.bincount(np.concatenate(owner), weights=s2.astype(np.float64), minlength=len(cells))
        for q, (j, i) in enumerate(cells):
            if n_in[q] > 0:
                env[j, i] = f(f(int(hits[q])) / f(n_in[q]))
    R = 2
    filled = env.copy()
    for j, i in np.argwhere(env < 0.0):
        win = env[max(0, j - R):j + R + 1, max(0, i - R):i + R + 1]
        good = win[win >= 0.0]
        filled[j, i] = max(0.0, float(good.max())) if good.size else 0.0
    return filled
