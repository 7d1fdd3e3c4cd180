"""Geometric PSF grid renderer (render_psf, bind_renderers.cpp:148-191;
PSFRenderer.cpp:260-465).

Host side: the chief-ray pre-pass that anchors every source's tile, then one
device launch (gl_render_psf_tiles) that traces every (source, pupil sample,
wavelength) and splats it relative to the chief landing.

The pre-pass traces on the device: each Newton iterate of every source is one
lane of a batched primary trace, so a whole grid of sources advances in
lockstep (three landings per iterate: the point and its two finite-difference
neighbours), with the per-source control flow of newton_aim
(newton_aim.h:19-59) kept on the host.
"""
from __future__ import annotations

import enum
import ctypes as C
import math
from typing import Optional

import numpy as np

from . import _abi
from .config import PSFCenterMode, PSFConfig, render_params
from .lens import OpticalSystem
from .render import trace_ghost_rays
from .runtime import ctx_handle, lib

PSF_CELL_OK = 0
PSF_CELL_CHIEF_VIGNETTED = 1
PSF_CELL_DARK = 2


class PSFCellStatus(enum.IntEnum):
    """Per-cell status in the `status` array of render_psf (PSFRenderer.h:22-26)."""
    OK = PSF_CELL_OK                            # the chief ray reached the cell centre
    CHIEF_VIGNETTED = PSF_CELL_CHIEF_VIGNETTED  # chief blocked, part of the pupil transmits
    DARK = PSF_CELL_DARK                        # no transmission near the cell

CHIEF_LAMBDA_NM = 587.56


def _rays(lens: OpticalSystem, ax, ay, u, v):
    """Primary rays from the spawn plane (spawn_plane.h: 20 mm ahead of the
    front vertex, disc shifted with the beam) in float32 arithmetic."""
    f = np.float32
    ax = np.asarray(ax, np.float32)
    ay = np.asarray(ay, np.float32)
    bx = np.tan(ax).astype(np.float32)
    by = np.tan(ay).astype(np.float32)
    l = np.sqrt(bx * bx + by * by + f(1.0)).astype(np.float32)
    inv = (f(1.0) / l).astype(np.float32)
    fR = f(lens.surfaces[0].semi_aperture)
    z0 = f(lens.surfaces[0].z) - f(20.0)
    r = np.zeros((len(ax), 7), np.float32)
    r[:, 0] = np.asarray(u, np.float32) * fR + f(-20.0) * bx
    r[:, 1] = np.asarray(v, np.float32) * fR + f(-20.0) * by
    r[:, 2] = z0
    r[:, 3] = bx * inv
    r[:, 4] = by * inv
    r[:, 5] = inv
    r[:, 6] = f(CHIEF_LAMBDA_NM)
    return r


def _land(lens, ax, ay, u, v, ctx, fp64):
    """Landings (x, y, ok) of one primary ray per entry (CPU t >= 0 rule)."""
    if len(ax) == 0:
        return np.zeros(0), np.zeros(0), np.zeros(0, bool)
    pos, _w, st, _sg = trace_ghost_rays(lens, -1, -1, _rays(lens, ax, ay, u, v), t_ge0=True,
                                        fp64_radius_min=fp64, ctx=ctx)
    ok = (st == 0) & np.isfinite(pos[:, 0]) & np.isfinite(pos[:, 1])
    return pos[:, 0].astype(np.float64), pos[:, 1].astype(np.float64), ok


def _disk(N):
    c = -1.0 + (np.arange(N) + 0.5) * (2.0 / N)
    U, V = np.meshgrid(c, c)
    m = U * U + V * V <= 1.0
    return U[m].astype(np.float32), V[m].astype(np.float32)


def pupil_probe(lens, ax, ay, N, ctx, fp64):
    """N x N unit-disk probe per source (PSFRenderer.cpp:73-104): hits,
    fraction, survivor-mean landing."""
    du, dv = _disk(N)
    k = len(du)
    n = len(ax)
    AX = np.repeat(np.asarray(ax, np.float32), k)
    AY = np.repeat(np.asarray(ay, np.float32), k)
    x, y, ok = _land(lens, AX, AY, np.tile(du, n), np.tile(dv, n), ctx, fp64)
    ok = ok.reshape(n, k)
    hits = ok.sum(1)
    xs = np.where(ok, x.reshape(n, k), 0.0).sum(1)
    ys = np.where(ok, y.reshape(n, k), 0.0).sum(1)
    mx = np.where(hits > 0, xs / np.maximum(hits, 1), 0.0)
    my = np.where(hits > 0, ys / np.maximum(hits, 1), 0.0)
    return hits, hits / float(k), mx, my


def newton_aim_batch(landing, tx, ty, ax, ay, max_iter, fd, step_cap, tol):
    """newton_aim (newton_aim.h:19-59) for many sources at once.  landing(ax,
    ay, idx) -> (x, y, ok) evaluates the listed sources.  Returns (ax, ay,
    residual) of each source's best iterate."""
    n = len(tx)
    ax = ax.astype(np.float64).copy()
    ay = ay.astype(np.float64).copy()
    best_ax, best_ay = ax.copy(), ay.copy()
    best = np.full(n, 1e30)
    idx_all = np.arange(n)
    cx, cy, ok = landing(ax, ay, idx_all)
    alive = ok.copy()
    res0 = np.full(n, 1e30)

    def record(i, x, y):
        r = np.hypot(x - tx[i], y - ty[i])
        better = r < best[i]
        best[i[better]] = r[better]
        best_ax[i[better]] = ax[i[better]]
        best_ay[i[better]] = ay[i[better]]
        return r

    for _ in range(max_iter):
        i = np.nonzero(alive)[0]
        if len(i) == 0:
            break
        r = record(i, cx[i], cy[i])
        conv = r < tol
        alive[i[conv]] = False
        i = i[~conv]
        if len(i) == 0:
            break
        xx, xy, okx = landing(ax[i] + fd, ay[i], i)
        yx, yy, oky = landing(ax[i], ay[i] + fd, i)
        good = okx & oky
        alive[i[~good]] = False
        i, xx, xy, yx, yy = i[good], xx[good], xy[good], yx[good], yy[good]
        if len(i) == 0:
            break
        Jxx, Jyx = (xx - cx[i]) / fd, (xy - cy[i]) / fd
        Jxy, Jyy = (yx - cx[i]) / fd, (yy - cy[i]) / fd
        det = Jxx * Jyy - Jxy * Jyx
        sing = np.abs(det) < 1e-12
        alive[i[sing]] = False
        i, Jxx, Jyx, Jxy, Jyy, det = i[~sing], Jxx[~sing], Jyx[~sing], Jxy[~sing], Jyy[~sing], det[~sing]
        if len(i) == 0:
            break
        rx, ry = cx[i] - tx[i], cy[i] - ty[i]
        dax = np.clip((-rx * Jyy + ry * Jxy) / det, -step_cap[i], step_cap[i])
        day = np.clip((rx * Jyx - ry * Jxx) / det, -step_cap[i], step_cap[i])
        ax[i] += dax
        ay[i] += day
        nx, ny, okn = landing(ax[i], ay[i], i)
        cx[i], cy[i] = nx, ny
        alive[i[~okn]] = False
    # fold in the final landing of every source that started, then revert to the best iterate
    st = np.nonzero(ok)[0]
    if len(st):
        fx, fy, okf = landing(ax[st], ay[st], st)
        record(st[okf], fx[okf], fy[okf])
    res0[:] = best
    return best_ax, best_ay, res0, ok


def chief_prepass(lens, sources, cfg: PSFConfig, ctx=None):
    """Chief landing, status, pupil fraction, aim residual and the traced
    beam direction of every source (PSFRenderer.cpp:370-440)."""
    fp64 = float(getattr(cfg, "fp64_radius_min", 1000.0))
    n = len(sources)
    ax = sources[:, 0].astype(np.float64)
    ay = sources[:, 1].astype(np.float64)
    chief_x = np.zeros(n)
    chief_y = np.zeros(n)
    status = np.full(n, PSF_CELL_OK, np.uint8)
    frac = np.zeros(n, np.float32)
    resid = np.zeros(n, np.float32)
    beam_ax, beam_ay = ax.copy(), ay.copy()
    if int(cfg.center_mode) != int(PSFCenterMode.PSF_CENTER_FIXED_TARGET):
        hits, fr, mx, my = pupil_probe(lens, ax, ay, 7, ctx, fp64)
        chief_x, chief_y = mx, my
        status = np.where(hits > 0, PSF_CELL_OK, PSF_CELL_DARK).astype(np.uint8)
        frac = fr.astype(np.float32)
        # A deep pupil between the probe's samples: retry in the tracked axial
        # window (PSFRenderer.cpp:383-410); frac stays the full-disc measure.
        dark = np.nonzero(hits == 0)[0]
        if len(dark):
            found = deep_pupil(lens, ax[dark], ay[dark], None, ctx)
            for j, i in enumerate(dark):
                if found[j, 5] > 0:
                    chief_x[i], chief_y[i] = found[j, 2], found[j, 3]
                    status[i] = PSF_CELL_OK
        return chief_x, chief_y, status, frac, resid, beam_ax, beam_ay
    tx = sources[:, 5].astype(np.float64)
    ty = sources[:, 6].astype(np.float64)
    cap = np.maximum(0.02, 0.5 * np.hypot(ax, ay))
    # clean chief: the u = v = 0 primary aimed at the target to 0.1 um
    def land_center(a, b, idx):
        zero = np.zeros(len(a), np.float32)
        return _land(lens, a, b, zero, zero, ctx, fp64)
    bax, bay, res, started = newton_aim_batch(land_center, tx, ty, ax, ay, 8, 1e-4, cap, 1e-4)
    done = started.copy()
    if done.any():
        i = np.nonzero(done)[0]
        x, y, _ = land_center(bax[i], bay[i], i)
        chief_x[i], chief_y[i] = x, y
        beam_ax[i], beam_ay[i] = bax[i], bay[i]
        resid[i] = res[i]
        status[i] = PSF_CELL_OK
        frac[i] = pupil_probe(lens, bax[i], bay[i], 17, ctx, fp64)[1]
    # fallback: the survivor mean of a 9 x 9 probe aimed to 5 um
    rest = np.nonzero(~done)[0]
    if len(rest):
        def land_mean(a, b, idx):
            h, _f, mx, my = pupil_probe(lens, a, b, 9, ctx, fp64)
            return mx, my, h >= 4
        bax2, bay2, res2, started2 = newton_aim_batch(land_mean, tx[rest], ty[rest], ax[rest], ay[rest], 8, 1e-4,
                                                      cap[rest], 5e-3)
        for j, i in enumerate(rest):
            if started2[j]:
                h, f, mx, my = pupil_probe(lens, bax2[j:j + 1], bay2[j:j + 1], 9, ctx, fp64)
                chief_x[i], chief_y[i] = mx[0], my[0]
                beam_ax[i], beam_ay[i] = bax2[j], bay2[j]
                resid[i] = res2[j]
                status[i] = PSF_CELL_CHIEF_VIGNETTED
                frac[i] = pupil_probe(lens, bax2[j:j + 1], bay2[j:j + 1], 17, ctx, fp64)[1][0]
            else:
                chief_x[i], chief_y[i] = tx[i], ty[i]
                status[i] = PSF_CELL_DARK
                resid[i] = 1e30
                frac[i] = 0.0
        # Deep pupil: aim the tracked axial window's survivor mean instead
        # (PSFRenderer.cpp:203-247).
        dark = np.array([i for i in rest if status[i] == PSF_CELL_DARK], np.int64)
        if len(dark):
            found = deep_pupil(lens, ax[dark], ay[dark], np.stack([tx[dark], ty[dark]], 1), ctx)
            for j, i in enumerate(dark):
                if found[j, 5] > 0:
                    beam_ax[i], beam_ay[i] = found[j, 0], found[j, 1]
                    chief_x[i], chief_y[i] = found[j, 2], found[j, 3]
                    resid[i] = found[j, 4]
                    status[i] = PSF_CELL_CHIEF_VIGNETTED
                    frac[i] = pupil_probe(lens, beam_ax[i:i + 1], beam_ay[i:i + 1], 17, ctx, fp64)[1][0]
    return chief_x, chief_y, status, frac, resid, beam_ax, beam_ay


def deep_pupil(lens, ax, ay, targets=None, ctx=None) -> np.ndarray:
    """gl_psf_deep_pupil: per source [beam ax, ay, landing x, y, residual,
    found] from the axial window tracked out to the source angle (targets
    None: survivor mean at the angle; else the angle aimed at the target)."""
    ang = np.ascontiguousarray(np.stack([np.asarray(ax, np.float32), np.asarray(ay, np.float32)], 1))
    out = np.zeros((len(ang), 6), np.float32)
    tg = None if targets is None else np.ascontiguousarray(targets, np.float32).reshape(-1, 2)
    dl = lens.device_lens()
    _abi.check(lib().gl_psf_deep_pupil(ctx_handle(ctx), dl.handle, len(ang), _abi.fptr(ang),
                                       _abi.fptr(tg) if tg is not None else None, int(tg is not None),
                                       _abi.fptr(out)), "render_psf")
    return out


def render_psf(sources, lens: OpticalSystem, config: PSFConfig, ctx=None, timing: bool = False) -> dict:
    """render_psf(sources, lens, config) -> dict (bind_renderers.cpp:148-191).
    sources: (N, 5) [angle_x, angle_y, r, g, b] or (N, 7) with target_x_mm,
    target_y_mm appended (FIXED_TARGET mode; angles are the aim seed)."""
    cfg = config
    if cfg.grid_nx <= 0 or cfg.grid_ny <= 0 or cfg.tile_w <= 0 or cfg.tile_h <= 0:
        raise ValueError("render_psf: grid and tile dimensions must be positive")
    if not (cfg.tile_extent_mm > 0.0) or not math.isfinite(cfg.tile_extent_mm):
        raise ValueError("render_psf: tile_extent_mm must be finite and positive")
    s = np.asarray(sources, np.float32)
    if s.ndim != 2 or s.shape[1] not in (5, 7):
        raise ValueError("render_psf: sources must be a (N, 5) array of [angle_x, angle_y, r, g, b] rows, or a "
                         "(N, 7) array with [.., target_x_mm, target_y_mm] appended")
    if s.shape[1] == 5:
        s = np.concatenate([s, np.zeros((len(s), 2), np.float32)], 1)
    if not np.all(np.isfinite(s)):
        raise RuntimeError("render_psf: HIP error: source values must be finite")
    n = len(s)
    if n > cfg.grid_nx * cfg.grid_ny:
        raise RuntimeError(f"render_psf: HIP error: {n} sources exceeds tile capacity {cfg.grid_nx * cfg.grid_ny}")
    cx, cy, status, frac, resid, bax, bay = chief_prepass(lens, s, cfg, ctx) if n else \
        (np.zeros(0),) * 2 + (np.zeros(0, np.uint8), np.zeros(0, np.float32), np.zeros(0, np.float32),
                              np.zeros(0), np.zeros(0))
    arr = (_abi.gl_psf_source * max(1, n))()
    for i in range(n):
        a = arr[i]
        a.angle_x, a.angle_y = float(bax[i]), float(bay[i])
        a.r, a.g, a.b = float(s[i, 2]), float(s[i, 3]), float(s[i, 4])
        a.chief_x_mm, a.chief_y_mm = float(cx[i]), float(cy[i])
        a.tile_x0, a.tile_y0 = (i % cfg.grid_nx) * cfg.tile_w, (i // cfg.grid_nx) * cfg.tile_h
    pp = _abi.gl_psf_params()
    pp.grid_nx, pp.grid_ny, pp.tile_w, pp.tile_h = cfg.grid_nx, cfg.grid_ny, cfg.tile_w, cfg.tile_h
    pp.tile_extent_mm = float(cfg.tile_extent_mm)
    pp.monochromatic = int(bool(cfg.monochromatic))
    pp.splat_sigma_um = float(cfg.splat_sigma_um)
    rp = render_params(cfg)
    W, H = cfg.grid_nx * cfg.tile_w, cfg.grid_ny * cfg.tile_h
    planes = [np.zeros((H, W), np.float32) for _ in range(3)]
    ms = C.c_float(0.0)
    dl = lens.device_lens()
    _abi.check(lib().gl_render_psf_tiles(ctx_handle(ctx), dl.handle, arr, n, C.byref(rp),
                                         C.byref(pp), _abi.fptr(planes[0]), _abi.fptr(planes[1]),
                                         _abi.fptr(planes[2]), C.byref(ms)), "render_psf")
    out = {"composite_w": W, "composite_h": H, "tile_w": cfg.tile_w, "tile_h": cfg.tile_h,
           "grid_nx": cfg.grid_nx, "grid_ny": cfg.grid_ny, "tile_extent_mm": float(cfg.tile_extent_mm),
           "r": planes[0], "g": planes[1], "b": planes[2],
           "chief_x_mm": np.asarray(cx, np.float32), "chief_y_mm": np.asarray(cy, np.float32),
           "status": np.asarray(status, np.uint8), "pupil_fraction": np.asarray(frac, np.float32),
           "aim_residual_mm": np.asarray(resid, np.float32)}
    if timing:
        out["kernel_ms"] = ms.value
    return out
