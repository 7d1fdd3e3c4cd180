"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py).
PSF grid splat (psf_kernel, ghostlight/src/psf_render.cu:34-218) on the CPU:
primary traces through the FP32 oracle tracer (GPU t > 0 rule), bilinear or
normalised-Gaussian splat relative to the given chief landing, clipped to
the source's tile; accumulation in float64.  The chief pre-pass probe
(pupil_probe, ghostlight/src/PSFRenderer.cpp:73-104) is restated too."""
from __future__ import annotations

import math

import numpy as np

from oracle import pyoracle


def _rays(lens, ax, ay, u, v, lam):
    f = np.float32
    bx, by = f(math.tan(f(ax))), f(math.tan(f(ay)))
    l = f(np.sqrt(bx * bx + by * by + f(1.0)))
    inv = f(f(1.0) / l)
    r = np.zeros((len(u), 7), np.float32)
    r[:, 0] = np.asarray(u, np.float32) * f(lens.surfaces[0].semi_aperture) + f(-20.0) * bx
    r[:, 1] = np.asarray(v, np.float32) * f(lens.surfaces[0].semi_aperture) + f(-20.0) * by
    r[:, 2] = f(lens.surfaces[0].z) - f(20.0)
    r[:, 3], r[:, 4], r[:, 5] = bx * inv, by * inv, inv
    r[:, 6] = lam
    return r


def pupil_probe_mean(lens, ax, ay, N=7, fp64=1000.0):
    c = -1.0 + (np.arange(N) + 0.5) * (2.0 / N)
    U, V = np.meshgrid(c, c)
    m = U * U + V * V <= 1.0
    r = _rays(lens, ax, ay, U[m], V[m], 587.56)
    pos, _w, st, _ = pyoracle.trace_ghost(lens, -1, -1, r, dbl=False, t_ge0=True, fp64_radius_min=fp64)
    ok = (st == 0) & np.isfinite(pos[:, 0]) & np.isfinite(pos[:, 1])
    if not ok.any():
        return 0, 0.0, 0.0
    return int(ok.sum()), float(pos[ok, 0].mean()), float(pos[ok, 1].mean())


def psf_splat(lens, sources, chief, grid, spec, tile_w, tile_h, grid_nx, grid_ny, mm_per_px, sigma_px, radius,
              mono=False, fp64=1000.0):
    """sources: (N, 5) [beam angle_x, angle_y, r, g, b]; chief: (N, 2) mm."""
    W, H = grid_nx * tile_w, grid_ny * tile_h
    out = np.zeros((3, H, W))
    ray_w = 1.0 / len(grid)
    for i, s in enumerate(sources):
        tx0, ty0 = (i % grid_nx) * tile_w, (i // grid_nx) * tile_h
        for lam, cr_, cg_, cb_ in spec:
            r = _rays(lens, s[0], s[1], grid[:, 0], grid[:, 1], lam)
            pos, w, st, _ = pyoracle.trace_ghost(lens, -1, -1, r, dbl=False, t_ge0=False, fp64_radius_min=fp64)
            ok = (st == 0) & np.isfinite(pos[:, 0]) & np.isfinite(pos[:, 1])
            tx = (pos[ok, 0] - chief[i, 0]) / mm_per_px + tile_w * 0.5
            ty = (pos[ok, 1] - chief[i, 1]) / mm_per_px + tile_h * 0.5
            ww = w[ok] * ray_w
            margin = radius + 1 if radius > 0 else 1
            keep = (tx >= -margin) & (tx <= tile_w + margin) & (ty >= -margin) & (ty <= tile_h + margin)
            tx, ty, ww = tx[keep], ty[keep], ww[keep]
            px, py = tx + tx0, ty + ty0
            avg = (s[2] + s[3] + s[4]) / 3.0
            cols = [avg * cg_ * ww] if mono else [s[2] * cr_ * ww, s[3] * cg_ * ww, s[4] * cb_ * ww]
            if sigma_px <= 0:
                x0, y0 = np.floor(px - 0.5).astype(int), np.floor(py - 0.5).astype(int)
                fx, fy = (px - 0.5) - x0, (py - 0.5) - y0
                taps = [(x0, y0, (1 - fx) * (1 - fy)), (x0 + 1, y0, fx * (1 - fy)), (x0, y0 + 1, (1 - fx) * fy),
                        (x0 + 1, y0 + 1, fx * fy)]
            else:
                cx, cy = np.floor(px).astype(int), np.floor(py).astype(int)
                ox, oy = px - (cx + 0.5), py - (cy + 0.5)
                d = np.arange(-radius, radius + 1)
                g = np.exp(-((d[None, :, None] - ox[:, None, None]) ** 2 + (d[None, None, :] - oy[:, None, None]) ** 2)
                           / (2.0 * sigma_px ** 2))            # [ray, dx, dy]
                g = g / g.sum(axis=(1, 2), keepdims=True)
                taps = []
                for a, dx in enumerate(d):
                    for b, dy in enumerate(d):
                        wt = g[:, a, b]
                        taps.append((cx + dx, cy + dy, np.where(wt < 1e-6, 0.0, wt)))
            for xs, ys, wt in taps:
                inside = (xs >= tx0) & (xs < tx0 + tile_w) & (ys >= ty0) & (ys < ty0 + tile_h)
                for c, col in enumerate(cols):
                    np.add.at(out[c], (ys[inside], xs[inside]), (col * wt)[inside])
    return out
