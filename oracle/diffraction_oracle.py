"""ORACLE — test infrastructure only (tests/, __graft_entry__.smoke(), bench.py).
numpy FP64 restatement of the reference's diffraction passes:
  build_pupil_kernel        ghostlight/src/starburst_render.cu:155-191
  baffle_stack_blocks       ghostlight/src/baffle.h:49-73
  psf_from_fft + unit sum   ghostlight/src/starburst_render.cu:202-232, 976-996
  bake_sprite_kernel        ghostlight/src/starburst_render.cu:255-278
  downsample_kernel         ghostlight/src/starburst_render.cu:283-299
  splat_resample_kernel     ghostlight/src/starburst_render.cu:316-384
  veil reference + splat    ghostlight/src/veil_render.cu:94-212, 243-414
  mdft_starburst            ghostlight/src/mdft_render.cu:77-340
  veil glare-spread kernel  ghostlight/src/veil_render.cu:94-106, applied to a
                            whole image by linear FFT convolution
  the pupil -> PSF matrix DFT (MFT) of the precision starburst workload.
The FFT is numpy's (pocketfft, complex128); the reference's own numpy
mirrors (validation/validate_starburst.py:49-60,
validation/validate_starburst_mdft.py:52-69) define the same transforms.
"""
from __future__ import annotations

import math

import numpy as np


def mask_normals(n_blades: int, rot_rad: float):
    """Side normals at rot + sector (k + 1/2), rounded once from FP64, and the
    unit apothem cos(pi / n) (the half-plane polygon test)."""
    if n_blades < 3:
        return None
    sector = 2.0 * math.pi / n_blades
    phi = np.float64(np.float32(rot_rad)) + sector * (np.arange(n_blades) + 0.5)
    return (np.cos(phi).astype(np.float32), np.sin(phi).astype(np.float32),
            np.float32(math.cos(math.pi / n_blades)))


def mask_contains(u, v, aspect, normals):
    u = np.asarray(u, np.float32)
    v = np.asarray(v, np.float32)
    ux = (u / np.float32(aspect)).astype(np.float32)
    r2 = ux * ux + v * v
    ok = r2 <= np.float32(1.0)
    if normals is not None:
        c, s, ap = normals
        m = ux[..., None] * c + v[..., None] * s
        ok &= m.max(-1) <= ap
    return ok


def sample_env(env, u, v):
    """Bilinear lookup of the ME x ME envelope at pupil (u, v), clamped to the
    cell centres, in float32 with the device's operation order."""
    f = np.float32
    env = np.asarray(env, np.float32)
    ME = env.shape[0]
    u = np.asarray(u, np.float32)
    v = np.asarray(v, np.float32)
    fx = ((u * f(0.5) + f(0.5)) * f(ME) - f(0.5)).astype(np.float32)
    fy = ((v * f(0.5) + f(0.5)) * f(ME) - f(0.5)).astype(np.float32)
    fx = np.minimum(np.maximum(fx, f(0.0)), f(ME - 1)).astype(np.float32)
    fy = np.minimum(np.maximum(fy, f(0.0)), f(ME - 1)).astype(np.float32)
    x0 = np.floor(fx).astype(int)
    y0 = np.floor(fy).astype(int)
    x1 = np.minimum(x0 + 1, ME - 1)
    y1 = np.minimum(y0 + 1, ME - 1)
    tx = (fx - x0.astype(np.float32)).astype(np.float32)
    ty = (fy - y0.astype(np.float32)).astype(np.float32)
    a = (env[y0, x0] * (f(1) - tx) + env[y0, x1] * tx).astype(np.float32)
    b = (env[y1, x0] * (f(1) - tx) + env[y1, x1] * tx).astype(np.float32)
    return (a * (f(1) - ty) + b * ty).astype(np.float32)


def baffles_block(baffles, px, py, tan_ax, tan_ay):
    """baffle_stack_blocks (baffle.h:62-73) in float32: rows of (shape, dz, top,
    bottom, left, right, cx, cy, radius, aspect); True where any plane blocks."""
    f = np.float32
    px = np.asarray(px, np.float32)
    py = np.asarray(py, np.float32)
    out = np.zeros(np.broadcast(px, py).shape, bool)
    for b in baffles:
        shape, dz, top, bottom, left, right, cx, cy, radius, aspect = (f(x) for x in b)
        mx = (px + f(tan_ax) * dz).astype(np.float32)
        my = (py + f(tan_ay) * dz).astype(np.float32)
        if int(shape) == 0:
            out |= (mx > right) | (mx < -left) | (my > top) | (my < -bottom)
        else:
            dx = ((mx - cx) / radius).astype(np.float32)
            dy = ((my - cy) / (radius * aspect)).astype(np.float32)
            out |= (dx * dx + dy * dy) > f(1.0)
    return out


def image_bilinear(img, u, v):
    """tex2D with linear filtering, normalised coordinates and clamp addressing
    (texel centres at (i + 1/2) / W), float32 operations in the device's order."""
    f = np.float32
    H, W = img.shape
    x = (np.asarray(u, np.float32) * f(W) - f(0.5)).astype(np.float32)
    y = (np.asarray(v, np.float32) * f(H) - f(0.5)).astype(np.float32)
    fx, fy = np.floor(x), np.floor(y)
    a, b = (x - fx).astype(np.float32), (y - fy).astype(np.float32)
    i0 = np.clip(fx.astype(np.int64), 0, W - 1); i1 = np.clip(fx.astype(np.int64) + 1, 0, W - 1)
    j0 = np.clip(fy.astype(np.int64), 0, H - 1); j1 = np.clip(fy.astype(np.int64) + 1, 0, H - 1)
    img = np.asarray(img, np.float32)
    top = ((f(1) - a) * img[j0, i0] + a * img[j0, i1]).astype(np.float32)
    bot = ((f(1) - a) * img[j1, i0] + a * img[j1, i1]).astype(np.float32)
    return ((f(1) - b) * top + b * bot).astype(np.float32)


def build_pupil(N, fill, normals, env=None, baffles=None, ep=(1.0, 1.0), tan_a=(0.0, 0.0), tex=None):
    """4x4-supersampled stop x envelope x baffle clip x front bitmap, float32
    sample positions as the kernel; a baffle sees the pupil ray at
    (u ep_x, v ep_y) heading tan_a; tex = (image, aspect, semi, d0) is sampled
    where that ray crosses the bitmap's surface d0 downstream."""
    i = np.arange(N, dtype=np.float32)
    acc = np.zeros((N, N), np.float64)
    f = np.float32
    for sy in range(4):
        v = ((((i + (f(sy) + f(0.5)) * f(0.25)) / f(N)) * f(2.0) - f(1.0)) / f(fill)).astype(np.float32)
        for sx in range(4):
            u = ((((i + (f(sx) + f(0.5)) * f(0.25)) / f(N)) * f(2.0) - f(1.0)) / f(fill)).astype(np.float32)
            U, V = np.meshgrid(u, v)
            m = mask_contains(U, V, 1.0, normals).astype(np.float64)
            if env is not None:
                m *= sample_env(env, U, V).astype(np.float64)
            if baffles:
                m[baffles_block(baffles, U * np.float32(ep[0]), V * np.float32(ep[1]), tan_a[0], tan_a[1])] = 0.0
            if tex is not None:
                img, t_asp, t_semi, t_d0 = tex
                g = np.float32
                hx = ((U * g(ep[0]) + g(tan_a[0]) * g(t_d0)) / g(t_asp)).astype(np.float32)
                hy = (V * g(ep[1]) + g(tan_a[1]) * g(t_d0)).astype(np.float32)
                tu = (g(0.5) + hx / (g(2.0) * g(t_semi))).astype(np.float32)
                tv = (g(0.5) + hy / (g(2.0) * g(t_semi))).astype(np.float32)
                m *= image_bilinear(img, tu, tv).astype(np.float64)
            acc += m
    return acc / 16.0


def psf_from_pupil(pupil):
    F = np.fft.fft2(pupil.astype(np.complex128))
    psf = np.fft.fftshift(np.abs(F) ** 2)
    s = psf.sum()
    return psf / s if s > 0 else psf


def bake_sprite(psf, spec, lambda_ref=550.0):
    N = psf.shape[0]
    half = N // 2
    d = np.arange(N) - half
    DX, DY = np.meshgrid(d.astype(np.float64), d.astype(np.float64))
    out = np.zeros((N, N, 3))
    for lam, cr, cg, cb in spec:
        k = lambda_ref / lam
        fx, fy = half + DX * k, half + DY * k
        ok = (fx >= 0) & (fy >= 0) & (fx < N - 1) & (fy < N - 1)
        x0 = np.clip(np.floor(fx).astype(int), 0, N - 2); y0 = np.clip(np.floor(fy).astype(int), 0, N - 2)
        tx, ty = fx - x0, fy - y0
        a = psf[y0, x0] * (1 - tx) + psf[y0, x0 + 1] * tx
        b = psf[y0 + 1, x0] * (1 - tx) + psf[y0 + 1, x0 + 1] * tx
        val = np.where(ok, (a * (1 - ty) + b * ty) * k * k, 0.0)
        out[..., 0] += cr * val; out[..., 1] += cg * val; out[..., 2] += cb * val
    return out


def downsample(sprite, ds):
    if ds <= 1:
        return sprite
    N = sprite.shape[0]
    M = N // ds
    C = sprite.shape[2] if sprite.ndim == 3 else 1
    s = sprite.reshape(M, ds, M, ds, C) if sprite.ndim == 3 else sprite.reshape(M, ds, M, ds)
    return s.sum(axis=(1, 3))


def po2_floor(x, hi):
    p = 1
    while p * 2 <= x and p * 2 <= hi:
        p *= 2
    return p


def resample(sprite, sx_ds, sy_ds, px_mm_x, px_mm_y, sources, gain, width, height):
    """Exact area-weighted box resample (splat_resample_kernel semantics)."""
    M = sprite.shape[0]
    mono = sprite.ndim == 2
    out = np.zeros((3, height, width))
    rx = min(int(math.ceil(0.5 * M * sx_ds / px_mm_x)) + 1, width)
    ry = min(int(math.ceil(0.5 * M * sy_ds / px_mm_y)) + 1, height)
    # cumulative sums give the box integrals in closed form
    S = sprite if not mono else sprite[..., None]
    for (px, py, r, g, b) in sources:
        cx, cy = int(math.floor(px)), int(math.floor(py))
        ix = np.arange(max(0, cx - rx), min(width, cx + rx + 1))
        iy = np.arange(max(0, cy - ry), min(height, cy + ry + 1))
        if len(ix) == 0 or len(iy) == 0:
            continue
        xlo = ((ix - px) * px_mm_x) / sx_ds + 0.5 * M
        xhi = ((ix + 1.0 - px) * px_mm_x) / sx_ds + 0.5 * M
        ylo = ((iy - py) * px_mm_y) / sy_ds + 0.5 * M
        yhi = ((iy + 1.0 - py) * px_mm_y) / sy_ds + 0.5 * M
        cells = np.arange(M)
        wx = np.clip(np.minimum(xhi[:, None], cells + 1) - np.maximum(xlo[:, None], cells), 0, None)
        wy = np.clip(np.minimum(yhi[:, None], cells + 1) - np.maximum(ylo[:, None], cells), 0, None)
        for ch, col in enumerate((r, g, b)):
            plane = S[..., ch if not mono else 0]
            out[ch][np.ix_(iy, ix)] += (wy @ plane @ wx.T) * col * gain
    return out


def veil_reference(N=1024, a_texels=32.0, p=1.5):
    d = np.arange(N) - N // 2
    DX, DY = np.meshgrid(d, d)
    g = (a_texels ** 2 / (DX ** 2 + DY ** 2 + a_texels ** 2)) ** p
    return g / g.sum()


def mft(pupil, M, q_x, q_y=None):
    """Matrix DFT of an N x N pupil to an M x M field, sampling q times finer
    than the FFT grid (q may differ per axis): E = Ay @ P @ Ax^T with
    A[k, n] = exp(-2 pi i (k - M/2)(n - N/2) / (q N)).  complex128."""
    N = pupil.shape[0]
    q_y = q_x if q_y is None else q_y
    n = np.arange(N) - N / 2
    k = np.arange(M) - M / 2
    Ax = np.exp(-2j * np.pi * np.outer(k, n) / (q_x * N))
    Ay = np.exp(-2j * np.pi * np.outer(k, n) / (q_y * N))
    return Ay @ pupil.astype(np.complex128) @ Ax.T


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel(); b = np.asarray(b, np.float64).ravel()
    d = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / d) if d > 0 else float(np.linalg.norm(a - b))


def autocorr_C(pupil, fill):
    """Compact autocorrelation C of the pupil, unit-total normalised
    (mdft_render.cu:240-262), complex128."""
    N = pupil.shape[0]
    mc = min(int(math.ceil(fill * N)) + 4, N // 2 - 1)
    F = np.fft.fft2(pupil.astype(np.complex128))
    G = np.fft.fft2(np.abs(F) ** 2)
    idx = (np.arange(-mc, mc + 1)) % N
    C = np.conj(G[np.ix_(idx, idx)])
    return C / (C[mc, mc].real * N * N), mc


def mdft_window(C, mc, spec_w, Sx, Sy, px_mm_x, px_mm_y, src, width, height, knorm_gain, lambda_ref=550.0):
    """Reference MDFT formula for one source (mdft_render.cu:264-330) in FP64:
    returns the (3, H, W) contribution."""
    px, py, r, g, b = src
    half_x = max(1, int(math.floor(0.5 * Sx / px_mm_x)))
    half_y = max(1, int(math.floor(0.5 * Sy / px_mm_y)))
    kx0 = max(0, int(math.floor(px)) - half_x); kx1 = min(width, int(math.floor(px)) + half_x + 1)
    ky0 = max(0, int(math.floor(py)) - half_y); ky1 = min(height, int(math.floor(py)) + half_y + 1)
    out = np.zeros((3, height, width))
    if kx1 <= kx0 or ky1 <= ky0:
        return out
    m = np.arange(-mc, mc + 1, dtype=np.float64)
    xr = ((np.arange(kx0, kx1) + 0.5) - px) * px_mm_x
    yr = ((np.arange(ky0, ky1) + 0.5) - py) * px_mm_y
    win = np.zeros((3, ky1 - ky0, kx1 - kx0))
    for lam, cr, cg, cb in spec_w:
        s = lam / lambda_ref
        k2 = (lambda_ref / lam) ** 2
        Sxl, Syl = Sx * s, Sy * s
        Px = np.sinc(m * px_mm_x / Sxl)[:, None] * np.exp(-2j * np.pi * np.outer(m, xr) / Sxl)
        Py = (np.sinc(m * px_mm_y / Syl)[:, None] * np.exp(-2j * np.pi * np.outer(m, yr) / Syl)).T
        R = np.real(Py @ (C @ Px))
        win[0] += cr * k2 * R; win[1] += cg * k2 * R; win[2] += cb * k2 * R
    for ch, col in enumerate((r, g, b)):
        out[ch, ky0:ky1, kx0:kx1] += win[ch] * knorm_gain * col
    return out


def veil_pad(W, H):
    P = 2
    while P < 2 * max(W, H):
        P *= 2
    return P


def veil_kernel(P, px_mm_x, px_mm_y, a_mm, p):
    """Unit-sum g(r) = (a^2/(r^2+a^2))^p at integer pixel offsets, origin at
    [0, 0] with negative offsets wrapped (FP64)."""
    d = np.arange(P)
    d = np.where(d < P // 2, d, d - P).astype(np.float64)
    rx, ry = d * px_mm_x, d * px_mm_y
    g = (a_mm ** 2 / (rx[None, :] ** 2 + ry[:, None] ** 2 + a_mm ** 2)) ** p
    return g / g.sum()


def veil_convolve(planes, px_mm_x, px_mm_y, a_mm, p, gain):
    """Linear convolution of each plane with veil_kernel over the zero-padded
    P x P grid (numpy FFT, complex128)."""
    H, W = planes[0].shape
    P = veil_pad(W, H)
    G = np.fft.fft2(veil_kernel(P, px_mm_x, px_mm_y, a_mm, p))
    out = []
    for x in planes:
        z = np.zeros((P, P))
        z[:H, :W] = x
        out.append(gain * np.real(np.fft.ifft2(np.fft.fft2(z) * G))[:H, :W])
    return out


def veil_convolve_direct(plane, px_mm_x, px_mm_y, a_mm, p, gain):
    """Brute-force sum over every input pixel (tiny images only)."""
    H, W = plane.shape
    P = veil_pad(W, H)
    g = veil_kernel(P, px_mm_x, px_mm_y, a_mm, p)
    out = np.zeros((H, W))
    for y in range(H):
        for x in range(W):
            if plane[y, x] != 0.0:
                dy = (np.arange(H)[:, None] - y) % P
                dx = (np.arange(W)[None, :] - x) % P
                out += plane[y, x] * g[dy, dx]
    return gain * out
