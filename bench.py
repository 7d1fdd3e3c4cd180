"""Benchmark: ghost-ray segments/s on the batched 24-surface workload.

Workload (one step = one ghost render, everything resident on the device):
seeded 24-surface wide-angle lens with 2 cemented doublets and 3-layer TMM AR
coatings, 276 ghost pairs pre-filtered then probed / culled / concentrated,
256x256 pupil grid, 31 wavelengths, 1920x1080 FP32 RGB output, one point
source at (12 deg, 4 deg).  A step runs the probe passes, the trace + splat
kernel and the spectral correction; the image stays in HBM (no D2H).

Prints ONE JSON line.  `value` = executed ghost-ray segments per second
(segment = one surface interaction reaching intersect_surface, counted exactly
by the device for the same render).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_VECTOR_PEAK_TFLOPS = 157.3   # MI355X published FP32 vector peak
# Algorithmic FP32 ops per traced segment on this workload (adds, multiplies,
# divides, square roots and compares of the tracer's arithmetic), counted by
# instrumenting the FP32 restatement of the trace on a sample of its rays;
# tests/test_bench_constants.py recounts it.
OPS_PER_SEGMENT = 116.08
# HBM bytes per launch of the trace kernel on this workload, from rocprofv3
# FETCH_SIZE / WRITE_SIZE runs scaled by an in-place 4096^2 complex64
# transpose (128 MiB each way): 4.8 MiB read (lens, wavelength records,
# pupil lattice, work list) + 849 MiB written (the 55 M deposit records of
# 16 B).
TRACE_TRAFFIC_BYTES = 8.95e8
HBM_PEAK_GBS = 8000.0


def workload():
    from ghostlight_amd import FlareConfig, synthetic
    lens = synthetic.config2_lens()
    W, H = 1920, 1080
    hw, hh = synthetic.sensor_for(lens, W, H, half_field_deg=20.0)
    cfg = FlareConfig(ray_grid=256, spectral_samples=31)
    src = np.array([[math.radians(12.0), math.radians(4.0), 1.0, 1.0, 1.0]], np.float32)
    return lens, cfg, src, W, H, hw, hh


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def host_threads() -> int:
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


def ops_per_segment(lens, pairs, cfg, src, n_pairs_sample=12, n_rays=256):
    """Algorithmic FP32 ops per traced segment, counted by the oracle's
    counting scalar on a sample of this workload's rays (the count behind
    OPS_PER_SEGMENT; called by the test that keeps the constant honest)."""
    import ctypes as C
    from oracle import pyoracle
    from ghostlight_amd import hostprep
    surfs, arena = lens.pack()
    n = lens.num_surfaces()
    spec = hostprep.spectral_table(cfg.spectral_samples)
    rng = np.random.default_rng(0)
    fR, z0 = lens.surfaces[0].semi_aperture, lens.surfaces[0].z - 20.0
    bx, by = math.tan(float(src[0, 0])), math.tan(float(src[0, 1]))
    d = np.array([bx, by, 1.0]); d /= np.linalg.norm(d)
    tot = np.zeros(3, np.uint64)
    idx = np.linspace(0, len(pairs) - 1, min(n_pairs_sample, len(pairs))).astype(int)
    for i in idx:
        a, b = pairs[i]
        uv = rng.uniform(-1, 1, (n_rays, 2))
        uv = uv[(uv ** 2).sum(1) <= 1]
        lam = spec[rng.integers(0, len(spec), len(uv)), 0]
        rays = np.zeros((len(uv), 7), np.float32)
        rays[:, 0] = uv[:, 0] * fR - 20.0 * bx
        rays[:, 1] = uv[:, 1] * fR - 20.0 * by
        rays[:, 2] = z0
        rays[:, 3:6] = d
        rays[:, 6] = lam
        sp = np.zeros((len(rays), 4), np.float32); sp[:, 0] = lam
        at, bf = pyoracle.ior_tables(surfs, n, sp)
        out = np.zeros(3, np.uint64)
        pyoracle.lib().orc_count_ops(surfs, n, pyoracle._f(arena), a, b, pyoracle._f(rays), len(rays),
                                     pyoracle._f(at), pyoracle._f(bf), out.ctypes.data_as(C.POINTER(C.c_uint64)))
        tot += out
    return float(tot[0]) / max(1.0, float(tot[1]))


def cpu_baseline(lens, cfg, src, pairs, boosts, W, H, hw, hh, target_s=15.0):
    """The FP32 oracle (OpenMP over all host cores) on a bounded sample of
    the same workload: the same lens, grid, spectra and source, on a subset
    of the ghost pairs sized to ~target_s seconds."""
    from oracle import pyoracle
    from ghostlight_amd import hostprep
    from ghostlight_amd.config import to_params
    spec = hostprep.spectral_table(cfg.spectral_samples)
    C = hostprep.spectral_correction(spec)
    p, _keep = to_params(cfg)
    threads = host_threads()
    k = 2
    t0 = time.perf_counter()
    _, info = pyoracle.render_ghosts(lens, pairs[:k], boosts[:k], src, spec, p, hw, hh, W, H, correction=C,
                                     threads=threads)
    dt = time.perf_counter() - t0
    k2 = int(min(len(pairs), max(k, k * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    _, info = pyoracle.render_ghosts(lens, pairs[:k2], boosts[:k2], src, spec, p, hw, hh, W, H, correction=C,
                                     threads=threads)
    dt = time.perf_counter() - t0
    return {"value": info["segments"] / dt, "unit": "segments/s", "cores": threads, "kind": "port",
            "sample": f"FP32 oracle render of the first {k2} of {len(pairs)} ghost pairs of the same workload "
                      f"({info['segments']} segments in {dt:.2f} s, {cpu_model()})"}


def precision_sample(lens, cfg, src, pairs, boosts, W, H, hw, hh, k, ctx):
    """Error of the device render against the FP32 and FP64 oracles on the
    first k pairs of the workload (the same launcher semantics on the CPU)."""
    from oracle import pyoracle
    from ghostlight_amd import hostprep, render_ghosts
    from ghostlight_amd.config import to_params
    spec = hostprep.spectral_table(cfg.spectral_samples)
    C = hostprep.spectral_correction(spec)
    p, _keep = to_params(cfg)
    gpu = render_ghosts(lens, src, W, H, cfg, hw, hh, pairs=pairs[:k], boosts=boosts[:k], ctx=ctx)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    r32, _ = pyoracle.render_ghosts(lens, pairs[:k], boosts[:k], src, spec, p, hw, hh, W, H, correction=C,
                                    threads=threads)
    r64, _ = pyoracle.render_ghosts(lens, pairs[:k], boosts[:k], src, spec, p, hw, hh, W, H, correction=C,
                                    dbl=True, threads=threads)
    return {"rel_l2_vs_fp32_oracle": pyoracle.rel_l2(gpu, r32), "rel_l2_vs_fp64_oracle": pyoracle.rel_l2(gpu, r64),
            "fp32_oracle_vs_fp64_oracle": pyoracle.rel_l2(r32, r64),
            "sample": f"ghost image of the first {k} of {len(pairs)} pairs of the same workload"}


def veil_pad(W, H):
    """Side of the zero-padded FFT grid of the veil convolution: the least
    power of two >= 2 max(W, H) (linear, not circular, convolution)."""
    P = 2
    while P < 2 * max(W, H):
        P *= 2
    return P


def veil_pass_bytes(W, H):
    """HBM bytes the six-pass veil pipeline streams for a W x H, 3-plane
    image: per plane pair, pass 1 reads the planes and writes the H x P row
    transforms (A), pass 2 reads A and writes the P x P grid (B), passes 3
    and 4 read and write B (pass 3 also reads the P x P float kernel), pass 5
    reads H columns of B and writes A, pass 6 reads A and writes the planes."""
    P = veil_pad(W, H)
    A, B, G, plane = H * P * 8, P * P * 8, P * P * 4, W * H * 4
    return sum(5 * A + 5 * B + G + 2 * k * plane for k in (2, 1)), P


def fft_rooflines(sizes=(1024, 2048, 4096), reps=20):
    """In-place N x N forward transforms (rows, in-place transpose, rows,
    in-place transpose), HIP-event timed; each pass reads and writes the
    N^2 complex64 array once, 4 passes = 64 N^2 bytes."""
    from ghostlight_amd import diffraction as dfr
    out = {}
    for N in sizes:
        ms = dfr.fft2d_time(N, reps)
        b = 64 * N * N
        gbs = b / (ms * 1e-3) / 1e9
        out[str(N)] = {"ms": round(ms, 4), "bytes": b, "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)}
    return out


def cpu_baseline_starburst(P, M, q, lams, ghost_planes, hw, hh):
    """NumPy on the host cores, the forms the reference's validation uses:
    the matrix DFT as two complex64 matmuls per lambda (threaded BLAS), and
    the veil convolution as complex64 fft2 / ifft2 of each zero-padded plane
    (numpy's FFT is single-threaded)."""
    from ghostlight_amd import FlareConfig
    N = P.shape[0]
    n = np.arange(N) - N / 2
    k = np.arange(M) - M / 2
    P64 = P.astype(np.complex64)
    t0 = time.perf_counter()
    for lam in lams:
        A = np.exp(-2j * np.pi * np.outer(k, n) / (q * float(lam) / 550.0 * N)).astype(np.complex64)
        E = A @ P64 @ A.T
        _ = (E.real * E.real + E.imag * E.imag)
    mft_s = time.perf_counter() - t0
    H, W = ghost_planes[0].shape
    Pp = veil_pad(W, H)
    d = FlareConfig().diffraction
    px, py, a = 2 * hw / W, 2 * hh / H, d.veil_spread * hh
    t0 = time.perf_counter()
    dd = np.arange(Pp)
    dd = np.where(dd < Pp // 2, dd, dd - Pp).astype(np.float32)
    g = (a * a / ((dd[None, :] * px) ** 2 + (dd[:, None] * py) ** 2 + a * a)) ** d.veil_falloff
    G = np.fft.fft2((g / g.sum()).astype(np.complex64))
    for x in ghost_planes:
        z = np.zeros((Pp, Pp), np.complex64)
        z[:H, :W] = x
        _ = d.veil_gain * np.real(np.fft.ifft2(np.fft.fft2(z) * G))[:H, :W]
    veil_s = time.perf_counter() - t0
    return {"mft": {"value": len(lams) / mft_s, "unit": "PSF/s", "ms_per_lambda": round(mft_s * 1e3 / len(lams), 2),
                    "cores": host_threads(), "kind": "port",
                    "sample": f"numpy complex64 matrix DFT {N}^2 -> {M}^2 (two matmuls), all {len(lams)} lambda, "
                              f"{cpu_model()}"},
            "veil": {"value": 1.0 / veil_s, "unit": "images/s", "ms": round(veil_s * 1e3, 1), "cores": 1,
                     "kind": "port",
                     "sample": f"numpy complex64 fft2 convolution of the {W}x{H} 3-plane image on the {Pp}^2 grid"}}


def precision_starburst_leg(ghost_planes, W, H, hw, hh, with_cpu=True):
    """Matrix-DFT of the 1024^2 precision pupil (9-blade rounded polygon,
    Perlin dust, 20 scratches) to a 2048^2 PSF at 31 wavelengths with
    lambda-scaled sampling, FP32 MFMA and 3xBF16 MFMA; per-lambda error vs
    the FP64 (numpy complex128) matrix DFT at every lambda; then the FFT veil
    convolution of the 1920x1080 ghost image (warm: the kernel transform and
    buffers of the first call are reused) with its HBM roofline, the LDS row
    FFT's HBM roofline at sprite sizes, and the numpy CPU baselines."""
    from ghostlight_amd import FlareConfig, diffraction as dfr, hostprep, synthetic
    if with_cpu:
        from oracle import diffraction_oracle as dor
    N, M, q = 1024, 2048, 2.0
    P = synthetic.precision_pupil(N, seed=3)
    lams = np.array(hostprep.build_spectral_lambdas(31), np.float32)
    flops = 8.0 * len(lams) * (N * N * M + M * M * N)
    out = {"pupil": f"{N}x{N} complex64, 9-blade rounded polygon, Perlin dust 0.9-1.0, 20 scratches",
           "psf": f"{M}x{M} per lambda, {len(lams)} lambda, q = {q} x lambda / 550 nm", "gflop": round(flops / 1e9, 1)}
    for name, prec, peak in (("fp32", dfr.MDFT_FP32, 157.3), ("bf16x3", dfr.MDFT_BF16X3, 2500.0)):
        dfr.mft_psf(P, M, q, lams[:2], precision=prec)                 # warm the buffers
        psf, ms = dfr.mft_psf(P, M, q, lams, precision=prec, timing=True)
        tf = flops / (ms * 1e-3) / 1e12
        issued = tf * (3.0 if prec == dfr.MDFT_BF16X3 else 1.0)       # hi*hi + hi*lo + lo*hi products
        out[name] = {"ms": round(ms, 3), "tflops": round(tf, 2), "mfma_tflops_issued": round(issued, 2),
                     "peak": peak, "frac": round(issued / peak, 4)}
        if with_cpu:
            errs = {}
            for i, lam in enumerate(lams):
                ref = np.abs(dor.mft(P, M, q * float(lam) / 550.0)) ** 2
                errs[f"{float(lam):.1f}"] = float(dor.rel_l2(psf[i], ref))
                del ref
            out[name].update(rel_l2_vs_fp64=errs, rel_l2_vs_fp64_max=max(errs.values()))
            print(f"bench: MDFT {name} {ms:.3f} ms, worst rel-L2 vs FP64 over {len(lams)} lambda "
                  f"{max(errs.values()):.2e}", file=sys.stderr, flush=True)
        del psf
    cfg = FlareConfig()
    dfr.veil_convolve(ghost_planes, hw, hh, cfg)                       # kernel transform + buffers
    vms = min(dfr.veil_convolve(ghost_planes, hw, hh, cfg, timing=True)[1] for _ in range(3))
    vb, P2 = veil_pass_bytes(W, H)
    vgbs = vb / (vms * 1e-3) / 1e9
    out["veil_fft"] = {"image": f"{W}x{H} ghost image, 3 planes", "fft": f"{P2}x{P2}", "ms": round(vms, 3),
                       "roofline": {"bound": "hbm", "achieved": round(vgbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": round(vgbs / HBM_PEAK_GBS, 4), "bytes": vb}}
    if with_cpu:
        vo = dfr.veil_convolve(ghost_planes, hw, hh, cfg)
        d = cfg.diffraction
        vref = dor.veil_convolve([np.asarray(x, np.float64) for x in ghost_planes], 2 * hw / W, 2 * hh / H,
                                 d.veil_spread * hh, d.veil_falloff, d.veil_gain)
        out["veil_fft"]["rel_l2_vs_fp64"] = float(dor.rel_l2(np.stack(vo), np.stack(vref)))
    out["fft_rows"] = fft_rooflines()
    print(f"bench: veil {vms:.3f} ms ({vgbs:.0f} GB/s on {vb / 1e9:.2f} GB), FFT {out['fft_rows']}",
          file=sys.stderr, flush=True)
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_starburst(P, M, q, lams, ghost_planes, hw, hh)
    return out


def plain_flare_leg(with_cpu=True, reps=5):
    """The plain CPU-runnable flare: the 12-surface double-Gauss-like lens,
    all two-bounce ghosts, 64^2 pupil grid, 8 wavelengths, a point source at
    10 deg, 512^2 output, plus the 256^2 6-blade starburst; one render through
    the public path (ghosts + starburst layer), median of `reps` after a
    warm-up.  CPU baseline: the FP32 oracle's render of the ghost layer."""
    from ghostlight_amd import Context, PointFlareConfig, synthetic
    from ghostlight_amd.render import filter_ghost_pairs, render_sources
    lens = synthetic.config1_lens()
    cal = lens.calibration()
    W = H = 512
    cfg = PointFlareConfig(ray_grid=64, spectral_samples=8, aperture_blades=6, collect_stats=True)
    cfg.diffraction.starburst = True
    cfg.diffraction.starburst_grid = 256
    cfg.diffraction.starburst_grid_cap = 256
    src = np.array([[math.radians(10.0), 0.0, 1.0, 1.0, 1.0]], np.float32)
    ctx = Context()
    out = render_sources(src, W, H, lens, cal, cfg, ctx=ctx)
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        out = render_sources(src, W, H, lens, cal, cfg, ctx=ctx)
        ts.append((time.perf_counter() - t0) * 1e3)
    ms = float(np.median(ts))
    segs = int(out["stats"]["segments"])
    res = {"lens": "12 surfaces, 6-blade stop, MgF2", "output": f"{W}x{H}", "ray_grid": 64, "lambda": 8,
           "starburst_grid": 256, "render_ms": round(ms, 3), "segments": segs,
           "segments_per_s": segs / (ms * 1e-3)}
    if with_cpu:
        from oracle import pyoracle
        from ghostlight_amd import hostprep
        from ghostlight_amd.config import to_params
        hw, hh = cal.sensor_half_w, cal.sensor_half_h
        pairs, boosts = filter_ghost_pairs(lens, hw, hh, cfg, ctx=ctx)
        spec = hostprep.spectral_table(cfg.spectral_samples)
        Cm = hostprep.spectral_correction(spec)
        p, _keep = to_params(cfg)
        threads = host_threads()
        cts = []
        for _ in range(3):
            t0 = time.perf_counter()
            o32, info = pyoracle.render_ghosts(lens, pairs, boosts, src, spec, p, hw, hh, W, H, correction=Cm,
                                               threads=threads)
            cts.append(time.perf_counter() - t0)
        o64, _ = pyoracle.render_ghosts(lens, pairs, boosts, src, spec, p, hw, hh, W, H, correction=Cm, dbl=True,
                                        threads=threads)
        got = [out["ghost_r"], out["ghost_g"], out["ghost_b"]]
        dt = float(np.median(cts))
        res["cpu_baseline"] = {"value": info["segments"] / dt, "unit": "segments/s", "cores": threads, "kind": "port",
                               "sample": f"FP32 oracle render of the whole ghost layer ({len(pairs)} pairs, "
                                         f"{info['segments']} segments in {dt:.3f} s, {cpu_model()})"}
        res["ghost_rel_l2_vs_fp32_oracle"] = pyoracle.rel_l2(got, o32)
        res["ghost_rel_l2_vs_fp64"] = pyoracle.rel_l2(got, o64)
    ctx.close()
    return res


def sweep_leg(n_sources=120, with_cpu=True):
    """Anamorphic source sweep: the 30-surface lens with 4 cylindrical
    surfaces (1.5x squeeze), seeded decenter/tilt, 192^2 pupil, 16 lambda,
    sources on a diagonal, each into its own 2048x858 frame, one launch."""
    import dataclasses
    from ghostlight_amd import Context, FlareConfig, filter_ghost_pairs, render_ghosts, synthetic
    lens = synthetic.config4_lens()
    W, H = 2048, 858
    hw, hh = synthetic.sensor_for(lens, W, H, half_field_deg=12.0)
    t = np.radians(np.linspace(-14.0, 14.0, n_sources))
    src = np.stack([t, t * (H / W), np.ones(n_sources), np.ones(n_sources), np.ones(n_sources)], 1).astype(np.float32)
    # Schedule candidate 1 (all 16 wavelengths in one work item, so a pass of
    # 64 lanes holds four samples x 16 wavelengths; eight workgroups per CU):
    # the fastest candidate on this lens, fixed so that the first render needs
    # no autotuning trials.
    cfg = FlareConfig(ray_grid=192, spectral_samples=16, frame_per_source=True, batch_mode=1, schedule=1)
    ctx = Context()
    pairs, boosts = filter_ghost_pairs(lens, hw, hh, cfg, ctx=ctx)
    t0 = time.perf_counter()
    render_ghosts(lens, src, W, H, cfg, hw, hh, pairs=pairs, boosts=boosts, ctx=ctx, device_out=True)
    ctx.synchronize()
    first = time.perf_counter() - t0
    t0 = time.perf_counter()
    render_ghosts(lens, src, W, H, cfg, hw, hh, pairs=pairs, boosts=boosts, ctx=ctx, device_out=True)
    ctx.synchronize()
    dt = time.perf_counter() - t0
    kms = ctx.last_kernel_ms()
    _, st = render_ghosts(lens, src, W, H, cfg, hw, hh, pairs=pairs, boosts=boosts, ctx=ctx, device_out=True,
                          stats=True)
    ctx.close()
    out = {"sources": n_sources, "frame": f"{W}x{H}", "pairs_prefiltered": len(pairs), "render_ms": round(dt * 1e3, 2),
           "first_render_ms": round(first * 1e3, 2), "trace_kernel_ms": round(kms, 2),
           "segments": st["segments"], "segments_per_s": st["segments"] / dt, "frames_per_s": n_sources / dt,
           "output_bytes": 3 * 4 * W * H * n_sources}
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline_sweep(lens, cfg, src, pairs, boosts, W, H, hw, hh)
    return out


def cpu_baseline_sweep(lens, cfg, src, pairs, boosts, W, H, hw, hh, target_s=10.0):
    """The FP32 oracle (OpenMP over the host cores) on a bounded sample of
    the sweep: 4 of its sources, the first k of its pairs, sized to about
    target_s seconds; segments/s."""
    from oracle import pyoracle
    from ghostlight_amd import hostprep
    from ghostlight_amd.config import to_params
    spec = hostprep.spectral_table(cfg.spectral_samples)
    C = hostprep.spectral_correction(spec)
    p, _keep = to_params(cfg)
    threads = host_threads()
    s4 = src[np.linspace(0, len(src) - 1, 4).astype(int)]
    k = 2
    t0 = time.perf_counter()
    pyoracle.render_ghosts(lens, pairs[:k], boosts[:k], s4, spec, p, hw, hh, W, H, correction=C, threads=threads)
    dt = time.perf_counter() - t0
    k2 = int(min(len(pairs), max(k, k * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    _, info = pyoracle.render_ghosts(lens, pairs[:k2], boosts[:k2], s4, spec, p, hw, hh, W, H, correction=C,
                                     threads=threads)
    dt = time.perf_counter() - t0
    return {"value": info["segments"] / dt, "unit": "segments/s", "cores": threads, "kind": "port",
            "sample": f"FP32 oracle render of 4 of the {len(src)} sources x the first {k2} of {len(pairs)} pairs "
                      f"({info['segments']} segments in {dt:.2f} s, {cpu_model()})"}


def psf_grid_leg():
    """Geometric PSF grid at the validation shape: 7 x 7 field points on the
    12-surface lens, 96^2 tiles, 256^2 pupil grid, 16 lambda, 4 um Gaussian
    splat, chief-centroid anchoring."""
    from ghostlight_amd import PSFConfig, synthetic
    from ghostlight_amd.psf import render_psf
    lens = synthetic.config1_lens()
    cfg = PSFConfig(grid_nx=7, grid_ny=7, tile_w=96, tile_h=96, tile_extent_mm=0.08, ray_grid=256,
                    spectral_samples=16, splat_sigma_um=4.0)
    a = np.radians(np.linspace(-10.0, 10.0, 7))
    AX, AY = np.meshgrid(a, a)
    src = np.stack([AX.ravel(), AY.ravel(), np.ones(49), np.ones(49), np.ones(49)], 1).astype(np.float32)
    render_psf(src, lens, cfg)                                     # warm
    t0 = time.perf_counter()
    out = render_psf(src, lens, cfg, timing=True)
    wall = (time.perf_counter() - t0) * 1e3
    from ghostlight_amd import hostprep
    n_grid = len(hostprep.aperture_samples(lens, 256))
    traces = 49 * n_grid * 16
    return {"sources": 49, "tile": "96x96", "pupil_samples": n_grid, "lambda": 16, "splat_sigma_um": 4.0,
            "kernel_ms": round(out["kernel_ms"], 3), "render_ms": round(wall, 3),
            "traces_per_s": traces / (out["kernel_ms"] * 1e-3), "dark_cells": int((out["status"] == 2).sum())}


def served_ghost_error(req, out):
    """rel-L2 of a served request's ghost layer against the FP64 and FP32
    oracles' renders of the same lens, pair set, source and spectra, and the
    FP32 oracle's own distance from the FP64 one."""
    from oracle import pyoracle
    from ghostlight_amd import hostprep
    from ghostlight_amd.config import to_params
    from ghostlight_amd.render import filter_ghost_pairs, source_rgb
    from ghostlight_amd.source_map import solve_source_map
    lens, cfg = req.lens, req.config
    cal = lens.calibration()
    hw = cfg.sensor_half_w if cfg.sensor_half_w > 0 else cal.sensor_half_w
    hh = cfg.sensor_half_h if cfg.sensor_half_h > 0 else cal.sensor_half_h
    smap = solve_source_map(lens, cal, (cfg.source_x - 0.5) * 2.0, (cfg.source_y - 0.5) * 2.0, hw, hh)
    rgb = source_rgb(cfg)
    src = np.array([[np.float32(smap.angle_x), np.float32(smap.angle_y), rgb[0], rgb[1], rgb[2]]], np.float32)
    pairs, boosts = filter_ghost_pairs(lens, hw, hh, cfg)
    spec = hostprep.spectral_table(cfg.spectral_samples)
    p, _keep = to_params(cfg)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    C = hostprep.spectral_correction(spec)
    o64, _ = pyoracle.render_ghosts(lens, pairs, boosts, src, spec, p, hw, hh, req.width, req.height,
                                    correction=C, dbl=True, threads=threads)
    o32, _ = pyoracle.render_ghosts(lens, pairs, boosts, src, spec, p, hw, hh, req.width, req.height,
                                    correction=C, threads=threads)
    got = [out["ghost_r"], out["ghost_g"], out["ghost_b"]]
    return pyoracle.rel_l2(got, o64), pyoracle.rel_l2(got, o32), pyoracle.rel_l2(o32, o64)


def serving_leg(n_requests, rate_hz=40.0, slots=4):
    """Mixed-request serving stream (seeded lenses of 8-32 surfaces, 512^2 to
    3840x2160, 8-31 wavelengths, 5-11 blades; ghosts + starburst + veil at
    the default 1024^2 starburst grid).  The ghost layers of a sample of the
    served requests (the first of each output size) are checked against the
    FP64 oracle."""
    from ghostlight_amd.serving import (FlareServer, MixSpec, make_mix, run_one_at_a_time, run_stream,
                                        serving_lens_pool)
    pool = serving_lens_pool(5)
    mix = make_mix(MixSpec(n_requests=n_requests, rate_hz=rate_hz, seed=5), pool)
    warm = make_mix(MixSpec(n_requests=2 * slots, rate_hz=1e6, seed=99), pool)
    burst = [(0.0, req) for _, req in mix]
    sample = {}
    for i, (_, req) in enumerate(mix):
        sample.setdefault((req.width, req.height), i)
    keep = sorted(sample.values())
    with FlareServer(slots=slots) as srv:
        run_stream(srv, warm)
        served = run_stream(srv, mix, keep=keep)
        saturated = run_stream(srv, burst)
    outs = served.pop("outputs")
    errors = []
    for i in keep:
        req = mix[i][1]
        e64, e32, base = served_ghost_error(req, outs[i])
        errors.append({"request": i, "size": f"{req.width}x{req.height}", "surfaces": req.lens.num_surfaces(),
                       "lambda": req.config.spectral_samples, "ghost_rel_l2_vs_fp64": float(e64),
                       "ghost_rel_l2_vs_fp32_oracle": float(e32), "fp32_oracle_vs_fp64": float(base)})
    del outs
    plain = run_one_at_a_time(mix)
    r = lambda d: {k: (round(v, 3) if isinstance(v, float) else v) for k, v in d.items()}  # noqa: E731
    return {"requests": n_requests, "rate_hz": rate_hz, "slots": slots, "starburst_grid": 1024,
            "served": r(served), "served_all_at_once": r(saturated), "one_at_a_time": r(plain),
            "served_error": errors, "served_error_max": max(x["ghost_rel_l2_vs_fp64"] for x in errors)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true",
                    help="skip the CPU baselines and every FP64 / FP32 oracle error check (device timing only)")
    ap.add_argument("--no-starburst", action="store_true", help="skip the precision-starburst and PSF legs")
    ap.add_argument("--sweep-sources", type=int, default=120, help="sources in the sweep leg (0 skips it)")
    ap.add_argument("--serving-requests", type=int, default=200,
                    help="requests in the serving leg (0 skips it)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              f"(multi-GPU sharding is a design note only)", file=sys.stderr)
        sys.exit(2)

    from ghostlight_amd import Context, filter_ghost_pairs, render_ghosts
    lens, cfg, src, W, H, hw, hh = workload()
    ctx = Context()
    pairs, boosts = filter_ghost_pairs(lens, hw, hh, cfg, ctx=ctx)
    print(f"bench: C2 lens {lens.num_surfaces()} surfaces, {len(pairs)} pre-filtered pairs, "
          f"ray_grid {cfg.ray_grid}, {cfg.spectral_samples} lambda, {W}x{H}", file=sys.stderr, flush=True)

    def step():
        return render_ghosts(lens, src, W, H, cfg, hw, hh, pairs=pairs, boosts=boosts, ctx=ctx, device_out=True)

    for _ in range(a.warmup):
        step()
    # The first renders of a workload shape try the launch-schedule
    # candidates; warm up until the autotuner has settled.
    extra = 0
    while ctx.last_schedule()["tuning"] and extra < 16:
        step()
        extra += 1
    ctx.synchronize()
    sched = ctx.last_schedule()
    kms, sms = [], []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
        kms.append(ctx.last_kernel_ms())
        sms.append(ctx.last_splat_ms())
    ctx.synchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / a.steps
    # Exact segment count and survivor telemetry of the same render (untimed).
    _, st = render_ghosts(lens, src, W, H, cfg, hw, hh, pairs=pairs, boosts=boosts, ctx=ctx, device_out=True,
                          stats=True)
    segs = st["segments"]
    live_pairs = sum(1 for x in st["pair_on_sensor"] if x > 0)
    value = segs / (ms_per_step * 1e-3)
    graph_kernel_ms = float(np.mean(kms))
    splat_ms = float(np.mean(sms))
    # Roofline timing: the same kernel with the tuned schedule issued eagerly,
    # HIP events around the launch on the context's stream.
    import dataclasses
    cfg_ev = dataclasses.replace(cfg, batch_mode=1, schedule=sched["candidate"])
    kev = []
    for _ in range(3):
        render_ghosts(lens, src, W, H, cfg_ev, hw, hh, pairs=pairs, boosts=boosts, ctx=ctx, device_out=True)
        kev.append(ctx.last_kernel_ms())
    kernel_ms = float(np.mean(kev))
    print(f"bench: schedule {sched}, {extra} autotune warmup renders", file=sys.stderr, flush=True)
    print(f"bench: {segs} segments/step, {ms_per_step:.3f} ms/step, trace kernel {kernel_ms:.3f} ms "
          f"(graph replay wall clock {graph_kernel_ms:.3f} ms), "
          f"tiled splat {splat_ms:.3f} ms, {live_pairs} pairs reach the sensor, "
          f"{st.get('traces_on_sensor', 0)} on-sensor traces", file=sys.stderr, flush=True)

    sc = ctx.last_splat_counts()
    ops = OPS_PER_SEGMENT
    achieved = segs * ops / (kernel_ms * 1e-3) / 1e12
    image_bytes = 3 * 4 * W * H
    roof = {"bound": "valu", "achieved": round(achieved, 3), "peak": FP32_VECTOR_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": round(achieved / FP32_VECTOR_PEAK_TFLOPS, 4),
            "traffic": TRACE_TRAFFIC_BYTES,
            "kernel": "ghost_persist_kernel (trace + deposit records)", "kernel_ms": round(kernel_ms, 4),
            "kernel_ms_in_graph": round(graph_kernel_ms, 4), "schedule": sched,
            "ops_per_segment": round(ops, 2), "segments_per_launch": segs,
            "splat": {"resolve_ms": round(splat_ms, 4), "records": sc["records"], "binned": sc["binned"],
                      "chunks": sc["chunks"], "staged_tiles": sc["staged_tiles"],
                      "record_bytes": 16 * sc["records"], "image_bytes": image_bytes,
                      # trace writes the records; bin_count and bin_scatter read them, scatter writes the
                      # binned copy, resolve reads it; plus the image's read-add-write
                      "step_bytes": 16 * (3 * sc["records"] + 2 * sc["binned"]) + 2 * image_bytes}}

    cpu = prec = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(lens, cfg, src, pairs, boosts, W, H, hw, hh)
        except Exception as e:
            print(f"bench: cpu baseline unavailable ({e})", file=sys.stderr)
        try:
            prec = precision_sample(lens, cfg, src, pairs, boosts, W, H, hw, hh, 24, ctx)
        except Exception as e:
            print(f"bench: precision sample unavailable ({e})", file=sys.stderr)
    starburst = None
    if not a.no_starburst:
        print("bench: precision-starburst leg", file=sys.stderr, flush=True)
        ghost = render_ghosts(lens, src, W, H, cfg, hw, hh, pairs=pairs, boosts=boosts, ctx=ctx)
        starburst = precision_starburst_leg(ghost, W, H, hw, hh, with_cpu=not a.no_cpu)
    psf = None
    if not a.no_starburst:
        print("bench: PSF grid leg", file=sys.stderr, flush=True)
        psf = psf_grid_leg()
    plain = None
    if not a.no_starburst:
        print("bench: plain 12-surface flare leg", file=sys.stderr, flush=True)
        plain = plain_flare_leg(with_cpu=not a.no_cpu)
    sweep = None
    if a.sweep_sources > 0:
        print(f"bench: source-sweep leg, {a.sweep_sources} sources", file=sys.stderr, flush=True)
        sweep = sweep_leg(a.sweep_sources, with_cpu=not a.no_cpu)
    serving = None
    if a.serving_requests > 0:
        print(f"bench: serving leg, {a.serving_requests} requests", file=sys.stderr, flush=True)
        serving = serving_leg(a.serving_requests)

    out = {
        "metric": "ghost-ray segments/s",
        "value": value,
        "unit": "segments/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded lens, point source)",
        "config": {"workload": "batched ghost trace: 24-surface lens, 2 cemented doublets, 3-layer TMM AR, "
                               "276 pairs -> filter/probe/cull/concentrate, 256x256 pupil, 31 lambda, "
                               "1920x1080 FP32 RGB", "pairs_prefiltered": len(pairs),
                   "pairs_on_sensor": live_pairs, "segments_per_step": segs, "parallelism": "single GPU"},
        "roofline": roof,
        "cpu_baseline": cpu,
        "precision": prec,
        "precision_starburst": starburst,
        "psf_grid": psf,
        "plain_flare": plain,
        "source_sweep": sweep,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
