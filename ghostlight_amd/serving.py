"""Flare serving loop: continuous batching of point-flare renders on one GPU.

Requests (lens, output size, config) arrive at any time.  A fixed set of
slots, one render context (HIP stream + buffer cache) each, executes them; a
request is admitted as soon as a slot is free *and* the HBM it may need fits
the budget, so renders of different lenses, resolutions and wavelength counts
overlap on the device.  Lenses stay resident (one device copy per
prescription), and the first-order source mapping, colour matrices and pupil
grids are cached per context.

Admission accounting.  Each context keeps grow-only device buffers.  The
device memory in use is the sum of the contexts' caches; admitting request r
on slot s may grow s's cache to est(r).  The scheduler admits the oldest
request whose growth fits `budget - in_use`; if nothing fits, it trims the
caches of idle slots before waiting.  A request larger than the whole budget
is rejected with MemoryError.

The reference has no serving layer (its renderers serialise behind process
mutexes, ghost_render.cu:504,656); this module is build-specific.
"""
from __future__ import annotations

import math
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from .config import PointFlareConfig
from .lens import LensCalibration, OpticalSystem


@dataclass
class FlareRequest:
    lens: OpticalSystem
    width: int
    height: int
    config: PointFlareConfig
    calibration: Optional[LensCalibration] = None
    rid: int = -1
    t_submit: float = 0.0          # perf_counter() at submission (the arrival)


@dataclass
class FlareResult:
    rid: int
    output: Optional[dict]
    t_submit: float
    t_start: float
    t_done: float
    slot: int
    error: Optional[BaseException] = None

    @property
    def latency_ms(self) -> float:
        return (self.t_done - self.t_submit) * 1e3


class _Pending:
    __slots__ = ("req", "est", "event", "result")

    def __init__(self, req: FlareRequest, est: int):
        self.req = req
        self.est = est
        self.event = threading.Event()
        self.result: Optional[FlareResult] = None


def estimate_bytes(req: FlareRequest) -> int:
    """Upper estimate of the device bytes one render holds in its context:
    output planes of every pass, splat records, diffraction grids."""
    cfg = req.config
    npx = req.width * req.height
    b = 3 * 4 * npx                                   # ghost planes
    n = req.lens.num_surfaces()
    pairs = n * (n - 1) // 2
    ns = max(3, int(cfg.spectral_samples))
    grid = int(cfg.ray_grid) ** 2
    # deposit records: a fraction of (pairs x samples x lambda) lands; the
    # context sizes the buffer from the previous render with 25% headroom
    rec = min(pairs * grid * ns, 1 << 24)
    b += int(rec * 16 * 2.5)
    # fixed-point staging of the splat's multi-chunk tiles (24 KiB a 32^2 tile)
    b += min(npx // 1024 + 1, int(rec * 1.5) // 8192 + 1) * 3 * 1024 * 8
    d = cfg.diffraction
    if d.starburst:
        N = max(256, int(d.starburst_grid))
        b += N * N * (8 + 8 + 4 + 12 + 12) + 3 * 4 * npx
    if d.veil:
        b += 1024 * 1024 * 8 + 3 * 4 * npx
    return b + (16 << 20)


def _default_render(ctx, req: FlareRequest) -> dict:
    from .render import render_point_flare
    cal = req.calibration if req.calibration is not None else req.lens.calibration()
    return render_point_flare(req.width, req.height, req.lens, cal, req.config, ctx=ctx)


class FlareServer:
    """Continuous-batching flare server.

    slots          concurrent renders (contexts / HIP streams)
    hbm_budget     device bytes the server may hold; default 80% of the
                   free device memory at start
    render_fn      (ctx, request) -> output dict; the default renders
                   ghosts + enabled starburst / veil through the C ABI
    context_fn     slot index -> context object with memory_bytes() / trim()
    """

    def __init__(self, slots: int = 4, hbm_budget: Optional[int] = None,
                 render_fn: Optional[Callable] = None, context_fn: Optional[Callable] = None,
                 estimate_fn: Callable[[FlareRequest], int] = estimate_bytes, starve_s: float = 0.25):
        if slots < 1:
            raise ValueError("slots must be >= 1")
        if context_fn is None:
            from .runtime import Context
            context_fn = lambda i: Context()  # noqa: E731
        if hbm_budget is None:
            from .runtime import device_memory
            free, _total = device_memory()
            hbm_budget = int(free * 0.8)
        self.budget = int(hbm_budget)
        self.starve_s = float(starve_s)
        self.render_fn = render_fn or _default_render
        self.estimate_fn = estimate_fn
        self.ctxs = [context_fn(i) for i in range(slots)]
        self._held = [0] * slots            # bytes each slot's context holds (last reading / reservation)
        self._busy = [False] * slots
        self._queue: deque = deque()
        self._cv = threading.Condition()
        self._stop = False
        self._next_id = 0
        self.completed: List[FlareResult] = []
        self.max_in_use = 0
        self._threads = [threading.Thread(target=self._worker, args=(i,), daemon=True) for i in range(slots)]
        for t in self._threads:
            t.start()

    # ---- client side --------------------------------------------------
    def submit(self, req: FlareRequest) -> _Pending:
        est = int(self.estimate_fn(req))
        if est > self.budget:
            raise MemoryError(f"request needs ~{est} device bytes, over the {self.budget}-byte budget")
        with self._cv:
            if self._stop:
                raise RuntimeError("server is closed")
            req.rid = self._next_id
            self._next_id += 1
            req.t_submit = time.perf_counter()
            p = _Pending(req, est)
            self._queue.append(p)
            self._cv.notify_all()
        return p

    @staticmethod
    def wait(p: _Pending, timeout: Optional[float] = None) -> FlareResult:
        if not p.event.wait(timeout):
            raise TimeoutError("request did not complete in time")
        return p.result

    def close(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify_all()
        for t in self._threads:
            t.join()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---- scheduler ------------------------------------------------------
    def _in_use(self) -> int:
        return sum(self._held)

    def _pick(self, slot: int) -> Optional[_Pending]:
        """Oldest queued request whose growth of this slot's cache fits.  A
        head request that has waited longer than `starve_s` blocks the ones
        behind it, so large renders are not overtaken forever."""
        free = self.budget - self._in_use()
        head_only = (time.perf_counter() - self._queue[0].req.t_submit) > self.starve_s
        for i, p in enumerate(self._queue):
            if head_only and i > 0:
                break
            growth = max(0, p.est - self._held[slot])
            if growth <= free:
                del self._queue[i]
                return p
        return None

    def _trim_idle(self) -> bool:
        freed = False
        for j, busy in enumerate(self._busy):
            if not busy and self._held[j] > 0:
                self.ctxs[j].trim()
                self._held[j] = 0
                freed = True
        return freed

    def _worker(self, slot: int) -> None:
        ctx = self.ctxs[slot]
        while True:
            with self._cv:
                while True:
                    if self._stop and not self._queue:
                        return
                    p = self._pick(slot) if self._queue else None
                    if p is not None:
                        break
                    if self._queue and self._trim_idle():
                        continue
                    self._cv.wait(0.05)
                self._busy[slot] = True
                self._held[slot] = max(self._held[slot], p.est)
                self.max_in_use = max(self.max_in_use, self._in_use())
            t0 = time.perf_counter()
            out, err = None, None
            try:
                out = self.render_fn(ctx, p.req)
            except BaseException as e:  # reported to the caller through the result
                err = e
            t1 = time.perf_counter()
            res = FlareResult(p.req.rid, out, p.req.t_submit, t0, t1, slot, err)
            with self._cv:
                self._busy[slot] = False
                try:
                    self._held[slot] = int(ctx.memory_bytes())
                except Exception:
                    pass
                self.completed.append(res)
                self._cv.notify_all()
            p.result = res
            p.event.set()


# ---- workload: a seeded Poisson stream of mixed requests --------------------

RESOLUTIONS = [(512, 512), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]


def serving_lens_pool(seed: int = 5) -> List[OpticalSystem]:
    """Seeded lenses spanning 8 to 32 surfaces (plus the stop)."""
    from .synthetic import make_lens
    specs = [(8, 0), (12, 0), (16, 2), (20, 2), (24, 2), (28, 2), (32, 2)]
    pool = []
    for i, (ng, nd) in enumerate(specs):
        pool.append(make_lens(seed=seed * 100 + i, n_glass_surfaces=ng, n_doublets=nd, blades=7,
                              coating_layers=1, name=f"serve-{ng}s"))
    return pool


@dataclass
class MixSpec:
    rate_hz: float = 40.0
    n_requests: int = 200
    seed: int = 5
    ray_grid: int = 64
    starburst: bool = True
    veil: bool = True
    resolutions: list = field(default_factory=lambda: list(RESOLUTIONS))


def make_mix(spec: MixSpec, pool: Optional[List[OpticalSystem]] = None):
    """[(arrival_offset_s, FlareRequest)] for a seeded Poisson stream: lenses
    from the pool, output 512^2 .. 3840x2160, 8..31 wavelengths, random
    source positions, 5..11-blade apertures."""
    rng = np.random.default_rng(spec.seed)
    pool = pool or serving_lens_pool(spec.seed)
    t = 0.0
    out = []
    for _ in range(spec.n_requests):
        t += float(rng.exponential(1.0 / spec.rate_hz))
        lens = pool[int(rng.integers(len(pool)))]
        w, h = spec.resolutions[int(rng.integers(len(spec.resolutions)))]
        cfg = PointFlareConfig(ray_grid=spec.ray_grid, spectral_samples=int(rng.integers(8, 32)))
        cfg.source_x = float(rng.uniform(0.1, 0.9))
        cfg.source_y = float(rng.uniform(0.1, 0.9))
        cfg.aperture_blades = int(rng.integers(5, 12))
        cfg.diffraction.starburst = spec.starburst
        cfg.diffraction.veil = spec.veil
        out.append((t, FlareRequest(lens=lens, width=w, height=h, config=cfg)))
    return out


def percentile(xs, q):
    if not xs:
        return float("nan")
    return float(np.percentile(np.asarray(xs, np.float64), q))


def run_stream(server: FlareServer, mix, time_scale: float = 1.0, keep=()) -> Dict[str, float]:
    """Submit the mix at its arrival times and wait for every result.
    Returns throughput and latency percentiles (ms), overall and for 1080p;
    with `keep` (request indices), also "outputs": {index: output dict}."""
    t0 = time.perf_counter()
    pend = []
    for (ta, req) in mix:
        now = time.perf_counter() - t0
        if ta * time_scale > now:
            time.sleep(ta * time_scale - now)
        pend.append((req, server.submit(req)))
    results = [FlareServer.wait(p) for _, p in pend]
    t1 = time.perf_counter()
    errs = [r.error for r in results if r.error is not None]
    if errs:
        raise errs[0]
    lat = [r.latency_ms for r in results]
    lat1080 = [r.latency_ms for (req, _), r in zip(pend, results) if (req.width, req.height) == (1920, 1080)]
    span = max(r.t_done for r in results) - t0
    return {"requests": len(results), "seconds": span, "flares_per_s": len(results) / span,
            "p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
            "p50_1080p_ms": percentile(lat1080, 50), "p99_1080p_ms": percentile(lat1080, 99),
            "max_in_use_bytes": server.max_in_use, "wall_s": t1 - t0,
            **({"outputs": {i: results[i].output for i in keep}} if keep else {})}


def run_one_at_a_time(mix, render_fn=None, ctx=None) -> Dict[str, float]:
    """The plain path: the same requests rendered back to back on one
    context, each waiting for the previous one (no overlap)."""
    from .runtime import Context
    ctx = ctx or Context()
    render_fn = render_fn or _default_render
    t0 = time.perf_counter()
    service = []
    for _, req in mix:
        a = time.perf_counter()
        render_fn(ctx, req)
        service.append((time.perf_counter() - a) * 1e3)
    span = time.perf_counter() - t0
    # Latency of the same arrival process through a single FIFO server.
    clock, lat = 0.0, []
    for (ta, _), s in zip(mix, service):
        start = max(clock, ta * 1e3)
        clock = start + s
        lat.append(clock - ta * 1e3)
    return {"requests": len(mix), "seconds": span, "flares_per_s": len(mix) / span,
            "service_p50_ms": percentile(service, 50), "p50_ms": percentile(lat, 50),
            "p99_ms": percentile(lat, 99)}
