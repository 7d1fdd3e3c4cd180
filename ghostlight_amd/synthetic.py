"""Seeded synthetic lenses and workloads for the benchmark configurations.

Glasses are picked from published catalogue Sellmeier coefficients (Schott
datasheets); radii, thicknesses and glasses are drawn from a seeded
``numpy.random.Generator``.  A candidate prescription is accepted only when its
paraxial d-line design is sane (positive focal length and back focus), so each
seed deterministically yields one lens.  Semi-apertures follow the paraxial
marginal + chief ray heights; the stop sets the f-number.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from .lens import (ApertureShape, Coating, CoatingModel, CylinderAxis, DispersionModel,
                   OpticalSystem, Surface, SurfaceForm, rot_from_tilt)

# name: (B1, B2, B3, C1, C2, C3)  -- Schott catalogue Sellmeier coefficients
GLASSES = {
    "N-BAK1": (1.12365662, 0.309276848, 0.881511957, 0.00644742752, 0.0222284402, 107.297751),
    "N-SK2": (1.28189012, 0.257738258, 0.96818604, 0.0072719164, 0.0242823527, 110.377773),
    "N-SK16": (1.34317774, 0.241144399, 0.994317969, 0.00704687339, 0.0229005, 92.7508526),
    "N-SSK8": (1.44857867, 0.117965926, 1.06937528, 0.00869310149, 0.0421566593, 111.300666),
    "N-BAF10": (1.5851495, 0.143559385, 1.08521269, 0.00926681282, 0.0424489805, 105.613573),
    "N-LAK9": (1.46231905, 0.344399589, 1.15508372, 0.00724270156, 0.0243353131, 85.4686868),
    "N-LAF2": (1.80984227, 0.15729555, 1.0930037, 0.0101711622, 0.0442431765, 100.687748),
    "N-F2": (1.39757037, 0.159201403, 1.2686543, 0.00995906143, 0.0546931752, 119.248346),
    "N-SF2": (1.47343127, 0.163681849, 1.36920899, 0.0109019098, 0.0585683687, 127.404933),
    "N-SF5": (1.52481889, 0.187085527, 1.42729015, 0.011254756, 0.0588995392, 129.141675),
    "N-SF11": (1.73759695, 0.313747346, 1.89878101, 0.013188707, 0.0623068142, 155.23629),
    "N-KF9": (1.19286778, 0.0893346571, 0.920819805, 0.00839154696, 0.0404010786, 112.572446),
}


def sellmeier(coef, lam_nm: float) -> float:
    l2 = (lam_nm * 1e-3) ** 2
    B1, B2, B3, C1, C2, C3 = coef
    return math.sqrt(1 + B1 * l2 / (l2 - C1) + B2 * l2 / (l2 - C2) + B3 * l2 / (l2 - C3))


def glass_nd_vd(coef) -> Tuple[float, float]:
    nd, nF, nC = sellmeier(coef, 587.56), sellmeier(coef, 486.13), sellmeier(coef, 656.27)
    return nd, (nd - 1.0) / (nF - nC)


_GLASS_OK = [g for g, c in GLASSES.items() if 1.5 <= glass_nd_vd(c)[0] <= 1.8 and 25 <= glass_nd_vd(c)[1] <= 60]


def _n(glass: Optional[str], lam=587.56) -> float:
    return 1.0 if glass is None else sellmeier(GLASSES[glass], lam)


@dataclass
class _Elem:
    glasses: List[str]          # 1 (singlet) or 2 (cemented doublet)
    radii: List[float]          # len(glasses) + 1
    thick: List[float]          # centre thicknesses, len(glasses)


def _paraxial(radii, media, gaps, lam=587.56):
    """Paraxial trace (y, nu) of a ray from infinity at height 1 through the
    surfaces; gaps[i] = distance after surface i (last = back focus, unused).
    Returns (efl, bfd, heights)."""
    y, u = 1.0, 0.0
    n = 1.0
    ys = []
    for i, R in enumerate(radii):
        n2 = media[i]
        c = 0.0 if abs(R) < 1e-9 else 1.0 / R
        phi = (n2 - n) * c
        nu = n * u - y * phi
        u = nu / n2
        n = n2
        ys.append(y)
        if i < len(radii) - 1:
            y = y + u * gaps[i]
    if abs(u) < 1e-12:
        return math.inf, math.inf, ys
    efl = -1.0 / u
    bfd = -y / u
    return efl, bfd, ys


def _chief_heights(radii, media, gaps, stop_idx, angle):
    """Paraxial chief ray through the stop centre at field `angle` (rad):
    solve the entrance height with two trial rays."""
    def trace(y0):
        y, u, n = y0, math.tan(angle), 1.0
        ys = []
        for i, R in enumerate(radii):
            n2 = media[i]
            c = 0.0 if abs(R) < 1e-9 else 1.0 / R
            nu = n * u - y * (n2 - n) * c
            u = nu / n2
            n = n2
            ys.append(y)
            if i < len(radii) - 1:
                y = y + u * gaps[i]
        return ys
    a, b = trace(0.0), trace(1.0)
    denom = b[stop_idx] - a[stop_idx]
    y0 = -a[stop_idx] / denom if abs(denom) > 1e-12 else 0.0
    return trace(y0)


def make_lens(seed: int, n_glass_surfaces: int, n_doublets: int = 0, stop_after_elem: Optional[int] = None,
              blades: int = 6, f_number: float = 2.8, coating_layers: int = 1, tmm_layers: int = 0,
              n_cyl: int = 0, squeeze: float = 1.0, decenter_max: float = 0.0, tilt_max_deg: float = 0.0,
              max_field_deg: float = 20.0, name: str = "") -> OpticalSystem:
    """Seeded spherical lens of `n_glass_surfaces` glass surfaces (+1 stop).

    Radii |R| ~ U[20, 80] mm, centre thickness ~ U[1, 8] mm, air gaps
    ~ U[1, 8] mm, glasses with n_d in [1.5, 1.8] and V_d in [25, 60].  The
    stop is a flat air-air surface after element `stop_after_elem` (default:
    the middle), a `blades`-gon (0 = circle)."""
    n_singlets = (n_glass_surfaces - 3 * n_doublets)
    if n_singlets < 0 or n_singlets % 2:
        raise ValueError("n_glass_surfaces must equal 3*doublets + 2*singlets")
    n_singlets //= 2
    n_elem = n_singlets + n_doublets
    if stop_after_elem is None:
        stop_after_elem = n_elem // 2
    for attempt in range(2000):
        rng = np.random.default_rng([seed, attempt])
        kinds = ["s"] * n_singlets + ["d"] * n_doublets
        rng.shuffle(kinds)
        elems = []
        for k in kinds:
            ng = 1 if k == "s" else 2
            gl = [str(rng.choice(_GLASS_OK)) for _ in range(ng)]
            # Alternate meniscus orientation about the stop (double-Gauss-like).
            sign = 1.0 if len(elems) < stop_after_elem else -1.0
            mags = rng.uniform(20.0, 80.0, ng + 1)
            signs = np.where(rng.random(ng + 1) < 0.7, sign, -sign)
            elems.append(_Elem(gl, list(mags * signs), list(rng.uniform(1.0, 8.0, ng))))
        gaps_air = list(rng.uniform(1.0, 8.0, n_elem))
        radii, media, gaps, stop_idx = [], [], [], -1
        for ei, e in enumerate(elems):
            for j, R in enumerate(e.radii):
                radii.append(R)
                media.append(_n(e.glasses[j]) if j < len(e.glasses) else 1.0)
                gaps.append(e.thick[j] if j < len(e.glasses) else gaps_air[ei])
            if ei == stop_after_elem - 1:
                radii.append(0.0); media.append(1.0); gaps.append(float(rng.uniform(2.0, 6.0)))
                stop_idx = len(radii) - 1
        efl, bfd, ys = _paraxial(radii, media, gaps)
        if not (15.0 < efl < 150.0 and 8.0 < bfd < 120.0):
            continue
        if max(abs(y) for y in ys) > 2.5:
            continue
        # Entrance pupil radius -> marginal scale; stop radius from it.
        ep = efl / (2.0 * f_number)
        marg = [abs(y) * ep for y in ys]
        chief = [abs(y) for y in _chief_heights(radii, media, gaps, stop_idx, math.radians(max_field_deg))]
        semi = []
        ok = True
        for i, R in enumerate(radii):
            h = marg[i] + chief[i]
            if i == stop_idx:
                semi.append(marg[i])
                continue
            sa = max(3.0, 1.12 * h)
            if abs(R) > 1e-9 and sa > 0.92 * abs(R):
                ok = False
                break
            semi.append(sa)
        if not ok:
            continue
        break
    else:
        raise RuntimeError(f"no valid lens for seed {seed}")

    surfs: List[Surface] = []
    gi = 0
    glass_seq = []
    for ei, e in enumerate(elems):
        for j in range(len(e.radii)):
            glass_seq.append(e.glasses[j] if j < len(e.glasses) else None)
        if ei == stop_after_elem - 1:
            glass_seq.append(None)
    cyl_left = n_cyl
    for i, R in enumerate(radii):
        g = glass_seq[i]
        s = Surface(radius=float(R), thickness=float(gaps[i] if i < len(radii) - 1 else bfd),
                    semi_aperture=float(semi[i]))
        if g is None:
            s.ior, s.abbe_v, s.disp_model = 1.0, 0.0, DispersionModel.AIR
        else:
            c = GLASSES[g]
            nd, vd = glass_nd_vd(c)
            s.ior, s.abbe_v, s.disp_model = float(np.float32(nd)), float(np.float32(vd)), DispersionModel.SELLMEIER
            s.sellmeier_B, s.sellmeier_C = tuple(float(x) for x in c[:3]), tuple(float(x) for x in c[3:])
        if i == stop_idx:
            s.is_stop = True
            if blades >= 3:
                s.aperture_shape = ApertureShape.POLYGON
                s.aperture_blades = blades
                s.aperture_rotation_rad = float(np.float32(math.radians(90.0 / blades)))
            if squeeze != 1.0:
                s.aperture_aspect = float(1.0 / squeeze)
        else:
            s.coating = Coating(model=CoatingModel.SIMPLE, ar_layers=coating_layers)
            if cyl_left > 0 and i % 3 == 1:
                s.form = SurfaceForm.CYLINDRICAL
                s.cyl_axis = CylinderAxis.Y
                cyl_left -= 1
        surfs.append(s)
    lens = OpticalSystem(surfs, name=name or f"synthetic-{n_glass_surfaces}s-seed{seed}")
    lens.focal_length = float(efl)
    lens.entrance_pupil_semi = float(ep)
    lens.finalize()
    if decenter_max > 0 or tilt_max_deg > 0:
        rng = np.random.default_rng([seed, 7777])
        for s in lens.surfaces:
            if s.is_stop:
                continue
            s.decenter_x = float(rng.uniform(-decenter_max, decenter_max))
            s.decenter_y = float(rng.uniform(-decenter_max, decenter_max))
            s.rot = rot_from_tilt(rng.uniform(-tilt_max_deg, tilt_max_deg), rng.uniform(-tilt_max_deg, tilt_max_deg))
    if tmm_layers > 0:
        # Quarter-wave style AR stack: MgF2 / ZrO2 / Al2O3-like layers with
        # constant indices (nk tables with one row per layer).
        stack_n = [1.38, 2.05, 1.63, 1.46][:tmm_layers]
        layers = [{"material": f"L{k}", "thickness_nm": float(550.0 / (4.0 * n)),
                   "nk_table": np.array([[0.4, n, 0.0], [0.7, n, 0.0]], np.float32)}
                  for k, n in enumerate(stack_n)]
        for i, s in enumerate(lens.surfaces):
            if not s.is_stop:
                lens.set_coating_layers(i, layers)
    return lens


def calibration_for(lens: OpticalSystem, width: int, height: int, half_field_deg: float = 14.0):
    """First-order calibration record of a synthetic lens (paraxial focal
    length and entrance pupil; the sensor covers `half_field_deg`)."""
    from .lens import LensCalibration
    hw, hh = sensor_for(lens, width, height, half_field_deg)
    f = lens.focal_length
    ep = getattr(lens, "entrance_pupil_semi", f / 5.6)
    return LensCalibration(sensor_half_w=hw, sensor_half_h=hh, max_half_angle_h=math.atan(hw / f),
                           max_half_angle_v=math.atan(hh / f), image_circle_semi_w=hw, image_circle_semi_h=hh,
                           focal_length_x=f, focal_length_y=f, entrance_pupil_semi_x=ep, entrance_pupil_semi_y=ep,
                           f_number_x=f / (2 * ep), f_number_y=f / (2 * ep), pupil_area_frac=1.0)


def calibration_from_paraxial(lens: OpticalSystem, sensor_half_w: float = 18.0, sensor_half_h: float = 12.0):
    """First-order calibration of any lens at the d line: paraxial focal
    length from a unit-height ray from infinity, entrance pupil = stop
    semi-aperture over that ray's height at the stop, and a full-frame sensor."""
    from .lens import LensCalibration
    y, u, n = 1.0, 0.0, 1.0
    y_stop = None
    surfs = [s for s in lens.surfaces if s.is_active or s.is_stop]
    for i, s in enumerate(surfs):
        n2 = s.ior_at(587.56) if s.is_active else n
        c = 0.0 if abs(s.radius) < 1e-9 else 1.0 / s.radius
        if s.is_stop and y_stop is None:
            y_stop = abs(y)
        u = (n * u - y * (n2 - n) * c) / n2
        n = n2
        if i + 1 < len(surfs):
            y = y + u * (surfs[i + 1].z - s.z)
    if abs(u) < 1e-12:
        raise ValueError("calibration: afocal lens (no paraxial focal length)")
    f = -1.0 / u
    stop = next((s for s in lens.surfaces if s.is_stop), None)
    ep = stop.semi_aperture / y_stop if (stop is not None and y_stop) else lens.surfaces[0].semi_aperture
    return LensCalibration(sensor_half_w=sensor_half_w, sensor_half_h=sensor_half_h,
                           max_half_angle_h=math.atan(sensor_half_w / f), max_half_angle_v=math.atan(sensor_half_h / f),
                           image_circle_semi_w=sensor_half_w, image_circle_semi_h=sensor_half_h,
                           focal_length_x=f, focal_length_y=f, entrance_pupil_semi_x=ep, entrance_pupil_semi_y=ep,
                           f_number_x=f / (2 * ep), f_number_y=f / (2 * ep), pupil_area_frac=1.0)


def sensor_for(lens: OpticalSystem, width: int, height: int, half_field_deg: float = 14.0):
    """Sensor half extents (mm) so the frame's half-height covers `half_field_deg`."""
    hh = lens.focal_length * math.tan(math.radians(half_field_deg))
    return hh * width / height, hh


# ---- named workloads --------------------------------------------------------

def config1_lens() -> OpticalSystem:
    """C1: 12 glass surfaces, 6 elements, stop after element 3 (index 6),
    6-blade polygon stop, single-layer MgF2 coating."""
    return make_lens(seed=1, n_glass_surfaces=12, n_doublets=0, stop_after_elem=3, blades=6,
                     coating_layers=1, name="C1-double-gauss-12")


def config2_lens() -> OpticalSystem:
    """C2: 24 glass surfaces with 2 cemented doublets, 3-layer TMM AR coatings."""
    return make_lens(seed=2, n_glass_surfaces=24, n_doublets=2, blades=7, coating_layers=0,
                     tmm_layers=3, max_field_deg=30.0, name="C2-wide-24")


def config4_lens() -> OpticalSystem:
    """C4: 30 glass surfaces, 4 cylindrical (1.5x squeeze), decenter <= 0.2 mm,
    tilt <= 0.3 deg."""
    return make_lens(seed=4, n_glass_surfaces=30, n_doublets=2, blades=9, coating_layers=1,
                     n_cyl=4, squeeze=1.5, decenter_max=0.2, tilt_max_deg=0.3, name="C4-anamorphic-30")


# ---- precision-starburst pupil ----------------------------------------------

def _perlin(shape, cells, rng):
    """2-D gradient noise (K. Perlin, "Improving Noise", SIGGRAPH 2002:
    random unit gradients on a lattice, quintic fade 6t^5 - 15t^4 + 10t^3),
    values roughly in [-0.7, 0.7]."""
    H, W = shape
    ang = rng.uniform(0.0, 2.0 * np.pi, (cells + 1, cells + 1))
    gx, gy = np.cos(ang), np.sin(ang)
    y = np.linspace(0.0, cells, H, endpoint=False)
    x = np.linspace(0.0, cells, W, endpoint=False)
    X, Y = np.meshgrid(x, y)
    x0, y0 = np.floor(X).astype(int), np.floor(Y).astype(int)
    fx, fy = X - x0, Y - y0
    fade = lambda t: t * t * t * (t * (t * 6.0 - 15.0) + 10.0)  # noqa: E731
    def dot(ix, iy, dx, dy):
        return gx[iy, ix] * dx + gy[iy, ix] * dy
    n00 = dot(x0, y0, fx, fy)
    n10 = dot(x0 + 1, y0, fx - 1.0, fy)
    n01 = dot(x0, y0 + 1, fx, fy - 1.0)
    n11 = dot(x0 + 1, y0 + 1, fx - 1.0, fy - 1.0)
    u, v = fade(fx), fade(fy)
    return (n00 * (1 - u) + n10 * u) * (1 - v) + (n01 * (1 - u) + n11 * u) * v


def rounded_polygon_mask(N, blades, curvature, rotation=0.0, fill=0.9, ss=4):
    """ss x ss supersampled coverage of an n-blade stop whose edges bow
    outward: per half-sector 1/r = A cos(phi - phi_w) + C with 1/r(waist) =
    1/r_w and r(tip) = 1, r_w = apothem + curvature (1 - apothem) (curvature
    0: the regular polygon, 1: the circle)."""
    apo = math.cos(math.pi / blades)
    r_w = apo + curvature * (1.0 - apo)
    half = math.pi / blades
    A = (1.0 / r_w - 1.0) / (1.0 - math.cos(half)) if half > 0 else 0.0
    Cc = 1.0 / r_w - A
    acc = np.zeros((N, N))
    i = np.arange(N, dtype=np.float64)
    for sy in range(ss):
        v = ((i + (sy + 0.5) / ss) / N * 2.0 - 1.0) / fill
        for sx in range(ss):
            u = ((i + (sx + 0.5) / ss) / N * 2.0 - 1.0) / fill
            U, V = np.meshgrid(u, v)
            r = np.hypot(U, V)
            phi = np.mod(np.arctan2(V, U) - rotation, 2.0 * half)   # 0 at a tip, waist at `half`
            rb = 1.0 / (A * np.cos(phi - half) + Cc)
            acc += r <= rb
    return acc / (ss * ss)


def precision_pupil(N=1024, blades=9, curvature=0.35, seed=3, n_scratches=20):
    """The precision-starburst pupil: a 9-blade rounded polygon, seeded
    Perlin dust (transmission 0.9..1.0) and thin scratch lines
    (transmission 0.2), as complex64."""
    rng = np.random.default_rng(seed)
    mask = rounded_polygon_mask(N, blades, curvature, rotation=float(rng.uniform(0, 2 * math.pi / blades)))
    n = _perlin((N, N), 24, rng) + 0.5 * _perlin((N, N), 96, rng)
    n = (n - n.min()) / max(n.max() - n.min(), 1e-12)
    dust = 0.9 + 0.1 * n
    scr = np.ones((N, N))
    yy, xx = np.mgrid[0:N, 0:N].astype(np.float64)
    for _ in range(n_scratches):
        x0, y0 = rng.uniform(0.2 * N, 0.8 * N, 2)
        th = rng.uniform(0, math.pi)
        L = rng.uniform(0.05 * N, 0.3 * N)
        w = rng.uniform(0.6, 1.8)
        dx, dy = math.cos(th), math.sin(th)
        t = (xx - x0) * dx + (yy - y0) * dy
        d = np.abs(-(xx - x0) * dy + (yy - y0) * dx)
        hit = (np.abs(t) <= L / 2) & (d <= w / 2)
        scr[hit] = 0.2
    return (mask * dust * scr).astype(np.complex64)
