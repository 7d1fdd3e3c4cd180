"""Prescription types: Coating, Surface, OpticalSystem, LensCalibration.

Attribute names, enum values and defaults follow the reference's bound types
(ghostlight/src/optical_system.h:86-249, ghostlight/src/lens_calibration.h:11-57,
ghostlight/bindings/python/src/bind_geometry.cpp:144-520).  The prescription is
packed into the C ABI's ``gl_surface`` array plus one float arena for
table-backed coatings (``OpticalSystem.pack``).
"""
from __future__ import annotations

import enum
import math
import threading
from dataclasses import dataclass, field
from typing import List, Optional

import ctypes as C

import numpy as np

from . import _abi

# Serialises the creation of device lenses (OpticalSystem.device_lens).
_DEVICE_LENS_LOCK = threading.Lock()


class SurfaceForm(enum.IntEnum):
    SPHERE = 0
    ASPHERE = 1
    CYLINDRICAL = 2


class CylinderAxis(enum.IntEnum):
    X = 0
    Y = 1


class ApertureShape(enum.IntEnum):
    CIRCLE = 0
    POLYGON = 1
    IMAGE = 2


class DispersionModel(enum.IntEnum):
    AIR = 0
    ABBE = 1
    SELLMEIER = 2


class CoatingModel(enum.IntEnum):
    SIMPLE = 0
    SPECTRAL = 1
    ANGULAR = 2
    SPECTRAL_ANGULAR = 3
    ATTENUATOR_GAUSS = 4
    ARTIST = 5


class TraceStatus(enum.IntEnum):
    OK = 0
    VIGNETTED = 1
    TIR = 2
    MISSED_SURFACE = 3
    INVALID_INPUT = 4


@dataclass
class Coating:
    """Per-surface coating descriptor (optical_system.h:109-147)."""
    model: CoatingModel = CoatingModel.SIMPLE
    ar_layers: int = 0
    out_of_range_discard: bool = False
    gauss_sigma: float = 0.0
    gauss_background: float = 0.0
    gauss_peak: float = 0.0
    gauss_decenter_x: float = 0.0
    gauss_decenter_y: float = 0.0
    angle_ref_ior: float = 1.0
    tint_r: float = 1.0
    tint_g: float = 1.0
    tint_b: float = 1.0
    tint_strength: float = 0.04
    # table storage (CoatingTables, optical_system.h:289-301)
    table: Optional[np.ndarray] = None        # (K, 2) key, r
    sa_wavelengths: Optional[np.ndarray] = None
    sa_angles: Optional[np.ndarray] = None
    sa_r: Optional[np.ndarray] = None         # (nw, na)


_IDENT = (1.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0)


@dataclass
class Vec3f:
    """World-space point or direction (Vec3f, bind_geometry.cpp:30-37)."""
    x: float = 0.0
    y: float = 0.0
    z: float = 0.0

    def __iter__(self):
        return iter((self.x, self.y, self.z))


def dot(a, b) -> float:
    """Dot product of two 3-vectors (Vec3f or any 3-sequence), float32 arithmetic."""
    ax, ay, az = (np.float32(v) for v in a)
    bx, by, bz = (np.float32(v) for v in b)
    return float(ax * bx + ay * by + az * bz)


def cross(a, b) -> Vec3f:
    """Cross product of two 3-vectors as a Vec3f, float32 arithmetic."""
    ax, ay, az = (np.float32(v) for v in a)
    bx, by, bz = (np.float32(v) for v in b)
    return Vec3f(float(ay * bz - az * by), float(az * bx - ax * bz), float(ax * by - ay * bx))


class ApertureProfile:
    """Read-only derived blade geometry of a surface's aperture
    (ApertureProfile, bind_geometry.cpp:121-141), built by the host library's
    gl_blade_profile_build from the surface's authored blade controls;
    radius_at(theta) is the boundary radius as a fraction of the tip radius."""

    def __init__(self, s: "Surface"):
        from . import _abi
        self._lib = _abi.load_library()
        self._p = _abi.gl_blade_profile()
        blades = int(s.aperture_blades) if int(s.aperture_shape) == int(ApertureShape.POLYGON) else 0
        _abi.check(self._lib.gl_blade_profile_build(blades, float(s.aperture_rotation_rad),
                                                    float(s.aperture_curvature), float(s.aperture_twist),
                                                    float(s.aperture_notch_rad), float(s.aperture_notch_angle_rad),
                                                    _abi.C.byref(self._p)), "aperture_profile")

    def __getattr__(self, name):
        if name.startswith("_"):
            raise AttributeError(name)
        return getattr(self._p, name)

    def deformed(self) -> bool:
        return self._p.blades >= 3 and not self._p.plain

    def radius_at(self, theta: float) -> float:
        from . import _abi
        return float(self._lib.gl_blade_radius_at(_abi.C.byref(self._p), float(theta)))

    def dr_dtheta(self, theta: float) -> float:
        from . import _abi
        return float(self._lib.gl_blade_dr_dtheta(_abi.C.byref(self._p), float(theta)))


@dataclass
class Surface:
    """One optical surface (optical_system.h:157-249)."""
    radius: float = 0.0
    thickness: float = 0.0
    ior: float = 1.0
    abbe_v: float = 0.0
    semi_aperture: float = 10.0
    coating: Coating = field(default_factory=Coating)
    is_stop: bool = False
    is_active: bool = True
    z: float = 0.0
    aperture_shape: ApertureShape = ApertureShape.CIRCLE
    aperture_blades: int = 0
    aperture_rotation_rad: float = 0.0
    aperture_aspect: float = 1.0
    aperture_semi_diameter: float = 0.0
    # Blade-shape controls of a polygonal stop (optical_system.h:196-205);
    # only the plain polygon (all zero) is rendered by this build.
    aperture_curvature: float = 0.0
    aperture_twist: float = 0.0
    aperture_notch_rad: float = 0.0
    aperture_notch_angle_rad: float = 0.0
    decenter_x: float = 0.0
    decenter_y: float = 0.0
    rot: tuple = _IDENT
    form: SurfaceForm = SurfaceForm.SPHERE
    conic_k: float = 0.0
    asphere_terms: tuple = ()
    cyl_axis: CylinderAxis = CylinderAxis.X
    disp_model: DispersionModel = DispersionModel.ABBE
    sellmeier_B: tuple = (0.0, 0.0, 0.0)
    sellmeier_C: tuple = (0.0, 0.0, 0.0)

    @property
    def aperture_profile(self) -> ApertureProfile:
        """Blade geometry derived from the authored controls, rebuilt on each
        read so it always follows the current fields."""
        return ApertureProfile(self)

    @property
    def n_asphere_terms(self) -> int:
        return len(self.asphere_terms)

    def ior_at(self, lambda_nm: float) -> float:
        """Wavelength-dependent index (optical_system.h:230-235), float32 arithmetic."""
        f = np.float32
        lam = f(lambda_nm)
        if self.disp_model == DispersionModel.SELLMEIER:
            B = [f(b) for b in self.sellmeier_B]
            Cc = [f(c) for c in self.sellmeier_C]
            l = lam * f(1e-3)
            l2 = l * l
            n2 = f(1.0)
            for i in range(3):
                n2 = f(n2 + f(B[i] * l2) / f(l2 - Cc[i]))
            return float(np.sqrt(n2)) if n2 > 0 else 1.0
        n_d, V = f(self.ior), f(self.abbe_v)
        if V < f(0.1) or n_d <= f(1.0001):
            return float(n_d)
        dn = (n_d - f(1)) / V
        iF = f(1) / (f(486.13) * f(486.13))
        iC = f(1) / (f(656.27) * f(656.27))
        idl = f(1) / (f(587.56) * f(587.56))
        B = dn / (iF - iC)
        A = n_d - B * idl
        return float(A + B / (lam * lam))


@dataclass
class LensCalibration:
    """Calibration record (lens_calibration.h:11-57)."""
    sensor_half_w: float = 18.0
    sensor_half_h: float = 12.0
    max_half_angle_h: float = 0.0
    max_half_angle_v: float = 0.0
    image_circle_semi_w: float = 0.0
    image_circle_semi_h: float = 0.0
    focal_length_x: float = 0.0
    focal_length_y: float = 0.0
    entrance_pupil_semi_x: float = 0.0
    entrance_pupil_semi_y: float = 0.0
    f_number_x: float = 0.0
    f_number_y: float = 0.0
    pupil_area_frac: float = 1.0

    FIELDS = ("sensor_half_w", "sensor_half_h", "max_half_angle_h", "max_half_angle_v",
              "image_circle_semi_w", "image_circle_semi_h", "focal_length_x", "focal_length_y",
              "entrance_pupil_semi_x", "entrance_pupil_semi_y", "f_number_x", "f_number_y",
              "pupil_area_frac")

    @classmethod
    def from_array(cls, a):
        return cls(**{k: float(v) for k, v in zip(cls.FIELDS, a)})


@dataclass
class ApertureImage:
    """Bitmap behind an IMAGE aperture (ApertureImage, optical_system.h:262-269):
    row-major [0, 1] luminance, height x width; empty pixels = not loaded (the
    surface then passes every ray).  The loader fills source_path and
    semi_diameter; the caller decodes the file and sets the pixels."""
    width: int = 0
    height: int = 0
    pixels: Optional[np.ndarray] = None
    semi_diameter: float = 0.0
    source_path: str = ""

    def __setattr__(self, name, value):
        # assigning a 2-D array to `pixels` sets the size too, as the
        # reference's pixels property does (bind_geometry.cpp:327-351)
        if name == "pixels" and isinstance(value, np.ndarray) and value.ndim == 2:
            self.set_pixels(value)
            return
        object.__setattr__(self, name, value)

    def set_pixels(self, img) -> None:
        a = np.ascontiguousarray(img, np.float32)
        if a.ndim != 2:
            raise ValueError("aperture image must be a 2-D (height, width) array")
        self.height, self.width = a.shape
        self.pixels = a.ravel().copy()


class OpticalSystem:
    """Flat surface sequence with the sensor at z = 0 (optical_system.h:305-416)."""

    def __init__(self, surfaces: Optional[List[Surface]] = None, name: str = ""):
        self.name = name
        self.focal_length = 0.0
        self.surfaces: List[Surface] = list(surfaces or [])
        # one per surface, parallel to `surfaces` (OpticalSystem::aperture_images)
        self.aperture_images: List[ApertureImage] = [ApertureImage() for _ in self.surfaces]
        self._dev = None        # cached device handle (gl_lens*) keyed by pack digest
        self._dev_key = None

    def num_surfaces(self) -> int:
        return len(self.surfaces)

    def finalize(self) -> bool:
        """Lay the vertices from the thickness chain so the chain ends at
        z = 0 (OpticalSystem::finalize, optical_system.cpp:1427-1453): total =
        sum of thicknesses, z starts at -total and advances by each
        thickness, float32 throughout."""
        if not self.surfaces:
            return False
        while len(self.aperture_images) < len(self.surfaces):
            self.aperture_images.append(ApertureImage())
        total = np.float32(0.0)
        for s in self.surfaces:
            total = np.float32(total + np.float32(s.thickness))
        z = np.float32(-total)
        for s in self.surfaces:
            s.z = float(z)
            z = np.float32(z + np.float32(s.thickness))
        return True

    @classmethod
    def load(cls, path) -> "OpticalSystem":
        """Load a ``ghostlight-optical`` JSON lens file (optical_system.py:89-101)."""
        from .lens_file import load_lens
        return load_lens(path)

    def ior_before(self, idx: int) -> float:
        idx = min(idx, len(self.surfaces))
        for k in range(idx - 1, -1, -1):
            if self.surfaces[k].is_active:
                return self.surfaces[k].ior
        return 1.0

    def ior_before_at(self, idx: int, lambda_nm: float) -> float:
        idx = min(idx, len(self.surfaces))
        for k in range(idx - 1, -1, -1):
            if self.surfaces[k].is_active:
                return self.surfaces[k].ior_at(lambda_nm)
        return 1.0

    # ---- coatings ---------------------------------------------------------
    def set_coating_layers(self, i: int, layers) -> None:
        """Bake a thin-film stack to a SPECTRAL_ANGULAR table (31 x 18) with the
        transfer-matrix method (OpticalSystem::bake_coating_layers,
        optical_system.cpp:1483-1570).  ``layers``: list of dicts with
        ``thickness_nm`` and ``nk_table`` (K, 3) of (lambda_um, n, k)."""
        from . import hostprep
        wl = np.arange(31, dtype=np.float32) * np.float32(10.0) + np.float32(400.0)
        n_amb = np.array([self.ior_before_at(i, float(l)) for l in wl], np.float32)
        n_sub = np.array([self.surfaces[i].ior_at(float(l)) for l in wl], np.float32)
        n_re = np.zeros((len(layers), 31), np.float32)
        n_im = np.zeros((len(layers), 31), np.float32)
        thick = np.zeros(len(layers), np.float32)
        for li, d in enumerate(layers):
            nk = np.asarray(d["nk_table"], np.float32)
            thick[li] = d["thickness_nm"]
            for w, lam in enumerate(wl):
                n_re[li, w], n_im[li, w] = _interp_nk(nk, np.float32(lam * np.float32(1e-3)))
        table = hostprep.bake_tmm_table(n_re, n_im, thick, n_amb, n_sub)
        c = self.surfaces[i].coating
        c.model = CoatingModel.SPECTRAL_ANGULAR
        c.out_of_range_discard = False
        c.angle_ref_ior = self.ior_before_at(i, 587.56)
        c.sa_wavelengths = table[:31].copy()
        c.sa_angles = table[31:49].copy()
        c.sa_r = table[49:].reshape(31, 18).copy()

    # ---- packing for the C ABI -------------------------------------------
    def pack(self):
        """Return (gl_surface array, float32 arena)."""
        n = len(self.surfaces)
        if n > _abi.GL_MAX_SURFACES:
            raise ValueError(f"lens has {n} surfaces; at most {_abi.GL_MAX_SURFACES} supported")
        arr = (_abi.gl_surface * max(n, 1))()
        # The arena as float32 chunks, joined once at the end; `used` is its
        # running length (the offset of the next chunk).
        chunks: List[np.ndarray] = []
        used = 0

        def put(a) -> int:
            nonlocal used
            a = np.asarray(a, np.float32).ravel()
            off = used
            chunks.append(a)
            used += a.size
            return off
        stop_index = next((i for i, s in enumerate(self.surfaces) if s.is_stop), -1)
        for i, s in enumerate(self.surfaces):
            g = arr[i]
            g.radius, g.thickness, g.ior, g.abbe_v = s.radius, s.thickness, s.ior, s.abbe_v
            g.semi_aperture, g.z = s.semi_aperture, s.z
            g.is_stop, g.is_active = int(bool(s.is_stop)), int(bool(s.is_active))
            g.aperture_shape, g.aperture_blades = int(s.aperture_shape), int(s.aperture_blades)
            g.aperture_rotation_rad, g.aperture_aspect = s.aperture_rotation_rad, s.aperture_aspect
            g.aperture_curvature, g.aperture_twist = s.aperture_curvature, s.aperture_twist
            g.aperture_notch_rad, g.aperture_notch_angle_rad = s.aperture_notch_rad, s.aperture_notch_angle_rad
            g.decenter_x, g.decenter_y = s.decenter_x, s.decenter_y
            for k in range(9):
                g.rot[k] = s.rot[k]
            g.form, g.conic_k = int(s.form), s.conic_k
            at = list(s.asphere_terms)[: _abi.GL_MAX_ASPHERE_TERMS]
            for k, t in enumerate(at):
                g.asphere_terms[k] = t
            g.n_asphere_terms = len(at)
            g.cyl_axis, g.disp_model = int(s.cyl_axis), int(s.disp_model)
            for k in range(3):
                g.sellmeier_B[k], g.sellmeier_C[k] = s.sellmeier_B[k], s.sellmeier_C[k]
            c = s.coating
            g.coating_model, g.ar_layers = int(c.model), int(c.ar_layers)
            g.out_of_range_discard, g.angle_ref_ior = int(bool(c.out_of_range_discard)), c.angle_ref_ior
            g.tint_r, g.tint_g, g.tint_b, g.tint_strength = c.tint_r, c.tint_g, c.tint_b, c.tint_strength
            g.gauss_sigma, g.gauss_background, g.gauss_peak = c.gauss_sigma, c.gauss_background, c.gauss_peak
            g.gauss_decenter_x, g.gauss_decenter_y = c.gauss_decenter_x, c.gauss_decenter_y
            g.table_offset = g.table_count = g.sa_offset = g.sa_n_wavelengths = g.sa_n_angles = 0
            if c.table is not None and len(c.table):
                t = np.asarray(c.table, np.float32).reshape(-1, 2)
                g.table_count = len(t)
                g.table_offset = put(t)
            if c.sa_r is not None:
                wl = np.asarray(c.sa_wavelengths, np.float32).ravel()
                an = np.asarray(c.sa_angles, np.float32).ravel()
                r = np.asarray(c.sa_r, np.float32).reshape(len(wl), len(an))
                g.sa_n_wavelengths, g.sa_n_angles = len(wl), len(an)
                g.sa_offset = put(wl)
                put(an)
                put(r)
            g.aperture_semi_diameter = float(s.aperture_semi_diameter)
            g.image_offset = g.image_width = g.image_height = g.image_sdf = g.image_sdf_offset = 0
            img = self.aperture_images[i] if i < len(self.aperture_images) else None
            if (s.aperture_shape == ApertureShape.IMAGE and img is not None and img.pixels is not None
                    and len(img.pixels)):
                px = np.asarray(img.pixels, np.float32).ravel()
                if img.width <= 0 or img.height <= 0 or len(px) != img.width * img.height:
                    raise ValueError(f"surface {i}: aperture image dimensions do not match its pixel data")
                g.image_width, g.image_height = int(img.width), int(img.height)
                g.image_offset = put(px)
                if s.is_stop and i == stop_index:
                    # the stop's bitmap also carries its edge distance field for HURB
                    # (find_sdf_target_surface / bake_aperture_sdf, aperture_sdf.cpp:94-189)
                    sdf = np.zeros(4 * len(px), np.float32)
                    if s.aperture_semi_diameter > 0 and _abi.load_library().gl_bake_aperture_sdf(
                            _abi.fptr(px), int(img.width), int(img.height), float(s.aperture_semi_diameter),
                            float(s.aperture_aspect), _abi.fptr(sdf)) == 0:
                        g.image_sdf = 1
                        g.image_sdf_offset = put(sdf)
        return arr, (np.concatenate(chunks) if used else np.zeros(1, np.float32))

    @classmethod
    def from_packed(cls, surfaces, arena=None, name: str = "") -> "OpticalSystem":
        """The inverse of pack(): a lens from the C ABI's gl_surface array
        (a ctypes array or its raw bytes) and coating arena."""
        if isinstance(surfaces, (bytes, bytearray, memoryview, np.ndarray)):
            raw = bytes(np.asarray(surfaces, np.uint8)) if isinstance(surfaces, np.ndarray) else bytes(surfaces)
            n = len(raw) // C.sizeof(_abi.gl_surface)
            surfaces = (_abi.gl_surface * n).from_buffer_copy(raw[: n * C.sizeof(_abi.gl_surface)])
        ar = np.asarray(arena if arena is not None else [0.0], np.float32)
        out = []
        for g in surfaces:
            c = Coating(model=CoatingModel(g.coating_model), ar_layers=int(g.ar_layers),
                        out_of_range_discard=bool(g.out_of_range_discard), gauss_sigma=g.gauss_sigma,
                        gauss_background=g.gauss_background, gauss_peak=g.gauss_peak,
                        gauss_decenter_x=g.gauss_decenter_x, gauss_decenter_y=g.gauss_decenter_y,
                        angle_ref_ior=g.angle_ref_ior, tint_r=g.tint_r, tint_g=g.tint_g, tint_b=g.tint_b,
                        tint_strength=g.tint_strength)
            if g.table_count > 0:
                c.table = ar[g.table_offset:g.table_offset + 2 * g.table_count].reshape(-1, 2).copy()
            if g.sa_n_wavelengths > 0 and g.sa_n_angles > 0:
                o, nw, na = g.sa_offset, g.sa_n_wavelengths, g.sa_n_angles
                c.sa_wavelengths = ar[o:o + nw].copy()
                c.sa_angles = ar[o + nw:o + nw + na].copy()
                c.sa_r = ar[o + nw + na:o + nw + na + nw * na].reshape(nw, na).copy()
            out.append(Surface(radius=g.radius, thickness=g.thickness, ior=g.ior, abbe_v=g.abbe_v,
                               semi_aperture=g.semi_aperture, coating=c, is_stop=bool(g.is_stop),
                               is_active=bool(g.is_active), z=g.z, aperture_shape=ApertureShape(g.aperture_shape),
                               aperture_blades=int(g.aperture_blades), aperture_rotation_rad=g.aperture_rotation_rad,
                               aperture_aspect=g.aperture_aspect, decenter_x=g.decenter_x, decenter_y=g.decenter_y,
                               aperture_curvature=g.aperture_curvature, aperture_twist=g.aperture_twist,
                               aperture_notch_rad=g.aperture_notch_rad,
                               aperture_notch_angle_rad=g.aperture_notch_angle_rad,
                               aperture_semi_diameter=g.aperture_semi_diameter,
                               rot=tuple(g.rot[k] for k in range(9)), form=SurfaceForm(g.form), conic_k=g.conic_k,
                               asphere_terms=tuple(g.asphere_terms[k] for k in range(g.n_asphere_terms)),
                               cyl_axis=CylinderAxis(g.cyl_axis), disp_model=DispersionModel(g.disp_model),
                               sellmeier_B=tuple(g.sellmeier_B[k] for k in range(3)),
                               sellmeier_C=tuple(g.sellmeier_C[k] for k in range(3))))
        lens = cls(out, name=name)
        for i, g in enumerate(surfaces):
            if g.image_width > 0 and g.image_height > 0:
                im = lens.aperture_images[i]
                im.semi_diameter = g.aperture_semi_diameter
                im.set_pixels(ar[g.image_offset:g.image_offset + g.image_width * g.image_height]
                              .reshape(g.image_height, g.image_width))
        return lens

    def pack_key(self) -> bytes:
        arr, arena = self.pack()
        import hashlib
        h = hashlib.blake2b(bytes(arr), digest_size=16)
        h.update(arena.tobytes())
        return h.digest()

    def device_lens(self):
        """Device-resident handle for this prescription (re-created on change).

        Callers keep the returned object while they use its handle: a
        replaced DeviceLens frees its device lens when the last reference
        goes.  Creation is serialised, so threads rendering the same lens
        share one DeviceLens instead of replacing each other's."""
        from . import runtime
        key = self.pack_key()
        with _DEVICE_LENS_LOCK:
            dev = self._dev
            if dev is None or self._dev_key != key:
                dev = runtime.DeviceLens(self)
                self._dev = dev
                self._dev_key = key
        return dev

    # ---- renderer wrappers (optical_system.py:334-429) -------------------
    def calibration(self, d_line_nm: float = 587.56) -> "LensCalibration":
        """The lens's calibration record (optical_system.py:334-346): the one
        attached with set_calibration, else calibrate_lens traced on the
        device, cached until the prescription changes."""
        cal = getattr(self, "_calibration", None)
        if cal is not None:
            return cal
        key = (self.pack_key(), float(d_line_nm))
        cached = getattr(self, "_calib_cache", None)
        if cached is not None and cached[0] == key:
            return cached[1]
        from .source_map import calibrate_lens
        cal = calibrate_lens(self, d_line_nm)
        self._calib_cache = (key, cal)
        return cal

    def set_calibration(self, calib: "LensCalibration") -> None:
        self._calibration = calib

    def render_point_flare(self, width: int, height: int, config, *, calib: Optional["LensCalibration"] = None,
                           ctx=None) -> dict:
        from .render import render_point_flare
        return render_point_flare(width, height, self, calib if calib is not None else self.calibration(), config,
                                  ctx=ctx)

    def render_source_flare(self, offsets, width: int, height: int, config, *,
                            calib: Optional["LensCalibration"] = None, ctx=None) -> dict:
        from .render import render_source_flare
        arr = np.ascontiguousarray(offsets, dtype=np.float32)
        if arr.ndim != 2 or arr.shape[1] != 3 or arr.shape[0] < 1:
            raise ValueError("render_source_flare: offsets must be a (N, 3) float32 array "
                             "of [d_angle_x, d_angle_y, weight] rows with N >= 1")
        return render_source_flare(arr, width, height, self, calib if calib is not None else self.calibration(),
                                   config, ctx=ctx)

    def render_psf(self, sources, config, *, weights=None, targets_mm=None, ctx=None) -> dict:
        """Geometric PSF grid (optical_system.py:431-481): sources (N, 2)
        field angles or (N, 5) rows; weights scale the source colour;
        targets_mm (N, 2) switches on FIXED_TARGET anchoring."""
        from .psf import render_psf
        s = np.asarray(sources, np.float32)
        if s.ndim != 2 or s.shape[1] not in (2, 5, 7):
            raise ValueError("render_psf: sources must be (N, 2), (N, 5) or (N, 7)")
        if s.shape[1] == 2:
            s = np.concatenate([s, np.ones((len(s), 3), np.float32)], 1)
        if weights is not None:
            w = np.asarray(weights, np.float32).reshape(-1, 1)
            s = s.copy()
            s[:, 2:5] *= w
        if targets_mm is not None:
            t = np.asarray(targets_mm, np.float32).reshape(-1, 2)
            s = np.concatenate([s[:, :5], t], 1)
        return render_psf(s, self, config, ctx=ctx)

    def __repr__(self) -> str:
        return f"OpticalSystem(name={self.name!r}, surfaces={len(self.surfaces)})"


def _interp_nk(nk: np.ndarray, lu: np.float32):
    """Complex index of a layer at lambda (um), linear in the table
    (optical_system.cpp:1510-1529), float32 arithmetic."""
    xs, ns, ks = nk[:, 0], nk[:, 1], nk[:, 2]
    if lu <= xs[0]:
        return ns[0], ks[0]
    if lu >= xs[-1]:
        return ns[-1], ks[-1]
    for k in range(1, len(xs)):
        if lu <= xs[k]:
            span = np.float32(xs[k] - xs[k - 1])
            f = np.float32((lu - xs[k - 1]) / span) if span > 0 else np.float32(0)
            return (np.float32(ns[k - 1] + f * np.float32(ns[k] - ns[k - 1])),
                    np.float32(ks[k - 1] + f * np.float32(ks[k] - ks[k - 1])))
    return ns[-1], ks[-1]


def rot_from_tilt(tilt_x_deg: float, tilt_y_deg: float) -> tuple:
    """Row-major local->world rotation R = Ry(tilt_y) @ Rx(tilt_x)."""
    ax, ay = math.radians(tilt_x_deg), math.radians(tilt_y_deg)
    cx, sx, cy, sy = math.cos(ax), math.sin(ax), math.cos(ay), math.sin(ay)
    Rx = np.array([[1, 0, 0], [0, cx, -sx], [0, sx, cx]])
    Ry = np.array([[cy, 0, sy], [0, 1, 0], [-sy, 0, cy]])
    return tuple(float(v) for v in (Ry @ Rx).ravel())
