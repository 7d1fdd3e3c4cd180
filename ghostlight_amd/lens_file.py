"""The ``ghostlight-optical`` lens file loader (OpticalSystem::load,
optical_system.cpp:1299-1425, with the flatten of optical_system.cpp:48-1293).

A lens file is JSON: a glass catalogue, a list of elements (each a group of
surfaces with the materials between them and a rigid-body transform), and an
optional rig of pivots that move groups of elements.  Loading flattens it to the
sequential surface list the tracer reads, with the sensor at z = 0:

1. glasses are keyed by their catalogue key; Sellmeier glasses get n_d / V_d
   backfilled from the formula at the d, F and C lines;
2. element transforms resolve in file order (``relative_to_preceding`` adds
   the previous element's z), then each pivot rotates/translates its elements
   about their centroid (or a manual point), pivots stacking in order;
3. every element emits one surface per entry, the vertex mapped through the
   element transform; the medium after a surface is the next material, air
   after the element's last surface;
4. the air gap after each element is rewritten from the nominal (pre-tilt) z
   of the next element, and everything is shifted so the chain ends at z = 0
   under the pre-pivot layout;
5. physical thin-film stacks are baked to spectral-angular tables once the
   media are final.

All arithmetic is float32 in the reference's operation order; the rotation
builder calls the C library's cosf/sinf, as the reference does, so a loaded
lens packs to the same bytes as the reference's loader produces.
"""
from __future__ import annotations

import ctypes
import ctypes.util
import json
import os
import sys
from typing import Dict, List, Optional, Tuple

import numpy as np

from .lens import (ApertureImage, ApertureShape, Coating, CoatingModel, CylinderAxis, DispersionModel, OpticalSystem,
                   Surface, SurfaceForm)

f32 = np.float32
LENS_FORMAT_MAJOR = 1
LENS_FORMAT_MINOR = 0
MAX_ASPHERE_TERMS = 8

_libm = ctypes.CDLL(ctypes.util.find_library("m") or "libm.so.6")
_libm.cosf.restype = ctypes.c_float
_libm.cosf.argtypes = [ctypes.c_float]
_libm.sinf.restype = ctypes.c_float
_libm.sinf.argtypes = [ctypes.c_float]


def _cosf(x) -> np.float32:
    return f32(_libm.cosf(float(x)))


def _sinf(x) -> np.float32:
    return f32(_libm.sinf(float(x)))


class LensFileError(ValueError):
    """A lens file that cannot be loaded (the reference's load() == false)."""


def _warn(path, msg):
    print(f"WARNING: {path}: {msg}", file=sys.stderr)


def _fl(v) -> np.float32:
    """A JSON number as float32 (nlohmann get<float>: double, then float)."""
    return f32(float(v))


def _val(d: dict, key: str, default):
    return d[key] if isinstance(d, dict) and key in d else default


# ---- glasses ---------------------------------------------------------------

def _sellmeier_n(B, Cc, lam_nm) -> np.float32:
    l = f32(f32(lam_nm) * f32(1e-3))
    l2 = f32(l * l)
    n2 = f32(1.0)
    for i in range(3):
        n2 = f32(n2 + f32(f32(B[i] * l2) / f32(l2 - Cc[i])))
    return f32(np.sqrt(n2)) if n2 > f32(0) else f32(1.0)


def _parse_glass_catalogue(j_cat, path) -> Dict[str, tuple]:
    if not isinstance(j_cat, dict):
        raise LensFileError(f"{path}: glass_catalogue must be an object")
    out = {}
    for key, je in j_cat.items():
        jd = je.get("dispersion") if isinstance(je, dict) else None
        if not isinstance(jd, dict):
            raise LensFileError(f"{path}: glass '{key}' missing dispersion block")
        model = jd.get("model", "")
        if model == "sellmeier":
            Bj, Cj = jd.get("B"), jd.get("C")
            if not (isinstance(Bj, list) and isinstance(Cj, list) and len(Bj) == 3 and len(Cj) == 3):
                raise LensFileError(f"{path}: glass '{key}' sellmeier needs B[3] and C[3]")
            B = tuple(_fl(v) for v in Bj)
            Cc = tuple(_fl(v) for v in Cj)
            n_d = _sellmeier_n(B, Cc, 587.56)
            n_F = _sellmeier_n(B, Cc, 486.13)
            n_C = _sellmeier_n(B, Cc, 656.27)
            dFC = f32(n_F - n_C)
            V_d = f32(f32(n_d - f32(1.0)) / dFC) if dFC > f32(1e-9) else f32(0.0)
            out[key] = (DispersionModel.SELLMEIER, n_d, V_d, B, Cc)
        elif model == "abbe":
            if "nd" not in jd or "Vd" not in jd:
                raise LensFileError(f"{path}: glass '{key}' abbe needs nd + Vd")
            out[key] = (DispersionModel.ABBE, _fl(jd["nd"]), _fl(jd["Vd"]), (f32(0),) * 3, (f32(0),) * 3)
        else:
            raise LensFileError(f"{path}: glass '{key}' unknown dispersion model '{model}'")
    return out


# ---- transforms (row-major 3x3, float32) -----------------------------------

def _make_rotation(tilt_x_deg, tilt_y_deg, roll_deg) -> List[np.float32]:
    """R = Ry(tilt_y) Rx(tilt_x) Rz(roll) (optical_system.cpp:183-198)."""
    PI = f32(3.14159265358979323846)
    d = f32(PI / f32(180.0))
    tx, ty, rz = f32(f32(tilt_x_deg) * d), f32(f32(tilt_y_deg) * d), f32(f32(roll_deg) * d)
    cx, sx = _cosf(tx), _sinf(tx)
    cy, sy = _cosf(ty), _sinf(ty)
    cz, sz = _cosf(rz), _sinf(rz)
    return [f32(f32(cy * cz) + f32(f32(sy * sx) * sz)), f32(f32(-cy * sz) + f32(f32(sy * sx) * cz)), f32(sy * cx),
            f32(cx * sz), f32(cx * cz), f32(-sx),
            f32(f32(-sy * cz) + f32(f32(cy * sx) * sz)), f32(f32(sy * sz) + f32(f32(cy * sx) * cz)), f32(cy * cx)]


def _mat_vec(A, b) -> List[np.float32]:
    return [f32(f32(f32(A[3 * i] * b[0]) + f32(A[3 * i + 1] * b[1])) + f32(A[3 * i + 2] * b[2])) for i in range(3)]


def _mat_mat(A, B) -> List[np.float32]:
    out = []
    for i in range(3):
        for j in range(3):
            s = f32(0.0)
            for k in range(3):
                s = f32(s + f32(A[3 * i + k] * B[3 * k + j]))
            out.append(s)
    return out


class _Transform:
    __slots__ = ("pos", "rot", "piv_corr")

    def __init__(self):
        self.pos = [f32(0)] * 3
        self.rot = [f32(1), f32(0), f32(0), f32(0), f32(1), f32(0), f32(0), f32(0), f32(1)]
        self.piv_corr = [f32(0)] * 3


def _parse_transform(jx) -> _Transform:
    t = _Transform()
    if not isinstance(jx, dict):
        return t
    tilt = [f32(0)] * 3
    pivot = [f32(0)] * 3
    p = jx.get("position")
    if isinstance(p, dict):
        t.pos = [_fl(p[k]) if k in p else f32(0) for k in ("x", "y", "z")]
    r = jx.get("rotation")
    if isinstance(r, dict):
        tilt = [_fl(r[k]) if k in r else f32(0) for k in ("tilt_x", "tilt_y", "roll")]
    v = jx.get("pivot")
    if isinstance(v, dict):
        pivot = [_fl(v[k]) if k in v else f32(0) for k in ("x", "y", "z")]
    t.rot = _make_rotation(*tilt)
    if any(c != f32(0) for c in pivot):
        rp = _mat_vec(t.rot, pivot)
        t.piv_corr = [f32(pivot[i] - rp[i]) for i in range(3)]
    return t


def _apply_pivot(p: dict, point, t: _Transform) -> None:
    """P_e' = R_off (P_e - P_p) + P_p + dp, R_e' = R_off R_e
    (optical_system.cpp:1016-1048)."""
    R_off = _make_rotation(*p["off_rot"])
    rel = [f32(t.pos[i] - point[i]) for i in range(3)]
    rot = _mat_vec(R_off, rel)
    t.pos = [f32(f32(rot[i] + point[i]) + p["off_pos"][i]) for i in range(3)]
    t.rot = _mat_mat(R_off, t.rot)
    t.piv_corr = _mat_vec(R_off, t.piv_corr)


def _parse_pivots(j_piv, path) -> List[dict]:
    if not isinstance(j_piv, list):
        raise LensFileError(f"{path}: pivots must be an array")
    out = []
    for jp in j_piv:
        if not isinstance(jp, dict):
            raise LensFileError(f"{path}: pivot entry must be an object")
        name = jp.get("name", "")
        if not isinstance(jp.get("elements"), list):
            raise LensFileError(f"{path}: pivot '{name}' missing 'elements' array")
        p = {"name": name, "elements": [str(e) for e in jp["elements"]], "manual": False,
             "xyz": [f32(0)] * 3, "off_pos": [f32(0)] * 3, "off_rot": [f32(0)] * 3}
        pp = jp.get("pivot_point")
        if isinstance(pp, dict):
            p["manual"] = pp.get("mode", "centroid") == "manual"
            p["xyz"] = [_fl(_val(pp, k, 0.0)) for k in ("x", "y", "z")]
        off = jp.get("offset")
        if isinstance(off, dict):
            if isinstance(off.get("position"), dict):
                p["off_pos"] = [_fl(_val(off["position"], k, 0.0)) for k in ("x", "y", "z")]
            if isinstance(off.get("rotation"), dict):
                p["off_rot"] = [_fl(_val(off["rotation"], k, 0.0)) for k in ("tilt_x", "tilt_y", "roll")]
        out.append(p)
    return out


# ---- surface modifiers -----------------------------------------------------

def _clamp01(v) -> np.float32:
    v = f32(v)
    return f32(0) if v < f32(0) else (f32(1) if v > f32(1) else v)


def _extract_coating(mods, path) -> Tuple[Coating, Optional[list]]:
    """Coating modifier (last one wins, optical_system.cpp:338-520).  Returns
    the coating and, for a physical layer stack, the layer specs to bake."""
    c = Coating()
    layers = None
    if not isinstance(mods, list):
        return c, None
    for m in mods:
        if not isinstance(m, dict) or m.get("type", "") != "coating":
            continue
        model = m.get("model", "")
        if not model:
            raise LensFileError(f"{path}: coating modifier has no 'model' field")
        if model == "layers":
            if not isinstance(m.get("layers"), list):
                raise LensFileError(f"{path}: coating model 'layers' has no 'layers' array")
            specs = []
            for jl in m["layers"]:
                if not isinstance(jl, dict):
                    specs = None
                    break
                nk = [(_fl(e["lambda_um"]), _fl(_val(e, "n", 1.0)), _fl(_val(e, "k", 0.0)))
                      for e in jl.get("nk_table", []) if isinstance(e, dict) and "lambda_um" in e]
                th = _fl(_val(jl, "thickness_nm", 0.0))
                if not nk or th <= f32(0):
                    specs = None
                    break
                specs.append({"thickness_nm": th, "nk_table": np.array(nk, np.float32)})
            if not specs:
                raise LensFileError(f"{path}: coating 'layers' stack is empty or has a layer with no nk_table / "
                                    "non-positive thickness")
            c, layers = Coating(), specs
            continue
        if model == "simple":
            c, layers = Coating(), None
            c.ar_layers = int(_val(m, "ar_layers", 0))
        elif model == "artist":
            c, layers = Coating(), None
            c.model = CoatingModel.ARTIST
            tint = m.get("tint")
            if isinstance(tint, list) and len(tint) == 3:
                c.tint_r, c.tint_g, c.tint_b = (float(_clamp01(_fl(v))) for v in tint)
            c.tint_strength = float(_clamp01(_fl(_val(m, "strength", 0.04))))
        elif model in ("spectral", "angular"):
            key = "lambda_nm" if model == "spectral" else "angle_deg"
            rows = [(_fl(e[key]), _clamp01(_fl(e["r"]))) for e in (m.get("data") or [])
                    if isinstance(e, dict) and key in e and "r" in e] if isinstance(m.get("data"), list) else []
            if not rows:
                _warn(path, "coating modifier has an empty/invalid 'data' table; treating as uncoated.")
                continue
            rows.sort(key=lambda kr: kr[0])
            c, layers = Coating(), None
            c.model = CoatingModel.SPECTRAL if model == "spectral" else CoatingModel.ANGULAR
            c.out_of_range_discard = _val(m, "out_of_range", "clamp") == "discard"
            if model == "angular":
                c.angle_ref_ior = float(_fl(_val(m, "angle_ref_ior", 1.0)))
            c.table = np.array(rows, np.float32)
        elif model == "spectral_angular":
            wl = [_fl(v) for v in m.get("wavelengths_nm", [])] if isinstance(m.get("wavelengths_nm"), list) else []
            ang = [_fl(v) for v in m.get("angles_deg", [])] if isinstance(m.get("angles_deg"), list) else []
            r = []
            if isinstance(m.get("r"), list):
                for row in m["r"]:
                    if isinstance(row, list):
                        r.extend(_clamp01(_fl(v)) for v in row)
            if (not wl or not ang or len(r) != len(wl) * len(ang) or wl != sorted(wl) or ang != sorted(ang)):
                _warn(path, "spectral_angular coating has inconsistent or unsorted wavelengths/angles/r arrays; "
                            "treating as uncoated.")
                continue
            c, layers = Coating(), None
            c.model = CoatingModel.SPECTRAL_ANGULAR
            c.out_of_range_discard = _val(m, "out_of_range", "clamp") == "discard"
            c.angle_ref_ior = float(_fl(_val(m, "angle_ref_ior", 1.0)))
            c.sa_wavelengths = np.array(wl, np.float32)
            c.sa_angles = np.array(ang, np.float32)
            c.sa_r = np.array(r, np.float32).reshape(len(wl), len(ang))
        elif model == "attenuator_gaussian":
            c, layers = Coating(), None
            c.model = CoatingModel.ATTENUATOR_GAUSS
            c.gauss_sigma = float(_fl(_val(m, "sigma", 0.0)))
            c.gauss_background = float(_fl(_val(m, "attenuation_background", 0.0)))
            c.gauss_peak = float(_fl(_val(m, "attenuation_peak", 0.0)))
            c.gauss_decenter_x = float(_fl(_val(m, "decenter_x", 0.0)))
            c.gauss_decenter_y = float(_fl(_val(m, "decenter_y", 0.0)))
        else:
            raise LensFileError(f"{path}: coating model '{model}' is not recognised")
    return c, layers


def _set_circular(s: Surface, aspect) -> None:
    s.aperture_shape = ApertureShape.CIRCLE
    s.aperture_blades = 0
    s.aperture_rotation_rad = 0.0
    s.aperture_aspect = float(aspect)
    s.aperture_curvature = s.aperture_twist = 0.0
    s.aperture_notch_rad = s.aperture_notch_angle_rad = 0.0


def _blade_control(m, key, lo, hi, path) -> np.float32:
    if key not in m:
        return f32(0)
    v = _fl(m[key])
    if v < f32(lo) or v > f32(hi):
        c = f32(lo) if v < f32(lo) else f32(hi)
        _warn(path, f"aperture {key} must be within [{lo}, {hi}]; got {v} -- clamping to {c}.")
        return c
    return v


def _apply_aperture(mods, s: Surface, path) -> Optional[dict]:
    """Aperture modifiers (last one wins, optical_system.cpp:575-690).
    Returns the image-aperture record (path, semi diameter) if any."""
    image = None
    if not isinstance(mods, list):
        return None
    PI = f32(3.14159265358979323846)
    DEG = f32(PI / f32(180.0))
    for m in mods:
        if not isinstance(m, dict) or m.get("type", "") != "aperture":
            continue
        shape = m.get("shape", "")
        aspect = f32(1.0)
        if "aperture_aspect" in m:
            aspect = _fl(m["aperture_aspect"])
            if not aspect > f32(0):
                _warn(path, f"aperture_aspect must be > 0; got {aspect} -- clamping to 1.0.")
                aspect = f32(1.0)
        if shape == "circular":
            _set_circular(s, aspect)
        elif shape == "polygon":
            blades = int(_val(m, "blades", 0))
            if blades < 3:
                _warn(path, f"polygon aperture needs blades >= 3; got {blades} -- falling back to circular.")
                _set_circular(s, aspect)
                continue
            rot_deg = _fl(_val(m, "rotation_deg", 0.0))
            _set_circular(s, aspect)
            s.aperture_shape = ApertureShape.POLYGON
            s.aperture_blades = blades
            s.aperture_rotation_rad = float(f32(rot_deg * DEG))
            s.aperture_curvature = float(_blade_control(m, "curvature", -1.0, 1.0, path))
            s.aperture_twist = float(_blade_control(m, "twist", -1.0, 1.0, path))
            s.aperture_notch_rad = float(f32(_blade_control(m, "notch_deg", -45.0, 45.0, path) * DEG))
            s.aperture_notch_angle_rad = float(f32(_blade_control(m, "notch_angle_deg", 0.0, 45.0, path) * DEG))
        elif shape == "image":
            if "image_path" not in m or "semi_diameter" not in m:
                _warn(path, "image aperture missing 'image_path' or 'semi_diameter'; falling back to circular.")
                _set_circular(s, aspect)
                continue
            sd = _fl(m["semi_diameter"])
            if not sd > f32(0):
                _warn(path, "image aperture semi_diameter must be > 0; falling back to circular.")
                _set_circular(s, aspect)
                continue
            _set_circular(s, aspect)
            s.aperture_shape = ApertureShape.IMAGE
            s.aperture_semi_diameter = float(sd)
            image = {"source_path": str(m["image_path"]), "semi_diameter": float(sd)}
        else:
            _warn(path, f"unknown aperture shape '{shape}'; falling back to circular.")
            _set_circular(s, aspect)
    return image


def _parse_form(jf, s: Surface, path) -> None:
    if not isinstance(jf, dict):
        raise LensFileError(f"{path}: surface 'form' must be an object")
    t = jf.get("type", "")
    if t == "sphere":
        s.form = SurfaceForm.SPHERE
        s.radius = float(_fl(_val(jf, "radius", 0.0)))
    elif t == "asphere":
        s.form = SurfaceForm.ASPHERE
        s.radius = float(_fl(_val(jf, "radius", 0.0)))
        s.conic_k = float(_fl(_val(jf, "conic_constant", 0.0)))
        terms = jf.get("terms")
        if isinstance(terms, list):
            if len(terms) > MAX_ASPHERE_TERMS:
                _warn(path, f"asphere has {len(terms)} terms; truncating to {MAX_ASPHERE_TERMS}.")
            s.asphere_terms = tuple(float(_fl(v)) for v in terms[:MAX_ASPHERE_TERMS])
    elif t == "cylindrical":
        s.form = SurfaceForm.CYLINDRICAL
        s.radius = float(_fl(_val(jf, "radius", 0.0)))
        ax = jf.get("axis", "x")
        s.cyl_axis = CylinderAxis.Y if ax in ("y", "Y") else CylinderAxis.X
    else:
        raise LensFileError(f"{path}: unknown surface form '{t}'")


def _apply_air(s: Surface) -> None:
    s.ior, s.abbe_v, s.disp_model = 1.0, 0.0, DispersionModel.AIR
    s.sellmeier_B = (0.0, 0.0, 0.0)
    s.sellmeier_C = (0.0, 0.0, 0.0)


def _apply_glass(name, catalogue, s: Surface, path) -> None:
    if not name or name in ("air", "AIR"):
        _apply_air(s)
        return
    if name not in catalogue:
        raise LensFileError(f"{path}: glass '{name}' not present in glass_catalogue")
    model, n_d, V_d, B, Cc = catalogue[name]
    s.ior, s.abbe_v, s.disp_model = float(n_d), float(V_d), model
    s.sellmeier_B = tuple(float(b) for b in B)
    s.sellmeier_C = tuple(float(c) for c in Cc)


# ---- flatten ---------------------------------------------------------------

def _flatten_element(j_el, M: _Transform, nominal_z, catalogue, surfaces, layer_specs, images, ids, path):
    js_list, jm_list = j_el.get("surfaces"), j_el.get("materials")
    if not isinstance(js_list, list):
        raise LensFileError(f"{path}: element missing 'surfaces' array")
    if not isinstance(jm_list, list):
        raise LensFileError(f"{path}: element missing 'materials' array")
    if len(js_list) != len(jm_list) + 1:
        raise LensFileError(f"{path}: element '{j_el.get('name', '')}' has {len(js_list)} surfaces and "
                            f"{len(jm_list)} materials (surfaces must equal materials + 1)")
    cum_local = f32(0)
    cum_nominal = f32(nominal_z)
    last_nominal = f32(nominal_z)
    n_el = len(js_list)
    for i, js in enumerate(js_list):
        s = Surface(semi_aperture=0.0, disp_model=DispersionModel.AIR)
        s.is_stop = bool(_val(js, "is_stop", False))
        s.is_active = bool(_val(js, "is_active", True))
        s.semi_aperture = float(_fl(_val(js, "semi_aperture", 0.0)))
        mods = _val(js, "modifiers", [])
        s.coating, layers = _extract_coating(mods, path)
        _parse_form(_val(js, "form", {}), s, path)
        image = _apply_aperture(mods, s, path)
        thick = _fl(_val(js, "thickness", 0.0))
        s.thickness = float(thick)
        s.decenter_x = float(f32(f32(M.pos[0] + M.piv_corr[0]) + f32(M.rot[2] * cum_local)))
        s.decenter_y = float(f32(f32(M.pos[1] + M.piv_corr[1]) + f32(M.rot[5] * cum_local)))
        s.z = float(f32(f32(M.pos[2] + M.piv_corr[2]) + f32(M.rot[8] * cum_local)))
        s.rot = tuple(float(v) for v in M.rot)
        if i < n_el - 1:
            _apply_glass(_val(jm_list[i], "glass", ""), catalogue, s, path)
        else:
            _apply_air(s)
        if s.semi_aperture <= 0.0:
            _warn(path, f"element '{j_el.get('name', '')}' surface {i} has semi_aperture <= 0; the tracer will "
                        "reject all rays.")
        surfaces.append(s)
        layer_specs.append(layers)
        images.append(image)
        sid = _val(js, "id", "") or f"auto-surface-{len(surfaces) - 1}"
        ids.append(sid)
        last_nominal = cum_nominal
        cum_local = f32(cum_local + thick)
        cum_nominal = f32(cum_nominal + thick)
    return last_nominal


def load_lens(path) -> OpticalSystem:
    """Load a ``ghostlight-optical`` JSON lens file (OpticalSystem::load,
    optical_system.cpp:1299-1425).  Raises LensFileError where the reference's
    loader returns false."""
    path = os.fspath(path)
    try:
        with open(path, "r", encoding="utf-8") as fh:
            j = json.load(fh)
    except OSError as e:
        raise LensFileError(f"cannot open lens file: {path} ({e})") from None
    except ValueError as e:
        raise LensFileError(f"{path}: invalid JSON: {e}") from None
    if not isinstance(j, dict) or j.get("format", "") != "ghostlight-optical":
        raise LensFileError(f"{path}: unexpected format (expected 'ghostlight-optical')")
    major = minor = 0
    if isinstance(j.get("version"), dict):
        major = int(_val(j["version"], "major", 0))
        minor = int(_val(j["version"], "minor", 0))
    if major != LENS_FORMAT_MAJOR:
        raise LensFileError(f"{path}: unsupported lens format major version {major} "
                            f"(this build reads version {LENS_FORMAT_MAJOR} only)")
    if minor > LENS_FORMAT_MINOR:
        _warn(path, f"lens format minor version {minor} is newer than this build understands "
                    f"({LENS_FORMAT_MINOR}); unknown fields will be ignored.")
    name, focal = "", 0.0
    if isinstance(j.get("metadata"), dict):
        name = str(_val(j["metadata"], "name", ""))
        focal = float(_fl(_val(j["metadata"], "focal_length_mm", 0.0)))
    catalogue = _parse_glass_catalogue(j["glass_catalogue"], path) if "glass_catalogue" in j else {}
    if not isinstance(j.get("optical_system"), list):
        raise LensFileError(f"{path}: optical_system must be an array")

    # Pass 1: element transforms in file order.
    records = []
    prev_z, have_prev = f32(0), False
    for item in j["optical_system"]:
        if not isinstance(item, dict) or item.get("type", "") != "element":
            raise LensFileError(f"{path}: optical_system item has unknown type (expected 'element')")
        jx = _val(item, "transform", {})
        t = _parse_transform(jx)
        mode = "absolute"
        if isinstance(jx, dict) and isinstance(jx.get("position"), dict):
            mode = jx["position"].get("mode", "absolute")
        if mode == "relative_to_preceding":
            if not have_prev:
                _warn(path, f"first element '{item.get('name', '')}' has position.mode='relative_to_preceding'; "
                            "treating as absolute.")
            else:
                t.pos[2] = f32(t.pos[2] + prev_z)
        elif mode != "absolute":
            _warn(path, f"element '{item.get('name', '')}' has unknown position.mode='{mode}'; treating as absolute.")
        records.append({"id": str(_val(item, "id", "")), "node": item, "xf": t, "pre_pivot_z": t.pos[2]})
        prev_z, have_prev = t.pos[2], True

    # Pass 2: pivots, stacking in file order.
    if "pivots" in j:
        id_to_idx = {}
        for i, r in enumerate(records):
            if r["id"] and r["id"] not in id_to_idx:
                id_to_idx[r["id"]] = i
        for p in _parse_pivots(j["pivots"], path):
            targets = []
            for eid in p["elements"]:
                if eid not in id_to_idx:
                    raise LensFileError(f"{path}: pivot '{p['name']}' references unknown element id '{eid}'")
                targets.append(id_to_idx[eid])
            if not targets:
                continue
            if p["manual"]:
                point = list(p["xyz"])
            else:
                # Centroid of the current element origins, summed in double.
                acc = [0.0, 0.0, 0.0]
                for k in targets:
                    for a in range(3):
                        acc[a] += float(records[k]["xf"].pos[a])
                n = float(len(targets))
                point = [f32(acc[a] / n) for a in range(3)]
            for k in targets:
                _apply_pivot(p, point, records[k]["xf"])

    # Pass 3: emit surfaces.
    surfaces: List[Surface] = []
    layer_specs, images, ids, emits = [], [], [], []
    for r in records:
        nominal_z = r["xf"].pos[2]
        first = len(surfaces)
        last_nominal = _flatten_element(r["node"], r["xf"], nominal_z, catalogue, surfaces, layer_specs, images, ids,
                                        path)
        emits.append((nominal_z, first, len(surfaces) - first, last_nominal))
    for k in range(len(emits) - 1):
        nz, first, n, last_nominal = emits[k]
        surfaces[first + n - 1].thickness = float(f32(emits[k + 1][0] - last_nominal))
    if emits and records:
        internal = f32(0)
        for js in records[-1]["node"]["surfaces"]:
            internal = f32(internal + _fl(_val(js, "thickness", 0.0)))
        total = f32(records[-1]["pre_pivot_z"] + internal)
        for s in surfaces:
            s.z = float(f32(f32(s.z) - total))
    if not surfaces:
        raise LensFileError(f"{path}: optical_system produced no surfaces")

    lens = OpticalSystem(surfaces, name=name)
    lens.focal_length = focal
    lens.surface_ids = ids
    lens.aperture_images = [ApertureImage(semi_diameter=im["semi_diameter"], source_path=im["source_path"])
                            if im else ApertureImage() for im in images]
    # Post-flatten: bake physical layer stacks now that every medium is final.
    for i, specs in enumerate(layer_specs):
        if specs:
            lens.set_coating_layers(i, specs)
    lens._source_path = path
    return lens
