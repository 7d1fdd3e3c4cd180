"""Write the golden vectors in tests/golden/ from the reference's own CPU code.

Runs only where oracle/_ref/libglref.so was built from the reference sources
(oracle/Makefile `ref`).  The fixtures are data (inputs and the reference's
outputs); they pin the oracle and the product's host preparation where the
reference sources are not available (e.g. on the GPU box).

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ghostlight_amd import _abi, synthetic  # noqa: E402
from oracle import pyoracle  # noqa: E402


def rays_for(lens, n, seed, max_deg=8.0):
    rng = np.random.default_rng(seed)
    fR = lens.surfaces[0].semi_aperture
    z0 = lens.surfaces[0].z - 20.0
    ang = np.radians(rng.uniform(-max_deg, max_deg, (n, 2)))
    d = np.stack([np.tan(ang[:, 0]), np.tan(ang[:, 1]), np.ones(n)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 7), np.float32)
    rays[:, :2] = rng.uniform(-1, 1, (n, 2)) * fR
    rays[:, 2] = z0
    rays[:, 3:6] = d
    rays[:, 6] = rng.uniform(400, 700, n)
    return rays


def main():
    ref = pyoracle.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libglref.so not built (needs the reference sources)")
    out = {}
    # Spectral tables + correction matrices (spectral.cpp:317, :419).
    for n in (8, 16, 31, 64):
        for cs in (0, 1, 2, 3, 4):
            tab = (_abi.gl_spec * n)()
            ref.ref_build_spectral_table(n, cs, None, tab)
            t = np.frombuffer(bytes(tab), np.float32).reshape(n, 4).copy()
            C9 = (C.c_float * 9)()
            ref.ref_spectral_correction(tab, n, cs, None, C9)
            out[f"spec_n{n}_cs{cs}"] = t
            out[f"corr_n{n}_cs{cs}"] = np.array(C9[:], np.float32)
    # Colour matrices.
    for ics in (0, 1):
        for ocs in (0, 1, 2, 3, 4):
            Mi, Mo = (C.c_float * 9)(), (C.c_float * 9)()
            ref.ref_color_matrices(ics, None, ocs, None, Mi, Mo)
            out[f"cmat_in{ics}_out{ocs}"] = np.array(Mi[:] + Mo[:], np.float32)
    lenses = {"c1": synthetic.config1_lens(), "c2": synthetic.config2_lens(), "c4": synthetic.config4_lens()}
    for name, lens in lenses.items():
        surfs, arena = lens.pack()
        n = lens.num_surfaces()
        out[f"lens_{name}_surfs"] = np.frombuffer(bytes(surfs), np.uint8).copy()
        out[f"lens_{name}_arena"] = arena
        # Pupil samples (aperture_sampler.h:181) for the render / probe / refine grids.
        for (N, j, seed, bo, rd) in [(64, 0, 0, 0, 0.0), (64, 1, 3, 0, 0.0), (64, 2, 0, 0, 0.0),
                                     (40, 1, int(np.uint32(0x9E3779B9).astype(np.int32)), 0, 0.0),
                                     (128, 1, int(np.uint32(0xB5297A4D).astype(np.int32)), 0, 0.0),
                                     (48, 0, 0, 5, 12.0)]:
            cap = N * N
            o = (_abi.gl_sample * cap)()
            k = C.c_int()
            ref.ref_build_aperture_samples(surfs, n, N, j, seed, bo, C.c_float(rd), o, cap, C.byref(k))
            out[f"samples_{name}_N{N}_j{j}_s{seed}_b{bo}"] = np.frombuffer(bytes(o), np.float32).reshape(-1, 2)[: k.value].copy()
        # Pair pre-filter (ghost.cpp:105).
        hw, hh = synthetic.sensor_for(lens, 1920, 1080)
        cap = n * n
        o = (_abi.gl_pair * cap)()
        k = C.c_int()
        ref.ref_filter_ghost_pairs(surfs, n, pyoracle._f(arena), C.c_float(hw), C.c_float(hh), C.c_float(1e-7), 1,
                                   C.c_float(100.0), o, cap, C.byref(k))
        out[f"filter_{name}_pairs"] = np.array([[o[i].surf_a, o[i].surf_b] for i in range(k.value)], np.int32)
        out[f"filter_{name}_boosts"] = np.array([o[i].area_boost for i in range(k.value)], np.float32)
        out[f"filter_{name}_sensor"] = np.array([hw, hh], np.float32)
        # Single-ray ghost traces (trace.cpp:171) through a spread of pairs.
        rays = rays_for(lens, 1000, seed=11)
        out[f"rays_{name}"] = rays
        pair_list = [(0, 1), (1, n - 1), (2, n // 2), (n // 3, n - 2), (n - 3, n - 1)]
        out[f"trace_{name}_pairs"] = np.array(pair_list, np.int32)
        for (a, b) in pair_list:
            pos = np.zeros((len(rays), 3), np.float32)
            w = np.zeros(len(rays), np.float32)
            st = np.zeros(len(rays), np.int32)
            ref.ref_trace_ghost_rays(surfs, n, pyoracle._f(arena), a, b, pyoracle._f(rays), len(rays),
                                     pyoracle._f(pos), pyoracle._f(w), pyoracle._i(st))
            out[f"trace_{name}_{a}_{b}_pos"] = pos
            out[f"trace_{name}_{a}_{b}_w"] = w
            out[f"trace_{name}_{a}_{b}_st"] = st
        pos = np.zeros((len(rays), 3), np.float32)
        w = np.zeros(len(rays), np.float32)
        st = np.zeros(len(rays), np.int32)
        ref.ref_trace_primary_rays(surfs, n, pyoracle._f(arena), pyoracle._f(rays), len(rays), pyoracle._f(pos),
                                   pyoracle._f(w), pyoracle._i(st))
        out[f"primary_{name}_pos"], out[f"primary_{name}_w"], out[f"primary_{name}_st"] = pos, w, st
    # TMM single-layer and 3-layer reflectances (tmm.h:89).
    tm = []
    for layers in ([(1.38, 0.0, 99.64)], [(1.38, 0.0, 99.6), (2.05, 0.0, 67.1), (1.63, 0.0, 84.4)]):
        nre = np.array([l[0] for l in layers], np.float32)
        nim = np.array([l[1] for l in layers], np.float32)
        th = np.array([l[2] for l in layers], np.float32)
        for lam in (400.0, 550.0, 700.0):
            for aoi in (0.0, 30.0, 60.0, 85.0):
                r = ref.ref_tmm_reflectance(len(layers), pyoracle._f(nre), pyoracle._f(nim), pyoracle._f(th),
                                            C.c_float(1.0), C.c_float(1.5168), C.c_float(lam), C.c_float(aoi))
                tm.append([len(layers), lam, aoi, r])
    out["tmm"] = np.array(tm, np.float32)
    path = os.path.join(HERE, "reference_vectors.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB, {len(out)} arrays)")


REF_ROOT = "/root/reference"
PF_CASES = [("double_gauss", "lenses/DoubleGauss.lens"),
            ("sirui", "lenses/anamorphic/Sirui_75mm_F1.8_1.33x_EP3936919NWA1.lens")]
# (source_x, source_y, blades override) of validation/point_flare_sampling_baseline.py:25
PF_CONFIGS = [(0.5, 0.5, 0), (1.3, 0.5, 0), (1.6, 0.9, 0), (1.3, 0.5, 6)]
PF_SH = 14.0
PF_DS = 8


def box(img, k):
    H, W, Cc = img.shape
    return img.reshape(H // k, k, W // k, k, Cc).sum(axis=(1, 3))


def point_flare_goldens():
    """The reference's point-flare goldens (validation/goldens/point_flare_sampling,
    320x320x3 float32, produced by the reference's GPU renderer) in reduced
    form -- 8x8 box sums, per-channel totals, peak -- together with what it takes
    to re-render them: each lens as the C ABI's surface array (OpticalSystem::load,
    optical_system.cpp:1299), its calibrate_lens record and the solve_source_angle
    angles (source_map.cpp:182) of every configuration, from the reference's CPU
    code."""
    ref = pyoracle.ref()
    out = {}
    gold = os.path.join(REF_ROOT, "validation", "goldens", "point_flare_sampling")
    for name, rel in PF_CASES:
        arr = (_abi.gl_surface * 64)()
        arena = np.zeros(1 << 20, np.float32)
        al = C.c_int()
        n = ref.ref_load_lens(os.path.join(REF_ROOT, rel).encode(), arr, 64,
                              arena.ctypes.data_as(C.POINTER(C.c_float)), len(arena), C.byref(al))
        if n <= 0:
            raise SystemExit(f"could not load {rel}")
        raw = bytes(arr)[: n * C.sizeof(_abi.gl_surface)]
        out[f"pf_{name}_surfaces"] = np.frombuffer(raw, np.uint8).copy()
        out[f"pf_{name}_arena"] = arena[: max(1, al.value)].copy()
        cal = np.zeros(13, np.float32)
        ref.ref_calibrate_lens(arr, n, arena.ctypes.data_as(C.POINTER(C.c_float)),
                               cal.ctypes.data_as(C.POINTER(C.c_float)))
        out[f"pf_{name}_calib"] = cal
        for sx, sy, blades in PF_CONFIGS:
            ang = np.zeros(2, np.float32)
            ref.ref_solve_source_angle(arr, n, arena.ctypes.data_as(C.POINTER(C.c_float)),
                                       cal.ctypes.data_as(C.POINTER(C.c_float)), C.c_float((sx - 0.5) * 2.0),
                                       C.c_float((sy - 0.5) * 2.0), C.c_float(PF_SH), C.c_float(PF_SH),
                                       ang.ctypes.data_as(C.POINTER(C.c_float)))
            for cull in (0, 1):
                key = f"{name}_{sx}_{sy}_b{blades}_c{cull}"
                img = np.load(os.path.join(gold, key + ".npy"))          # allow_pickle=False (default)
                out[f"pf_{key}_angles"] = ang
                out[f"pf_{key}_ds"] = box(img.astype(np.float64), PF_DS).astype(np.float32)
                out[f"pf_{key}_sum"] = img.astype(np.float64).sum(axis=(0, 1))
                out[f"pf_{key}_peak"] = np.array([float(img.sum(-1).max())])
    path = os.path.join(HERE, "point_flare_goldens.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


SM_LENSES = [("double_gauss", "lenses/DoubleGauss.lens"), ("helios44", "lenses/Helios44.lens"),
             ("atlas_40_f3", "lenses/anamorphic/AtlasLensCo_40mm_F3.0_1.5x_US12379576B2_Table1.lens"),
             ("atlas_40_f35", "lenses/anamorphic/AtlasLensCo_40mm_F3.5_1.5x_US20240053589A1_Table2_InfFocus.lens"),
             ("cooke_50_2x", "lenses/anamorphic/Cooke_50mm_F3_2x_US9341827B2.lens"),
             ("sirui_24", "lenses/anamorphic/Sirui_24mm_F1.8_1.33x_US10969568B1.lens"),
             ("sirui", "lenses/anamorphic/Sirui_75mm_F1.8_1.33x_EP3936919NWA1.lens"),
             ("sirui_venus", "lenses/anamorphic/Sirui_Venus_50mm_T2.9_1.6X_US20220196995A1.lens"),
             ("exp_helios_2x", "lenses/experiments/Experimental_Helios44SphericalBlock_60mm_2x_"
                               "DoubletsCounterRotatedDeAstigmatiser.lens"),
             ("exp_panatar", "lenses/experiments/Experimental_PanavisionPanatarInspired_50mm_2x_"
                             "CounterRotatingDeAstigmatiser.lens"),
             ("artisans_50", "lenses/spherical/7Artisans_50mm_F0.95.lens"),
             ("canon_100", "lenses/spherical/Canon_100mm.lens"),
             ("canon_24", "lenses/spherical/Canon_24mm_F1.4.lens"),
             ("cooke_pancro", "lenses/spherical/Cooke_SpeedPancro_50mm_F2_US2959102A.lens"),
             ("summicron_100", "lenses/spherical/Leitz_Summicron_100mm.lens"),
             ("hex_stop", "validation/goldens/aperture_baseline/hex_stop.lens"),
             ("hex_rot30", "validation/goldens/aperture_baseline/hex_stop_rot30.lens"),
             ("blade_shape", "validation/goldens/aperture_baseline/blade_shape.lens")]
# Screen NDC requests; None in the last two slots = the calibrated half extents.
SM_REQUESTS = [(0.0, 0.0, None, None), (0.5, 0.0, None, None), (0.0, -0.5, None, None), (0.7, -0.6, None, None),
               (-0.95, 0.9, None, None), (1.6, 0.2, None, None), (-2.5, -1.5, None, None), (0.3, 0.3, 14.0, 14.0)]


def load_ref_lens(ref, rel):
    arr = (_abi.gl_surface * 64)()
    arena = np.zeros(1 << 20, np.float32)
    al = C.c_int()
    n = ref.ref_load_lens(os.path.join(REF_ROOT, rel).encode(), arr, 64,
                          arena.ctypes.data_as(C.POINTER(C.c_float)), len(arena), C.byref(al))
    if n <= 0:
        raise SystemExit(f"could not load {rel}")
    return arr, n, arena[: max(1, al.value)].copy()


def source_map_goldens():
    """calibrate_lens (lens_calibration.cpp:344), solve_source_angle
    (source_map.cpp:182, the whole SourceMapSolve), source_map_landing (:487)
    and source_map_screen (:456) from the reference's CPU code, for the
    reference's own lens files; plus the calibration manifest the reference
    keeps beside its aperture goldens."""
    import json
    ref = pyoracle.ref()
    fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    out = {}
    for name, rel in SM_LENSES:
        arr, n, arena = load_ref_lens(ref, rel)
        out[f"sm_{name}_surfaces"] = np.frombuffer(bytes(arr)[: n * C.sizeof(_abi.gl_surface)], np.uint8).copy()
        out[f"sm_{name}_arena"] = arena
        cal = np.zeros(13, np.float32)
        ref.ref_calibrate_lens(arr, n, fp(arena), fp(cal))
        out[f"sm_{name}_calib"] = cal
        reqs, solves = [], []
        for nx, ny, hw, hh in SM_REQUESTS:
            hw = float(cal[0]) if hw is None else hw
            hh = float(cal[1]) if hh is None else hh
            r = np.zeros(17, np.float32)
            ref.ref_solve_source_map(arr, n, fp(arena), fp(cal), C.c_float(nx), C.c_float(ny), C.c_float(hw),
                                     C.c_float(hh), fp(r))
            reqs.append((nx, ny, hw, hh))
            solves.append(r)
        out[f"sm_{name}_requests"] = np.array(reqs, np.float32)
        out[f"sm_{name}_solves"] = np.array(solves, np.float32)
        # The forward map of every solve at its own angle and at two offsets.
        offs = np.array([[0.0, 0.0], [0.01, -0.02], [0.3, 0.1]], np.float32)
        scr = np.zeros((len(solves), len(offs), 2), np.float32)
        for i, r in enumerate(solves):
            for j, (dx, dy) in enumerate(offs):
                ref.ref_source_map_screen(fp(r), C.c_float(np.float32(r[0] + dx)), C.c_float(np.float32(r[1] + dy)),
                                          fp(scr[i, j]))
        out[f"sm_{name}_screen_offsets"] = offs
        out[f"sm_{name}_screen"] = scr
        mh, mv = float(cal[2]), float(cal[3])
        angs = np.array([[0.0, 0.0], [0.5 * mh, 0.0], [0.0, -0.8 * mv], [0.6 * mh, 0.5 * mv], [1.3 * mh, 0.4 * mv]],
                        np.float32)
        lands = np.zeros((len(angs), 3), np.float32)
        for i, (ax, ay) in enumerate(angs):
            xy = np.zeros(2, np.float32)
            ok = ref.ref_source_map_landing(arr, n, fp(arena), fp(cal), C.c_float(ax), C.c_float(ay), fp(xy))
            lands[i] = (ok, xy[0], xy[1])
        out[f"sm_{name}_landing_angles"] = angs
        out[f"sm_{name}_landings"] = lands
        print(f"  {name}: {n} surfaces, calib sensor_half ({cal[0]:.4f}, {cal[1]:.4f}), statuses "
              f"{[int(r[2]) for r in solves]}", flush=True)
    man = json.load(open(os.path.join(REF_ROOT, "validation", "goldens", "aperture_baseline", "manifest.json")))
    fields = [f for f in _abi.gl_calibration._fields_]
    for key in ("double_gauss", "sirui", "hex_stop", "hex_rot30", "blade_shape"):
        out[f"manifest_{key}"] = np.array([man[key][f[0]] for f in fields], np.float32)
    path = os.path.join(HERE, "source_map_goldens.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


LENS_DIR = os.path.join(HERE, "lenses")


def lens_file_goldens():
    """What the reference's loader (OpticalSystem::load, optical_system.cpp:1299)
    makes of every lens file under tests/golden/lenses (copies of the
    reference's lens files and its Python tests' fixtures), exported as the C
    ABI's surface bytes and coating arena; -1 marks a file it refuses."""
    ref = pyoracle.ref()
    out = {}
    for fn in sorted(os.listdir(LENS_DIR)):
        if not fn.endswith(".lens"):
            continue
        arr = (_abi.gl_surface * 64)()
        arena = np.zeros(1 << 20, np.float32)
        al = C.c_int()
        n = ref.ref_load_lens(os.path.join(LENS_DIR, fn).encode(), arr, 64,
                              arena.ctypes.data_as(C.POINTER(C.c_float)), len(arena), C.byref(al))
        key = fn[:-5]
        out[f"lf_{key}_n"] = np.array([n], np.int32)
        if n > 0:
            out[f"lf_{key}_surfaces"] = np.frombuffer(bytes(arr)[: n * C.sizeof(_abi.gl_surface)], np.uint8).copy()
            out[f"lf_{key}_arena"] = arena[: max(1, al.value)].copy()
    path = os.path.join(HERE, "lens_file_goldens.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


AB_CASES = ("double_gauss", "sirui", "hex_stop", "hex_rot30", "blade_shape")
AB_LAYERS = ("ghost_axial", "ghost_offaxis", "ghost_corner", "starburst_sprite", "starburst_mdft", "psf",
             "source_chunk", "gate")
AB_LENSES = {"double_gauss": "lenses/DoubleGauss.lens",
             "sirui": "lenses/anamorphic/Sirui_75mm_F1.8_1.33x_EP3936919NWA1.lens",
             "hex_stop": "validation/goldens/aperture_baseline/hex_stop.lens",
             "hex_rot30": "validation/goldens/aperture_baseline/hex_stop_rot30.lens",
             "blade_shape": "validation/goldens/aperture_baseline/blade_shape.lens"}
# Source position of the gate layer per lens (validation/aperture_baseline.py:49-58).
AB_GATE_SX = {"double_gauss": 1.03, "sirui": 1.05, "hex_stop": 1.03, "hex_rot30": 1.03, "blade_shape": 1.03}


def aperture_baseline_goldens():
    """The reference's aperture-baseline goldens (validation/goldens/
    aperture_baseline/<lens>__<layer>.npy: 192x192x3 renders, 144x144x3 PSF
    grid) for the five lenses, reduced to 4x4 box sums, per-channel totals and
    the peak."""
    gold = os.path.join(REF_ROOT, "validation", "goldens", "aperture_baseline")
    out = {}
    for case in AB_CASES:
        for layer in AB_LAYERS:
            img = np.load(os.path.join(gold, f"{case}__{layer}.npy")).astype(np.float64)
            k = f"ab_{case}__{layer}"
            out[k + "_ds"] = box(img, 4).astype(np.float32)
            out[k + "_sum"] = img.sum(axis=(0, 1))
            out[k + "_peak"] = np.array([float(np.abs(img).max())])
            out[k + "_shape"] = np.array(img.shape, np.int32)
        # The gate layer's source: the reference's solve of (gate_sx, 0.5) on a
        # frame of 0.7 x the calibrated half extents (aperture_baseline.py:103-111).
        ref = pyoracle.ref()
        fp = lambda a: a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
        arr, n, arena = load_ref_lens(ref, AB_LENSES[case])
        cal = np.zeros(13, np.float32)
        ref.ref_calibrate_lens(arr, n, fp(arena), fp(cal))
        hw, hh = np.float32(0.70 * float(cal[0])), np.float32(0.70 * float(cal[1]))
        r = np.zeros(17, np.float32)
        ref.ref_solve_source_map(arr, n, fp(arena), fp(cal), C.c_float((AB_GATE_SX[case] - 0.5) * 2.0),
                                 C.c_float(0.0), C.c_float(hw), C.c_float(hh), fp(r))
        out[f"ab_{case}__gate_angles"] = r[:2].copy()
        out[f"ab_{case}__gate_half"] = np.array([hw, hh], np.float32)
    path = os.path.join(HERE, "aperture_goldens.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


# (blades, rotation, curvature, twist, notch, notch angle): plain polygons,
# each control alone and together, both clamps, blade_shape's own stop.
BLADE_CASES = [
    (6, 0.0, 0.0, 0.0, 0.0, 0.0), (5, 0.3, 0.0, 0.0, 0.0, 0.4), (9, 0.1, 0.35, 0.0, 0.0, 0.0),
    (9, 0.0, -0.6, 0.0, 0.0, 0.0), (7, 0.2, 0.0, 0.7, 0.0, 0.0), (7, 0.0, 0.0, -1.0, 0.0, 0.0),
    (6, 0.0, 0.0, 0.0, 0.5, 0.0), (6, 0.4, 0.0, 0.0, -0.3, 0.6), (11, 0.0, 0.5, -0.4, 0.3, 0.2),
    (3, 1.0, -1.0, 0.9, 0.78, 0.78), (8, -0.5, 1.0, 0.0, 0.2, 0.0), (4, 0.0, 2.0, -3.0, 1.5, 2.0),
    (13, 2.5, -0.2, 0.3, -0.6, 0.1), (32, 0.0, 0.3, 0.2, 0.1, 0.05),
]


def blade_goldens():
    """The reference's blade silhouette (make_aperture_profile and radius_at,
    aperture_profile.h:116-363) for BLADE_CASES and the blade_shape stop, the
    reference's trace statuses of primary rays through the blade_shape lens
    aimed at its stop edge, and its pupil sample grid for that lens."""
    ref = pyoracle.ref()
    out = {}
    prof_fields = [f for f, _ in _abi.gl_blade_profile._fields_]
    arr, n, arena = load_ref_lens(ref, AB_LENSES["blade_shape"])
    stop = next(i for i in range(n) if arr[i].is_stop)
    g = arr[stop]
    cases = BLADE_CASES + [(g.aperture_blades, g.aperture_rotation_rad, g.aperture_curvature, g.aperture_twist,
                            g.aperture_notch_rad, g.aperture_notch_angle_rad)]
    theta = np.concatenate([np.linspace(-7.0, 7.0, 4001), np.random.default_rng(5).uniform(-40, 40, 999)])
    theta = theta.astype(np.float32)
    params, profs, radii = [], [], []
    for c in cases:
        pr = _abi.gl_blade_profile()
        ref.ref_blade_profile(C.c_int(int(c[0])), *[C.c_float(float(x)) for x in c[1:]], C.byref(pr))
        r = np.zeros(len(theta), np.float32)
        ref.ref_blade_radius(C.c_int(int(c[0])), *[C.c_float(float(x)) for x in c[1:]],
                             theta.ctypes.data_as(C.POINTER(C.c_float)), len(theta),
                             r.ctypes.data_as(C.POINTER(C.c_float)))
        params.append(np.array(c, np.float32))
        profs.append(np.frombuffer(bytes(pr), np.uint8).copy())
        radii.append(r)
    out["bp_params"] = np.stack(params)
    out["bp_profile_bytes"] = np.stack(profs)
    out["bp_theta"] = theta
    out["bp_radius"] = np.stack(radii)
    out["bp_fields"] = np.array(prof_fields)
    # Primary rays at the stop edge: trace a fan at 0, 3 and 6 degrees whose
    # pupil positions straddle the deformed silhouette.
    rng = np.random.default_rng(11)
    m = 6000
    fR = float(arr[0].semi_aperture)
    z0 = float(arr[0].z) - 20.0
    rays = np.zeros((m, 7), np.float32)
    ang = np.radians(rng.choice([0.0, 3.0, 6.0], m))
    phi = rng.uniform(-np.pi, np.pi, m)
    rho = rng.uniform(0.3, 0.75, m)
    rays[:, 0] = rho * np.cos(phi) * fR - 20.0 * np.tan(ang)
    rays[:, 1] = rho * np.sin(phi) * fR
    rays[:, 2] = z0
    d = np.stack([np.tan(ang), np.zeros(m), np.ones(m)], 1)
    rays[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 6] = 587.56
    pos = np.zeros((m, 3), np.float32)
    w = np.zeros(m, np.float32)
    st = np.zeros(m, np.int32)
    ref.ref_trace_primary_rays(arr, n, pyoracle._f(arena), pyoracle._f(rays), m, pyoracle._f(pos), pyoracle._f(w),
                               st.ctypes.data_as(C.POINTER(C.c_int)))
    out["bs_rays"] = rays
    out["bs_status"] = st
    out["bs_pos"] = pos
    # The reference's pupil sample grid for the deformed stop (no override).
    for N, jit in ((64, 0), (96, 1), (80, 2)):
        cap = N * N
        buf = (_abi.gl_sample * cap)()
        cnt = C.c_int()
        ref.ref_build_aperture_samples(arr, n, N, jit, 12345, 0, C.c_float(0.0), buf, cap, C.byref(cnt))
        out[f"bs_grid_{N}_{jit}"] = np.frombuffer(bytes(buf), np.float32)[: 2 * cnt.value].reshape(-1, 2).copy()
    # The reference's signed distance field bake (aperture_sdf.cpp:110-189)
    # of a few bitmaps: a disc, an anisotropic star, a soft random matte.
    sdf_cases = []
    c = (np.arange(48) + 0.5) / 48 * 2 - 1
    X, Y = np.meshgrid(c, c)
    sdf_cases.append(((np.hypot(X, Y) < 0.6).astype(np.float32), 10.0, 1.0))
    th = np.arctan2(Y, X)
    sdf_cases.append(((np.hypot(X, Y) < 0.55 + 0.2 * np.cos(5 * th)).astype(np.float32)[:, :40].copy(), 7.5, 1.33))
    sdf_cases.append((np.random.default_rng(4).random((20, 28)).astype(np.float32), 3.0, 0.8))
    for q, (img, semi, asp) in enumerate(sdf_cases):
        H_, W_ = img.shape
        o4 = np.zeros(4 * W_ * H_, np.float32)
        ref.ref_bake_aperture_sdf(pyoracle._f(np.ascontiguousarray(img)), W_, H_, C.c_float(semi), C.c_float(asp),
                                  pyoracle._f(o4))
        out[f"sdf{q}_img"] = img
        out[f"sdf{q}_par"] = np.array([semi, asp], np.float32)
        out[f"sdf{q}_out"] = o4
    path = os.path.join(HERE, "blade_goldens.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path)} bytes, {len(out)} arrays)")


def trace_path_goldens():
    """The diagnostic traces (trace.cpp:264-395): per-surface events and the
    landing of 96 rays through five pairs and the primary path of each
    synthetic lens (and of the 12-surface one with an interior surface
    muted), from the reference's own RayPath overloads."""
    from ghostlight_amd.render import TRACE_EVENT_DTYPE
    ref = pyoracle.ref()
    if ref is None:
        raise SystemExit("oracle/_ref/libglref.so not built (needs the reference sources)")
    muted = synthetic.config1_lens()
    muted.surfaces[5].is_active = False
    lenses = {"c1": synthetic.config1_lens(), "c2": synthetic.config2_lens(), "c4": synthetic.config4_lens(),
              "c1m": muted}
    out = {}
    for name, lens in lenses.items():
        surfs, arena = lens.pack()
        n = lens.num_surfaces()
        out[f"lens_{name}_surfs"] = np.frombuffer(bytes(surfs), np.uint8).copy()
        out[f"lens_{name}_arena"] = arena
        rays = rays_for(lens, 96, seed=23, max_deg=12.0)
        out[f"rays_{name}"] = rays
        m = 3 * n
        pairs = [(0, 1), (1, n - 1), (2, n // 2), (n // 3, n - 2), (n - 3, n - 1), (-1, -1)]
        out[f"pairs_{name}"] = np.array(pairs, np.int32)
        for (a, b) in pairs:
            ev = np.zeros((len(rays), m), TRACE_EVENT_DTYPE)
            cnt = np.zeros(len(rays), np.int32)
            pos = np.zeros((len(rays), 3), np.float32)
            w = np.zeros(len(rays), np.float32)
            st = np.zeros(len(rays), np.int32)
            ref.ref_trace_ray_paths(surfs, n, pyoracle._f(arena), a, b, pyoracle._f(rays), len(rays), m,
                                    ev.ctypes.data_as(C.c_void_p), pyoracle._i(cnt), pyoracle._f(pos), pyoracle._f(w),
                                    pyoracle._i(st))
            key = f"path_{name}_{a}_{b}"
            out[key + "_ev"], out[key + "_n"], out[key + "_pos"], out[key + "_w"], out[key + "_st"] = ev, cnt, pos, w, st
    path = os.path.join(HERE, "trace_path_goldens.npz")
    np.savez_compressed(path, **out)
    print(f"wrote {path} ({os.path.getsize(path) / 1024:.0f} KiB, {len(out)} arrays)")


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "point_flare":
        point_flare_goldens()
    elif len(sys.argv) > 1 and sys.argv[1] == "source_map":
        source_map_goldens()
    elif len(sys.argv) > 1 and sys.argv[1] == "lens_files":
        lens_file_goldens()
    elif len(sys.argv) > 1 and sys.argv[1] == "aperture":
        aperture_baseline_goldens()
    elif len(sys.argv) > 1 and sys.argv[1] == "blades":
        blade_goldens()
    elif len(sys.argv) > 1 and sys.argv[1] == "trace_paths":
        trace_path_goldens()
    else:
        main()
        point_flare_goldens()
        source_map_goldens()
        lens_file_goldens()
        aperture_baseline_goldens()
        blade_goldens()
        trace_path_goldens()
