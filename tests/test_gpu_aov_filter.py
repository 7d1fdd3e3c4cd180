"""PER_PAIR AOV layers, GhostFilter pair selection and surface muting on the
device, through the public render_point_flare.

  AOV        the properties of the reference's test_ghost_aov.py:19-130 (no
             extra keys in NONE mode, r/g/b triplets of (H, W) float32, layers
             non-negative up to the CMF lobes, layers sum to the combined
             image, the aov_max_pairs cap and a combined image from the capped
             layers only), plus each layer against the FP32 oracle's render of
             that pair alone (the reference renders one launch per pair,
             ghost_render.cu:1526-1567; this build makes all layers in one)
  filter     test_ghost_filter.py:76-218: ALL ignores the list, an empty
             INCLUDE list is a no-op, INCLUDE / EXCLUDE are pointwise below
             the baseline and partition it, AOV layers follow the filter, a
             pair that is not there renders nothing; and an INCLUDE render
             against the oracle's render of the same two pairs
  muting     validate_culling_edgecases.py: with an interior surface muted (and
             with the 6-blade override) cull-on equals cull-off within the
             8-bit quantum, on DoubleGauss and the 1.33x Sirui at (1.3, 0.5),
             and the muted render matches the FP32 oracle"""
import dataclasses

import numpy as np
import pytest

import golden_flare as gf
from ghostlight_amd import hostprep
from ghostlight_amd.config import GhostAovMode, GhostFilter, GhostFilterMode, PointFlareConfig, to_params
from ghostlight_amd.render import filter_ghost_pairs, source_rgb
from oracle import pyoracle

pytestmark = pytest.mark.gpu

_LENS = {}


def _lens(name="double_gauss"):
    if name not in _LENS:
        _LENS[name] = (gf.lens(name), gf.calibration(name))
    return _LENS[name]


def _aov_cfg(**kw):
    base = dict(ray_grid=16, spectral_samples=4)
    base.update(kw)
    return PointFlareConfig(**base)


def _layers(out, ch="r"):
    return {k: v for k, v in out.items() if k.startswith("ghost_s") and k.endswith("_" + ch)}


# ---------------------------------------------------------------- AOV


def test_aov_none_adds_no_keys_and_matches_default():
    lens, cal = _lens()
    a = lens.render_point_flare(32, 32, _aov_cfg(), calib=cal)
    b = lens.render_point_flare(32, 32, _aov_cfg(aov_mode=GhostAovMode.NONE), calib=cal)
    assert not any(k.startswith("ghost_s") for k in a)
    np.testing.assert_allclose(a["ghost_r"], b["ghost_r"], rtol=1e-5, atol=1e-7)


def test_aov_per_pair_layers():
    lens, cal = _lens()
    out = lens.render_point_flare(32, 32, _aov_cfg(aov_mode=GhostAovMode.PER_PAIR), calib=cal)
    r = _layers(out)
    assert len(r) > 0
    for k in r:
        for ch in "gb":
            assert k[:-1] + ch in out
    for k in (k for k in out if k.startswith("ghost_s")):
        assert out[k].shape == (32, 32)
        assert out[k].dtype == np.float32
        assert np.all(out[k] >= -1e-6)
    for ch in ("ghost_r", "ghost_g", "ghost_b"):
        assert ch in out
        total = sum(v for k, v in out.items() if k.startswith("ghost_s") and k.endswith("_" + ch[-1]))
        np.testing.assert_allclose(total, out[ch], rtol=1e-5, atol=1e-6)
    # one layer per filtered pair, in filter order
    pairs, _ = filter_ghost_pairs(lens, cal.sensor_half_w, cal.sensor_half_h, _aov_cfg())
    assert list(r) == [f"ghost_s{a}_s{b}_r" for a, b in pairs]


def test_aov_max_pairs_cap():
    lens, cal = _lens()
    out = lens.render_point_flare(32, 32, _aov_cfg(aov_mode=GhostAovMode.PER_PAIR, aov_max_pairs=2), calib=cal)
    r = _layers(out)
    assert len(r) == 2
    np.testing.assert_allclose(sum(r.values()), out["ghost_r"], rtol=1e-5, atol=1e-6)
    none = lens.render_point_flare(32, 32, _aov_cfg(aov_mode=GhostAovMode.PER_PAIR, aov_max_pairs=0), calib=cal)
    assert not _layers(none) and float(np.abs(none["ghost_r"]).max()) == 0.0


def test_aov_layers_match_single_pair_oracle_renders():
    """Each layer = the oracle's render of its pair alone (one pair: no
    probes, the reference's per-pair launch)."""
    name, sx, sy = "double_gauss", 1.3, 0.5
    lens, cal = _lens(name)
    W = H = 96
    cfg = dataclasses.replace(gf.config(sx, sy, 0, 1), ray_grid=32, spectral_samples=8,
                              aov_mode=GhostAovMode.PER_PAIR, aov_max_pairs=8)
    ax, ay = gf.angles(name, sx, sy, 0, 1)
    rgb = source_rgb(cfg)
    src = np.array([[ax, ay, rgb[0], rgb[1], rgb[2]]], np.float32)
    from ghostlight_amd import render_sources
    out = render_sources(src, W, H, lens, cal, cfg)
    pairs, boosts = pyoracle.filter_ghost_pairs(lens, gf.SH, gf.SH, min_intensity=0.0)
    spec = hostprep.spectral_table(8)
    p, _ = to_params(cfg)
    Cm = hostprep.spectral_correction(spec)
    worst = 0.0
    for (a, b), bo in list(zip(pairs, boosts))[:8]:
        ref, _ = pyoracle.render_ghosts(lens, [(a, b)], [bo], src, spec, p, gf.SH, gf.SH, W, H, correction=Cm)
        got = [out[f"ghost_s{a}_s{b}_{ch}"] for ch in "rgb"]
        tot = sum(float(x.sum()) for x in ref)
        e = pyoracle.rel_l2(got, ref) if tot > 0 else float(max(np.abs(g).max() for g in got))
        worst = max(worst, e)
        print(f"layer ({a},{b}): energy {tot:.4g}, rel-L2 vs oracle {e:.2e}", flush=True)
    assert worst <= 1e-4


# ---------------------------------------------------------------- GhostFilter


def _flt_cfg(**kw):
    base = dict(ray_grid=12, spectral_samples=3, min_ghost_intensity=0.0, ghost_normalize=False)
    base.update(kw)
    return PointFlareConfig(**base)


def _with_filter(mode, pairs, **kw):
    return _flt_cfg(ghost_filter=GhostFilter(mode=mode, pairs=list(pairs)), **kw)


def test_ghost_filter_selection():
    lens, cal = _lens()
    base = _flt_cfg()
    pairs, _ = filter_ghost_pairs(lens, cal.sensor_half_w, cal.sensor_half_h, base)
    assert len(pairs) >= 2
    ref = lens.render_point_flare(32, 32, base, calib=cal)["ghost_r"]
    r = lambda c: lens.render_point_flare(32, 32, c, calib=cal)  # noqa: E731
    np.testing.assert_allclose(r(_with_filter(GhostFilterMode.ALL, [(0, 1)]))["ghost_r"], ref, rtol=1e-5, atol=1e-7)
    np.testing.assert_allclose(r(_with_filter(GhostFilterMode.INCLUDE, []))["ghost_r"], ref, rtol=1e-5, atol=1e-7)
    pick = pairs[0]
    inc = r(_with_filter(GhostFilterMode.INCLUDE, [pick]))["ghost_r"]
    exc = r(_with_filter(GhostFilterMode.EXCLUDE, [pick]))["ghost_r"]
    assert np.all(inc <= ref + 1e-5) and float(inc.sum()) > 0.0
    assert np.all(exc <= ref + 1e-5)
    np.testing.assert_allclose(inc + exc, ref, rtol=1e-4, atol=1e-5)
    empty = r(_with_filter(GhostFilterMode.INCLUDE, [(999, 9999)]))
    assert all(float(empty[k].max()) == 0.0 for k in ("ghost_r", "ghost_g", "ghost_b"))
    ai = r(_with_filter(GhostFilterMode.INCLUDE, [pick], aov_mode=GhostAovMode.PER_PAIR))
    assert list(_layers(ai)) == [f"ghost_s{pick[0]}_s{pick[1]}_r"]
    ae = r(_with_filter(GhostFilterMode.EXCLUDE, [pick], aov_mode=GhostAovMode.PER_PAIR))
    assert len(_layers(ae)) == len(pairs) - 1
    assert f"ghost_s{pick[0]}_s{pick[1]}_r" not in ae


def test_ghost_filter_include_matches_oracle_pairs():
    """INCLUDE two pairs = the oracle's render of exactly those two pairs."""
    name, sx, sy = "double_gauss", 1.3, 0.5
    lens, cal = _lens(name)
    W = H = 96
    allp, allb = pyoracle.filter_ghost_pairs(lens, gf.SH, gf.SH, min_intensity=0.0)
    keep = [allp[3], allp[7]]
    base = dataclasses.replace(gf.config(sx, sy, 0, 1), ray_grid=32, spectral_samples=8)
    cfg = dataclasses.replace(base, ghost_filter=GhostFilter(mode=GhostFilterMode.INCLUDE, pairs=keep))
    ax, ay = gf.angles(name, sx, sy, 0, 1)
    rgb = source_rgb(cfg)
    src = np.array([[ax, ay, rgb[0], rgb[1], rgb[2]]], np.float32)
    from ghostlight_amd import render_sources
    out = render_sources(src, W, H, lens, cal, cfg)
    spec = hostprep.spectral_table(8)
    p, _ = to_params(base)
    bo = [allb[allp.index(q)] for q in keep]
    ref, _ = pyoracle.render_ghosts(lens, keep, bo, src, spec, p, gf.SH, gf.SH, W, H,
                                    correction=hostprep.spectral_correction(spec))
    e = pyoracle.rel_l2([out["ghost_r"], out["ghost_g"], out["ghost_b"]], ref)
    print(f"INCLUDE {keep}: rel-L2 vs oracle {e:.2e}", flush=True)
    assert sum(float(x.sum()) for x in ref) > 0.0
    assert e <= 1e-4


# ---------------------------------------------------------------- muting


def _mute_one(lens):
    S = list(lens.surfaces)
    cands = [i for i, s in enumerate(S) if s.is_active and not s.is_stop and 0 < i < len(S) - 1]
    idx = cands[len(cands) // 2]
    lens.surfaces[idx].is_active = False
    return idx


QUANTUM_PPM = 1.0 / 255.0 * 1e6


@pytest.mark.parametrize("name", gf.LENSES)
@pytest.mark.parametrize("variant", ["plain", "hex_stop", "muted"])
def test_cull_lossless_edge_cases(name, variant):
    lens = gf.lens(name)
    idx = _mute_one(lens) if variant == "muted" else None
    cal = lens.calibration()
    blades = 6 if variant == "hex_stop" else 0

    def cfg(cull):
        c = PointFlareConfig(source_x=1.3, source_y=0.5, source_r=10.0, source_g=10.0, source_b=10.0,
                             flare_gain=5000.0, ray_grid=256, spectral_samples=16, pupil_jitter=2, jitter_seed=12345,
                             sensor_half_w=gf.SH, sensor_half_h=gf.SH, min_ghost_intensity=0.0, cull_dead_pairs=cull)
        if blades:
            c = dataclasses.replace(c, aperture_blades=6, aperture_rotation=0.0)
        return c

    def render(cull):
        o = lens.render_point_flare(320, 320, cfg(cull), calib=cal)
        return np.stack([o["ghost_r"], o["ghost_g"], o["ghost_b"]], -1).astype(np.float64)

    full, cull = render(False), render(True)
    peak = float(full.sum(-1).max())
    maxpix = float(np.abs(full - cull).sum(-1).max()) / (peak + 1e-20) * 1e6
    energy = float(cull.sum() / (full.sum() + 1e-20))
    print(f"{name} {variant}{'' if idx is None else f' (surface {idx} muted)'}: max pixel change {maxpix:.1f} ppm, "
          f"energy {energy:.6f}", flush=True)
    assert peak > 0.0
    assert maxpix < QUANTUM_PPM
    assert abs(energy - 1.0) < 1e-4


def test_muted_surface_render_matches_oracle():
    """The IOR 'before' walk and the trace skip a muted surface the same way
    on the device and in the oracle (which follows optical_system.h:386-402)."""
    name, sx, sy = "double_gauss", 1.3, 0.5
    lens = gf.lens(name)
    idx = _mute_one(lens)
    W = H = 128
    cfg = dataclasses.replace(gf.config(sx, sy, 0, 1), ray_grid=48, spectral_samples=8)
    ax, ay = gf.angles(name, sx, sy, 0, 1)
    rgb = source_rgb(cfg)
    src = np.array([[ax, ay, rgb[0], rgb[1], rgb[2]]], np.float32)
    from ghostlight_amd import render_ghosts
    pairs, boosts = pyoracle.filter_ghost_pairs(lens, gf.SH, gf.SH, min_intensity=0.0)
    assert all(idx not in q for q in pairs)
    got = render_ghosts(lens, src, W, H, cfg, gf.SH, gf.SH, pairs=pairs, boosts=boosts)
    spec = hostprep.spectral_table(8)
    p, _ = to_params(cfg)
    ref, _ = pyoracle.render_ghosts(lens, pairs, boosts, src, spec, p, gf.SH, gf.SH, W, H,
                                    correction=hostprep.spectral_correction(spec))
    e = pyoracle.rel_l2(got, ref)
    print(f"muted surface {idx}: {len(pairs)} pairs, rel-L2 vs oracle {e:.2e}", flush=True)
    assert sum(float(x.sum()) for x in ref) > 0.0
    assert e <= 1e-4
