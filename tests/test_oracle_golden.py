"""CPU: the oracle and the product's host preparation against the reference's
own outputs (tests/golden/reference_vectors.npz, written from the reference's
CPU code by tests/golden/make_golden.py) and, where oracle/_ref is built here,
against the reference library directly."""
import ctypes as C
import os

import numpy as np
import pytest

from ghostlight_amd import _abi, hostprep
from oracle import pyoracle

GOLD = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))
NAMES = ("c1", "c2", "c4")


def packed(name):
    return pyoracle.PackedLens(GOLD[f"lens_{name}_surfs"], GOLD[f"lens_{name}_arena"])


@pytest.mark.parametrize("n", [8, 16, 31, 64])
@pytest.mark.parametrize("cs", [0, 1, 2, 3, 4])
def test_spectral_table_and_correction_bitexact(n, cs):
    t = hostprep.spectral_table(n, cs)
    assert np.array_equal(t, GOLD[f"spec_n{n}_cs{cs}"])
    Cm = hostprep.spectral_correction(t, cs)
    assert np.array_equal(Cm.ravel(), GOLD[f"corr_n{n}_cs{cs}"])


@pytest.mark.parametrize("ics", [0, 1])
@pytest.mark.parametrize("ocs", [0, 1, 2, 3, 4])
def test_color_matrices_bitexact(ics, ocs):
    Mi, Mo = hostprep.color_matrices(ics, ocs)
    assert np.array_equal(np.concatenate([Mi.ravel(), Mo.ravel()]), GOLD[f"cmat_in{ics}_out{ocs}"])


class _HostLens:
    """Adapter: product host prep over a packed fixture lens."""

    def __init__(self, p):
        self.p = p

    def pack(self):
        return self.p.pack()

    def num_surfaces(self):
        return self.p.num_surfaces()


@pytest.mark.parametrize("name", NAMES)
def test_pupil_samples_bitexact(name):
    lens = packed(name)
    surfs, _ = lens.pack()
    keys = [k for k in GOLD.files if k.startswith(f"samples_{name}_")]
    assert len(keys) == 6
    for k in keys:
        parts = k.split("_")
        N, j, seed, bo = int(parts[2][1:]), int(parts[3][1:]), int(parts[4][1:]), int(parts[5][1:])
        rd = 12.0 if bo else 0.0
        want = GOLD[k]
        got_orc = pyoracle.build_samples(lens, N, j, seed, bo, rd)
        assert np.array_equal(got_orc, want), k
        cap = N * N
        out = (_abi.gl_sample * cap)()
        cnt = C.c_int32()
        _abi.check(_abi.load_library().gl_build_aperture_samples(surfs, lens.num_surfaces(), N, j, seed, bo, rd, out,
                                                                 cap, C.byref(cnt)), "samples")
        got = np.frombuffer(bytes(out), np.float32).reshape(-1, 2)[: cnt.value]
        assert np.array_equal(got, want), k


@pytest.mark.parametrize("name", NAMES)
def test_oracle_trace_matches_reference_cpu_tracer(name):
    """oracle trace_ghost (CPU t >= 0 rule) == reference trace_ghost_ray,
    bit for bit: statuses, positions and weights."""
    lens = packed(name)
    rays = GOLD[f"rays_{name}"]
    n_ok = 0
    for a, b in GOLD[f"trace_{name}_pairs"]:
        pos, w, st, _ = pyoracle.trace_ghost(lens, int(a), int(b), rays, dbl=False, t_ge0=True)
        assert np.array_equal(st, GOLD[f"trace_{name}_{a}_{b}_st"]), (a, b)
        ok = st == 0
        n_ok += int(ok.sum())
        assert np.array_equal(pos[ok].astype(np.float32), GOLD[f"trace_{name}_{a}_{b}_pos"][ok])
        assert np.array_equal(w[ok].astype(np.float32), GOLD[f"trace_{name}_{a}_{b}_w"][ok])
    assert n_ok > 0, "no surviving ghost ray in the fixture set"
    pos, w, st, _ = pyoracle.trace_ghost(lens, -1, -1, rays, dbl=False, t_ge0=True)
    assert np.array_equal(st, GOLD[f"primary_{name}_st"])
    ok = st == 0
    assert ok.sum() > 10
    assert np.array_equal(pos[ok].astype(np.float32), GOLD[f"primary_{name}_pos"][ok])
    assert np.array_equal(w[ok].astype(np.float32), GOLD[f"primary_{name}_w"][ok])


@pytest.mark.parametrize("name", NAMES)
def test_fp64_oracle_close_to_fp32(name):
    """The FP64 oracle agrees with the FP32 one to float precision on the
    surviving set (guards a dropped term in the double instantiation)."""
    lens = packed(name)
    rays = GOLD[f"rays_{name}"][:200]
    a, b = GOLD[f"trace_{name}_pairs"][2]
    p32, w32, s32, _ = pyoracle.trace_ghost(lens, int(a), int(b), rays, dbl=False)
    p64, w64, s64, _ = pyoracle.trace_ghost(lens, int(a), int(b), rays, dbl=True)
    both = (s32 == 0) & (s64 == 0)
    assert (s32 != s64).mean() < 0.02
    np.testing.assert_allclose(p32[both], p64[both], atol=2e-3)
    np.testing.assert_allclose(w32[both], w64[both], rtol=1e-4)


@pytest.mark.parametrize("name", NAMES)
def test_oracle_filter_matches_reference(name):
    lens = packed(name)
    hw, hh = GOLD[f"filter_{name}_sensor"]
    pairs, boosts = pyoracle.filter_ghost_pairs(lens, float(hw), float(hh))
    assert np.array_equal(np.array(pairs, np.int32).reshape(-1, 2), GOLD[f"filter_{name}_pairs"])
    assert np.array_equal(np.array(boosts, np.float32), GOLD[f"filter_{name}_boosts"])


def test_tmm_bake_matches_reference():
    """gl_bake_tmm_table (3-layer stack, constant indices) vs tmm::reflectance."""
    for row in GOLD["tmm"]:
        nl, lam, aoi, want = int(row[0]), float(row[1]), float(row[2]), float(row[3])
        layers = [(1.38, 99.64)] if nl == 1 else [(1.38, 99.6), (2.05, 67.1), (1.63, 84.4)]
        n_re = np.repeat(np.array([[l[0]] for l in layers], np.float32), 31, axis=1)
        n_im = np.zeros_like(n_re)
        th = np.array([l[1] for l in layers], np.float32)
        tab = hostprep.bake_tmm_table(n_re, n_im, th, np.ones(31, np.float32), np.full(31, 1.5168, np.float32))
        wi, ai = int(round((lam - 400) / 10)), int(round(aoi / 5))
        got = tab[49 + wi * 18 + ai]
        assert got == np.float32(want), (nl, lam, aoi, got, want)


def test_oracle_trace_vs_reference_library_random():
    """Beyond the fixtures: 20k fresh random rays through the C1 lens."""
    # Looked up inside the test, never at collection: the compiled reference
    # exists only where its sources are, and is not shipped with the tree.
    if pyoracle.ref() is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    from ghostlight_amd import synthetic
    lens = synthetic.config1_lens()
    surfs, arena = lens.pack()
    n = lens.num_surfaces()
    rng = np.random.default_rng(123)
    N = 4000
    rays = np.zeros((N, 7), np.float32)
    ang = np.radians(rng.uniform(-15, 15, (N, 2)))
    d = np.stack([np.tan(ang[:, 0]), np.tan(ang[:, 1]), np.ones(N)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, :2] = rng.uniform(-1, 1, (N, 2)) * lens.surfaces[0].semi_aperture
    rays[:, 2] = lens.surfaces[0].z - 20
    rays[:, 3:6] = d
    rays[:, 6] = rng.uniform(400, 700, N)
    for a, b in [(0, 3), (4, 10), (7, 12)]:
        pos = np.zeros((N, 3), np.float32); w = np.zeros(N, np.float32); st = np.zeros(N, np.int32)
        pyoracle.ref().ref_trace_ghost_rays(surfs, n, pyoracle._f(arena), a, b, pyoracle._f(rays), N,
                                            pyoracle._f(pos), pyoracle._f(w), pyoracle._i(st))
        op, ow, ost, _ = pyoracle.trace_ghost(lens, a, b, rays)
        assert np.array_equal(ost, st)
        ok = st == 0
        assert np.array_equal(op[ok].astype(np.float32), pos[ok])
        assert np.array_equal(ow[ok].astype(np.float32), w[ok])
