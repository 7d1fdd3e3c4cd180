"""The device film-gate layer (gl_render_gate) against its CPU restatement on
the same inputs: identical trace and scrape counts (the primary traces are
bit-exact, so every wall decision is the same), images within float rounding
of the scatter draws' transcendentals."""
import numpy as np
import pytest

import test_oracle_gate as tg
from ghostlight_amd.render import render_gate_layer

pytestmark = pytest.mark.gpu
TOL = 1e-4


@pytest.mark.parametrize("case", list(tg.FILES))
def test_gate_matches_oracle(case):
    cfg = tg._cfg()
    lens, src, hw, hh, ref, ctr = tg.gate_case(case, cfg)
    out = render_gate_layer(src, 192, 192, lens, cfg, hw, hh, counters=True)
    img = np.stack([out["gate_r"], out["gate_g"], out["gate_b"]], -1).astype(np.float64)
    err = float(np.linalg.norm(img - ref) / np.linalg.norm(ref))
    print(f"{case}: counters {out['gate_counters']} vs {ctr}, rel-L2 {err:.2e}", flush=True)
    assert out["gate_counters"] == ctr
    assert err <= TOL


def test_gate_lobes_and_sources():
    """Isotropic Cauchy lobe, a rotated groove, two sources of different
    colour, asymmetric offsets."""
    cfg = tg._cfg()
    cfg.gate.lobe = 0
    cfg.gate.groove_angle_deg = 30.0
    cfg.gate.offset_left_mm = -0.5
    cfg.gate.offset_top_mm = 0.3
    cfg.gate.scatter_samples = 7
    lens, src, hw, hh, _, _ = tg.gate_case("double_gauss", cfg)
    src2 = np.concatenate([src, src * np.array([[1, 1, 0.2, 1.0, 0.5]], np.float32) +
                           np.array([[-0.35, 0.02, 0, 0, 0]], np.float32)])
    from ghostlight_amd import hostprep
    from ghostlight_amd.config import gate_params
    from oracle import pyoracle
    grid = pyoracle.build_samples(lens, cfg.ray_grid, cfg.pupil_jitter, cfg.jitter_seed)
    spec = hostprep.spectral_table(8).copy()
    spec[:, 1:] *= np.float32(300.0 / 8)
    ref, ctr = pyoracle.render_gate(lens, src2, grid, spec, gate_params(cfg.gate), cfg.jitter_seed, hw, hh, 192, 192)
    out = render_gate_layer(src2, 192, 192, lens, cfg, hw, hh, counters=True)
    a = np.stack([out["gate_r"], out["gate_g"], out["gate_b"]], -1).astype(np.float64)
    b = np.stack(ref, -1).astype(np.float64)
    err = float(np.linalg.norm(a - b) / np.linalg.norm(b))
    print(f"two sources: counters {out['gate_counters']} vs {ctr}, rel-L2 {err:.2e}", flush=True)
    assert out["gate_counters"] == ctr and ctr[1] > 0
    assert err <= TOL
