"""The FP32 op count per traced segment that the benchmark prices the ghost
tracer's roofline with is recounted here: the FP32 restatement of the trace,
run with a counting scalar over a sample of the benchmark workload's rays
(12 pairs spread over the pre-filtered list, 256 pupil rays each), must give
the constant to 0.1%."""
import os
import sys

import pytest

from oracle import pyoracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    sys.path.insert(0, ROOT)
    import bench as b
    return b


def test_ops_per_segment_constant(bench):
    if pyoracle.lib() is None:
        pytest.skip("oracle library not built")
    lens, cfg, src, W, H, hw, hh = bench.workload()
    pairs, _boosts = pyoracle.filter_ghost_pairs(lens, hw, hh)
    ops = bench.ops_per_segment(lens, pairs, cfg, src)
    print(f"ops per segment {ops:.3f} (constant {bench.OPS_PER_SEGMENT})")
    assert abs(ops / bench.OPS_PER_SEGMENT - 1.0) < 1e-3
