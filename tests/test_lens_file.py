"""The lens-file loader against the reference's own loader
(OpticalSystem::load, optical_system.cpp:1299): every lens file under
tests/golden/lenses packs to the same gl_surface bytes and coating arena the
reference's loader exports (fixture from tests/golden/make_golden.py), and
the files it refuses are refused."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from ghostlight_amd import OpticalSystem, _abi
from ghostlight_amd.lens_file import LensFileError

HERE = os.path.dirname(__file__)
LENS_DIR = os.path.join(HERE, "golden", "lenses")
GOLD = np.load(os.path.join(HERE, "golden", "lens_file_goldens.npz"))
FILES = sorted(f[:-5] for f in os.listdir(LENS_DIR) if f.endswith(".lens"))


@pytest.mark.parametrize("name", FILES)
def test_load_matches_reference_loader(name):
    path = os.path.join(LENS_DIR, name + ".lens")
    n = int(GOLD[f"lf_{name}_n"][0])
    if n <= 0:
        with pytest.raises(LensFileError):
            OpticalSystem.load(path)
        return
    lens = OpticalSystem.load(path)
    assert lens.num_surfaces() == n
    arr, arena = lens.pack()
    got = np.frombuffer(bytes(arr)[: n * C.sizeof(_abi.gl_surface)], np.uint8)
    assert np.array_equal(got, GOLD[f"lf_{name}_surfaces"]), "surface bytes differ"
    assert np.array_equal(arena, GOLD[f"lf_{name}_arena"]), "coating arena differs"


def _doublet(tmp_path, **patch):
    j = json.load(open(os.path.join(LENS_DIR, "example_doublet.lens")))
    for k, v in patch.items():
        j[k] = v
    p = tmp_path / "t.lens"
    p.write_text(json.dumps(j))
    return str(p), j


def test_loader_refusals(tmp_path):
    p, _ = _doublet(tmp_path, format="something-else")
    with pytest.raises(LensFileError, match="format"):
        OpticalSystem.load(p)
    p, _ = _doublet(tmp_path, version={"major": 2, "minor": 0})
    with pytest.raises(LensFileError, match="major version"):
        OpticalSystem.load(p)
    p, j = _doublet(tmp_path)
    j["optical_system"][0]["materials"].append({"glass": "air"})
    (tmp_path / "t.lens").write_text(json.dumps(j))
    with pytest.raises(LensFileError, match="materials"):
        OpticalSystem.load(p)
    with pytest.raises(LensFileError, match="cannot open"):
        OpticalSystem.load(str(tmp_path / "missing.lens"))


def test_loaded_chain_ends_at_sensor():
    """The rebased chain: last vertex + back focal distance lands at z = 0
    (up to float rounding), and vertices increase along the chain."""
    for name in ("DoubleGauss", "cooketriplet", "relative_chain"):
        lens = OpticalSystem.load(os.path.join(LENS_DIR, name + ".lens"))
        z = np.array([s.z for s in lens.surfaces], np.float64)
        assert np.all(np.diff(z) > 0)
        assert abs(z[-1] + lens.surfaces[-1].thickness) < 1e-3
