"""Shared loader for tests/golden/trace_path_goldens.npz: the reference's
diagnostic traces (RayPath overloads, trace.cpp:264-395) of 96 rays through
five pairs and the primary path of the synthetic 12-, 24- and 30-surface
lenses and the 12-surface one with surface 5 muted."""
import os

import numpy as np

from oracle.pyoracle import PackedLens

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "trace_path_goldens.npz")
NAMES = ("c1", "c2", "c4", "c1m")


def load():
    return np.load(GOLDEN)


def cases(z):
    """(name, a, b, rays) for every stored trace; b = -1 is the primary path."""
    out = []
    for name in NAMES:
        for a, b in z[f"pairs_{name}"]:
            out.append((name, int(a), int(b)))
    return out


def packed_lens(z, name):
    return PackedLens(z[f"lens_{name}_surfs"], z[f"lens_{name}_arena"])


def golden(z, name, a, b):
    k = f"path_{name}_{a}_{b}"
    return z[k + "_ev"], z[k + "_n"], z[k + "_pos"], z[k + "_w"], z[k + "_st"]
