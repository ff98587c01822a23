"""Loaders for the committed golden fixtures (tests/golden/)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _fx(v):
    return float.fromhex(v) if isinstance(v, str) else float(v)


def arr(rows):
    """Hex-float nested lists -> float64 array (exact)."""
    if rows and isinstance(rows[0], list):
        return np.array([[_fx(v) for v in r] for r in rows], dtype=np.float64)
    return np.array([_fx(v) for v in rows], dtype=np.float64)


def golden():
    with open(os.path.join(GOLDEN, "golden_vectors.json")) as f:
        return json.load(f)


def ref_fixture(name):
    with open(os.path.join(GOLDEN, "ref_fixtures", name)) as f:
        return json.load(f)


def bits(a):
    """float64 array -> uint64 view, for bit-exact comparisons."""
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.uint64)
