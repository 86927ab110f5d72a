"""Shared helpers for the tests: fixture loading, hashing, comparisons."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def load_seq(name: str):
    p = os.path.join(GOLDEN, f"seq_{name}.json")
    if not os.path.exists(p):
        return None
    with open(p) as fh:
        return json.load(fh)


def load_units():
    return np.load(os.path.join(GOLDEN, "unit_vectors.npz"))


def compare_frame(res: dict, fx: dict, label: str):
    """Asserts one encoded frame matches its reference fixture, field by field."""
    assert sha(res["phys"]) == fx["phys_sha"], f"{label}: physical blocks differ"
    assert sha(res["motion"]) == fx["motion_sha"], f"{label}: motion bytes differ"
    assert sha(res["unique"]) == fx["unique_sha"], f"{label}: unique palette differs"
    assert len(res["unique"]) == fx["n_unique"], f"{label}: unique count differs"
    if "eps" in res and res["eps"] is not None:
        assert sha(res["eps"]) == fx["eps_sha"], f"{label}: logical endpoints differ"


def compare_results(a: dict, b: dict, label: str):
    for key in ("phys", "motion", "unique", "eps"):
        if key in a and key in b and a[key] is not None and b[key] is not None:
            if not np.array_equal(a[key], b[key]):
                bad = np.flatnonzero(np.asarray(a[key]).ravel() != np.asarray(b[key]).ravel())
                raise AssertionError(f"{label}: {key} differs at {len(bad)} entries, first {bad[:8]}")
