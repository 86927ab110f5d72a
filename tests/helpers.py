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


def pack565(r: int, g: int, b: int) -> int:
    return ((r & 0xF8) << 8) | ((g & 0xFC) << 3) | (b >> 3)


def block_pixels(rgb: np.ndarray, bx: int, by: int) -> np.ndarray:
    return np.ascontiguousarray(rgb[4 * by:4 * by + 4, 4 * bx:4 * bx + 4]).reshape(48)


def check_reuse_frame(O, rgb, own_fit, phys, motion, unique, eps, is_intra, ref_phys, S, thr, label=""):
    """Checks an encoded frame as a decoder and as the reference's accept test
    see it, whatever search chose its reuses: the unique palette holds the own
    words of the unique blocks in raster order; every reused block's index word
    equals the word at its motion offset (the previous frame's for inter, this
    frame's final word of an earlier block for intra); its error delta, by the
    oracle, passes the threshold; its endpoints are the oracle's re-fit (or
    its own fit when the word is its own).  Returns (n_inter, n_intra)."""
    h, w, _ = rgb.shape
    bw, bh = w // 4, h // 4
    words = (np.asarray(phys, np.uint64) >> np.uint64(32)).astype(np.uint32)
    own_words = (np.asarray(own_fit, np.uint64) >> np.uint64(32)).astype(np.uint32)
    mx = np.asarray(motion[0::2]).astype(np.int64)
    my = np.asarray(motion[1::2]).astype(np.int64)
    uniq = (mx == 255) & (my == 255)
    assert np.array_equal(np.asarray(phys)[uniq], np.asarray(own_fit)[uniq]), f"{label}: unique block changed"
    assert np.array_equal(np.asarray(unique), own_words[uniq]), f"{label}: unique palette"
    n_inter = n_intra = 0
    for k in np.flatnonzero(~uniq):
        bx, by = int(k % bw), int(k // bw)
        if mx[k] & 0x80:
            assert not is_intra, f"{label}: inter reuse in an intra frame"
            x, y = int(mx[k] & 0x7F), int(my[k] & 0x7F)
            assert 0 <= x < 2 * S and 0 <= y < 2 * S, f"{label}: block {k} inter offset out of window"
            j, i = by + y - S, bx + x - S
            assert 0 <= j < bh and 0 <= i < bw, f"{label}: block {k} inter offset out of frame"
            src = int(np.asarray(ref_phys)[j * bw + i]) >> 32
            n_inter += 1
        else:
            x, y = int(mx[k]), int(my[k])
            j, i = by + y - 2 * S + 1, bx + x - S
            assert 0 <= j < bh and 0 <= i < bw, f"{label}: block {k} intra offset out of frame"
            assert j < by or (j == by and i < bx), f"{label}: block {k} intra offset not causal"
            assert by - j <= 2 * S - 1 and -S <= i - bx < S, f"{label}: block {k} intra offset out of window"
            src = int(words[j * bw + i])
            n_intra += 1
        assert int(words[k]) == src, f"{label}: block {k} word differs from its source"
        px = block_pixels(rgb, bx, by)
        ok, e, reassign, ep = O.eval_candidate(int(own_fit[k]), px, src)
        assert ok and e <= thr, f"{label}: block {k} accepted with delta {e} > {thr} (valid {ok})"
        ep = [int(v) for v in ep]
        if src == int(own_words[k]):
            assert int(phys[k]) == int(own_fit[k]), f"{label}: block {k} own word re-fit"
        else:
            exp = pack565(*ep[:3]) | (pack565(*ep[3:]) << 16) | (src << 32)
            assert int(phys[k]) == exp, f"{label}: block {k} endpoints differ from the re-fit"
            if eps is not None:
                assert [int(v) for v in eps[6 * k:6 * k + 6]] == ep, f"{label}: block {k} eps"
    return n_inter, n_intra
