"""The CPU entropy stage and .mptc format (mptc_amd/host/codec.cpp) and the
decoder's index-word reconstruction, against the reference's own bitstreams
(tests/golden/bitstreams.json, written by the reference's CompressMultiUnique
compiled from its sources) and against what the format fixes independently.

The encoded frames come from the CPU oracle, which the sequence fixtures pin
to the reference encoder bit for bit."""
from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

from mptc_amd import codec
from mptc_amd.workloads import SEQUENCES, synth_cfg
from tests.helpers import GOLDEN


def _cases():
    with open(os.path.join(GOLDEN, "bitstreams.json")) as fh:
        return json.load(fh)


CASES = {c["name"]: c for c in _cases()}


@pytest.fixture(scope="module")
def encoded(orc):
    """The oracle's encode of each bitstream case's frames."""
    out = {}
    for c in CASES.values():
        spec = SEQUENCES[c["sequence"]]
        frames = orc.synth_frames(synth_cfg(spec), c["frames"])
        out[c["name"]] = (spec, frames, orc.encode_sequence(frames, spec["S"], spec["thr"], spec["intra_interval"]))
    return out


@pytest.mark.parametrize("name", sorted(CASES))
def test_stream_matches_reference_bitstream(encoded, name):
    c = CASES[name]
    spec, _, res = encoded[name]
    data = codec.write_stream(res, spec["width"], spec["height"], spec["S"], c["unique_interval"])
    assert data[:34].hex() == c["header"], "header (sizes, group count, maxima)"
    assert len(data) == c["size"]
    assert hashlib.sha256(data).hexdigest() == c["sha256"]


@pytest.mark.parametrize("name", sorted(CASES))
def test_stream_round_trip(encoded, name):
    """The reader rebuilds every written frame: motion bytes and palettes
    exactly, index words through the reconstruction, endpoints through the
    inverse wavelet (exact: no coefficient of these frames leaves int8)."""
    c = CASES[name]
    spec, _, res = encoded[name]
    ui = c["unique_interval"]
    data = codec.write_stream(res, spec["width"], spec["height"], spec["S"], ui)
    info, frames = codec.read_stream(data)
    assert info["width"] == spec["width"] and info["height"] == spec["height"]
    assert info["groups"] == len(res) // ui and info["unique_interval"] == ui
    assert len(frames) == (len(res) // ui) * ui  # frames after the last whole group are not written
    for f, got in enumerate(frames):
        want = res[f]
        assert np.array_equal(got["motion"], want["motion"]), f"frame {f} motion"
        assert np.array_equal(got["unique"], want["unique"]), f"frame {f} palette"
        assert np.array_equal(got["phys"], want["phys"]), f"frame {f} blocks"


def test_reconstruction_rebuilds_index_words(encoded):
    """ReconstructDXTFrame on every frame: motion bytes + unique palette +
    previous frame give back the encoder's index words (endpoints kept)."""
    for name in ("bs_c1_s8_t8_u4", "bs_w512_t8_u2"):
        spec, _, res = encoded[name]
        bw, bh = spec["width"] // 4, spec["height"] // 4
        prev = None
        for f, r in enumerate(res):
            blank = np.asarray(r["phys"], np.uint64) & np.uint64(0xFFFFFFFF)
            got = codec.reconstruct_indices(r["motion"], bw, bh, spec["S"], r["unique"], prev, blank)
            assert np.array_equal(got, r["phys"]), f"{name} frame {f}"
            prev = r["phys"]


def test_reconstruction_rejects_bad_streams(encoded):
    spec, _, res = encoded["bs_c1_s8_t8_u4"]
    bw, bh = spec["width"] // 4, spec["height"] // 4
    r = res[1]
    with pytest.raises(codec.CodecError):  # a unique block without a palette word
        codec.reconstruct_indices(r["motion"], bw, bh, spec["S"], r["unique"][:-1], res[0]["phys"], r["phys"])
    with pytest.raises(codec.CodecError):  # inter reuse without a previous frame
        codec.reconstruct_indices(r["motion"], bw, bh, spec["S"], r["unique"], None, r["phys"])
    m = np.array(r["motion"])
    m[0:2] = [0, 0]  # block 0 pointing at an intra source outside the frame
    with pytest.raises(codec.CodecError):
        codec.reconstruct_indices(m, bw, bh, spec["S"], r["unique"], res[0]["phys"], r["phys"])


def test_writer_rejects_grids_the_wavelet_cannot_take():
    with pytest.raises(codec.CodecError):  # 480x270 blocks (1080p): not a multiple of 64
        codec.write_stream([], 1920, 1080, 8, 8)


def test_ycocg_and_wavelet_are_lossless_where_coefficients_fit():
    """Closed-form pins of the endpoint pipeline, through the writer and reader:
    random endpoint planes survive the stream when their wavelet coefficients
    fit in int8 (smooth planes), and a plane that overflows it comes back
    different (the narrowing the reference applies)."""
    rng = np.random.default_rng(3)
    bw = bh = 64
    yy, xx = np.mgrid[0:bh, 0:bw]
    smooth = ((xx // 8 + yy // 16) % 32).astype(np.uint64)  # r = b: small chroma
    e1 = (smooth << np.uint64(11)) | (np.uint64(20) << np.uint64(5)) | smooth
    e2 = (((31 - smooth) << np.uint64(11)) | (((xx + yy) // 4 % 64).astype(np.uint64) << np.uint64(5)) |
          (31 - smooth))
    words = rng.integers(0, 2**32, bw * bh, dtype=np.uint64)
    phys = (e1.reshape(-1) | (e2.reshape(-1) << np.uint64(16)) | (words << np.uint64(32))).astype(np.uint64)
    motion = np.full(2 * bw * bh, 255, np.uint8)
    frame = dict(phys=phys, motion=motion, unique=(phys >> np.uint64(32)).astype(np.uint32))
    _, back = codec.read_stream(codec.write_stream([frame], 256, 256, 8, 1))
    assert np.array_equal(back[0]["phys"], phys)
    # a checkerboard of magenta (31, 0, 31) and green (0, 63, 0): Cg alternates
    # -63 / +63, so the high-high coefficients reach +-252, beyond int8
    checker = np.where((xx + yy) % 2 == 0, (31 << 11) | 31, 63 << 5).astype(np.uint64).reshape(-1)
    rough = checker | (words << np.uint64(32))
    _, back = codec.read_stream(codec.write_stream([dict(frame, phys=rough)], 256, 256, 8, 1))
    assert not np.array_equal(back[0]["phys"] & np.uint64(0xFFFF), rough & np.uint64(0xFFFF))


def test_arithmetic_coder_round_trips_skewed_data():
    """The coder and its decoder agree on byte streams of very different
    statistics (constant, two symbols, uniform, geometric, a change of
    statistics half way, 0 and 1 symbol), and the code stays within a few
    percent of the data's order-0 entropy where that is small."""
    rng = np.random.default_rng(5)
    n = 50000
    streams = {
        "constant": np.full(n, 200, np.uint8),
        "binary": np.where(rng.random(n) < 0.9, 3, 250).astype(np.uint8),
        "uniform": rng.integers(0, 256, n).astype(np.uint8),
        "geometric": np.minimum(rng.geometric(0.3, n), 255).astype(np.uint8),
        "switch": np.concatenate([np.full(n // 2, 7, np.uint8), rng.integers(0, 256, n // 2).astype(np.uint8)]),
        "one": np.array([255], np.uint8),
        "empty": np.zeros(0, np.uint8),
    }
    for kind, sym in streams.items():
        code = codec.entropy_encode(sym)
        assert np.array_equal(codec.entropy_decode(code, len(sym)), sym), kind
        if len(sym) > 1000:
            p = np.bincount(sym, minlength=256) / len(sym)
            ent = -(p[p > 0] * np.log2(p[p > 0])).sum() * len(sym) / 8
            # the adaptive model never drives a symbol's probability to 1
            # (counts are halved at 2^15): allow ~0.1 bit per symbol on top
            bound = {"uniform": 1.01, "switch": 1.1}.get(kind, 1.05)
            assert len(code) <= ent * bound + len(sym) / 80 + 64, (kind, len(code), ent)
            assert len(code) < 0.9 * len(sym) or kind == "uniform", kind  # skewed data compresses
