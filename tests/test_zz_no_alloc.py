"""The allocation-free entry point (mptc_encode_frame_noalloc) against the
context entry point (mptc_encode_frame_device), through the C-ABI: the same
frames give the same bytes, on consecutive calls with the larger input first
(1280x720: intra then inter), then a smaller one (640x360, two thresholds and
search areas), and arguments beyond its limits are rejected.  Named to be
collected last."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from mptc_amd.workloads import synth_cfg
from tests.helpers import compare_results

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _bufs(hip, nb):
    return dict(phys=hip.Buffer(8 * nb), motion=hip.Buffer(2 * nb), unique=hip.Buffer(4 * nb),
                n_unique=hip.Buffer(4), eps=hip.Buffer(6 * nb), kind=hip.Buffer(nb))


def _read(b, nb):
    n = int(b["n_unique"].download(np.uint32, 1)[0])
    return dict(phys=b["phys"].download(np.uint64, nb), motion=b["motion"].download(np.uint8, 2 * nb),
                unique=b["unique"].download(np.uint32, n), eps=b["eps"].download(np.uint8, 6 * nb),
                kind=b["kind"].download(np.uint8, nb))


def test_noalloc_matches_context_entry_point():
    from mptc_amd import hip, lib
    hip.set_device(0)
    cases = [(1280, 720, 8, 0, 11), (640, 360, 8, 3, 12), (640, 360, 5, 8, 13)]
    for w, h, S, thr, seed in cases:
        nb = (w // 4) * (h // 4)
        frames = lib.synth_frames(synth_cfg(dict(width=w, height=h, seed=seed, n_rects=24, vel_mode=1, vel=4,
                                                 cut_permille=10)), 0, 2)
        d_rgb = hip.Buffer.from_array(frames)
        fb = w * h * 3
        enc = lib.Encoder(w, h, S, thr)
        want = [_bufs(hip, nb) for _ in range(2)]
        got = [_bufs(hip, nb) for _ in range(2)]
        stream = hip.Stream()
        for f in range(2):
            log(f"noalloc {w}x{h} S={S} thr={thr} frame {f}")
            a, b = want[f], got[f]
            enc.encode_frame_device(d_rgb.offset(f * fb), w, h, f == 0, None if f == 0 else want[0]["phys"],
                                    a["phys"], a["motion"], a["unique"], a["n_unique"], a["eps"], a["kind"])
            lib.encode_frame_noalloc(0, d_rgb.offset(f * fb), w, h, S, thr, f == 0,
                                     None if f == 0 else got[0]["phys"], b["phys"], b["motion"], b["unique"],
                                     b["n_unique"], b["eps"], b["kind"], stream=stream.handle)
            ra, rb = _read(a, nb), _read(b, nb)
            compare_results(rb, ra, f"noalloc {w}x{h} frame {f}")
            assert np.array_equal(rb["kind"], ra["kind"]), f"noalloc {w}x{h} frame {f}: kinds"
            assert (rb["kind"] > 0).any() or f == 0
        stream.destroy()
        enc.close()


def test_noalloc_rejects_inputs_beyond_its_limits():
    from mptc_amd import hip, lib
    hip.set_device(0)
    nb = 480 * 270
    b = _bufs(hip, nb)
    rgb = hip.Buffer(1920 * 1080 * 3)
    with pytest.raises(lib.MptcError):  # 129,600 blocks > 65,536
        lib.encode_frame_noalloc(0, rgb, 1920, 1080, 8, 0, True, None, b["phys"], b["motion"], b["unique"],
                                 b["n_unique"])
    with pytest.raises(lib.MptcError):  # S above 8
        lib.encode_frame_noalloc(0, rgb, 640, 360, 9, 0, True, None, b["phys"], b["motion"], b["unique"],
                                 b["n_unique"])
    with pytest.raises(lib.MptcError):  # inter frame without its reference
        lib.encode_frame_noalloc(0, rgb, 640, 360, 8, 0, False, None, b["phys"], b["motion"], b["unique"],
                                 b["n_unique"])
    with pytest.raises(lib.MptcError):  # width not a multiple of 4
        lib.encode_frame_noalloc(0, rgb, 642, 360, 8, 0, True, None, b["phys"], b["motion"], b["unique"],
                                 b["n_unique"])
