"""Named synthetic workloads (inputs from mptc_amd/csrc/framegen.h).

Each entry fixes the frame generator settings and the encoder settings of
one workload: frame size, seed, rectangles, motion, scene cuts, search area S
(window side 2S), error threshold (in the reference's floor(SSE/48) units,
codec/dxt_image.cpp:275) and the intra-frame interval (codec/codec.cpp:1462).
`golden_frames` is how many leading frames carry reference fixtures.
"""
from __future__ import annotations

import ctypes as C


class SynthCfg(C.Structure):
    """mptc_synth_cfg of mptc_amd/csrc/framegen.h"""
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("seed", C.c_uint64),
                ("n_rects", C.c_int32), ("vel_mode", C.c_int32), ("vel", C.c_int32),
                ("noise_amp", C.c_int32), ("cut_permille", C.c_int32)]


def _seq(w, h, frames, seed, rects, vel_mode, vel, cut, S, thr, intra=8, golden=None, full=False):
    return dict(width=w, height=h, frames=frames, seed=seed, n_rects=rects, vel_mode=vel_mode,
                vel=vel, cut_permille=cut, S=S, thr=thr, intra_interval=intra,
                golden_frames=frames if golden is None else golden, full_arrays=full)


SEQUENCES = {
    # 8 x 256x256, 4 rectangles moving 2 px/frame, 16-block window (S=8)
    "c1_s8_t0": _seq(256, 256, 8, 1, 4, 0, 2, 0, 8, 0, full=True),
    "c1_s8_t8": _seq(256, 256, 8, 1, 4, 0, 2, 0, 8, 8),
    "c1_s16_t0": _seq(256, 256, 8, 1, 4, 0, 2, 0, 16, 0),
    # 64 x 1920x1080, 32 rectangles with velocities in [-4, 4], 1 % scene cuts
    "c2_t0": _seq(1920, 1080, 64, 2, 32, 1, 4, 10, 8, 0),
    "c2_t8": _seq(1920, 1080, 64, 2, 32, 1, 4, 10, 8, 8),
    # 32 x 3840x2160, 64x64-block window (S=32), threshold sweep; fixtures for 2 frames
    "c3_t0": _seq(3840, 2160, 32, 3, 32, 1, 4, 10, 32, 0, golden=2),
    "c3_t8": _seq(3840, 2160, 32, 3, 32, 1, 4, 10, 32, 8, golden=2),
    "c3_t32": _seq(3840, 2160, 32, 3, 32, 1, 4, 10, 32, 32, golden=2),
    "c3_t128": _seq(3840, 2160, 32, 3, 32, 1, 4, 10, 32, 128, golden=2),
}

# 4 x 512x256 (a 128x64 block grid: two endpoint-wavelet blocks per row), for the .mptc stream tests
SEQUENCES["w512_t8"] = _seq(512, 256, 4, 7, 8, 1, 3, 0, 8, 8, golden=0)

# 12 uneven streams encoded together: 3 each of 640x360, 1280x720, 1920x1080, 3840x2160
STREAM_SIZES = [(640, 360), (1280, 720), (1920, 1080), (3840, 2160)]
for _i in range(12):
    _w, _h = STREAM_SIZES[_i // 3]
    SEQUENCES[f"c4_s{_i:02d}"] = _seq(_w, _h, 30, 100 + _i, 32, 1, 4, 10, 8, 8, golden=3)

BATCH_STREAMS = [f"c4_s{i:02d}" for i in range(12)]


def synth_cfg(spec: dict) -> SynthCfg:
    return SynthCfg(spec["width"], spec["height"], spec["seed"], spec["n_rects"],
                    spec["vel_mode"], spec["vel"], 8, spec["cut_permille"])
