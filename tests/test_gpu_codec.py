"""End to end on the GPU: frames encoded by the HIP path (through the C-ABI),
written by the CPU entropy stage into the .mptc stream, must give the
reference encoder's own file byte for byte (tests/golden/bitstreams.json),
and the stream must decode back to the GPU's blocks."""
from __future__ import annotations

import hashlib
import json
import os
import sys

import numpy as np
import pytest

from mptc_amd.workloads import SEQUENCES, synth_cfg
from tests.helpers import GOLDEN

pytestmark = pytest.mark.gpu

with open(os.path.join(GOLDEN, "bitstreams.json")) as _fh:
    CASES = {c["name"]: c for c in json.load(_fh)}


@pytest.mark.parametrize("name", sorted(CASES))
def test_gpu_frames_give_the_reference_bitstream(orc, name):
    from mptc_amd import codec, lib
    c = CASES[name]
    spec = SEQUENCES[c["sequence"]]
    frames = orc.synth_frames(synth_cfg(spec), c["frames"])
    print(f"bitstream {name}: {c['frames']} frames {spec['width']}x{spec['height']} S={spec['S']} "
          f"thr={spec['thr']} unique interval {c['unique_interval']}", flush=True)
    sys.stdout.flush()
    enc = lib.Encoder(spec["width"], spec["height"], spec["S"], spec["thr"])
    res = enc.encode_sequence(frames, spec["intra_interval"])
    enc.close()
    data = codec.write_stream(res, spec["width"], spec["height"], spec["S"], c["unique_interval"])
    assert len(data) == c["size"] and hashlib.sha256(data).hexdigest() == c["sha256"]
    _, back = codec.read_stream(data)
    for f, got in enumerate(back):
        assert np.array_equal(got["phys"], res[f]["phys"]), f"frame {f}"


def test_directory_encoder_writes_the_reference_file(orc, tmp_path):
    """The mptc_encode command (the reference's encoder arguments, PNG frames
    of a directory, GPU encode, CPU entropy stage) writes the reference's
    .mptc file for the config-1 frames."""
    import ctypes as C
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    host = os.path.join(root, "mptc_amd", "host")
    L = C.CDLL(os.path.join(host, "libmptc_host.so"))
    L.mptc_png_save.argtypes = [C.c_char_p, C.c_void_p, C.c_int32, C.c_int32]
    L.mptc_png_save.restype = C.c_int32
    c = CASES["bs_c1_s8_t0_u8"]
    spec = SEQUENCES[c["sequence"]]
    frames = orc.synth_frames(synth_cfg(spec), c["frames"])
    d = tmp_path / "frames"
    d.mkdir()
    for i, f in enumerate(frames):
        a = np.ascontiguousarray(f)
        assert L.mptc_png_save(str(d / f"frame{i:03d}.png").encode(), a.ctypes.data, spec["width"],
                               spec["height"]) == 0
    out = tmp_path / "out.mptc"
    cmd = [os.path.join(host, "mptc_encode"), str(spec["S"]), str(spec["intra_interval"]), str(c["unique_interval"]),
           str(spec["thr"]), "1", str(d), str(out)]
    print("running", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    data = out.read_bytes()
    assert len(data) == c["size"] and hashlib.sha256(data).hexdigest() == c["sha256"]
