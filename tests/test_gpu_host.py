"""The C++ host layer (mptc_amd/host): the sequence encoder with the
reference's intra schedule, run as a separate program, must reproduce the
reference's per-frame fixtures."""
from __future__ import annotations

import os
import subprocess
import sys

import numpy as np
import pytest

from mptc_amd.workloads import SEQUENCES
from tests.helpers import load_seq, sha

pytestmark = pytest.mark.gpu

BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mptc_amd", "host", "encode_seq")


@pytest.mark.parametrize("name", ["c1_s8_t0", "c1_s8_t8", "c1_s16_t0"])
def test_host_sequence_encoder_matches_reference(tmp_path, name):
    spec = SEQUENCES[name]
    fx = load_seq(name)
    assert fx is not None and os.path.exists(BIN)
    print(f"host encoder {name}", flush=True)
    sys.stdout.flush()
    args = [BIN, str(tmp_path), spec["width"], spec["height"], spec["seed"], spec["n_rects"], spec["vel_mode"],
            spec["vel"], spec["cut_permille"], spec["frames"], spec["S"], spec["thr"], spec["intra_interval"]]
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for f, fr in enumerate(fx["frames"]):
        base = tmp_path / f"frame{f:03d}"
        phys = np.fromfile(f"{base}.phys", np.uint64)
        motion = np.fromfile(f"{base}.motion", np.uint8)
        unique = np.fromfile(f"{base}.unique", np.uint32)
        eps = np.fromfile(f"{base}.eps", np.uint8)
        assert sha(phys) == fr["phys_sha"], f"frame {f}"
        assert sha(motion) == fr["motion_sha"], f"frame {f}"
        assert sha(unique) == fr["unique_sha"] and len(unique) == fr["n_unique"], f"frame {f}"
        assert sha(eps) == fr["eps_sha"], f"frame {f}"
