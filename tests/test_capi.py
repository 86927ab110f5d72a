"""The C-ABI library loads, exports every function include/mptc.h declares,
and its CPU-side pieces behave; no GPU compute here."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

from mptc_amd import lib
from mptc_amd.workloads import SEQUENCES, synth_cfg


def test_library_exports_every_declared_symbol():
    L = lib.load()
    names = lib.header_functions()
    assert len(names) >= 10
    for n in names:
        assert hasattr(L, n), n


def test_codec_library_exports_every_declared_symbol():
    import re

    from mptc_amd import codec
    L = codec.load()
    text = open(os.path.join(os.path.dirname(lib.HEADER), "mptc_codec.h")).read()
    names = re.findall(r"^\s*(?:int32_t|void|const char\s*\*)\s*\**\s*(mptc_\w+)\s*\(", text, re.M)
    assert len(names) >= 9
    for n in names:
        assert hasattr(L, n), n


def test_generator_matches_oracle_copy(orc):
    for name in ("c1_s8_t0", "c2_t0", "c4_s00"):
        cfg = synth_cfg(SEQUENCES[name])
        a = lib.synth_frames(cfg, 2, 2)
        b = np.stack([orc.synth_frame(cfg, 2), orc.synth_frame(cfg, 3)])
        assert np.array_equal(a, b), name


def test_generator_properties(orc):
    from oracle import SynthCfg
    c = SynthCfg(64, 32, 1, 0, 0, 2, 0, 0)  # gradient only, no noise
    f0, f1 = orc.synth_frame(c, 0), orc.synth_frame(c, 1)
    assert np.array_equal(f0, f1)  # no rectangles, no noise, no cut: static
    c.noise_amp = 8
    g = orc.synth_frame(c, 0).astype(int)
    d = g - f0.astype(int)
    assert d.min() >= -8 and d.max() <= 8 and len(np.unique(d)) > 10


def test_create_rejects_bad_config():
    L = lib.load()
    h = C.c_void_p()
    bad = lib.Config(0, 6, 8, 8, 0)  # width not a multiple of 4
    assert L.mptc_create(C.byref(bad), C.byref(h)) == lib.MPTC_EINVAL
    bad = lib.Config(0, 8, 8, 0, 0)  # search area out of range
    assert L.mptc_create(C.byref(bad), C.byref(h)) == lib.MPTC_EINVAL
    bad = lib.Config(0, 8, 8, 8, 40000)  # threshold out of range
    assert L.mptc_create(C.byref(bad), C.byref(h)) == lib.MPTC_EINVAL
    assert b"threshold" in L.mptc_last_error()


def test_null_arguments_are_rejected():
    L = lib.load()
    assert L.mptc_fit(None, None, 4, 4, None) == lib.MPTC_EINVAL
    assert L.mptc_reencode(None, None, 4, 4, 1, None, None, None, None, None, None, None) == lib.MPTC_EINVAL
    assert L.mptc_get_stats(None, None) == lib.MPTC_EINVAL


def test_host_layer_is_built_and_runs_without_a_gpu():
    """The C++ host library and its driver link against libmptc_hip.so; the
    driver starts (no device call) and rejects bad usage."""
    import subprocess
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(here, "mptc_amd", "host", "encode_seq")
    assert os.path.exists(os.path.join(here, "mptc_amd", "host", "libmptc_host.so"))
    r = subprocess.run([exe, "--version"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "mptc" in r.stdout
    r = subprocess.run([exe, "x"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr
