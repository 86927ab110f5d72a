"""PNG frame input of the host layer (mptc_amd/host/png.cpp): pixels as the
reference's stbi_load(path, ..., 3) gives them (codec/dxt_image.cpp:396), for
the PNG colour types and depths a frame may come in, checked against PIL's
decoder and against the files the reference's own writer makes."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import pytest

HOST_LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mptc_amd", "host",
                        "libmptc_host.so")


@pytest.fixture(scope="module")
def L():
    L = C.CDLL(HOST_LIB)
    L.mptc_png_load.argtypes = [C.c_char_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
    L.mptc_png_load.restype = C.c_int32
    L.mptc_png_save.argtypes = [C.c_char_p, C.c_void_p, C.c_int32, C.c_int32]
    L.mptc_png_save.restype = C.c_int32
    return L


def load(L, path):
    buf = np.zeros(4096 * 4096 * 3, np.uint8)
    w, h = C.c_int32(), C.c_int32()
    rc = L.mptc_png_load(str(path).encode(), buf.ctypes.data, buf.size, C.byref(w), C.byref(h))
    if rc:
        raise ValueError(rc)
    return buf[:w.value * h.value * 3].reshape(h.value, w.value, 3)


def test_colour_types_match_pil(L, tmp_path):
    from PIL import Image
    rng = np.random.default_rng(2)
    rgb = rng.integers(0, 256, (37, 53, 3)).astype(np.uint8)
    cases = {
        "rgb": Image.fromarray(rgb, "RGB"),
        "rgba": Image.fromarray(np.dstack([rgb, rng.integers(0, 256, (37, 53))]).astype(np.uint8), "RGBA"),
        "grey": Image.fromarray(rgb[..., 0], "L"),
        "grey_alpha": Image.merge("LA", [Image.fromarray(rgb[..., 0]), Image.fromarray(rgb[..., 1])]),
        "palette": Image.fromarray(rgb, "RGB").quantize(colors=37),
        "bits1": Image.fromarray((rgb[..., 0] > 127).astype(np.uint8) * 255, "L").convert("1"),
    }
    for name, im in cases.items():
        for opt in ({}, {"optimize": True}):
            p = tmp_path / f"{name}.png"
            im.save(p, **opt)  # PIL picks the row filters (optimize: adaptive)
            want = np.asarray(Image.open(p).convert("RGB"))
            assert np.array_equal(load(L, p), want), (name, opt)


def test_sixteen_bit_takes_the_high_byte(L, tmp_path):
    from PIL import Image
    rng = np.random.default_rng(4)
    g16 = rng.integers(0, 65536, (20, 30)).astype(np.uint16)
    p = tmp_path / "g16.png"
    Image.fromarray(g16).save(p)  # uint16 -> 16-bit greyscale PNG
    got = load(L, p)
    hi = (g16 >> 8).astype(np.uint8)
    assert np.array_equal(got, np.dstack([hi, hi, hi]))


def test_own_writer_round_trips(L, tmp_path):
    rng = np.random.default_rng(6)
    rgb = np.ascontiguousarray(rng.integers(0, 256, (64, 96, 3)).astype(np.uint8))
    p = tmp_path / "w.png"
    assert L.mptc_png_save(str(p).encode(), rgb.ctypes.data, 96, 64) == 0
    assert np.array_equal(load(L, p), rgb)
    from PIL import Image
    assert np.array_equal(np.asarray(Image.open(p).convert("RGB")), rgb)


def test_reference_writer_files(L, tmp_path):
    """Frames written by the reference's own PNG writer (stb_image_write,
    compiled from its sources) read back exactly."""
    import oracle
    if not os.path.exists(oracle.REF_CODEC_SO):
        pytest.skip("reference harness not built here")
    R = C.CDLL(oracle.REF_CODEC_SO)
    from mptc_amd.workloads import SEQUENCES, synth_cfg
    O = oracle.Oracle()
    frame = np.ascontiguousarray(O.synth_frame(synth_cfg(SEQUENCES["c1_s8_t0"]), 3))
    p = tmp_path / "ref.png"
    assert R.ref_write_png(str(p).encode(), frame.ctypes.data_as(C.c_void_p), 256, 256) == 0
    assert np.array_equal(load(L, p), frame)


def test_bad_files_are_rejected(L, tmp_path):
    p = tmp_path / "junk.png"
    p.write_bytes(b"not a png at all")
    with pytest.raises(ValueError):
        load(L, p)
    with pytest.raises(ValueError):
        load(L, tmp_path / "missing.png")
