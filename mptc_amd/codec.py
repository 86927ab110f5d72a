"""ctypes binding of libmptc_codec.so (include/mptc_codec.h): the CPU entropy
stage and .mptc file format around the encode path, the stream reader and
the decoder's index-word reconstruction.  No GPU."""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
CODEC_LIB = os.path.join(HERE, "host", "libmptc_codec.so")

_vp = C.c_void_p
_lib = None


class CodecError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mptc codec error {code}: {msg}")
        self.code = code


class StreamInfo(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("unique_interval", C.c_int32),
                ("search_area", C.c_int32), ("groups", C.c_int32), ("max_palette_bytes", C.c_uint32),
                ("max_palette_code", C.c_uint32), ("max_motion_code", C.c_uint32), ("max_y_code", C.c_uint32),
                ("max_c_code", C.c_uint32)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(CODEC_LIB):
        raise ImportError(f"{CODEC_LIB} is not built (run python mptc_amd/build.py)")
    L = C.CDLL(CODEC_LIB)
    L.mptc_codec_last_error.restype = C.c_char_p
    L.mptc_writer_create.argtypes = [C.c_int32] * 5 + [C.POINTER(_vp)]
    L.mptc_writer_add_frame.argtypes = [_vp, _vp, _vp, _vp, C.c_uint32]
    L.mptc_writer_finish.argtypes = [_vp, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_uint64)]
    L.mptc_writer_destroy.argtypes = [_vp]
    L.mptc_writer_destroy.restype = None
    L.mptc_reader_open.argtypes = [_vp, C.c_uint64, C.POINTER(_vp), C.POINTER(StreamInfo)]
    L.mptc_reader_next_frame.argtypes = [_vp, _vp, _vp, _vp, C.POINTER(C.c_uint32)]
    L.mptc_reader_destroy.argtypes = [_vp]
    L.mptc_reader_destroy.restype = None
    L.mptc_reconstruct_indices.argtypes = [_vp, C.c_int32, C.c_int32, C.c_int32, _vp, C.c_uint32, _vp, _vp]
    L.mptc_entropy_encode.argtypes = [_vp, C.c_uint64, _vp, C.c_uint64, C.POINTER(C.c_uint64)]
    L.mptc_entropy_decode.argtypes = [_vp, C.c_uint64, C.c_uint64, _vp]
    for name in ("mptc_writer_create", "mptc_writer_add_frame", "mptc_writer_finish", "mptc_reader_open",
                 "mptc_reader_next_frame", "mptc_reconstruct_indices", "mptc_entropy_encode",
                 "mptc_entropy_decode"):
        getattr(L, name).restype = C.c_int32
    _lib = L
    return L


def _check(rc: int):
    if rc < 0:
        raise CodecError(rc, load().mptc_codec_last_error().decode())
    return rc


def _p(a):
    return None if a is None else a.ctypes.data


def write_stream(frames, width: int, height: int, search_area: int, unique_interval: int) -> bytes:
    """frames: sequence of dicts with phys (u64), motion (u8) and unique (u32)
    arrays, in encode order.  Returns the .mptc stream."""
    L = load()
    w = _vp()
    _check(L.mptc_writer_create(width, height, search_area, unique_interval, len(frames), C.byref(w)))
    try:
        for f in frames:
            phys = np.ascontiguousarray(f["phys"], np.uint64)
            motion = np.ascontiguousarray(f["motion"], np.uint8)
            unique = np.ascontiguousarray(f["unique"], np.uint32)
            _check(L.mptc_writer_add_frame(w, _p(phys), _p(motion), _p(unique) if len(unique) else None,
                                           len(unique)))
        data = C.POINTER(C.c_uint8)()
        size = C.c_uint64()
        _check(L.mptc_writer_finish(w, C.byref(data), C.byref(size)))
        return C.string_at(data, size.value)
    finally:
        L.mptc_writer_destroy(w)


def read_stream(data: bytes):
    """Decodes a .mptc stream; returns (header info, list of frames with
    phys, motion and unique)."""
    L = load()
    buf = np.frombuffer(data, np.uint8)
    r = _vp()
    info = StreamInfo()
    _check(L.mptc_reader_open(_p(buf), len(data), C.byref(r), C.byref(info)))
    try:
        nb = (info.width // 4) * (info.height // 4)
        out = []
        while True:
            phys = np.zeros(nb, np.uint64)
            motion = np.zeros(2 * nb, np.uint8)
            unique = np.zeros(nb, np.uint32)
            n = C.c_uint32()
            if _check(L.mptc_reader_next_frame(r, _p(phys), _p(motion), _p(unique), C.byref(n))) == 1:
                break
            out.append(dict(phys=phys, motion=motion, unique=unique[:n.value].copy()))
        return info.as_dict(), out
    finally:
        L.mptc_reader_destroy(r)


def reconstruct_indices(motion, blocks_x: int, blocks_y: int, search_area: int, unique, prev_phys, phys):
    """ReconstructDXTFrame: returns phys with its index words rebuilt from the
    motion bytes, the unique palette and the previous frame."""
    out = np.array(phys, np.uint64)
    motion = np.ascontiguousarray(motion, np.uint8)
    unique = np.ascontiguousarray(unique, np.uint32)
    prev = None if prev_phys is None else np.ascontiguousarray(prev_phys, np.uint64)
    _check(load().mptc_reconstruct_indices(_p(motion), blocks_x, blocks_y, search_area,
                                           _p(unique) if len(unique) else None, len(unique), _p(prev), _p(out)))
    return out


def entropy_encode(symbols) -> bytes:
    """EntropyEncode of a byte vector: the adaptive arithmetic code."""
    sym = np.ascontiguousarray(symbols, np.uint8)
    cap = 2 * len(sym) + 64
    out = np.zeros(cap, np.uint8)
    n = C.c_uint64()
    _check(load().mptc_entropy_encode(_p(sym) if len(sym) else None, len(sym), _p(out), cap, C.byref(n)))
    return out[:n.value].tobytes()


def entropy_decode(code: bytes, n: int) -> np.ndarray:
    buf = np.frombuffer(code, np.uint8)
    out = np.zeros(n, np.uint8)
    _check(load().mptc_entropy_decode(_p(buf), len(code), n, _p(out) if n else None))
    return out
