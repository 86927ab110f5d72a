"""ctypes binding of libmptc_hip.so (include/mptc.h).

The library is built in-tree by mptc_amd/build.py; loading fails loudly if it
is missing -- there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmptc_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mptc.h")

MPTC_OK, MPTC_EINVAL, MPTC_EHIP, MPTC_ENOMEM, MPTC_ESTATE = 0, -1, -2, -3, -4


class MptcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"mptc error {code}: {msg}")
        self.code = code


class Config(C.Structure):
    _fields_ = [("device", C.c_int32), ("max_width", C.c_int32), ("max_height", C.c_int32),
                ("search_area", C.c_int32), ("err_threshold", C.c_int32),
                ("max_batch_frames", C.c_int32), ("max_batch_blocks", C.c_int32)]


class Tuning(C.Structure):
    _fields_ = [("wg_size", C.c_int32), ("tile_w", C.c_int32), ("pass_batch", C.c_int32),
                ("use_graph", C.c_int32), ("win_chunks", C.c_int32), ("prune", C.c_int32),
                ("itile_w", C.c_int32)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class FrameIO(C.Structure):
    _fields_ = [("d_rgb", C.c_void_p), ("width", C.c_int32), ("height", C.c_int32),
                ("is_intra", C.c_int32), ("d_ref_phys", C.c_void_p), ("d_phys_out", C.c_void_p),
                ("d_motion_out", C.c_void_p), ("d_unique_out", C.c_void_p),
                ("d_n_unique_out", C.c_void_p), ("d_eps_out", C.c_void_p), ("d_kind_out", C.c_void_p)]


class FrameStats(C.Structure):
    _fields_ = [("n_blocks", C.c_int64), ("n_inter", C.c_int64), ("n_intra", C.c_int64),
                ("n_unique", C.c_int64), ("inter_candidates", C.c_int64),
                ("intra_candidates", C.c_int64), ("reeval_candidates", C.c_int64),
                ("intra_iterations", C.c_int32), ("is_intra", C.c_int32),
                ("exact_inter_candidates", C.c_int64), ("exact_intra_candidates", C.c_int64)]

    def as_dict(self):
        return {f[0]: getattr(self, f[0]) for f in self._fields_}


class TrainCfg(C.Structure):
    _fields_ = [("batch", C.c_int32), ("steps", C.c_int32), ("lr", C.c_float),
                ("seed_weight", C.c_float), ("seed", C.c_uint64), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("lr_final", C.c_float)]


NET_INTER, NET_INTRA = 0, 1
NET_IN, NET_HIDDEN, NET_OUT, NET_LOGITS, NET_PARAMS = 48, 256, 262, 256, 145670
SIM_BINS = 17  # differing pixel indices 0..16 (mptc_net_set_similarity)
# offsets in the flat parameter vector: W1 [48][256], b1, W2 [256][256], b2, W3 [256][262], b3
NET_OFF_W3 = NET_IN * NET_HIDDEN + NET_HIDDEN + NET_HIDDEN * NET_HIDDEN + NET_HIDDEN
NET_OFF_B3 = NET_OFF_W3 + NET_HIDDEN * NET_OUT
assert NET_OFF_B3 + NET_OUT == NET_PARAMS


def similarity_scores(tries: np.ndarray, accepts: np.ndarray) -> np.ndarray:
    """Log of the smoothed acceptance rate by number of differing indices:
    an additive ranking score (log P(accept | h)), from similarity_counts."""
    t = np.asarray(tries, np.float64)
    a = np.asarray(accepts, np.float64)
    return np.log((a + 0.5) / (t + 1.0)).astype(np.float32)


def zero_logit_layer(params: np.ndarray) -> np.ndarray:
    """A copy of a parameter vector with the logit columns of W3 and the logit
    biases of b3 at zero (the network's logits then equal its base scores)."""
    p = np.array(params, np.float32, copy=True)
    p[NET_OFF_W3:NET_OFF_B3].reshape(NET_HIDDEN, NET_OUT)[:, :NET_LOGITS] = 0.0
    p[NET_OFF_B3:NET_OFF_B3 + NET_LOGITS] = 0.0
    return p
SEARCH_EXHAUSTIVE, SEARCH_LEARNED_INTER, SEARCH_LEARNED = 0, 1, 2

_u8p = C.POINTER(C.c_uint8)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)
_vp = C.c_void_p

_lib = None


def header_functions() -> list[str]:
    """Names of the functions declared in include/mptc.h."""
    text = open(HEADER).read()
    return re.findall(r"^\s*(?:int32_t|void|const char\s*\*)\s*\**\s*(mptc_\w+)\s*\(", text, re.M)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built (run python mptc_amd/build.py)")
    L = C.CDLL(LIB_PATH)
    L.mptc_last_error.restype = C.c_char_p
    L.mptc_create.argtypes = [C.POINTER(Config), C.POINTER(C.c_void_p)]
    L.mptc_destroy.argtypes = [_vp]
    L.mptc_destroy.restype = None
    L.mptc_set_params.argtypes = [_vp, C.c_int32, C.c_int32]
    L.mptc_fit.argtypes = [_vp, _vp, C.c_int32, C.c_int32, _vp]
    L.mptc_reencode.argtypes = [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, _vp, _vp, _vp, _vp,
                                _vp, _vp, _vp]
    L.mptc_encode_frame_device.argtypes = [_vp, _vp, C.c_int32, C.c_int32, C.c_int32, _vp, _vp,
                                           _vp, _vp, _vp, _vp, _vp, _vp]
    L.mptc_get_stats.argtypes = [_vp, C.POINTER(FrameStats)]
    L.mptc_synchronize.argtypes = [_vp]
    L.mptc_profile_enable.argtypes = [_vp, C.c_int32]
    L.mptc_profile_read.argtypes = [_vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double)]
    L.mptc_synth_frames.argtypes = [_vp, C.c_int32, C.c_int32, _vp]
    L.mptc_encode_batch_device.argtypes = [_vp, C.c_int32, C.POINTER(FrameIO), _vp]
    L.mptc_get_tuning.argtypes = [_vp, C.POINTER(Tuning)]
    L.mptc_set_tuning.argtypes = [_vp, C.POINTER(Tuning)]
    L.mptc_autotune.argtypes = [_vp, C.c_int32, C.POINTER(FrameIO), C.c_int32, C.POINTER(Tuning),
                                C.POINTER(C.c_double)]
    L.mptc_net_init.argtypes = [_vp, C.c_int32, C.c_uint64]
    L.mptc_net_get_params.argtypes = [_vp, C.c_int32, _vp]
    L.mptc_net_set_params.argtypes = [_vp, C.c_int32, _vp]
    L.mptc_net_forward.argtypes = [_vp, C.c_int32, _vp, C.c_int32, _vp]
    L.mptc_net_gradients.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, C.c_int32, C.c_float, _vp,
                                     C.POINTER(C.c_float)]
    L.mptc_net_train.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, C.c_int32, C.POINTER(TrainCfg), _vp]
    L.mptc_record_pairs.argtypes = [_vp, C.c_int32]
    L.mptc_collect_pairs.argtypes = [_vp, C.c_int32, _vp, _vp, _vp, _vp, C.c_int32]
    L.mptc_set_search_mode.argtypes = [_vp, C.c_int32, C.c_int32]
    L.mptc_set_prior.argtypes = [_vp, C.c_int32, _vp]
    L.mptc_get_proposals.argtypes = [_vp, C.c_int32, _vp, C.c_int64]
    L.mptc_net_set_base.argtypes = [_vp, C.c_int32, _vp]
    L.mptc_net_set_similarity.argtypes = [_vp, C.c_int32, _vp]
    L.mptc_get_similarity_counts.argtypes = [_vp, C.c_int32, _vp]
    L.mptc_frame_sse.argtypes = [_vp, C.c_int32, C.c_int32, _vp, _vp, _vp]
    L.mptc_encode_frame_noalloc.argtypes = [C.c_int32, _vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32,
                                            C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]
    for name in header_functions():
        if name not in ("mptc_destroy", "mptc_last_error"):
            getattr(L, name).restype = C.c_int32
    _lib = L
    return L


def _check(rc: int):
    if rc != MPTC_OK:
        raise MptcError(rc, load().mptc_last_error().decode())


def _ptr(a) -> int | None:
    if a is None:
        return None
    if isinstance(a, np.ndarray):
        assert a.flags.c_contiguous
        return a.ctypes.data
    return a.data_ptr()  # torch tensor (device or host)


def synth_frames(cfg, first: int, n: int) -> np.ndarray:
    out = np.zeros((n, cfg.height, cfg.width, 3), np.uint8)
    _check(load().mptc_synth_frames(C.byref(cfg), first, n, out.ctypes.data))
    return out


class Encoder:
    """One context: device work buffers for frames up to max_width x max_height."""

    def __init__(self, max_width: int, max_height: int, search_area: int = 8,
                 err_threshold: int = 0, device: int = 0, max_batch_frames: int = 1,
                 max_batch_blocks: int = 0):
        L = load()
        self.L = L
        self.cfg = Config(device, max_width, max_height, search_area, err_threshold,
                          max_batch_frames, max_batch_blocks)
        h = C.c_void_p()
        _check(L.mptc_create(C.byref(self.cfg), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            self.L.mptc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, search_area: int, err_threshold: int):
        _check(self.L.mptc_set_params(self.h, search_area, err_threshold))

    # ---- host-buffer entry points (the DXTImage ctor fit loop and Reencode) ----
    def fit(self, rgb: np.ndarray) -> np.ndarray:
        h, w, _ = rgb.shape
        rgb = np.ascontiguousarray(rgb, np.uint8)
        out = np.zeros((h // 4) * (w // 4), np.uint64)
        _check(self.L.mptc_fit(self.h, rgb.ctypes.data, w, h, out.ctypes.data))
        return out

    def reencode(self, rgb: np.ndarray, phys_fit: np.ndarray, is_intra: bool,
                 ref_phys: np.ndarray | None = None):
        h, w, _ = rgb.shape
        nb = (h // 4) * (w // 4)
        rgb = np.ascontiguousarray(rgb, np.uint8)
        phys = np.array(phys_fit, dtype=np.uint64)
        ref = None if ref_phys is None else np.ascontiguousarray(ref_phys, np.uint64)
        motion = np.zeros(2 * nb, np.uint8)
        unique = np.zeros(nb, np.uint32)
        n_unique = np.zeros(1, np.uint32)
        eps = np.zeros(6 * nb, np.uint8)
        kind = np.zeros(nb, np.uint8)
        _check(self.L.mptc_reencode(self.h, rgb.ctypes.data, w, h, int(bool(is_intra)),
                                    None if ref is None else ref.ctypes.data, phys.ctypes.data,
                                    motion.ctypes.data, unique.ctypes.data, n_unique.ctypes.data,
                                    eps.ctypes.data, kind.ctypes.data))
        return dict(phys=phys, motion=motion, unique=unique[:int(n_unique[0])].copy(), eps=eps,
                    kind=kind)

    # ---- device-resident step ----
    def encode_frame_device(self, d_rgb, w: int, h: int, is_intra: bool, d_ref, d_phys,
                            d_motion, d_unique, d_n_unique, d_eps=None, d_kind=None,
                            stream: int | None = None):
        _check(self.L.mptc_encode_frame_device(self.h, _ptr(d_rgb), w, h, int(bool(is_intra)),
                                               _ptr(d_ref), _ptr(d_phys), _ptr(d_motion),
                                               _ptr(d_unique), _ptr(d_n_unique), _ptr(d_eps),
                                               _ptr(d_kind), stream))
        if stream is None:
            self.synchronize()

    def encode_batch_device(self, frames: list[dict], stream: int | None = None):
        """frames: dicts with d_rgb, width, height, is_intra, d_ref, d_phys, d_motion,
        d_unique, d_n_unique and optional d_eps, d_kind (device buffers)."""
        arr = (FrameIO * len(frames))()
        for i, f in enumerate(frames):
            arr[i] = FrameIO(_ptr(f["d_rgb"]), f["width"], f["height"], int(bool(f["is_intra"])),
                             _ptr(f.get("d_ref")), _ptr(f["d_phys"]), _ptr(f["d_motion"]),
                             _ptr(f["d_unique"]), _ptr(f["d_n_unique"]), _ptr(f.get("d_eps")),
                             _ptr(f.get("d_kind")))
        _check(self.L.mptc_encode_batch_device(self.h, len(frames), arr, stream))
        if stream is None:
            self.synchronize()

    def frame_io_array(self, frames: list[dict]):
        arr = (FrameIO * len(frames))()
        for i, f in enumerate(frames):
            arr[i] = FrameIO(_ptr(f["d_rgb"]), f["width"], f["height"], int(bool(f["is_intra"])),
                             _ptr(f.get("d_ref")), _ptr(f["d_phys"]), _ptr(f["d_motion"]),
                             _ptr(f["d_unique"]), _ptr(f["d_n_unique"]), _ptr(f.get("d_eps")),
                             _ptr(f.get("d_kind")))
        return arr

    def synchronize(self):
        """Waits for the work queued on the context's own stream."""
        _check(self.L.mptc_synchronize(self.h))

    def get_tuning(self) -> dict:
        t = Tuning()
        _check(self.L.mptc_get_tuning(self.h, C.byref(t)))
        return t.as_dict()

    def set_tuning(self, wg_size=256, tile_w=16, pass_batch=2, use_graph=0, win_chunks=4, prune=1, itile_w=0):
        t = Tuning(wg_size, tile_w, pass_batch, use_graph, win_chunks, prune, itile_w)
        _check(self.L.mptc_set_tuning(self.h, C.byref(t)))

    def autotune(self, frames: list[dict], steps: int = 3):
        arr = self.frame_io_array(frames)
        best, ms = Tuning(), C.c_double()
        _check(self.L.mptc_autotune(self.h, len(frames), arr, steps, C.byref(best), C.byref(ms)))
        return best.as_dict(), ms.value

    def stats(self) -> dict:
        s = FrameStats()
        _check(self.L.mptc_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def profile_enable(self, on: bool = True):
        _check(self.L.mptc_profile_enable(self.h, int(on)))

    def profile_read(self, name: str):
        n, ms = C.c_int64(), C.c_double()
        _check(self.L.mptc_profile_read(self.h, name.encode(), C.byref(n), C.byref(ms)))
        return n.value, ms.value

    # ---- learned proposals ----
    def net_init(self, net: int, seed: int):
        _check(self.L.mptc_net_init(self.h, net, seed))

    def net_params(self, net: int) -> np.ndarray:
        out = np.zeros(NET_PARAMS, np.float32)
        _check(self.L.mptc_net_get_params(self.h, net, out.ctypes.data))
        return out

    def net_set_params(self, net: int, params: np.ndarray):
        p = np.ascontiguousarray(params, np.float32)
        assert p.size == NET_PARAMS
        _check(self.L.mptc_net_set_params(self.h, net, p.ctypes.data))

    def net_forward(self, net: int, d_px, n: int, d_out):
        _check(self.L.mptc_net_forward(self.h, net, _ptr(d_px), n, _ptr(d_out)))

    def net_gradients(self, net: int, d_px, d_target, d_eps, n: int, seed_weight: float):
        g = np.zeros(NET_PARAMS, np.float32)
        loss = C.c_float()
        _check(self.L.mptc_net_gradients(self.h, net, _ptr(d_px), _ptr(d_target), _ptr(d_eps), n,
                                         seed_weight, g.ctypes.data, C.byref(loss)))
        return g, loss.value

    def net_train(self, net: int, d_px, d_target, d_eps, n_pairs: int, batch=4096, steps=1000,
                  lr=1e-3, seed_weight=1.0, seed=1, lr_final=0.0) -> np.ndarray:
        cfg = TrainCfg(batch, steps, lr, seed_weight, seed, 0.0, 0.0, 0.0, lr_final)
        losses = np.zeros(steps, np.float32)
        _check(self.L.mptc_net_train(self.h, net, _ptr(d_px), _ptr(d_target), _ptr(d_eps), n_pairs,
                                     C.byref(cfg), losses.ctypes.data))
        return losses

    def record_pairs(self, on: bool = True):
        _check(self.L.mptc_record_pairs(self.h, int(on)))

    def collect_pairs(self, net: int, d_px, d_target, d_eps_out, d_count, capacity: int):
        """Appends the last (recording, exhaustive) step's pairs for `net`:
        pixels (48 B), target set (8 x uint32), endpoints (6 B) per pair."""
        _check(self.L.mptc_collect_pairs(self.h, net, _ptr(d_px), _ptr(d_target), _ptr(d_eps_out),
                                         _ptr(d_count), capacity))

    def set_search_mode(self, mode: int, top_k: int = 8):
        _check(self.L.mptc_set_search_mode(self.h, mode, top_k))

    def proposals(self, net: int, n_blocks: int, top_k: int) -> np.ndarray:
        out = np.zeros(n_blocks * top_k, np.int16)
        _check(self.L.mptc_get_proposals(self.h, net, out.ctypes.data, n_blocks))
        return out.reshape(n_blocks, top_k)

    def net_set_base(self, net: int, scores: np.ndarray | None):
        """Fixed scores added under the network's logits (None removes them)."""
        if scores is None:
            _check(self.L.mptc_net_set_base(self.h, net, None))
        else:
            a = np.ascontiguousarray(scores, np.float32)
            assert a.shape == (256,)
            _check(self.L.mptc_net_set_base(self.h, net, a.ctypes.data))

    def net_set_similarity(self, net: int, scores: np.ndarray | None):
        """Scores by the number of differing pixel indices (17) added in the
        learned search's ranking (None removes them)."""
        if scores is None:
            _check(self.L.mptc_net_set_similarity(self.h, net, None))
        else:
            a = np.ascontiguousarray(scores, np.float32)
            assert a.shape == (SIM_BINS,)
            _check(self.L.mptc_net_set_similarity(self.h, net, a.ctypes.data))

    def similarity_counts(self, net: int) -> tuple[np.ndarray, np.ndarray]:
        """(tries, accepts) by number of differing indices, gathered by
        collect_pairs since recording was enabled."""
        out = np.zeros(2 * SIM_BINS, np.uint64)
        _check(self.L.mptc_get_similarity_counts(self.h, net, out.ctypes.data))
        return out[:SIM_BINS].copy(), out[SIM_BINS:].copy()

    def set_prior(self, net: int, scores: np.ndarray | None):
        if scores is None:
            _check(self.L.mptc_set_prior(self.h, net, None))
        else:
            p = np.ascontiguousarray(scores, np.float32)
            assert p.size == NET_LOGITS
            _check(self.L.mptc_set_prior(self.h, net, p.ctypes.data))

    def encode_sequence(self, frames: np.ndarray, intra_interval: int):
        """The reference frame schedule (codec/codec.cpp:1379-1508, entropy stage
        excluded) on host buffers: fit, then Reencode against the previous frame."""
        out = []
        prev = None
        set_intra, num = True, 1
        for rgb in frames:
            fit = self.fit(rgb)
            r = self.reencode(rgb, fit, set_intra, None if set_intra else prev)
            r["is_intra"] = set_intra
            out.append(r)
            prev = r["phys"]
            set_intra, num = next_intra(set_intra, num, intra_interval)
        return out


def encode_frame_noalloc(device: int, d_rgb, w: int, h: int, search_area: int, err_threshold: int,
                         is_intra: bool, d_ref, d_phys, d_motion, d_unique, d_n_unique, d_eps=None, d_kind=None,
                         stream: int | None = None):
    """mptc_encode_frame_noalloc: one frame, no context, no allocation;
    returns once the frame is done."""
    _check(load().mptc_encode_frame_noalloc(device, _ptr(d_rgb), w, h, search_area, err_threshold,
                                            int(bool(is_intra)), _ptr(d_ref), _ptr(d_phys), _ptr(d_motion),
                                            _ptr(d_unique), _ptr(d_n_unique), _ptr(d_eps), _ptr(d_kind), stream))


def frame_sse(d_rgb, w: int, h: int, d_phys, d_out, stream: int | None = None) -> None:
    """Queues mptc_frame_sse: the frame's squared error against the decoding
    of its physical blocks, into d_out (device uint64)."""
    _check(load().mptc_frame_sse(_ptr(d_rgb), w, h, _ptr(d_phys), _ptr(d_out), stream))


def psnr_from_sse(sse: int, w: int, h: int) -> float:
    """PSNR in dB of a frame from its squared error (DXTImage::PSNR,
    codec/dxt_image.cpp:363-383: mean over pixels, peak 3 * 255^2)."""
    mse = float(sse) / float(w * h)
    return 10.0 * math.log10(3.0 * 255.0 * 255.0 / mse) if mse > 0 else float("inf")


def next_intra(set_intra: bool, num: int, interval: int):
    """Intra-frame schedule of codec/codec.cpp:1462-1469."""
    if num <= interval:
        set_intra = False
        if num == interval:
            set_intra, num = True, 1
        else:
            num += 1
    return set_intra, num
