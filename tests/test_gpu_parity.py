"""GPU parity: the HIP path, called through the C-ABI (libmptc_hip.so), must be
bit-identical to the reference encoder (golden fixtures) and to the CPU oracle."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

from mptc_amd.workloads import SEQUENCES, synth_cfg
from tests.helpers import compare_frame, compare_results, load_seq, load_units

pytestmark = pytest.mark.gpu


def log(msg: str):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def enc():
    from mptc_amd import lib
    e = lib.Encoder(3840, 2160, search_area=32, err_threshold=0)
    yield e
    e.close()


def _frame_from_blocks(px: np.ndarray, bw: int, bh: int) -> np.ndarray:
    """Tiles 48-byte blocks (raster order) into a (4bh, 4bw, 3) RGB frame."""
    img = np.zeros((4 * bh, 4 * bw, 3), np.uint8)
    for k in range(bw * bh):
        bx, by = k % bw, k // bw
        img[4 * by:4 * by + 4, 4 * bx:4 * bx + 4] = px[k].reshape(4, 4, 3)
    return img


def test_fit_matches_reference_vectors(enc):
    u = load_units()
    px = u["fit_px"]
    n = len(px)
    bw = 50
    bh = (n + bw - 1) // bw
    pad = np.zeros((bw * bh - n, 48), np.uint8)
    img = _frame_from_blocks(np.concatenate([px, pad]), bw, bh)
    log(f"fit vectors: {n} blocks as a {4 * bw}x{4 * bh} frame")
    got = enc.fit(img)[:n]
    assert np.array_equal(got, u["fit_out"])


def test_fit_matches_oracle_on_frames(enc, orc):
    for name in ("c1_s8_t0", "c2_t0"):
        spec = SEQUENCES[name]
        rgb = orc.synth_frame(synth_cfg(spec), 3)
        log(f"fit {name} frame 3")
        assert np.array_equal(enc.fit(rgb), orc.fit_frame(rgb))


def _run_sequence(enc, orc, name, n_frames=None, check_eps=True):
    fx = load_seq(name)
    if fx is None:
        pytest.skip(f"fixture seq_{name}.json not generated")
    spec = SEQUENCES[name]
    n = spec["golden_frames"] if n_frames is None else min(n_frames, spec["golden_frames"])
    enc.set_params(spec["S"], spec["thr"])
    frames = orc.synth_frames(synth_cfg(spec), n)
    log(f"sequence {name}: {n} frames {spec['width']}x{spec['height']} S={spec['S']} thr={spec['thr']}")
    res = enc.encode_sequence(frames, spec["intra_interval"])
    for f, (r, want) in enumerate(zip(res, fx["frames"])):
        if not check_eps:
            r = dict(r, eps=None)
        compare_frame(r, want, f"{name} frame {f}")
    return res


@pytest.mark.parametrize("name", ["c1_s8_t0", "c1_s8_t8", "c1_s16_t0"])
def test_config1_sequences(enc, orc, name):
    _run_sequence(enc, orc, name)


def test_config1_full_arrays(enc, orc):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "seq_c1_s8_t0.npz"))
    spec = SEQUENCES["c1_s8_t0"]
    enc.set_params(spec["S"], spec["thr"])
    frames = orc.synth_frames(synth_cfg(spec), spec["frames"])
    res = enc.encode_sequence(frames, spec["intra_interval"])
    off = 0
    for f, r in enumerate(res):
        n = int(z["n_unique"][f])
        compare_results(r, dict(phys=z["phys"][f], motion=z["motion"][f], eps=z["eps"][f],
                                unique=z["unique"][off:off + n]), f"c1_s8_t0 frame {f}")
        off += n


@pytest.mark.parametrize("name", ["c2_t0", "c2_t8"])
def test_config2_1080p(enc, orc, name):
    res = _run_sequence(enc, orc, name)
    # PSNR of the GPU's frames (oracle formula, pinned in test_oracle.py)
    # against the reference's DXTImage::PSNR of its own frames
    import json
    with open(os.path.join(os.path.dirname(__file__), "golden", "psnr_ref.json")) as fh:
        want = json.load(fh)[name]
    spec = SEQUENCES[name]
    frames = orc.synth_frames(synth_cfg(spec), len(want))
    from mptc_amd import hip, lib
    d_sse = hip.Buffer(8)
    for f, wf in enumerate(want):
        assert orc.psnr_logical(frames[f], res[f]["phys"], res[f]["eps"]) == pytest.approx(wf["psnr_enc"], abs=1e-9)
        assert orc.psnr_physical(frames[f], res[f]["phys"]) == pytest.approx(wf["psnr_dec"], abs=1e-9)
        # the device-side measure (mptc_frame_sse) gives the reference's decoder PSNR
        d_rgb, d_phys = hip.Buffer.from_array(frames[f]), hip.Buffer.from_array(res[f]["phys"])
        lib.frame_sse(d_rgb, spec["width"], spec["height"], d_phys, d_sse)
        hip.synchronize()
        sse = int(d_sse.download(np.uint64, 1)[0])
        assert lib.psnr_from_sse(sse, spec["width"], spec["height"]) == pytest.approx(wf["psnr_dec"], abs=1e-9)
        d_rgb.free()
        d_phys.free()
    d_sse.free()


@pytest.mark.parametrize("w,h", [(4, 4), (36, 20), (640, 360)])
def test_frame_sse_matches_oracle_on_random_words(orc, w, h):
    """Arbitrary physical words (both colour modes, every index) against the
    oracle's decoder-side PSNR and a numpy decode of the same words."""
    from mptc_amd import hip, lib
    rng = np.random.default_rng(w * h)
    rgb = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    nb = (w // 4) * (h // 4)
    phys = rng.integers(0, 2 ** 63, nb, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, nb, dtype=np.uint64)
    d_rgb, d_phys, d_sse = hip.Buffer.from_array(rgb), hip.Buffer.from_array(phys), hip.Buffer(8)
    lib.frame_sse(d_rgb, w, h, d_phys, d_sse)
    hip.synchronize()
    sse = int(d_sse.download(np.uint64, 1)[0])
    assert lib.psnr_from_sse(sse, w, h) == pytest.approx(orc.psnr_physical(rgb, phys), abs=1e-9)
    # numpy decode: 565 endpoints, 4-colour (e1 > e2) or 3-colour + black palettes
    e1 = (phys & np.uint64(0xFFFF)).astype(np.int64)
    e2 = ((phys >> np.uint64(16)) & np.uint64(0xFFFF)).astype(np.int64)

    def rgb565(e):
        r, g, b = e >> 11, (e >> 5) & 63, e & 31
        return np.stack([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)], -1)

    c1, c2 = rgb565(e1), rgb565(e2)
    four = (e1 > e2)[:, None]
    pal2 = np.where(four, (2 * c1 + c2) // 3, (c1 + c2) // 2)
    pal3 = np.where(four, (c1 + 2 * c2) // 3, 0)
    pal = np.stack([c1, c2, pal2, pal3], 1)  # nb x 4 x 3
    idx = ((phys >> np.uint64(32))[:, None] >> (2 * np.arange(16, dtype=np.uint64))) & np.uint64(3)
    dec = pal[np.arange(nb)[:, None], idx.astype(np.int64)]  # nb x 16 x 3
    px = rgb.reshape(h // 4, 4, w // 4, 4, 3).transpose(0, 2, 1, 3, 4).reshape(nb, 16, 3).astype(np.int64)
    assert sse == int(((px - dec) ** 2).sum())
    for b in (d_rgb, d_phys, d_sse):
        b.free()


@pytest.mark.parametrize("name", ["c3_t0", "c3_t8", "c3_t32", "c3_t128"])
def test_config3_4k_s32(enc, orc, name):
    _run_sequence(enc, orc, name)


@pytest.mark.parametrize("i", range(12))
def test_config4_streams(enc, orc, i):
    _run_sequence(enc, orc, f"c4_s{i:02d}")


# ---- against the oracle on small and edge-case inputs -------------------------

def _oracle_vs_gpu(enc, orc, frames, S, thr, intra_interval, label):
    enc.set_params(S, thr)
    log(f"{label}: {len(frames)} frames {frames.shape[2]}x{frames.shape[1]} S={S} thr={thr}")
    want = orc.encode_sequence(frames, S, thr, intra_interval)
    got = enc.encode_sequence(frames, intra_interval)
    for f, (a, b) in enumerate(zip(got, want)):
        compare_results(a, b, f"{label} frame {f}")
        assert np.array_equal(a["kind"], b["kind"]), f"{label} frame {f}: kinds differ"


@pytest.mark.parametrize("w,h,S", [(4, 4, 1), (8, 4, 1), (4, 8, 2), (16, 16, 3), (64, 8, 5),
                                   (8, 64, 8), (40, 36, 2), (100, 60, 4)])
def test_small_frames_random(enc, orc, w, h, S):
    rng = np.random.default_rng(w * 1000 + h * 10 + S)
    base = rng.integers(0, 256, (h, w, 3))
    frames = np.stack([np.clip(base + rng.integers(-6, 7, (h, w, 3)), 0, 255) for _ in range(4)]).astype(np.uint8)
    for thr in (0, 5):
        _oracle_vs_gpu(enc, orc, frames, S, thr, 3, f"random {w}x{h}")


def test_negative_and_large_thresholds(enc, orc):
    from oracle import SynthCfg
    frames = orc.synth_frames(SynthCfg(96, 64, 77, 3, 0, 2, 8, 0), 4)
    for thr in (-1, 1, 50, 30000):
        _oracle_vs_gpu(enc, orc, frames, 4, thr, 4, f"threshold {thr}")


def test_pruned_search_thresholds(enc, orc):
    """Thresholds 0..4 with S <= 8 take the pruned search (bound + exact
    evaluation of the survivors); it must match the oracle bit for bit."""
    from oracle import SynthCfg
    frames = orc.synth_frames(SynthCfg(160, 96, 91, 5, 1, 3, 8, 50), 5)
    for S in (1, 3, 8):
        for thr in (0, 2, 4):
            _oracle_vs_gpu(enc, orc, frames, S, thr, 3, f"S {S} threshold {thr}")
            st = enc.stats()
            assert st["exact_inter_candidates"] <= st["inter_candidates"]
            assert st["exact_intra_candidates"] <= st["intra_candidates"]


def test_constant_and_two_colour_frames(enc, orc):
    h, w = 32, 48
    const = np.full((3, h, w, 3), 97, np.uint8)
    _oracle_vs_gpu(enc, orc, const, 3, 0, 2, "constant")
    yy, xx = np.mgrid[0:h, 0:w]
    check = np.where(((xx // 2 + yy // 3) % 2)[..., None] == 0, np.array([10, 200, 30]), np.array([240, 20, 120]))
    frames = np.stack([np.roll(check, s, axis=1) for s in range(3)]).astype(np.uint8)
    _oracle_vs_gpu(enc, orc, frames, 3, 0, 8, "two-colour")


def test_gradient_chains(enc, orc):
    # smooth gradients and flat steps make long chains of intra reuse (each
    # block may copy a word a neighbour itself copied), exercising the
    # fixpoint iterations
    h, w = 64, 128
    yy, xx = np.mgrid[0:h, 0:w]
    frames = []
    for s in range(3):
        img = np.stack([(xx * 2 + s) % 256, (yy * 3 + xx // 4) % 256, (xx + yy + 7 * s) % 256], -1)
        frames.append(img)
    frames = np.array(frames, np.uint8)
    flat = np.clip(frames.astype(int) // 16 * 16 + 3, 0, 255).astype(np.uint8)  # wide flat steps
    for thr in (0, 2, 8, 64):
        _oracle_vs_gpu(enc, orc, frames, 6, thr, 8, f"gradient thr={thr}")
        _oracle_vs_gpu(enc, orc, flat, 5, thr, 8, f"flat steps thr={thr}")


def test_synthetic_medium_sequences(enc, orc):
    from oracle import SynthCfg
    for seed, S, thr in ((5, 8, 0), (6, 8, 8), (7, 3, 1), (8, 12, 4)):
        frames = orc.synth_frames(SynthCfg(160, 96, seed, 6, 1, 4, 8, 100), 5)
        _oracle_vs_gpu(enc, orc, frames, S, thr, 4, f"synth seed={seed}")


def test_device_entry_point_matches_host(enc, orc):
    from mptc_amd import hip
    from mptc_amd.lib import next_intra
    spec = SEQUENCES["c1_s8_t8"]
    enc.set_params(spec["S"], spec["thr"])
    frames = orc.synth_frames(synth_cfg(spec), 4)
    host = enc.encode_sequence(frames, spec["intra_interval"])
    h, w = spec["height"], spec["width"]
    nb = (h // 4) * (w // 4)
    stream = hip.Stream()
    d_frames = hip.Buffer.from_array(frames)
    phys = [hip.Buffer(8 * nb) for _ in range(2)]
    motion, unique, n_unique = hip.Buffer(2 * nb), hip.Buffer(4 * nb), hip.Buffer(4)
    eps, kind = hip.Buffer(6 * nb), hip.Buffer(nb)
    set_intra, num = True, 1
    for f in range(4):
        cur, prev = phys[f % 2], phys[(f + 1) % 2]
        enc.encode_frame_device(d_frames.offset(f * h * w * 3), w, h, set_intra, None if set_intra else prev,
                                cur, motion, unique, n_unique, eps, kind, stream=stream.handle)
        stream.synchronize()
        n = int(n_unique.download(np.uint32, 1)[0])
        got = dict(phys=cur.download(np.uint64, nb), motion=motion.download(np.uint8, 2 * nb),
                   unique=unique.download(np.uint32, n), eps=eps.download(np.uint8, 6 * nb))
        compare_results(got, host[f], f"device frame {f}")
        assert np.array_equal(kind.download(np.uint8, nb), host[f]["kind"])
        st = enc.stats()
        assert st["n_unique"] == n and st["n_inter"] + st["n_intra"] + n == nb
        set_intra, num = next_intra(set_intra, num, spec["intra_interval"])
    stream.destroy()


def test_invalid_arguments_are_rejected(enc):
    from mptc_amd import lib
    with pytest.raises(lib.MptcError):
        enc.set_params(0, 0)
    with pytest.raises(lib.MptcError):
        enc.set_params(64, 0)
    with pytest.raises(lib.MptcError):
        enc.fit(np.zeros((6, 8, 3), np.uint8))  # height not a multiple of 4
    with pytest.raises(lib.MptcError):
        enc.fit(np.zeros((2200, 4000, 3), np.uint8))  # more blocks than the context holds
