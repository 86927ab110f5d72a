"""Batched multi-stream steps, launch-shape tuning and graph replay: every
configuration must give the same bytes as the single-stream path (which the
parity tests pin to the reference)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from mptc_amd.workloads import BATCH_STREAMS, SEQUENCES, synth_cfg
from tests.helpers import compare_frame, compare_results, load_seq

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Stream:
    """Device buffers of one stream: all its frames, two output slots."""

    def __init__(self, hip, frames: np.ndarray):
        n, h, w, _ = frames.shape
        self.w, self.h, self.n = w, h, n
        self.nb = (w // 4) * (h // 4)
        self.rgb = hip.Buffer.from_array(frames)
        self.phys = [hip.Buffer(8 * self.nb) for _ in range(2)]
        self.motion = hip.Buffer(2 * self.nb)
        self.unique = hip.Buffer(4 * self.nb)
        self.n_unique = hip.Buffer(4)
        self.eps = hip.Buffer(6 * self.nb)
        self.kind = hip.Buffer(self.nb)

    def io(self, f: int, is_intra: bool):
        return dict(d_rgb=self.rgb.offset(f * self.w * self.h * 3), width=self.w, height=self.h,
                    is_intra=is_intra, d_ref=None if is_intra else self.phys[(f + 1) % 2],
                    d_phys=self.phys[f % 2], d_motion=self.motion, d_unique=self.unique,
                    d_n_unique=self.n_unique, d_eps=self.eps, d_kind=self.kind)

    def result(self, f: int):
        n = int(self.n_unique.download(np.uint32, 1)[0])
        return dict(phys=self.phys[f % 2].download(np.uint64, self.nb),
                    motion=self.motion.download(np.uint8, 2 * self.nb),
                    unique=self.unique.download(np.uint32, n), eps=self.eps.download(np.uint8, 6 * self.nb),
                    kind=self.kind.download(np.uint8, self.nb))


def _encode_streams(enc, streams, n_frames, intra_interval):
    """Encodes n_frames steps of every stream as batches; returns per-step results."""
    from mptc_amd.lib import next_intra
    out = []
    set_intra, num = True, 1
    for f in range(n_frames):
        enc.encode_batch_device([s.io(f, set_intra) for s in streams])
        out.append([s.result(f) for s in streams])
        set_intra, num = next_intra(set_intra, num, intra_interval)
    return out


def test_batch_of_uneven_streams_matches_reference(orc):
    from mptc_amd import hip, lib
    names = BATCH_STREAMS
    specs = [SEQUENCES[n] for n in names]
    n_frames = min(s["golden_frames"] for s in specs)
    fxs = [load_seq(n) for n in names]
    if any(fx is None for fx in fxs):
        pytest.skip("stream fixtures not generated")
    total = sum((s["width"] // 4) * (s["height"] // 4) for s in specs)
    enc = lib.Encoder(3840, 2160, specs[0]["S"], specs[0]["thr"], max_batch_frames=len(specs),
                      max_batch_blocks=total)
    streams = [Stream(hip, orc.synth_frames(synth_cfg(s), n_frames)) for s in specs]
    log(f"batch: {len(streams)} streams, {total} blocks per step, {n_frames} steps")
    res = _encode_streams(enc, streams, n_frames, specs[0]["intra_interval"])
    for f in range(n_frames):
        for i, name in enumerate(names):
            compare_frame(res[f][i], fxs[i]["frames"][f], f"{name} frame {f}")
    enc.close()


@pytest.mark.parametrize("wg,tw,pb,graph,chunks,prune,itw", [(256, 64, 2, 1, 2, 1, 8), (64, 8, 1, 0, 1, 1, 0),
                                                             (128, 16, 2, 1, 4, 0, 32), (256, 32, 3, 0, 8, 1, 64),
                                                             (64, 64, 16, 1, 6, 0, 16), (256, 16, 1, 1, 3, 1, 32),
                                                             (256, 8, 2, 0, 4, 1, 0), (256, 16, 2, 0, 4, 0, 64)])
def test_tunings_are_bit_identical(orc, wg, tw, pb, graph, chunks, prune, itw):
    from mptc_amd import hip, lib
    from oracle import SynthCfg
    sizes = [(160, 96, 11), (96, 64, 12), (256, 128, 13)]
    frames = [orc.synth_frames(SynthCfg(w, h, sd, 6, 1, 4, 8, 100), 4) for w, h, sd in sizes]
    S, thr = 6, 2
    want = [orc.encode_sequence(fr, S, thr, 3) for fr in frames]
    total = sum((w // 4) * (h // 4) for w, h, _ in sizes)
    enc = lib.Encoder(256, 128, S, thr, max_batch_frames=3, max_batch_blocks=total)
    enc.set_tuning(wg, tw, pb, graph, chunks, prune, itw)
    streams = [Stream(hip, fr) for fr in frames]
    log(f"tuning wg={wg} tile_w={tw} passes={pb} graph={graph} chunks={chunks} prune={prune} itile_w={itw}")
    res = _encode_streams(enc, streams, 4, 3)
    st = enc.stats()
    exact = st["exact_inter_candidates"] + st["exact_intra_candidates"]
    if prune:  # S = 6, threshold 2: the bound rules out candidates
        assert 0 < exact < st["inter_candidates"] + st["intra_candidates"]
    else:
        assert exact == 0
    for f in range(4):
        for i in range(3):
            compare_results(res[f][i], want[i][f], f"stream {i} frame {f}")
            assert np.array_equal(res[f][i]["kind"], want[i][f]["kind"])
    enc.close()


def test_autotune_picks_a_valid_shape(orc):
    from mptc_amd import hip, lib
    spec = SEQUENCES["c1_s8_t8"]
    frames = orc.synth_frames(synth_cfg(spec), 2)
    enc = lib.Encoder(256, 256, spec["S"], spec["thr"])
    st = Stream(hip, frames)
    enc.encode_batch_device([st.io(0, True)])
    best, ms = enc.autotune([st.io(1, False)], steps=2)
    log(f"autotune best {best} {ms:.3f} ms/step")
    assert best["wg_size"] in (64, 128, 256) and best["tile_w"] in (64, 32, 16, 8) and ms > 0
    assert best["itile_w"] in (0, 8, 16, 32, 64)
    assert enc.get_tuning() == best
    # the tuned shape still encodes bit-exactly
    enc.encode_batch_device([st.io(0, True)])
    enc.encode_batch_device([st.io(1, False)])
    want = orc.encode_sequence(frames, spec["S"], spec["thr"], spec["intra_interval"])
    compare_results(st.result(1), want[1], "autotuned frame 1")
    enc.close()


@pytest.mark.parametrize("wg,tw,pb,graph,chunks,prune,itw", [(256, 8, 2, 0, 4, 1, 0), (64, 64, 1, 1, 1, 0, 64),
                                                             (128, 32, 3, 1, 8, 1, 16), (256, 16, 2, 0, 2, 1, 32)])
def test_tunings_at_1080p_match_reference(orc, wg, tw, pb, graph, chunks, prune, itw):
    """Launch shapes at the 1080p size the bench times (the first one is the
    shape the bench's autotuner picks there), against the reference's fixture."""
    from mptc_amd import hip, lib
    name = "c2_t0"
    fx = load_seq(name)
    spec = SEQUENCES[name]
    n = 4
    frames = orc.synth_frames(synth_cfg(spec), n)
    enc = lib.Encoder(spec["width"], spec["height"], spec["S"], spec["thr"])
    enc.set_tuning(wg, tw, pb, graph, chunks, prune, itw)
    st = Stream(hip, frames)
    log(f"1080p tuning wg={wg} tile_w={tw} passes={pb} graph={graph} chunks={chunks} prune={prune} itile_w={itw}")
    res = _encode_streams(enc, [st], n, spec["intra_interval"])
    for f in range(n):
        compare_frame(res[f][0], fx["frames"][f], f"{name} frame {f}")
    enc.close()


def test_calls_on_alternating_streams(orc):
    """One context driven from several streams in turn with no host wait
    between the calls (each frame has its own output buffers): a call on
    another stream waits for the previous call's work, so no step overwrites
    the buffers the previous step's commit still reads, and every frame
    matches the oracle."""
    from mptc_amd import hip, lib
    from mptc_amd.lib import next_intra
    spec = SEQUENCES["c1_s8_t8"]
    n = 6
    frames = orc.synth_frames(synth_cfg(spec), n)
    want = orc.encode_sequence(frames, spec["S"], spec["thr"], 3)
    w, h = spec["width"], spec["height"]
    nb = (w // 4) * (h // 4)
    enc = lib.Encoder(w, h, spec["S"], spec["thr"])
    sa, sb = hip.Stream(), hip.Stream()
    rgb = hip.Buffer.from_array(frames)
    outs = [dict(phys=hip.Buffer(8 * nb), motion=hip.Buffer(2 * nb), unique=hip.Buffer(4 * nb),
                 n_unique=hip.Buffer(4), eps=hip.Buffer(6 * nb)) for _ in range(n)]
    set_intra, num = True, 1
    for f in range(n):
        handle = [sa.handle, sb.handle, None][f % 3]
        o = outs[f]
        io = dict(d_rgb=rgb.offset(f * w * h * 3), width=w, height=h, is_intra=set_intra,
                  d_ref=None if set_intra else outs[f - 1]["phys"], d_phys=o["phys"], d_motion=o["motion"],
                  d_unique=o["unique"], d_n_unique=o["n_unique"], d_eps=o["eps"])
        log(f"alternating streams: frame {f} on {'ab-'[f % 3]}")
        enc.encode_batch_device([io], stream=handle)
        set_intra, num = next_intra(set_intra, num, 3)
    sa.synchronize()
    sb.synchronize()
    enc.synchronize()
    for f in range(n):
        o = outs[f]
        k = int(o["n_unique"].download(np.uint32, 1)[0])
        got = dict(phys=o["phys"].download(np.uint64, nb), motion=o["motion"].download(np.uint8, 2 * nb),
                   unique=o["unique"].download(np.uint32, k), eps=o["eps"].download(np.uint8, 6 * nb))
        compare_results(got, want[f], f"alternating streams frame {f}")
    sa.destroy()
    sb.destroy()
    enc.close()


def test_learned_batch_matches_single_streams(orc):
    """Learned mode with index-similarity scores on both networks: a batch of
    uneven streams gives each stream the same bytes as encoding it alone (a
    block's proposals and intra decisions depend only on its own frame).  In
    the batch the fused top-k's tiles of 32 blocks span two rows of the
    200x120 frame (50 blocks per row) and the boundaries between frames."""
    from mptc_amd import hip, lib
    from oracle import SynthCfg
    sizes = [(200, 120, 21), (96, 64, 22), (256, 128, 23)]
    frames = [orc.synth_frames(SynthCfg(w, h, sd, 6, 1, 4, 8, 100), 3) for w, h, sd in sizes]
    S, thr, k = 8, 2, 4
    rng = np.random.default_rng(5)
    sims = [rng.normal(0, 3, lib.SIM_BINS).astype(np.float32) for _ in (0, 1)]

    def learned(enc):
        for net in (0, 1):
            enc.net_init(net, 61 + net)
            enc.net_set_similarity(net, sims[net])
        enc.set_search_mode(lib.SEARCH_LEARNED, k)

    total = sum((w // 4) * (h // 4) for w, h, _ in sizes)
    enc = lib.Encoder(256, 128, S, thr, max_batch_frames=3, max_batch_blocks=total)
    learned(enc)
    log(f"learned batch: {len(sizes)} streams, {total} blocks per step")
    res = _encode_streams(enc, [Stream(hip, fr) for fr in frames], 3, 2)
    enc.close()
    reused = 0
    for i, fr in enumerate(frames):
        one_enc = lib.Encoder(fr.shape[2], fr.shape[1], S, thr)
        learned(one_enc)
        one = _encode_streams(one_enc, [Stream(hip, fr)], 3, 2)
        one_enc.close()
        for f in range(3):
            compare_results(res[f][i], one[f][0], f"stream {i} frame {f}")
            assert np.array_equal(res[f][i]["kind"], one[f][0]["kind"])
            reused += int((res[f][i]["kind"] != 0).sum())
    assert reused > 0
