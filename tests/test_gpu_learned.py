"""The learned-proposal search modes.  They are not bit-exact to the
reference by design (they search fewer offsets), so every frame they encode
is checked as the reference's own accept test and a decoder see it
(tests/helpers.check_reuse_frame): each reuse points at the right word, passes
the threshold by the oracle's error, and carries the oracle's re-fit.  The
exhaustive mode must stay bit-exact after learned encodes."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from mptc_amd.workloads import SEQUENCES, synth_cfg
from tests.helpers import check_reuse_frame, load_seq, sha

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def nearest_prior(net: int) -> np.ndarray:
    """Scores that propose the nearest offsets first."""
    L = np.arange(256)
    dx = L % 16 - 8
    dy = L // 16 - 8 if net == 0 else L // 16
    return -(dx * dx + dy * dy).astype(np.float32) - 1e-3 * L


class Seq:
    def __init__(self, hip, frames):
        n, h, w, _ = frames.shape
        self.w, self.h, self.n = w, h, n
        self.nb = (w // 4) * (h // 4)
        self.rgb = hip.Buffer.from_array(frames)
        self.phys = [hip.Buffer(8 * self.nb) for _ in range(n)]
        self.motion = [hip.Buffer(2 * self.nb) for _ in range(n)]
        self.unique = hip.Buffer(4 * self.nb)
        self.n_unique = hip.Buffer(4)
        self.eps = [hip.Buffer(6 * self.nb) for _ in range(n)]
        self.frames = frames

    def rgb_at(self, f):
        return self.rgb.offset(f * self.w * self.h * 3)


def encode(enc, lib, seq, intra_interval, check=None, collect=None):
    """Encodes the whole sequence; returns per-frame stats.  check = (O, S, thr)
    verifies every frame; collect = callback(f) after each frame."""
    stats = []
    set_intra, num = True, 1
    for f in range(seq.n):
        enc.encode_frame_device(seq.rgb_at(f), seq.w, seq.h, set_intra, None if set_intra else seq.phys[f - 1],
                                seq.phys[f], seq.motion[f], seq.unique, seq.n_unique, seq.eps[f], None)
        st = enc.stats()
        stats.append(st)
        if check is not None:
            O, S, thr = check
            phys = seq.phys[f].download(np.uint64, seq.nb)
            motion = seq.motion[f].download(np.uint8, 2 * seq.nb)
            nu = int(seq.n_unique.download(np.uint32, 1)[0])
            unique = seq.unique.download(np.uint32, nu)
            eps = seq.eps[f].download(np.uint8, 6 * seq.nb)
            ref = None if set_intra else seq.phys[f - 1].download(np.uint64, seq.nb)
            own = O.fit_frame(seq.frames[f])
            ni, na = check_reuse_frame(O, seq.frames[f], own, phys, motion, unique, eps, set_intra, ref, S, thr,
                                       label=f"frame {f}")
            assert (ni, na) == (st["n_inter"], st["n_intra"])
            assert st["n_unique"] == nu
        if collect is not None:
            collect(f)
        set_intra, num = lib.next_intra(set_intra, num, intra_interval)
    return stats


@pytest.fixture(scope="module")
def env():
    from mptc_amd import hip, lib
    hip.set_device(0)
    return hip, lib


@pytest.mark.parametrize("mode,S,thr,k", [(1, 8, 0, 8), (2, 8, 0, 8), (2, 8, 8, 16), (2, 4, 3, 1), (1, 6, 30, 4),
                                          (2, 11, 0, 8)])
def test_learned_modes_accept_only_exact_passes(env, orc, mode, S, thr, k):
    hip, lib = env
    log(f"learned mode {mode} S={S} thr={thr} k={k}")
    spec = dict(SEQUENCES["c1_s8_t8"])
    frames = lib.synth_frames(synth_cfg(spec), 0, 5)
    seq = Seq(hip, frames)
    enc = lib.Encoder(seq.w, seq.h, S, thr)
    enc.set_prior(0, nearest_prior(0))
    enc.set_prior(1, nearest_prior(1))
    enc.set_search_mode(mode, k)
    stats = encode(enc, lib, seq, 3, check=(orc, S, thr))
    assert sum(s["n_inter"] + s["n_intra"] for s in stats) > 0
    # candidate counts: at most k per block and search
    for s in stats:
        assert s["inter_candidates"] <= k * s["n_blocks"]
        if mode == 2:
            assert s["intra_candidates"] <= k * s["n_blocks"]
    enc.close()


def test_exhaustive_mode_stays_exact_after_learned_encodes(env, orc):
    hip, lib = env
    spec = SEQUENCES["c1_s8_t0"]
    fx = load_seq("c1_s8_t0")
    frames = lib.synth_frames(synth_cfg(spec), 0, spec["frames"])
    seq = Seq(hip, frames)
    enc = lib.Encoder(seq.w, seq.h, spec["S"], spec["thr"])
    enc.set_prior(0, nearest_prior(0))
    enc.set_prior(1, nearest_prior(1))
    enc.set_search_mode(2, 8)
    encode(enc, lib, seq, spec["intra_interval"])
    enc.set_search_mode(0, 8)
    encode(enc, lib, seq, spec["intra_interval"])
    for f in range(seq.n):
        fr = fx["frames"][f]
        assert sha(seq.phys[f].download(np.uint64, seq.nb)) == fr["phys_sha"], f"frame {f}"
        assert sha(seq.motion[f].download(np.uint8, 2 * seq.nb)) == fr["motion_sha"], f"frame {f}"
        assert sha(seq.eps[f].download(np.uint8, 6 * seq.nb)) == fr["eps_sha"], f"frame {f}"
    last = fx["frames"][-1]
    nu = int(seq.n_unique.download(np.uint32, 1)[0])
    assert nu == last["n_unique"]
    assert sha(seq.unique.download(np.uint32, nu)) == last["unique_sha"]
    enc.close()


def test_learned_mode_rejects_missing_networks(env):
    hip, lib = env
    enc = lib.Encoder(64, 64, 8, 0)
    with pytest.raises(lib.MptcError):
        enc.set_search_mode(1, 8)
    enc.net_init(0, 1)
    enc.set_search_mode(1, 8)
    with pytest.raises(lib.MptcError):
        enc.set_search_mode(2, 8)
    with pytest.raises(lib.MptcError):
        enc.set_search_mode(2, 17)
    enc.close()


def _collect(enc, lib, hip, frames, S, thr, cap):
    """Plain-search training pairs of both networks from a sequence."""
    seq = Seq(hip, frames)
    bufs = {}
    for net in (0, 1):
        bufs[net] = dict(px=hip.Buffer(48 * cap), tgt=hip.Buffer(4 * cap), eps=hip.Buffer(6 * cap),
                         count=hip.Buffer.from_array(np.zeros(1, np.uint32)))

    def grab(f):
        for net in (0, 1):
            b = bufs[net]
            enc.collect_pairs(net, seq.rgb_at(f), seq.w, seq.h, seq.motion[f], seq.eps[f], S, b["px"], b["tgt"],
                              b["eps"], b["count"], cap)

    enc.set_search_mode(0, 8)
    encode(enc, lib, seq, 8, collect=grab)
    for net in (0, 1):
        bufs[net]["n"] = min(cap, int(bufs[net]["count"].download(np.uint32, 1)[0]))
    return bufs


def test_trained_network_beats_untrained_and_fixed_proposals(env, orc):
    """Trains both networks on plain-search pairs of one sequence and encodes
    another (other seed): the trained proposals must find more reuses than
    an untrained network's, and every reuse must pass the exact check."""
    hip, lib = env
    S, thr = 8, 0
    w, h = 640, 360
    train_cfg = synth_cfg(dict(width=w, height=h, seed=901, n_rects=24, vel_mode=1, vel=4, cut_permille=10))
    test_cfg = synth_cfg(dict(width=w, height=h, seed=902, n_rects=24, vel_mode=1, vel=4, cut_permille=10))
    enc = lib.Encoder(w, h, S, thr)
    train = lib.synth_frames(train_cfg, 0, 24)
    bufs = _collect(enc, lib, hip, train, S, thr, 1 << 19)
    log(f"pairs inter {bufs[0]['n']} intra {bufs[1]['n']}")
    assert bufs[0]["n"] > 2000 and bufs[1]["n"] > 2000
    test = lib.synth_frames(test_cfg, 0, 9)
    seq = Seq(hip, test)

    def reuse(mode, k=8, check=False):
        enc.set_search_mode(mode, k)
        st = encode(enc, lib, seq, 8, check=(orc, S, thr) if check else None)
        return sum(s["n_inter"] + s["n_intra"] for s in st)

    plain = reuse(0)
    for net in (0, 1):
        enc.net_init(net, 77 + net)
    untrained = reuse(2)
    enc.set_prior(0, nearest_prior(0))
    enc.set_prior(1, nearest_prior(1))
    fixed = reuse(2)
    enc.set_prior(0, None)
    enc.set_prior(1, None)
    for net in (0, 1):
        b = bufs[net]
        losses = enc.net_train(net, b["px"], b["tgt"], b["eps"], b["n"], batch=2048, steps=400, lr=2e-3,
                               seed_weight=1.0, seed=5, lr_final=2e-4)
        log(f"net {net} loss {losses[:3].mean():.3f} -> {losses[-20:].mean():.3f}")
        assert losses[-20:].mean() < losses[:5].mean() - 0.5
    trained = reuse(2, check=True)
    log(f"reused blocks: plain {plain}, untrained {untrained}, nearest-first {fixed}, trained {trained}")
    assert trained > untrained
    assert trained >= 0.5 * plain
    enc.close()
