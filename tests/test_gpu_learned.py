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
    """Exhaustive-search training pairs of both networks from a sequence."""
    seq = Seq(hip, frames)
    bufs = {}
    for net in (0, 1):
        bufs[net] = dict(px=hip.Buffer(48 * cap), tgt=hip.Buffer(32 * cap), eps=hip.Buffer(6 * cap),
                         count=hip.Buffer.from_array(np.zeros(1, np.uint32)))

    def grab(f):
        for net in (0, 1):
            b = bufs[net]
            enc.collect_pairs(net, b["px"], b["tgt"], b["eps"], b["count"], cap)

    enc.set_search_mode(0, 8)
    enc.record_pairs(True)
    encode(enc, lib, seq, 8, collect=grab)
    enc.record_pairs(False)
    for net in (0, 1):
        bufs[net]["n"] = min(cap, int(bufs[net]["count"].download(np.uint32, 1)[0]))
    return bufs


@pytest.mark.parametrize("S,thr", [(8, 0), (8, 150), (5, 40)])
def test_recorded_sets_match_the_accept_test(env, orc, S, thr):
    """Every offset of a recorded set passes the accept test by the oracle and
    every passing offset is recorded; blocks are told apart by their noise
    pixels.  With thr >= 0 a block has a non-empty inter (intra) set exactly
    when the search reused an inter (intra) word."""
    hip, lib = env
    rng = np.random.default_rng(S * 1000 + thr)
    h, w = 64, 96
    base = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    f1 = np.clip(base.astype(int) + rng.integers(-6, 7, (h, w, 3)), 0, 255).astype(np.uint8)
    # flat patches so that some words repeat and pass at thr = 0
    base[8:24, 8:40] = 120
    f1[8:24, 12:44] = 121
    frames = np.stack([base, f1])
    seq = Seq(hip, frames)
    enc = lib.Encoder(w, h, S, thr)
    enc.record_pairs(True)
    bw, bh = w // 4, h // 4
    nb = bw * bh
    cap = nb
    kinds = hip.Buffer(nb)
    for f in range(2):
        enc.encode_frame_device(seq.rgb_at(f), w, h, f == 0, None if f == 0 else seq.phys[0], seq.phys[f],
                                seq.motion[f], seq.unique, seq.n_unique, seq.eps[f], kinds)
    st = enc.stats()
    kind = kinds.download(np.uint8, nb)
    own = orc.fit_frame(f1)
    ref = seq.phys[0].download(np.uint64, nb)
    out = seq.phys[1].download(np.uint64, nb)
    blocks = {f1[4 * by:4 * by + 4, 4 * bx:4 * bx + 4].tobytes(): (bx, by) for by in range(bh) for bx in range(bw)}
    assert len(blocks) > nb // 2
    for net in (0, 1):
        px, tgt, eps = hip.Buffer(48 * cap), hip.Buffer(32 * cap), hip.Buffer(6 * cap)
        cnt = hip.Buffer.from_array(np.zeros(1, np.uint32))
        enc.collect_pairs(net, px, tgt, eps, cnt, cap)
        n = int(cnt.download(np.uint32, 1)[0])
        assert n == (st["n_inter"] if net == 0 else st["n_intra"]), (net, n, st)
        P = px.download(np.uint8, 48 * n).reshape(n, 48)
        T = tgt.download(np.uint32, 8 * n).reshape(n, 8)
        checked = 0
        for p in range(n):
            key = P[p].reshape(4, 4, 3).tobytes()
            if key not in blocks:
                continue
            bx, by = blocks[key]
            k = by * bw + bx
            if f1[4 * by:4 * by + 4, 4 * bx:4 * bx + 4].reshape(-1).tolist().count(121) > 40:
                continue  # flat patch: pixels do not identify the block
            exp = np.zeros(256, bool)
            for L in range(256):
                dx = L % 16 - 8
                dy = L // 16 - 8 if net == 0 else L // 16
                j, i = (by + dy, bx + dx) if net == 0 else (by - dy, bx + dx)
                if not (-S <= dx < S) or not (0 <= i < bw and 0 <= j < bh):
                    continue
                if net == 0 and not (-S <= dy < S):
                    continue
                if net == 1 and (dy > 2 * S - 1 or (dy == 0 and dx >= 0)):
                    continue
                word = int((ref if net == 0 else out)[j * bw + i]) >> 32
                ok, e, _, _ = orc.eval_candidate(int(own[k]), f1[4 * by:4 * by + 4, 4 * bx:4 * bx + 4], word)
                exp[L] = ok and e <= thr
            got = np.array([(int(T[p, L // 32]) >> (L % 32)) & 1 for L in range(256)], bool)
            assert np.array_equal(got, exp), (net, bx, by, np.flatnonzero(got), np.flatnonzero(exp))
            checked += 1
        log(f"net {net}: {n} pairs, {checked} checked")
        assert n == 0 or checked > 0
    enc.close()


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


def rank_keys(y, L):
    """The proposal ranking: fp32 scores (NaN as -FLT_MAX) as order-preserving
    integers with the low 8 bits replaced by 255 - L (larger first)."""
    y = np.where(np.isnan(y), np.float32(-3.402823466e38), np.asarray(y, np.float32)).astype(np.float32)
    u = y.view(np.uint32).astype(np.int64)
    u = np.where(u & 0x80000000, (~u) & 0xFFFFFFFF, u | 0x80000000)
    return (u & 0xFFFFFF00) | (255 - np.asarray(L, np.int64))


def _admitted(net, S, bx, by, bw, bh):
    """Offsets (logit indices) the search of block (bx, by) admits."""
    L = np.arange(256)
    dx = L % 16 - 8
    lo, hi = -min(S, 8), min(S, 8) - 1
    ok = (dx >= lo) & (dx <= hi) & (bx + dx >= 0) & (bx + dx < bw)
    if net == 0:
        dy = L // 16 - 8
        return ok & (dy >= lo) & (dy <= hi) & (by + dy >= 0) & (by + dy < bh)
    dy = L // 16
    return ok & (dy <= min(2 * S - 1, 15)) & (by - dy >= 0) & ((dy > 0) | (dx < 0))


@pytest.mark.parametrize("k,S", [(8, 8), (3, 5), (16, 8)])
def test_fused_topk_matches_network_logits(env, k, S):
    """The step's proposals are the top-k admitted logits of the network
    (ties to the lower index), as computed from the standalone forward pass."""
    hip, lib = env
    spec = dict(SEQUENCES["c1_s8_t8"])
    frames = lib.synth_frames(synth_cfg(spec), 0, 2)
    seq = Seq(hip, frames)
    enc = lib.Encoder(seq.w, seq.h, S, 0)
    for net in (0, 1):
        enc.net_init(net, 31 + net)
    enc.set_search_mode(2, k)
    kinds = hip.Buffer(seq.nb)
    for f in range(2):
        enc.encode_frame_device(seq.rgb_at(f), seq.w, seq.h, f == 0, None if f == 0 else seq.phys[0], seq.phys[f],
                                seq.motion[f], seq.unique, seq.n_unique, seq.eps[f], kinds)
    kind = kinds.download(np.uint8, seq.nb)
    bw, bh = seq.w // 4, seq.h // 4
    px = frames[1].reshape(bh, 4, bw, 4, 3).transpose(0, 2, 1, 3, 4).reshape(seq.nb, 48).copy()
    d_px, d_y = hip.Buffer.from_array(px), hip.Buffer(seq.nb * 262 * 4)
    for net in (0, 1):
        props = enc.proposals(net, seq.nb, k)
        enc.net_forward(net, d_px, seq.nb, d_y)
        y = d_y.download(np.float32, seq.nb * 262).reshape(seq.nb, 262)[:, :256]
        checked = 0
        for g in range(seq.nb):
            if net == 1 and kind[g] == 2:
                continue  # inter-reused blocks do not search the intra window
            bx, by = g % bw, g // bw
            ok = _admitted(net, S, bx, by, bw, bh)
            cand = np.flatnonzero(ok)
            order = cand[np.argsort(-rank_keys(y[g, cand], cand), kind="stable")][:k]
            exp = np.full(k, -1)
            exp[:len(order)] = order
            assert np.array_equal(props[g], exp), (net, g, props[g], exp)
            checked += 1
        assert checked > 0
    enc.close()


def test_collect_pairs_after_a_step_on_a_caller_stream(env):
    """Acceptance sets recorded by a step queued on the caller's stream are
    collected only after that step's work is done: the pairs equal those of
    the same step run on the context's own stream."""
    hip, lib = env
    spec = dict(SEQUENCES["c1_s8_t8"])
    frames = lib.synth_frames(synth_cfg(spec), 0, 2)
    S, thr = spec["S"], spec["thr"]
    got = []
    for use_caller in (False, True):
        seq = Seq(hip, frames)
        enc = lib.Encoder(seq.w, seq.h, S, thr)
        enc.record_pairs(True)
        stream = hip.Stream() if use_caller else None
        enc.encode_frame_device(seq.rgb_at(0), seq.w, seq.h, True, None, seq.phys[0], seq.motion[0], seq.unique,
                                seq.n_unique, seq.eps[0], None)
        enc.encode_frame_device(seq.rgb_at(1), seq.w, seq.h, False, seq.phys[0], seq.phys[1], seq.motion[1],
                                seq.unique, seq.n_unique, seq.eps[1], None,
                                stream=stream.handle if stream else None)
        out = []
        for net in (0, 1):
            cap = seq.nb
            px, tgt, eps = hip.Buffer(48 * cap), hip.Buffer(32 * cap), hip.Buffer(6 * cap)
            cnt = hip.Buffer.from_array(np.zeros(1, np.uint32))
            enc.collect_pairs(net, px, tgt, eps, cnt, cap)  # no wait on `stream` before this call
            n = int(cnt.download(np.uint32, 1)[0])
            out.append((n, px.download(np.uint8, 48 * n), tgt.download(np.uint32, 8 * n),
                        eps.download(np.uint8, 6 * n)))
        if stream:
            stream.synchronize()
            stream.destroy()
        enc.close()
        got.append(out)
    for net in (0, 1):
        a, b = got[0][net], got[1][net]
        assert a[0] == b[0] and a[0] > 0, (net, a[0], b[0])
        # pairs are appended in an order set by the workgroups: compare as sets of rows
        ra = sorted(bytes(a[1][48 * i:48 * i + 48]) + a[2][8 * i:8 * i + 8].tobytes() + bytes(a[3][6 * i:6 * i + 6])
                    for i in range(a[0]))
        rb = sorted(bytes(b[1][48 * i:48 * i + 48]) + b[2][8 * i:8 * i + 8].tobytes() + bytes(b[3][6 * i:6 * i + 6])
                    for i in range(b[0]))
        assert ra == rb, net


def test_learned_1080p_with_network_inference(env, orc):
    """Config 5 at its size: both networks trained briefly on the exhaustive
    search's pairs of three 1080p frames, then the learned search (mode 2,
    k = 8) encodes the same three frames with network inference over all
    129,600 blocks per frame.  Every reuse is checked by the oracle's accept
    test and as a decoder reads it."""
    hip, lib = env
    spec = SEQUENCES["c2_t0"]
    S, thr = spec["S"], spec["thr"]
    n = 3
    frames = lib.synth_frames(synth_cfg(spec), 0, n)
    enc = lib.Encoder(spec["width"], spec["height"], S, thr)
    log("1080p learned: collecting pairs")
    bufs = _collect(enc, lib, hip, frames, S, thr, 1 << 19)
    log(f"pairs inter {bufs[0]['n']} intra {bufs[1]['n']}")
    assert bufs[0]["n"] > 1000 and bufs[1]["n"] > 1000
    for net in (0, 1):
        enc.net_init(net, 5 + net)
        b = bufs[net]
        enc.net_train(net, b["px"], b["tgt"], b["eps"], b["n"], batch=4096, steps=150, lr=2e-3,
                      seed_weight=1.0, seed=3, lr_final=2e-4)
    enc.set_search_mode(2, 8)
    seq = Seq(hip, frames)
    enc.profile_enable(True)
    log("1080p learned: encoding with checks")
    stats = encode(enc, lib, seq, spec["intra_interval"], check=(orc, S, thr))
    prof = enc.profile_read("net_l3")
    enc.profile_enable(False)
    assert prof[0] >= 2 * n - 1  # inference ran: inter (frames 1, 2) and intra (every frame)
    reused = sum(s["n_inter"] + s["n_intra"] for s in stats)
    log(f"1080p learned: {reused} reused blocks over {n} frames")
    assert reused > 0
    for s in stats:
        assert s["n_blocks"] == 129600
        assert s["inter_candidates"] <= 8 * s["n_blocks"] and s["intra_candidates"] <= 8 * s["n_blocks"]
    enc.close()


def test_base_only_network_proposes_like_the_prior(env, orc):
    """A network whose logit weights and biases are zero, under base scores
    equal to a prior, proposes exactly what that prior proposes (the start of
    training on top of a fixed ranking); the frames it encodes are identical."""
    from oracle import mlp
    hip, lib = env
    spec = dict(SEQUENCES["c1_s8_t8"])
    frames = lib.synth_frames(synth_cfg(spec), 0, 3)
    S, thr, k = spec["S"], spec["thr"], 8
    outs = []
    for use_base in (False, True):
        seq = Seq(hip, frames)
        enc = lib.Encoder(seq.w, seq.h, S, thr)
        for net in (0, 1):
            if use_base:
                P = mlp.unpack(mlp.he_params(40 + net))
                P["W3"][:, :256] = 0.0
                P["b3"][:256] = 0.0
                enc.net_set_params(net, mlp.pack(P["W1"], P["b1"], P["W2"], P["b2"], P["W3"], P["b3"]).astype(np.float32))
                enc.net_set_base(net, nearest_prior(net))
            else:
                enc.set_prior(net, nearest_prior(net))
        enc.set_search_mode(2, k)
        stats = encode(enc, lib, seq, 3, check=(orc, S, thr) if use_base else None)
        props = [enc.proposals(net, seq.nb, k) for net in (0, 1)]
        outs.append(([seq.phys[f].download(np.uint64, seq.nb) for f in range(seq.n)],
                     [seq.eps[f].download(np.uint8, 6 * seq.nb) for f in range(seq.n)],
                     [seq.motion[f].download(np.uint8, 2 * seq.nb) for f in range(seq.n)], props,
                     [(s["n_inter"], s["n_intra"], s["n_unique"]) for s in stats]))
        enc.close()
    a, b = outs
    for f in range(len(frames)):
        assert np.array_equal(a[0][f], b[0][f]) and np.array_equal(a[1][f], b[1][f]), f"frame {f}"
        assert np.array_equal(a[2][f], b[2][f]), f"frame {f}"
    assert np.array_equal(a[3][0], b[3][0]) and np.array_equal(a[3][1], b[3][1])
    assert a[4] == b[4]


@pytest.mark.parametrize("name,thr", [("c1_s8_t8", 8), ("c1_s8_t0", 0)])
def test_learned_intra_tile_shapes_agree(env, orc, name, thr):
    """The learned intra fixpoint settles chains inside 8x8, 16x4, 32x2 or
    64x1 block tiles per pass; the frames do not depend on the tile shape."""
    hip, lib = env
    spec = dict(SEQUENCES[name])
    frames = lib.synth_frames(synth_cfg(spec), 0, 4)
    S = spec["S"]
    outs = {}
    for ltw in (8, 16, 32, 64, 0):
        log(f"learned intra tile {ltw} x {64 // max(ltw, 8)}, {name} thr={thr}")
        seq = Seq(hip, frames)
        enc = lib.Encoder(seq.w, seq.h, S, thr)
        enc.set_prior(0, nearest_prior(0))
        enc.set_prior(1, nearest_prior(1))
        enc.set_search_mode(2, 8)
        enc.set_tuning(itile_w=ltw)
        assert enc.get_tuning()["itile_w"] == ltw
        stats = encode(enc, lib, seq, 3, check=(orc, S, thr) if ltw == 32 else None)
        outs[ltw] = ([seq.phys[f].download(np.uint64, seq.nb) for f in range(seq.n)],
                     [seq.eps[f].download(np.uint8, 6 * seq.nb) for f in range(seq.n)],
                     [seq.motion[f].download(np.uint8, 2 * seq.nb) for f in range(seq.n)],
                     [(s["n_inter"], s["n_intra"], s["n_unique"]) for s in stats])
        enc.close()
    for ltw in (16, 32, 64, 0):
        for f in range(len(frames)):
            for part in range(3):
                assert np.array_equal(outs[8][part][f], outs[ltw][part][f]), (ltw, f, part)
        assert outs[8][3] == outs[ltw][3]
    with pytest.raises(lib.MptcError):
        e = lib.Encoder(64, 64, 8, 0)
        try:
            e.set_tuning(itile_w=12)
        finally:
            e.close()


def _index_diff(a, b):
    """Pixels whose 2-bit index differs between index words a and b."""
    x = (np.asarray(a, np.uint32) ^ np.asarray(b, np.uint32)).astype(np.uint32)
    v = (x | (x >> np.uint32(1))) & np.uint32(0x55555555)
    return np.array([bin(int(t)).count("1") for t in np.ravel(v)]).reshape(np.shape(v))


def _cand_block(net, bx, by, bw, L):
    dx, r = L % 16 - 8, L // 16
    y = by + r - 8 if net == 0 else by - r
    return y * bw + bx + dx


@pytest.mark.parametrize("size", [(256, 256), (200, 120)])
@pytest.mark.parametrize("source", ["network", "prior"])
def test_similarity_scores_enter_the_ranking(env, source, size):
    """With index-similarity scores set, the proposals are the top-k admitted
    offsets of (logit or prior score) + sim[h], h = the pixels whose index
    differs between the block's own word and the candidate's word (inter: the
    previous frame's word; intra: the word the intra search starts from).
    At 200x120 (50 blocks per row) the fused top-k's tiles of 32 blocks span
    two rows of the frame."""
    hip, lib = env
    spec = dict(SEQUENCES["c1_s8_t8"])
    spec["width"], spec["height"] = size
    frames = lib.synth_frames(synth_cfg(spec), 0, 2)
    seq = Seq(hip, frames)
    S, k = 8, 8
    enc = lib.Encoder(seq.w, seq.h, S, 0)
    rng = np.random.default_rng(77)
    sims = [rng.normal(0, 4, 17).astype(np.float32) for _ in (0, 1)]
    for net in (0, 1):
        if source == "network":
            enc.net_init(net, 51 + net)
        else:
            enc.set_prior(net, nearest_prior(net))
        enc.net_set_similarity(net, sims[net])
    enc.set_search_mode(2, k)
    kinds = hip.Buffer(seq.nb)
    for f in range(2):
        enc.encode_frame_device(seq.rgb_at(f), seq.w, seq.h, f == 0, None if f == 0 else seq.phys[0], seq.phys[f],
                                seq.motion[f], seq.unique, seq.n_unique, seq.eps[f], kinds)
    kind = kinds.download(np.uint8, seq.nb)
    bw, bh = seq.w // 4, seq.h // 4
    own = (enc.fit(frames[1]) >> np.uint64(32)).astype(np.uint32)
    prev = (seq.phys[0].download(np.uint64, seq.nb) >> np.uint64(32)).astype(np.uint32)
    cur = (seq.phys[1].download(np.uint64, seq.nb) >> np.uint64(32)).astype(np.uint32)
    start = np.where(kind == 2, cur, own)  # inter-reused blocks start the intra search at their reused word
    px = frames[1].reshape(bh, 4, bw, 4, 3).transpose(0, 2, 1, 3, 4).reshape(seq.nb, 48).copy()
    d_px, d_y = hip.Buffer.from_array(px), hip.Buffer(seq.nb * 262 * 4)
    for net in (0, 1):
        props = enc.proposals(net, seq.nb, k)
        if source == "network":
            enc.net_forward(net, d_px, seq.nb, d_y)
            y = d_y.download(np.float32, seq.nb * 262).reshape(seq.nb, 262)[:, :256]
        else:
            y = np.tile(nearest_prior(net), (seq.nb, 1))
        words = prev if net == 0 else start
        checked = 0
        for g in range(0, seq.nb, 3):
            if net == 1 and kind[g] == 2:
                continue
            bx, by = g % bw, g // bw
            cand = np.flatnonzero(_admitted(net, S, bx, by, bw, bh))
            q = np.array([_cand_block(net, bx, by, bw, L) for L in cand], dtype=np.int64)
            h = _index_diff(np.full(len(q), own[g], np.uint32), words[q]).astype(np.int64)
            score = (y[g, cand] + sims[net][h]).astype(np.float32)
            order = cand[np.argsort(-rank_keys(score, cand), kind="stable")][:k]
            exp = np.full(k, -1)
            exp[:len(order)] = order
            assert np.array_equal(props[g], exp), (source, net, g, props[g], exp)
            checked += 1
        assert checked > 0
    enc.close()


def test_similarity_counts_match_the_oracle(env, orc):
    """The statistics gathered while recording: per number of differing
    indices h, the admitted offsets of every searching block (tries) and
    those passing the accept test, against the oracle's exact evaluation of
    the same candidates (inter: previous frame's words; intra: acceptance
    against the final words, h against the starting words)."""
    from oracle import SynthCfg
    hip, lib = env
    S, thr = 6, 6
    frames = orc.synth_frames(SynthCfg(128, 64, 31, 5, 1, 3, 8, 0), 2)
    seq = Seq(hip, frames)
    enc = lib.Encoder(seq.w, seq.h, S, thr)
    enc.record_pairs(True)
    kinds = hip.Buffer(seq.nb)
    cap = seq.nb
    bufs = [dict(px=hip.Buffer(48 * cap), tgt=hip.Buffer(32 * cap), eps=hip.Buffer(6 * cap),
                 cnt=hip.Buffer.from_array(np.zeros(1, np.uint32))) for _ in (0, 1)]
    kind_f = []
    for f in range(2):
        enc.encode_frame_device(seq.rgb_at(f), seq.w, seq.h, f == 0, None if f == 0 else seq.phys[0], seq.phys[f],
                                seq.motion[f], seq.unique, seq.n_unique, seq.eps[f], kinds)
        kind_f.append(kinds.download(np.uint8, seq.nb))
        for net in (0, 1):
            b = bufs[net]
            enc.collect_pairs(net, b["px"], b["tgt"], b["eps"], b["cnt"], cap)
    bw, bh = seq.w // 4, seq.h // 4
    want = {net: (np.zeros(17, np.int64), np.zeros(17, np.int64)) for net in (0, 1)}
    for f in range(2):
        own = orc.fit_frame(frames[f])
        ow = (own >> np.uint64(32)).astype(np.uint32)
        fin = (seq.phys[f].download(np.uint64, seq.nb) >> np.uint64(32)).astype(np.uint32)
        start = np.where(kind_f[f] == 2, fin, ow)
        prev = None if f == 0 else (seq.phys[0].download(np.uint64, seq.nb) >> np.uint64(32)).astype(np.uint32)
        for g in range(seq.nb):
            bx, by = g % bw, g // bw
            pxg = frames[f][4 * by:4 * by + 4, 4 * bx:4 * bx + 4].reshape(48)
            for net in (0, 1):
                if net == 0 and f == 0:
                    continue
                if net == 1 and kind_f[f][g] == 2:
                    continue
                for L in np.flatnonzero(_admitted(net, S, bx, by, bw, bh)):
                    q = _cand_block(net, bx, by, bw, L)
                    hw = prev[q] if net == 0 else start[q]
                    h = int(_index_diff(ow[g], hw))
                    cw = prev[q] if net == 0 else fin[q]
                    ok, e = orc.eval_candidate(int(own[g]), pxg, int(cw))[:2]
                    want[net][0][h] += 1
                    if ok and e <= thr:
                        want[net][1][h] += 1
    for net in (0, 1):
        tries, acc = enc.similarity_counts(net)
        log(f"net {net}: tries {tries.tolist()} accepts {acc.tolist()}")
        assert np.array_equal(tries.astype(np.int64), want[net][0]), net
        assert np.array_equal(acc.astype(np.int64), want[net][1]), net
        assert acc.sum() > 0
    enc.close()
