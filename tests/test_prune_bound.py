"""The pruning rule of the pruned exhaustive search (mptc_amd/csrc/prune_kernels.h),
restated in numpy with the kernel's fp32 operation order, checked against the
oracle's exact candidate evaluation (codec/dxt_image.cpp:677-698).

Soundness: a candidate the rule prunes never passes (valid and delta <= thr).
If it did, the pruned search would drop a reuse the reference makes.  The
GPU parity tests then compare whole pruned encodes with the oracle."""
from __future__ import annotations

import numpy as np
import pytest

from mptc_amd.workloads import SEQUENCES, synth_cfg

SHIFTS = np.arange(16, dtype=np.uint64) * np.uint64(2)


def bound_w(px48: np.ndarray, words: np.ndarray) -> np.ndarray:
    """W_c = sum |p|^2 - sum_g |S_g|^2 / n_g per candidate word, in fp32 as the
    kernel evaluates it: integer group sums, (float)q_g * (1.0f / n_g) summed
    left to right over g = 0..3 without contraction."""
    px = px48.reshape(16, 3).astype(np.int64)
    idx = ((np.asarray(words, np.uint64)[:, None] >> SHIFTS) & np.uint64(3)).astype(np.int64)  # (N, 16)
    onehot = idx[:, :, None] == np.arange(4)[None, None, :]                                   # (N, 16, 4)
    sums = np.einsum("npg,pc->ngc", onehot.astype(np.int64), px)                               # (N, 4, 3)
    cnt = onehot.sum(1)                                                                        # (N, 4)
    q = (sums * sums).sum(2)                                                                   # exact int
    inv = np.where(cnt > 0, np.float32(1.0) / np.maximum(cnt, 1).astype(np.float32), np.float32(0.0))
    T = np.zeros(len(words), np.float32)
    for g in range(4):
        T = (T + q[:, g].astype(np.float32) * inv[:, g].astype(np.float32)).astype(np.float32)
    Q = np.float32(int((px * px).sum()))
    return (Q - T).astype(np.float32)


def prunes(px48, words, err_own, thr, margin=64):
    return bound_w(px48, words) >= np.float32(48 * (thr + err_own + 1) + margin)


def exact_w(px48, word) -> float:
    px = px48.reshape(16, 3).astype(np.int64)
    idx = (int(word) >> (2 * np.arange(16))) & 3
    w = float((px * px).sum())
    for g in range(4):
        m = idx == g
        if m.any():
            w -= float((px[m].sum(0) ** 2).sum()) / m.sum()
    return w


def _check_block(orc, px48, own, words, thr):
    """(candidates, pruned, passing) for one block; asserts soundness."""
    eo = orc.error(orc.p2l(int(own)), px48)
    pr = prunes(px48, words, eo, thr)
    n_pass = 0
    for c, p in zip(words, pr):
        ok, e, _, _ = orc.eval_candidate(int(own), px48, int(c))
        passing = ok and e <= thr
        n_pass += passing
        assert not (p and passing), f"pruned a passing candidate: word {int(c):#010x} delta {e} thr {thr}"
    return len(words), int(pr.sum()), n_pass


@pytest.mark.parametrize("thr", [0, 2, 4])
def test_bound_never_prunes_a_passing_candidate(orc, thr):
    spec = SEQUENCES["c2_t0"]
    fr = orc.synth_frames(synth_cfg(spec), 2)
    W, H = spec["width"], spec["height"]
    bw, bh, S = W // 4, H // 4, spec["S"]
    ref = orc.fit_frame(fr[0])
    rgb = fr[1]
    own = orc.fit_frame(rgb)
    rng = np.random.default_rng(10 + thr)
    tot = pruned = passing = 0
    for k in rng.choice(bw * bh, 60, replace=False):
        bx, by = int(k % bw), int(k // bw)
        px48 = np.ascontiguousarray(rgb[4 * by:4 * by + 4, 4 * bx:4 * bx + 4]).reshape(48)
        words = [int(ref[j * bw + i]) >> 32 for j in range(max(0, by - S), min(bh, by + S))
                 for i in range(max(0, bx - S), min(bw, bx + S))]
        words.append(int(own[k]) >> 32)  # the own word passes (delta 0): never pruned
        a, b, c = _check_block(orc, px48, own[k], np.array(words, np.uint64), thr)
        tot, pruned, passing = tot + a, pruned + b, passing + c
    # the rule must do work and the sample must contain passing candidates
    assert pruned > tot // 10 and passing > 0, (tot, pruned, passing)


def test_bound_on_adversarial_blocks(orc):
    """Flat, two-colour and saturated blocks, whose candidates sit near the
    threshold, against random and structured words."""
    rng = np.random.default_rng(3)
    blocks = [np.full(48, 128, np.uint8), np.zeros(48, np.uint8), np.full(48, 255, np.uint8)]
    two = np.where(np.arange(16)[:, None] % 2 == 0, [255, 0, 255], [0, 255, 0]).astype(np.uint8).reshape(48)
    blocks += [two, rng.integers(0, 256, 48, dtype=np.uint8), (np.arange(48) * 5 % 256).astype(np.uint8)]
    blocks += [np.clip(100 + rng.integers(-2, 3, 48), 0, 255).astype(np.uint8) for _ in range(4)]
    words = np.concatenate([rng.integers(0, 2**32, 300, dtype=np.uint64),
                            np.array([0, 0xFFFFFFFF, 0x55555555, 0xAAAAAAAA, 0x0000FFFF, 0xFFFF0000], np.uint64)])
    for px48 in blocks:
        own = orc.fit_block(px48)
        for thr in (0, 4):
            _check_block(orc, px48, own, np.append(words, np.uint64(int(own) >> 32)), thr)


def test_fp32_bound_stays_within_margin():
    """The fp32 evaluation of W_c differs from the exact rational value by far
    less than the 64 margin, also for extreme pixel values."""
    rng = np.random.default_rng(7)
    worst = 0.0
    for trial in range(300):
        px48 = rng.choice(np.array([0, 1, 127, 128, 254, 255], np.uint8), 48) if trial % 2 else \
            rng.integers(0, 256, 48, dtype=np.uint8)
        words = rng.integers(0, 2**32, 8, dtype=np.uint64)
        got = bound_w(px48, words)
        for c, v in zip(words, got):
            worst = max(worst, abs(float(v) - exact_w(px48, c)))
    assert worst < 8.0, worst


def _w_variant(px48, word, variant):
    """W_c with a planted mistake (float64; the mistakes dwarf rounding)."""
    px = px48.reshape(16, 3).astype(np.int64)
    idx = (int(word) >> (2 * np.arange(16))) & 3
    if variant == "transposed":  # pixel (x, y) given the index of pixel (y, x)
        idx = idx.reshape(4, 4).T.reshape(16)
    w = float((px * px).sum())
    for g in range(3 if variant == "dropped_group" else 4):
        m = idx == g
        if m.any():
            w -= float((px[m].sum(0) ** 2).sum()) / m.sum()
    return w


@pytest.mark.parametrize("variant", ["dropped_group", "transposed", "limit_without_own_error"])
def test_planted_bound_mistakes_are_caught(orc, variant):
    """The soundness check has teeth: each plausible mistake in the rule
    prunes candidates the oracle accepts."""
    spec = SEQUENCES["c2_t0"]
    fr = orc.synth_frames(synth_cfg(spec), 2)
    bw, bh, thr = spec["width"] // 4, spec["height"] // 4, 0
    ref, rgb = orc.fit_frame(fr[0]), fr[1]
    own = orc.fit_frame(rgb)
    rng = np.random.default_rng(5)
    caught = 0
    for k in rng.choice(bw * bh, 60, replace=False):
        bx, by = int(k % bw), int(k // bw)
        px48 = np.ascontiguousarray(rgb[4 * by:4 * by + 4, 4 * bx:4 * bx + 4]).reshape(48)
        words = [int(own[k]) >> 32] + [int(ref[(by + dy) * bw + bx + dx]) >> 32
                                       for dy in range(-3, 4) for dx in range(-3, 4)
                                       if 0 <= by + dy < bh and 0 <= bx + dx < bw]
        eo = orc.error(orc.p2l(int(own[k])), px48)
        for c in words:
            ok, e, _, _ = orc.eval_candidate(int(own[k]), px48, int(c))
            if not (ok and e <= thr):
                continue
            if variant == "limit_without_own_error":
                caught += float(bound_w(px48, np.array([c], np.uint64))[0]) >= 48 * (thr + 1) + 64
            else:
                caught += _w_variant(px48, c, variant) >= 48 * (thr + eo + 1) + 64
    assert caught > 0, variant


# ---- E and passing-set offsets of the pruned intra search --------------------

K_PRUNE_BATCH = 16  # kPruneBatch (prune_kernels.h)
K_PASS_WORDS = 8    # kPassWords (batch_kernels.h)
K_MAX_CHUNKS = 8    # kMaxChunks (mptc_hip.hip)


def chunks_for(S: int, want: int) -> int:
    """mptc_hip.hip chunks_for: the largest chunk count up to `want` that
    divides the 2S window rows and keeps every chunk a whole number of
    four-position E groups."""
    for k in range(min(want, K_MAX_CHUNKS), 1, -1):
        if (2 * S) % k == 0 and ((2 * S // k) * 2 * S) % 4 == 0:
            return k
    return 1


def pruned_intra_writes(S: int, n_chunks: int):
    """Every (E group, passing-set word) the pruned intra search writes for one
    block, following k_pruned_search<true>'s loops (prune_kernels.h): chunk c
    scans positions [p0, p1) in batches of kPruneBatch; each batch stores
    E4[(pb + k) >> 2] for k = 0, 4, .. < nb and ORs its bits into passing-set
    word pb >> 5 and, when the batch straddles a 32-bit boundary, the next."""
    W2 = 2 * S
    rows = W2 // n_chunks
    groups, words = [], []
    for c in range(n_chunks):
        p0, p1 = c * rows * W2, (c + 1) * rows * W2
        for pb in range(p0, p1, K_PRUNE_BATCH):
            nb = min(K_PRUNE_BATCH, p1 - pb)
            assert nb % 4 == 0, (S, n_chunks, pb, nb)  # the kernel packs four entries per store
            groups += [(pb + k) >> 2 for k in range(0, nb, 4)]
            words.append(pb >> 5)
            sh = pb & 31
            if sh > 16 and pb + nb > ((pb >> 5) + 1) * 32:
                words.append((pb >> 5) + 1)
    return groups, words


@pytest.mark.parametrize("S", range(1, 9))
def test_pruned_search_offsets_stay_in_bounds(S):
    """The E groups and passing-set words the pruned intra search writes,
    for every window-chunk setting the tuning allows, lie inside the buffers
    the context allocates (S^2 groups of four int16 deltas per block,
    kPassWords words per block) and cover every window position exactly once.
    The first, uncommitted version of this kernel faulted the GPU with an E
    store past its group; this pins the index arithmetic on the CPU.  The
    re-evaluation pass writes slot s of group s >> 2 for s < 4S^2, inside the
    same S^2 groups."""
    for want in range(1, K_MAX_CHUNKS + 1):
        K = chunks_for(S, want)
        groups, words = pruned_intra_writes(S, K)
        assert max(groups) < S * S and min(groups) >= 0, (S, K)
        assert sorted(groups) == list(range(S * S)), (S, K)  # each group once
        assert max(words) < K_PASS_WORDS and (4 * S * S - 1) >> 5 < K_PASS_WORDS, (S, K)
    # a planted off-by-one (a batch that writes one group past its end) is caught
    groups, _ = pruned_intra_writes(S, 1)
    bad = groups + [max(groups) + 1]
    assert max(bad) >= S * S
