"""The proposal network's GPU kernels (inference; the fused training step:
BF16 MFMA forward, loss and backward products, weight gradients, bias sums,
Adam) against the CPU restatement in oracle/mlp.py."""
from __future__ import annotations

import sys

import numpy as np
import pytest

from oracle import mlp

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def env():
    from mptc_amd import hip, lib
    hip.set_device(0)
    enc = lib.Encoder(64, 64, 8, 0)
    yield hip, lib, enc
    enc.close()


def _blocks(sl):
    return {"W1": sl(mlp.OFF_W1, mlp.OFF_B1), "b1": sl(mlp.OFF_B1, mlp.OFF_W2),
            "W2": sl(mlp.OFF_W2, mlp.OFF_B2), "b2": sl(mlp.OFF_B2, mlp.OFF_W3),
            "W3": sl(mlp.OFF_W3, mlp.OFF_B3), "b3": sl(mlp.OFF_B3, mlp.N_PARAMS)}


@pytest.mark.parametrize("n", [1, 77, 1000, 4099])
def test_forward_matches_oracle(env, n):
    hip, lib, enc = env
    log(f"forward n={n}")
    rng = np.random.default_rng(n)
    p = mlp.he_params(7)
    enc.net_set_params(0, p)
    assert np.array_equal(enc.net_params(0), p)
    px = rng.integers(0, 256, (n, 48), dtype=np.uint8)
    d_px = hip.Buffer.from_array(px)
    d_y = hip.Buffer(n * 262 * 4)
    enc.net_forward(0, d_px, n, d_y)
    y = d_y.download(np.float32, n * 262).reshape(n, 262)
    ref = mlp.forward(p, px)["Y"]
    err = np.abs(y - ref)
    # fp32 accumulation in another order can flip a bf16 rounding of a hidden
    # activation (2^-8 relative); everything else agrees to fp32 rounding
    assert err.max() <= 2e-2 * np.abs(ref).max(), err.max()
    assert err.mean() <= 1e-3 * np.abs(ref).mean(), err.mean()


def test_forward_layout_detects_transposes(env):
    """Structured weights: each output depends on one known input, so a
    transposed or shifted operand layout gives visibly wrong values."""
    hip, lib, enc = env
    W1 = np.zeros((48, 256))
    W1[np.arange(48), np.arange(48)] = 1.0          # h1[c] = x[c] for c < 48
    W2 = np.zeros((256, 256))
    W2[np.arange(48), 255 - np.arange(48)] = 2.0    # h2[255 - c] = 2 x[c]
    W3 = np.zeros((256, 262))
    W3[255 - np.arange(48), np.arange(48) * 5] = 1.0  # y[5c] = 2 x[c]
    b3 = np.arange(262) * 0.001
    p = mlp.pack(W1, np.zeros(256), W2, np.zeros(256), W3, b3).astype(np.float32)
    enc.net_set_params(1, p)
    n = 130
    px = np.random.default_rng(1).integers(0, 256, (n, 48), dtype=np.uint8)
    d_px = hip.Buffer.from_array(px)
    d_y = hip.Buffer(n * 262 * 4)
    enc.net_forward(1, d_px, n, d_y)
    y = d_y.download(np.float32, n * 262).reshape(n, 262)
    xb = mlp.bf16(px.astype(np.float32) * np.float32(1 / 255.0))
    exp = np.tile(b3.astype(np.float32), (n, 1))
    exp[:, np.arange(48) * 5] += 2 * xb
    assert np.allclose(y, exp, atol=1e-6), np.abs(y - exp).max()


def set_words(rng, n, density):
    """Random target sets as n x 8 uint32 words; density 0 = one offset per row."""
    if density == 0:
        sets = np.zeros((n, 256), bool)
        sets[np.arange(n), rng.integers(0, 256, n)] = True
    else:
        sets = rng.random((n, 256)) < density
        sets[np.arange(n), rng.integers(0, 256, n)] = True
    return np.packbits(sets.reshape(n, 8, 32)[:, :, ::-1], axis=2, bitorder="big").view(">u4").reshape(n, 8).astype(np.uint32)


def test_set_words_layout():
    rng = np.random.default_rng(0)
    w = set_words(rng, 3, 0.0)
    q = mlp.target_sets(w, 3)
    assert np.array_equal((q > 0).sum(axis=1), [1, 1, 1])


@pytest.mark.parametrize("n,seed_w,density", [(1, 1.0, 0.2), (47, 2.0, 0.1), (64, 1.0, 0.0), (513, 0.25, 0.05),
                                             (3000, 4.0, 0.3)])
def test_gradients_match_oracle(env, n, seed_w, density):
    """Gradients of the training step (batch rows padded to 32: 1 and 47 rows
    leave whole and partial workgroups of padding) against the oracle."""
    hip, lib, enc = env
    log(f"gradients n={n}")
    rng = np.random.default_rng(100 + n)
    p = mlp.he_params(n)
    enc.net_set_params(0, p)
    px = rng.integers(0, 256, (n, 48), dtype=np.uint8)
    tgt = set_words(rng, n, density)
    eps = rng.integers(0, 256, (n, 6), dtype=np.uint8)
    g, loss = enc.net_gradients(0, hip.Buffer.from_array(px), hip.Buffer.from_array(tgt),
                                hip.Buffer.from_array(eps), n, seed_w)
    rl, rg = mlp.loss_grad(p, px, tgt, eps, seed_w)
    assert abs(loss - rl) <= 1e-3 * abs(rl), (loss, rl)
    for name, (a, b) in _blocks(lambda a, b: (a, b)).items():
        d = np.linalg.norm(g[a:b] - rg[a:b])
        assert d <= 1e-2 * np.linalg.norm(rg[a:b]) + 1e-9, (name, d, np.linalg.norm(rg[a:b]))


def test_adam_step_matches_oracle(env):
    hip, lib, enc = env
    n = 256
    rng = np.random.default_rng(9)
    p = mlp.he_params(9)
    px = rng.integers(0, 256, (n, 48), dtype=np.uint8)
    tgt = set_words(rng, n, 0.02)
    eps = rng.integers(0, 256, (n, 6), dtype=np.uint8)
    d = [hip.Buffer.from_array(a) for a in (px, tgt, eps)]
    enc.net_set_params(0, p)
    losses = enc.net_train(0, *d, n, batch=n, steps=1, lr=1e-3, seed_weight=1.0, seed=3)
    p1 = enc.net_params(0)
    rl, rg = mlp.loss_grad(p, px, tgt, eps, 1.0)
    assert abs(losses[0] - rl) <= 1e-3 * abs(rl)
    w1, _, _ = mlp.adam_step(p.astype(np.float64), rg, 0.0, 0.0, 1, 1e-3)
    # the first step moves each weight by lr * g / (|g| + eps): elements whose
    # gradient is tiny may take another sign on the GPU, all others agree
    big = np.abs(rg) > 1e-6
    assert np.abs(p1 - w1)[big].max() <= 2e-5
    assert np.abs(p1 - p).max() <= 1e-3 * (1 + 1e-4)  # |step| <= lr, up to fp32 rounding of w


def test_training_fits_a_learnable_rule(env):
    """Targets that are a function of the pixels (the first pixel's red value,
    in 16 bands, picks the offset): the loss must fall well below log(256)."""
    hip, lib, enc = env
    n = 20000
    rng = np.random.default_rng(4)
    px = rng.integers(0, 256, (n, 48), dtype=np.uint8)
    idx = (px[:, 0] // 16).astype(np.int64) * 16 + 7
    sets = np.zeros((n, 256), bool)
    sets[np.arange(n), idx] = True
    tgt = np.packbits(sets.reshape(n, 8, 32)[:, :, ::-1], axis=2, bitorder="big").view(">u4").reshape(n, 8).astype(np.uint32)
    eps = px[:, :6].copy()
    d = [hip.Buffer.from_array(a) for a in (px, tgt, eps)]
    enc.net_init(1, 5)
    losses = enc.net_train(1, *d, n, batch=1024, steps=600, lr=2e-3, seed_weight=1.0, seed=1)
    log(f"loss {losses[:3]} ... {losses[-3:]}")
    assert np.all(np.isfinite(losses))
    assert losses[:5].mean() > 4.0
    assert losses[-20:].mean() < 1.0


def test_base_scores_add_under_the_logits(env):
    """Fixed base scores: the forward pass and the loss/gradients equal the
    oracle's with the same base; removing the base restores the plain net."""
    hip, lib, enc = env
    n = 700
    rng = np.random.default_rng(21)
    p = mlp.he_params(21)
    base = rng.normal(0, 3, 256).astype(np.float32)
    enc.net_set_params(0, p)
    enc.net_set_base(0, base)
    px = rng.integers(0, 256, (n, 48), dtype=np.uint8)
    d_px = hip.Buffer.from_array(px)
    d_y = hip.Buffer(n * 262 * 4)
    enc.net_forward(0, d_px, n, d_y)
    y = d_y.download(np.float32, n * 262).reshape(n, 262)
    ref = mlp.forward(p, px, base=base)["Y"]
    err = np.abs(y - ref)
    assert err.max() <= 2e-2 * np.abs(ref).max(), err.max()
    assert err.mean() <= 1e-3 * np.abs(ref).mean(), err.mean()
    tgt = set_words(rng, n, 0.05)
    eps = rng.integers(0, 256, (n, 6), dtype=np.uint8)
    g, loss = enc.net_gradients(0, d_px, hip.Buffer.from_array(tgt), hip.Buffer.from_array(eps), n, 1.0)
    rl, rg = mlp.loss_grad(p, px, tgt, eps, 1.0, base=base)
    assert abs(loss - rl) <= 1e-3 * abs(rl), (loss, rl)
    # a plain net would give a different loss: the base is really in it
    assert abs(mlp.loss_grad(p, px, tgt, eps, 1.0)[0] - rl) > 1e-2 * abs(rl)
    for name, (a, b) in _blocks(lambda a, b: (a, b)).items():
        d = np.linalg.norm(g[a:b] - rg[a:b])
        assert d <= 1e-2 * np.linalg.norm(rg[a:b]) + 1e-9, (name, d, np.linalg.norm(rg[a:b]))
    enc.net_set_base(0, None)
    enc.net_forward(0, d_px, n, d_y)
    y0 = d_y.download(np.float32, n * 262).reshape(n, 262)
    ref0 = mlp.forward(p, px)["Y"]
    assert np.abs(y0 - ref0).max() <= 2e-2 * np.abs(ref0).max()
    assert np.allclose(y[:, :256] - y0[:, :256], base[None, :], atol=1e-3 * (1 + np.abs(base).max()))
    assert np.array_equal(y[:, 256:], y0[:, 256:])  # seed outputs carry no base
