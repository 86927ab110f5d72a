"""TEST INFRASTRUCTURE ONLY -- CPU restatement of the learned-proposal network.

Only tests/, __graft_entry__.smoke() and bench.py's CPU baseline use this
module, as the checker of the GPU kernels; no product path imports it.

The proposal network has no counterpart in the reference (it is an addition
beside the reference's exhaustive search), so there is no reference file to
follow.  Its definition is the one include/mptc.h states:

  x  = bf16(pixel * (1/255))                          48 inputs, pixel-major RGB
  h1 = bf16(relu(x  @ bf16(W1) + b1))                 256
  h2 = bf16(relu(h1 @ bf16(W2) + b2))                 256
  y  = h2 @ bf16(W3) + b3                             262: 256 logits, 6 seeds
  loss = mean over rows of [ -log sum_{L in T} softmax(y[:256])_L + seed_w * mean_c (y[256 + c] - eps[c] / 255)^2 ]
  T  = the row's target set (one index = the usual cross-entropy; an empty set
       contributes only the seed term)

with W[k][n] (input k, output n) and FP32 biases.  The backward pass rounds
dY, dH2 and dH1 to bf16 before they enter a product, as the GPU kernels do:

  gW3 = h2^T bf16(dY)            gb3 = colsum(dY)
  dH2 = (bf16(dY) bf16(W3)^T) * [h2 > 0]
  gW2 = h1^T bf16(dH2)           gb2 = colsum(dH2)
  dH1 = (bf16(dH2) bf16(W2)^T) * [h1 > 0]
  gW1 = x^T bf16(dH1)            gb1 = colsum(dH1)

`exact=True` removes every bf16 rounding and computes in float64: that is the
function whose finite differences pin the backward formulas
(tests/test_mlp_oracle.py).  Products accumulate in float64 here; the GPU
accumulates in fp32 in another order, so the two agree to a tolerance, not
bit for bit.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_LOGITS, N_SEEDS = 48, 256, 256, 6
N_OUT = N_LOGITS + N_SEEDS
OFF_W1 = 0
OFF_B1 = OFF_W1 + N_IN * N_HID
OFF_W2 = OFF_B1 + N_HID
OFF_B2 = OFF_W2 + N_HID * N_HID
OFF_W3 = OFF_B2 + N_HID
OFF_B3 = OFF_W3 + N_HID * N_OUT
N_PARAMS = OFF_B3 + N_OUT


def bf16(x) -> np.ndarray:
    """float32 -> nearest bfloat16 (ties to even), returned as float32 values."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    nan = ((u & 0x7F800000) == 0x7F800000) & ((u & 0x7FFFFF) != 0)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF
    r = np.where(nan, ((u >> 16) | 0x40) & 0xFFFF, r)
    return (r.astype(np.uint32) << 16).view(np.float32).reshape(a.shape)


def unpack(p: np.ndarray) -> dict:
    p = np.asarray(p)
    return dict(W1=p[OFF_W1:OFF_B1].reshape(N_IN, N_HID), b1=p[OFF_B1:OFF_W2],
                W2=p[OFF_W2:OFF_B2].reshape(N_HID, N_HID), b2=p[OFF_B2:OFF_W3],
                W3=p[OFF_W3:OFF_B3].reshape(N_HID, N_OUT), b3=p[OFF_B3:N_PARAMS])


def pack(W1, b1, W2, b2, W3, b3) -> np.ndarray:
    return np.concatenate([np.ravel(W1), b1, np.ravel(W2), b2, np.ravel(W3), b3])


def _r(x, exact):
    return np.asarray(x, np.float64) if exact else bf16(np.asarray(x, np.float32)).astype(np.float64)


def forward(params: np.ndarray, px: np.ndarray, exact: bool = False, base=None) -> dict:
    """px: n x 48 uint8 (or, with exact, any real inputs already divided by 255).
    base: optional 256 fixed scores added to the logit columns (not learned)."""
    P = unpack(params)
    if exact:
        X = np.asarray(px, np.float64)
    else:
        X = _r(np.asarray(px, np.float32) * np.float32(1.0 / 255.0), False)
    W1, W2, W3 = (_r(P[k], exact) for k in ("W1", "W2", "W3"))
    b1, b2, b3 = (np.asarray(P[k], np.float64) for k in ("b1", "b2", "b3"))
    H1 = _r(np.maximum(X @ W1 + b1, 0.0), exact)
    H2 = _r(np.maximum(H1 @ W2 + b2, 0.0), exact)
    Y = H2 @ W3 + b3
    if base is not None:
        Y[:, :N_LOGITS] += np.asarray(base, np.float32).astype(np.float64)
    return dict(X=X, H1=H1, H2=H2, Y=Y, W1=W1, W2=W2, W3=W3)


def target_sets(target, n: int) -> np.ndarray:
    """n x 256 booleans: from logit indices (one per row), from n x 256
    booleans, or from n x 8 uint32 words (bit L of word L // 32)."""
    t = np.asarray(target)
    if t.ndim == 1:
        m = np.zeros((n, N_LOGITS), bool)
        m[np.arange(n), t.astype(np.int64)] = True
        return m
    if t.shape[1] == 8:
        bits = (t.astype(np.uint64)[:, np.arange(N_LOGITS) // 32] >> (np.arange(N_LOGITS) % 32).astype(np.uint64)) & 1
        return bits.astype(bool)
    return t.astype(bool)


def loss_grad(params: np.ndarray, px: np.ndarray, target, eps: np.ndarray,
              seed_w: float, exact: bool = False, base=None):
    """Mean loss over the rows and its gradient (flat, parameter layout)."""
    f = forward(params, px, exact, base)
    Y = f["Y"]
    n = Y.shape[0]
    logits = Y[:, :N_LOGITS]
    mx = logits.max(axis=1, keepdims=True)
    lse = np.log(np.exp(logits - mx).sum(axis=1, keepdims=True))
    prob = np.exp(logits - mx - lse)
    T = target_sets(target, n)
    seeds_t = np.asarray(eps, np.float64).reshape(n, N_SEEDS) * (1.0 / 255.0)
    d = Y[:, N_LOGITS:] - seeds_t
    ex = np.exp(logits - mx)
    ss = (ex * T).sum(axis=1, keepdims=True)
    has = ss[:, 0] > 0
    set_term = np.where(has, lse[:, 0] - np.log(np.where(has, ss[:, 0], 1.0)), 0.0)
    loss = float(np.mean(set_term + seed_w * (d * d).mean(axis=1)))
    r = np.divide(ex * T, ss, out=np.zeros_like(ex), where=ss > 0)
    dY = np.zeros_like(Y)
    dY[:, :N_LOGITS] = np.where(has[:, None], prob - r, 0.0)
    dY[:, N_LOGITS:] = 2.0 * seed_w * d / N_SEEDS
    dY /= n
    dYr = _r(dY, exact)
    gW3 = f["H2"].T @ dYr
    gb3 = dY.sum(axis=0)
    dH2 = (dYr @ f["W3"].T) * (f["H2"] > 0)
    dH2r = _r(dH2, exact)
    gW2 = f["H1"].T @ dH2r
    gb2 = dH2.sum(axis=0)
    dH1 = (dH2r @ f["W2"].T) * (f["H1"] > 0)
    dH1r = _r(dH1, exact)
    gW1 = f["X"].T @ dH1r
    gb1 = dH1.sum(axis=0)
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3)


def adam_step(w, g, m, v, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (t = steps taken including this one); returns w, m, v."""
    m = b1 * m + (1.0 - b1) * g
    v = b2 * v + (1.0 - b2) * g * g
    w = w - lr * (m / (1.0 - b1 ** t)) / (np.sqrt(v / (1.0 - b2 ** t)) + eps)
    return w, m, v


def he_params(seed: int) -> np.ndarray:
    """Random parameters of the network's shape (for tests; not the library's init)."""
    rng = np.random.default_rng(seed)
    W1 = rng.normal(0, np.sqrt(2 / N_IN), (N_IN, N_HID))
    W2 = rng.normal(0, np.sqrt(2 / N_HID), (N_HID, N_HID))
    W3 = rng.normal(0, np.sqrt(1 / N_HID), (N_HID, N_OUT))
    b1 = rng.normal(0, 0.05, N_HID)
    b2 = rng.normal(0, 0.05, N_HID)
    b3 = rng.normal(0, 0.05, N_OUT)
    return pack(W1, b1, W2, b2, W3, b3).astype(np.float32)
