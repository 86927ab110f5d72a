"""CPU checks of the network restatement (oracle/mlp.py): bf16 rounding
against torch's, the backward formulas against finite differences of the
loss, Adam against its closed form after one step."""
import numpy as np
import pytest

from oracle import mlp


def test_bf16_rounding_matches_torch(tmp_path):
    """torch's float32 -> bfloat16 conversion (round to nearest even) as the
    independent check.  It runs in a child process: torch loads its own
    HIP/HSA runtime, which must stay out of processes that run GPU tests."""
    import subprocess
    import sys
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.normal(0, 1, 20000).astype(np.float32),
        rng.normal(0, 1e-30, 2000).astype(np.float32),
        (rng.integers(0, 2**31, 5000, dtype=np.int64).astype(np.uint32)).view(np.float32),
        # exact ties: low 16 bits 0x8000, even and odd upper halves
        (np.arange(1000, dtype=np.uint32) << 16 | 0x8000).view(np.float32),
        np.array([0.0, -0.0, np.inf, -np.inf, 3.0e38, -3.0e38, 1e-45], np.float32),
    ])
    x = x[np.isfinite(x) | np.isinf(x)]
    np.save(tmp_path / "x.npy", x)
    code = ("import numpy as np, torch, sys; x = np.load(sys.argv[1]); "
            "np.save(sys.argv[2], torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy())")
    r = subprocess.run([sys.executable, "-c", code, str(tmp_path / "x.npy"), str(tmp_path / "y.npy")],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    ref = np.load(tmp_path / "y.npy")
    ours = mlp.bf16(x)
    assert np.array_equal(ours.view(np.uint32), ref.view(np.uint32))
    nan = mlp.bf16(np.array([np.nan], np.float32))
    assert np.isnan(nan[0])


def _fd_setup(seed=3, n=6, sets=False):
    rng = np.random.default_rng(seed)
    p = mlp.he_params(seed).astype(np.float64)
    x = rng.uniform(0, 1, (n, mlp.N_IN))
    if sets:
        t = rng.random((n, mlp.N_LOGITS)) < 0.05
        t[:, 7] = True
    else:
        t = rng.integers(0, mlp.N_LOGITS, n)
    eps = rng.integers(0, 256, (n, mlp.N_SEEDS))
    return p, x, t, eps


@pytest.mark.parametrize("sets", [False, True])
@pytest.mark.parametrize("block", ["W1", "b1", "W2", "b2", "W3", "b3"])
def test_backward_matches_finite_differences(block, sets):
    p, x, t, eps = _fd_setup(sets=sets)
    sw = 0.7
    _, g = mlp.loss_grad(p, x, t, eps, sw, exact=True)
    offs = {"W1": (mlp.OFF_W1, mlp.OFF_B1), "b1": (mlp.OFF_B1, mlp.OFF_W2),
            "W2": (mlp.OFF_W2, mlp.OFF_B2), "b2": (mlp.OFF_B2, mlp.OFF_W3),
            "W3": (mlp.OFF_W3, mlp.OFF_B3), "b3": (mlp.OFF_B3, mlp.N_PARAMS)}[block]
    rng = np.random.default_rng(11)
    idx = rng.choice(np.arange(*offs), size=min(40, offs[1] - offs[0]), replace=False)
    # include the seed outputs of the last layer
    if block in ("W3", "b3"):
        extra = [offs[0] + mlp.N_LOGITS + 2] if block == "b3" else [offs[0] + 5 * mlp.N_OUT + 258]
        idx = np.concatenate([idx, extra])
    h = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        lp, _ = mlp.loss_grad(pp, x, t, eps, sw, exact=True)
        lm, _ = mlp.loss_grad(pm, x, t, eps, sw, exact=True)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-5 * abs(fd), (block, i, fd, g[i])


def test_gradient_is_not_trivially_zero():
    p, x, t, eps = _fd_setup()
    _, g = mlp.loss_grad(p, x, t, eps, 0.5, exact=True)
    for a, b in [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_W2, mlp.OFF_B2), (mlp.OFF_W3, mlp.OFF_B3)]:
        assert np.abs(g[a:b]).max() > 1e-4


def test_bf16_path_close_to_exact():
    rng = np.random.default_rng(5)
    p = mlp.he_params(5)
    px = rng.integers(0, 256, (64, mlp.N_IN)).astype(np.uint8)
    t = rng.integers(0, 256, 64)
    eps = rng.integers(0, 256, (64, 6))
    l0, g0 = mlp.loss_grad(p, px, t, eps, 1.0, exact=False)
    l1, g1 = mlp.loss_grad(p.astype(np.float64), px.astype(np.float64) / 255.0, t, eps, 1.0, exact=True)
    assert abs(l0 - l1) < 0.05 * abs(l1)
    assert np.linalg.norm(g0 - g1) < 0.05 * np.linalg.norm(g1)


def test_loss_of_uniform_logits_is_log256():
    p = np.zeros(mlp.N_PARAMS)
    px = np.zeros((4, mlp.N_IN), np.uint8)
    eps = np.zeros((4, 6))
    loss, g = mlp.loss_grad(p, px, np.array([0, 5, 255, 17]), eps, 3.0)
    assert abs(loss - np.log(256.0)) < 1e-12


def test_adam_first_step_is_signed_lr():
    rng = np.random.default_rng(2)
    w = rng.normal(size=100)
    g = rng.normal(size=100)
    w1, m, v = mlp.adam_step(w, g, np.zeros(100), np.zeros(100), 1, 0.01, eps=0.0)
    assert np.allclose(w1, w - 0.01 * np.sign(g))
    assert np.allclose(m, 0.1 * g) and np.allclose(v, 0.001 * g * g)


def test_target_set_encodings_agree():
    rng = np.random.default_rng(8)
    n = 5
    sets = rng.random((n, 256)) < 0.1
    sets[:, 3] = True
    words = np.zeros((n, 8), np.uint32)
    for r in range(n):
        for L in np.flatnonzero(sets[r]):
            words[r, L // 32] |= np.uint32(1) << np.uint32(L % 32)
    q1, q2 = mlp.target_sets(sets, n), mlp.target_sets(words, n)
    assert np.array_equal(q1, q2) and np.array_equal(q1, sets)
    # one index is the one-hot set
    idx = rng.integers(0, 256, n)
    one = np.zeros((n, 256), bool)
    one[np.arange(n), idx] = True
    assert np.array_equal(mlp.target_sets(idx, n), mlp.target_sets(one, n))


def test_set_loss_closed_form():
    """Zero weights: uniform softmax, so a set of m offsets costs log(256 / m);
    an empty set costs nothing and gives no logit gradient."""
    p = np.zeros(mlp.N_PARAMS)
    sets = np.zeros((4, 256), bool)
    sets[0, :5] = True
    sets[1, 100] = True
    sets[2, ::2] = True
    loss, g = mlp.loss_grad(p, np.zeros((4, 48), np.uint8), sets, np.zeros((4, 6)), 1.0)
    exp = np.mean([np.log(256 / 5), np.log(256.0), np.log(2.0), 0.0])
    assert abs(loss - exp) < 1e-12
    # W3 logit columns: the gradient flows only through rows with a set
    assert np.abs(g[mlp.OFF_B3:mlp.OFF_B3 + 256]).sum() > 0


def test_set_loss_is_cross_entropy_for_one_offset():
    rng = np.random.default_rng(4)
    p = mlp.he_params(4)
    px = rng.integers(0, 256, (7, 48)).astype(np.uint8)
    t = rng.integers(0, 256, 7)
    Y = mlp.forward(p, px)["Y"][:, :256]
    ce = np.mean(np.log(np.exp(Y).sum(axis=1)) - Y[np.arange(7), t])
    loss, _ = mlp.loss_grad(p, px, t, np.zeros((7, 6)), 0.0)
    assert abs(loss - ce) < 1e-9
