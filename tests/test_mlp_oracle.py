"""The surrogate's reference arithmetic (oracle/mlp_oracle.py) pinned by
mathematics: finite-difference gradients of the exact float64 network, the
BF16 rounding rule on known bit patterns, and Adam's first step."""
import numpy as np

import mlp_oracle as M


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e38, np.inf], dtype=np.float32)
    r = M.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: to even (1.0);
    # 1 + 3 * 2^-9 rounds up to 1 + 2^-7
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == np.float32(1 + 2 ** -7)
    assert r[3] == -2.5 and np.isinf(r[5])
    assert M.NPARAM == 32 * 256 + 256 + 3 * (256 * 256 + 256) + 256 + 1


def test_gradients_finite_differences():
    rng = np.random.default_rng(3)
    p = M.init_params(5).astype(np.float64)
    x = rng.normal(size=(64, M.IN))
    t = rng.normal(size=64)
    g, loss = M.grads(p, x, t, bf=False)
    h = 1e-6
    for name in ["W1", "b1", "W2", "b3", "W4", "W5", "b5"]:
        o, n = M.OFF[name]
        idx = [o] if n == 1 else list(o + rng.choice(n if name != "W1" else M.IN * M.H, 4, replace=False))
        for i in idx:
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            lp = np.mean((M.forward(pp, x, False)[0] - t) ** 2)
            lm = np.mean((M.forward(pm, x, False)[0] - t) ** 2)
            fd = (lp - lm) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (name, i, fd, g[i])


def test_adam_first_step_is_signed_lr():
    g = np.array([0.5, -2.0, 1e-3], dtype=np.float32)
    p = np.zeros(3, dtype=np.float32)
    p1, m, v = M.adam(p, g, np.zeros(3), np.zeros(3), 0, lr=1e-3)
    # bias-corrected first step: m_hat = g, v_hat = g^2 -> -lr sign(g)
    np.testing.assert_allclose(p1, -1e-3 * np.sign(g), rtol=1e-4)


def test_bf16_forward_close_to_exact():
    rng = np.random.default_rng(4)
    p = M.init_params(6)
    x = rng.normal(size=(32, M.IN))
    y1 = M.forward(p, x, True)[0]
    y2 = M.forward(p.astype(np.float64), x, False)[0]
    assert np.max(np.abs(y1 - y2)) < 0.05 * (np.max(np.abs(y2)) + 1)
