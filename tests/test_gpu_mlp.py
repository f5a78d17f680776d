"""The surrogate network's GPU kernels (BF16 matrix-core GEMMs, backward
pass, Adam, fused inference) against the reference arithmetic in
oracle/mlp_oracle.py, which mirrors the GPU's BF16 rounding points and
accumulates in float64; the GPU accumulates in FP32 on the matrix cores, so
the bars below are relative to each block's scale."""
import ctypes as C

import numpy as np
import pytest

import mlp_oracle as M
from pyjurassic import lib as L

pytestmark = pytest.mark.gpu

F = C.POINTER(C.c_float)


@pytest.fixture(scope="module")
def glib():
    lib = L.load()
    assert lib.jur_mlp_nparam() == M.NPARAM
    return lib


def fp(a):
    return np.ascontiguousarray(a, dtype=np.float32).ctypes.data_as(F)


def set_params(glib, p):
    p = np.ascontiguousarray(p, dtype=np.float32)
    glib.jur_mlp_set_params(p.ctypes.data_as(F))


def get_params(glib):
    p = np.zeros(M.NPARAM, dtype=np.float32)
    glib.jur_mlp_get_params(p.ctypes.data_as(F))
    return p


@pytest.mark.parametrize("variant", [0, 1])
def test_forward_matches_oracle(glib, variant):
    """Both fused inference kernels (option mlp_infer: 0 a wave per 64
    samples x 128 neurons, 1 per 128 samples x 64 neurons, the default)."""
    glib.jur_mlp_init(11)  # identity normalisation
    p = M.init_params(21)
    set_params(glib, p)
    rng = np.random.default_rng(0)
    glib.jur_set_option(b"mlp_infer", float(variant))
    try:
        for n in [1, 63, 64, 65, 127, 128, 129, 1000]:
            x = rng.normal(size=(n, M.IN)).astype(np.float32)
            y = np.zeros(n, dtype=np.float32)
            glib.jur_mlp_forward(fp(x), n, y.ctypes.data_as(F))
            ref = M.forward(p, x.astype(np.float64), True)[0]
            err = np.abs(y - ref)
            print("variant %d n=%d forward max err %.3e (scale %.3e)" % (variant, n, err.max(), np.abs(ref).max()),
                  flush=True)
            assert err.max() <= 1e-2 * (np.abs(ref).max() + 1e-3)
    finally:
        glib.jur_set_option(b"mlp_infer", 1.0)


def test_gradients_match_oracle(glib):
    glib.jur_mlp_init(11)
    p = M.init_params(22)
    set_params(glib, p)
    rng = np.random.default_rng(1)
    B = 256
    x = rng.normal(size=(B, M.IN)).astype(np.float32)
    t = rng.normal(size=B).astype(np.float32)
    g = np.zeros(M.NPARAM, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    glib.jur_mlp_grads(fp(x), fp(t), B, g.ctypes.data_as(F), loss.ctypes.data_as(F))
    gr, lr = M.grads(p, x.astype(np.float64), t.astype(np.float64), True)
    assert loss[0] == pytest.approx(lr, rel=1e-3)
    for name, (o, n) in M.OFF.items():
        a, b = g[o:o + n], gr[o:o + n]
        rel = np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-30)
        print("%s grad rel err %.2e (norm %.3e)" % (name, rel, np.linalg.norm(b)), flush=True)
        if name == "W1":
            assert np.all(a.reshape(M.INP, M.H)[M.IN:] == 0)
        assert rel < 1e-2, name


def test_adam_step(glib):
    """One training step replayed from the captured graph = Adam applied to
    the gradients of the same batch."""
    glib.jur_mlp_init(11)
    p = M.init_params(23)
    set_params(glib, p)
    rng = np.random.default_rng(2)
    B = 256
    x = rng.normal(size=(B, M.IN)).astype(np.float32)
    t = rng.normal(size=B).astype(np.float32)
    g = np.zeros(M.NPARAM, dtype=np.float32)
    glib.jur_mlp_grads(fp(x), fp(t), B, g.ctypes.data_as(F), None)
    glib.jur_mlp_set_data(fp(x), fp(t), B, 0, 0)
    hist = np.zeros(1, dtype=np.float32)
    glib.jur_mlp_train(1, B, C.c_float(1e-3), hist.ctypes.data_as(F))
    p1 = get_params(glib)
    ref, _, _ = M.adam(p, g, np.zeros(M.NPARAM), np.zeros(M.NPARAM), 0, 1e-3)
    np.testing.assert_allclose(p1, ref, rtol=1e-5, atol=1e-7)


def test_training_reduces_loss_and_save_load(glib, tmp_path):
    rng = np.random.default_rng(5)
    n = 16384
    x = rng.uniform(-2, 2, size=(n, M.IN)).astype(np.float32)
    y = (np.sin(x[:, 0]) + 0.5 * x[:, 1] * x[:, 2] + 0.1 * x[:, 3:].sum(axis=1)).astype(np.float32)
    glib.jur_mlp_init(12)
    glib.jur_mlp_set_data(fp(x), fp(y), n, 1, 3)
    steps = 400
    hist = np.zeros(steps, dtype=np.float32)
    glib.jur_mlp_train(steps, 1024, C.c_float(2e-3), hist.ctypes.data_as(F))
    print("loss first %.4f last %.4f, %.2f ms per step" % (hist[:10].mean(), hist[-10:].mean(),
                                                         glib.jur_stat(b"train_ms") / steps), flush=True)
    assert hist[-10:].mean() < 0.05 * hist[:10].mean()
    yp = np.zeros(n, dtype=np.float32)
    glib.jur_mlp_forward(fp(x), n, yp.ctypes.data_as(F))
    assert np.sqrt(np.mean((yp - y) ** 2)) < 0.25 * np.std(y)
    path = tmp_path / "net.bin"
    glib.jur_mlp_save(str(path).encode())
    glib.jur_mlp_init(99)
    glib.jur_mlp_load(str(path).encode())
    y2 = np.zeros(n, dtype=np.float32)
    glib.jur_mlp_forward(fp(x), n, y2.ctypes.data_as(F))
    np.testing.assert_array_equal(y2, yp)
