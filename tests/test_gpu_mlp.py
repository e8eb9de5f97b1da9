"""SGS model training kernels (FP32 MLP forward, backward, Adam) and the model-space sample builder on the GPU,
against the FP64 oracle restatement (itself pinned by finite differences and torch in test_oracle_mlp.py).

Tolerances are those of FP32 arithmetic: loss and gradient to 5e-5 relative (2-norm), predictions to 1e-5, Adam
parameters to 2e-6 absolute after three steps; the FP64 sample builder to 1e-11 (inputs to FP32 rounding).
"""
import sys

import numpy as np
import pytest

import pyoracle as O
from pyhonee import problems as PR
from pyhonee import train as TR
from pyhonee.driver import context_from_dict

pytestmark = pytest.mark.gpu
ALPHA = 0.1


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _f64(a):
    return np.asarray(a, dtype=np.float32).astype(np.float64)


@pytest.mark.parametrize("n", [1000, 3001])
def test_loss_grad_matches_oracle(n):
    _say(f"mlp loss/grad n={n}")
    rng = np.random.default_rng(n)
    mlp = TR.Mlp(alpha=ALPHA, seed=3)
    X, T = rng.normal(size=(n, 6)).astype(np.float32), rng.normal(size=(n, 6)).astype(np.float32)
    loss, g = mlp.loss_grad(X, T)
    loss_ref, g_ref = O.mlp_loss_grad(_f64(mlp.params), _f64(X), _f64(T), float(np.float32(ALPHA)))
    assert abs(loss - loss_ref) < 1e-5 * loss_ref
    assert np.linalg.norm(g - g_ref) < 5e-5 * np.linalg.norm(g_ref), np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    mlp.close()


def test_predict_matches_oracle():
    _say("mlp predict")
    rng = np.random.default_rng(4)
    mlp = TR.Mlp(alpha=ALPHA, seed=5)
    X = rng.normal(size=(777, 6)).astype(np.float32)
    Y = mlp.predict(X)
    Y_ref = O.mlp_forward(_f64(mlp.params), _f64(X), float(np.float32(ALPHA)))
    assert np.abs(Y - Y_ref).max() < 1e-5 * max(1.0, np.abs(Y_ref).max())
    mlp.close()


def test_adam_steps_match_oracle():
    _say("mlp adam")
    rng = np.random.default_rng(6)
    mlp = TR.Mlp(alpha=ALPHA, seed=7)
    X, T = rng.normal(size=(2048, 6)).astype(np.float32), rng.normal(size=(2048, 6)).astype(np.float32)
    opts = TR.options(lr=1e-3)
    th = _f64(mlp.params)
    m, v = np.zeros_like(th), np.zeros_like(th)
    for t in range(1, 4):
        mlp.adam_step(X, T, opts)
        _, g = O.mlp_loss_grad(th, _f64(X), _f64(T), float(np.float32(ALPHA)))
        th, m, v = O.adam_step(th, g, m, v, t, 1e-3)
    assert np.abs(mlp.params - th).max() < 2e-6, np.abs(mlp.params - th).max()
    mlp.close()


def test_training_fits_a_teacher_and_is_deterministic():
    _say("mlp training")
    rng = np.random.default_rng(8)
    teacher = TR.Mlp(alpha=ALPHA, seed=100)
    X = rng.uniform(-1, 1, size=(20000, 6)).astype(np.float32)
    T = teacher.predict(X)
    opts = TR.options(lr=3e-3, batch=1024, epochs=40, seed=5)
    params = []
    for _ in range(2):
        student = TR.Mlp(alpha=ALPHA, seed=1)
        hist = student.train(X, T, opts)
        params.append(student.params)
        student.close()
    assert np.all(np.isfinite(hist)) and hist[-1] < 0.05 * hist[0], (hist[0], hist[-1])
    assert np.array_equal(params[0], params[1])
    teacher.close()


def test_model_samples_match_oracle():
    _say("sgs model samples")
    ctx = context_from_dict(PR.tgv_context_dict(state_var=1))
    gas = O.newt_from_dict(PR.tgv_context_dict(state_var=1)).gas
    rng = np.random.default_rng(9)
    n = 300
    filtered = np.zeros((n, 11))
    filtered[:, 0] = 71.0 * (1 + 0.01 * rng.uniform(-1, 1, n))
    filtered[:, 1:4] = rng.normal(size=(n, 3))
    filtered[:, 4] = 0.25
    filtered[:, 5:] = rng.normal(size=(n, 6))
    G = rng.normal(size=(n, 9))
    h = rng.uniform(0.1, 1.0, size=(n, 3))
    A = np.concatenate([np.concatenate([hh / np.linalg.norm(hh), [0, 0, 0], [np.linalg.norm(hh)]])[None] for hh in h])
    train = np.zeros((n, 12))
    train[:, 6:] = rng.normal(size=(n, 6)) * 0.1
    X, S = TR.model_samples(train, filtered, G, A, ctx.gas)
    X_ref, S_ref = O.sgs_model_samples(gas, train, filtered, G, A)
    assert np.abs(X - X_ref).max() < 1e-6 * np.abs(X_ref).max()
    assert np.linalg.norm(S - S_ref) < 1e-11 * np.linalg.norm(S_ref)


# ---- the 6-20-20-6 network (a second hidden layer) ----
DIMS2 = (6, 20, 20, 6)


@pytest.mark.parametrize("n", [1000, 3001])
def test_two_hidden_layers_loss_grad_matches_oracle(n):
    _say(f"mlp 6-20-20-6 loss/grad n={n}")
    rng = np.random.default_rng(10 + n)
    mlp = TR.Mlp(alpha=ALPHA, seed=13, dims=DIMS2)
    assert mlp.nparams == 20 * 6 + 20 + 20 * 20 + 20 + 6 * 20 + 6
    X, T = rng.normal(size=(n, 6)).astype(np.float32), rng.normal(size=(n, 6)).astype(np.float32)
    loss, g = mlp.loss_grad(X, T)
    loss_ref, g_ref = O.mlp_loss_grad(_f64(mlp.params), _f64(X), _f64(T), float(np.float32(ALPHA)), DIMS2)
    assert abs(loss - loss_ref) < 1e-5 * loss_ref
    assert np.linalg.norm(g - g_ref) < 5e-5 * np.linalg.norm(g_ref), np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref)
    Y = mlp.predict(X[:500])
    Y_ref = O.mlp_forward(_f64(mlp.params), _f64(X[:500]), float(np.float32(ALPHA)), DIMS2)
    assert np.abs(Y - Y_ref).max() < 1e-5 * max(1.0, np.abs(Y_ref).max())
    mlp.close()


def test_two_hidden_layers_adam_and_training():
    _say("mlp 6-20-20-6 adam and training")
    rng = np.random.default_rng(16)
    mlp = TR.Mlp(alpha=ALPHA, seed=17, dims=DIMS2)
    X, T = rng.normal(size=(2048, 6)).astype(np.float32), rng.normal(size=(2048, 6)).astype(np.float32)
    th = _f64(mlp.params)
    m, v = np.zeros_like(th), np.zeros_like(th)
    for t in range(1, 4):
        mlp.adam_step(X, T, TR.options(lr=1e-3))
        _, g = O.mlp_loss_grad(th, _f64(X), _f64(T), float(np.float32(ALPHA)), DIMS2)
        th, m, v = O.adam_step(th, g, m, v, t, 1e-3)
    assert np.abs(mlp.params - th).max() < 2e-6, np.abs(mlp.params - th).max()
    mlp.close()
    teacher = TR.Mlp(alpha=ALPHA, seed=101, dims=DIMS2)
    Xt = rng.uniform(-1, 1, size=(20000, 6)).astype(np.float32)
    Tt = teacher.predict(Xt)
    student = TR.Mlp(alpha=ALPHA, seed=2, dims=DIMS2)
    hist = student.train(Xt, Tt, TR.options(lr=3e-3, batch=1024, epochs=40, seed=5))
    assert np.all(np.isfinite(hist)) and hist[-1] < 0.05 * hist[0], (hist[0], hist[-1])
    student.close()
    teacher.close()
