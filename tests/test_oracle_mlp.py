"""The oracle's MLP training restatement (FP64) against independent checks (no GPU).

* Back-propagation gradient against central finite differences of the loss (relative 1e-7) and against torch
  autograd of an independently built torch.nn network with the same parameters (1e-12).
* Adam steps against torch.optim.Adam (three steps, 1e-12).
* Model-space samples: rotating the eigenframe stress back, E^T S E delta^2 |grad u|^2, recovers the filtered SGS
  stress exactly as the closure's output stage does, so training targets and inference agree.
"""
import numpy as np
import pytest

import pyoracle as O
from pyhonee import problems as PR


def _data(seed, n=64):
    rng = np.random.default_rng(seed)
    theta = rng.normal(size=20 * 6 + 20 + 6 * 20 + 6) * 0.5
    return theta, rng.normal(size=(n, 6)), rng.normal(size=(n, 6))


def test_gradient_finite_differences():
    theta, X, T = _data(1)
    alpha = 0.1
    _, g = O.mlp_loss_grad(theta, X, T, alpha)
    h = 1e-6
    gfd = np.zeros_like(theta)
    for k in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        gfd[k] = (O.mlp_loss_grad(tp, X, T, alpha)[0] - O.mlp_loss_grad(tm, X, T, alpha)[0]) / (2 * h)
    assert np.linalg.norm(g - gfd) < 1e-7 * np.linalg.norm(g)


DIMS2 = (6, 20, 20, 6)


def _data2(seed, n=64):
    rng = np.random.default_rng(seed)
    theta = rng.normal(size=20 * 6 + 20 + 20 * 20 + 20 + 6 * 20 + 6) * 0.4
    return theta, rng.normal(size=(n, 6)), rng.normal(size=(n, 6))


def _torch_net(theta, alpha, dims=O.MLP_DIMS):
    """An independently built torch.nn network with the oracle's parameters."""
    torch = pytest.importorskip("torch")
    mods = []
    for n_in, n_out in zip(dims[:-1], dims[1:]):
        mods += [torch.nn.Linear(n_in, n_out), torch.nn.LeakyReLU(alpha)]
    net = torch.nn.Sequential(*mods[:-1]).double()
    with torch.no_grad():
        for lin, (W, b) in zip([m for m in net if isinstance(m, torch.nn.Linear)], O.mlp_unpack(theta, dims)):
            lin.weight.copy_(torch.tensor(W))
            lin.bias.copy_(torch.tensor(b))
    return torch, net


def _torch_params(torch, net):
    return [p for m in net if isinstance(m, torch.nn.Linear) for p in (m.weight, m.bias)]


def _torch_grad(torch, net):
    return np.concatenate([p.grad.detach().numpy().ravel() for p in _torch_params(torch, net)])


@pytest.mark.parametrize("dims", [O.MLP_DIMS, DIMS2])
def test_gradient_matches_torch_autograd(dims):
    theta, X, T = _data(2) if len(dims) == 3 else _data2(2)
    alpha = 0.2
    loss, g = O.mlp_loss_grad(theta, X, T, alpha, dims)
    torch, net = _torch_net(theta, alpha, dims)
    lt = torch.nn.functional.mse_loss(net(torch.tensor(X)), torch.tensor(T))
    lt.backward()
    assert abs(lt.item() - loss) < 1e-12 * loss
    assert np.allclose(_torch_grad(torch, net), g, rtol=1e-12, atol=1e-14)
    assert np.allclose(O.mlp_forward(theta, X, alpha, dims), net(torch.tensor(X)).detach().numpy(), rtol=1e-13, atol=1e-13)


def test_gradient_finite_differences_two_hidden_layers():
    theta, X, T = _data2(5, n=32)
    alpha = 0.1
    _, g = O.mlp_loss_grad(theta, X, T, alpha, DIMS2)
    h = 1e-6
    gfd = np.zeros_like(theta)
    for k in range(theta.size):
        tp, tm = theta.copy(), theta.copy()
        tp[k] += h
        tm[k] -= h
        gfd[k] = (O.mlp_loss_grad(tp, X, T, alpha, DIMS2)[0] - O.mlp_loss_grad(tm, X, T, alpha, DIMS2)[0]) / (2 * h)
    assert np.linalg.norm(g - gfd) < 1e-7 * np.linalg.norm(g)


def test_closure_inference_with_two_hidden_layers_matches_the_network():
    """The C inference (o_dd_inference with the second hidden layer) equals mlp_forward of the same parameters: with
    bounds (0, 1) the de-normalization is the identity, and with an identity eigenframe the stress is the network output
    scaled by delta^2 |grad u|^2."""
    theta, _, _ = _data2(9)
    alpha = 0.05
    (W1, b1), (Wh, bh), (W2, b2) = O.mlp_unpack(theta, DIMS2)
    bounds = np.tile([0.0, 1.0], (6, 1))
    m = O.SgsModel({"W1": W1, "b1": b1, "Wh": Wh, "bh": bh, "W2": W2, "b2": b2, "bounds": bounds, "alpha": alpha})
    G = np.diag([0.3, -0.1, -0.2])  # diagonal strain, decreasing: identity eigenframe, no vorticity
    kmA = np.array([1.0, 1.0, 1.0, 0.0, 0.0, 0.0])
    r = O.sgs_point(m, G, kmA, 0.5, 1e-3)
    y = O.mlp_forward(theta, r["inputs"][None, :], alpha, DIMS2)[0]
    s = y * 0.5 ** 2 * r["gmag"] ** 2
    km = np.array([s[0], s[1], s[2], np.sqrt(2) * s[5], np.sqrt(2) * s[4], np.sqrt(2) * s[3]])
    assert np.allclose(r["km_sgs"], km, rtol=1e-13, atol=1e-15)


def test_adam_matches_torch():
    theta, X, T = _data(3)
    alpha, lr = 0.05, 3e-3
    torch, net = _torch_net(theta, alpha)
    opt = torch.optim.Adam(net.parameters(), lr=lr, betas=(0.9, 0.999), eps=1e-8)
    m, v, th = np.zeros_like(theta), np.zeros_like(theta), theta.copy()
    for t in range(1, 4):
        _, g = O.mlp_loss_grad(th, X, T, alpha)
        th, m, v = O.adam_step(th, g, m, v, t, lr)
        opt.zero_grad()
        torch.nn.functional.mse_loss(net(torch.tensor(X)), torch.tensor(T)).backward()
        opt.step()
    tt = np.concatenate([p.detach().numpy().ravel() for p in _torch_params(torch, net)])
    assert np.allclose(th, tt, rtol=1e-12, atol=1e-13)


def test_model_samples_round_trip():
    gas = O.newt_from_dict(PR.tgv_context_dict(state_var=1)).gas
    rng = np.random.default_rng(7)
    n = 8
    filtered = np.zeros((n, 11))
    filtered[:, 0] = 71.0
    filtered[:, 1:4] = rng.normal(size=(n, 3))
    filtered[:, 4] = 0.25
    G = rng.normal(size=(n, 9))
    A = np.tile(np.concatenate([[1, 1, 1, 0, 0, 0] / np.sqrt(3.0), [0.2]]), (n, 1))
    train = np.zeros((n, 12))
    train[:, 6:] = rng.normal(size=(n, 6))
    X, S = O.sgs_model_samples(gas, train, filtered, G, A)
    zero = O.SgsModel({"b1": np.zeros(20), "b2": np.zeros(6), "W1": np.zeros((20, 6)), "W2": np.zeros((6, 20)), "bounds": np.zeros((6, 2)),
                       "alpha": 0.0})
    for i in range(n):
        rho = filtered[i, 0] / ((gas.cp - gas.cv) * filtered[i, 4])
        r = O.sgs_point(zero, G[i].reshape(3, 3), A[i, :6], A[i, 6], gas.mu / rho)
        assert np.array_equal(X[i], r["inputs"])
        s = S[i]
        Sm = np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]]) * A[i, 6] ** 2 * r["gmag"] ** 2
        tau = r["evec"].T @ Sm @ r["evec"]
        t = train[i, 6:]
        assert np.allclose([tau[0, 0], tau[1, 1], tau[2, 2], tau[1, 2], tau[0, 2], tau[0, 1]], t, rtol=1e-12, atol=1e-13)
