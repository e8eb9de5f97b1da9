"""The oracle's data-driven SGS closure against what the mathematics fixes (no GPU).

* Jacobi eigen solver (utils_eigensolver_jacobi.h): eigenvalues against numpy.linalg.eigvalsh, decreasing order,
  orthonormal eigenvector rows that reconstruct the matrix, convergence inside the sweep budget.
* Anisotropy tensor (grid_anisotropy_tensor.h:14-31): closed form diag(h)/|h|, delta = |h| for box elements, and
  (g^{-1/2})/|eig| for a rotated, stretched element against numpy.
* Model (sgs_dd_utils.h, sgs_dd_model.h): the MLP against an independent torch.nn evaluation with the reference
  weights (tests/golden/dd_sgs_data, read with the product's reader); frame invariance of the whole closure (a proper
  rotation R of the velocity gradient and of the anisotropy tensor rotates the stress to R tau R^T); the stress is
  symmetric and traceless-free (no constraint), and vanishes with the velocity gradient.
* Projections: the lumped velocity-gradient projection is exact for a linear velocity field; the CG anisotropy
  projection is exact (constant) on a uniform box mesh.
* IFunction_NodalSgs: discrete conservation (the residual sums to zero over a periodic box), zero residual for a
  uniform stress and state, linearity in the stress.
"""
import os

import numpy as np
import pytest

import pyoracle as O
from pyhonee import mesh as M
from pyhonee import problems as PR
from pyhonee import sgs as S

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dd_sgs_data")


def _rot(rng):
    Qm, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    if np.linalg.det(Qm) < 0:
        Qm[:, 0] *= -1
    return Qm


def _km_pack(A):
    w = np.sqrt(2.0)
    return np.array([A[0, 0], A[1, 1], A[2, 2], w * A[2, 1], w * A[2, 0], w * A[1, 0]])


def _km_unpack(v):
    w = 1 / np.sqrt(2.0)
    return np.array([[v[0], w * v[5], w * v[4]], [w * v[5], v[1], w * v[3]], [w * v[4], w * v[3], v[2]]])


def test_jacobi_eigensolver():
    rng = np.random.default_rng(11)
    for k in range(500):
        B = rng.normal(size=(3, 3)) * rng.uniform(0.01, 100)
        A = B + B.T
        ev, E, it = O.diagonalize3(A, 5)
        assert it > 0
        ref = np.sort(np.linalg.eigvalsh(A))[::-1]
        assert np.allclose(ev, ref, rtol=1e-12, atol=1e-13 * np.abs(ref).max())
        assert ev[0] >= ev[1] >= ev[2]
        assert np.allclose(E @ E.T, np.eye(3), atol=1e-13)
        assert np.allclose(E.T @ np.diag(ev) @ E, A, atol=1e-12 * np.abs(A).max())
    # already diagonal: no rotation, sorted
    ev, E, it = O.diagonalize3(np.diag([1.0, 3.0, 2.0]), 5)
    assert np.array_equal(ev, [3.0, 2.0, 1.0]) and np.array_equal(np.abs(E), np.eye(3)[[1, 2, 0]])


def test_anisotropy_tensor_closed_forms():
    h = np.array([0.3, 0.05, 1.2])
    kmA, delta = O.anisotropy_tensor(np.diag(2 / h))
    assert abs(delta - np.linalg.norm(h)) < 1e-14 * delta
    assert np.allclose(_km_unpack(kmA), np.diag(h) / np.linalg.norm(h), atol=1e-14)
    rng = np.random.default_rng(3)
    for _ in range(20):
        R = _rot(rng)
        J = R @ np.diag(rng.uniform(0.1, 3, 3)) @ _rot(rng)  # dx/dX
        dXdx = np.linalg.inv(J)
        g = dXdx.T @ dXdx
        w, V = np.linalg.eigh(g)
        isq = V @ np.diag(1 / np.sqrt(w)) @ V.T
        nrm = np.linalg.norm(1 / np.sqrt(w))
        kmA, delta = O.anisotropy_tensor(dXdx)
        assert abs(delta - 2 * nrm) < 1e-12 * delta
        assert np.allclose(_km_unpack(kmA), isq / nrm, atol=1e-12)


@pytest.fixture(scope="module")
def model():
    p = S.load_dd_model(GOLDEN, alpha=0.3)
    return p, O.SgsModel(p)


def test_mlp_matches_torch(model):
    torch = pytest.importorskip("torch")
    p, om = model
    net = torch.nn.Sequential(torch.nn.Linear(6, 20), torch.nn.LeakyReLU(0.3), torch.nn.Linear(20, 6)).double()
    with torch.no_grad():
        net[0].weight.copy_(torch.tensor(p["W1"]))
        net[0].bias.copy_(torch.tensor(p["b1"]))
        net[2].weight.copy_(torch.tensor(p["W2"]))
        net[2].bias.copy_(torch.tensor(p["b2"]))
    rng = np.random.default_rng(5)
    for _ in range(50):
        gu = rng.normal(size=(3, 3))
        kmA, delta = O.anisotropy_tensor(np.diag(2 / rng.uniform(0.1, 1, 3)))
        r = O.sgs_point(om, gu, kmA, delta, 1e-3)
        with torch.no_grad():
            ref = net(torch.tensor(r["inputs"])[None])[0].numpy()
        assert np.allclose(r["outputs"], ref, rtol=1e-13, atol=1e-14)
    # the reader's transposes: w1.dat holds [input][neuron]
    w1 = S.read_dat(os.path.join(GOLDEN, "w1.dat")).ravel()
    assert p["W1"][3, 2] == w1[2 * 20 + 3]


def test_closure_frame_invariance(model):
    _, om = model
    rng = np.random.default_rng(9)
    for _ in range(30):
        gu = rng.normal(size=(3, 3))
        Aphys = _km_unpack(O.anisotropy_tensor(np.linalg.inv(np.diag(rng.uniform(0.2, 1, 3)) @ _rot(rng)))[0])
        delta = rng.uniform(0.05, 0.5)
        R = _rot(rng)
        r0 = O.sgs_point(om, gu, _km_pack(Aphys), delta, 2e-3)
        r1 = O.sgs_point(om, R @ gu @ R.T, _km_pack(R @ Aphys @ R.T), delta, 2e-3)
        T0, T1 = _km_unpack(r0["km_sgs"]), _km_unpack(r1["km_sgs"])
        assert np.allclose(r0["inputs"], r1["inputs"], atol=1e-12)
        assert np.allclose(R @ T0 @ R.T, T1, atol=1e-11 * np.abs(T0).max())
    r = O.sgs_point(om, np.zeros((3, 3)), _km_pack(np.eye(3) / np.sqrt(3)), 0.1, 1e-3)
    assert np.all(r["km_sgs"] == 0)


def _box(periodic, degree=2, shape=(3, 2, 2), perturb=0.0):
    mesh = M.box_mesh(shape, degree, lengths=(1.0, 2.0, 1.5), periodic=periodic, perturb=perturb, seed=4)
    om = O.Mesh(mesh, M.tabulate(degree, 0))
    return mesh, om, om.setup_qdata()


def test_velocity_gradient_projection_exact_for_linear_velocity():
    mesh, om, qd = _box((False, False, False), perturb=0.15)
    c = O.newt_from_dict(PR.tgv_context_dict(state_var=1))
    rng = np.random.default_rng(2)
    A, b = rng.normal(size=(3, 3)), rng.normal(size=3)
    Y = PR.nodal_field(mesh, lambda x: np.concatenate([np.full((len(x), 1), 71.0), x @ A.T + b, np.full((len(x), 1), 0.25)], axis=1))
    G = O.velocity_gradient_projection(om, c, qd, Y.ravel()).reshape(-1, 3, 3)
    assert np.allclose(G, A[None], atol=1e-11 * np.abs(A).max())


def test_grid_anisotropy_projection_uniform_box():
    mesh, om, qd = _box((True, True, True), degree=2, shape=(4, 2, 3))
    X, its = O.grid_anisotropy_projection(om, qd)
    h = np.array([1.0 / 4, 2.0 / 2, 1.5 / 3]) / 1.0
    ref = np.concatenate([_km_pack(np.diag(h) / np.linalg.norm(h)), [np.linalg.norm(h)]])
    assert 0 < its < 200
    assert np.allclose(X.reshape(-1, 7), ref[None], atol=1e-9)


def test_ifunction_nodal_sgs_invariants(model):
    mesh, om, qd = _box((True, True, True), degree=2, perturb=0.1)
    c = O.newt_from_dict(PR.tgv_context_dict(state_var=1))
    rng = np.random.default_rng(8)
    Y = PR.nodal_field(mesh, lambda x: PR.taylor_green_prim(x, lx=2.0))
    km = rng.normal(size=(mesh.nnodes, 6)).ravel()
    v = O.ifunction_nodal_sgs(om, c, qd, Y.ravel(), km)
    assert np.allclose(v.reshape(-1, 5).sum(0), 0, atol=1e-12 * np.abs(v).max())
    assert np.allclose(v.reshape(-1, 5)[:, 0], 0)
    v2 = O.ifunction_nodal_sgs(om, c, qd, Y.ravel(), 2 * km)
    assert np.allclose(v2, 2 * v, rtol=1e-13, atol=1e-14 * np.abs(v).max())
    Yu = np.tile([71.0, 0.3, -0.2, 0.1, 0.25], mesh.nnodes)
    ku = np.tile(rng.normal(size=6), mesh.nnodes)
    assert np.abs(O.ifunction_nodal_sgs(om, c, qd, Yu, ku)).max() < 1e-12


def test_sgs_nodal_is_pointwise(model):
    _, om = model
    rng = np.random.default_rng(12)
    n = 7
    gas = O.newt_from_dict(PR.tgv_context_dict(state_var=1)).gas
    q = np.tile([71.0, 0.5, 0.1, -0.2, 0.25], n) * (1 + 0.01 * rng.uniform(-1, 1, 5 * n))
    G = rng.normal(size=9 * n)
    A = np.concatenate([np.concatenate([_km_pack(np.diag(h) / np.linalg.norm(h)), [np.linalg.norm(h)]]) for h in rng.uniform(0.1, 1, (n, 3))])
    km = O.sgs_nodal(gas, 1, om, q, G, A).reshape(n, 6)
    rho = q.reshape(n, 5)[:, 0] / ((gas.cp - gas.cv) * q.reshape(n, 5)[:, 4])
    for i in range(n):
        r = O.sgs_point(om, G[9 * i: 9 * i + 9].reshape(3, 3), A[7 * i: 7 * i + 6], A[7 * i + 6], gas.mu / rho[i])
        assert np.allclose(km[i], r["km_sgs"], rtol=1e-14, atol=1e-16)
