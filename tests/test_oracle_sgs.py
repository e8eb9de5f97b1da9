"""The oracle's data-driven SGS closure against what the mathematics fixes (no GPU).

* Jacobi eigen solver (utils_eigensolver_jacobi.h): eigenvalues against numpy.linalg.eigvalsh, decreasing order,
  orthonormal eigenvector rows that reconstruct the matrix, convergence inside the sweep budget.
* Anisotropy tensor (grid_anisotropy_tensor.h:14-31): closed form diag(h)/|h|, delta = |h| for box elements, and
  (g^{-1/2})/|eig| for a rotated, stretched element against numpy.
* Model (sgs_dd_utils.h, sgs_dd_model.h): the MLP against an independent torch.nn evaluation with the reference
  weights (tests/golden/dd_sgs_data, read with the product's reader); frame invariance of the whole closure (a proper
  rotation R of the velocity gradient and of the anisotropy tensor rotates the stress to R tau R^T); the stress is
  symmetric and traceless-free (no constraint), and vanishes with the velocity gradient.
* Input normalization and output de-normalization (sgs_dd_utils.h:40-127) by closed forms: a diagonal (pure strain)
  gradient gives its sorted entries, zero vorticity and nu / delta^2 over |grad u|; a simple shear through a diagonal
  anisotropy tensor gives (1/2, 0, -1/2, 0, +-1, nu / (delta^2 s)); a rigid rotation gives the oriented vorticity
  (|w0|, sign(w0) sign(w2) w1) over |w| / sqrt 2; a zero-weight network gives the stress
  delta^2 |grad u|^2 ((hi - lo)(b2 - 1) + hi) on the permuted diagonal.
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


# ---- closed forms of the model's input normalization and output de-normalization (sgs_dd_utils.h:40-127) ----
def _zero_net(b2, bounds, alpha=0.3):
    """A 6-20-6 network whose output is its bias b2 (all weights zero), with min-max bounds [6][2]."""
    return O.SgsModel({"b1": np.linspace(-1, 1, 20), "b2": np.asarray(b2, float), "W1": np.zeros((20, 6)), "W2": np.zeros((6, 20)),
                       "bounds": np.asarray(bounds, float).ravel(), "alpha": alpha})


_B2 = [0.15, 0.35, 0.55, 0.75, 0.95, 1.25]
_BOUNDS = [[-0.7, 0.4], [-0.2, 0.9], [-1.3, -0.1], [0.05, 0.6], [-0.4, 0.2], [-0.9, 1.7]]


def test_inputs_pure_strain_closed_form():
    """Diagonal velocity gradient G = gamma diag(1, 3, 2) with grid anisotropy A = diag(0.9, 1.1, 0.7): the isotropic
    gradient G A is diagonal, so the strain-rate eigenvalues are its entries sorted decreasing, the vorticity vanishes,
    |grad u| is their Euclidean norm and the sixth input is nu / delta^2 over it (a slip to nu / delta, or to A G for a
    non-commuting pair, is caught by the shear case below)."""
    gamma, h, delta, nu = 2.5, np.array([0.9, 1.1, 0.7]), 0.3, 1.7e-3
    G = gamma * np.diag([1.0, 3.0, 2.0])
    r = O.sgs_point(_zero_net(_B2, _BOUNDS), G, [*h, 0, 0, 0], delta, nu)
    lam = np.sort(gamma * np.array([1.0, 3.0, 2.0]) * h)[::-1]
    gmag = np.linalg.norm(lam)
    assert r["gmag"] == pytest.approx(gmag, rel=1e-14)
    expect = np.array([*lam, 0.0, 0.0, nu / delta**2]) / (gmag + 2.0**-52)
    np.testing.assert_allclose(r["inputs"], expect, rtol=1e-13, atol=1e-15)
    # eigenvector rows: e0 = +-y, e1 = +-z, e2 = +-x (sorted 3.3, 1.4, 0.9 times gamma)
    np.testing.assert_allclose(np.abs(r["evec"]), np.eye(3)[[1, 2, 0]], atol=1e-15)


@pytest.mark.parametrize("g", [0.8, -1.9])
def test_inputs_simple_shear_closed_form(g):
    """du/dy = g with A = diag(0.8, 1.25, 1): the isotropic gradient is G A, whose only entry is s = g * 1.25 (A G would
    give g * 0.8).  Its strain rate 1/2 (G A + (G A)^T) has eigenvalues |s|/2, 0, -|s|/2; the vorticity -2 R_01 = -s lies
    along the zero-strain eigenvector; |grad u| = sqrt(|S|^2 + |omega|^2 / 2) = |s|.  So the inputs are (1/2, 0, -1/2,
    0, +-1, nu / (delta^2 |s|)): a missing 1/2 in the strain rate, a missing 2 in the vorticity or a weight other than 1/2 on
    the vorticity in |grad u| changes the first three, the order of the product changes the last."""
    h, delta, nu = np.array([0.8, 1.25, 1.0]), 0.05, 1.5e-5
    G = np.zeros((3, 3))
    G[0, 1] = g
    r = O.sgs_point(_zero_net(_B2, _BOUNDS), G, [*h, 0, 0, 0], delta, nu)
    s = abs(g * h[1])
    assert r["gmag"] == pytest.approx(s, rel=1e-14)
    inp = r["inputs"]
    np.testing.assert_allclose(inp[[0, 1, 2, 3]], np.array([0.5, 0.0, -0.5, 0.0]) * s / (s + 2.0**-52), atol=1e-14)
    assert abs(abs(inp[4]) - s / (s + 2.0**-52)) < 1e-14
    assert inp[5] == pytest.approx(nu / delta**2 / (s + 2.0**-52), rel=1e-13)


@pytest.mark.parametrize("omega", [(0.3, -0.7, 1.1), (-0.3, -0.7, 1.1), (0.4, 0.9, -0.6)])
def test_inputs_rigid_rotation_closed_form(omega):
    """u = Omega x x: the strain rate vanishes (the eigen solver leaves the identity frame, evec rows x, y, z), the
    vorticity is curl u = 2 Omega, and OrientBasisWithVector flips e0 and e2 to align with it and sets e1 = e2 x e0.  So
    the S-frame vorticity is (|w0|, sign(w0) sign(w2) w1, |w2|), |grad u| = |w| / sqrt 2, and the inputs are
    (0, 0, 0, |w0|, sign(w0) sign(w2) w1, nu / delta^2) / |grad u|: a swapped or sign-flipped vorticity component fails."""
    Om = np.asarray(omega)
    G = np.array([[0, -Om[2], Om[1]], [Om[2], 0, -Om[0]], [-Om[1], Om[0], 0]])
    delta, nu = 0.2, 3e-3
    r = O.sgs_point(_zero_net(_B2, _BOUNDS), G, [1, 1, 1, 0, 0, 0], delta, nu)
    w = 2 * Om
    gmag = np.linalg.norm(w) / np.sqrt(2)
    assert r["gmag"] == pytest.approx(gmag, rel=1e-14)
    sg = lambda v: 1.0 if v >= 0 else -1.0  # noqa: E731
    vs = [abs(w[0]), sg(w[0]) * sg(w[2]) * w[1]]
    expect = np.array([0, 0, 0, *vs, nu / delta**2]) / (gmag + 2.0**-52)
    np.testing.assert_allclose(r["inputs"], expect, rtol=1e-13, atol=1e-15)
    e0, e2 = sg(w[0]) * np.array([1.0, 0, 0]), sg(w[2]) * np.array([0, 0, 1.0])
    np.testing.assert_allclose(r["evec"], np.array([e0, np.cross(e2, e0), e2]), atol=1e-15)


def test_outputs_denormalization_and_rotation_closed_form():
    """With a zero-weight network the outputs are b2; ComputeSgsDDOutputs maps them to (hi - lo)(b2 - 1) + hi, scales
    by delta^2 |grad u|^2 and rotates the S-frame tensor [[o0, o3, o4], [o3, o1, o5], [o4, o5, o2]] back with E^T S E.
    For the diagonal gradient of the pure-strain case E's rows are +-y, +-z, +-x, so tau_yy = o0, tau_zz = o1,
    tau_xx = o2 exactly and the shear components are +-o3 (yz), +-o4 (xy), +-o5 (xz).  delta instead of delta^2, |grad u|
    instead of its square, a different de-normalization or E S E^T (which puts o1 at xx) all fail."""
    gamma, delta, nu = 1.3, 0.11, 1e-3
    G = gamma * np.diag([1.0, 3.0, 2.0])
    r = O.sgs_point(_zero_net(_B2, _BOUNDS), G, [1, 1, 1, 0, 0, 0], delta, nu)
    np.testing.assert_allclose(r["outputs"], _B2, rtol=0, atol=0)
    gmag = gamma * np.sqrt(14.0)
    lo, hi = np.array(_BOUNDS)[:, 0], np.array(_BOUNDS)[:, 1]
    o = ((hi - lo) * (np.array(_B2) - 1) + hi) * delta**2 * gmag**2
    tau = _km_unpack(r["km_sgs"])
    np.testing.assert_allclose([tau[1, 1], tau[2, 2], tau[0, 0]], o[:3], rtol=1e-13)
    np.testing.assert_allclose(np.abs([tau[1, 2], tau[0, 1], tau[0, 2]]), np.abs(o[3:]), rtol=1e-13)
    np.testing.assert_allclose(tau, tau.T, atol=0)


def test_closed_forms_catch_a_planted_delta_slip():
    """The pins must bite: evaluating the de-normalized stress with delta in place of delta^2 (a plausible slip in the
    re-dimensionalization) moves tau_xx off the closed form by the factor 1 / delta."""
    gamma, delta = 1.3, 0.11
    G = gamma * np.diag([1.0, 3.0, 2.0])
    r = O.sgs_point(_zero_net(_B2, _BOUNDS), G, [1, 1, 1, 0, 0, 0], delta, 1e-3)
    lo, hi = np.array(_BOUNDS)[2]
    right = ((hi - lo) * (_B2[2] - 1) + hi) * delta**2 * 14 * gamma**2
    slipped = ((hi - lo) * (_B2[2] - 1) + hi) * delta * 14 * gamma**2
    tau_xx = _km_unpack(r["km_sgs"])[0, 0]
    assert tau_xx == pytest.approx(right, rel=1e-13) and abs(tau_xx - slipped) > 1e-3 * abs(right)
