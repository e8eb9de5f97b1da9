"""The oracle's differential filter and SGS training pairs against what the mathematics fixes (no GPU).

* A constant state is reproduced exactly by the filter (the diffusion term vanishes, R = M c = LHS c).
* On an affine periodic box the discrete filter acts on a Fourier mode like the continuous operator
  (I - div(alpha grad)): sin(k x) -> sin(k x) / (1 + alpha_xx k^2), with alpha = kernel_scaling diag(width_scaling)^2
  for a grid-independent width and alpha = kernel_scaling diag(width_scaling h)^2 for the grid-based width (the
  anisotropy tensor of a box element with sides h is Delta = diag(h)); checked per direction on a stretched box.
* Wall damping: van Driest with a vanishing friction length and the MMS damping far from the wall reduce to the
  undamped filter with the same widths.
* The LHS diagonal equals the diagonal of the operator assembled column by column.
* Training pairs: the six inputs are the closure inputs of the filtered state at the node, and the targets are the
  filtered velocity products minus the products of the filtered velocity (XX YY ZZ YZ XZ XY).
"""
import numpy as np

import pyoracle as O
from pyhonee import mesh as M
from pyhonee import problems as PR


def _box(shape, degree, lengths, periodic=(True, True, True), perturb=0.0):
    mesh = M.box_mesh(shape, degree, lengths=lengths, periodic=periodic, perturb=perturb, seed=3)
    om = O.Mesh(mesh, M.tabulate(degree, 0))
    return mesh, om, om.setup_qdata()


def _prim_state(mesh, fu=None, fv=None):
    def f(x):
        Y = np.zeros((len(x), 5))
        Y[:, 0] = 71.0
        Y[:, 4] = 0.25
        if fu is not None:
            Y[:, 1] = fu(x)
        if fv is not None:
            Y[:, 2] = fv(x)
        return Y
    return PR.nodal_field(mesh, f)


def _nodal(mesh, fn):
    """A scalar nodal field from the element nodes' coordinates (multiplicity-averaged, consistent on periodic meshes)."""
    xyz = mesh.element_node_coords().reshape(-1, 3)
    v = np.zeros(mesh.nnodes)
    np.add.at(v, mesh.offsets.ravel(), fn(xyz))
    return v / mesh.multiplicity()


def _newt():
    return O.newt_from_dict(PR.tgv_context_dict(state_var=1))


def test_constant_state_is_preserved():
    mesh, om, qd = _box((3, 2, 2), 2, (1.0, 2.0, 1.5), perturb=0.15)
    c = _newt()
    Y = np.tile([71.0, 0.3, -0.2, 0.1, 0.25], mesh.nnodes)
    Acol = O.anisotropy_collocate(om, qd)
    R = O.diff_filter_rhs(om, c, qd, Y)
    for f in (O.diff_filter(), O.diff_filter(grid_based_width=True, width_scaling=(2.0, 0.5, 1.0))):
        X, its = O.diff_filter_solve(om, f, qd, Acol, 11, R)
        u = Y[1:4]
        ref = np.concatenate([Y[:5], [u[0] * u[0], u[1] * u[1], u[2] * u[2], u[1] * u[2], u[0] * u[2], u[0] * u[1]]])
        assert 0 < its < 500
        assert np.allclose(X.reshape(-1, 11), ref[None], rtol=1e-8, atol=1e-10)


def _response(om, qd, f, comp, mode):
    c = _newt()
    Acol = O.anisotropy_collocate(om, qd)
    R = O.diff_filter_rhs(om, c, qd, mode)
    X, _ = O.diff_filter_solve(om, f, qd, Acol, 11, R, rtol=1e-13)
    return X.reshape(-1, 11)[:, comp]


def test_fourier_response_isotropic():
    mesh, om, qd = _box((8, 2, 2), 3, (1.0, 1.0, 1.0))
    k = 2 * np.pi
    f = O.diff_filter(width_scaling=(0.5, 0.5, 0.5))  # alpha = 0.1 * 0.25
    Y = _prim_state(mesh, fu=lambda x: np.sin(k * x[:, 0]))
    got = _response(om, qd, f, 1, Y.ravel())
    ref = _nodal(mesh, lambda x: np.sin(k * x[:, 0])) / (1 + 0.025 * k * k)
    assert np.abs(got - ref).max() < 2e-4, np.abs(got - ref).max()
    # a wrong width factor would move the response far more than the discretization error
    assert abs(1 / (1 + 0.025 * k * k) - 1 / (1 + 0.1 * k * k)) > 0.2


def test_fourier_response_grid_based_anisotropic():
    shape, lengths = (8, 4, 2), (1.0, 2.0, 1.0)
    mesh, om, qd = _box(shape, 3, lengths)
    h = np.array(lengths) / np.array(shape)
    w = np.array([4.0, 1.0, 1.0])
    f = O.diff_filter(grid_based_width=True, width_scaling=tuple(w))
    alpha = 0.1 * (w * h) ** 2
    kx, ky = 2 * np.pi / lengths[0], 2 * np.pi / lengths[1]
    Y = _prim_state(mesh, fu=lambda x: np.sin(kx * x[:, 0]), fv=lambda x: np.cos(ky * x[:, 1]))
    gu = _response(om, qd, f, 1, Y.ravel())
    gv = _response(om, qd, f, 2, Y.ravel())
    eu = np.abs(gu - _nodal(mesh, lambda x: np.sin(kx * x[:, 0])) / (1 + alpha[0] * kx * kx)).max()
    ev = np.abs(gv - _nodal(mesh, lambda x: np.cos(ky * x[:, 1])) / (1 + alpha[1] * ky * ky)).max()
    assert eu < 2e-4 and ev < 2e-3, (eu, ev)


def test_wall_damping_limits():
    mesh, om, qd = _box((3, 3, 2), 2, (1.0, 1.0, 1.0), periodic=(True, False, True))
    rng = np.random.default_rng(4)
    Acol = O.anisotropy_collocate(om, qd)
    u = rng.normal(size=3 * mesh.nnodes)
    base = O.diff_filter_lhs(om, O.diff_filter(grid_based_width=True, width_scaling=(1.5, 0.7, 2.0)), qd, Acol, 3, u)
    vd = O.diff_filter_lhs(om, O.diff_filter(grid_based_width=True, width_scaling=(1.5, 0.7, 2.0), damping=1, friction_length=1e-300), qd, Acol, 3, u)
    assert np.allclose(vd, base, rtol=1e-13, atol=1e-13 * np.abs(base).max())
    # MMS damping tanh(60 y) is 1 in double precision for y >= 1: the mesh is shifted to y in [1, 2]
    mesh2 = M.box_mesh((3, 3, 2), 2, lengths=(1.0, 1.0, 1.0), periodic=(True, False, True), perturb=0.0, seed=3)
    mesh2.xcoord[:, 1] += 1.0
    om2 = O.Mesh(mesh2, M.tabulate(2, 0))
    qd2 = om2.setup_qdata()
    A2 = O.anisotropy_collocate(om2, qd2)
    mms = O.diff_filter_lhs(om2, O.diff_filter(grid_based_width=True, damping=2), qd2, A2, 3, u)
    ref = O.diff_filter_lhs(om2, O.diff_filter(grid_based_width=True, width_scaling=(1.0, 1.0, 1.0)), qd2, A2, 3, u)
    assert np.allclose(mms, ref, rtol=1e-13, atol=1e-13 * np.abs(ref).max())


def test_lhs_symmetric_and_diagonal():
    mesh, om, qd = _box((2, 2, 2), 2, (1.0, 1.5, 1.0), perturb=0.1)
    f = O.diff_filter(grid_based_width=True, width_scaling=(1.2, 0.8, 1.0))
    Acol = O.anisotropy_collocate(om, qd)
    n = mesh.nnodes
    Amat = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        Amat[:, j] = O.diff_filter_lhs(om, f, qd, Acol, 1, e)
    # the reference form alpha (dXdx grad q) is symmetric only where alpha and dXdx commute; the mass part is symmetric
    d = O.diff_filter_lhs_diagonal(om, f, qd, Acol, 1)
    assert np.allclose(d, np.diag(Amat), rtol=1e-12, atol=1e-15)
    d3 = O.diff_filter_lhs_diagonal(om, f, qd, Acol, 3).reshape(-1, 3)
    assert np.array_equal(d3[:, 0], d3[:, 2])


def test_training_pairs_pointwise():
    p = PR.tgv_context_dict(state_var=1)
    gas = O.newt_from_dict(p).gas
    model = O.SgsModel({"b1": np.zeros(20), "b2": np.zeros(6), "W1": np.zeros((20, 6)), "W2": np.zeros((6, 20)),
                        "bounds": np.zeros((6, 2)), "alpha": 0.0})
    rng = np.random.default_rng(21)
    n = 6
    Yf = np.zeros((n, 11))
    Yf[:, 0] = 71.0 * (1 + 0.01 * rng.uniform(-1, 1, n))
    Yf[:, 1:4] = rng.normal(size=(n, 3))
    Yf[:, 4] = 0.25 * (1 + 0.01 * rng.uniform(-1, 1, n))
    Yf[:, 5:] = rng.normal(size=(n, 6))
    G = rng.normal(size=(n, 9))
    A = np.concatenate([np.concatenate([[1, 1, 1, 0, 0, 0] / np.sqrt(3.0), [0.3]])[None] for _ in range(n)])
    out = O.sgs_training_nodal(gas, Yf.ravel(), G.ravel(), A.ravel()).reshape(n, 12)
    R = gas.cp - gas.cv
    for i in range(n):
        rho = Yf[i, 0] / (R * Yf[i, 4])
        r = O.sgs_point(model, G[i].reshape(3, 3), A[i, :6], A[i, 6], gas.mu / rho)
        assert np.allclose(out[i, :6], r["inputs"], rtol=1e-14, atol=1e-15)
        u = Yf[i, 1:4]
        uu = np.array([u[0] * u[0], u[1] * u[1], u[2] * u[2], u[1] * u[2], u[0] * u[2], u[0] * u[1]])
        assert np.allclose(out[i, 6:], Yf[i, 5:] - uu, rtol=1e-15, atol=1e-15)
