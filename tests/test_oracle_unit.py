"""Pinning the CPU oracle beyond the golden fixtures.

* State transforms: the reference's own known-answer tests (problems/newtonian.c:947-987): U<->Y<->V round trips at
  20 eps and finite-difference checks of every StateFrom*_fwd at 5e-6 (FD step 4e-7), with its fixed state
  T=200, rho=1.2, u=(40, 44, 48).
* Variable independence of the IFunction QFunction: in primitive or entropy variables it must equal the conservative
  QFunction evaluated on the same physical state and tangents (only StateFromQ/StateFromQ_fwd differ between the
  variants, and the conservative path is pinned by the kinetic-energy fixture through IFunction = M_SUPG qdot - RHS).
* IJacobian: exact derivative of the IFunction (central FD) without stabilization on a non-uniform state, and with
  SUPG at a uniform state (where the frozen-Tau terms vanish).
* Discrete conservation on a periodic box: sum over nodes of the residual equals the integral of U_dot.
* Basis/geometry closed forms: polynomial exactness of B and D, weights, affine qdata.
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import make_case, oracle_setup
from pyhonee import mesh as M
from pyhonee import problems as PR

EPS = 2.220446049250313e-16
GAS_DEFAULT = {"lambda": -2.0 / 3.0, "mu": 1.8e-5, "k": 0.02638, "cv": 717.0, "cp": 1004.0}


def _gas(d):
    return O.OGas(d["lambda"], d["mu"], d["k"], d["cv"], d["cp"])


def _kat_states(gas):
    T, rho = 200.0, 1.2
    gamma = gas["cp"] / gas["cv"]
    P = (gamma - 1) * rho * gas["cv"] * T
    u = np.array([40.0, 44.0, 48.0])
    ek = 0.5 * u @ u
    et = ek + gas["cv"] * T
    entropy = np.log(P) - gamma * np.log(rho)
    rdp = rho / P
    Y0 = np.array([P, *u, T])
    U0 = np.array([rho, *(rho * u), rho * et])
    V0 = np.array([(gamma - entropy) / (gamma - 1) - rdp * ek, *(rdp * u), -rdp])
    return {0: U0, 1: Y0, 2: V0}


def _check(Q_s, Q_a, Q_b, rtol):
    """CheckQWithTolerance (problems/newtonian.c:858-885)."""
    thr = 10 * EPS
    rel = np.zeros(5)
    rel[0] = (Q_a[0] - Q_b[0]) / (Q_s[0] if abs(Q_s[0]) > thr else 1)
    rel[4] = (Q_a[4] - Q_b[4]) / (Q_s[4] if abs(Q_s[4]) > thr else 1)
    um = np.linalg.norm(Q_s[1:4])
    rel[1:4] = (Q_a[1:4] - Q_b[1:4]) / (um if um > thr else 1)
    assert np.all(np.abs(rel) < rtol), rel


@pytest.mark.parametrize("a,b", [(1, 0), (1, 2), (0, 1), (0, 2), (2, 0), (2, 1)])
def test_state_round_trip(a, b):
    gas = _gas(GAS_DEFAULT)
    A0 = _kat_states(GAS_DEFAULT)[a]
    U, Y = O.state_from_q(gas, A0, a)
    B0 = O.state_to_q(gas, U, Y, b)
    U, Y = O.state_from_q(gas, B0, b)
    A1 = O.state_to_q(gas, U, Y, a)
    _check(A0, A1, A0, 20 * EPS)


@pytest.mark.parametrize("a,b,fac", [(1, 0, (1, 1, 1)), (1, 2, (1, 1, 1)), (0, 1, (1, 1, 1)), (0, 2, (10, 1, 1)), (2, 0, (5, 1, 1)),
                                     (2, 1, (5, 5, 5))])
def test_state_fwd_finite_difference(a, b, fac):
    gas = _gas(GAS_DEFAULT)
    A0 = _kat_states(GAS_DEFAULT)[a]
    eps = 4e-7
    U0, Y0 = O.state_from_q(gas, A0, a)
    for i in range(5):
        dA = np.zeros(5)
        dA[i] = A0[i]
        dU, dY = O.state_from_q_fwd(gas, A0, dA, a)
        dB = O.state_to_q_fwd(gas, U0, Y0, dU, dY, b)
        A1 = A0 * (1 + eps * (np.arange(5) == i))
        U1, Y1 = O.state_from_q(gas, A1, a)
        dB_fd = (O.state_to_q(gas, U1, Y1, b) - O.state_to_q(gas, U0, Y0, b)) / eps
        rtol = 5e-6 * np.array(fac)
        thr = 10 * EPS
        rel0 = (dB[0] - dB_fd[0]) / (dB_fd[0] if abs(dB_fd[0]) > thr else 1)
        rel4 = (dB[4] - dB_fd[4]) / (dB_fd[4] if abs(dB_fd[4]) > thr else 1)
        um = np.linalg.norm(dB_fd[1:4])
        relu = (dB[1:4] - dB_fd[1:4]) / (um if um > thr else 1)
        assert abs(rel0) < rtol[0] and np.all(np.abs(relu) < rtol[1]) and abs(rel4) < rtol[2], (i, rel0, relu, rel4)


def _random_point(rng, gas):
    Y = np.array([71.4 * (1 + 0.1 * rng.uniform(-1, 1)), *rng.uniform(-1, 1, 3), 0.249 * (1 + 0.1 * rng.uniform(-1, 1))])
    gY = rng.uniform(-1, 1, (3, 5)) * np.array([1.0, 1, 1, 1, 0.01])
    Ydot = rng.uniform(-1, 1, 5) * np.array([1.0, 1, 1, 1, 0.01])
    dXdx = np.eye(3) * 2.0 + 0.3 * rng.uniform(-1, 1, (3, 3))
    return Y, gY, Ydot, dXdx


@pytest.mark.parametrize("var", [1, 2])
@pytest.mark.parametrize("stab", [0, 1, 2])
def test_ifunction_qf_variable_independence(var, stab):
    rng = np.random.default_rng(100 + 3 * var + stab)
    ctx = PR.tgv_context_dict(state_var=0, stabilization=stab, dt=1e-3)
    gas = _gas(ctx["gas"])
    c = O.newt_from_dict(ctx)
    for _ in range(5):
        Y, gY, Ydot, dXdx = _random_point(rng, gas)
        U, _ = O.state_from_q(gas, Y, 1)
        q_var = O.state_to_q(gas, U, Y, var)
        # tangents expressed in the chosen variables and in conservative variables
        dvar = lambda dy: O.state_to_q_fwd(gas, U, Y, *O.state_from_q_fwd(gas, Y, dy, 1), var)  # noqa: E731
        dcons = lambda dy: O.state_from_q_fwd(gas, Y, dy, 1)[0]  # noqa: E731
        g_var = np.array([dvar(gY[d]) for d in range(3)])
        g_cons = np.array([dcons(gY[d]) for d in range(3)])
        v_ref, gv_ref, _ = O.qf_ifunction(c, U, g_cons, dcons(Ydot), 0.7, dXdx)
        c.state_var = var
        v, gv, _ = O.qf_ifunction(c, q_var, g_var, dvar(Ydot), 0.7, dXdx)
        c.state_var = 0
        scale = np.abs(gv_ref).max()
        assert np.abs(v - v_ref).max() <= 1e-10 * np.abs(v_ref).max() + 1e-13
        assert np.abs(gv - gv_ref).max() <= 1e-10 * scale, np.abs(gv - gv_ref).max() / scale


def test_ifunction_equals_stabilized_mass_minus_rhs():
    """Conservative variables: IFunction(q, qdot) = M_SUPG(q) qdot - RHS(q) (newtonian.h:44-82, 133-202, 205-257)."""
    case = make_case(shape=(2, 2, 2), degree=2, state_var=0, stabilization=2, seed=5)
    om, oc, qd = oracle_setup(case)
    v, _ = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    m = O.mass_newtonian(om, oc, qd, case["q"], case["qdot"])
    r = O.rhs(om, oc, qd, case["q"])
    np.testing.assert_allclose(v, m - r, rtol=0, atol=1e-11 * np.abs(v).max())


def test_ijacobian_is_derivative_without_stabilization():
    """Primitive variables, no stabilization, q_dot = 0: the viscous and inviscid fluxes depend on q and on grad(q)
    linearly through grad Y, so IJacobian_Newtonian is the exact derivative.  (In conservative or entropy variables
    the reference's linearization omits the dependence of grad Y on the state, and with q_dot != 0 the dependence of
    dU/dq on q; those terms vanish at a uniform state, checked below with SUPG.)"""
    case = make_case(shape=(2, 2, 2), degree=2, state_var=1, stabilization=0, seed=21, g=(0.0, -1.0, 0.5))
    om, oc, qd = oracle_setup(case)
    q, qdot = case["q"], np.zeros_like(case["q"])
    _, jac = O.ifunction(om, oc, qd, q, qdot)
    shift = 37.0
    oc.ijacobian_time_shift = shift
    q5 = q.reshape(-1, 5)
    dq = (case["rng"].uniform(-1, 1, q5.shape) * np.abs(q5).mean(0)).ravel()
    Jdq = O.ijacobian(om, oc, qd, jac, dq)
    h = 1e-6
    vp, _ = O.ifunction(om, oc, qd, q + h * dq, qdot + shift * h * dq)
    vm, _ = O.ifunction(om, oc, qd, q - h * dq, qdot - shift * h * dq)
    fd = (vp - vm) / (2 * h)
    assert np.linalg.norm(Jdq - fd) / np.linalg.norm(fd) < 1e-7


@pytest.mark.parametrize("var", [0, 1, 2])
def test_ijacobian_supg_at_uniform_state(var):
    mesh = M.box_mesh((2, 2, 2), 2, perturb=0.2, seed=3)
    tab = M.tabulate(2, 0)
    om = O.Mesh(mesh, tab)
    ctx = PR.tgv_context_dict(state_var=var, stabilization=2, dt=1e-3, tau=PR.default_tau(2))
    oc = O.newt_from_dict(ctx)
    qd = om.setup_qdata()
    Y = np.tile([71.4, 0.3, -0.2, 0.1, 0.249], (mesh.nnodes, 1))
    q = PR.q_from_prim(PR.TGV_GAS, Y, var).ravel()
    qdot = np.zeros_like(q)
    _, jac = O.ifunction(om, oc, qd, q, qdot)
    shift = 50.0
    oc.ijacobian_time_shift = shift
    rng = np.random.default_rng(var)
    q5 = q.reshape(-1, 5)
    dq = (rng.uniform(-1, 1, q5.shape) * np.abs(q5).mean(0)).ravel()
    Jdq = O.ijacobian(om, oc, qd, jac, dq)
    h = 1e-6
    vp, _ = O.ifunction(om, oc, qd, q + h * dq, qdot + shift * h * dq)
    vm, _ = O.ifunction(om, oc, qd, q - h * dq, qdot - shift * h * dq)
    fd = (vp - vm) / (2 * h)
    assert np.linalg.norm(Jdq - fd) / np.linalg.norm(fd) < 1e-7


@pytest.mark.parametrize("var", [0, 1, 2])
@pytest.mark.parametrize("stab", [0, 2])
def test_periodic_conservation(var, stab):
    case = make_case(shape=(2, 3, 2), degree=3, state_var=var, stabilization=stab, seed=40 + var)
    om, oc, qd = oracle_setup(case)
    v, _ = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    # integral of U_dot: sum over elements and points of wdetJ * U_dot (interpolated through the oracle mass term)
    zero_flux = O.ifunction(om, oc, qd, case["q"], np.zeros_like(case["q"]))[0]
    total = v.reshape(-1, 5).sum(0)
    flux_total = zero_flux.reshape(-1, 5).sum(0)
    scale = np.abs(v).max()
    assert np.all(np.abs(flux_total) < 1e-11 * scale * np.sqrt(v.size)), flux_total / scale


def test_uniform_state_zero_residual():
    mesh = M.box_mesh((2, 2, 2), 3, perturb=0.2, seed=4)
    om = O.Mesh(mesh, M.tabulate(3, 0))
    qd = om.setup_qdata()
    for var in (0, 1, 2):
        oc = O.newt_from_dict(PR.tgv_context_dict(state_var=var, stabilization=2))
        q = PR.q_from_prim(PR.TGV_GAS, np.tile([71.4, 0.3, -0.2, 0.1, 0.249], (mesh.nnodes, 1)), var).ravel()
        v, _ = O.ifunction(om, oc, qd, q, np.zeros_like(q))
        # each node's test function sees a constant flux: sum of grad(phi_i) over its support is zero
        assert np.abs(v).max() < 1e-11 * 250.0


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_basis_polynomial_exactness(degree):
    tab = M.tabulate(degree, 1)
    nodes = M.gll_nodes(degree + 1)
    for k in range(degree + 1):
        np.testing.assert_allclose(tab.B @ nodes**k, tab.qref**k, atol=1e-14)
        d = k * tab.qref ** (k - 1) if k else np.zeros_like(tab.qref)
        np.testing.assert_allclose(tab.D @ nodes**k, d, atol=1e-13)
    assert abs(tab.qweight.sum() - 2.0) < 1e-14
    # Gauss rule exact to degree 2Q-1
    assert abs(tab.qweight @ tab.qref ** (2 * tab.Q - 2) - 2.0 / (2 * tab.Q - 1)) < 1e-14


def test_gll_nodes_closed_form():
    np.testing.assert_allclose(M.gll_nodes(3), [-1, 0, 1], atol=1e-15)
    np.testing.assert_allclose(M.gll_nodes(4), [-1, -1 / np.sqrt(5), 1 / np.sqrt(5), 1], atol=1e-15)
    np.testing.assert_allclose(M.gll_nodes(5), [-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1], atol=1e-15)


def test_affine_qdata_closed_form():
    L = (2.0, 3.0, 5.0)
    mesh = M.box_mesh((2, 3, 4), 2, lengths=L, periodic=(False, False, False))
    om = O.Mesh(mesh, M.tabulate(2, 0))
    qd = om.setup_qdata()
    h = np.array([L[0] / 2, L[1] / 3, L[2] / 4])
    w = M.tabulate(2, 0).qweight
    W = np.einsum("k,j,i->kji", w, w, w).ravel()
    np.testing.assert_allclose(qd[:, 0, :], np.broadcast_to(W * np.prod(h / 2), qd[:, 0, :].shape), rtol=1e-14)
    dXdx = qd[:, 1:, :].reshape(mesh.nelem, 3, 3, -1)
    np.testing.assert_allclose(dXdx, np.broadcast_to(np.diag(2 / h)[None, :, :, None], dXdx.shape), atol=1e-14)
    assert abs(O.volume(om, qd) - np.prod(L)) < 1e-12


def test_ijacobian_diagonal_matches_columns():
    case = make_case(shape=(2, 2, 2), degree=1, state_var=1, stabilization=2, seed=9)
    om, oc, qd = oracle_setup(case)
    _, jac = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    oc.ijacobian_time_shift = 10.0
    diag = O.ijacobian_diagonal(om, oc, qd, jac)
    n = case["q"].size
    cols = np.empty(n)
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        cols[i] = O.ijacobian(om, oc, qd, jac, e)[i]
    np.testing.assert_allclose(diag, cols, rtol=1e-12, atol=1e-12 * np.abs(cols).max())


def test_oracle_point_block_diagonal_brute_force():
    """The oracle's point-block diagonal equals the node blocks of the assembled operator's unit columns, and its
    block diagonals equal the point diagonal."""
    from helpers import make_case, oracle_setup

    case = make_case(shape=(2, 2, 2), degree=1, state_var=1, seed=3)
    om, oc, qd = oracle_setup(case)
    _, jac = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    oc.ijacobian_time_shift = 50.0
    pb = O.ijacobian_point_block_diagonal(om, oc, qd, jac)
    nn = om.nnodes
    ref = np.zeros((nn, 5, 5))
    for node in range(nn):
        for ci in range(5):
            e = np.zeros(5 * nn)
            e[node * 5 + ci] = 1
            ref[node, :, ci] = O.ijacobian(om, oc, qd, jac, e)[node * 5: node * 5 + 5]
    assert np.abs(ref - pb).max() <= 1e-13 * np.abs(ref).max()
    np.testing.assert_allclose(np.einsum("nkk->nk", pb).ravel(), O.ijacobian_diagonal(om, oc, qd, jac), rtol=0, atol=1e-13 * np.abs(ref).max())


@pytest.mark.parametrize("degree", [1, 2, 3])
def test_q_extra_quadrature_integrates_the_volume(degree):
    """With Q_1d = P_1d + q_extra the quadrature still integrates det J exactly: the trilinear map's det J is of degree
    at most 2 in each direction, inside every Gauss rule here, so a perturbed mesh's volume is the same for q_extra
    0, 1 and 2, and an affine box's is the box volume."""
    mesh = M.box_mesh((2, 2, 2), degree, perturb=0.15, seed=3)
    vols = []
    for q_extra in (0, 1, 2):
        om = O.Mesh(mesh, M.tabulate(degree, q_extra))
        vols.append(O.volume(om, om.setup_qdata()))
    assert max(vols) - min(vols) < 1e-14 * vols[0]
    box = M.box_mesh((2, 2, 2), degree)
    om = O.Mesh(box, M.tabulate(degree, 2))
    assert abs(O.volume(om, om.setup_qdata()) - (2 * np.pi) ** 3) < 1e-13 * (2 * np.pi) ** 3  # 1728 positive terms summed
