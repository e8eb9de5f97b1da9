"""Pinning the CPU oracle's outflow (Riemann and weak-pressure, qfunctions/bc_outflow.h) and slip-wall
(qfunctions/bc_slip.h) face operators by what the mathematics fixes, and through the boundary integral, which the
reference's Taylor-Green goldens pin:

* Slip wall: the HLLC flux against the mirror state has a zero contact speed, so the wall is impermeable for any
  interior state (no mass or energy flux at any node), and on a plane wall the momentum flux is normal to it.
* Riemann outflow: for a uniform state at the outflow pressure and temperature (recirc 1, no velocity correction) the
  exterior state equals the interior one, HLLC reduces to F(U).n, and the face residual equals the consistent
  boundary integral's.
* Weak-pressure outflow: the interior state at the prescribed pressure; where the interior pressure already is that
  pressure, the residual equals the boundary integral's (also with gradients), and for a fluid at rest it is the
  boundary integral's scaled by p_out / p.
* Jacobians: central finite differences of the face residuals (primitive variables, where the gradient linearization
  is exact; the slip wall has no gradient terms and is checked in all three variable sets).  The Riemann outflow is
  checked at recirc 1: below 1 the reference's Jacobian uses Softplus_fwd's value, the derivative factor without the
  dx it is passed (bc_outflow.h:23-27), which the restatement keeps.
* Explicit mode negates the residual.
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import make_case, oracle_setup, rel_err
from pyhonee import mesh as M


def _faces(mesh, sides):
    fe, fi = M.box_boundary(mesh, sides)
    return O.Faces(fe, fi)


def _fd(res, jac, q, dq, h):
    return rel_err(jac(dq), (res(q + h * dq) - res(q - h * dq)) / (2 * h))[0]


def _dq(case, q, scale=1e-3):
    rng = np.random.default_rng(5)
    return rng.uniform(-1, 1, q.shape) * np.abs(q.reshape(-1, 5)).mean(0)[None, :].repeat(case["mesh"].nnodes, 0).ravel() * scale


@pytest.mark.parametrize("state_var", [0, 1, 2])
def test_slip_wall_is_impermeable(state_var):
    case = make_case(shape=(2, 2, 2), degree=2, periodic=(False, False, False), perturb=0.15, seed=31, state_var=state_var)
    om, oc, _ = oracle_setup(case)
    f = _faces(case["mesh"], range(6))
    qds = O.setup_qdata_boundary(om, f)
    v, jac = O.slip(om, f, oc, qds, case["q"])
    v = v.reshape(-1, 5)
    scale = np.abs(v[:, 1:4]).max()
    assert scale > 1
    assert np.abs(v[:, 0]).max() < 1e-12 * scale and np.abs(v[:, 4]).max() < 1e-12 * scale
    assert np.allclose(jac.reshape(f.nfaces, 5, -1).shape, (f.nfaces, 5, om.tab.Q ** 2))


def test_slip_momentum_flux_is_normal_on_a_plane_wall():
    case = make_case(shape=(2, 3, 2), degree=3, periodic=(False, True, True), perturb=0.0, seed=33, state_var=1)
    om, oc, _ = oracle_setup(case)
    f = _faces(case["mesh"], [1])  # +x plane
    qds = O.setup_qdata_boundary(om, f)
    v, _ = O.slip(om, f, oc, qds, case["q"])
    v = v.reshape(-1, 5)
    assert np.abs(v[:, 1]).max() > 1
    assert np.abs(v[:, 2:4]).max() < 1e-12 * np.abs(v[:, 1]).max()


@pytest.mark.parametrize("implicit", [True, False])
def test_riemann_outflow_equals_boundary_integral_for_uniform_state(implicit):
    case = make_case(shape=(2, 2, 3), degree=2, periodic=(False, False, False), perturb=0.12, seed=35, state_var=1, implicit=implicit)
    om, oc, _ = oracle_setup(case)
    f = _faces(case["mesh"], range(6))
    qds = O.setup_qdata_boundary(om, f)
    Y = np.array([71.0, 3.0, -2.0, 1.5, 0.25])
    q = np.tile(Y, case["mesh"].nnodes)
    c = O.outflow_ctx(oc, Y[0], Y[4], recirc=1.0)
    v_out, jac_out = O.outflow(om, f, c, qds, q)
    v_bi, jac_bi = O.boundary_integral(om, f, oc, qds, q)
    assert np.abs(v_bi).max() > 1
    assert rel_err(v_out, v_bi)[1] < 1e-12
    if implicit:
        assert rel_err(jac_out, jac_bi)[0] < 1e-15


def test_pressure_outflow_equals_boundary_integral_at_the_outflow_pressure():
    case = make_case(shape=(2, 2, 2), degree=3, periodic=(False, True, False), perturb=0.12, seed=37, state_var=1)
    om, oc, _ = oracle_setup(case)
    f = _faces(case["mesh"], [0, 1, 4, 5])
    qds = O.setup_qdata_boundary(om, f)
    q = case["q"].reshape(-1, 5).copy()
    q[:, 0] = 70.5  # uniform pressure, varying velocity and temperature (non-zero gradients)
    q = q.ravel()
    c = O.outflow_ctx(oc, 70.5, 0.25, pressure_outflow=True)
    v_po, jac_po = O.outflow(om, f, c, qds, q)
    v_bi, jac_bi = O.boundary_integral(om, f, oc, qds, q)
    assert rel_err(v_po, v_bi)[1] < 1e-12
    assert rel_err(jac_po, jac_bi)[0] < 1e-15


def test_pressure_outflow_at_rest_scales_with_pressure():
    case = make_case(shape=(2, 2, 2), degree=2, periodic=(False, False, False), perturb=0.1, seed=39, state_var=1)
    om, oc, _ = oracle_setup(case)
    f = _faces(case["mesh"], range(6))
    qds = O.setup_qdata_boundary(om, f)
    q = np.tile(np.array([70.0, 0.0, 0.0, 0.0, 0.25]), case["mesh"].nnodes)
    v_bi, _ = O.boundary_integral(om, f, oc, qds, q)
    v_po, _ = O.outflow(om, f, O.outflow_ctx(oc, 84.0, 0.3, pressure_outflow=True), qds, q)
    assert np.abs(v_bi.reshape(-1, 5)[:, [0, 4]]).max() < 1e-12 * np.abs(v_bi).max()
    assert rel_err(v_po, 1.2 * v_bi)[1] < 1e-12


@pytest.mark.parametrize("pressure", [False, True])
def test_outflow_jacobian_matches_fd(pressure):
    case = make_case(shape=(2, 2, 2), degree=2, periodic=(False, True, True), perturb=0.1, seed=41, state_var=1)
    om, oc, _ = oracle_setup(case)
    f = _faces(case["mesh"], [0, 1])
    qds = O.setup_qdata_boundary(om, f)
    c = O.outflow_ctx(oc, 71.0, 0.26, recirc=1.0, pressure_outflow=pressure)
    q = case["q"]
    _, jac = O.outflow(om, f, c, qds, q)
    err = _fd(lambda x: O.outflow(om, f, c, qds, x)[0], lambda d: O.outflow_jacobian(om, f, c, qds, jac, d), q, _dq(case, q), 1e-3)
    assert err < 1e-7, err


@pytest.mark.parametrize("state_var", [0, 1, 2])
def test_slip_jacobian_matches_fd(state_var):
    case = make_case(shape=(2, 2, 2), degree=2, periodic=(False, True, True), perturb=0.1, seed=43, state_var=state_var)
    om, oc, _ = oracle_setup(case)
    f = _faces(case["mesh"], [0, 1])
    qds = O.setup_qdata_boundary(om, f)
    q = case["q"]
    _, jac = O.slip(om, f, oc, qds, q)
    err = _fd(lambda x: O.slip(om, f, oc, qds, x)[0], lambda d: O.slip_jacobian(om, f, oc, qds, jac, d), q, _dq(case, q), 1e-3)
    assert err < 1e-6, err


@pytest.mark.parametrize("kind", ["riemann", "pressure", "slip"])
def test_explicit_residual_is_negated(kind):
    case = make_case(shape=(2, 2, 2), degree=2, periodic=(True, True, False), seed=45, state_var=0)
    om, oc, _ = oracle_setup(case)
    f = _faces(case["mesh"], [4, 5])
    qds = O.setup_qdata_boundary(om, f)

    def run():
        if kind == "slip":
            return O.slip(om, f, oc, qds, case["q"])[0]
        return O.outflow(om, f, O.outflow_ctx(oc, 71.0, 0.25, recirc=0.5, pressure_outflow=kind == "pressure"), qds, case["q"])[0]

    v_impl = run()
    oc.is_implicit = 0
    v_expl = run()
    assert np.array_equal(v_expl, -v_impl)
