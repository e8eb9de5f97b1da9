"""Quadrature with more points than nodes: Q_1d = P_1d + q_extra (HONEE's -q_extra, src/cloptions.c:152-153; e.g.
examples/channel.yaml uses q_extra 2), on the HIP path against the oracle at the volume parity tolerances (relative
2-norm 1e-12, max-norm 1e-11; FP32 Jacobian 1e-5).  Every q_extra 1, 2 at degrees 1-3 for the IFunction, the IJacobian
and its Jacobi diagonal; the point-block diagonal, the FP32 Jacobian, the explicit RHS and the boundary faces
(consistent boundary integral, freestream HLLC) at some of them.  The oracle's quadrature is pinned independently:
the quadrature weights times det J integrate the volume exactly at every Q_1d
(test_oracle_unit.py::test_q_extra_quadrature_integrates_the_volume)."""
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import rel_err
from pyhonee import mesh as M
from pyhonee import problems as PR
from pyhonee.driver import NewtonianProblem, context_from_dict

pytestmark = pytest.mark.gpu

FP32 = "/gpu/hip/mi355x:jacobian_precision=fp32"


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _check(x, ref, fp32=False, what=""):
    e2, emax = rel_err(x, ref)
    _say(f"{what} rel2 {e2:.3e} max {emax:.3e}")
    if fp32:
        assert e2 < 1e-5, (e2, emax)
    else:
        assert e2 < 1e-12 and emax < 1e-11, (e2, emax)


def _case(degree, q_extra, state_var=1, stabilization=2, periodic=(True, True, True), implicit=True, seed=0, shape=(2, 2, 2)):
    mesh = M.box_mesh(shape, degree, periodic=periodic, perturb=0.1, seed=seed)
    tab = M.tabulate(degree, q_extra)
    d = PR.tgv_context_dict(state_var=state_var, stabilization=stabilization, implicit=implicit, dt=1e-3, tau=PR.default_tau(degree))
    rng = np.random.default_rng(seed)
    Y = PR.nodal_field(mesh, lambda x: PR.taylor_green_prim(x)) * (1 + 0.05 * rng.uniform(-1, 1, (mesh.nnodes, 5)))
    q = PR.q_from_prim(PR.TGV_GAS, Y, state_var).ravel()
    qdot = (rng.uniform(-1, 1, Y.shape) * np.abs(Y).mean(0) * 0.1).ravel()
    om = O.Mesh(mesh, tab)
    return mesh, tab, d, q, qdot, rng, om, O.newt_from_dict(d), om.setup_qdata()


@pytest.mark.parametrize("degree,q_extra", [(1, 1), (1, 2), (2, 1), (2, 2), (3, 1), (3, 2)])
def test_volume_operators_q_extra(degree, q_extra):
    _say(f"volume operators degree={degree} q_extra={q_extra}")
    mesh, tab, d, q, qdot, rng, om, oc, qd = _case(degree, q_extra, seed=500 + 10 * degree + q_extra)
    assert (tab.P, tab.Q) == (degree + 1, degree + 1 + q_extra)
    v_ref, jac = O.ifunction(om, oc, qd, q, qdot)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1)
    _check(prob.qdata(), qd, what="qdata")
    g = np.zeros_like(q)
    prob.ifunction(q, qdot, g)
    _check(g, v_ref, what="ifunction")
    _check(prob.jac_data(), jac, what="jac_data")
    shift = 300.0
    oc.ijacobian_time_shift = shift
    dq = rng.uniform(-1e-3, 1e-3, q.shape) * np.abs(q.reshape(-1, 5)).mean(0)[None, :].repeat(mesh.nnodes, 0).ravel()
    v = np.zeros_like(q)
    prob.ijacobian(dq, v, shift)
    _check(v, O.ijacobian(om, oc, qd, jac, dq), what="ijacobian")
    dg = np.zeros_like(q)
    prob.ijacobian_diagonal(dg, shift)
    _check(dg, O.ijacobian_diagonal(om, oc, qd, jac), what="jacobi diagonal")
    if (degree, q_extra) in ((1, 2), (2, 1), (3, 1)):
        pb = np.zeros((mesh.nnodes, 5, 5))
        prob.ijacobian_point_block_diagonal(pb, shift)
        _check(pb, O.ijacobian_point_block_diagonal(om, oc, qd, jac), what="point block")
    prob.close()


def test_fp32_jacobian_q_extra():
    _say("fp32 jacobian degree=3 q_extra=1")
    mesh, tab, d, q, qdot, rng, om, oc, qd = _case(3, 1, seed=531)
    _, jac = O.ifunction(om, oc, qd, q, qdot)
    shift = 300.0
    oc.ijacobian_time_shift = shift
    dq = rng.uniform(-1e-3, 1e-3, q.shape)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1, FP32)
    prob.ifunction(q, qdot, np.zeros_like(q))
    v = np.zeros_like(q)
    prob.ijacobian(dq, v, shift)
    _check(v, O.ijacobian(om, oc, qd, jac, dq), fp32=True, what="fp32 ijacobian")
    prob.close()


def test_explicit_rhs_q_extra():
    """The channel example's discretization (degree 1, q_extra 2) for the explicit conservative RHS.  On a 3 x 3 x 3 mesh:
    the degree-1 nodes of a 2 x 2 x 2 periodic box sit where the Taylor-Green velocity vanishes, and the RHS there is
    ill-conditioned (a one-ulp relative change of the state moves it by 2e-12 relative); on 3 x 3 x 3 the same change
    moves it by 3e-15."""
    _say("explicit rhs degree=1 q_extra=2")
    mesh, tab, d, q, qdot, rng, om, oc, qd = _case(1, 2, state_var=0, stabilization=0, implicit=False, seed=541, shape=(3, 3, 3))
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 0)
    r = np.zeros_like(q)
    prob.rhs(q, r)
    _check(r, O.rhs(om, oc, qd, q), what="rhs")
    prob.close()


@pytest.mark.parametrize("kind", [0, 2])
def test_faces_q_extra(kind):
    """Boundary faces with Q_1d = P_1d + 2 at degree 2: the consistent boundary integral (kind 0) and the freestream
    HLLC flux (kind 2), volume and faces of the IFunction and the IJacobian together."""
    _say(f"faces degree=2 q_extra=2 kind={kind}")
    mesh, tab, d, q, qdot, rng, om, oc, qd = _case(2, 2, periodic=(False, True, True), seed=550 + kind)
    fe, fi = M.box_boundary(mesh, [0, 1])
    f = O.Faces(fe, fi)
    qds = O.setup_qdata_boundary(om, f)
    v_vol, jac_vol = O.ifunction(om, oc, qd, q, qdot)
    Y_inf = np.array([PR.TGV_REFERENCE_Y[0], 1.0, 0.1, 0.0, PR.TGV_REFERENCE_Y[4]])
    if kind == 0:
        v_f, jac_f = O.boundary_integral(om, f, oc, qds, q)
    else:
        U_inf, _ = O.state_from_q(oc.gas, Y_inf, 1)
        fc = O.freestream_ctx(oc, U_inf, Y_inf, True)
        v_f, jac_f = O.freestream(om, f, fc, qds, q)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1)
    prob.add_boundary(kind, fe, fi, None if kind == 0 else Y_inf)
    g = np.zeros_like(q)
    prob.ifunction(q, qdot, g)
    _check(g, v_vol + v_f, what="ifunction with faces")
    shift = 200.0
    oc.ijacobian_time_shift = shift
    dq = rng.uniform(-1e-3, 1e-3, q.shape)
    vj_f = O.boundary_integral_jacobian(om, f, oc, qds, jac_f, dq) if kind == 0 else O.freestream_jacobian(om, f, fc, qds, jac_f, dq)
    v = np.zeros_like(q)
    prob.ijacobian(dq, v, shift)
    _check(v, O.ijacobian(om, oc, qd, jac_vol, dq) + vj_f, what="ijacobian with faces")
    prob.close()
