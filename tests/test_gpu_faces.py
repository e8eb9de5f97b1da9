"""Boundary-face operators on the HIP path (consistent boundary integral, freestream HLL/HLLC, Riemann and weak-pressure
outflow, slip wall) against the CPU oracle.

The IFunction and IJacobian composites (volume + faces) are applied through the C-ABI driver and compared with the
oracle's volume operator plus its face operators on the same seeded states.  Tolerances as for the volume parity
tests: relative 2-norm <= 1e-12 and relative max-norm <= 1e-11 in FP64; surface qdata and surface Jacobian data at
1e-12; FP32 Jacobian at 1e-5 relative to the FP64 oracle.
"""
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import make_case, oracle_setup, product_context, rel_err
from pyhonee import mesh as M
from pyhonee.driver import NewtonianProblem

pytestmark = pytest.mark.gpu

Y_INF = np.array([71.4, 0.9, 0.05, -0.02, 0.25])
# outflow {pressure, temperature, recirc, softplus_velocity}: recirc below 1 exercises the softplus regularization
OUTFLOW = np.array([71.0, 0.25, 0.5, 1e-2])
# recirc 1 for assembled diagonals: below 1 the reference's outflow Jacobian body is affine, not linear, in dq (its
# Softplus_fwd returns the derivative factor without the dx it is passed, bc_outflow.h:23-27), so diagonals assembled
# from per-point unit inputs and the operator's global unit columns legitimately differ
OUTFLOW_LINEAR = np.array([71.0, 0.25, 1.0, 1e-2])


def _bc_params(kind, linear=False):
    return {0: None, 1: Y_INF, 2: Y_INF, 3: OUTFLOW_LINEAR if linear else OUTFLOW, 4: OUTFLOW, 5: None}[kind]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _faces_setup(case, sides):
    fe, fi = M.box_boundary(case["mesh"], sides)
    return fe, fi, O.Faces(fe, fi)


def _oracle_boundary(om, oc, f, kind, q, linear=False):
    qds = O.setup_qdata_boundary(om, f)
    if kind == 0:
        v, jac = O.boundary_integral(om, f, oc, qds, q)
        return v, jac, qds, None
    if kind in (3, 4):
        c = O.outflow_ctx(oc, *_bc_params(kind, linear), pressure_outflow=kind == 4)
        v, jac = O.outflow(om, f, c, qds, q)
        return v, jac, qds, c
    if kind == 5:
        v, jac = O.slip(om, f, oc, qds, q)
        return v, jac, qds, None
    U_inf, _ = O.state_from_q(oc.gas, Y_INF, 1)
    fc = O.freestream_ctx(oc, U_inf, Y_INF, kind == 2)
    v, jac = O.freestream(om, f, fc, qds, q)
    return v, jac, qds, fc


def _oracle_boundary_jac(om, oc, f, kind, qds, jac, fc, dq):
    if kind == 0:
        return O.boundary_integral_jacobian(om, f, oc, qds, jac, dq)
    if kind in (3, 4):
        fc.newt.ijacobian_time_shift = oc.ijacobian_time_shift
        return O.outflow_jacobian(om, f, fc, qds, jac, dq)
    if kind == 5:
        return O.slip_jacobian(om, f, oc, qds, jac, dq)
    return O.freestream_jacobian(om, f, fc, qds, jac, dq)


@pytest.mark.parametrize("kind,state_var,resource", [
    (0, 1, "/gpu/hip/mi355x"),
    (0, 0, "/gpu/hip/mi355x:scatter=atomic"),
    (0, 2, "/gpu/hip/mi355x"),
    (1, 1, "/gpu/hip/mi355x"),
    (2, 1, "/gpu/hip/mi355x"),
    (2, 0, "/gpu/hip/mi355x:scatter=atomic"),
    (1, 2, "/gpu/hip/mi355x"),
    (3, 1, "/gpu/hip/mi355x"),
    (3, 0, "/gpu/hip/mi355x:scatter=atomic"),
    (3, 2, "/gpu/hip/mi355x"),
    (4, 1, "/gpu/hip/mi355x"),
    (4, 0, "/gpu/hip/mi355x"),
    (5, 1, "/gpu/hip/mi355x"),
    (5, 2, "/gpu/hip/mi355x:scatter=atomic"),
    (5, 0, "/gpu/hip/mi355x"),
])
def test_face_operators_match_oracle(kind, state_var, resource):
    _say(f"faces kind={kind} var={state_var} {resource}")
    sides = [0, 1, 4] if kind in (0, 3, 4, 5) else [0, 1]
    case = make_case(shape=(3, 2, 2), degree=3, periodic=(False, True, False), perturb=0.12, state_var=state_var, stabilization=2,
                     seed=100 + 10 * kind + state_var)
    om, oc, qd = oracle_setup(case)
    fe, fi, f = _faces_setup(case, sides)
    q, qdot = case["q"], case["qdot"]
    v_vol, jac_vol = O.ifunction(om, oc, qd, q, qdot)
    v_bnd, jac_bnd, qds, fc = _oracle_boundary(om, oc, f, kind, q)

    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), state_var, resource)
    b = prob.add_boundary(kind, fe, fi, _bc_params(kind))
    assert rel_err(prob.boundary_data(b, 0), qds)[0] < 1e-13
    g = np.zeros_like(q)
    prob.ifunction(q, qdot, g, time=0.0)
    e2, emax = rel_err(g, v_vol + v_bnd)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    assert rel_err(prob.boundary_data(b, 1), jac_bnd)[0] < 1e-12

    shift = 250.0
    oc.ijacobian_time_shift = shift
    dq = case["rng"].uniform(-1, 1, q.shape) * np.abs(q.reshape(-1, 5)).mean(0)[None, :].repeat(case["mesh"].nnodes, 0).ravel() * 1e-3
    vj_ref = O.ijacobian(om, oc, qd, jac_vol, dq) + _oracle_boundary_jac(om, oc, f, kind, qds, jac_bnd, fc, dq)
    vj = np.zeros_like(dq)
    prob.ijacobian(dq, vj, shift)
    e2, emax = rel_err(vj, vj_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    prob.close()


@pytest.mark.parametrize("kind", [0, 2, 3, 4, 5])
def test_face_operators_fp32_jacobian(kind):
    _say(f"faces fp32 jacobian kind={kind}")
    case = make_case(shape=(2, 2, 2), degree=3, periodic=(False, True, True), perturb=0.1, state_var=1, stabilization=2, seed=130 + kind)
    om, oc, qd = oracle_setup(case)
    fe, fi, f = _faces_setup(case, [0, 1])
    q, qdot = case["q"], case["qdot"]
    _, jac_vol = O.ifunction(om, oc, qd, q, qdot)
    _, jac_bnd, qds, fc = _oracle_boundary(om, oc, f, kind, q)
    shift = 250.0
    oc.ijacobian_time_shift = shift
    dq = case["rng"].uniform(-1e-3, 1e-3, q.shape)
    vj_ref = O.ijacobian(om, oc, qd, jac_vol, dq) + _oracle_boundary_jac(om, oc, f, kind, qds, jac_bnd, fc, dq)
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 1, "/gpu/hip/mi355x:jacobian_precision=fp32")
    prob.add_boundary(kind, fe, fi, _bc_params(kind))
    g = np.zeros_like(q)
    prob.ifunction(q, qdot, g)
    vj = np.zeros_like(dq)
    prob.ijacobian(dq, vj, shift)
    e2, _ = rel_err(vj, vj_ref)
    print(f"fp32 IJacobian with faces: rel2={e2:.3e}", flush=True)
    assert e2 < 1e-5
    prob.close()


def test_closed_box_uniform_state_balances():
    """All six sides with the boundary integral: a uniform state gives a zero residual at every node."""
    _say("closed box uniform state")
    case = make_case(shape=(2, 2, 2), degree=2, periodic=(False, False, False), perturb=0.1, state_var=1, stabilization=2, seed=141)
    fe, fi = M.box_boundary(case["mesh"], range(6))
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 1)
    prob.add_boundary(0, fe, fi)
    q = np.tile(np.array([71.0, 3.0, -2.0, 1.5, 0.25]), case["mesh"].nnodes)
    g = np.zeros_like(q)
    prob.ifunction(q, np.zeros_like(q), g)
    scale = 71.0 * 2 * np.pi  # pressure x face length scale of the residual
    assert np.abs(g).max() < 1e-10 * scale, np.abs(g).max()
    prob.close()


@pytest.mark.parametrize("kind,degree", [(0, 1), (1, 2), (2, 1), (3, 2), (4, 1), (5, 2)])
def test_jacobian_diagonal_with_faces(kind, degree):
    """Point-Jacobi diagonal of the IJacobian composite (volume + faces) against the oracle's unit columns."""
    _say(f"diagonal with faces kind={kind} degree={degree}")
    case = make_case(shape=(2, 2, 2), degree=degree, periodic=(False, True, True), perturb=0.1, state_var=1, stabilization=2, seed=150 + kind)
    om, oc, qd = oracle_setup(case)
    fe, fi, f = _faces_setup(case, [0, 1])
    q, qdot = case["q"], case["qdot"]
    _, jac_vol = O.ifunction(om, oc, qd, q, qdot)
    _, jac_bnd, qds, fc = _oracle_boundary(om, oc, f, kind, q, linear=True)
    shift = 120.0
    oc.ijacobian_time_shift = shift
    d_ref = O.ijacobian_diagonal(om, oc, qd, jac_vol)
    n = q.size
    for i in range(n):  # face part by unit columns
        e = np.zeros(n)
        e[i] = 1.0
        d_ref[i] += _oracle_boundary_jac(om, oc, f, kind, qds, jac_bnd, fc, e)[i]
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 1)
    prob.add_boundary(kind, fe, fi, _bc_params(kind, linear=True))
    g = np.zeros_like(q)
    prob.ifunction(q, qdot, g)
    d = np.zeros_like(q)
    prob.ijacobian_diagonal(d, shift)
    e2, emax = rel_err(d, d_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    prob.close()


@pytest.mark.parametrize("kind,state_var,fp32", [(0, 1, False), (2, 2, False), (1, 0, True), (3, 1, False), (4, 2, False), (5, 0, False),
                                                 (3, 1, True)])
def test_point_block_diagonal_with_faces(kind, state_var, fp32):
    """Point-block diagonal [node][c_out][c_in] of the IJacobian composite (volume + faces) against the oracle's
    volume point blocks plus the face operator's unit columns; FP32 Jacobian mode at 1e-5."""
    _say(f"point-block diagonal with faces kind={kind} var={state_var} fp32={fp32}")
    case = make_case(shape=(2, 2, 2), degree=2, periodic=(False, True, True), perturb=0.1, state_var=state_var, stabilization=2,
                     seed=170 + kind)
    om, oc, qd = oracle_setup(case)
    fe, fi, f = _faces_setup(case, [0, 1])
    q, qdot = case["q"], case["qdot"]
    _, jac_vol = O.ifunction(om, oc, qd, q, qdot)
    _, jac_bnd, qds, fc = _oracle_boundary(om, oc, f, kind, q, linear=True)
    shift = 80.0
    oc.ijacobian_time_shift = shift
    pb_ref = O.ijacobian_point_block_diagonal(om, oc, qd, jac_vol)
    nn = om.nnodes
    for node in range(nn):  # face part by unit columns
        for ci in range(5):
            e = np.zeros(5 * nn)
            e[node * 5 + ci] = 1.0
            pb_ref[node, :, ci] += _oracle_boundary_jac(om, oc, f, kind, qds, jac_bnd, fc, e)[node * 5: node * 5 + 5]
    res = "/gpu/hip/mi355x:jacobian_precision=fp32" if fp32 else "/gpu/hip/mi355x"
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), state_var, res)
    prob.add_boundary(kind, fe, fi, _bc_params(kind, linear=True))
    g = np.zeros_like(q)
    prob.ifunction(q, qdot, g)
    pb = np.zeros((nn, 5, 5))
    prob.ijacobian_point_block_diagonal(pb, shift)
    e2, emax = rel_err(pb, pb_ref)
    print(f"point-block rel2 {e2:.3e} max {emax:.3e}", flush=True)
    if fp32:
        assert e2 < 1e-5
    else:
        assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    # the diagonal of each block is the point diagonal; in FP32 mode the two kernels evaluate the FP32 actions with
    # different instruction sequences, and each is within 1e-5 of the FP64 oracle
    d = np.zeros_like(q)
    prob.ijacobian_diagonal(d, shift)
    assert rel_err(np.einsum("nkk->nk", pb).ravel(), d)[0] < (2e-5 if fp32 else 1e-13)
    prob.close()
