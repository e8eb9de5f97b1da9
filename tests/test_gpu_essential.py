"""Essential (strong) boundary conditions on the HIP path: no-slip walls (velocity components on the y sides of a
channel-like box, periodic in x and z) with HONEE's DMPlex semantics (SetBCsFromICs and
DMPlexInsertBoundaryValues_FromICs, src/misc.c:51-95; IFunction_NS, src/setupts.c:77-119): the operator inputs are the
local vectors q .* mask + Qbc and q_dot .* mask, the constrained residual rows are dropped, and the Jacobian is the
identity on the constrained entries.  Compared with the oracle's volume and boundary-integral operators on those local
vectors, at the volume parity tolerances (relative 2-norm 1e-12, max-norm 1e-11).  A BDF2 run keeps the wall
velocity at its boundary value exactly."""
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import make_case, oracle_setup, product_context, rel_err
from pyhonee import mesh as M
from pyhonee import ts as TS
from pyhonee.driver import NewtonianProblem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _channel(state_var, degree=2, seed=201):
    case = make_case(shape=(2, 3, 2), degree=degree, periodic=(True, False, True), perturb=0.1, state_var=state_var, stabilization=2, seed=seed)
    mesh = case["mesh"]
    dofs = M.essential_dofs(mesh, [2, 3], [1, 2, 3])
    N = 5 * mesh.nnodes
    mask = np.ones(N)
    mask[dofs] = 0.0
    return case, dofs, mask


@pytest.mark.parametrize("state_var", [1, 0])
def test_essential_operators_match_oracle(state_var):
    _say(f"essential operators var={state_var}")
    case, dofs, mask = _channel(state_var)
    om, oc, qd = oracle_setup(case)
    fe, fi = M.box_boundary(case["mesh"], [2, 3])
    f = O.Faces(fe, fi)
    qds = O.setup_qdata_boundary(om, f)
    q, qdot = case["q"], case["qdot"]
    q_loc, qdot_loc = q * mask, qdot * mask  # boundary values zero (no-slip)
    v_vol, jac_vol = O.ifunction(om, oc, qd, q_loc, qdot_loc)
    v_bnd, jac_bnd = O.boundary_integral(om, f, oc, qds, q_loc)
    g_ref = mask * (v_vol + v_bnd)

    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), state_var)
    prob.add_boundary(0, fe, fi)
    prob.set_essential(dofs)
    g = np.zeros_like(q)
    prob.ifunction(q, qdot, g)
    e2, emax = rel_err(g, g_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    assert np.all(g[dofs] == 0.0)

    shift = 150.0
    oc.ijacobian_time_shift = shift
    dq = case["rng"].uniform(-1, 1, q.shape) * np.abs(q.reshape(-1, 5)).mean(0)[None, :].repeat(case["mesh"].nnodes, 0).ravel() * 1e-3
    vj_ref = mask * (O.ijacobian(om, oc, qd, jac_vol, dq * mask) + O.boundary_integral_jacobian(om, f, oc, qds, jac_bnd, dq * mask)) + (1 - mask) * dq
    vj = np.zeros_like(dq)
    prob.ijacobian(dq, vj, shift)
    e2, emax = rel_err(vj, vj_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)

    d_ref = O.ijacobian_diagonal(om, oc, qd, jac_vol)
    n = q.size
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        d_ref[i] += O.boundary_integral_jacobian(om, f, oc, qds, jac_bnd, e)[i]
    d_ref = mask * d_ref + (1 - mask)
    d = np.zeros_like(q)
    prob.ijacobian_diagonal(d, shift)
    e2, emax = rel_err(d, d_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)

    pb = np.zeros((om.nnodes, 5, 5))
    prob.ijacobian_point_block_diagonal(pb, shift)
    m5 = mask.reshape(-1, 5)
    assert np.array_equal(np.einsum("nkk->nk", pb).ravel(), d) or rel_err(np.einsum("nkk->nk", pb).ravel(), d)[0] < 1e-13
    assert np.all(pb[:, 1:4, :][(m5[:, 1:4] == 0)] == np.eye(5)[1:4][None].repeat(om.nnodes, 0)[(m5[:, 1:4] == 0)])
    prob.close()


def test_essential_bdf2_keeps_wall_values():
    _say("essential BDF2 channel")
    case, dofs, mask = _channel(1, degree=2, seed=211)
    fe, fi = M.box_boundary(case["mesh"], [2, 3])
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 1)
    prob.add_boundary(0, fe, fi)
    prob.set_essential(dofs)
    q0 = case["q"] * mask  # no-slip initial condition
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_BDF, bdf_order=2, dt=2e-3, ksp_pc=TS.PC_PBJACOBI))
    ts.set_solution(q0, 0.0)
    ts.step(4)
    q = ts.solution()
    st = ts.stats()
    _say(f"newton {st['newton_its']} gmres {st['krylov_its']} fnorm {st['last_fnorm']:.3e}")
    assert np.all(np.isfinite(q))
    assert np.all(q[dofs] == 0.0)
    assert np.abs(q - q0).max() > 0  # the flow evolved
    # the final state solves the constrained BDF2 system: Newton stops on rtol 1e-8 or on PETSc's step test
    # (||dx|| <= 1e-8 ||x||), which leaves the masked residual below 1e-6 of its initial value
    assert st["last_fnorm"] <= 1e-6 * st["last_fnorm0"]
    ts.close()
    prob.close()
