"""The reference's implicit fixtures tests/output/gaussianwave-IDL.bin (primitive) and gaussianwave-IDL-entropy.bin
(entropy) reproduced with the HIP path's operators, compared in order: the composite IFunction (volume + HLLC freestream
faces, essential rows) and IJacobian of libceed.so, called through the C-ABI, drive the PETSc solver restatement of
petsc_ts.py (the runs are described in gaussianwave_case.py; tests/test_oracle_gaussianwave_implicit.py reproduces
the same fixtures with the oracle).  Newton rtol 1e-4 and GMRES rtol 1e-2 make the final state depend on every
residual and Jacobian evaluation, so the comparison at the reference's 2e-11 pins the HIP primitive- and
entropy-variable IFunction and IJacobian with the damping layer and freestream faces at p=1."""
import sys

import numpy as np
import pytest

import gaussianwave_case as GW
import petsc_ts as PA
from pyhonee.driver import NewtonianProblem, context_from_dict

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("state_var", [1, 2])
def test_gaussianwave_idl_reproduces_reference_fixture(state_var):
    _say(f"gaussian wave IDL implicit p=1, state_var {state_var}: reference fixture")
    mesh, tab, d, (fe, fi), ess, q0, free, pattern = GW.setup_implicit(state_var)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), state_var)
    prob.add_boundary(prob.FREESTREAM_HLLC, fe, fi, GW.REF_Y)
    prob.set_essential(ess, q0[ess])
    g = np.zeros_like(q0)
    v = np.zeros_like(q0)

    def residual(t, dt, X, V):
        prob.ifunction(np.ascontiguousarray(X), np.ascontiguousarray(V), g, time=t, dt=dt)
        return g.copy()

    def jacobian(shift, dq):
        prob.ijacobian(np.ascontiguousarray(dq), v, shift)
        return v.copy()

    fixed = np.zeros_like(q0)
    fixed[ess] = q0[ess]
    newton = PA.Newton(**GW.IMPLICIT_SOLVER)
    x = PA.ts_alpha(PA.Operators(residual, jacobian, free, fixed, pattern), q0, GW.DT, GW.STEPS, GW.IMPLICIT[state_var][1], newton)
    prob.close()
    err = GW.regression_metric(x, GW.read_implicit(state_var))
    _say(f"{GW.IMPLICIT[state_var][0]} regression metric {err:.3e}; Newton {newton.newton_its}, GMRES {newton.linear_its}, "
         f"Jacobian assemblies {newton.assemblies}")
    assert err < GW.IMPLICIT_TOLERANCE


def test_gaussianwave_shell_reproduces_reference_fixture():
    """tests/output/gaussianwave-shell.bin (p=3, primitive, matrix-free Jacobian with point-block Jacobi): the HIP
    IJacobian apply is the operator and the HIP point-block diagonal assembly of the composite (volume + freestream
    faces) the preconditioner, both at the lagged Jacobian evaluations."""
    _say("gaussian wave shell p=3 primitive: reference fixture")
    mesh, tab, d, (fe, fi), ess, q0, free = GW.setup_shell()
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1)
    prob.add_boundary(prob.FREESTREAM_HLLC, fe, fi, GW.REF_Y)
    prob.set_essential(ess, q0[ess])
    g = np.zeros_like(q0)
    v = np.zeros_like(q0)
    pb = np.zeros(25 * mesh.nnodes)

    def residual(t, dt, X, V):
        prob.ifunction(np.ascontiguousarray(X), np.ascontiguousarray(V), g, time=t, dt=dt)
        return g.copy()

    def jacobian(shift, dq):
        prob.ijacobian(np.ascontiguousarray(dq), v, shift)
        return v.copy()

    def point_blocks(shift):
        prob.ijacobian_point_block_diagonal(pb, shift)
        return pb.reshape(-1, 5, 5).copy()

    fixed = np.zeros_like(q0)
    fixed[ess] = q0[ess]
    ops = PA.Operators(residual, jacobian, free, fixed, point_blocks=point_blocks)
    newton = PA.Newton(**GW.IMPLICIT_SOLVER)
    x = PA.ts_alpha(ops, q0, GW.DT, GW.STEPS, GW.SHELL_RADIUS, newton)
    prob.close()
    err, err0 = GW.compare_shell(mesh, ops.to_local(x, fixed), q0)
    _say(f"gaussianwave-shell.bin regression metric {err:.3e} (initial state {err0:.1e}); Newton {newton.newton_its}, "
         f"GMRES {newton.linear_its}")
    assert err0 > 1e-3
    assert err < GW.TOLERANCE
