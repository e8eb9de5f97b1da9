"""The reference's implicit channel fixtures tests/output/channel.bin (conservative) and channel-prim.bin (primitive)
reproduced with the HIP path's operators (libceed.so through the C-ABI: p=1 with q_extra 2, viscous IFunction and
IJacobian without stabilization, body force, essential wall and symmetry rows) driving the PETSc TSBEULER restatement
of petsc_ts.py, compared in order (the runs are described in channel_case.py; tests/test_oracle_channel_golden.py
reproduces the same fixtures with the oracle)."""
import sys

import numpy as np
import pytest

import channel_case as CH
import petsc_ts as PA
from pyhonee.driver import NewtonianProblem, context_from_dict

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("state_var", [0, 1])
def test_channel_reproduces_reference_fixture(state_var):
    _say(f"channel implicit p=1 q_extra 2, state_var {state_var}: reference fixture")
    mesh, tab, d, ess, q0, free, pattern = CH.setup(state_var)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), state_var)
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
    newton = PA.Newton(**CH.SOLVER)
    x = PA.ts_beuler(PA.Operators(residual, jacobian, free, fixed, pattern), q0, CH.DT, CH.STEPS, newton)
    prob.close()
    ref = CH.read(state_var)
    err = float(np.abs(x - ref).max() / np.abs(ref).max())
    _say(f"{CH.FIXTURES[state_var]} regression metric {err:.3e}; Newton {newton.newton_its}, GMRES {newton.linear_its}")
    assert err < CH.TOLERANCE
    assert err < 1e-13
