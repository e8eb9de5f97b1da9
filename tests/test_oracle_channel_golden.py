"""The reference's implicit channel fixtures tests/output/channel.bin (conservative) and channel-prim.bin (primitive)
reproduced by the oracle alone (no GPU), compared in order (the runs are described in channel_case.py; the PETSc
TSBEULER / Newton / GMRES / ILU(0) stack is restated in petsc_ts.py).  They pin the oracle's viscous implicit IFunction
and IJacobian without stabilization in conservative and primitive variables, the channel's body force, the essential
wall and symmetry rows and the DMPlex vertex order of a box periodic in x."""
import numpy as np

import channel_case as CH
import petsc_ts as PA
import pyoracle as O


def oracle_operators(state_var, q_extra=2, body_force=True):
    mesh, tab, d, ess, q0, free, pattern = CH.setup(state_var, q_extra)
    if not body_force:
        d["g"] = (0.0, 0.0, 0.0)
    om = O.Mesh(mesh, tab)
    oc = O.newt_from_dict(d)
    qd = om.setup_qdata()
    kept = {}

    def residual(t, dt, X, V):
        oc.dt = dt
        v, kept["jac"] = O.ifunction(om, oc, qd, X, V)
        return v

    def jacobian(shift, dq):
        oc.ijacobian_time_shift = shift
        return O.ijacobian(om, oc, qd, kept["jac"], dq)

    fixed = np.zeros_like(q0)
    fixed[ess] = q0[ess]
    return PA.Operators(residual, jacobian, free, fixed, pattern), q0, free


def _metric(state_var, **kw):
    ops, q0, free = oracle_operators(state_var, **kw)
    newton = PA.Newton(**CH.SOLVER)
    x = PA.ts_beuler(ops, q0, CH.DT, CH.STEPS, newton)
    ref = CH.read(state_var)
    return float(np.abs(x - ref).max() / np.abs(ref).max()), float(np.abs(q0[free] - ref).max() / np.abs(ref).max()), newton


def test_oracle_reproduces_channel_fixtures():
    for state_var in (0, 1):
        err, err0, newton = _metric(state_var)
        print(f"{CH.FIXTURES[state_var]}: metric {err:.2e} (initial state {err0:.1e}); Newton {newton.newton_its}, "
              f"GMRES {newton.linear_its}")
        assert err0 > 1e-5
        assert err < CH.TOLERANCE
        # measured 1.1e-14 for both; the tighter bound also separates q_extra 2 from q_extra 0 (1.5e-12, 1.2e-11)
        assert err < 1e-13


def test_fixtures_reject_plausible_mistakes():
    """Without the body force the runs miss by 1.2e-4 (conservative) and 8.2e-6 (primitive); with q_extra 0 by
    1.5e-12 and 1.2e-11 (inside the reference's 2e-11, outside the 1e-13 this build reaches).  q_extra 1 agrees with
    q_extra 2 to 1e-14: the fixture does not separate 3 from 4 quadrature points per direction."""
    for state_var in (0, 1):
        assert _metric(state_var, body_force=False)[0] > 1e5 * CH.TOLERANCE
        assert _metric(state_var, q_extra=0)[0] > 1e-12
