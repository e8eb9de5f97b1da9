"""The reference's implicit fixtures tests/output/gaussianwave-IDL.bin (primitive variables) and
gaussianwave-IDL-entropy.bin (entropy variables) reproduced by the oracle alone (no GPU), compared in order.

The runs are described in gaussianwave_case.py; the PETSc solver stack they use (TSALPHA with its restart, Newton with
the cubic line search and a persistent lagged Jacobian, GMRES with ILU(0) on the assembled Jacobian) is restated in
petsc_ts.py.  With Newton rtol 1e-4 and GMRES rtol 1e-2 the final state depends on every iterate, so matching it to
the reference's 2e-11 pins:
* the oracle's implicit IFunction (SUPG, damping layer) and IJacobian in primitive and entropy variables, the HLLC
  freestream faces and their Jacobian, and the essential rows;
* the solver restatement;
* the p=1 DMPlex vertex order (x fastest), since the vectors are compared entry by entry.
"""
import numpy as np

import gaussianwave_case as GW
import petsc_ts as PA
import pyoracle as O
from pyhonee import mesh as M


def oracle_operators(state_var, jacobian_shift_scale=1.0):
    mesh, tab, d, (fe, fi), ess, q0, free, pattern = GW.setup_implicit(state_var)
    om = O.Mesh(mesh, tab)
    oc = O.newt_from_dict(d)
    qd = om.setup_qdata()
    f = O.Faces(fe, fi)
    qds = O.setup_qdata_boundary(om, f)
    U_inf, Y_inf = O.state_from_q(oc.gas, GW.REF_Y, 1)
    fc = O.freestream_ctx(oc, U_inf, Y_inf, True)
    kept = {}

    def residual(t, dt, X, V):
        oc.dt = dt
        v, kept["jac"] = O.ifunction(om, oc, qd, X, V)
        w, kept["jac_sur"] = O.freestream(om, f, fc, qds, X)
        return v + w

    def jacobian(shift, dq):
        oc.ijacobian_time_shift = shift * jacobian_shift_scale
        return O.ijacobian(om, oc, qd, kept["jac"], dq) + O.freestream_jacobian(om, f, fc, qds, kept["jac_sur"], dq)

    fixed = np.zeros_like(q0)
    fixed[ess] = q0[ess]
    return PA.Operators(residual, jacobian, free, fixed, pattern), q0


def _run(state_var, radius=None, jacobian_shift_scale=1.0):
    ops, q0 = oracle_operators(state_var, jacobian_shift_scale)
    newton = PA.Newton(**GW.IMPLICIT_SOLVER)
    x = PA.ts_alpha(ops, q0, GW.DT, GW.STEPS, GW.IMPLICIT[state_var][1] if radius is None else radius, newton)
    return x, newton


def test_oracle_reproduces_gaussianwave_idl_fixtures():
    for state_var in (1, 2):
        ref = GW.read_implicit(state_var)
        x, newton = _run(state_var)
        err = GW.regression_metric(x, ref)
        print(f"state_var {state_var}: metric {err:.2e}; Newton {newton.newton_its}, GMRES {newton.linear_its}, "
              f"assemblies {newton.assemblies}")
        assert newton.assemblies == 2  # iterations 0 and 20 of the persistent lag
        _, _, _, _, _, q0, free, _ = GW.setup_implicit(state_var)
        assert GW.regression_metric(q0[free], ref) > 1e-3  # the five steps moved the state far beyond the tolerance
        assert err < GW.IMPLICIT_TOLERANCE, err  # measured 4.7e-16 (primitive) and 9.7e-16 (entropy)


def test_fixtures_reject_plausible_mistakes():
    """The pins bite: the other variable set's generalized-alpha radius (2.4e-4, 1.7e-5), an IJacobian whose shift is
    10% off (3.3e-6), and the y-fastest node order (0.21) all fail the 2e-11 tolerance by orders of magnitude."""
    tol = GW.IMPLICIT_TOLERANCE
    ref1 = GW.read_implicit(1)
    assert GW.regression_metric(_run(1, radius=1.0)[0], ref1) > 1e4 * tol
    assert GW.regression_metric(_run(2, radius=0.5)[0], GW.read_implicit(2)) > 1e4 * tol
    assert GW.regression_metric(_run(1, jacobian_shift_scale=0.9)[0], ref1) > 1e4 * tol
    x = _run(1)[0]
    assert GW.regression_metric(x.reshape(2, 6, 6, 4).transpose(0, 2, 1, 3).ravel(), ref1) > 1e-2


def test_oracle_reproduces_gaussianwave_shell_fixture():
    """tests/output/gaussianwave-shell.bin: the p=3 implicit primitive run with the matrix-free Jacobian and point-block
    Jacobi (gaussianwave_case.py).  The preconditioner's node blocks are the oracle's volume point blocks plus the
    freestream faces' blocks taken column by column from the face Jacobian.  Measured 3.3e-9 against the test's 1e-8,
    the level of the p=3 explicit fixture."""
    mesh, tab, d, (fe, fi), ess, q0, free = GW.setup_shell()
    om = O.Mesh(mesh, tab)
    oc = O.newt_from_dict(d)
    qd = om.setup_qdata()
    f = O.Faces(fe, fi)
    qds = O.setup_qdata_boundary(om, f)
    U_inf, Y_inf = O.state_from_q(oc.gas, GW.REF_Y, 1)
    fc = O.freestream_ctx(oc, U_inf, Y_inf, True)
    kept = {}

    def residual(t, dt, X, V):
        oc.dt = dt
        v, kept["jac"] = O.ifunction(om, oc, qd, X, V)
        w, kept["jac_sur"] = O.freestream(om, f, fc, qds, X)
        return v + w

    def jacobian(shift, dq):
        oc.ijacobian_time_shift = shift
        return O.ijacobian(om, oc, qd, kept["jac"], dq) + O.freestream_jacobian(om, f, fc, qds, kept["jac_sur"], dq)

    face_nodes = np.unique(M.face_node_tables(mesh, fe, fi)[0])

    def point_blocks(shift):
        oc.ijacobian_time_shift = shift
        pb = O.ijacobian_point_block_diagonal(om, oc, qd, kept["jac"])
        e = np.zeros_like(q0)
        for n in face_nodes:
            for j in range(5):
                e[5 * n + j] = 1.0
                pb[n, :, j] += O.freestream_jacobian(om, f, fc, qds, kept["jac_sur"], e)[5 * n:5 * n + 5]
                e[5 * n + j] = 0.0
        return pb

    fixed = np.zeros_like(q0)
    fixed[ess] = q0[ess]
    ops = PA.Operators(residual, jacobian, free, fixed, point_blocks=point_blocks)
    newton = PA.Newton(**GW.IMPLICIT_SOLVER)
    x = PA.ts_alpha(ops, q0, GW.DT, GW.STEPS, GW.SHELL_RADIUS, newton)
    err, err0 = GW.compare_shell(mesh, ops.to_local(x, fixed), q0)
    print(f"gaussianwave-shell.bin: metric {err:.2e} (initial state {err0:.1e}); Newton {newton.newton_its}, GMRES {newton.linear_its}")
    assert err0 > 1e-3
    assert err < GW.TOLERANCE
