"""The reference run behind tests/output/gaussianwave-explicit.bin (examples/navierstokes.c:233, "Gaussian Wave,
explicit, supg, IDL"), shared by the HIP-path test and the oracle test: examples/gaussianwave.yaml on a 2 x 2 x 1 box
[0,1] x [0,1] x [0,0.025], degree 3, 5 explicit steps of dt 2e-3 with TSRK5F (src/setupts.c:354-355; adapt none),
conservative variables, SUPG with the default Tau coefficients of NS_NEWTONIAN_IG at degree 3
(problems/newtonian.c:588-610), mu 0, the internal damping layer (decay time 2e-3, start 0, length 0.25, pressure 70),
freestream faces with the default HLLC flux on the four lateral sides (face sets 3-6) against the reference state
(pressure 71.75, velocity (2, 2, 0), temperature 0.25), essential z-momentum on the two z sides (face sets 1-2), and
the consistent mass solved by GMRES with diagonal Jacobi (KSP defaults: rtol 1e-5, restart 30).

The initial condition restates IC_GaussianWave (qfunctions/gaussianwave.h:20-53) here, in the tests only.

The fixture is a PETSc global vector: DMPlex point order, each node's components together, and the constrained
z-momentum of the 98 nodes on the z sides left out (882 = 196 x 5 - 98).  Since the constrained component sits between
the y-momentum (about 2) and the energy (about 180) and the free z-momentum is of order 1e-9, the records split
unambiguously into nodes of 5 and of 4 values; they are matched to this build's nodes by a minimum-cost assignment
(order-invariant), constrained records to the z-side nodes only.  The pass criterion is HONEE's RegressionTest metric
(src/misc.c:98-126) max|Q - Qref| / max|Qref| against the test's tolerance 1e-8."""
import os

import numpy as np
from scipy.optimize import linear_sum_assignment

from pyhonee import honee_file
from pyhonee import mesh as M
from pyhonee import problems as PR

FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "gaussianwave-explicit.bin")
GAS = {"cv": 717.0, "cp": 1004.0, "lambda": -2.0 / 3.0, "mu": 0.0, "k": 0.02638}  # NS_NEWTONIAN_IG defaults, mu from the YAML
REF_Y = np.array([71.75, 2.0, 2.0, 0.0, 0.25])  # reference pressure, velocity, temperature (examples/gaussianwave.yaml)
EPICENTER, AMPLITUDE, WIDTH = (0.33, 0.75), 2.0, 0.05
DT, STEPS = 2e-3, 5
TOLERANCE = 1e-8  # -compare_final_state_atol (examples/navierstokes.c:233)


def gaussian_wave_ic(xyz):
    """IC_GaussianWave in conservative variables: density and total energy scaled by the Gaussian perturbation, the
    momentum following the density at the reference velocity, and the kinetic energy of the unperturbed reference state
    (as the reference writes it)."""
    R = GAS["cp"] - GAS["cv"]
    gamma = GAS["cp"] / GAS["cv"]
    rho_inf = REF_Y[0] / (R * REF_Y[4])
    u = REF_Y[1:4]
    e_kinetic = 0.5 * rho_inf * float(u @ u)
    x0, y0 = xyz[:, 0] - EPICENTER[0], xyz[:, 1] - EPICENTER[1]
    pert = 1 + AMPLITUDE * np.exp(-(x0 * x0 + y0 * y0) / (2 * WIDTH * WIDTH))
    U = np.zeros((xyz.shape[0], 5))
    U[:, 0] = rho_inf * pert
    U[:, 1:4] = u[None, :] * U[:, 0:1]
    U[:, 4] = REF_Y[0] / (gamma - 1) * pert + e_kinetic
    return U


def setup():
    """Mesh, basis tables, context dict, lateral faces (element, face id), essential dofs and the initial state."""
    mesh = M.box_mesh((2, 2, 1), 3, lengths=(1.0, 1.0, 0.025), periodic=(False, False, False))
    tab = M.tabulate(3, 0)
    d = {"gas": dict(GAS), "tau": PR.default_tau(3), "state_var": 0, "stabilization": 2, "is_implicit": False, "dt": DT,
         "idl_enable": True, "idl_amplitude": 1.0 / 2e-3, "idl_start": 0.0, "idl_length": 0.25, "idl_pressure": 70.0}
    fe, fi = M.box_boundary(mesh, [0, 1, 2, 3])
    ess = M.essential_dofs(mesh, [4, 5], [3])
    q0 = PR.nodal_field(mesh, gaussian_wave_ic).ravel()
    return mesh, tab, d, (fe, fi), ess, q0


def _records(vals, free_z_bound=1.0):
    """Split the global vector into node records of 5 values, or of 4 with the z-component left out: a free z-momentum
    (or z-velocity) is below free_z_bound, the energy (or temperature) that takes its place in a record of 4 above."""
    recs, constrained, i = [], [], 0
    while i < len(vals):
        if abs(vals[i + 3]) < free_z_bound:
            recs.append(vals[i:i + 5])
            constrained.append(False)
            i += 5
        else:
            recs.append(np.array([vals[i], vals[i + 1], vals[i + 2], 0.0, vals[i + 3]]))
            constrained.append(True)
            i += 4
    return np.array(recs), np.array(constrained)


def compare(mesh, q, q0, t, fixture=FIXTURE, free_z_bound=1.0):
    """RegressionTest metric of the state q against the fixture (order-invariant match), and the initial state's."""
    vals, step, t_ref = honee_file.read_vec(fixture)
    assert step == STEPS and abs(t_ref - STEPS * DT) < 1e-15 and abs(t - t_ref) < 1e-15
    ref, constrained = _records(vals, free_z_bound)
    zside = np.zeros(mesh.nnodes, dtype=bool)
    zside[M.boundary_nodes(mesh, [4, 5])] = True
    mine = np.asarray(q).reshape(-1, 5)
    assert ref.shape[0] == mesh.nnodes and constrained.sum() == zside.sum() == 98
    assert np.all(mine[zside, 3] == 0.0)
    scale = np.abs(vals).max()  # max|Qref|
    cost = np.abs(ref[:, None, :] - mine[None, :, :]).max(-1) / scale
    cost[constrained[:, None] != zside[None, :]] = 1e3
    r, c = linear_sum_assignment(cost)
    err = float(np.abs(ref[r] - mine[c]).max() / scale)
    err0 = float(np.abs(ref[r] - np.asarray(q0).reshape(-1, 5)[c]).max() / scale)
    return err, err0


# ---- the implicit runs behind tests/output/gaussianwave-IDL.bin and gaussianwave-IDL-entropy.bin ----
# examples/navierstokes.c:245 ("Gaussian Wave, with IDL": -ts_alpha_radius 0.5, primitive variables from the YAML) and
# :230 ("Gaussian Wave, IDL and Entropy variables": -state_var entropy, the default radius): examples/gaussianwave.yaml
# on a 5 x 5 x 1 box at the default degree 1, implicit SUPG, TSALPHA with dt 2e-3 for 5 steps, Newton rtol 1e-4 with
# the Jacobian lagged 20 iterations across solves, GMRES rtol 1e-2 with the default ILU(0), and the same damping layer,
# faces and essential z-velocity as the explicit run.  Every node lies on a z side, so the global vector holds 4 values
# per node in DMPlex vertex order (x fastest, then y, then z: the box mesh's vertex numbering).
IMPLICIT = {1: ("gaussianwave-IDL.bin", 0.5), 2: ("gaussianwave-IDL-entropy.bin", 1.0)}
IMPLICIT_TOLERANCE = 2e-11  # -compare_final_state_atol of both tests
IMPLICIT_SOLVER = dict(rtol=1e-4, ksp_rtol=1e-2, lag=20)


def prim_from_cons(U):
    R = GAS["cp"] - GAS["cv"]
    rho = U[:, 0]
    u = U[:, 1:4] / rho[:, None]
    p = (GAS["cp"] / GAS["cv"] - 1.0) * (U[:, 4] - 0.5 * rho * np.sum(u * u, axis=1))
    return np.column_stack([p, u, p / (rho * R)])


def setup_implicit(state_var):
    """Mesh, tables, context dict, lateral faces, essential dofs, the initial state in state_var, the global order of
    the free entries and the global matrix pattern (element-coupled nodes, all free component pairs)."""
    mesh = M.box_mesh((5, 5, 1), 1, lengths=(1.0, 1.0, 0.025), periodic=(False, False, False))
    tab = M.tabulate(1, 0)
    d = {"gas": dict(GAS), "tau": PR.default_tau(1), "state_var": state_var, "stabilization": 2, "is_implicit": True, "dt": DT,
         "idl_enable": True, "idl_amplitude": 1.0 / 2e-3, "idl_start": 0.0, "idl_length": 0.25, "idl_pressure": 70.0}
    fe, fi = M.box_boundary(mesh, [0, 1, 2, 3])
    ess = M.essential_dofs(mesh, [4, 5], [3])
    U0 = PR.nodal_field(mesh, gaussian_wave_ic)
    q0 = PR.q_from_prim(GAS, prim_from_cons(U0), state_var).ravel()
    q0[ess] = 0.0  # the IC's z-velocity component, zero in every variable set
    free = honee_file.global_indices(mesh, ess)  # DMPlex global order of the free entries
    node_adj = np.zeros((mesh.nnodes, mesh.nnodes), dtype=bool)
    for e in np.asarray(mesh.offsets).reshape(mesh.nelem, -1):
        node_adj[np.ix_(e, e)] = True
    pattern = node_adj[np.ix_(free // 5, free // 5)]
    return mesh, tab, d, (fe, fi), ess, q0, free, pattern


def read_implicit(state_var):
    vals, step, t = honee_file.read_vec(os.path.join(os.path.dirname(__file__), "golden", IMPLICIT[state_var][0]))
    assert step == STEPS and abs(t - STEPS * DT) < 1e-15
    return vals


def regression_metric(x, ref):
    """RegressionTest (src/misc.c:98-126): max|Q - Qref| / max|Qref| over the global vector, in order."""
    return float(np.abs(np.asarray(x) - ref).max() / np.abs(ref).max())


# ---- the run behind tests/output/gaussianwave-shell.bin (examples/navierstokes.c:240, "Gaussian Wave, using
# MatShell"): the YAML's implicit primitive SUPG run (no damping layer) on the p=3 2 x 2 x 1 box with -amat_type shell
# -pc_type vpbjacobi -ts_alpha_radius 0.5: the Jacobian stays matrix-free and the preconditioner is the point-block
# Jacobi of its node blocks, both from the persistent lagged Jacobian evaluations.  The records are matched by
# assignment as for the explicit fixture; a free z-velocity is below 1e-2, the temperature of a record of 4 above.
SHELL_FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "gaussianwave-shell.bin")
SHELL_RADIUS = 0.5


def setup_shell():
    """Mesh, tables, context dict, lateral faces, essential dofs, the primitive initial state and the free entries."""
    mesh, tab, d, faces, ess, U0 = setup()
    d = dict(d, state_var=1, is_implicit=True, idl_enable=False)
    q0 = PR.q_from_prim(GAS, prim_from_cons(U0.reshape(-1, 5)), 1).ravel()
    q0[ess] = 0.0
    return mesh, tab, d, faces, ess, q0, np.setdiff1d(np.arange(q0.size), ess)


def compare_shell(mesh, q, q0):
    return compare(mesh, q, q0, STEPS * DT, fixture=SHELL_FIXTURE, free_z_bound=1e-2)
