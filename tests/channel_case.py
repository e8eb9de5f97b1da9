"""The reference runs behind tests/output/channel.bin and channel-prim.bin (examples/navierstokes.c:252-253,
"Channel" and "Channel, Primitive"), shared by the oracle and HIP-path tests.

examples/channel.yaml with -dm_plex_box_faces 5,5,1 and -ts_max_steps 5: a box [0, 0.01] x [0, 0.01] x [0, 0.001]
periodic in x, degree 1 with q_extra 2 (4 quadrature points per direction), mu 0.01 on NS_NEWTONIAN_IG's gas, no
stabilization, the Poiseuille state of Exact_Channel (umax 40, theta0 300, P0 1e5; qfunctions/channel.h:24-53) as the
initial condition, and its body force as the gravity vector (B / rho0, 0, 0) with B = 2 umax mu / H^2
(problems/channel.c:128-137).  Essential z-velocity on the z sides (face sets 1-2), essential zero velocity on the
walls y = 0 and y = 0.01 (face sets 3-4), no natural boundary terms.  TSBEULER with dt 5e-6, PETSc's default Newton
(rtol 1e-8) and GMRES (rtol 1e-5) tolerances and ILU(0) on the assembled Jacobian.

The fixtures are PETSc global vectors in DMPlex vertex order (x fastest; 5 x 6 x 2 vertices), each node's free
components in order: 2 values on the wall nodes (velocity constrained), 4 elsewhere (z-velocity constrained)."""
import os

import numpy as np

from pyhonee import honee_file
from pyhonee import mesh as M
from pyhonee import problems as PR

GAS = {"cv": 717.0, "cp": 1004.0, "lambda": -2.0 / 3.0, "mu": 0.01, "k": 0.02638}
UMAX, THETA0, P0 = 40.0, 300.0, 1.0e5
LENGTHS = (0.01, 0.01, 0.001)
DT, STEPS = 5e-6, 5
FIXTURES = {0: "channel.bin", 1: "channel-prim.bin"}
TOLERANCE = 2e-11  # -compare_final_state_atol of both tests
SOLVER = dict(rtol=1e-8, ksp_rtol=1e-5, lag=1)


def setup(state_var, q_extra=2):
    """Mesh, tables, context dict, essential dofs, the initial state, the global order of the free entries and the
    global matrix pattern."""
    mesh = M.box_mesh((5, 5, 1), 1, lengths=LENGTHS, periodic=(True, False, False))
    tab = M.tabulate(1, q_extra)
    H = 0.5 * LENGTHS[1]
    R = GAS["cp"] - GAS["cv"]
    B = 2 * UMAX * GAS["mu"] / (H * H)
    d = {"gas": dict(GAS), "tau": PR.default_tau(1), "state_var": state_var, "stabilization": 0, "is_implicit": True, "dt": DT,
         "g": (B / (P0 / (R * THETA0)), 0.0, 0.0)}
    ess = np.union1d(M.essential_dofs(mesh, [4, 5], [3]), M.essential_dofs(mesh, [2, 3], [1, 2, 3]))
    Y0 = PR.nodal_field(mesh, lambda xyz: PR.channel_prim(xyz, GAS, umax=UMAX, theta0=THETA0, P0=P0, center=H, H=H))
    q0 = PR.q_from_prim(GAS, Y0, state_var).ravel()
    free = honee_file.global_indices(mesh, ess)  # DMPlex global order of the free entries
    node_adj = np.zeros((mesh.nnodes, mesh.nnodes), dtype=bool)
    for e in np.asarray(mesh.offsets).reshape(mesh.nelem, -1):
        node_adj[np.ix_(e, e)] = True
    pattern = node_adj[np.ix_(free // 5, free // 5)]
    return mesh, tab, d, ess, q0, free, pattern


def read(state_var):
    vals, step, t = honee_file.read_vec(os.path.join(os.path.dirname(__file__), "golden", FIXTURES[state_var]))
    assert step == STEPS and abs(t - STEPS * DT) < 1e-18
    return vals
