"""The wall-bounded configuration on the HIP path: a channel box [0, 2pi] x [0, 2] x [0, 2pi/3] with wall-normal
geometric stretching (ratio 1.1, clustered at both walls), periodic in x and z, no-slip walls as essential velocity
components (problems/channel.c, essential_comps 1,2,3), the body force B = 2 umax mu / H^2 as g, primitive FP64 state,
SUPG, BDF2 with Newton-GMRES and point-block Jacobi.

* The laminar Poiseuille profile (Exact_Channel, qfunctions/channel.h:24-53) balances momentum exactly: the body
  force cancels the viscous stress, and p=2 represents the quadratic profile exactly on the stretched (still
  box-shaped) elements, so the momentum residual of the IFunction at the profile (q_dot = 0, no stabilization)
  vanishes against the size of the viscous term alone (body force off), up to the body force's imbalance from the
  density variation rho0 / rho - 1 of the reference's temperature profile (amplitude Pr Ec / 3 = 1.5e-8 here;
  measured 1.5e-8, bounded by 1e-7).  In time the profile drifts only
  at the level viscous heating causes: the reference's temperature profile is not in thermal equilibrium
  (mu u'^2 / (rho cv) = 6e-5 per unit time at the walls), measured 7.9e-6 after 10 BDF2 steps with or without SUPG and
  bounded by 1e-4.
* With seeded Fourier perturbations of 10% of the centerline velocity the run stays finite and the wall velocity stays
  exactly zero.
* The duct variant: the x direction closed by freestream HLL faces at the inflow and the outflow."""
import sys

import numpy as np
import pytest

from helpers import product_context
from pyhonee import mesh as M
from pyhonee import problems as PR
from pyhonee import ts as TS
from pyhonee.driver import NewtonianProblem, context_from_dict

pytestmark = pytest.mark.gpu

LENGTHS = (2 * np.pi, 2.0, 2 * np.pi / 3)
MU = 1e-2


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _setup(shape=(4, 8, 2), degree=2, periodic_x=True, seed=None, resource="/gpu/hip/mi355x:jacobian_precision=fp32", stabilization=2):
    mesh = M.box_mesh(shape, degree, lengths=LENGTHS, periodic=(periodic_x, False, True), stretch=(1.0, 1.1, 1.0),
                      symmetric_stretch=(False, True, False))
    tab = M.tabulate(degree, 0)
    d = PR.channel_context_dict(MU, tau=PR.default_tau(degree), stabilization=stabilization)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1, resource)
    fe, fi = M.box_boundary(mesh, [2, 3])
    prob.add_boundary(0, fe, fi)  # consistent boundary integral on the walls
    prob.set_essential(M.essential_dofs(mesh, [2, 3], [1, 2, 3]))
    Y = PR.channel_ic(mesh, d["gas"], seed=seed, lengths=LENGTHS)
    return mesh, prob, Y, d


def _run(prob, Y, steps, dt=2e-2):
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_BDF, bdf_order=2, dt=dt, ksp_pc=TS.PC_PBJACOBI))
    ts.set_solution(Y.ravel(), 0.0)
    ts.step(steps)
    q = ts.solution().reshape(-1, 5)
    st = ts.stats()
    ts.close()
    return q, st


def test_poiseuille_momentum_balance():
    _say("channel: Poiseuille momentum residual")
    res = {}
    for scale in (1.0, 0.0):
        mesh = M.box_mesh((4, 8, 2), 2, lengths=LENGTHS, periodic=(True, False, True), stretch=(1.0, 1.1, 1.0),
                          symmetric_stretch=(False, True, False))
        d = PR.channel_context_dict(MU, tau=PR.default_tau(2), stabilization=0, body_force_scale=scale)
        prob = NewtonianProblem(mesh, M.tabulate(2, 0), context_from_dict(d), 1)
        fe, fi = M.box_boundary(mesh, [2, 3])
        prob.add_boundary(0, fe, fi)
        prob.set_essential(M.essential_dofs(mesh, [2, 3], [1, 2, 3]))
        Y = PR.channel_ic(mesh, d["gas"], lengths=LENGTHS)
        g = np.zeros(Y.size)
        prob.ifunction(Y.ravel(), np.zeros(Y.size), g)
        res[scale] = g.reshape(-1, 5)
        prob.close()
    viscous = np.abs(res[0.0][:, 1]).max()
    _say(f"x-momentum residual {np.abs(res[1.0][:, 1]).max():.3e} against the viscous term {viscous:.3e}")
    assert viscous > 1e-4
    assert np.abs(res[1.0][:, 1:4]).max() < 1e-7 * viscous


@pytest.mark.parametrize("stabilization,bound", [(0, 1e-4), (2, 1e-4)])
def test_poiseuille_is_steady(stabilization, bound):
    _say(f"channel: laminar Poiseuille, stabilization {stabilization}")
    mesh, prob, Y, _ = _setup(stabilization=stabilization)
    wall = M.boundary_nodes(mesh, [2, 3])
    assert np.abs(Y[wall, 1:4]).max() < 1e-14
    q, st = _run(prob, Y, 10)
    du = np.abs(q[:, 1:4] - Y[:, 1:4]).max()
    _say(f"max velocity change {du:.3e} after 10 steps; newton {st['newton_its']} gmres {st['krylov_its']}")
    assert np.all(q[wall, 1:4] == 0.0)
    assert du < bound
    prob.close()


def test_perturbed_channel_runs():
    _say("channel: perturbed")
    mesh, prob, Y, _ = _setup(seed=180)
    wall = M.boundary_nodes(mesh, [2, 3])
    q, st = _run(prob, Y, 5)
    _say(f"newton {st['newton_its']} gmres {st['krylov_its']}")
    assert np.all(np.isfinite(q)) and np.all(q[wall, 1:4] == 0.0)
    assert np.abs(q[:, 1:4] - Y[:, 1:4]).max() > 1e-6  # the perturbation evolves
    prob.close()


def test_duct_with_freestream_inflow_outflow():
    _say("duct: freestream HLL inflow and outflow")
    mesh, prob, Y, d = _setup(shape=(4, 6, 2), periodic_x=False)
    Y_inf = np.array([PR.TGV_REFERENCE_Y[0], 2.0 / 3.0, 0.0, 0.0, PR.TGV_REFERENCE_Y[4]])  # bulk velocity of the profile
    for side in (0, 1):
        fe, fi = M.box_boundary(mesh, [side])
        prob.add_boundary(prob.FREESTREAM_HLL, fe, fi, Y_inf)
    wall = M.boundary_nodes(mesh, [2, 3])
    q, st = _run(prob, Y, 5)
    _say(f"newton {st['newton_its']} gmres {st['krylov_its']}")
    assert np.all(np.isfinite(q)) and np.all(q[wall, 1:4] == 0.0)
    prob.close()
