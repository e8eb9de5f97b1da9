"""Explicit-path operators on the HIP path against the CPU oracle: the RHS composite (RHSFunction_Newtonian plus
boundary faces), the mass matrices of the explicit "mass_" KSP (SUPG-stabilized and Mass_5) with their diagonals, and
the total kinetic energy monitor.

Tolerances as for the IFunction parity tests: relative 2-norm <= 1e-12 and relative max-norm <= 1e-11 in FP64;
the monitor sums at 1e-12 relative (StrainDissipation, MuVorticity2, TotalKineticEnergy) and 1e-12 of the kinetic
energy scale for the two roundoff-level columns on periodic meshes.
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


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("stabilization,gravity,degree", [(0, False, 3), (1, False, 3), (2, True, 2), (1, True, 1), (0, False, 4), (0, True, 2)])
def test_rhs_matches_oracle(stabilization, gravity, degree):
    _say(f"rhs stab={stabilization} gravity={gravity} degree={degree}")
    g = (0.0, -9.81, 0.3) if gravity else (0.0, 0.0, 0.0)
    case = make_case(shape=(3, 2, 2), degree=degree, perturb=0.12, state_var=0, stabilization=stabilization, g=g, implicit=False,
                     seed=200 + 10 * stabilization + degree)
    om, oc, qd = oracle_setup(case)
    q = case["q"]
    r_ref = O.rhs(om, oc, qd, q)
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 0)
    assert prob.explicit
    r = np.zeros_like(q)
    prob.rhs(q, r, time=0.25)
    e2, emax = rel_err(r, r_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    prob.close()


@pytest.mark.parametrize("kind", [0, 1, 2])
def test_rhs_with_faces_matches_oracle(kind):
    """Boundary faces join the RHS composite of an explicit problem (the face bodies carry the explicit sign)."""
    _say(f"rhs with faces kind={kind}")
    case = make_case(shape=(3, 2, 2), degree=2, periodic=(False, True, False), perturb=0.1, state_var=0, stabilization=1, implicit=False,
                     seed=230 + kind)
    om, oc, qd = oracle_setup(case)
    sides = [0, 1, 4] if kind == 0 else [0, 1]
    fe, fi = M.box_boundary(case["mesh"], sides)
    f = O.Faces(fe, fi)
    q = case["q"]
    qds = O.setup_qdata_boundary(om, f)
    if kind == 0:
        v_bnd, _ = O.boundary_integral(om, f, oc, qds, q)
    else:
        U_inf, _ = O.state_from_q(oc.gas, Y_INF, 1)
        v_bnd, _ = O.freestream(om, f, O.freestream_ctx(oc, U_inf, Y_INF, kind == 2), qds, q)
    r_ref = O.rhs(om, oc, qd, q) + v_bnd
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 0)
    prob.add_boundary(kind, fe, fi, Y_INF if kind else None)
    r = np.zeros_like(q)
    prob.rhs(q, r)
    e2, emax = rel_err(r, r_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    prob.close()


@pytest.mark.parametrize("stabilization,degree", [(2, 3), (2, 2), (0, 3), (1, 2)])
def test_mass_and_diagonal_match_oracle(stabilization, degree):
    """SUPG problems use the stabilized mass; the others Mass_5, which equals the unstabilized Newtonian mass."""
    _say(f"mass stab={stabilization} degree={degree}")
    case = make_case(shape=(2, 3, 2), degree=degree, perturb=0.1, state_var=0, stabilization=stabilization, implicit=False,
                     seed=260 + 10 * stabilization + degree)
    om, oc, qd = oracle_setup(case)
    q = case["q"]
    x = case["rng"].uniform(-1, 1, q.shape)
    y_ref = O.mass_newtonian(om, oc, qd, q, x)
    d_ref = O.mass_newtonian_diagonal(om, oc, qd, q)
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 0)
    y = np.zeros_like(q)
    prob.mass(q, x, y)
    e2, emax = rel_err(y, y_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    d = np.zeros_like(q)
    prob.mass_diagonal(q, d)
    e2, emax = rel_err(d, d_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    prob.close()


@pytest.mark.parametrize("state_var", [0, 1, 2])
def test_monitor_ke_matches_oracle(state_var):
    _say(f"monitor ke var={state_var}")
    case = make_case(shape=(2, 2, 3), degree=2, perturb=0.1, state_var=state_var, stabilization=0, seed=290 + state_var)
    om, oc, qd = oracle_setup(case)
    s_ref = O.monitor_ke(om, oc, qd, case["q"])
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), state_var)
    s = prob.monitor_ke(case["q"])
    assert abs(prob.volume - O.volume(om, qd)) < 1e-13 * O.volume(om, qd)
    for k in (0, 1, 4):
        assert abs(s[k] - s_ref[k]) <= 1e-12 * abs(s_ref[k]), (k, s[k], s_ref[k])
    for k in (2, 3):
        assert abs(s[k] - s_ref[k]) <= 1e-12 * abs(s_ref[0]) + 1e-12 * abs(s_ref[k]), (k, s[k], s_ref[k])
    prob.close()


@pytest.mark.parametrize("stabilization", [2, 0])
def test_mass_point_block_diagonal_matches_unit_columns(stabilization):
    _say(f"mass point-block stab={stabilization}")
    case = make_case(shape=(2, 2, 2), degree=2, perturb=0.1, state_var=0, stabilization=stabilization, implicit=False, seed=310 + stabilization)
    om, oc, qd = oracle_setup(case)
    q = case["q"]
    nn = om.nnodes
    pb_ref = np.zeros((nn, 5, 5))
    for node in range(nn):
        for ci in range(5):
            e = np.zeros(5 * nn)
            e[node * 5 + ci] = 1.0
            pb_ref[node, :, ci] = O.mass_newtonian(om, oc, qd, q, e)[node * 5: node * 5 + 5]
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 0)
    pb = np.zeros((nn, 5, 5))
    prob.mass_point_block_diagonal(q, pb)
    e2, emax = rel_err(pb, pb_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    prob.close()
