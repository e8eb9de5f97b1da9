"""Data-driven SGS closure on the HIP path against the CPU oracle (problems/sgs_dd_model.c and its projections).

* grid anisotropy projection (src/grid_anisotropy_tensor.c): AnisotropyTensorProjection + Mass_7, CG with diagonal
  Jacobi at rtol 1e-10 -- the same iteration count as the oracle's CG and the solution to 1e-11 relative;
* velocity-gradient projection (src/velocity_gradient_projection.c): VelocityGradientProjectionRHS + lumped Mass_9;
* nodal closure (ComputeSgsDDNodal): per-node evaluation against the oracle at every node, on periodic and
  non-periodic meshes (node multiplicities 1..8);
* SGS residual (IFunction_NodalSgs) for a given nodal stress, and the whole IFunction with the SGS term added
  (src/setupts.c:118-120) against the oracle's chain IFunction + projection + closure + residual.

FP64 tolerances: relative 2-norm <= 1e-12 (1e-11 where the result passes through the eigen solvers and the CG) and
relative max-norm <= 1e-10.
"""
import os
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import make_case, oracle_setup, product_context, rel_err
from pyhonee import sgs as S
from pyhonee.driver import NewtonianProblem

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dd_sgs_data")
ALPHA = 0.1


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def model():
    p = S.load_dd_model(GOLDEN, alpha=ALPHA)
    return p, O.SgsModel(p)


def _setup(state_var, seed, model, periodic=(True, True, True), degree=2, shape=(3, 2, 2)):
    case = make_case(shape=shape, degree=degree, periodic=periodic, perturb=0.12, state_var=state_var, stabilization=0, seed=seed)
    om, oc, qd = oracle_setup(case)
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), state_var)
    prob.setup_sgs_dd(model[0])
    return case, om, oc, qd, prob


@pytest.mark.parametrize("degree,periodic", [(2, (True, True, True)), (3, (False, True, False)), (1, (True, False, True))])
def test_grid_anisotropy_projection(model, degree, periodic):
    _say(f"sgs grid anisotropy degree={degree} periodic={periodic}")
    case, om, oc, qd, prob = _setup(1, 400 + degree, model, periodic=periodic, degree=degree)
    X, its = prob.sgs_grid_anisotropy()
    X_ref, its_ref = O.grid_anisotropy_projection(om, qd)
    assert its == its_ref, (its, its_ref)
    e2, emax = rel_err(X, X_ref)
    assert e2 < 1e-11 and emax < 1e-10, (e2, emax)
    prob.close()


@pytest.mark.parametrize("state_var", [0, 1, 2])
def test_velocity_gradient_projection(model, state_var):
    _say(f"sgs velocity gradient var={state_var}")
    case, om, oc, qd, prob = _setup(state_var, 410 + state_var, model, periodic=(True, False, True))
    q = case["q"]
    G = np.zeros(9 * om.nnodes)
    prob.velocity_gradient_projection(q, G)
    G_ref = O.velocity_gradient_projection(om, oc, qd, q)
    e2, emax = rel_err(G, G_ref)
    assert e2 < 1e-12 and emax < 1e-10, (e2, emax)
    prob.close()


@pytest.mark.parametrize("state_var,periodic", [(1, (True, True, True)), (0, (False, False, True)), (2, (False, True, False))])
def test_sgs_nodal_stress(model, state_var, periodic):
    _say(f"sgs nodal stress var={state_var} periodic={periodic}")
    case, om, oc, qd, prob = _setup(state_var, 420 + state_var, model, periodic=periodic)
    q = case["q"]
    G = O.velocity_gradient_projection(om, oc, qd, q)
    A, _ = prob.sgs_grid_anisotropy()
    km = np.zeros(6 * om.nnodes)
    prob.sgs_nodal_stress(q, G, km)
    km_ref = O.sgs_nodal(oc.gas, state_var, model[1], q, G, A.ravel())
    assert np.all(np.isfinite(km)) and np.abs(km_ref).max() > 0
    e2, emax = rel_err(km, km_ref)
    assert e2 < 1e-11 and emax < 1e-10, (e2, emax)
    prob.close()


def _two_layer_model(seed=31):
    """A 6-20-20-6 closure: the shipped model's first and output layers with a random second hidden layer."""
    p = dict(S.load_dd_model(GOLDEN, alpha=ALPHA))
    rng = np.random.default_rng(seed)
    p["Wh"] = rng.normal(size=(20, 20)) * 0.3
    p["bh"] = rng.normal(size=20) * 0.1
    return p, O.SgsModel(p)


@pytest.mark.parametrize("state_var", [1, 2])
def test_sgs_nodal_stress_two_hidden_layers(state_var):
    """The fused nodal closure with a second hidden layer (the context's num_layers = 2 extension) against the oracle."""
    _say(f"sgs nodal stress 6-20-20-6 var={state_var}")
    m2 = _two_layer_model()
    case, om, oc, qd, prob = _setup(state_var, 450 + state_var, m2, periodic=(True, False, True))
    q = case["q"]
    G = O.velocity_gradient_projection(om, oc, qd, q)
    A, _ = prob.sgs_grid_anisotropy()
    km = np.zeros(6 * om.nnodes)
    prob.sgs_nodal_stress(q, G, km)
    km_ref = O.sgs_nodal(oc.gas, state_var, m2[1], q, G, A.ravel())
    km_one = O.sgs_nodal(oc.gas, state_var, O.SgsModel(S.load_dd_model(GOLDEN, alpha=ALPHA)), q, G, A.ravel())
    assert np.linalg.norm(km_ref - km_one) > 0.1 * np.linalg.norm(km_one)  # the second layer changes the stress
    e2, emax = rel_err(km, km_ref)
    assert e2 < 1e-11 and emax < 1e-10, (e2, emax)
    prob.close()


@pytest.mark.parametrize("state_var", [0, 1, 2])
def test_sgs_residual(model, state_var):
    _say(f"sgs residual var={state_var}")
    case, om, oc, qd, prob = _setup(state_var, 430 + state_var, model, periodic=(True, True, False), degree=3)
    q = case["q"]
    km = case["rng"].normal(size=6 * om.nnodes)
    g = np.zeros_like(q)
    prob.apply_sgs_residual(q, km, g)
    g_ref = O.ifunction_nodal_sgs(om, oc, qd, q, km)
    e2, emax = rel_err(g, g_ref)
    assert e2 < 1e-12 and emax < 1e-10, (e2, emax)
    prob.close()


@pytest.mark.parametrize("state_var,degree", [(1, 2), (0, 3), (2, 2)])
def test_ifunction_with_sgs(model, state_var, degree):
    """The IFunction of an implicit problem with the SGS model: Newtonian volume residual plus the SGS residual of
    the nodal stress computed from the projected velocity gradient and grid anisotropy."""
    _say(f"ifunction with sgs var={state_var} degree={degree}")
    case, om, oc, qd, prob = _setup(state_var, 440 + state_var, model, degree=degree)
    q, qdot = case["q"], case["qdot"]
    g = np.zeros_like(q)
    prob.ifunction(q, qdot, g)
    v_ref, _ = O.ifunction(om, oc, qd, q, qdot)
    G = O.velocity_gradient_projection(om, oc, qd, q)
    A, _ = O.grid_anisotropy_projection(om, qd)
    km = O.sgs_nodal(oc.gas, state_var, model[1], q, G, A)
    v_sgs = O.ifunction_nodal_sgs(om, oc, qd, q, km)
    assert np.linalg.norm(v_sgs) > 1e-6 * np.linalg.norm(v_ref)  # the SGS term is not negligible here
    e2, emax = rel_err(g, v_ref + v_sgs)
    assert e2 < 1e-11 and emax < 1e-10, (e2, emax)
    prob.close()
