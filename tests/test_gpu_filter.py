"""Differential filter and SGS training pairs on the HIP path against the CPU oracle (src/differential_filter.c,
src/smartsim/sgs_dd_training.c).

* DifferentialFilter_RHS (three state-variable sets): relative 2-norm <= 1e-12, max-norm <= 1e-11;
* the LHS composite (DifferentialFilter_LHS_5 + _6 over the interleaved 11-component filtered field) and its Jacobi
  diagonal, isotropic, grid-based anisotropic and with van Driest damping on a wall-normal direction: 1e-12;
* the filter solve (KSPCG + Jacobi, natural norm, rtol 1e-10): the oracle's iteration count (+-1) and the solution to
  1e-9 relative;
* the training pairs (filter, projected filtered velocity gradient, closure inputs, filtered SGS stress) against the
  oracle's chain: 1e-8 relative (the CG tolerance propagates through the projection and the eigen solver).
"""
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import make_case, oracle_setup, product_context, rel_err
from pyhonee import problems as PR
from pyhonee.driver import NewtonianProblem

pytestmark = pytest.mark.gpu

FILTERS = {
    "iso": dict(width_scaling=(0.5, 0.5, 0.5)),
    "grid": dict(grid_based_width=True, width_scaling=(1.5, 0.7, 1.0)),
    "vandriest": dict(grid_based_width=True, width_scaling=(1.0, 0.7, 1.0), damping_function=1, friction_length=0.05),
}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _oracle_filter(opts):
    return O.diff_filter(grid_based_width=opts.get("grid_based_width", False), width_scaling=opts.get("width_scaling", (1.0, 1.0, 1.0)),
                         damping=opts.get("damping_function", 0), friction_length=opts.get("friction_length", 0.0))


def _setup(state_var, seed, opts, periodic=(True, True, True), degree=2, shape=(3, 2, 2)):
    case = make_case(shape=shape, degree=degree, periodic=periodic, perturb=0.1, state_var=state_var, stabilization=0, seed=seed)
    om, oc, qd = oracle_setup(case)
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), state_var)
    prob.setup_diff_filter(**opts)
    return case, om, oc, qd, prob


@pytest.mark.parametrize("state_var", [0, 1, 2])
def test_filter_rhs(state_var):
    _say(f"filter rhs var={state_var}")
    case, om, oc, qd, prob = _setup(state_var, 500 + state_var, FILTERS["iso"])
    r = prob.diff_filter_rhs(case["q"])
    r_ref = O.diff_filter_rhs(om, oc, qd, case["q"])
    e2, emax = rel_err(r, r_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    prob.close()


@pytest.mark.parametrize("kind,degree", [("iso", 2), ("grid", 3), ("vandriest", 2), ("grid", 1)])
def test_filter_lhs_and_diagonal(kind, degree):
    _say(f"filter lhs {kind} degree={degree}")
    case, om, oc, qd, prob = _setup(1, 510 + degree, FILTERS[kind], periodic=(True, False, True), degree=degree)
    u = case["rng"].normal(size=11 * om.nnodes)
    v, d = prob.diff_filter_lhs(u)
    Acol = O.anisotropy_collocate(om, qd)
    f = _oracle_filter(FILTERS[kind])
    v_ref = O.diff_filter_lhs(om, f, qd, Acol, 11, u)
    d_ref = O.diff_filter_lhs_diagonal(om, f, qd, Acol, 11)
    e2, emax = rel_err(v, v_ref)
    assert e2 < 1e-12 and emax < 1e-11, ("lhs", e2, emax)
    e2, emax = rel_err(d, d_ref)
    assert e2 < 1e-12 and emax < 1e-11, ("diag", e2, emax)
    prob.close()


@pytest.mark.parametrize("kind,state_var", [("grid", 1), ("iso", 0)])
def test_filter_solve(kind, state_var):
    _say(f"filter solve {kind} var={state_var}")
    case, om, oc, qd, prob = _setup(state_var, 520 + state_var, FILTERS[kind])
    X, its = prob.diff_filter_apply(case["q"])
    Acol = O.anisotropy_collocate(om, qd)
    R = O.diff_filter_rhs(om, oc, qd, case["q"])
    X_ref, its_ref = O.diff_filter_solve(om, _oracle_filter(FILTERS[kind]), qd, Acol, 11, R)
    assert abs(its - its_ref) <= 1, (its, its_ref)
    e2, emax = rel_err(X, X_ref)
    assert e2 < 1e-9 and emax < 1e-8, (e2, emax)
    prob.close()


def test_sgs_training_pairs():
    _say("sgs training pairs")
    opts = dict(grid_based_width=True, width_scaling=(1.0, 1.0, 1.0))
    case, om, oc, qd, prob = _setup(1, 530, opts, degree=2, shape=(3, 3, 2))
    prob.setup_sgs_training()
    q = case["q"]
    Acol = O.anisotropy_collocate(om, qd)
    R = O.diff_filter_rhs(om, oc, qd, q)
    A, _ = O.grid_anisotropy_projection(om, qd)
    c_prim = O.newt_from_dict(PR.tgv_context_dict(state_var=1))
    for scale in (1.2, 3.1):
        out, its = prob.sgs_training_data(q, scale)
        X_ref, its_ref = O.diff_filter_solve(om, _oracle_filter(dict(grid_based_width=True, width_scaling=(scale,) * 3)), qd, Acol, 11, R)
        Xf = X_ref.reshape(-1, 11)
        G = O.velocity_gradient_projection(om, c_prim, qd, np.ascontiguousarray(Xf[:, :5]).ravel())
        ref = O.sgs_training_nodal(c_prim.gas, X_ref, G, A).reshape(-1, 12)
        assert np.all(np.isfinite(out))
        for cols in (slice(0, 6), slice(6, 12)):
            e2, emax = rel_err(out[:, cols], ref[:, cols])
            assert e2 < 1e-8 and emax < 1e-7, (scale, cols, e2, emax)
    prob.close()
