"""HIP path (through the C-ABI libceed.so and the PETSc-free SetupLibceed driver) against the CPU oracle.

Tolerances (FP64 residual path): relative 2-norm error <= 1e-12 and relative max-norm error <= 1e-11 for the assembled
IFunction/IJacobian L-vectors; quadrature data (qdata, jac_data) at 1e-12.  The summation order of the element
scatter-add and of the collocated-gradient contraction differs from the oracle's, so agreement is to rounding, not
bitwise.  FP32 IJacobian: relative 2-norm error against the FP64 oracle <= 1e-5 (single-precision arithmetic on
O(1e3) cancellations).
"""
import sys

import numpy as np
import pytest

import pyoracle as O
from helpers import make_case, oracle_setup, product_context, rel_err
from pyhonee.driver import NewtonianProblem

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run_ifunction(case, resource="/gpu/hip/mi355x"):
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), case["ctx"]["state_var"], resource)
    g = np.zeros_like(case["q"])
    prob.ifunction(case["q"], case["qdot"], g, time=0.0)
    return prob, g


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_qdata(degree):
    _say(f"qdata degree={degree}")
    case = make_case(shape=(2, 3, 2), degree=degree, perturb=0.2, seed=degree)
    om, oc, qd = oracle_setup(case)
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 1)
    qd_gpu = prob.qdata()
    e2, emax = rel_err(qd_gpu, qd)
    assert e2 < 1e-13 and emax < 1e-13, (e2, emax)


@pytest.mark.parametrize("state_var", [0, 1, 2])
@pytest.mark.parametrize("stab", [0, 1, 2])
def test_ifunction_p3(state_var, stab):
    _say(f"ifunction p=3 var={state_var} stab={stab}")
    case = make_case(degree=3, state_var=state_var, stabilization=stab, seed=10 + 3 * state_var + stab)
    om, oc, qd = oracle_setup(case)
    v_ref, jac_ref = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    prob, g = _run_ifunction(case)
    e2, emax = rel_err(g, v_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    ej = rel_err(prob.jac_data(), jac_ref)
    assert ej[0] < 1e-12, ej


@pytest.mark.parametrize("degree", [1, 2, 4])
def test_ifunction_other_degrees(degree):
    _say(f"ifunction degree={degree}")
    case = make_case(shape=(2, 2, 3), degree=degree, state_var=1, stabilization=2, seed=30 + degree)
    om, oc, qd = oracle_setup(case)
    v_ref, jac_ref = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    prob, g = _run_ifunction(case)
    e2, emax = rel_err(g, v_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    assert rel_err(prob.jac_data(), jac_ref)[0] < 1e-12


def test_ifunction_idl_gravity_nonperiodic():
    _say("ifunction idl+gravity nonperiodic")
    case = make_case(shape=(3, 2, 2), degree=3, periodic=(False, True, False), state_var=0, stabilization=2, idl=True,
                     g=(0.1, -9.81, 0.3), seed=41)
    om, oc, qd = oracle_setup(case)
    v_ref, jac_ref = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    prob, g = _run_ifunction(case)
    e2, emax = rel_err(g, v_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    assert rel_err(prob.jac_data(), jac_ref)[0] < 1e-12


@pytest.mark.parametrize("state_var", [0, 1, 2])
def test_ijacobian_fp64(state_var):
    _say(f"ijacobian fp64 var={state_var}")
    case = make_case(degree=3, state_var=state_var, stabilization=2, seed=50 + state_var, idl=(state_var == 1))
    om, oc, qd = oracle_setup(case)
    _, jac_ref = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    shift = 250.0
    oc.ijacobian_time_shift = shift
    q5 = case["q"].reshape(-1, 5)
    dq = (case["rng"].uniform(-1, 1, q5.shape) * np.abs(q5).mean(0) * 1e-3).ravel()
    v_ref = O.ijacobian(om, oc, qd, jac_ref, dq)
    prob, _ = _run_ifunction(case)
    v = np.zeros_like(dq)
    prob.ijacobian(dq, v, shift)
    e2, emax = rel_err(v, v_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)


def test_ijacobian_fp32_variant():
    _say("ijacobian fp32")
    case = make_case(degree=3, state_var=1, stabilization=2, seed=61)
    om, oc, qd = oracle_setup(case)
    _, jac_ref = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    shift = 250.0
    oc.ijacobian_time_shift = shift
    dq = case["rng"].uniform(-1e-3, 1e-3, case["q"].shape)
    v_ref = O.ijacobian(om, oc, qd, jac_ref, dq)
    prob, _ = _run_ifunction(case, resource="/gpu/hip/mi355x:jacobian_precision=fp32")
    jd32 = prob.jac_data32()
    assert rel_err(jd32.astype(np.float64), jac_ref)[0] < 1e-6
    v = np.zeros_like(dq)
    prob.ijacobian(dq, v, shift)
    e2, emax = rel_err(v, v_ref)
    print(f"fp32 IJacobian error vs fp64 oracle: rel2={e2:.3e} relmax={emax:.3e}", flush=True)
    assert e2 < 1e-5, (e2, emax)


@pytest.mark.parametrize("state_var,degree", [(0, 2), (1, 3), (2, 3), (1, 1)])
def test_ijacobian_diagonal(state_var, degree):
    """Point-Jacobi diagonal (CeedOperatorLinearAssembleDiagonal) against the oracle's element-column assembly."""
    _say(f"ijacobian diagonal var={state_var} degree={degree}")
    case = make_case(shape=(2, 2, 2), degree=degree, state_var=state_var, stabilization=2, seed=80 + state_var, idl=(degree == 1))
    om, oc, qd = oracle_setup(case)
    _, jac_ref = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    shift = 120.0
    oc.ijacobian_time_shift = shift
    d_ref = O.ijacobian_diagonal(om, oc, qd, jac_ref)
    prob, _ = _run_ifunction(case)
    d = np.zeros_like(d_ref)
    prob.ijacobian_diagonal(d, shift)
    e2, emax = rel_err(d, d_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)


def test_repeated_apply_is_stable():
    """Two applies with updated context values reuse the bound kernel and read the new values."""
    _say("repeated apply")
    case = make_case(degree=2, shape=(2, 2, 2), state_var=1, stabilization=2, seed=71)
    om, oc, qd = oracle_setup(case)
    prob, g1 = _run_ifunction(case)
    g2 = np.zeros_like(g1)
    prob.ifunction(case["q"], case["qdot"], g2, dt=5e-4)
    oc.dt = 5e-4
    v_ref, _ = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    assert rel_err(g2, v_ref)[0] < 1e-12
    assert rel_err(g1, g2)[0] > 1e-10  # dt enters Tau, so the residual changed


@pytest.mark.parametrize("degree,periodic,scatter", [(3, (True, True, True), "twopass"), (2, (False, True, False), "twopass"),
                                                     (1, (True, False, True), "twopass"), (3, (True, True, True), "atomic"),
                                                     (2, (False, True, False), "atomic")])
def test_scatter_modes(degree, periodic, scatter):
    """Both element scatters (resource option scatter=twopass, the default, or scatter=atomic): oracle parity for the
    IFunction and the IJacobian; the deterministic two-pass scatter also gives bitwise-identical repeated applies."""
    _say(f"{scatter} scatter degree={degree} periodic={periodic}")
    case = make_case(shape=(3, 2, 2), degree=degree, periodic=periodic, state_var=1, stabilization=2, seed=90 + degree)
    om, oc, qd = oracle_setup(case)
    v_ref, jac_ref = O.ifunction(om, oc, qd, case["q"], case["qdot"])
    prob, g = _run_ifunction(case, resource=f"/gpu/hip/mi355x:scatter={scatter}")
    e2, emax = rel_err(g, v_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    g2 = np.zeros_like(g)
    prob.ifunction(case["q"], case["qdot"], g2, time=0.0)
    if scatter == "twopass":
        assert np.array_equal(g, g2)
    shift = 250.0
    oc.ijacobian_time_shift = shift
    dq = case["rng"].uniform(-1e-3, 1e-3, case["q"].shape)
    vj_ref = O.ijacobian(om, oc, qd, jac_ref, dq)
    vj = np.zeros_like(dq)
    prob.ijacobian(dq, vj, shift)
    e2, emax = rel_err(vj, vj_ref)
    assert e2 < 1e-12 and emax < 1e-11, (e2, emax)
    vj2 = np.zeros_like(dq)
    prob.ijacobian(dq, vj2, shift)
    if scatter == "twopass":
        assert np.array_equal(vj, vj2)
