"""Recomputed quadrature data (resource option qdata=recompute) against stored quadrature data (qdata=stored, the
default).  The volume kernels rebuild {w detJ, dXdx} from the element vertex coordinates that the Setup operator
used when its qdata vector is unmodified since, so the results equal the stored path's to rounding (the setup kernel and
the volume kernels evaluate the same device function).

* IFunction and FP64 IJacobian at p = 1..4 on a non-affine mesh (the kernels recompute for the IFunction at Q = 4, 5
  and the FP64 IJacobian at Q = 4, and read qdata elsewhere), and the explicit RHS at p = 2 (reads it in both modes);
* a write to the qdata vector (CeedVectorScale through the C-ABI) drops the recorded geometry: afterwards the kernels
  read the modified vector, in both modes alike, and the outputs change accordingly.
"""
import ctypes as ct
import sys

import numpy as np
import pytest

from helpers import make_case, product_context
from pyhonee.driver import NewtonianProblem, load_libs

pytestmark = pytest.mark.gpu

STORED = "/gpu/hip/mi355x:qdata=stored"
RECOMPUTE = "/gpu/hip/mi355x:qdata=recompute"


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _implicit(case, resource):
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 1, resource)
    g = np.zeros_like(case["q"])
    prob.ifunction(case["q"], case["qdot"], g)
    dq = np.random.default_rng(4).uniform(-1e-3, 1e-3, case["q"].shape)
    v = np.zeros_like(g)
    prob.ijacobian(dq, v, 250.0)
    return prob, g, v, dq


@pytest.mark.parametrize("degree", [1, 2, 3, 4])
def test_recomputed_qdata_matches_stored(degree):
    _say(f"qdata recompute vs stored, p={degree}")
    case = make_case(shape=(3, 2, 2), degree=degree, state_var=1, stabilization=2, perturb=0.15, seed=degree)
    p1, g1, v1, _ = _implicit(case, RECOMPUTE)
    p2, g2, v2, _ = _implicit(case, STORED)
    p1.close()
    p2.close()
    eg, ev = _rel(g1, g2), _rel(v1, v2)
    _say(f"IFunction {eg:.2e} (bitwise {np.array_equal(g1, g2)}), IJacobian {ev:.2e} (bitwise {np.array_equal(v1, v2)})")
    assert eg < 1e-14 and ev < 1e-14


def test_recomputed_qdata_explicit_rhs():
    _say("qdata recompute vs stored, explicit RHS p=2")
    case = make_case(shape=(3, 3, 2), degree=2, state_var=0, stabilization=0, perturb=0.1, seed=7, implicit=False)
    out = []
    for resource in (RECOMPUTE, STORED):
        prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 0, resource)
        r = np.zeros_like(case["q"])
        prob.rhs(case["q"], r)
        prob.close()
        out.append(r)
    e = _rel(out[0], out[1])
    _say(f"RHS {e:.2e} (bitwise {np.array_equal(out[0], out[1])})")
    assert e < 1e-14


def _scale_volume_qdata(prob, factor):
    """CeedVectorScale on the qdata vector of the IFunction composite's volume sub-operator (its first sub-operator)."""
    ceedlib, setup = load_libs()
    ceedlib.CeedOperatorCompositeGetSubList.argtypes = [ct.c_void_p, ct.POINTER(ct.POINTER(ct.c_void_p))]
    ceedlib.CeedOperatorGetFieldByName.argtypes = [ct.c_void_p, ct.c_char_p, ct.POINTER(ct.c_void_p)]
    ceedlib.CeedOperatorFieldGetVector.argtypes = [ct.c_void_p, ct.POINTER(ct.c_void_p)]
    ceedlib.CeedVectorScale.argtypes = [ct.c_void_p, ct.c_double]
    op = setup.HoneeMiProblemGetOperator(prob.handle, 0)
    subs = ct.POINTER(ct.c_void_p)()
    assert ceedlib.CeedOperatorCompositeGetSubList(op, ct.byref(subs)) == 0
    field, vec = ct.c_void_p(), ct.c_void_p()
    assert ceedlib.CeedOperatorGetFieldByName(subs[0], b"qdata", ct.byref(field)) == 0
    assert ceedlib.CeedOperatorFieldGetVector(field, ct.byref(vec)) == 0
    assert ceedlib.CeedVectorScale(vec, factor) == 0


def test_written_qdata_is_read_not_recomputed():
    _say("qdata written after setup: the kernels read it")
    case = make_case(shape=(2, 2, 2), degree=3, state_var=1, stabilization=2, perturb=0.15, seed=9)
    results = []
    for resource in (RECOMPUTE, STORED):
        prob, g0, v0, dq = _implicit(case, resource)
        _scale_volume_qdata(prob, 2.0)
        g = np.zeros_like(g0)
        prob.ifunction(case["q"], case["qdot"], g)
        v = np.zeros_like(g0)
        prob.ijacobian(dq, v, 250.0)
        prob.close()
        results.append((g0, g, v0, v))
    (g0r, gr, v0r, vr), (g0s, gs, v0s, vs) = results
    _say(f"after the write: IFunction recompute vs stored {_rel(gr, gs):.2e}, change {_rel(gr, g0r):.2e}")
    assert _rel(gr, g0r) > 1e-3  # the doubled qdata changed the residual
    assert _rel(gr, gs) < 1e-14 and _rel(vr, vs) < 1e-14
