"""The allocation-free operator apply (CeedMiOperatorApplyNoAlloc) against CeedOperatorApply, through the C-ABI, on the
IJacobian composite (its passive fields, quadrature data and jac_data, stay bound between calls).

Both entry points run the same kernels in the same order, so the outputs must be bitwise identical (the two-pass
scatter's E-vector lives in the library's fixed device scratch instead of the Ceed's workspace).  Two calls in a row,
the second on a smaller operator, so the second runs over the stale contents of the first.  Also: an operator that was
never applied, and one whose E-vector exceeds the fixed scratch, are refused with an error instead of allocating.
"""
import ctypes as ct
import sys

import numpy as np
import pytest

from helpers import make_case, product_context
from pyhonee.driver import CEED_MEM_DEVICE, NewtonianProblem, load_libs

pytestmark = pytest.mark.gpu

CEED_USE_POINTER = 1


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _vec(lib, ceed, t):
    v = ct.c_void_p()
    assert lib.CeedVectorCreate(ceed, ct.c_longlong(t.numel()), ct.byref(v)) == 0
    assert lib.CeedVectorSetArray(v, CEED_MEM_DEVICE, CEED_USE_POINTER, ct.c_void_p(t.data_ptr())) == 0
    return v


def _apply_both(shape, seed):
    import torch

    ceedlib, setup = load_libs()
    ceedlib.CeedVectorCreate.argtypes = [ct.c_void_p, ct.c_longlong, ct.POINTER(ct.c_void_p)]
    ceedlib.CeedVectorSetArray.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p]
    ceedlib.CeedVectorDestroy.argtypes = [ct.POINTER(ct.c_void_p)]
    ceedlib.CeedOperatorApply.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
    ceedlib.CeedMiOperatorApplyNoAlloc.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p]
    ceedlib.CeedGetErrorMessage.argtypes = [ct.c_void_p, ct.POINTER(ct.c_char_p)]
    case = make_case(shape=shape, degree=3, state_var=1, stabilization=2, seed=seed)
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 1)
    dev = torch.device("cuda:0")
    q = torch.tensor(case["q"], dtype=torch.float64, device=dev)
    qdot = torch.tensor(case["qdot"], dtype=torch.float64, device=dev)
    g0 = torch.zeros_like(q)
    prob.ifunction(q, qdot, g0, mem=CEED_MEM_DEVICE)  # writes jac_data
    dq = torch.tensor(case["rng"].uniform(-1e-3, 1e-3, q.shape), dtype=torch.float64, device=dev)
    prob.ijacobian(dq, g0, 300.0, mem=CEED_MEM_DEVICE)  # binds the IJacobian and sets its shift
    ceed = prob.ceed
    # the IJacobian composite: its passive fields (qdata, jac_data) stay bound between calls
    op = setup.HoneeMiProblemGetOperator(prob.handle, 1)
    g_ref = torch.full_like(q, 7.0)
    g_na = torch.full_like(q, -3.0)
    vin, vref, vna = _vec(ceedlib, ceed, dq), _vec(ceedlib, ceed, g_ref), _vec(ceedlib, ceed, g_na)
    assert ceedlib.CeedOperatorApply(op, vin, vref, None) == 0
    err = ceedlib.CeedMiOperatorApplyNoAlloc(op, vin, vna)
    if err:
        msg = ct.c_char_p()
        ceedlib.CeedGetErrorMessage(ceed, ct.byref(msg))
        raise AssertionError(f"CeedMiOperatorApplyNoAlloc failed ({err}): {msg.value.decode()}")
    prob.synchronize()
    a, b = g_ref.cpu().numpy(), g_na.cpu().numpy()
    for v in (vin, vref, vna):
        ceedlib.CeedVectorDestroy(ct.byref(v))
    prob.close()
    return a, b


def test_no_alloc_apply_matches_apply_twice_second_smaller():
    for shape, seed in (((4, 3, 3), 11), ((2, 2, 2), 12)):
        _say(f"no-alloc apply {shape} p=3 seed {seed}")
        a, b = _apply_both(shape, seed)
        assert np.isfinite(a).all() and np.abs(a).max() > 0
        assert np.array_equal(a, b), float(np.abs(a - b).max())


def test_no_alloc_apply_refuses_instead_of_allocating():
    import torch

    ceedlib, setup = load_libs()
    ceedlib.CeedMiOperatorApplyNoAlloc.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p]
    ceedlib.CeedVectorCreate.argtypes = [ct.c_void_p, ct.c_longlong, ct.POINTER(ct.c_void_p)]
    ceedlib.CeedVectorSetArray.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p]
    ceedlib.CeedVectorDestroy.argtypes = [ct.POINTER(ct.c_void_p)]
    _say("no-alloc apply on an operator that was never applied")
    case = make_case(shape=(2, 2, 2), degree=3, state_var=1, stabilization=2, seed=13)
    prob = NewtonianProblem(case["mesh"], case["tab"], product_context(case), 1)
    q = torch.tensor(case["q"], dtype=torch.float64, device="cuda:0")
    g = torch.zeros_like(q)
    vin, vout = _vec(ceedlib, prob.ceed, q), _vec(ceedlib, prob.ceed, g)
    op = setup.HoneeMiProblemGetOperator(prob.handle, 1)  # the IJacobian: never applied, so not bound
    assert ceedlib.CeedMiOperatorApplyNoAlloc(op, vin, vout) != 0
    for v in (vin, vout):
        ceedlib.CeedVectorDestroy(ct.byref(v))
    prob.close()


def test_no_alloc_apply_refuses_an_evector_beyond_the_fixed_scratch():
    """35^3 p=3 elements need 35^3 x 64 x 5 = 13.7 M doubles of E-vector, more than the 12 Mi-double scratch."""
    import torch

    from pyhonee import mesh as M
    from pyhonee import problems as PR

    ceedlib, setup = load_libs()
    ceedlib.CeedMiOperatorApplyNoAlloc.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p]
    ceedlib.CeedVectorCreate.argtypes = [ct.c_void_p, ct.c_longlong, ct.POINTER(ct.c_void_p)]
    ceedlib.CeedVectorSetArray.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p]
    ceedlib.CeedVectorDestroy.argtypes = [ct.POINTER(ct.c_void_p)]
    ceedlib.CeedGetErrorMessage.argtypes = [ct.c_void_p, ct.POINTER(ct.c_char_p)]
    _say("no-alloc apply beyond the fixed scratch (35^3 p=3)")
    mesh = M.box_mesh((35, 35, 35), 3)
    tab = M.tabulate(3, 0)
    prob = NewtonianProblem(mesh, tab, PR.tgv_context(state_var=1, stabilization=2, dt=2e-3), 1)
    Y = PR.nodal_field(mesh, lambda x: PR.taylor_green_prim(x))
    q = torch.tensor(Y.ravel(), dtype=torch.float64, device="cuda:0")
    g = torch.zeros_like(q)
    prob.ifunction(q, torch.zeros_like(q), g, mem=CEED_MEM_DEVICE)
    prob.ijacobian(q, g, 500.0, mem=CEED_MEM_DEVICE)
    vin, vout = _vec(ceedlib, prob.ceed, q), _vec(ceedlib, prob.ceed, g)
    op = setup.HoneeMiProblemGetOperator(prob.handle, 1)
    assert ceedlib.CeedMiOperatorApplyNoAlloc(op, vin, vout) != 0
    msg = ct.c_char_p()
    ceedlib.CeedGetErrorMessage(prob.ceed, ct.byref(msg))
    assert b"fixed scratch" in msg.value, msg.value
    for v in (vin, vout):
        ceedlib.CeedVectorDestroy(ct.byref(v))
    prob.close()
