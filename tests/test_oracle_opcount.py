"""The exact operation count of the QFunction bodies (oracle/opcount.cpp), from which the benchmark's FLOP figures come.

The oracle's point bodies (oracle/ns_qfunctions.h) are compiled a second time over an operation-counting scalar.  The
pins: the counted code returns the oracle's outputs bit for bit (so it is the same code on the same inputs); the
conversion from conservative to primitive variables (qfunctions/newtonian_state.h:64-73) gives its hand count; the
count does not depend on the point's data for the benchmark's configuration; and the figures stored in bench.py
equal the count."""
import os
import sys

import numpy as np
import pytest

import pyoracle as O
from pyhonee import problems as PR

sys.path.insert(0, os.path.join(os.path.dirname(__file__), ".."))
import bench  # noqa: E402


def _point(rng, c, var):
    Y = np.array([PR.TGV_REFERENCE_Y[0] * (1 + 0.05 * rng.uniform(-1, 1)), *rng.uniform(-0.5, 0.5, 3),
                  PR.TGV_REFERENCE_Y[4] * (1 + 0.05 * rng.uniform(-1, 1))])
    U, Yd = O.state_from_q(c.gas, Y, 1)
    q = O.state_to_q(c.gas, U, Yd, var)
    gref = rng.uniform(-1, 1, (3, 5)) * np.abs(q)[None, :] * 0.1
    qdot = rng.uniform(-1, 1, 5) * np.abs(q) * 0.01
    J = np.diag(rng.uniform(2, 4, 3)) + 0.2 * rng.uniform(-1, 1, (3, 3))
    return q, gref, qdot, float(rng.uniform(0.1, 1)), J, rng.uniform(0, 1, 3)


def test_hand_count_of_prim_from_cons():
    """Y_i = U_i / rho (3 div); e_k = .5 * u.u (3 mul, 2 add, 1 mul); e_t = E / rho (1 div); e_i = e_t - e_k (1 add);
    T = e_i / cv (1 div); p = (cp / cv - 1) * rho * e_i (1 div, 1 add, 2 mul): 4 additions, 6 multiplications and 6
    divisions."""
    gas = O.OGas(-2.0 / 3, 1.8e-5, 0.02638, 717.0, 1004.0)
    U = np.array([1.2, 0.3, -0.2, 0.1, 2.5e5])
    Y, n = O.opcount_prim_from_cons(gas, U)
    assert n == {"add": 4, "mul": 6, "div": 6, "sqrt": 0, "exp": 0, "log": 0}
    _, Yref = O.state_from_q(gas, U, 0)
    assert np.array_equal(Y, Yref)


@pytest.mark.parametrize("var", [0, 1, 2])
@pytest.mark.parametrize("stab", [0, 1, 2])
def test_counted_bodies_equal_the_oracle(var, stab):
    rng = np.random.default_rng(10 * var + stab)
    c = O.newt_from_dict(PR.tgv_context_dict(state_var=var, stabilization=stab, implicit=True, dt=1e-3, tau=PR.default_tau(3)))
    c.ijacobian_time_shift = 300.0
    for _ in range(5):
        q, gref, qdot, w, J, x = _point(rng, c, var)
        v, gv, jd, _ = O.opcount_ifunction(c, q, gref, qdot, w, J, x)
        v2, gv2, jd2 = O.qf_ifunction(c, q, gref, qdot, w, J, x)
        assert np.array_equal(v, v2) and np.array_equal(gv, gv2) and np.array_equal(jd[:14], jd2[:14])
        dv, dgv, _ = O.opcount_ijacobian(c, qdot, gref, w, J, jd2)
        dv2, dgv2 = O.qf_ijacobian(c, qdot, gref, w, J, jd2)
        assert np.array_equal(dv, dv2) and np.array_equal(dgv, dgv2)
    cc = O.newt_from_dict(PR.tgv_context_dict(state_var=0, stabilization=stab, implicit=False, dt=1e-3, tau=PR.default_tau(3)))
    q, gref, _, w, J, x = _point(rng, cc, 0)
    r, gr, _ = O.opcount_rhs(cc, q, gref, w, J, x)
    r2, gr2 = O.qf_rhs(cc, q, gref, w, J, x)
    assert np.array_equal(r, r2) and np.array_equal(gr, gr2)


def test_bench_figures_are_the_count():
    """The benchmark's context (primitive, SUPG, no damping layer, no gravity); the count is the same at every point."""
    rng = np.random.default_rng(7)
    c = O.newt_from_dict(PR.tgv_context_dict(state_var=1, stabilization=2, dt=2e-3))
    c.ijacobian_time_shift = 500.0
    seen_if, seen_ij = set(), set()
    for _ in range(50):
        q, gref, qdot, w, J, x = _point(rng, c, 1)
        _, _, jd, n_if = O.opcount_ifunction(c, q, gref, qdot, w, J, x)
        _, _, n_ij = O.opcount_ijacobian(c, qdot, gref, w, J, jd)
        seen_if.add(tuple(sorted(n_if.items())))
        seen_ij.add(tuple(sorted(n_ij.items())))
    assert len(seen_if) == 1 and len(seen_ij) == 1
    n_if, n_ij = dict(seen_if.pop()), dict(seen_ij.pop())
    assert n_if["exp"] == n_if["log"] == n_ij["exp"] == n_ij["log"] == 0
    assert bench.QF_OPS_PER_POINT["ifunction"] == {k: n_if[k] for k in ("add", "mul", "div", "sqrt")}
    assert bench.QF_OPS_PER_POINT["ijacobian"] == {k: n_ij[k] for k in ("add", "mul", "div", "sqrt")}


def test_basis_flops_at_p3():
    """4x4x4 points, collocated: one pass of 64 outputs of a length-4 contraction is 64 * 7 = 448 flops, three passes
    1344.  IFunction per component: interpolation and gradient of q (1344 + 1344), interpolation of q-dot (1344), the
    gradient transpose with the sums (1344 + 192) and the interpolation transpose (1344): 6912, times 5 components.
    IJacobian: 5568 per component."""
    assert bench.basis_flops_per_element(4, 4, "ifunction") == 5 * 6912
    assert bench.basis_flops_per_element(4, 4, "ijacobian") == 5 * 5568
    assert bench.basis_flops_per_element(4, 6, "ifunction") is None
    assert bench.flops_per_apply("ifunction", 1, 4, 4) == 64 * 919 + 5 * 6912
