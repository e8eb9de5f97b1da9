"""Host-side checks of the time integrator (no GPU): the explicit schemes' Butcher tableaus, and the CPU oracle's time
integrators against what the mathematics fixes.

* Tableaus of libhonee_ts.so: the order conditions of each scheme's order (up to 5 for Fehlberg's 5(4) weights), the
  failure of the next order's conditions, equality with the oracle's independent restatement, and for the SSP(n^2,3)
  schemes the stability function R(z) = 1 + z b^T (I - zA)^{-1} 1 against Ketcheson's low-storage recurrence run on
  y' = z y.
* Oracle time integrators: observed convergence orders on the Taylor-Green problem (BDF1 ~ 1, BDF2 ~ 2 including the
  half-step start, RK4 with the lumped mass ~ 4).  A plausible slip (a wrong BDF coefficient or stage weight) breaks
  the observed order.
"""
import numpy as np
import pytest

import pyoracle as O
from pyhonee import mesh as M
from pyhonee import problems as PR
from pyhonee import ts as TS

SCHEMES = [(TS.RK_4, "4", 4, 4), (TS.RK_5F, "5f", 6, 5), (TS.RK_SSP33, "ssp33", 3, 3), (TS.RK_EULER, "euler", 1, 1),
           (TS.RK_SSP_S3, "ssp3", 4, 3), (TS.RK_SSP_S3, "ssp3", 9, 3), (TS.RK_SSP_S3, "ssp3", 16, 3)]


def _order_conditions(A, b, c):
    """Residuals of the order conditions, grouped by order (Butcher trees up to order 5)."""
    e = np.ones_like(b)
    Ac, Ac2, Ac3 = A @ c, A @ c**2, A @ c**3
    AAc, AAc2, AAAc = A @ Ac, A @ Ac2, A @ (A @ Ac)
    return {
        1: [b @ e - 1],
        2: [b @ c - 1 / 2],
        3: [b @ c**2 - 1 / 3, b @ Ac - 1 / 6],
        4: [b @ c**3 - 1 / 4, b @ (c * Ac) - 1 / 8, b @ Ac2 - 1 / 12, b @ AAc - 1 / 24],
        5: [b @ c**4 - 1 / 5, b @ (c**2 * Ac) - 1 / 10, b @ (c * Ac2) - 1 / 15, b @ (c * AAc) - 1 / 30, b @ (Ac * Ac) - 1 / 20,
            b @ Ac3 - 1 / 20, b @ (A @ (c * Ac)) - 1 / 40, b @ AAc2 - 1 / 60, b @ AAAc - 1 / 120],
    }


@pytest.mark.parametrize("rk_type,name,stages,order", SCHEMES)
def test_tableau_order_conditions(rk_type, name, stages, order):
    A, b, c = TS.tableau(rk_type, stages)
    assert A.shape == (stages, stages)
    assert np.allclose(np.triu(A), 0), "explicit schemes have strictly lower-triangular A"
    conds = _order_conditions(A, b, c)
    for p in range(1, order + 1):
        assert np.max(np.abs(conds[p])) < 1e-14, (p, conds[p])
    if order < 5:
        assert np.max(np.abs(conds[order + 1])) > 1e-6, "the scheme should not reach the next order"


@pytest.mark.parametrize("rk_type,name,stages,order", SCHEMES)
def test_tableau_matches_oracle_restatement(rk_type, name, stages, order):
    A, b, c = TS.tableau(rk_type, stages)
    Ao, bo, co = O.tableau_ref(name, stages)
    np.testing.assert_allclose(A, Ao, rtol=0, atol=1e-15)
    np.testing.assert_allclose(b, bo, rtol=0, atol=1e-15)
    np.testing.assert_allclose(c, co, rtol=0, atol=1e-15)


def _ketcheson_recurrence(n, z):
    """One step of the low-storage SSP(n^2, 3) recurrence applied to y' = z y with h = 1, y0 = 1."""
    s, r = n * n, n * n - n
    q1 = 1.0 + 0j
    for _ in range(1, (n - 1) * (n - 2) // 2 + 1):
        q1 = q1 + z / r * q1
    q2 = q1
    for _ in range((n - 1) * (n - 2) // 2 + 1, n * (n + 1) // 2):
        q1 = q1 + z / r * q1
    q1 = (n * q2 + (n - 1) * (q1 + z / r * q1)) / (2 * n - 1)
    for _ in range(n * (n + 1) // 2 + 1, s + 1):
        q1 = q1 + z / r * q1
    return q1


@pytest.mark.parametrize("n", [2, 3, 4])
def test_ssp_stability_function(n):
    A, b, _ = TS.tableau(TS.RK_SSP_S3, n * n)
    for z in (-0.3, -1.7 + 0.4j, 0.2 - 2.1j, -3.0):
        R = 1 + z * b @ np.linalg.solve(np.eye(n * n) - z * A, np.ones(n * n))
        assert abs(R - _ketcheson_recurrence(n, z)) < 1e-13 * max(1, abs(R))


def test_invalid_ssp_stage_count_rejected():
    with pytest.raises(ValueError):
        TS.tableau(TS.RK_SSP_S3, 5)


def _tgv(implicit, stabilization=0, degree=2):
    L = 2 * np.pi
    mesh = M.box_mesh((2, 2, 2), degree, lengths=(L, L, L))
    om = O.Mesh(mesh, M.tabulate(degree, 0))
    c = O.newt_from_dict(PR.tgv_context_dict(state_var=0, stabilization=stabilization, implicit=implicit, dt=1e-2))
    q0 = PR.taylor_green_ic(mesh, 0, background=(0.3, -0.2, 0.1)).ravel()
    return om, c, om.setup_qdata(), q0


def _observed_order(solve, dts):
    sols = [solve(dt) for dt in dts]
    ref = sols[-1]
    errs = [np.linalg.norm(s - ref) for s in sols[:-1]]
    return [np.log2(errs[i] / errs[i + 1]) for i in range(len(errs) - 1)]


@pytest.mark.parametrize("order", [1, 2])
def test_oracle_bdf_convergence_order(order):
    om, c, qd, q0 = _tgv(implicit=True)
    T = 0.2

    def solve(dt):
        q, st = O.bdf(om, c, qd, q0, dt, int(round(T / dt)), order=order, snes_rtol=1e-12, ksp_rtol=1e-12)
        assert st["steps"] == int(round(T / dt)) and abs(st["time"] - T) < 1e-12
        return q

    rates = _observed_order(solve, [0.04, 0.02, 0.01, 0.0025] if order == 2 else [0.02, 0.01, 0.005, 0.000625])
    assert all(abs(r - order) < 0.25 for r in rates), rates


def test_oracle_rk4_lumped_convergence_order():
    om, c, qd, q0 = _tgv(implicit=False)
    A, b, cc = O.tableau_ref("4")
    T = 0.16

    def solve(dt):
        q, _, _ = O.rk(om, c, qd, q0, dt, int(round(T / dt)), A, b, cc, gmres=False, rowsum=True)
        return q

    rates = _observed_order(solve, [0.04, 0.02, 0.01, 0.0025])
    assert all(abs(r - 4) < 0.4 for r in rates), rates


def test_oracle_gmres_mass_solve_is_accurate():
    """With a tight tolerance the GMRES mass solve reproduces the exact consistent-mass solve (dense check)."""
    om, c, qd, q0 = _tgv(implicit=False, stabilization=2)
    n = q0.size
    dt = 1e-3
    c.dt = dt  # Tau (and so the stabilized mass) depends on the time step the integrator sets
    Mdense = np.stack([O.mass_newtonian(om, c, qd, q0, e) for e in np.eye(n)], axis=1)
    r = O.rhs(om, c, qd, q0)
    A, b, cc = O.tableau_ref("euler")
    q1, _, its = O.rk(om, c, qd, q0, dt, 1, A, b, cc, gmres=True, rowsum=False, rtol=1e-13, restart=30)
    q1_exact = q0 + dt * np.linalg.solve(Mdense, r)
    assert its > 0
    assert np.linalg.norm(q1 - q1_exact) < 1e-10 * np.linalg.norm(q1_exact - q0)


def test_options_layout_and_defaults():
    """The ctypes mirror of HoneeMiTSOptions agrees with the C struct: HoneeMiTSOptionsDefault's values arrive in the
    right fields, including the last ones (a missing or misplaced field would shift them)."""
    o = TS.default_options()
    assert o.type == TS.TS_BDF and o.bdf_order == 2 and o.dt == 1e-2
    assert o.snes_rtol == 1e-8 and o.snes_stol == 1e-8 and o.snes_max_it == 50
    assert o.ksp_rtol == 1e-5 and o.ksp_restart == 30 and o.ksp_max_it == 10000
    assert o.pc_lag == 1 and o.snes_mf_operator == 0
    assert o.alpha_radius == 1.0 and o.snes_lag_jacobian == 1 and o.snes_lag_jacobian_persists == 0
    assert o.snes_linesearch_type == 0  # PETSc's SNESLINESEARCHBT, HONEE's default
