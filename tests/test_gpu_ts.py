"""The GPU time integrator (libhonee_ts.so) against the reference's fixtures and the CPU oracle.

* Reference pin: the explicit p=2 Taylor-Green run (RK4, SUPG, consistent mass by GMRES with diagonal Jacobi,
  examples/navierstokes.c:222) on the GPU reproduces tests/output/taylorgreen_totalke.csv at the reference's CSV
  tolerance rtol 1e-9 (tests/junit.py:53) and its final state taylor-green.bin with RegressionTest's metric
  max|Q - Qref| / max|Qref| < 1e-10 (src/misc.c:98-126), node order matched by a minimum-cost assignment.
* Oracle parity of time-integrated fields (the same algorithms on the CPU): explicit schemes to 1e-11 relative after
  several steps; BDF with Newton-GMRES to 1e-9 relative with tight solver tolerances (the GPU's reductions differ from
  the oracle's in rounding, which the iterative solves carry to their tolerance); the FP32-Jacobian variant against
  the all-FP64 oracle to 1e-7 relative (its Newton iteration recovers the FP64 residual's accuracy), with the error
  printed.
"""
import os
import sys

import numpy as np
import pytest
from scipy.optimize import linear_sum_assignment

import pyoracle as O
from helpers import product_context
from pyhonee import honee_file
from pyhonee import mesh as M
from pyhonee import problems as PR
from pyhonee import ts as TS
from pyhonee.driver import NewtonianProblem, context_from_dict

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))


def test_rk4_consistent_mass_reproduces_reference_ke_and_state():
    _say("ts rk4 taylor-green fixture")
    L = 2 * np.pi
    mesh = M.box_mesh((2, 2, 2), 2, lengths=(L, L, L))
    tab = M.tabulate(2, 0)
    d = PR.tgv_context_dict(state_var=0, stabilization=2, implicit=False, dt=2e-3)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 0)
    opts = TS.default_options(type=TS.TS_RK, rk_type=TS.RK_4, dt=2e-3, mass_gmres=1, mass_jacobi_rowsum=0, mass_rtol=1e-5, mass_restart=30)
    ts = TS.TimeStepper(prob, opts)
    q0 = PR.taylor_green_ic(mesh, 0).ravel()
    ts.set_solution(q0, 0.0)
    vol = prob.volume
    rows = [[0, 0.0, *(ts.monitor_ke() / vol)]]
    ts.step(2)
    rows.append([2, 4e-3, *(ts.monitor_ke() / vol)])
    ts.step(1)
    rows.append([3, 6e-3, *(ts.monitor_ke() / vol)])
    rows = np.array(rows)
    lines = [l for l in open(os.path.join(GOLDEN, "taylorgreen_totalke.csv")) if l.strip() and not l.startswith("#")]
    gold = np.array([[float(v) for v in l.strip().split(",")] for l in lines[1:]])
    assert rows.shape == gold.shape
    np.testing.assert_allclose(rows[:, 1], gold[:, 1], rtol=1e-12)
    for col in (2, 3, 6):
        np.testing.assert_allclose(rows[:, col], gold[:, col], rtol=1e-9)
    for col in (4, 5):
        np.testing.assert_allclose(rows[:, col], gold[:, col], atol=1e-14, rtol=0)
    st = ts.stats()
    assert st["steps"] == 3 and st["mass_krylov_its"] > 0
    g, step, t = honee_file.read_vec(os.path.join(GOLDEN, "taylor-green.bin"))
    g = g.reshape(-1, 5)
    mine = ts.solution().reshape(-1, 5)
    scale = np.abs(g).max()
    cost = np.abs(g[:, None, :] - mine[None, :, :]).max(-1) / scale
    r, c = linear_sum_assignment(cost)
    err = np.abs(g[r] - mine[c]).max() / scale
    print(f"taylor-green.bin regression metric {err:.3e}", flush=True)
    assert err < 1e-10
    ts.close()
    prob.close()


def _tgv_case(shape, degree, state_var, stabilization, implicit, dt, perturb=0.0, seed=5, noise=1e-3):
    L = 2 * np.pi
    mesh = M.box_mesh(shape, degree, lengths=(L, L, L), perturb=perturb, seed=seed)
    tab = M.tabulate(degree, 0)
    d = PR.tgv_context_dict(state_var=state_var, stabilization=stabilization, implicit=implicit, dt=dt, tau=PR.default_tau(degree))
    rng = np.random.default_rng(seed)
    Y = PR.nodal_field(mesh, lambda x: PR.taylor_green_prim(x))
    Y[:, 1:4] += noise * rng.uniform(-1, 1, Y[:, 1:4].shape)
    q0 = PR.q_from_prim(PR.TGV_GAS, Y, state_var).ravel()
    om = O.Mesh(mesh, tab)
    return mesh, tab, d, q0, om, O.newt_from_dict(d), om.setup_qdata()


@pytest.mark.parametrize("rk_type,stages,gmres,rowsum,stab", [(TS.RK_SSP_S3, 4, 0, 1, 0), (TS.RK_5F, 6, 1, 0, 2), (TS.RK_SSP33, 3, 0, 0, 1)])
def test_explicit_schemes_match_oracle(rk_type, stages, gmres, rowsum, stab):
    _say(f"ts explicit rk={rk_type} stages={stages} gmres={gmres} rowsum={rowsum} stab={stab}")
    dt = 5e-3
    mesh, tab, d, q0, om, oc, qd = _tgv_case((3, 2, 2), 2, 0, stab, False, dt, perturb=0.1)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 0)
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_RK, rk_type=rk_type, ssp_stages=stages, dt=dt, mass_gmres=gmres,
                                                  mass_jacobi_rowsum=rowsum, mass_rtol=1e-10))
    ts.set_solution(q0, 0.1)
    ts.step(5)
    A, b, cc = TS.tableau(rk_type, stages)
    q_ref, _, _ = O.rk(om, oc, qd, q0, dt, 5, A, b, cc, t0=0.1, gmres=bool(gmres), rowsum=bool(rowsum), rtol=1e-10)
    e = _rel(ts.solution(), q_ref)
    print(f"explicit rel err after 5 steps {e:.3e}", flush=True)
    assert e < 1e-11
    assert abs(ts.stats()["time"] - 0.125) < 1e-14
    ts.close()
    prob.close()


@pytest.mark.parametrize("linesearch", ["bt", "basic", "norm_slope"])
def test_backtracking_line_search_matches_oracle(linesearch):
    """Inexact Newton (GMRES capped at 5 iterations, rtol 0.7) on a TGV case with a noisy velocity, where the full
    Newton step often fails the sufficient-decrease test: the GPU integrator's line search (cubic backtracking with
    the slope F . (J dx), as PETSc's SNESLINESEARCHBT) backtracks like the oracle's and ends at its state.  With
    SNESLINESEARCHBASIC both take full steps; with quadratic fits and the slope -||F||^2 both backtrack alike."""
    _say(f"ts bdf2 inexact Newton, {linesearch} line search")
    basic = linesearch == "basic"
    ls_type = {"bt": 0, "basic": 1, "norm_slope": 2}[linesearch]
    dt = 0.2
    mesh, tab, d, q0, om, oc, qd = _tgv_case((2, 2, 2), 2, 1, 2, True, dt, perturb=0.08, noise=0.3)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1)
    kw = dict(snes_rtol=1e-10, ksp_rtol=0.7, ksp_restart=5, ksp_max_it=5)
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_BDF, bdf_order=2, dt=dt, ksp_pc=TS.PC_PBJACOBI, snes_linesearch_type=ls_type,
                                                  **kw))
    ts.set_solution(q0, 0.0)
    ts.step(3)
    q_ref, st_ref = O.bdf(om, oc, qd, q0, dt, 3, order=2, pbjacobi=True, linesearch=linesearch, **kw)
    st = ts.stats()
    e = _rel(ts.solution(), q_ref)
    print(f"{linesearch} bdf2: rel err {e:.3e}; newton gpu {st['newton_its']} oracle {st_ref['newton_its']}; backtracks gpu "
          f"{st['line_search_backtracks']} oracle {st_ref['line_search_backtracks']}", flush=True)
    if basic:
        assert st["line_search_backtracks"] == 0 and st_ref["line_search_backtracks"] == 0
    else:
        assert st_ref["line_search_backtracks"] > 0 and st["line_search_backtracks"] > 0
        assert abs(st["line_search_backtracks"] - st_ref["line_search_backtracks"]) <= 2
    assert abs(st["newton_its"] - st_ref["newton_its"]) <= 3
    assert e < 1e-6  # both end within the step-length test's 1e-8 ||x|| of the same solution
    ts.close()
    prob.close()


@pytest.mark.parametrize("order,state_var,degree,pc", [(2, 1, 2, TS.PC_PBJACOBI), (1, 0, 2, TS.PC_JACOBI), (2, 2, 1, TS.PC_PBJACOBI),
                                                       (2, 0, 3, TS.PC_JACOBI)])
def test_bdf_newton_gmres_matches_oracle(order, state_var, degree, pc):
    """Point-Jacobi for conservative variables; primitive and entropy variables use the point-block Jacobi
    (PCVPBJACOBI) -- with point-Jacobi their preconditioned Jacobian is indefinite (the energy/temperature diagonal
    entry changes sign) and GMRES(30) stagnates."""
    _say(f"ts bdf order={order} var={state_var} degree={degree} pc={pc}")
    dt = 2e-2
    mesh, tab, d, q0, om, oc, qd = _tgv_case((2, 2, 3), degree, state_var, 2, True, dt, perturb=0.08)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), state_var)
    kw = dict(snes_rtol=1e-10, ksp_rtol=1e-9, ksp_restart=30)
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_BDF, bdf_order=order, dt=dt, ksp_pc=pc, **kw))
    ts.set_solution(q0, 0.0)
    ts.step(3)
    q_ref, st_ref = O.bdf(om, oc, qd, q0, dt, 3, order=order, pbjacobi=pc == TS.PC_PBJACOBI, **kw)
    st = ts.stats()
    e = _rel(ts.solution(), q_ref)
    print(f"bdf{order} rel err {e:.3e}; newton its gpu {st['newton_its']} oracle {st_ref['newton_its']}; "
          f"krylov its gpu {st['krylov_its']} oracle {st_ref['krylov_its']}", flush=True)
    assert st["steps"] == 3
    assert abs(st["newton_its"] - st_ref["newton_its"]) <= 1
    assert e < 1e-9
    ts.close()
    prob.close()


def test_bdf2_fp32_jacobian_against_fp64_oracle():
    _say("ts bdf2 fp32 jacobian")
    dt = 2e-2
    mesh, tab, d, q0, om, oc, qd = _tgv_case((2, 3, 2), 3, 1, 2, True, dt, perturb=0.05)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1, "/gpu/hip/mi355x:jacobian_precision=fp32")
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_BDF, bdf_order=2, dt=dt, snes_rtol=1e-10, ksp_pc=TS.PC_PBJACOBI))
    ts.set_solution(q0, 0.0)
    ts.step(3)
    q_ref, st_ref = O.bdf(om, oc, qd, q0, dt, 3, order=2, snes_rtol=1e-10, pbjacobi=True)
    st = ts.stats()
    e = _rel(ts.solution(), q_ref)
    print(f"bdf2 fp32-jacobian vs fp64 oracle: rel err {e:.3e}; newton its {st['newton_its']} (fp64 oracle {st_ref['newton_its']}); "
          f"final ||F|| {st['last_fnorm']:.3e} from {st['last_fnorm0']:.3e}", flush=True)
    assert e < 1e-7
    ts.close()
    prob.close()


def test_bdf2_matrix_free_operator_against_oracle():
    """-snes_mf_operator: GMRES applies the Jacobian as a Walker-Pernice finite difference of the residual while the
    preconditioner still comes from the assembled point blocks.  Newton converges to the same BDF2 solution as the exact
    Jacobian (the oracle's) within its tolerance; the differenced products cost one residual each."""
    _say("ts bdf2 matrix-free operator")
    dt = 2e-2
    mesh, tab, d, q0, om, oc, qd = _tgv_case((2, 2, 3), 2, 1, 2, True, dt, perturb=0.08)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1)
    kw = dict(snes_rtol=1e-10, ksp_rtol=1e-9, ksp_restart=30)
    opts = TS.default_options(type=TS.TS_BDF, bdf_order=2, dt=dt, ksp_pc=TS.PC_PBJACOBI, **kw)
    opts.snes_mf_operator = 1
    ts = TS.TimeStepper(prob, opts)
    ts.set_solution(q0, 0.0)
    ts.step(3)
    q_ref, st_ref = O.bdf(om, oc, qd, q0, dt, 3, order=2, pbjacobi=True, **kw)
    st = ts.stats()
    e = _rel(ts.solution(), q_ref)
    print(f"bdf2 matrix-free operator: rel err {e:.3e}; newton its {st['newton_its']} (oracle {st_ref['newton_its']}); "
          f"krylov its {st['krylov_its']} (oracle {st_ref['krylov_its']}); final ||F|| {st['last_fnorm']:.3e} from "
          f"{st['last_fnorm0']:.3e}", flush=True)
    assert st["steps"] == 3
    assert st["last_fnorm"] <= 1e-10 * st["last_fnorm0"]
    assert e < 1e-8
    ts.close()
    prob.close()


def test_alpha_integrator_reproduces_gaussianwave_shell_fixture():
    """The GPU integrator's TSALPHA (radius 0.5, with its backward-Euler restart) with Newton rtol 1e-4, GMRES rtol 1e-2
    and point-block Jacobi from the persistent lagged Jacobian (-snes_lag_jacobian 20 -snes_lag_jacobian_persists, the
    operator keeping the shift of the last evaluation as HONEE's MATCEED does) reproduces the reference's p=3 fixture
    tests/output/gaussianwave-shell.bin (examples/navierstokes.c:240) within its 1e-8."""
    import gaussianwave_case as GW
    _say("ts alpha: gaussianwave-shell fixture")
    mesh, tab, d, (fe, fi), ess, q0, free = GW.setup_shell()
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1)
    prob.add_boundary(prob.FREESTREAM_HLLC, fe, fi, GW.REF_Y)
    prob.set_essential(ess, q0[ess])
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_ALPHA, alpha_radius=GW.SHELL_RADIUS, dt=GW.DT, snes_rtol=1e-4, ksp_rtol=1e-2,
                                                  ksp_pc=TS.PC_PBJACOBI, snes_lag_jacobian=20, snes_lag_jacobian_persists=1))
    ts.set_solution(q0, 0.0)
    ts.step(GW.STEPS)
    q, st = ts.solution(), ts.stats()
    ts.close()
    prob.close()
    err, err0 = GW.compare_shell(mesh, q, q0)
    print(f"gaussianwave-shell.bin through the GPU integrator: metric {err:.3e} (initial {err0:.1e}); Newton {st['newton_its']}, "
          f"GMRES {st['krylov_its']}, preconditioner builds {st['diagonal_assemblies']}", flush=True)
    assert abs(st["time"] - GW.STEPS * GW.DT) < 1e-15
    assert st["diagonal_assemblies"] == 2  # Newton iterations 0 and 20 across the restart's and the steps' solves
    assert err0 > 1e-3
    assert err < GW.TOLERANCE


def test_beuler_integrator_reproduces_channel_prim_fixture():
    """The GPU integrator's TSBEULER at PETSc's default tolerances with point-block Jacobi reproduces the reference's
    tests/output/channel-prim.bin (examples/navierstokes.c:253; p=1 with q_extra 2, body force, essential walls)
    within its 2e-11.  The reference preconditions with ILU(0); the conservative fixture channel.bin is reproduced with
    it (tests/test_gpu_golden_channel.py), and with point-block Jacobi misses it at 1.8e-10."""
    import channel_case as CH
    _say("ts beuler: channel-prim fixture")
    mesh, tab, d, ess, q0, free, _ = CH.setup(1)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1)
    prob.set_essential(ess, q0[ess])
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_BEULER, dt=CH.DT, ksp_pc=TS.PC_PBJACOBI))
    ts.set_solution(q0, 0.0)
    ts.step(CH.STEPS)
    q, st = ts.solution(), ts.stats()
    ts.close()
    prob.close()
    ref = CH.read(1)
    err = float(np.abs(q[free] - ref).max() / np.abs(ref).max())
    print(f"channel-prim.bin through the GPU integrator: metric {err:.3e}; Newton {st['newton_its']}, GMRES {st['krylov_its']}", flush=True)
    assert err < CH.TOLERANCE


def test_mismatched_integrator_type_is_rejected():
    _say("ts type mismatch")
    mesh, tab, d, q0, om, oc, qd = _tgv_case((2, 2, 2), 1, 0, 0, True, 1e-2)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 0)
    with pytest.raises(RuntimeError, match="does not match"):
        TS.TimeStepper(prob, TS.default_options(type=TS.TS_RK))
    prob.close()
