"""The backtracking line search of the Newton solves (SNESLINESEARCHBT with the cubic order, as HONEE's SNESNEWTONLS
uses it; PETSc's SNESLineSearchApply_BT).  No reference fixture backtracks, so the step is pinned by what the
mathematics fixes:

* after the first rejected trial the next lambda minimizes the quadratic through f(0), f'(0) and f(lambda), so for an
  exactly quadratic merit function f it is the true minimizer;
* after later rejections it minimizes the cubic through f(0), f'(0) and the last two trials, so for an exactly cubic
  f it is the cubic's local minimizer (the root of f' where f'' > 0);
* both are clamped to [0.1, 0.5] of the rejected lambda.

f(l) = ||F(x - l dx)||^2 / 2, so f(0) = fnorm^2 / 2 and a trial's f = gnorm^2 / 2.  The oracle's BDF solve on an
inexact-Newton case that backtracks converges to the tightly solved state.
"""
import numpy as np
import pytest

import pyoracle as O
from test_gpu_ts import _tgv_case


def _norm_of(m):
    return float(np.sqrt(2.0 * m))


@pytest.mark.parametrize("lam_star", [0.13, 0.3, 0.47])
def test_quadratic_fit_is_exact_for_a_quadratic_merit_function(lam_star):
    m0, s = 0.7, -1.3  # f(0), f'(0)
    c = -s / (2.0 * lam_star)  # f(l) = m0 + s l + c l^2 has its minimum at lam_star
    lam = 1.0
    got = O.bt_next_lambda(_norm_of(m0), s, lam, _norm_of(m0 + s * lam + c * lam * lam), 0)
    assert abs(got - lam_star) < 1e-13


@pytest.mark.parametrize("lam_star,expect", [(0.02, 0.1), (0.9, 0.5)])
def test_quadratic_step_is_clamped(lam_star, expect):
    m0, s = 0.7, -1.3
    c = -s / (2.0 * lam_star)
    got = O.bt_next_lambda(_norm_of(m0), s, 1.0, _norm_of(m0 + s + c), 0)
    assert abs(got - expect) < 1e-15


@pytest.mark.parametrize("a,b,lam,lamprev", [(10.0, 2.0, 0.4, 1.0), (20.0, -0.5, 0.4, 0.6), (40.0, 1.0, 0.3, 0.5)])
def test_cubic_fit_is_exact_for_a_cubic_merit_function(a, b, lam, lamprev):
    m0, s = 0.9, -1.1
    f = lambda t: m0 + s * t + b * t * t + a * t ** 3  # noqa: E731
    lam_star = (-b + np.sqrt(b * b - 3.0 * a * s)) / (3.0 * a)  # f'(lam_star) = 0, f''(lam_star) > 0
    assert abs(s + 2 * b * lam_star + 3 * a * lam_star ** 2) < 1e-12 and 2 * b + 6 * a * lam_star > 0
    assert 0.1 * lam < lam_star < 0.5 * lam, "the case must lie inside the clamp"
    got = O.bt_next_lambda(_norm_of(m0), s, lam, _norm_of(f(lam)), 1, lamprev, _norm_of(f(lamprev)))
    assert abs(got - lam_star) < 1e-11 * max(1.0, lam_star)


def test_cubic_step_is_clamped():
    m0, s, a, b = 0.9, -1.1, 10.0, 2.0
    f = lambda t: m0 + s * t + b * t * t + a * t ** 3  # noqa: E731
    lam_star = (-b + np.sqrt(b * b - 3.0 * a * s)) / (3.0 * a)  # 0.128
    lam = 0.2  # lam_star > 0.5 lam: clamped to 0.1
    got = O.bt_next_lambda(_norm_of(m0), s, lam, _norm_of(f(lam)), 2, 0.5, _norm_of(f(0.5)))
    assert lam_star > 0.5 * lam and abs(got - 0.5 * lam) < 1e-15
    lam = 2.0  # lam_star < 0.1 lam: raised to 0.1 lam
    got = O.bt_next_lambda(_norm_of(m0), s, lam, _norm_of(f(lam)), 2, 3.0, _norm_of(f(3.0)))
    assert lam_star < 0.1 * lam and abs(got - 0.1 * lam) < 1e-15


def test_backtracking_bdf_converges_to_the_tight_solution():
    """Inexact Newton (GMRES capped at 5 iterations, rtol 0.7) on a TGV case with a noisy velocity: the full Newton
    step often fails the sufficient-decrease test.  The solve still converges to the state of a tight solve."""
    dt = 0.2
    mesh, tab, d, q0, om, oc, qd = _tgv_case((2, 2, 2), 2, 1, 2, True, dt, perturb=0.08, noise=0.3)
    q, st = O.bdf(om, oc, qd, q0, dt, 3, order=2, pbjacobi=True, snes_rtol=1e-10, ksp_rtol=0.7, ksp_restart=5, ksp_max_it=5)
    q_tight, st_tight = O.bdf(om, oc, qd, q0, dt, 3, order=2, pbjacobi=True, snes_rtol=1e-10, ksp_rtol=1e-9)
    e = float(np.abs(q - q_tight).max() / np.abs(q_tight).max())
    print(f"backtracks {st['line_search_backtracks']}, Newton {st['newton_its']} (tight: {st_tight['newton_its']}), rel diff {e:.2e}")
    assert st["line_search_backtracks"] > 0 and st_tight["line_search_backtracks"] == 0
    # the step-length test (||lambda dx|| <= 1e-8 ||x||) ends the inexact solves with short steps a little earlier
    assert e < 1e-6
