"""Pinning the CPU oracle's boundary-face operators (consistent boundary integral, freestream HLL/HLLC).

* Riemann solvers: the reference's own finite-difference unit tests (problems/bc_freestream.c:200-535) with their
  fixed states (T=200, rho=1.2, u=(40, 44, 48); right state scaled by 1.2; normal (1, 2, 3)/|.|), FD step 4e-7 and
  tolerance 5e-7 (the value problems/bc_freestream.c:186 passes), for RiemannFlux_HLL_fwd, RiemannFlux_HLLC_fwd and
  ComputeHLLSpeeds_Roe_fwd.
* Textbook properties (Toro, ch. 10): equal states give the exact flux F(U).n for HLL and HLLC; supersonic states
  are fully upwinded; HLLC resolves a stationary contact exactly (zero mass flux, pressure-only momentum flux) where
  HLL diffuses it.
* Boundary geometry: outward unit normals, face areas, and the divergence theorem on trilinear faces
  (closed-surface integral of x n_x equals the mesh volume, exactly for Gauss quadrature).
* Boundary integral vs volume residual: for a uniform state the volume IFunction's nodal residual is
  -(boundary flux) node by node, so volume + boundary integral vanish at every node.
* Jacobians: central finite differences of the face residuals (primitive variables, where the reference's
  linearization is exact).
"""
import numpy as np
import pytest

import pyoracle as O
from helpers import make_case, oracle_setup, rel_err
from pyhonee import mesh as M

EPS = 2.220446049250313e-16
GAS_DEFAULT = {"lambda": -2.0 / 3.0, "mu": 1.8e-5, "k": 0.02638, "cv": 717.0, "cp": 1004.0}


def _gas(d=GAS_DEFAULT):
    return O.OGas(d["lambda"], d["mu"], d["k"], d["cv"], d["cp"])


def _riemann_states(gas=GAS_DEFAULT):
    T, rho = 200.0, 1.2
    gamma = gas["cp"] / gas["cv"]
    p = (gamma - 1) * rho * gas["cv"] * T
    u = np.array([40.0, 44.0, 48.0])
    Y_left = np.array([p, *u, T])
    Y_right = 1.2 * Y_left
    n = np.array([1.0, 2.0, 3.0])
    return Y_left, Y_right, n / np.linalg.norm(n)


def _rel(S, A, B, thr=10 * EPS):
    return (A - B) / (S if abs(S) > thr else 1.0)


def _check_q(Q_s, Q_a, Q_b, rtol):
    """CheckQWithTolerance (problems/bc_freestream.c:205-232)."""
    um = np.linalg.norm(Q_s[1:4])
    errs = [_rel(Q_s[0], Q_a[0], Q_b[0])] + [_rel(um, Q_a[i], Q_b[i]) for i in (1, 2, 3)] + [_rel(Q_s[4], Q_a[4], Q_b[4])]
    assert np.all(np.abs(errs) < rtol), errs


@pytest.mark.parametrize("hllc", [False, True])
def test_riemann_flux_fwd_matches_fd(hllc):
    gas = _gas()
    Y0l, Y0r, n = _riemann_states()
    eps = 4e-7
    for i in range(10):
        dYl, dYr = np.zeros(5), np.zeros(5)
        Y1l, Y1r = Y0l.copy(), Y0r.copy()
        if i < 5:
            dYl[i] = Y0l[i]
            Y1l[i] *= 1 + eps
        else:
            dYr[i % 5] = Y0r[i % 5]
            Y1r[i % 5] *= 1 + eps
        dF = O.riemann_flux_fwd(gas, hllc, Y0l, dYl, Y0r, dYr, 1, n)
        F0 = O.riemann_flux(gas, hllc, Y0l, Y0r, 1, n)
        F1 = O.riemann_flux(gas, hllc, Y1l, Y1r, 1, n)
        dF_fd = (F1 - F0) / eps
        _check_q(dF_fd, dF, dF_fd, 5e-7)


def test_hll_speeds_fwd_matches_fd():
    gas = _gas()
    Y0l, Y0r, n = _riemann_states()
    eps = 4e-7
    sl0, sr0 = O.hll_speeds(gas, Y0l, Y0r, 1, n)
    for i in range(10):
        dYl, dYr = np.zeros(5), np.zeros(5)
        Y1l, Y1r = Y0l.copy(), Y0r.copy()
        if i < 5:
            dYl[i] = Y0l[i]
            Y1l[i] *= 1 + eps
        else:
            dYr[i % 5] = Y0r[i % 5]
            Y1r[i % 5] *= 1 + eps
        _, dsl, _, dsr = O.hll_speeds_fwd(gas, Y0l, dYl, Y0r, dYr, 1, n)
        sl1, sr1 = O.hll_speeds(gas, Y1l, Y1r, 1, n)
        dsl_fd, dsr_fd = (sl1 - sl0) / eps, (sr1 - sr0) / eps
        assert abs(_rel(dsl_fd, dsl, dsl_fd)) < 5e-7
        assert abs(_rel(dsr_fd, dsr, dsr_fd)) < 5e-7


def _exact_normal_flux(Y, n, gas=GAS_DEFAULT):
    gamma = gas["cp"] / gas["cv"]
    p, u, T = Y[0], Y[1:4], Y[4]
    rho = p / ((gas["cp"] - gas["cv"]) * T)
    E = rho * (gas["cv"] * T + 0.5 * u @ u)
    un = u @ n
    return np.array([rho * un, *(rho * u * un + p * n), (E + p) * un]), np.sqrt(gamma * p / rho)


@pytest.mark.parametrize("hllc", [False, True])
@pytest.mark.parametrize("scale_u", [0.0, 0.3, 1.0, 5.0, -5.0])
def test_riemann_equal_states_give_exact_flux(hllc, scale_u):
    gas = _gas()
    Y, _, n = _riemann_states()
    Y = Y.copy()
    Y[1:4] *= scale_u
    F = O.riemann_flux(gas, hllc, Y, Y, 1, n)
    F_exact, _ = _exact_normal_flux(Y, n)
    assert np.allclose(F, F_exact, rtol=1e-12, atol=1e-12 * np.abs(F_exact).max())


@pytest.mark.parametrize("hllc", [False, True])
def test_riemann_supersonic_upwinding(hllc):
    gas = _gas()
    Yl, Yr, n = _riemann_states()
    _, c = _exact_normal_flux(Yl, n)
    for sign in (+1, -1):
        Ya, Yb = Yl.copy(), Yr.copy()
        Ya[1:4] = sign * 4 * c * n
        Yb[1:4] = sign * 4 * c * n
        F = O.riemann_flux(gas, hllc, Ya, Yb, 1, n)
        expected, _ = _exact_normal_flux(Ya if sign > 0 else Yb, n)
        assert np.allclose(F, expected, rtol=1e-12, atol=1e-10 * np.abs(expected).max())


def test_hllc_resolves_stationary_contact():
    gas = _gas()
    Yl, _, n = _riemann_states()
    Yl = Yl.copy()
    Yl[1:4] = 0.0
    Yr = Yl.copy()
    Yr[4] *= 3.0  # same pressure, zero velocity, three times the temperature: a stationary contact
    F_hllc = O.riemann_flux(gas, True, Yl, Yr, 1, n)
    expected = np.array([0.0, *(Yl[0] * n), 0.0])
    assert np.allclose(F_hllc, expected, rtol=0, atol=1e-12 * Yl[0])
    F_hll = O.riemann_flux(gas, False, Yl, Yr, 1, n)
    assert abs(F_hll[0]) > 1e-3  # HLL smears the density jump


def _box_faces(mesh, sides):
    elems, fids = [], []
    for f in sides:
        e = M.boundary_faces(mesh, f)
        elems.append(e)
        fids.append(np.full(len(e), f, dtype=np.int32))
    return O.Faces(np.concatenate(elems), np.concatenate(fids))


def test_boundary_geometry_affine():
    L = (2.0, 3.0, 5.0)
    mesh = M.box_mesh((2, 3, 2), 2, lengths=L, periodic=(False, False, False), perturb=0.0, seed=0)
    tab = M.tabulate(2, 0)
    om = O.Mesh(mesh, tab)
    normals = {0: (-1, 0, 0), 1: (1, 0, 0), 2: (0, -1, 0), 3: (0, 1, 0), 4: (0, 0, -1), 5: (0, 0, 1)}
    for fid in range(6):
        f = _box_faces(mesh, [fid])
        qd = O.setup_qdata_boundary(om, f)
        d = fid // 2
        area = np.prod([L[k] for k in range(3) if k != d])
        assert np.isclose(qd[:, 0, :].sum(), area, rtol=1e-13)
        for k in range(3):
            assert np.allclose(qd[:, 1 + k, :], normals[fid][k], atol=1e-14)


def test_boundary_divergence_theorem_perturbed():
    mesh = M.box_mesh((3, 2, 2), 2, lengths=(1.0, 1.5, 2.0), periodic=(False, False, False), perturb=0.2, seed=5)
    tab = M.tabulate(2, 0)
    om = O.Mesh(mesh, tab)
    qd_vol = om.setup_qdata()
    volume = O.volume(om, qd_vol)
    f = _box_faces(mesh, range(6))
    qd = O.setup_qdata_boundary(om, f)
    # physical coordinates of the face quadrature points: x at qpts from the same trilinear map (numerically, via nodes)
    Q = tab.Q
    xq = _face_qpt_coords(mesh, tab, f)
    for d in range(3):
        flux = (qd[:, 0, :] * qd[:, 1 + d, :] * xq[:, d, :]).sum()
        assert np.isclose(flux, volume, rtol=1e-12), (d, flux, volume)
        assert abs((qd[:, 0, :] * qd[:, 1 + d, :]).sum()) < 1e-12  # closed surface


def _face_qpt_coords(mesh, tab, f):
    """Coordinates of the face quadrature points from the trilinear vertex field (x fastest node ordering)."""
    Q = tab.Q
    out = np.zeros((f.nfaces, 3, Q * Q))
    Bx = tab.Bx  # [Q][2]
    for k in range(f.nfaces):
        e, fid = f.elem[k], f.fid[k]
        xv = mesh.xcoord[mesh.xoffsets[e]].reshape(2, 2, 2, 3)  # [c][b][a][3]
        for qj in range(Q):
            for qi in range(Q):
                x = np.zeros(3)
                for j in range(2):
                    for i in range(2):
                        a, b, c = _face_abc(fid, 2, i, j)
                        x += Bx[qi, i] * Bx[qj, j] * xv[c, b, a]
                out[k, :, qi + Q * qj] = x
    return out


def _face_abc(fid, P, i, j):
    L = P - 1
    return {0: (0, j, i), 1: (L, i, j), 2: (i, 0, j), 3: (j, L, i), 4: (j, i, 0), 5: (i, j, L)}[fid]


def test_boundary_integral_cancels_volume_residual_for_uniform_state():
    """Uniform state: the volume residual at each node is minus the boundary flux at that node."""
    case = make_case(shape=(2, 2, 3), degree=2, periodic=(False, False, False), perturb=0.15, seed=7, state_var=1, stabilization=2)
    om, oc, qd = oracle_setup(case)
    Y = np.array([71.0, 3.0, -2.0, 1.5, 0.25])
    q = np.tile(Y, case["mesh"].nnodes)
    qdot = np.zeros_like(q)
    v_vol, _ = O.ifunction(om, oc, qd, q, qdot)
    f = _box_faces(case["mesh"], range(6))
    qds = O.setup_qdata_boundary(om, f)
    v_bnd, _ = O.boundary_integral(om, f, oc, qds, q)
    scale = np.abs(v_bnd).max()
    assert scale > 1
    assert np.abs(v_vol + v_bnd).max() < 1e-10 * scale


def _fd_check(apply_residual, apply_jac, q, dq, h):
    r_p, r_m = apply_residual(q + h * dq), apply_residual(q - h * dq)
    fd = (r_p - r_m) / (2 * h)
    jv = apply_jac(dq)
    return rel_err(jv, fd)[0]


def test_boundary_integral_jacobian_matches_fd():
    case = make_case(shape=(2, 2, 2), degree=2, periodic=(False, True, False), perturb=0.1, seed=11, state_var=1)
    om, oc, qd = oracle_setup(case)
    f = _box_faces(case["mesh"], [0, 1, 4, 5])
    qds = O.setup_qdata_boundary(om, f)
    q = case["q"]
    _, jac = O.boundary_integral(om, f, oc, qds, q)
    dq = case["rng"].uniform(-1, 1, q.shape) * np.abs(q.reshape(-1, 5)).mean(0).repeat(1).reshape(1, 5).repeat(case["mesh"].nnodes, 0).ravel() * 1e-2
    err = _fd_check(lambda x: O.boundary_integral(om, f, oc, qds, x)[0], lambda d: O.boundary_integral_jacobian(om, f, oc, qds, jac, d), q, dq, 1e-4)
    assert err < 1e-7, err


@pytest.mark.parametrize("hllc", [False, True])
@pytest.mark.parametrize("state_var", [0, 1, 2])
def test_freestream_jacobian_matches_fd(hllc, state_var):
    case = make_case(shape=(2, 2, 2), degree=2, periodic=(False, True, True), perturb=0.1, seed=13, state_var=state_var)
    om, oc, qd = oracle_setup(case)
    f = _box_faces(case["mesh"], [0, 1])
    qds = O.setup_qdata_boundary(om, f)
    Y_inf = np.array([71.4, 0.9, 0.05, -0.02, 0.25])
    U_inf, _ = O.state_from_q(oc.gas, Y_inf, 1)
    fc = O.freestream_ctx(oc, U_inf, Y_inf, hllc)
    q = case["q"]
    _, jac = O.freestream(om, f, fc, qds, q)
    assert np.allclose(jac.reshape(f.nfaces, 5, -1).std(), jac.std())  # stored interior states
    rng = np.random.default_rng(3)
    dq = rng.uniform(-1, 1, q.shape) * np.abs(q.reshape(-1, 5)).mean(0)[None, :].repeat(case["mesh"].nnodes, 0).ravel() * 1e-3
    err = _fd_check(lambda x: O.freestream(om, f, fc, qds, x)[0], lambda d: O.freestream_jacobian(om, f, fc, qds, jac, d), q, dq, 1e-3)
    assert err < 1e-6, err


def test_explicit_boundary_residual_is_negated():
    case = make_case(shape=(2, 2, 2), degree=2, periodic=(True, True, False), seed=17, state_var=0)
    om, oc, qd = oracle_setup(case)
    f = _box_faces(case["mesh"], [4, 5])
    qds = O.setup_qdata_boundary(om, f)
    v_impl, _ = O.boundary_integral(om, f, oc, qds, case["q"])
    oc.is_implicit = 0
    v_expl, _ = O.boundary_integral(om, f, oc, qds, case["q"])
    assert np.array_equal(v_expl, -v_impl)
