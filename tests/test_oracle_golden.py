"""The CPU oracle against the reference's own golden fixtures (copied under tests/golden/).

* taylorgreen_totalke.csv: kinetic-energy budget of the explicit p=2 Taylor-Green run on 2^3 periodic elements with
  RK4, SUPG and the stabilized consistent mass solved by GMRES (examples/navierstokes.c:222), compared at the
  reference's CSV tolerance rtol 1e-9 (tests/junit.py:53).  The two roundoff-level columns (DivergenceDissipation,
  VolumeExpansion ~1e-16..1e-35) are compared with an absolute tolerance.
* taylor-green.bin: the final state of the same run, in PETSc DMPlex order.  Node order is matched by a minimum-cost
  assignment (the comparison is permutation invariant) and checked with RegressionTest's metric max|Q - Qref|/max|Qref|
  against its atol 1e-10 (src/misc.c:98-126).
* taylor-green-IC.bin: the default-gas uniform Taylor-Green IC with background velocity (1,1,1) on 6^3 p=1 elements.
"""
import os

import numpy as np
import pytest
from scipy.optimize import linear_sum_assignment

import pyoracle as O
from pyhonee import honee_file
from pyhonee import mesh as M
from pyhonee import problems as PR

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _read_ke_csv():
    rows = [l for l in open(os.path.join(GOLDEN, "taylorgreen_totalke.csv")) if l.strip() and not l.startswith("#")]
    header = rows[0].strip().split(",")
    data = np.array([[float(v) for v in r.strip().split(",")] for r in rows[1:]])
    return header, data


@pytest.fixture(scope="module")
def tgv_run():
    L = 6.283185307179586
    mesh = M.box_mesh((2, 2, 2), 2, lengths=(L, L, L))
    tab = M.tabulate(2, 0)
    om = O.Mesh(mesh, tab)
    ctx = PR.tgv_context_dict(state_var=0, stabilization=2, implicit=False, dt=2e-3)
    c = O.newt_from_dict(ctx)
    ref_Y = PR.TGV_REFERENCE_Y
    q0 = PR.nodal_field(mesh, lambda x: O.ic_taylorgreen(c, ref_Y, [0, 0, 0], [L, L, L], x))
    qd = om.setup_qdata()
    qf, rows, its = O.rk4_stabilized(om, c, qd, q0.ravel(), 2e-3, 3, rtol=1e-5, restart=30, record_every=2)
    return dict(q0=q0, qf=qf.reshape(-1, 5), rows=rows, its=its)


def test_ke_csv_matches_reference(tgv_run):
    header, gold = _read_ke_csv()
    assert header[:3] == ["step", "time", "TotalKineticEnergy"]
    rows = tgv_run["rows"]
    assert rows.shape == gold.shape
    np.testing.assert_array_equal(rows[:, 0], gold[:, 0])
    np.testing.assert_allclose(rows[:, 1], gold[:, 1], rtol=1e-12)
    for col in (2, 3, 6):  # TotalKineticEnergy, StrainDissipation, MuVorticity2
        np.testing.assert_allclose(rows[:, col], gold[:, col], rtol=1e-9, err_msg=header[col])
    for col in (4, 5):  # roundoff-level columns
        np.testing.assert_allclose(rows[:, col], gold[:, col], atol=1e-14, rtol=0, err_msg=header[col])


def test_ke_csv_detects_wrong_tau(tgv_run):
    """A plausible slip (Ctau_M of the default instead of the YAML's value) must break the KE pin."""
    L = 6.283185307179586
    mesh = M.box_mesh((2, 2, 2), 2, lengths=(L, L, L))
    om = O.Mesh(mesh, M.tabulate(2, 0))
    ctx = PR.tgv_context_dict(state_var=0, stabilization=2, implicit=False, dt=2e-3)
    ctx["tau"]["Ctau_M"] = 0.125
    c = O.newt_from_dict(ctx)
    qd = om.setup_qdata()
    _, rows, _ = O.rk4_stabilized(om, c, qd, tgv_run["q0"].ravel(), 2e-3, 3, rtol=1e-5, restart=30, record_every=2)
    _, gold = _read_ke_csv()
    assert np.max(np.abs(rows[1:, 3] - gold[1:, 3]) / np.abs(gold[1:, 3])) > 1e-9


def test_final_state_matches_reference_bin(tgv_run):
    g, step, t = honee_file.read_vec(os.path.join(GOLDEN, "taylor-green.bin"))
    assert step == 3 and abs(t - 0.006) < 1e-15
    g = g.reshape(-1, 5)
    mine = tgv_run["qf"]
    assert g.shape == mine.shape
    scale = np.abs(g).max()
    cost = np.abs(g[:, None, :] - mine[None, :, :]).max(-1) / scale
    r, c = linear_sum_assignment(cost)
    err = np.abs(g[r] - mine[c]).max() / scale
    assert err < 1e-10, err


def test_uniform_ic_bin():
    vals, step, _ = honee_file.read_vec(os.path.join(GOLDEN, "taylor-green-IC.bin"))
    vals = vals.reshape(-1, 5)
    assert vals.shape == (343, 5)
    gas = {"lambda": -2.0 / 3.0, "mu": 1.8e-5, "k": 0.02638, "cv": 717.0, "cp": 1004.0}  # problems/newtonian.c:592-598
    c = O.newt_from_dict({"gas": gas, "tau": PR.default_tau(1), "state_var": 0})
    q = O.ic_taylorgreen(c, [1.01e5, 0, 0, 0, 288.15], [1, 1, 1], [1, 1, 1], np.random.default_rng(0).uniform(0, 1, (343, 3)))
    np.testing.assert_allclose(q, vals, rtol=1e-13)
    # the product-side numpy IC agrees with the oracle's
    qp = PR.q_from_prim(gas, PR.taylor_green_prim(np.zeros((1, 3)), (1.01e5, 0, 0, 0, 288.15), gas, (1, 1, 1), 1.0), 0)
    np.testing.assert_allclose(qp[0], vals[0], rtol=1e-13)


def test_product_ic_matches_oracle_ic():
    L = 2 * np.pi
    mesh = M.box_mesh((2, 3, 2), 3, lengths=(L, L, L), perturb=0.1, seed=3)
    for var in (0, 1, 2):
        c = O.newt_from_dict(PR.tgv_context_dict(state_var=var))
        qo = PR.nodal_field(mesh, lambda x: O.ic_taylorgreen(c, PR.TGV_REFERENCE_Y, [0.1, 0.2, 0.3], [L, L, L], x))
        qp = PR.taylor_green_ic(mesh, var, background=(0.1, 0.2, 0.3))
        np.testing.assert_allclose(qp, qo, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("name", ["taylor-green.bin", "taylor-green-IC.bin", "gaussianwave-explicit.bin"])
def test_writer_round_trips_reference_bin(name, tmp_path):
    """HoneeWriteBinaryVec's layout (src/honee-file.c:196-217): writing back what the reader returns reproduces the
    reference's file byte for byte (HONEE token, PetscInt step, float64 time, PETSc Vec record; the gaussian-wave file
    comes from a PETSc with 64-bit PetscInt, the Taylor-Green files from one with 32-bit)."""
    path = os.path.join(GOLDEN, name)
    vals, step, t = honee_file.read_vec(path)
    out = tmp_path / name
    honee_file.write_vec(str(out), vals, step=step, time=t, index_bits=honee_file.index_bits(path))
    assert out.read_bytes() == open(path, "rb").read()
