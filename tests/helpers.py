"""Shared builders for the parity tests: meshes, contexts and seeded states, expressed for both the MI355X driver
and the CPU oracle (which is only used here, as the checker)."""
from __future__ import annotations

import numpy as np

import pyoracle as O
from pyhonee import mesh as M
from pyhonee import problems as PR
from pyhonee.driver import context_from_dict


def make_case(shape=(3, 2, 2), degree=3, periodic=(True, True, True), perturb=0.15, seed=1, state_var=1, stabilization=2,
              idl=False, g=(0.0, 0.0, 0.0), noise=0.05, dt=1e-3, tau=None, lengths=(2 * np.pi, 2 * np.pi, 2 * np.pi),
              implicit=True):
    """A seeded, non-trivial Newtonian state on a (possibly non-affine) box mesh."""
    mesh = M.box_mesh(shape, degree, lengths=lengths, periodic=periodic, perturb=perturb, seed=seed)
    tab = M.tabulate(degree, 0)
    rng = np.random.default_rng(seed)
    d = PR.tgv_context_dict(state_var=state_var, stabilization=stabilization, implicit=implicit, dt=dt, tau=tau or PR.default_tau(degree))
    d["g"] = g
    d["ijacobian_time_shift"] = 0.0
    if idl:
        d.update(idl_enable=True, idl_pressure=71.0, idl_amplitude=1.0 / 2e-3, idl_start=0.5, idl_length=2.0)
    # primitive TGV state plus seeded perturbations, then converted to the requested variables
    xyz = mesh.element_node_coords()
    Y = PR.nodal_field(mesh, lambda x: PR.taylor_green_prim(x))
    Y = Y * (1 + noise * rng.uniform(-1, 1, Y.shape))
    q = PR.q_from_prim(PR.TGV_GAS, Y, state_var)
    qdot = rng.uniform(-1, 1, q.shape) * np.abs(q).mean(axis=0) * 0.1
    return dict(mesh=mesh, tab=tab, ctx=d, q=np.ascontiguousarray(q.ravel()), qdot=np.ascontiguousarray(qdot.ravel()), rng=rng, xyz=xyz)


def oracle_setup(case):
    om = O.Mesh(case["mesh"], case["tab"])
    oc = O.newt_from_dict(case["ctx"])
    qd = om.setup_qdata()
    return om, oc, qd


def rel_err(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)), float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def product_context(case):
    return context_from_dict(case["ctx"])
