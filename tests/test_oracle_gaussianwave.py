"""The reference's p=3 fixture tests/output/gaussianwave-explicit.bin reproduced by the oracle alone (no GPU): the run
of gaussianwave_case.py with the oracle's RHS (RHSFunction_Newtonian with SUPG and the damping layer), its HLLC
freestream faces, the essential z-momentum and its RK5F with the GMRES consistent-mass solve.  This pins the oracle's
p=3 basis, explicit SUPG residual, HLLC faces, damping layer and RK5F against the reference at its own tolerance."""
import numpy as np

import gaussianwave_case as GW
import pyoracle as O
from pyhonee import mesh as M


def _run(hllc=True, idl=True):
    mesh, tab, d, (fe, fi), ess, q0 = GW.setup()
    om = O.Mesh(mesh, tab)
    oc = O.newt_from_dict(d)
    oc.idl_enable = int(idl)
    qd = om.setup_qdata()
    f = O.Faces(fe, fi)
    qds = O.setup_qdata_boundary(om, f)
    U_inf, Y_inf = O.state_from_q(oc.gas, GW.REF_Y, 1)
    fc = O.freestream_ctx(oc, U_inf, Y_inf, hllc)
    A, b, cc = O.tableau_ref("5f")
    q, _, its = O.rk(om, oc, qd, q0, GW.DT, GW.STEPS, A, b, cc, gmres=True, rowsum=False, rtol=1e-5, restart=30, faces=f, freestream=fc,
                     qdata_sur=qds, essential=(ess, np.zeros(len(ess))))
    return GW.compare(mesh, q, q0, GW.STEPS * GW.DT) + (its,)


def test_oracle_reproduces_gaussianwave_explicit_fixture():
    err, err0, its = _run()
    assert its > 0
    assert err0 > 1e-4
    assert err < GW.TOLERANCE, err  # measured 3.5e-9


def test_fixture_rejects_plausible_mistakes():
    """The pin bites: the HLL flux in place of HLLC (1.4e-3) and the run without the damping layer (2.8e-2) both fail
    the fixture's tolerance."""
    assert _run(hllc=False)[0] > 100 * GW.TOLERANCE
    assert _run(idl=False)[0] > 100 * GW.TOLERANCE
