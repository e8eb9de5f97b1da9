"""The reference's p=3 fixture tests/output/gaussianwave-explicit.bin, reproduced on the HIP path (the run and the
comparison are described in gaussianwave_case.py).  The comparison pins the p=3 basis (node set, quadrature), the
explicit SUPG residual, the HLLC freestream faces, the damping layer and RK5F; tests/test_oracle_gaussianwave.py
reproduces the same fixture with the oracle."""
import sys

import pytest

import gaussianwave_case as GW
from pyhonee import ts as TS
from pyhonee.driver import NewtonianProblem, context_from_dict

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def test_gaussianwave_explicit_reproduces_reference_fixture():
    _say("gaussian wave explicit p=3: reference fixture")
    mesh, tab, d, (fe, fi), ess, q0 = GW.setup()
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 0)
    prob.add_boundary(prob.FREESTREAM_HLLC, fe, fi, GW.REF_Y)
    prob.set_essential(ess)
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_RK, rk_type=TS.RK_5F, dt=GW.DT, mass_gmres=1, mass_jacobi_rowsum=0, mass_rtol=1e-5,
                                                  mass_restart=30))
    ts.set_solution(q0, 0.0)
    ts.step(GW.STEPS)
    mine = ts.solution()
    st = ts.stats()
    ts.close()
    prob.close()
    err, err0 = GW.compare(mesh, mine, q0, st["time"])
    _say(f"gaussianwave-explicit.bin regression metric {err:.3e} (initial state against it: {err0:.3e}); "
         f"mass GMRES iterations {st['mass_krylov_its']}")
    assert err0 > 1e-4  # the five steps moved the state well beyond the tolerance
    assert err < GW.TOLERANCE
