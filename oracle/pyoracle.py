"""ctypes binding of the CPU oracle (liboracle.so) -- test infrastructure only.

Only the tests, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` load this module; the product
path never does.  See honee_oracle.h for layouts.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_dp = ct.POINTER(ct.c_double)
_ip = ct.POINTER(ct.c_int)


class OGas(ct.Structure):
    _fields_ = [("lambda_", ct.c_double), ("mu", ct.c_double), ("k", ct.c_double), ("cv", ct.c_double), ("cp", ct.c_double)]


class OTauCoef(ct.Structure):
    _fields_ = [(n, ct.c_double) for n in ("Ctau_t", "Ctau_v", "Ctau_C", "Ctau_M", "Ctau_E")]


class ONewt(ct.Structure):
    _fields_ = [
        ("gas", OGas),
        ("g", ct.c_double * 3),
        ("time", ct.c_double),
        ("ijacobian_time_shift", ct.c_double),
        ("idl_pressure", ct.c_double),
        ("idl_amplitude", ct.c_double),
        ("idl_start", ct.c_double),
        ("idl_length", ct.c_double),
        ("tau", OTauCoef),
        ("dt", ct.c_double),
        ("is_implicit", ct.c_int),
        ("state_var", ct.c_int),
        ("stabilization", ct.c_int),
        ("idl_enable", ct.c_int),
    ]


class OMesh(ct.Structure):
    _fields_ = [
        ("nelem", ct.c_int),
        ("P", ct.c_int),
        ("Q", ct.c_int),
        ("Px", ct.c_int),
        ("nnodes", ct.c_int),
        ("nxnodes", ct.c_int),
        ("offsets", _ip),
        ("xoffsets", _ip),
        ("xcoord", _dp),
        ("B", _dp),
        ("D", _dp),
        ("Bx", _dp),
        ("Dx", _dp),
        ("w1d", _dp),
    ]


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ct.CDLL(LIB_PATH)
        _lib.oracle_volume.restype = ct.c_double
    return _lib


def set_num_threads(n: int) -> None:
    lib().oracle_set_num_threads(int(n))


def max_threads() -> int:
    return int(lib().oracle_max_threads())


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


class Mesh:
    """Keeps the numpy arrays alive and exposes the OMesh struct."""

    def __init__(self, mesh, tab):
        self.mesh, self.tab = mesh, tab
        self._arrs = dict(
            offsets=np.ascontiguousarray(mesh.offsets, dtype=np.int32),
            xoffsets=np.ascontiguousarray(mesh.xoffsets, dtype=np.int32),
            xcoord=np.ascontiguousarray(mesh.xcoord, dtype=np.float64),
            B=np.ascontiguousarray(tab.B),
            D=np.ascontiguousarray(tab.D),
            Bx=np.ascontiguousarray(tab.Bx),
            Dx=np.ascontiguousarray(tab.Dx),
            w1d=np.ascontiguousarray(tab.qweight),
        )
        a = self._arrs
        self.s = OMesh(
            mesh.nelem, tab.P, tab.Q, 2, mesh.nnodes, a["xcoord"].shape[0],
            _i(a["offsets"]), _i(a["xoffsets"]), _d(a["xcoord"]), _d(a["B"]), _d(a["D"]), _d(a["Bx"]), _d(a["Dx"]), _d(a["w1d"]),
        )
        self.nelem, self.Q3, self.nnodes = mesh.nelem, tab.Q**3, mesh.nnodes

    def setup_qdata(self):
        qd = np.zeros((self.nelem, 10, self.Q3))
        lib().oracle_setup_qdata(ct.byref(self.s), _d(qd))
        return qd


def newt_from_dict(d: dict) -> ONewt:
    c = ONewt()
    g = d["gas"]
    c.gas = OGas(g["lambda"], g["mu"], g["k"], g["cv"], g["cp"])
    for i in range(3):
        c.g[i] = d.get("g", (0, 0, 0))[i]
    c.time = d.get("time", 0.0)
    c.ijacobian_time_shift = d.get("ijacobian_time_shift", 0.0)
    c.idl_pressure = d.get("idl_pressure", 0.0)
    c.idl_amplitude = d.get("idl_amplitude", 0.0)
    c.idl_start = d.get("idl_start", 0.0)
    c.idl_length = d.get("idl_length", 0.0)
    t = d["tau"]
    c.tau = OTauCoef(t["Ctau_t"], t["Ctau_v"], t["Ctau_C"], t["Ctau_M"], t["Ctau_E"])
    c.dt = d.get("dt", 1.0)
    c.is_implicit = int(d.get("is_implicit", 1))
    c.state_var = int(d.get("state_var", 0))
    c.stabilization = int(d.get("stabilization", 0))
    c.idl_enable = int(d.get("idl_enable", 0))
    return c


def njac(c: ONewt) -> int:
    return 14 + (1 if c.idl_enable else 0)


def ifunction(m: Mesh, c: ONewt, qdata, q, qdot):
    v = np.zeros(5 * m.nnodes)
    jac = np.zeros((m.nelem, njac(c), m.Q3))
    lib().oracle_ifunction(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), _d(qdot), _d(v), _d(jac))
    return v, jac


def ijacobian(m: Mesh, c: ONewt, qdata, jac, dq):
    v = np.zeros(5 * m.nnodes)
    lib().oracle_ijacobian(ct.byref(m.s), ct.byref(c), _d(qdata), _d(jac), _d(dq), _d(v))
    return v


def ijacobian_diagonal(m: Mesh, c: ONewt, qdata, jac):
    v = np.zeros(5 * m.nnodes)
    lib().oracle_ijacobian_diagonal(ct.byref(m.s), ct.byref(c), _d(qdata), _d(jac), _d(v))
    return v


def ijacobian_point_block_diagonal(m: Mesh, c: ONewt, qdata, jac):
    """[nnodes][5 (out)][5 (in)] point-block diagonal of the IJacobian."""
    pb = np.zeros(25 * m.nnodes)
    lib().oracle_ijacobian_point_block_diagonal(ct.byref(m.s), ct.byref(c), _d(qdata), _d(jac), _d(pb))
    return pb.reshape(m.nnodes, 5, 5)


def rhs(m: Mesh, c: ONewt, qdata, q):
    v = np.zeros(5 * m.nnodes)
    lib().oracle_rhs(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), _d(v))
    return v


def mass_newtonian(m: Mesh, c: ONewt, qdata, q, qdot):
    v = np.zeros(5 * m.nnodes)
    lib().oracle_mass_newtonian(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), _d(qdot), _d(v))
    return v


def mass_newtonian_diagonal(m: Mesh, c: ONewt, qdata, q):
    d = np.zeros(5 * m.nnodes)
    lib().oracle_mass_newtonian_diagonal(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), _d(d))
    return d


def monitor_ke(m: Mesh, c: ONewt, qdata, q):
    s = np.zeros(5)
    lib().oracle_monitor_ke(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), _d(s))
    return s


def volume(m: Mesh, qdata) -> float:
    return lib().oracle_volume(ct.byref(m.s), _d(qdata))


def ic_taylorgreen(c: ONewt, ref_Y, bg_u, L, xyz):
    xyz = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    q = np.zeros((xyz.shape[0], 5))
    lib().oracle_ic_taylorgreen(ct.byref(c), _d(np.asarray(ref_Y, float)), _d(np.asarray(bg_u, float)), _d(np.asarray(L, float)),
                                xyz.shape[0], _d(xyz), _d(q))
    return q


def rk4_stabilized(m: Mesh, c: ONewt, qdata, q, dt, nsteps, rtol=1e-5, restart=30, record_every=1):
    q = np.array(q, dtype=np.float64)
    rows = np.zeros((nsteps + 2, 7))
    nrows = ct.c_int(0)
    its = lib().oracle_rk4_stabilized(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), ct.c_double(dt), nsteps, ct.c_double(rtol), restart,
                                      record_every, _d(rows), ct.byref(nrows))
    return q, rows[: nrows.value], its


def state_from_q(gas: OGas, q, var):
    U, Y = np.zeros(5), np.zeros(5)
    lib().oracle_state_from_q(ct.byref(gas), _d(np.asarray(q, float)), var, _d(U), _d(Y))
    return U, Y


def state_from_q_fwd(gas: OGas, q, dq, var):
    dU, dY = np.zeros(5), np.zeros(5)
    lib().oracle_state_from_q_fwd(ct.byref(gas), _d(np.asarray(q, float)), _d(np.asarray(dq, float)), var, _d(dU), _d(dY))
    return dU, dY


def state_to_q(gas: OGas, U, Y, var):
    q = np.zeros(5)
    lib().oracle_state_to_q(ct.byref(gas), _d(np.asarray(U, float)), _d(np.asarray(Y, float)), var, _d(q))
    return q


def state_to_q_fwd(gas: OGas, U, Y, dU, dY, var):
    dq = np.zeros(5)
    lib().oracle_state_to_q_fwd(ct.byref(gas), _d(np.asarray(U, float)), _d(np.asarray(Y, float)), _d(np.asarray(dU, float)),
                                _d(np.asarray(dY, float)), var, _d(dq))
    return dq


def tau_diag(gas: OGas, tc: OTauCoef, q, var, dXdx, dt):
    T = np.zeros(3)
    lib().oracle_tau_diag(ct.byref(gas), ct.byref(tc), _d(np.asarray(q, float)), var, _d(np.ascontiguousarray(dXdx, float)),
                          ct.c_double(dt), _d(T))
    return T


def qf_ifunction(c: ONewt, q, gref, qdot, wdetJ, dXdx, x=(0.0, 0.0, 0.0)):
    v, gv, jd = np.zeros(5), np.zeros(15), np.zeros(15)
    a = lambda z: np.ascontiguousarray(z, dtype=np.float64)  # noqa: E731
    lib().oracle_qf_ifunction(ct.byref(c), _d(a(q)), _d(a(gref)), _d(a(qdot)), ct.c_double(wdetJ), _d(a(dXdx)), _d(a(x)), _d(v), _d(gv), _d(jd))
    return v, gv.reshape(3, 5), jd


def qf_ijacobian(c: ONewt, dq, gref, wdetJ, dXdx, jd):
    v, gv = np.zeros(5), np.zeros(15)
    a = lambda z: np.ascontiguousarray(z, dtype=np.float64)  # noqa: E731
    lib().oracle_qf_ijacobian(ct.byref(c), _d(a(dq)), _d(a(gref)), ct.c_double(wdetJ), _d(a(dXdx)), _d(a(jd)), _d(v), _d(gv))
    return v, gv.reshape(3, 5)


def qf_rhs(c: ONewt, U, gref, wdetJ, dXdx, x=(0.0, 0.0, 0.0)):
    v, gv = np.zeros(5), np.zeros(15)
    a = lambda z: np.ascontiguousarray(z, dtype=np.float64)  # noqa: E731
    lib().oracle_qf_rhs(ct.byref(c), _d(a(U)), _d(a(gref)), ct.c_double(wdetJ), _d(a(dXdx)), _d(a(x)), _d(v), _d(gv))
    return v, gv.reshape(3, 5)


# ---------------------------------------------------------------------------------------------------------------------
# Operation counts of the QFunction bodies (liboracle_opcount.so, opcount.cpp): the same code over a counting scalar
# ---------------------------------------------------------------------------------------------------------------------
OPCOUNT_PATH = os.path.join(HERE, "liboracle_opcount.so")
OP_KINDS = ("add", "mul", "div", "sqrt", "exp", "log")
_oplib = None


def oplib():
    global _oplib
    if _oplib is None:
        if not os.path.exists(OPCOUNT_PATH):
            build()
        _oplib = ct.CDLL(OPCOUNT_PATH)
    return _oplib


def _counts(n):
    return {k: int(n[i]) for i, k in enumerate(OP_KINDS)}


def opcount_prim_from_cons(gas: OGas, U):
    Y, n = np.zeros(5), np.zeros(6, dtype=np.int64)
    oplib().oracle_opcount_prim_from_cons(ct.byref(gas), _d(np.ascontiguousarray(U, float)), _d(Y), n.ctypes.data_as(ct.c_void_p))
    return Y, _counts(n)


def opcount_ifunction(c: ONewt, q, gref, qdot, wdetJ, dXdx, x=(0.0, 0.0, 0.0)):
    """IFunction body at one point over the counting scalar: (v, gv, jd, counts)."""
    v, gv, jd, n = np.zeros(5), np.zeros(15), np.zeros(15), np.zeros(6, dtype=np.int64)
    a = lambda z: np.ascontiguousarray(z, dtype=np.float64)  # noqa: E731
    oplib().oracle_opcount_ifunction(ct.byref(c), _d(a(q)), _d(a(gref)), _d(a(qdot)), ct.c_double(wdetJ), _d(a(dXdx)), _d(a(x)), _d(v), _d(gv),
                                     _d(jd), n.ctypes.data_as(ct.c_void_p))
    return v, gv.reshape(3, 5), jd, _counts(n)


def opcount_ijacobian(c: ONewt, dq, gref, wdetJ, dXdx, jd):
    v, gv, n = np.zeros(5), np.zeros(15), np.zeros(6, dtype=np.int64)
    a = lambda z: np.ascontiguousarray(z, dtype=np.float64)  # noqa: E731
    oplib().oracle_opcount_ijacobian(ct.byref(c), _d(a(dq)), _d(a(gref)), ct.c_double(wdetJ), _d(a(dXdx)), _d(a(jd)), _d(v), _d(gv),
                                     n.ctypes.data_as(ct.c_void_p))
    return v, gv.reshape(3, 5), _counts(n)


def opcount_rhs(c: ONewt, U, gref, wdetJ, dXdx, x=(0.0, 0.0, 0.0)):
    v, gv, n = np.zeros(5), np.zeros(15), np.zeros(6, dtype=np.int64)
    a = lambda z: np.ascontiguousarray(z, dtype=np.float64)  # noqa: E731
    oplib().oracle_opcount_rhs(ct.byref(c), _d(a(U)), _d(a(gref)), ct.c_double(wdetJ), _d(a(dXdx)), _d(a(x)), _d(v), _d(gv),
                               n.ctypes.data_as(ct.c_void_p))
    return v, gv.reshape(3, 5), _counts(n)


# ---------------------------------------------------------------------------------------------------------------------
# Boundary faces: consistent boundary integral and freestream Riemann boundary
# ---------------------------------------------------------------------------------------------------------------------
class OFreestream(ct.Structure):
    _fields_ = [("newt", ONewt), ("S_inf_U", ct.c_double * 5), ("S_inf_Y", ct.c_double * 5), ("riemann", ct.c_int)]


class OOutflow(ct.Structure):
    _fields_ = [("newt", ONewt), ("recirc", ct.c_double), ("softplus_velocity", ct.c_double), ("pressure", ct.c_double),
                ("temperature", ct.c_double), ("pressure_outflow", ct.c_int)]


class OFaces(ct.Structure):
    _fields_ = [("nfaces", ct.c_int), ("face_elem", _ip), ("face_id", _ip)]


class Faces:
    """Boundary faces given as (volume element, local face id 0:-x 1:+x 2:-y 3:+y 4:-z 5:+z)."""

    def __init__(self, face_elem, face_id):
        self.elem = np.ascontiguousarray(face_elem, dtype=np.int32)
        self.fid = np.ascontiguousarray(face_id, dtype=np.int32)
        self.s = OFaces(len(self.elem), _i(self.elem), _i(self.fid))
        self.nfaces = len(self.elem)


def freestream_ctx(c: ONewt, S_inf_U, S_inf_Y, hllc: bool) -> OFreestream:
    f = OFreestream()
    f.newt = c
    for i in range(5):
        f.S_inf_U[i] = S_inf_U[i]
        f.S_inf_Y[i] = S_inf_Y[i]
    f.riemann = 1 if hllc else 0
    return f


def setup_qdata_boundary(m: Mesh, f: Faces):
    qd = np.zeros((f.nfaces, 10, m.tab.Q**2))
    lib().oracle_setup_qdata_boundary(ct.byref(m.s), ct.byref(f.s), _d(qd))
    return qd


def boundary_integral(m: Mesh, f: Faces, c: ONewt, qdata_sur, q):
    v = np.zeros(5 * m.nnodes)
    jac = np.zeros((f.nfaces, 11, m.tab.Q**2))
    lib().oracle_boundary_integral(ct.byref(m.s), ct.byref(f.s), ct.byref(c), _d(qdata_sur), _d(q), _d(v), _d(jac))
    return v, jac


def boundary_integral_jacobian(m: Mesh, f: Faces, c: ONewt, qdata_sur, jac_sur, dq):
    v = np.zeros(5 * m.nnodes)
    lib().oracle_boundary_integral_jacobian(ct.byref(m.s), ct.byref(f.s), ct.byref(c), _d(qdata_sur), _d(jac_sur), _d(dq), _d(v))
    return v


def freestream(m: Mesh, f: Faces, fc: OFreestream, qdata_sur, q):
    v = np.zeros(5 * m.nnodes)
    jac = np.zeros((f.nfaces, 5, m.tab.Q**2))
    lib().oracle_freestream(ct.byref(m.s), ct.byref(f.s), ct.byref(fc), _d(qdata_sur), _d(q), _d(v), _d(jac))
    return v, jac


def freestream_jacobian(m: Mesh, f: Faces, fc: OFreestream, qdata_sur, jac_sur, dq):
    v = np.zeros(5 * m.nnodes)
    lib().oracle_freestream_jacobian(ct.byref(m.s), ct.byref(f.s), ct.byref(fc), _d(qdata_sur), _d(jac_sur), _d(dq), _d(v))
    return v


def outflow_ctx(c: ONewt, pressure, temperature, recirc=1.0, softplus_velocity=1e-2, pressure_outflow=False) -> OOutflow:
    """OutflowContext_ as problems/bc_outflow.c:154-160 fills it (defaults recirc 1, softplus width 1e-2)."""
    return OOutflow(c, float(recirc), float(softplus_velocity), float(pressure), float(temperature), int(bool(pressure_outflow)))


def outflow(m: Mesh, f: Faces, oc: OOutflow, qdata_sur, q):
    v = np.zeros(5 * m.nnodes)
    jac = np.zeros((f.nfaces, 11, m.tab.Q**2))
    lib().oracle_outflow(ct.byref(m.s), ct.byref(f.s), ct.byref(oc), _d(qdata_sur), _d(q), _d(v), _d(jac))
    return v, jac


def outflow_jacobian(m: Mesh, f: Faces, oc: OOutflow, qdata_sur, jac_sur, dq):
    v = np.zeros(5 * m.nnodes)
    lib().oracle_outflow_jacobian(ct.byref(m.s), ct.byref(f.s), ct.byref(oc), _d(qdata_sur), _d(jac_sur), _d(dq), _d(v))
    return v


def slip(m: Mesh, f: Faces, c: ONewt, qdata_sur, q):
    v = np.zeros(5 * m.nnodes)
    jac = np.zeros((f.nfaces, 5, m.tab.Q**2))
    lib().oracle_slip(ct.byref(m.s), ct.byref(f.s), ct.byref(c), _d(qdata_sur), _d(q), _d(v), _d(jac))
    return v, jac


def slip_jacobian(m: Mesh, f: Faces, c: ONewt, qdata_sur, jac_sur, dq):
    v = np.zeros(5 * m.nnodes)
    lib().oracle_slip_jacobian(ct.byref(m.s), ct.byref(f.s), ct.byref(c), _d(qdata_sur), _d(jac_sur), _d(dq), _d(v))
    return v


def _v(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def riemann_flux(gas: OGas, hllc: bool, qL, qR, var, n):
    F = np.zeros(5)
    lib().oracle_riemann_flux(ct.byref(gas), int(hllc), _d(_v(qL)), _d(_v(qR)), int(var), _d(_v(n)), _d(F))
    return F


def riemann_flux_fwd(gas: OGas, hllc: bool, qL, dqL, qR, dqR, var, n):
    dF = np.zeros(5)
    lib().oracle_riemann_flux_fwd(ct.byref(gas), int(hllc), _d(_v(qL)), _d(_v(dqL)), _d(_v(qR)), _d(_v(dqR)), int(var), _d(_v(n)), _d(dF))
    return dF


def hll_speeds(gas: OGas, qL, qR, var, n):
    sL, sR = ct.c_double(), ct.c_double()
    lib().oracle_hll_speeds(ct.byref(gas), _d(_v(qL)), _d(_v(qR)), int(var), _d(_v(n)), ct.byref(sL), ct.byref(sR))
    return sL.value, sR.value


def hll_speeds_fwd(gas: OGas, qL, dqL, qR, dqR, var, n):
    out = [ct.c_double() for _ in range(4)]
    lib().oracle_hll_speeds_fwd(ct.byref(gas), _d(_v(qL)), _d(_v(dqL)), _d(_v(qR)), _d(_v(dqR)), int(var), _d(_v(n)), *[ct.byref(o) for o in out])
    return tuple(o.value for o in out)


# ---- time integration (honee_oracle_solve.c) ----
class ORKMass(ct.Structure):
    _fields_ = [("gmres", ct.c_int), ("jacobi_rowsum", ct.c_int), ("rtol", ct.c_double), ("restart", ct.c_int), ("max_it", ct.c_int)]


class OBDFOptions(ct.Structure):
    _fields_ = [("order", ct.c_int), ("dt", ct.c_double), ("snes_rtol", ct.c_double), ("snes_atol", ct.c_double), ("snes_stol", ct.c_double),
                ("snes_max_it", ct.c_int), ("ksp_rtol", ct.c_double), ("ksp_atol", ct.c_double), ("ksp_restart", ct.c_int),
                ("ksp_max_it", ct.c_int), ("pc_pbjacobi", ct.c_int), ("linesearch", ct.c_int)]


class OBDFStats(ct.Structure):
    _fields_ = [("steps", ct.c_int), ("time", ct.c_double), ("ifunction_evals", ct.c_longlong), ("newton_its", ct.c_longlong),
                ("krylov_its", ct.c_longlong), ("line_search_backtracks", ct.c_longlong), ("last_newton_its", ct.c_int),
                ("last_krylov_its", ct.c_int), ("last_fnorm0", ct.c_double), ("last_fnorm", ct.c_double)]


def tableau_ref(name: str, stages: int = 4):
    """Butcher tableaus restated from their published definitions: classic RK4; Fehlberg's RK5(4) propagating the
    fifth-order weights; Shu-Osher SSP(3,3); forward Euler; Ketcheson's optimal third-order SSP with n^2 stages,
    whose stages are derived here by running its low-storage recurrence on symbolic stage coefficients."""
    if name == "4":
        A = [[0, 0, 0, 0], [0.5, 0, 0, 0], [0, 0.5, 0, 0], [0, 0, 1, 0]]
        b = [1 / 6, 1 / 3, 1 / 3, 1 / 6]
    elif name == "5f":
        A = [[0] * 6, [1 / 4, 0, 0, 0, 0, 0], [3 / 32, 9 / 32, 0, 0, 0, 0], [1932 / 2197, -7200 / 2197, 7296 / 2197, 0, 0, 0],
             [439 / 216, -8, 3680 / 513, -845 / 4104, 0, 0], [-8 / 27, 2, -3544 / 2565, 1859 / 4104, -11 / 40, 0]]
        b = [16 / 135, 0, 6656 / 12825, 28561 / 56430, -9 / 50, 2 / 55]
    elif name == "ssp33":
        A = [[0, 0, 0], [1, 0, 0], [0.25, 0.25, 0]]
        b = [1 / 6, 1 / 6, 2 / 3]
    elif name == "euler":
        A, b = [[0]], [1]
    elif name == "ssp3":
        n = int(round(np.sqrt(stages)))
        assert n * n == stages and n >= 2
        r = stages - n
        A = np.zeros((stages, stages))
        cur = np.zeros(stages)  # coefficients of the current register over h K_j (the u coefficient is 1)
        k = 0

        def lstep(cur, k):
            A[k] = cur
            cur = cur.copy()
            cur[k] += 1.0 / r
            return cur, k + 1

        for _ in range((n - 1) * (n - 2) // 2):
            cur, k = lstep(cur, k)
        saved = cur.copy()
        for _ in range((n - 1) * (n - 2) // 2 + 1, n * (n + 1) // 2):
            cur, k = lstep(cur, k)
        cur, k = lstep(cur, k)
        cur = (n * saved + (n - 1) * cur) / (2 * n - 1)
        for _ in range(n * (n + 1) // 2 + 1, n * n + 1):
            cur, k = lstep(cur, k)
        assert k == stages
        b = cur
    else:
        raise KeyError(name)
    A, b = np.array(A, dtype=np.float64), np.array(b, dtype=np.float64)
    return A, b, A.sum(axis=1)


class ORKExtra(ct.Structure):
    _fields_ = [("faces", ct.c_void_p), ("freestream", ct.c_void_p), ("qdata_sur", _dp), ("mask", _dp), ("qbc", _dp)]


def rk(m: Mesh, c: ONewt, qdata, q, dt, nsteps, A, b, cc, t0=0.0, gmres=False, rowsum=True, rtol=1e-5, restart=30, max_it=10000,
       record_every=0, faces=None, freestream=None, qdata_sur=None, essential=None):
    """Explicit RK with the mass KSP.  Optional: freestream faces (Faces, OFreestream with the explicit sign, their
    surface qdata) added to the RHS, and essential boundary conditions essential = (dofs, values)."""
    q = np.array(q, dtype=np.float64)
    A, b, cc = (np.ascontiguousarray(x, dtype=np.float64) for x in (A, b, cc))
    mo = ORKMass(int(gmres), int(rowsum), rtol, restart, max_it)
    rows = np.zeros((nsteps + 2, 7)) if record_every else None
    nrows = ct.c_int(0)
    its = ct.c_longlong(0)
    ex, keep = None, []
    if faces is not None or essential is not None:
        mask = qbc = None
        if essential is not None:
            dofs, values = essential
            mask, qbc = np.ones(q.size), np.zeros(q.size)
            mask[np.asarray(dofs)] = 0.0
            qbc[np.asarray(dofs)] = values
            keep += [mask, qbc]
        qds = np.ascontiguousarray(qdata_sur, dtype=np.float64) if faces is not None else None
        keep.append(qds)
        ex = ORKExtra(ct.cast(ct.byref(faces.s), ct.c_void_p) if faces is not None else None,
                      ct.cast(ct.byref(freestream), ct.c_void_p) if faces is not None else None,
                      _d(qds) if qds is not None else None, _d(mask) if mask is not None else None, _d(qbc) if qbc is not None else None)
    lib().oracle_rk_ex(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), ct.c_double(t0), ct.c_double(dt), nsteps, len(b), _d(A), _d(b), _d(cc),
                       ct.byref(mo), ct.byref(ex) if ex is not None else None, max(record_every, 1), _d(rows) if rows is not None else None,
                       ct.byref(nrows), ct.byref(its))
    return q, (rows[: nrows.value] if rows is not None else None), its.value


def bdf(m: Mesh, c: ONewt, qdata, q, dt, nsteps, order=2, t0=0.0, snes_rtol=1e-8, snes_atol=1e-50, snes_stol=1e-8, snes_max_it=50,
        ksp_rtol=1e-5, ksp_atol=1e-50, ksp_restart=30, ksp_max_it=10000, pbjacobi=False, linesearch="bt"):
    q = np.array(q, dtype=np.float64)
    o = OBDFOptions(order, dt, snes_rtol, snes_atol, snes_stol, snes_max_it, ksp_rtol, ksp_atol, ksp_restart, ksp_max_it, int(pbjacobi),
                    {"bt": 0, "basic": 1, "norm_slope": 2}[linesearch])
    st = OBDFStats()
    ierr = lib().oracle_bdf(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), ct.c_double(t0), nsteps, ct.byref(o), ct.byref(st))
    if ierr:
        raise RuntimeError(f"oracle BDF failed ({ierr})")
    return q, {k: getattr(st, k) for k, _ in OBDFStats._fields_}


def bt_next_lambda(fnorm, initslope, lam, gnorm, nfit, lamprev=0.0, gnormprev=0.0):
    """Next lambda of the backtracking line search after a rejected trial (oracle_bt_next_lambda)."""
    f = lib().oracle_bt_next_lambda
    f.restype = ct.c_double
    f.argtypes = [ct.c_double, ct.c_double, ct.c_double, ct.c_double, ct.c_int, ct.c_double, ct.c_double]
    return f(fnorm, initslope, lam, gnorm, nfit, lamprev, gnormprev)


# ---- data-driven SGS closure (honee_oracle.c, sgs_pointwise.h) ----
class OSgsModel(ct.Structure):
    _fields_ = [("num_inputs", ct.c_int), ("num_outputs", ct.c_int), ("num_neurons", ct.c_int), ("alpha", ct.c_double),
                ("b1", _dp), ("b2", _dp), ("W1", _dp), ("W2", _dp), ("bounds", _dp), ("bh", _dp), ("Wh", _dp)]


class SgsModel:
    """Keeps the parameter arrays alive for the C struct.  A model dict with "Wh" [neuron][neuron] and "bh" has a
    second hidden layer."""

    def __init__(self, params: dict):
        keys = ("b1", "b2", "W1", "W2", "bounds") + (("bh", "Wh") if "Wh" in params else ())
        self.arrays = {k: np.ascontiguousarray(params[k], dtype=np.float64) for k in keys}
        a = self.arrays
        self.s = OSgsModel(a["W1"].shape[1], a["W2"].shape[0], a["W1"].shape[0], float(params["alpha"]), _d(a["b1"]), _d(a["b2"]),
                           _d(a["W1"]), _d(a["W2"]), _d(a["bounds"]), _d(a["bh"]) if "bh" in a else None, _d(a["Wh"]) if "Wh" in a else None)


def diagonalize3(A, max_sweeps=5):
    A = np.ascontiguousarray(A, dtype=np.float64)
    ev, E = np.zeros(3), np.zeros((3, 3))
    it = lib().oracle_diagonalize3(_d(A), max_sweeps, _d(ev), _d(E))
    return ev, E, it


def anisotropy_tensor(dXdx):
    dXdx = np.ascontiguousarray(dXdx, dtype=np.float64)
    kmA, delta = np.zeros(6), ct.c_double(0)
    lib().oracle_anisotropy_tensor(_d(dXdx), _d(kmA), ct.byref(delta))
    return kmA, delta.value


def sgs_point(model: SgsModel, gu, kmA, delta, nu):
    gu = np.ascontiguousarray(gu, dtype=np.float64)
    kmA = np.ascontiguousarray(kmA, dtype=np.float64)
    inp, out, E, km = np.zeros(6), np.zeros(6), np.zeros((3, 3)), np.zeros(6)
    gmag = ct.c_double(0)
    lib().oracle_sgs_point(ct.byref(model.s), _d(gu), _d(kmA), ct.c_double(delta), ct.c_double(nu), _d(inp), _d(out), _d(E), ct.byref(gmag), _d(km))
    return dict(inputs=inp, outputs=out, evec=E, gmag=gmag.value, km_sgs=km)


def velocity_gradient_rhs(m: Mesh, c: ONewt, qdata, q):
    r = np.zeros(9 * m.nnodes)
    lib().oracle_velocity_gradient_rhs(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), _d(r))
    return r


def velocity_gradient_projection(m: Mesh, c: ONewt, qdata, q):
    G = np.zeros(9 * m.nnodes)
    lib().oracle_velocity_gradient_projection(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), _d(G))
    return G


def grid_anisotropy_rhs(m: Mesh, qdata):
    r = np.zeros(7 * m.nnodes)
    lib().oracle_grid_anisotropy_rhs(ct.byref(m.s), _d(qdata), _d(r))
    return r


def grid_anisotropy_projection(m: Mesh, qdata):
    X = np.zeros(7 * m.nnodes)
    its = lib().oracle_grid_anisotropy_projection(ct.byref(m.s), _d(qdata), _d(X))
    return X, its


def mass_n(m: Mesh, qdata, ncomp, u):
    v = np.zeros(ncomp * m.nnodes)
    lib().oracle_mass_n(ct.byref(m.s), _d(qdata), ncomp, _d(np.ascontiguousarray(u, dtype=np.float64)), _d(v))
    return v


def mass_n_diagonal(m: Mesh, qdata, ncomp):
    d = np.zeros(ncomp * m.nnodes)
    lib().oracle_mass_n_diagonal(ct.byref(m.s), _d(qdata), ncomp, _d(d))
    return d


def sgs_nodal(gas: OGas, state_var: int, model: SgsModel, q, G, A):
    n = q.size // 5
    km = np.zeros(6 * n)
    lib().oracle_sgs_nodal(ct.byref(gas), state_var, ct.byref(model.s), n, _d(q), _d(G), _d(A), _d(km))
    return km


def ifunction_nodal_sgs(m: Mesh, c: ONewt, qdata, q, km_sgs):
    v = np.zeros(5 * m.nnodes)
    lib().oracle_ifunction_nodal_sgs(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), _d(km_sgs), _d(v))
    return v


# ---- differential filter and SGS training pairs (honee_oracle.c) ----
class ODiffFilter(ct.Structure):
    _fields_ = [("grid_based_width", ct.c_int), ("width_scaling", ct.c_double * 3), ("kernel_scaling", ct.c_double),
                ("friction_length", ct.c_double), ("damping", ct.c_int), ("damping_constant", ct.c_double)]


def diff_filter(grid_based_width=False, width_scaling=(1.0, 1.0, 1.0), kernel_scaling=0.1, friction_length=0.0, damping=0,
                damping_constant=25.0) -> ODiffFilter:
    """Options of DifferentialFilterSetup (src/differential_filter.c:236-260), defaults included."""
    return ODiffFilter(int(grid_based_width), (ct.c_double * 3)(*width_scaling), kernel_scaling, friction_length, int(damping), damping_constant)


def anisotropy_collocate(m: Mesh, qdata):
    A = np.zeros(m.nelem * 7 * m.Q3)
    lib().oracle_anisotropy_collocate(ct.byref(m.s), _d(qdata), _d(A))
    return A


def diff_filter_rhs(m: Mesh, c: ONewt, qdata, q):
    r = np.zeros(11 * m.nnodes)
    lib().oracle_diff_filter_rhs(ct.byref(m.s), ct.byref(c), _d(qdata), _d(q), _d(r))
    return r


def diff_filter_lhs(m: Mesh, f: ODiffFilter, qdata, Acol, ncomp, u):
    v = np.zeros(ncomp * m.nnodes)
    lib().oracle_diff_filter_lhs(ct.byref(m.s), ct.byref(f), _d(qdata), _d(Acol), ncomp, _d(np.ascontiguousarray(u, dtype=np.float64)), _d(v))
    return v


def diff_filter_lhs_diagonal(m: Mesh, f: ODiffFilter, qdata, Acol, ncomp):
    d = np.zeros(ncomp * m.nnodes)
    lib().oracle_diff_filter_lhs_diagonal(ct.byref(m.s), ct.byref(f), _d(qdata), _d(Acol), ncomp, _d(d))
    return d


def diff_filter_solve(m: Mesh, f: ODiffFilter, qdata, Acol, ncomp, R, rtol=1e-10, atol=1e-50, max_it=10000):
    X = np.zeros(ncomp * m.nnodes)
    its = lib().oracle_diff_filter_solve(ct.byref(m.s), ct.byref(f), _d(qdata), _d(Acol), ncomp, _d(np.ascontiguousarray(R, dtype=np.float64)),
                                         ct.c_double(rtol), ct.c_double(atol), int(max_it), _d(X))
    return X, its


def sgs_training_nodal(gas: OGas, Yf, G, A):
    n = Yf.size // 11
    out = np.zeros(12 * n)
    lib().oracle_sgs_training_nodal(ct.byref(gas), n, _d(np.ascontiguousarray(Yf, dtype=np.float64)), _d(np.ascontiguousarray(G, dtype=np.float64)),
                                    _d(np.ascontiguousarray(A, dtype=np.float64)), _d(out))
    return out


# ---- SGS model training (FP64 restatement of the MLP, its gradient and Adam; the GPU trains in FP32) ----
MLP_DIMS = (6, 20, 6)


def mlp_unpack(theta, dims=MLP_DIMS):
    """Layers (W, b) of the parameter vector: dims (ni, nh, no) gives W1 [nh][ni], b1 [nh], W2 [no][nh], b2 [no];
    dims (ni, nh, nh2, no) gives W1 [nh][ni], b1 [nh], W2 [nh2][nh], b2 [nh2], W3 [no][nh2], b3 [no]."""
    theta = np.asarray(theta, dtype=np.float64)
    layers, o = [], 0
    for n_in, n_out in zip(dims[:-1], dims[1:]):
        W = theta[o:o + n_out * n_in].reshape(n_out, n_in)
        o += n_out * n_in
        b = theta[o:o + n_out]
        o += n_out
        layers.append((W, b))
    if o != theta.size:
        raise ValueError(f"parameter vector of {theta.size} entries does not match dims {dims}")
    return layers


def mlp_forward(theta, X, alpha, dims=MLP_DIMS):
    a = np.asarray(X, dtype=np.float64)
    layers = mlp_unpack(theta, dims)
    for W, b in layers[:-1]:
        h = a @ W.T + b
        a = np.where(h < 0, alpha * h, h)
    W, b = layers[-1]
    return a @ W.T + b


def mlp_loss_grad(theta, X, T, alpha, dims=MLP_DIMS):
    """Mean squared error over samples and outputs and its gradient by back-propagation (any number of hidden layers)."""
    layers = mlp_unpack(theta, dims)
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    n, no = T.shape
    acts, pre = [X], []
    for W, b in layers[:-1]:
        h = acts[-1] @ W.T + b
        pre.append(h)
        acts.append(np.where(h < 0, alpha * h, h))
    W, b = layers[-1]
    y = acts[-1] @ W.T + b
    e = y - T
    loss = float(np.sum(e * e) / (n * no))
    d = 2 * e / (n * no)
    grads = []
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        grads.append((d.T @ acts[li], d.sum(0)))
        if li > 0:
            d = (d @ W) * np.where(pre[li - 1] < 0, alpha, 1.0)
    grads.reverse()
    return loss, np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def adam_step(theta, grad, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam step t (1-based) with bias-corrected moments."""
    m = beta1 * m + (1 - beta1) * grad
    v = beta2 * v + (1 - beta2) * grad * grad
    mh = m / (1 - beta1 ** t)
    vh = v / (1 - beta2 ** t)
    return theta - lr * mh / (np.sqrt(vh) + eps), m, v


def sgs_model_samples(gas: OGas, train, filtered, G, A):
    """Model inputs and the eigenframe stress S = E tau E^T / (delta^2 |grad u|^2) (order S00 S11 S22 S01 S02 S12) of
    nodal training fields, with E and |grad u| from the oracle's closure inputs."""
    zero = SgsModel({"b1": np.zeros(20), "b2": np.zeros(6), "W1": np.zeros((20, 6)), "W2": np.zeros((6, 20)), "bounds": np.zeros((6, 2)), "alpha": 0.0})
    train, filtered, G, A = (np.asarray(a, dtype=np.float64).reshape(len(np.asarray(a).reshape(-1, k)), k)
                             for a, k in ((train, 12), (filtered, 11), (G, 9), (A, 7)))
    n = train.shape[0]
    X, S = np.zeros((n, 6)), np.zeros((n, 6))
    R = gas.cp - gas.cv
    for i in range(n):
        rho = filtered[i, 0] / (R * filtered[i, 4])
        r = sgs_point(zero, G[i].reshape(3, 3), A[i, :6], A[i, 6], gas.mu / rho)
        t = train[i, 6:]
        tau = np.array([[t[0], t[5], t[4]], [t[5], t[1], t[3]], [t[4], t[3], t[2]]])
        E = r["evec"]
        Sm = E @ tau @ E.T / (A[i, 6] ** 2 * r["gmag"] ** 2)
        X[i] = r["inputs"]
        S[i] = [Sm[0, 0], Sm[1, 1], Sm[2, 2], Sm[0, 1], Sm[0, 2], Sm[1, 2]]
    return X, S
