"""ctypes binding of libhonee_setup.so / libceed.so (the MI355X backend).

Loading fails loudly when the native libraries are missing: there is no Python or CPU fallback for the operators.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

from . import LIB_DIR
from .mesh import BoxMesh, Tabulation, face_node_tables

_dp, _ip, _fp = ct.POINTER(ct.c_double), ct.POINTER(ct.c_int), ct.POINTER(ct.c_float)

CEED_MEM_HOST, CEED_MEM_DEVICE = 0, 1
VAR_CONS, VAR_PRIM, VAR_ENTROPY = 0, 1, 2
STAB_NONE, STAB_SU, STAB_SUPG = 0, 1, 2


class Gas(ct.Structure):
    _fields_ = [("lambda_", ct.c_double), ("mu", ct.c_double), ("k", ct.c_double), ("cv", ct.c_double), ("cp", ct.c_double)]


class TauCoeffs(ct.Structure):
    _fields_ = [(n, ct.c_double) for n in ("Ctau_t", "Ctau_v", "Ctau_C", "Ctau_M", "Ctau_E")]


class NewtonianContext(ct.Structure):
    """Binary layout of struct NewtonianIdealGasContext_ (qfunctions/newtonian_types.h:29-47)."""

    _fields_ = [
        ("gas", Gas),
        ("g", ct.c_double * 3),
        ("time", ct.c_double),
        ("ijacobian_time_shift", ct.c_double),
        ("is_implicit", ct.c_bool),
        ("state_var", ct.c_int),
        ("stabilization", ct.c_int),
        ("idl_enable", ct.c_bool),
        ("idl_pressure", ct.c_double),
        ("idl_amplitude", ct.c_double),
        ("idl_start", ct.c_double),
        ("idl_length", ct.c_double),
        ("tau_coeffs", TauCoeffs),
        ("dt", ct.c_double),
        ("divFdiff_method", ct.c_int),
    ]


class HoneeMiMesh(ct.Structure):
    _fields_ = [
        ("nelem", ct.c_int),
        ("P", ct.c_int),
        ("Q", ct.c_int),
        ("nnodes", ct.c_int),
        ("nxnodes", ct.c_int),
        ("offsets", _ip),
        ("xoffsets", _ip),
        ("xcoord", _dp),
        ("B", _dp),
        ("D", _dp),
        ("Bx", _dp),
        ("Dx", _dp),
        ("qref", _dp),
        ("qweight", _dp),
    ]


def context_from_dict(d: dict) -> NewtonianContext:
    c = NewtonianContext()
    g = d["gas"]
    c.gas = Gas(g["lambda"], g["mu"], g["k"], g["cv"], g["cp"])
    for i in range(3):
        c.g[i] = d.get("g", (0.0, 0.0, 0.0))[i]
    c.time = d.get("time", 0.0)
    c.ijacobian_time_shift = d.get("ijacobian_time_shift", 0.0)
    c.is_implicit = bool(d.get("is_implicit", True))
    c.state_var = int(d.get("state_var", 0))
    c.stabilization = int(d.get("stabilization", 0))
    c.idl_enable = bool(d.get("idl_enable", False))
    c.idl_pressure = d.get("idl_pressure", 0.0)
    c.idl_amplitude = d.get("idl_amplitude", 0.0)
    c.idl_start = d.get("idl_start", 0.0)
    c.idl_length = d.get("idl_length", 0.0)
    t = d["tau"]
    c.tau_coeffs = TauCoeffs(t["Ctau_t"], t["Ctau_v"], t["Ctau_C"], t["Ctau_M"], t["Ctau_E"])
    c.dt = d.get("dt", 1.0)
    c.divFdiff_method = 0
    return c


_libs: dict = {}


def load_libs():
    """Load libceed.so and libhonee_setup.so from the package's lib/ directory (raises if they are not built)."""
    if "setup" not in _libs:
        lib_dir = os.environ.get("HONEE_MI355X_LIB_DIR", LIB_DIR)
        ceed_path = os.path.join(lib_dir, "libceed.so")
        setup_path = os.path.join(lib_dir, "libhonee_setup.so")
        for p in (ceed_path, setup_path):
            if not os.path.exists(p):
                raise RuntimeError(f"MI355X backend library {p} is not built; run __graft_entry__.build()")
        _libs["ceed"] = ct.CDLL(ceed_path, mode=ct.RTLD_GLOBAL)
        lib = ct.CDLL(setup_path)
        lib.HoneeMiLastError.restype = ct.c_char_p
        lib.HoneeMiProblemError.restype = ct.c_char_p
        lib.HoneeMiProblemError.argtypes = [ct.c_void_p]
        lib.HoneeMiProblemGetCeed.restype = ct.c_void_p
        lib.HoneeMiProblemGetCeed.argtypes = [ct.c_void_p]
        lib.HoneeMiProblemGetOperator.restype = ct.c_void_p
        lib.HoneeMiProblemGetOperator.argtypes = [ct.c_void_p, ct.c_int]
        lib.HoneeMiApplyIFunction.argtypes = [ct.c_void_p, ct.c_int, ct.c_double, ct.c_double, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiApplyIJacobian.argtypes = [ct.c_void_p, ct.c_int, ct.c_double, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiIJacobianDiagonal.argtypes = [ct.c_void_p, ct.c_int, ct.c_double, ct.c_void_p]
        lib.HoneeMiProblemAddBoundary.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiProblemSetEssential.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiGetBoundaryData.argtypes = [ct.c_void_p, ct.c_int, ct.c_int, ct.c_void_p]
        lib.HoneeMiIJacobianPointBlockDiagonal.argtypes = [ct.c_void_p, ct.c_int, ct.c_double, ct.c_void_p]
        lib.HoneeMiMassPointBlockDiagonal.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiApplyRHS.argtypes = [ct.c_void_p, ct.c_int, ct.c_double, ct.c_double, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiApplyMass.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiMassDiagonal.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiProblemIsExplicit.argtypes = [ct.c_void_p]
        lib.HoneeMiProblemVolume.argtypes = [ct.c_void_p]
        lib.HoneeMiProblemVolume.restype = ct.c_double
        lib.HoneeMiMonitorKE.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, _dp]
        lib.HoneeMiGetQData.argtypes = [ct.c_void_p, ct.c_void_p]
        lib.HoneeMiGetJacData.argtypes = [ct.c_void_p, ct.c_void_p]
        lib.HoneeMiGetJacData32.argtypes = [ct.c_void_p, ct.c_void_p]
        lib.HoneeMiProblemSetupSgsDD.argtypes = [ct.c_void_p, ct.c_void_p]
        lib.HoneeMiSgsGridAnisotropy.argtypes = [ct.c_void_p, ct.c_void_p, _ip]
        lib.HoneeMiVelocityGradientProjection.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiSgsNodalStress.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiApplySgsResidual.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiProblemSetupDiffFilter.argtypes = [ct.c_void_p, ct.c_void_p]
        lib.HoneeMiSetFilterWidthScaling.argtypes = [ct.c_void_p, _dp]
        lib.HoneeMiDiffFilterRHS.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiDiffFilterLHS.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        lib.HoneeMiDiffFilterApply.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_void_p, _ip]
        lib.HoneeMiProblemSetupSgsTraining.argtypes = [ct.c_void_p]
        lib.HoneeMiSgsTrainingData.argtypes = [ct.c_void_p, ct.c_int, ct.c_void_p, ct.c_double, ct.c_void_p, _ip]
        lib.HoneeMiSgsTrainingFields.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_void_p, ct.c_void_p]
        c = _libs["ceed"]
        c.CeedMiSetKernelTiming.argtypes = [ct.c_void_p, ct.c_bool]
        c.CeedMiResetKernelTimes.argtypes = [ct.c_void_p]
        c.CeedMiGetKernelTimes.argtypes = [ct.c_void_p, ct.c_int, ct.c_char_p, ct.c_size_t, _ip, _dp, _ip]
        c.CeedMiTimeRegionBegin.argtypes = [ct.c_void_p]
        c.CeedMiTimeRegionEnd.argtypes = [ct.c_void_p, ct.c_char_p]
        c.CeedMiSynchronize.argtypes = [ct.c_void_p]
        _libs["setup"] = lib
    return _libs["ceed"], _libs["setup"]


class HoneeMiSgsDDParams(ct.Structure):
    _fields_ = [("alpha", ct.c_double), ("b1", _dp), ("b2", _dp), ("w1", _dp), ("w2", _dp), ("out_scaling", _dp), ("bh", _dp), ("wh", _dp)]


class HoneeMiDiffFilterOptions(ct.Structure):
    _fields_ = [("grid_based_width", ct.c_int), ("width_scaling", ct.c_double * 3), ("kernel_scaling", ct.c_double),
                ("friction_length", ct.c_double), ("damping_function", ct.c_int), ("damping_constant", ct.c_double),
                ("ksp_rtol", ct.c_double), ("ksp_atol", ct.c_double), ("ksp_max_it", ct.c_int)]


def _ptr(a):
    """Pointer of a numpy array (host) or a torch tensor (device)."""
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(ct.c_void_p)
    return ct.c_void_p(a.data_ptr())


class NewtonianProblem:
    """Newtonian IFunction/IJacobian operators on the MI355X backend for a box mesh."""

    def __init__(self, mesh: BoxMesh, tab: Tabulation, ctx: NewtonianContext, state_var: int, resource: str = "/gpu/hip/mi355x"):
        self.ceedlib, self.lib = load_libs()
        self.mesh, self.tab = mesh, tab
        self.ctx = ctx
        self._keep = dict(
            offsets=np.ascontiguousarray(mesh.offsets, dtype=np.int32),
            xoffsets=np.ascontiguousarray(mesh.xoffsets, dtype=np.int32),
            xcoord=np.ascontiguousarray(mesh.xcoord, dtype=np.float64),
            B=np.ascontiguousarray(tab.B), D=np.ascontiguousarray(tab.D),
            Bx=np.ascontiguousarray(tab.Bx), Dx=np.ascontiguousarray(tab.Dx),
            qref=np.ascontiguousarray(tab.qref), qweight=np.ascontiguousarray(tab.qweight),
        )
        k = self._keep
        d = lambda a: a.ctypes.data_as(_dp)  # noqa: E731
        self.m = HoneeMiMesh(mesh.nelem, tab.P, tab.Q, mesh.nnodes, k["xcoord"].shape[0], k["offsets"].ctypes.data_as(_ip),
                             k["xoffsets"].ctypes.data_as(_ip), d(k["xcoord"]), d(k["B"]), d(k["D"]), d(k["Bx"]), d(k["Dx"]),
                             d(k["qref"]), d(k["qweight"]))
        self.handle = ct.c_void_p()
        ierr = self.lib.HoneeMiProblemCreate(resource.encode(), ct.byref(self.m), ct.byref(ctx), int(state_var), ct.byref(self.handle))
        if ierr:
            raise RuntimeError(f"HoneeMiProblemCreate failed ({ierr}): {self.lib.HoneeMiLastError().decode()}")
        self.njac = 14 + (1 if ctx.idl_enable else 0)
        self.nelem, self.Q3, self.nnodes = mesh.nelem, tab.Q**3, mesh.nnodes

    def _check(self, ierr, what):
        if ierr:
            raise RuntimeError(f"{what} failed ({ierr}): {self.lib.HoneeMiProblemError(self.handle).decode()}")

    @property
    def ceed(self):
        return self.lib.HoneeMiProblemGetCeed(self.handle)

    def ifunction(self, q, qdot, g, time=0.0, dt=None, mem=CEED_MEM_HOST):
        dt = self.ctx.dt if dt is None else dt
        self._check(self.lib.HoneeMiApplyIFunction(self.handle, mem, time, dt, _ptr(q), _ptr(qdot), _ptr(g)), "IFunction")

    def ijacobian(self, dq, v, shift, mem=CEED_MEM_HOST):
        self._check(self.lib.HoneeMiApplyIJacobian(self.handle, mem, shift, _ptr(dq), _ptr(v)), "IJacobian")

    def ijacobian_diagonal(self, diag, shift, mem=CEED_MEM_HOST):
        self._check(self.lib.HoneeMiIJacobianDiagonal(self.handle, mem, shift, _ptr(diag)), "IJacobian diagonal")

    def rhs(self, q, r, time=0.0, dt=None, mem=CEED_MEM_HOST):
        """Explicit RHS composite (volume RHSFunction_Newtonian + boundary faces); explicit problems only."""
        dt = self.ctx.dt if dt is None else dt
        self._check(self.lib.HoneeMiApplyRHS(self.handle, mem, time, dt, _ptr(q), _ptr(r)), "RHS")

    def mass(self, q, x, y, mem=CEED_MEM_HOST):
        """Mass matrix action y = M(q) x (SUPG-stabilized for SUPG problems, else Mass_5); explicit problems only."""
        self._check(self.lib.HoneeMiApplyMass(self.handle, mem, _ptr(q), _ptr(x), _ptr(y)), "Mass")

    def mass_diagonal(self, q, diag, mem=CEED_MEM_HOST):
        self._check(self.lib.HoneeMiMassDiagonal(self.handle, mem, _ptr(q), _ptr(diag)), "Mass diagonal")

    @property
    def explicit(self) -> bool:
        return bool(self.lib.HoneeMiProblemIsExplicit(self.handle))

    @property
    def volume(self) -> float:
        return float(self.lib.HoneeMiProblemVolume(self.handle))

    def monitor_ke(self, q, mem=CEED_MEM_HOST) -> np.ndarray:
        """Sums of the total kinetic energy monitor QFunction over the domain (not divided by the volume)."""
        out = np.zeros(5)
        self._check(self.lib.HoneeMiMonitorKE(self.handle, mem, _ptr(q), out.ctypes.data_as(_dp)), "MonitorKE")
        return out

    def ijacobian_point_block_diagonal(self, pb, shift, mem=CEED_MEM_HOST):
        """Point-block diagonal [nnodes][5 (out)][5 (in)] of the shifted IJacobian composite."""
        self._check(self.lib.HoneeMiIJacobianPointBlockDiagonal(self.handle, mem, shift, _ptr(pb)), "IJacobian point-block diagonal")

    def mass_point_block_diagonal(self, q, pb, mem=CEED_MEM_HOST):
        self._check(self.lib.HoneeMiMassPointBlockDiagonal(self.handle, mem, _ptr(q), _ptr(pb)), "Mass point-block diagonal")

    def setup_sgs_dd(self, model: dict):
        """Enable the data-driven SGS model (problems/sgs_dd_model.c:652-737) with the parameters of
        pyhonee.sgs.load_dd_model; every later IFunction adds the SGS residual."""
        arrs = dict(b1=np.ascontiguousarray(model["b1"], dtype=np.float64), b2=np.ascontiguousarray(model["b2"], dtype=np.float64),
                    w1=np.ascontiguousarray(np.asarray(model["W1"]).T, dtype=np.float64),  # w1.dat order [input][neuron]
                    w2=np.ascontiguousarray(np.asarray(model["W2"]).T, dtype=np.float64),  # w2.dat order [neuron][output]
                    out_scaling=np.ascontiguousarray(model["bounds"], dtype=np.float64))
        if "Wh" in model:  # second hidden layer: [first-layer neuron][second-layer neuron]
            arrs["bh"] = np.ascontiguousarray(model["bh"], dtype=np.float64)
            arrs["wh"] = np.ascontiguousarray(np.asarray(model["Wh"]).T, dtype=np.float64)
        ptr = lambda k: arrs[k].ctypes.data_as(_dp) if k in arrs else None  # noqa: E731
        prm = HoneeMiSgsDDParams(float(model["alpha"]), *(ptr(k) for k in ("b1", "b2", "w1", "w2", "out_scaling", "bh", "wh")))
        self._sgs_arrays = arrs  # alive until the setup has copied them
        self._check(self.lib.HoneeMiProblemSetupSgsDD(self.handle, ct.byref(prm)), "SetupSgsDD")

    def sgs_grid_anisotropy(self):
        """Projected grid anisotropy field [nnodes][7] (KM tensor, delta) and the CG iteration count."""
        out = np.zeros((self.nnodes, 7))
        its = ct.c_int(0)
        self._check(self.lib.HoneeMiSgsGridAnisotropy(self.handle, _ptr(out), ct.byref(its)), "SgsGridAnisotropy")
        return out, its.value

    def velocity_gradient_projection(self, q, grad_velo, mem=CEED_MEM_HOST):
        self._check(self.lib.HoneeMiVelocityGradientProjection(self.handle, mem, _ptr(q), _ptr(grad_velo)), "VelocityGradientProjection")

    def sgs_nodal_stress(self, q, grad_velo, km_sgs, mem=CEED_MEM_HOST):
        self._check(self.lib.HoneeMiSgsNodalStress(self.handle, mem, _ptr(q), _ptr(grad_velo), _ptr(km_sgs)), "SgsNodalStress")

    def apply_sgs_residual(self, q, km_sgs, g, mem=CEED_MEM_HOST):
        """g += the SGS residual of IFunction_NodalSgs for the nodal stress km_sgs."""
        self._check(self.lib.HoneeMiApplySgsResidual(self.handle, mem, _ptr(q), _ptr(km_sgs), _ptr(g)), "ApplySgsResidual")

    def setup_diff_filter(self, grid_based_width=False, width_scaling=(1.0, 1.0, 1.0), kernel_scaling=0.1, friction_length=0.0,
                          damping_function=0, damping_constant=25.0, ksp_rtol=0.0, ksp_atol=0.0, ksp_max_it=0):
        """Differential filter (src/differential_filter.c:177-275); defaults as DifferentialFilterSetup's."""
        o = HoneeMiDiffFilterOptions(int(grid_based_width), (ct.c_double * 3)(*width_scaling), kernel_scaling, friction_length,
                                     int(damping_function), damping_constant, ksp_rtol, ksp_atol, int(ksp_max_it))
        self._check(self.lib.HoneeMiProblemSetupDiffFilter(self.handle, ct.byref(o)), "SetupDiffFilter")

    def set_filter_width_scaling(self, scaling):
        arr = np.ascontiguousarray(scaling, dtype=np.float64)
        self._check(self.lib.HoneeMiSetFilterWidthScaling(self.handle, arr.ctypes.data_as(_dp)), "SetFilterWidthScaling")

    def diff_filter_rhs(self, q):
        out = np.zeros(11 * self.nnodes)
        self._check(self.lib.HoneeMiDiffFilterRHS(self.handle, CEED_MEM_HOST, _ptr(q), _ptr(out)), "DiffFilterRHS")
        return out

    def diff_filter_lhs(self, u):
        """The filter operator applied to u [nnodes][11], and its diagonal."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        v, d = np.zeros(11 * self.nnodes), np.zeros(11 * self.nnodes)
        self._check(self.lib.HoneeMiDiffFilterLHS(self.handle, _ptr(u), _ptr(v), _ptr(d)), "DiffFilterLHS")
        return v, d

    def diff_filter_apply(self, q, mem=CEED_MEM_HOST):
        out = np.zeros(11 * self.nnodes)
        its = ct.c_int(0)
        self._check(self.lib.HoneeMiDiffFilterApply(self.handle, mem, _ptr(q), _ptr(out), ct.byref(its)), "DiffFilterApply")
        return out, its.value

    def setup_sgs_training(self):
        self._check(self.lib.HoneeMiProblemSetupSgsTraining(self.handle), "SetupSgsTraining")

    def sgs_training_data(self, q, width_scale, mem=CEED_MEM_HOST):
        """Training pairs [nnodes][12] (six model inputs, six filtered SGS stress components) for one filter width."""
        out = np.zeros(12 * self.nnodes)
        its = ct.c_int(0)
        self._check(self.lib.HoneeMiSgsTrainingData(self.handle, mem, _ptr(q), float(width_scale), _ptr(out), ct.byref(its)), "SgsTrainingData")
        return out.reshape(self.nnodes, 12), its.value

    def sgs_training_fields(self):
        """Filtered field [nnodes][11], projected filtered velocity gradient [nnodes][9] and grid anisotropy [nnodes][7]
        of the last sgs_training_data call."""
        f, g, a = np.zeros((self.nnodes, 11)), np.zeros((self.nnodes, 9)), np.zeros((self.nnodes, 7))
        self._check(self.lib.HoneeMiSgsTrainingFields(self.handle, _ptr(f), _ptr(g), _ptr(a)), "SgsTrainingFields")
        return f, g, a

    BOUNDARY_INTEGRAL, FREESTREAM_HLL, FREESTREAM_HLLC, RIEMANN_OUTFLOW, PRESSURE_OUTFLOW, SLIP = 0, 1, 2, 3, 4, 5

    def add_boundary(self, kind: int, face_elem, face_id, S_inf_Y=None) -> int:
        """Add boundary faces (volume element, local face id) of the given kind to the IFunction/IJacobian composites.
        S_inf_Y: the exterior primitive state (freestream), or {pressure, temperature, recirc, softplus_velocity}
        (outflows)."""
        nodes, xnodes = face_node_tables(self.mesh, face_elem, face_id)
        self._keep.setdefault("faces", []).append((nodes, xnodes))
        yinf = None if S_inf_Y is None else np.ascontiguousarray(S_inf_Y, dtype=np.float64)
        if yinf is not None:
            self._keep["faces"].append(yinf)
        self._check(self.lib.HoneeMiProblemAddBoundary(self.handle, int(kind), len(nodes), _ptr(nodes), _ptr(xnodes),
                                                       None if yinf is None else _ptr(yinf)), "AddBoundary")
        self.nbnd = getattr(self, "nbnd", 0) + 1
        self._bnd_sizes = getattr(self, "_bnd_sizes", []) + [(len(nodes), 11 if kind in (0, 3, 4) else 5)]
        return self.nbnd - 1

    def set_essential(self, dofs, values=None):
        """Essential (strong) boundary conditions on the L-vector entries dofs (5 node + component) with the given
        values (default zero); see HoneeMiProblemSetEssential."""
        d = np.ascontiguousarray(dofs, dtype=np.int32)
        v = None if values is None else np.ascontiguousarray(np.broadcast_to(values, d.shape), dtype=np.float64)
        self._keep["essential"] = (d, v)
        self._check(self.lib.HoneeMiProblemSetEssential(self.handle, len(d), _ptr(d), None if v is None else _ptr(v)), "SetEssential")

    def boundary_data(self, boundary: int, which: int):
        """Surface qdata (which=0, [nfaces][10][Q^2]) or surface Jacobian data (which=1) of a boundary."""
        nf, nj = self._bnd_sizes[boundary]
        Q2 = self.tab.Q ** 2
        out = np.zeros((nf, 10 if which == 0 else nj, Q2))
        self._check(self.lib.HoneeMiGetBoundaryData(self.handle, int(boundary), int(which), _ptr(out)), "GetBoundaryData")
        return out

    def qdata(self):
        out = np.zeros((self.nelem, 10, self.Q3))
        self._check(self.lib.HoneeMiGetQData(self.handle, _ptr(out)), "GetQData")
        return out

    def jac_data(self):
        out = np.zeros((self.nelem, self.njac, self.Q3))
        self._check(self.lib.HoneeMiGetJacData(self.handle, _ptr(out)), "GetJacData")
        return out

    def jac_data32(self):
        out = np.zeros((self.nelem, self.njac, self.Q3), dtype=np.float32)
        self._check(self.lib.HoneeMiGetJacData32(self.handle, _ptr(out)), "GetJacData32")
        return out

    def set_timing(self, enable: bool):
        self.ceedlib.CeedMiSetKernelTiming(self.ceed, enable)

    def region_begin(self):
        """Start a timed region on the backend stream (two events around back-to-back work; see kernel_times)."""
        if self.ceedlib.CeedMiTimeRegionBegin(self.ceed):
            raise RuntimeError("CeedMiTimeRegionBegin failed")

    def region_end(self, name: str):
        if self.ceedlib.CeedMiTimeRegionEnd(self.ceed, name.encode()):
            raise RuntimeError("CeedMiTimeRegionEnd failed")

    def reset_times(self):
        self.ceedlib.CeedMiResetKernelTimes(self.ceed)

    def kernel_times(self) -> dict:
        names = ct.create_string_buffer(4096)
        launches = (ct.c_int * 32)()
        total = (ct.c_double * 32)()
        n = ct.c_int(0)
        self.ceedlib.CeedMiGetKernelTimes(self.ceed, 32, names, 4096, launches, total, ct.byref(n))
        keys = names.value.decode().split(",") if n.value else []
        return {k: (launches[i], total[i]) for i, k in enumerate(keys)}

    def synchronize(self):
        self.ceedlib.CeedMiSynchronize(self.ceed)

    def close(self):
        if self.handle:
            self.lib.HoneeMiProblemDestroy(ct.byref(self.handle))
            self.handle = ct.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
