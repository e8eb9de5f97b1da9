"""Benchmark: GDoF/s of HONEE's Newtonian residual (IFunction) + Jacobian (IJacobian) apply on one MI355X.

Workload (the implicit SUPG Taylor-Green configuration): 32^3 periodic hexahedra on [0, 2pi]^3, p=3, Q=4 Gauss points,
primitive FP64 state, SUPG with the Taylor-Green YAML Tau coefficients, velocity perturbed by seeded uniform noise of
amplitude 1e-3 U0 (seed 1600).  DoF = 5 x 884,736 nodes = 4,423,680.

A step is one IFunction apply followed by one IJacobian apply (FP64), each through CeedOperatorApply of the composite
operators HONEE builds, on vectors resident in HBM.  value = DoF / (seconds per step): whole residual+Jacobian pairs
per second, expressed in DoF.  The per-kernel figures (HIP events on the backend stream) give each operator's
GDoF/s and the HBM roofline fraction of the dominant kernel; the FP32-Jacobian variant is timed alongside.
The CPU baseline times this repository's FP64 restatement of the reference's /cpu/self path (the oracle) on a bounded
sample of the same workload on the host cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "honee-4aiad_amd"))

HBM_PEAK_GBS = 8000.0
# HBM bytes of one IFunction apply at the default workload (fused kernel 551.9 MB + node gather 139.2 MB, two-pass
# scatter), from two rocprofv3 --pmc runs (FETCH_SIZE scaled x2.0, WRITE_SIZE x1.0, the factors the counters show for
# the known-traffic kernels of --calibrate-traffic); the counts and commands are in
# profiles/ifunction_newtonian_tgv32_p3_traffic.md.  A stored figure, not measured by the run that reports it (the
# line says so); other workloads report traffic null.
# With --qdata recompute the same two runs give 520.6 MB (fused kernel 381.5 MB, node gather 139.1 MB).
PMC_TRAFFIC_BYTES = {((32, 32, 32), 3, "twopass", "stored"): 691.1e6, ((32, 32, 32), 3, "twopass", "recompute"): 520.6e6}
METRIC = "GDoF/s Newtonian residual+Jacobian apply (p=3); LES error per GPU-second"
FP64_PEAK_TFLOPS, FP32_PEAK_TFLOPS = 78.6, 157.3
# Floating-point operations per quadrature point of the QFunction bodies at the default workload (primitive
# variables, SUPG, no damping layer, no gravity), counted exactly by running the oracle's restatement over an
# operation-counting scalar; tests/test_oracle_opcount.py checks these figures against that count.
QF_OPS_PER_POINT = {
    "ifunction": {"add": 432, "mul": 472, "div": 12, "sqrt": 3},
    "ijacobian": {"add": 405, "mul": 463, "div": 9, "sqrt": 3},
}
# Executed floating-point operations of the fused kernels at the default workload, per dispatch, from rocprofv3 --pmc
# runs of SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 / _F32 and SQ_INSTS_VALU_MFMA_MOPS_F64 (64 per VALU instruction,
# 2 x 64 per FMA, 512 per MFMA op unit); the counts and commands are in profiles/flops_newtonian_tgv32_p3.md.  A stored
# figure, not measured by the run that reports it.
PMC_EXECUTED_FLOPS = {((32, 32, 32), 3, "ifunction"): 2883584000.0}


def basis_flops_per_element(P, Q, kind):
    """Flops of the tensor-product basis passes on one element with P = Q (collocated quadrature, the workloads timed
    here), five components.  A contraction of length n costs n multiplications and n - 1 additions per output; the
    interpolation and each gradient direction are three passes of n^3 outputs; the transposes are the same passes
    plus three additions per point to sum the value term and the three gradient terms.  IFunction: interpolation and
    gradient of q, interpolation of q-dot, gradient and interpolation transposes.  IJacobian: the same without
    q-dot.  Returns None for P != Q."""
    if P != Q:
        return None
    n3, c = Q**3, 2 * Q - 1
    pass3 = 3 * n3 * c
    per_comp = {"ifunction": 5 * pass3 + 3 * n3, "ijacobian": 4 * pass3 + 3 * n3}[kind]
    return 5 * per_comp


def flops_per_apply(kind, nelem, P, Q):
    qf = sum(QF_OPS_PER_POINT[kind].values())
    b = basis_flops_per_element(P, Q, kind)
    return None if b is None else nelem * (Q**3 * qf + b)


def algorithmic_bytes(nelem, P, Q, nnodes, njac=14, jac_bytes=8, vec_bytes=8):
    """Compulsory HBM traffic of one apply, each stream counted once: L-vectors in and out, qdata, jac_data, offsets, coordinates."""
    Q3, P3 = Q**3, P**3
    lvec = nnodes * 5 * vec_bytes
    qdata = nelem * Q3 * 10 * 8
    offsets = nelem * P3 * 4
    coords = nelem * 8 * 3 * 8
    ifunc = 2 * lvec + qdata + coords + offsets + lvec + nelem * Q3 * njac * jac_bytes
    ijac = lvec + qdata + nelem * Q3 * njac * jac_bytes + offsets + lvec
    return ifunc, ijac


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(mesh_shape_full, degree, seconds=10.0):
    """Time the oracle's IFunction + IJacobian on a 1/64 sample (8^3 elements), with all host threads and then with
    one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle as O
    from pyhonee import mesh as M
    from pyhonee import problems as PR

    shape = tuple(max(1, s // 4) for s in mesh_shape_full)
    mesh = M.box_mesh(shape, degree)
    tab = M.tabulate(degree, 0)
    om = O.Mesh(mesh, tab)
    ctx = PR.tgv_context_dict(state_var=1, stabilization=2, dt=2e-3)
    c = O.newt_from_dict(ctx)
    qd = om.setup_qdata()
    q = PR.taylor_green_ic(mesh, 1).ravel()
    qdot = np.zeros_like(q)
    dq = np.random.default_rng(0).uniform(-1e-3, 1e-3, q.shape)
    c.ijacobian_time_shift = 1.0 / 2e-3

    def timed(threads):
        O.set_num_threads(threads)
        n, t0 = 0, time.perf_counter()
        while True:
            _, jac = O.ifunction(om, c, qd, q, qdot)
            O.ijacobian(om, c, qd, jac, dq)
            n += 1
            if time.perf_counter() - t0 > seconds:
                break
        return n, (time.perf_counter() - t0) / n

    threads = O.max_threads()
    n, dt = timed(threads)
    n1, dt1 = timed(1)
    O.set_num_threads(threads)
    dof = 5 * mesh.nnodes
    return {"value": dof / dt / 1e9, "unit": "GDoF/s", "cores": threads, "kind": "port",
            "value_1thread": dof / dt1 / 1e9, "cpu_model": _cpu_model(), "host_cpus": len(os.sched_getaffinity(0)),
            "sample": f"TGV {shape[0]}x{shape[1]}x{shape[2]} p={degree} (1/{int(np.prod(mesh_shape_full) // np.prod(shape))} of the "
                      f"workload), {n} residual+Jacobian pairs in {n * dt:.1f} s on {threads} threads and {n1} in {n1 * dt1:.1f} s on "
                      f"1 thread, oracle (FP64 C restatement of /cpu/self, OpenMP over elements, gcc -O2)"}


def time_integration_c2(mesh, tab, Y, steps, warm_steps=1):
    """BDF2 on the implicit SUPG workload: Newton (rtol 1e-8) with GMRES(30) (rtol 1e-5) and point-block Jacobi, dt 2e-3,
    all-FP64 and with the FP32 Jacobian inside GMRES.  Timed after warm_steps (which include the BDF start-up) with the
    state resident on the GPU; returns seconds per step and iteration counts per step."""
    from pyhonee import problems as PR
    from pyhonee import ts as TS
    from pyhonee.driver import NewtonianProblem, context_from_dict

    d = PR.tgv_context_dict(state_var=1, stabilization=2, implicit=True, dt=2e-3)
    out, sols = {}, {}
    for label, resource in (("fp64", "/gpu/hip/mi355x"), ("fp32_jacobian", "/gpu/hip/mi355x:jacobian_precision=fp32")):
        print(f"# time integration {label}", file=sys.stderr, flush=True)
        prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1, resource)
        ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_BDF, bdf_order=2, dt=2e-3, ksp_pc=TS.PC_PBJACOBI))
        ts.set_solution(Y.ravel(), 0.0)
        ts.step(warm_steps)
        s0 = ts.stats()
        t0 = time.perf_counter()
        ts.step(steps)
        elapsed = time.perf_counter() - t0
        s1 = ts.stats()
        sols[label] = ts.solution()
        out[label] = {"s_per_step": elapsed / steps, "newton_its_per_step": (s1["newton_its"] - s0["newton_its"]) / steps,
                      "krylov_its_per_step": (s1["krylov_its"] - s0["krylov_its"]) / steps, "final_fnorm": s1["last_fnorm"]}
        ts.close()
        prob.close()
    out["fp32_jacobian_vs_fp64_solution_rel"] = float(np.linalg.norm(sols["fp32_jacobian"] - sols["fp64"]) / np.linalg.norm(sols["fp64"]))
    out["config"] = (f"TGV {mesh.nelem} hex p={tab.P - 1}, primitive, SUPG, BDF2 dt=2e-3, Newton rtol 1e-8 + GMRES(30) rtol 1e-5 + "
                     f"point-block Jacobi; {warm_steps} warm-up step(s), {steps} timed")
    return out


def explicit_c1(steps=100, cfl=0.3, cpu=True):
    """The CPU-runnable explicit configuration: TGV 8^3 periodic hex p=2 Q=3, conservative, SUPG off, Shu-Osher SSP-RK3
    with the lumped mass (HONEE's default mass KSP), `steps` steps at CFL 0.3 with
    dt = CFL / max_q (sqrt(u^T g u) + c sqrt(lambda_max(g))), g = dXdx^T dXdx / 4 (monitor_cfl.h plus the acoustic speed).
    GPU seconds for the run, and the oracle's seconds on the host for the same run."""
    from pyhonee import mesh as M
    from pyhonee import problems as PR
    from pyhonee import ts as TS
    from pyhonee.driver import NewtonianProblem, context_from_dict

    L = 2 * np.pi
    mesh = M.box_mesh((8, 8, 8), 2, lengths=(L, L, L))
    tab = M.tabulate(2, 0)
    q0 = PR.taylor_green_ic(mesh, 0).ravel()
    gas = PR.TGV_GAS
    Yn = PR.nodal_field(mesh, lambda x: PR.taylor_green_prim(x))
    c_max = float(np.sqrt(gas["cp"] / gas["cv"] * (gas["cp"] - gas["cv"]) * Yn[:, 4]).max())
    u_max = float(np.linalg.norm(Yn[:, 1:4], axis=1).max())
    sqrt_g = 0.5 * 2.0 / (L / 8)  # affine cube elements: dXdx = 2/h I
    dt = cfl / (sqrt_g * (u_max + c_max))
    d = PR.tgv_context_dict(state_var=0, stabilization=0, implicit=False, dt=dt)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 0)
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_RK, rk_type=TS.RK_SSP33, dt=dt))
    ts.set_solution(q0, 0.0)
    ts.step(1)
    ts.set_solution(q0, 0.0)
    t0 = time.perf_counter()
    ts.step(steps)
    gpu_s = time.perf_counter() - t0
    ke = ts.monitor_ke()[0] / prob.volume
    out = {"config": f"TGV 8^3 p=2 conservative, SUPG off, SSP-RK3, lumped mass, {steps} steps, dt={dt:.5g} (CFL {cfl})",
           "gpu_s": gpu_s, "gpu_ms_per_step": gpu_s / steps * 1e3, "final_kinetic_energy": ke}
    q_gpu = ts.solution()
    ts.close()
    prob.close()
    if cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle as O

        om = O.Mesh(mesh, tab)
        oc = O.newt_from_dict(d)
        qd = om.setup_qdata()
        A, b, cc = O.tableau_ref("ssp33")
        t0 = time.perf_counter()
        q_cpu, _, _ = O.rk(om, oc, qd, q0, dt, steps, A, b, cc, gmres=False, rowsum=True)
        out["cpu_oracle_s"] = time.perf_counter() - t0
        out["cpu_threads"] = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["gpu_vs_oracle_rel"] = float(np.linalg.norm(q_gpu - q_cpu) / np.linalg.norm(q_cpu))
    return out


def faces_c3(steps=20, warmup=3):
    """Applies with boundary faces on the wall-bounded configuration: channel 4pi x 2 x 4pi/3 on 48 x 32 x 24 hexes,
    wall-normal stretching 1.1, p=3 Q=4, primitive FP64, SUPG, Re_tau 180 (mu = 2 / 180^2 at umax = H = 1), Poiseuille
    state with 10% seeded Fourier perturbations, consistent boundary-integral faces on both walls.  Reports the
    IFunction and IJacobian composite applies (volume + faces) in GDoF/s and the per-kernel times; the volume-only
    figures are the periodic workload's.  Then, for every face kind (boundary integral, freestream HLL and HLLC,
    Riemann and pressure outflow, slip) on the same walls, the per-launch times of its face kernels in the IFunction,
    the IJacobian and the Jacobi diagonal assembly."""
    import torch

    from pyhonee import mesh as M
    from pyhonee import problems as PR
    from pyhonee.driver import CEED_MEM_DEVICE, NewtonianProblem, context_from_dict

    lengths = (4 * np.pi, 2.0, 4 * np.pi / 3)
    mesh = M.box_mesh((48, 32, 24), 3, lengths=lengths, periodic=(True, False, True), stretch=(1.0, 1.1, 1.0),
                      symmetric_stretch=(False, True, False))
    tab = M.tabulate(3, 0)
    d = PR.channel_context_dict(2.0 / 180**2, tau=PR.default_tau(3), dt=2e-3)
    Y = PR.channel_ic(mesh, d["gas"], seed=180, lengths=lengths)
    dof = Y.size
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(180)
    q = torch.tensor(Y.ravel(), dtype=torch.float64, device=dev)
    qdot = torch.tensor(rng.uniform(-1e-3, 1e-3, Y.size), dtype=torch.float64, device=dev)
    dq = torch.tensor(rng.uniform(-1e-3, 1e-3, Y.size), dtype=torch.float64, device=dev)
    g, v = torch.zeros_like(q), torch.zeros_like(q)
    prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1)
    fe, fi = M.box_boundary(mesh, [2, 3])
    prob.add_boundary(0, fe, fi)
    out = {"config": "channel 48x32x24 p=3 Q=4, stretching 1.1, primitive, SUPG, boundary-integral faces on both walls (3072 faces)",
           "dof": dof}
    out.update(_apply_times(prob, q, qdot, g, dq, v, dof, steps, warmup))
    prob.close()
    # every face kind on the same walls: per-launch times of the face kernels inside the composite applies and of the
    # composite's diagonal assembly (volume + faces)
    P0, theta0 = float(Y[..., 0].mean()), float(Y[..., 4].mean())
    params = {1: np.array([P0, 1.0, 0.0, 0.0, theta0]), 2: np.array([P0, 1.0, 0.0, 0.0, theta0]),
              3: np.array([P0, theta0, 1.0, 1e-2]), 4: np.array([P0, theta0, 1.0, 1e-2]), 5: None}
    names = {0: "boundary_integral", 1: "freestream_hll", 2: "freestream_hllc", 3: "riemann_outflow", 4: "pressure_outflow",
             5: "slip"}
    diag = torch.zeros_like(q)
    kinds = {}
    for kind in range(6):
        prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1)
        prob.add_boundary(kind, fe, fi, params.get(kind))
        r = _apply_times(prob, q, qdot, g, dq, v, dof, steps, warmup, diag=diag)
        kinds[names[kind]] = {k: {n: t for n, t in r[k]["kernel_ms"].items() if "newtonian" not in n and "gather" not in n}
                              for k in ("ifunction", "ijacobian", "jacobi_diagonal")}
        prob.close()
    out["face_kernel_ms"] = kinds
    return out


def _apply_times(prob, q, qdot, g, dq, v, dof, steps, warmup, diag=None):
    """Steady-state composite apply times (host clock over `steps` back-to-back applies) and per-launch kernel times
    (per-kernel events, a separate pass) of the IFunction and IJacobian, and of the Jacobi diagonal when diag is
    given."""
    from pyhonee.driver import CEED_MEM_DEVICE

    out = {}
    fns = [("ifunction", lambda: prob.ifunction(q, qdot, g, mem=CEED_MEM_DEVICE)),
           ("ijacobian", lambda: prob.ijacobian(dq, v, 500.0, mem=CEED_MEM_DEVICE))]
    if diag is not None:
        fns.append(("jacobi_diagonal", lambda: prob.ijacobian_diagonal(diag, 500.0, mem=CEED_MEM_DEVICE)))
    for name, fn in fns:
        for _ in range(warmup):
            fn()
        prob.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        prob.synchronize()
        t_apply = (time.perf_counter() - t0) / steps
        prob.set_timing(True)
        prob.reset_times()
        for _ in range(steps):
            fn()
        prob.synchronize()
        kt = prob.kernel_times()
        prob.set_timing(False)
        # ms per launch by kernel: the fused volume kernel, the face kernel and the node gathers of both scatters
        per = {k: v_[1] / v_[0] for k, v_ in kt.items()}
        out[name] = {"apply_ms": t_apply * 1e3, "gdofs_with_faces": dof / t_apply / 1e9, "kernel_ms": per}
    return out


def c5_scale(steps=10, ts_steps=2, warm_steps=1):
    """The scale configuration's residual and Jacobian: TGV at Re 1600 on 96^3 periodic hexahedra, p=4, Q=5
    (56,623,104 nodes, 283,115,520 DoF), primitive FP64 state with seeded velocity noise 1e-3 U0, SUPG.  Steady-state
    IFunction and IJacobian (FP64 and FP32 Jacobian) applies in GDoF/s with the IFunction's roofline fraction, then
    BDF2 (dt 1e-3) with the FP32 Jacobian inside GMRES(30) and point-block Jacobi: seconds per step after a warm-up
    step, iteration counts, and property checks (every Newton solve converged, the state finite, the kinetic energy
    not increasing)."""
    import torch

    from pyhonee import mesh as M
    from pyhonee import problems as PR
    from pyhonee import ts as TS
    from pyhonee.driver import CEED_MEM_DEVICE, NewtonianProblem, context_from_dict

    shape, degree, dt = (96, 96, 96), 4, 1e-3
    t0 = time.perf_counter()
    mesh = M.box_mesh(shape, degree)
    tab = M.tabulate(degree, 0)
    rng = np.random.default_rng(1600)
    Y = PR.nodal_field(mesh, lambda x: PR.taylor_green_prim(x), average=False)
    Y[:, 1:4] += rng.uniform(-1e-3, 1e-3, (mesh.nnodes, 3))
    dof = Y.size
    print(f"# C5 mesh and state: {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    d = PR.tgv_context_dict(state_var=1, stabilization=2, dt=dt, tau=PR.default_tau(degree))
    dev = torch.device("cuda:0")
    q = torch.tensor(Y.ravel(), dtype=torch.float64, device=dev)
    qdot = torch.empty_like(q).uniform_(-1e-3, 1e-3, generator=torch.Generator(device=dev).manual_seed(96))
    dq = torch.empty_like(q).uniform_(-1e-3, 1e-3, generator=torch.Generator(device=dev).manual_seed(97))
    g, v = torch.zeros_like(q), torch.zeros_like(q)
    out = {"config": f"TGV 96^3 periodic hex p=4 Q=5, primitive FP64, SUPG, dt {dt}", "dof": dof, "elements": mesh.nelem}
    b_if, b_ij = algorithmic_bytes(mesh.nelem, tab.P, tab.Q, mesh.nnodes)
    _, b32 = algorithmic_bytes(mesh.nelem, tab.P, tab.Q, mesh.nnodes, jac_bytes=4)
    shift = 1.0 / dt
    for label, resource in (("fp64", "/gpu/hip/mi355x"), ("fp32_jacobian", "/gpu/hip/mi355x:jacobian_precision=fp32")):
        print(f"# C5 {label} applies", file=sys.stderr, flush=True)
        prob = NewtonianProblem(mesh, tab, context_from_dict(d), 1, resource)
        r = {}
        for name, fn in (("ifunction", lambda: prob.ifunction(q, qdot, g, mem=CEED_MEM_DEVICE)),
                         ("ijacobian", lambda: prob.ijacobian(dq, v, shift, mem=CEED_MEM_DEVICE))):
            for _ in range(2):
                fn()
            prob.synchronize()
            prob.reset_times()
            prob.region_begin()
            for _ in range(steps):
                fn()
            prob.region_end(name)
            r[name] = prob.kernel_times()[name][1] / steps * 1e-3
        if label == "fp64":
            out["ifunction_gdofs"] = dof / r["ifunction"] / 1e9
            out["ifunction_roofline_frac"] = b_if / r["ifunction"] / 1e9 / HBM_PEAK_GBS
            out["ifunction_apply_ms"] = r["ifunction"] * 1e3
            out["ifunction_tflops"] = flops_per_apply("ifunction", mesh.nelem, tab.P, tab.Q) / r["ifunction"] / 1e12
            out["ijacobian_tflops"] = flops_per_apply("ijacobian", mesh.nelem, tab.P, tab.Q) / r["ijacobian"] / 1e12
            out["ijacobian_gdofs"] = dof / r["ijacobian"] / 1e9
            out["ijacobian_roofline_frac"] = b_ij / r["ijacobian"] / 1e9 / HBM_PEAK_GBS
            prob.close()
            continue
        out["ijacobian_fp32_gdofs"] = dof / r["ijacobian"] / 1e9
        out["ijacobian_fp32_roofline_frac"] = b32 / r["ijacobian"] / 1e9 / HBM_PEAK_GBS
        if ts_steps > 0:
            print("# C5 BDF2 with the FP32 Jacobian", file=sys.stderr, flush=True)
            ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_BDF, bdf_order=2, dt=dt, ksp_pc=TS.PC_PBJACOBI))
            ts.set_solution(Y.ravel(), 0.0)
            ke0 = float(ts.monitor_ke()[0])
            ts.step(warm_steps)
            s0 = ts.stats()
            t1 = time.perf_counter()
            ts.step(ts_steps)
            el = time.perf_counter() - t1
            s1 = ts.stats()
            ke1 = float(ts.monitor_ke()[0])
            qf = ts.solution()
            out["bdf2_fp32_jacobian"] = {
                "s_per_step": el / ts_steps, "newton_its_per_step": (s1["newton_its"] - s0["newton_its"]) / ts_steps,
                "krylov_its_per_step": (s1["krylov_its"] - s0["krylov_its"]) / ts_steps,
                "newton_converged": bool(s1["last_fnorm"] <= 1e-8 * s1["last_fnorm0"] or s1["last_fnorm"] < 1e-8),
                "last_fnorm": s1["last_fnorm"], "last_fnorm0": s1["last_fnorm0"], "finite": bool(np.isfinite(qf).all()),
                "kinetic_energy_ratio": ke1 / ke0, "steps": warm_steps + ts_steps}
            ts.close()
        prob.close()
    return out


def les_study():
    """The learned closure a posteriori: decaying isotropic turbulence (E(k) ~ k^4 exp(-2 (k/4)^2), u_rms 1, Mach 0.1),
    plain fine run 96^3 p=2 (SSP-RK3) differentially filtered to the coarse element size, network trained on the GPU
    on a second flow's filtered pairs at the coarse nodes, coarse 24^3 p=2 BDF2 LES with the FP32 Jacobian with and
    without the closure.  Reports the coarse runs' shell-spectrum error against the filtered fine run at t = 1 (relative
    l2 over 1 <= k <= 12), their GPU seconds against the fine run's, and LES error x GPU seconds."""
    from pyhonee import les as LES

    out = LES.study(log=lambda m: print(m, file=sys.stderr, flush=True))
    res = {"config": out["config"], "fine_gpu_s": out["fine_gpu_s"], "fine_steps": out["fine_steps"],
           "training_flow_gpu_s": out["training_flow_gpu_s"], "training_samples": out["training_samples"], "training_s": out["training_s"],
           "training_loss": out["training_loss"], "filtered_ke_t_end": out["filtered_ke"][-1], "spectra_t_end": out["spectra_t_end"]}
    for name in ("no_model", "closure"):
        r = out[name]
        res[name] = {"gpu_s": r["gpu_s"], "spectrum_error_t_end": r["spectrum_error"][-1], "spectrum_error": r["spectrum_error"],
                     "ke_error_t_end": r["ke_error"][-1], "speedup_vs_fine": r["speedup_vs_fine"],
                     "les_error_x_gpu_s": r["les_error_x_gpu_s"], "newton_its": r["newton_its"], "krylov_its": r["krylov_its"]}
        for k in ("speedup_vs_fine_incl_training", "speedup_vs_fine_incl_training_and_data"):
            if k in r:
                res[name][k] = r[k]
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--shape", type=int, nargs=3, default=(32, 32, 32))
    ap.add_argument("--degree", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--scatter", choices=("atomic", "twopass"), default="twopass",
                    help="element scatter-add: FP64 atomics, or the deterministic two-pass node gather")
    ap.add_argument("--qdata", choices=("recompute", "stored"), default="stored",
                    help="volume kernels read the quadrature data (default) or recompute it from the element vertices")
    ap.add_argument("--ts-steps", type=int, default=5, help="timed BDF2 steps of the implicit workload (0 skips time integration)")
    ap.add_argument("--no-explicit", action="store_true", help="skip the explicit 8^3 p=2 SSP-RK3 run")
    ap.add_argument("--no-les", action="store_true", help="skip the a-posteriori LES study of the learned closure")
    ap.add_argument("--no-faces", action="store_true", help="skip the channel applies with boundary faces")
    ap.add_argument("--no-c5", action="store_true", help="skip the 96^3 p=4 scale-configuration applies and BDF2 steps")
    ap.add_argument("--calibrate-traffic", action="store_true",
                    help="only run kernels of known HBM traffic (3 reductions reading 1 GiB, 3 fills writing 1 GiB) to scale "
                         "the FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc")
    args = ap.parse_args()
    if args.calibrate_traffic:
        import torch

        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        for _ in range(3):
            s_ = x.sum()
        for _ in range(3):
            x.fill_(2.0)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "3 x sum reading 1073741824 B, 3 x fill writing 1073741824 B", "check": float(s_)}), flush=True)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this benchmark runs on exactly one GPU; multi-GPU runs are not part of this build", "n_gpus": args.gpus}))
        sys.exit(2)

    import torch

    from pyhonee import mesh as M
    from pyhonee import problems as PR
    from pyhonee.driver import CEED_MEM_DEVICE, NewtonianProblem

    shape, degree = tuple(args.shape), args.degree
    print(f"# building mesh {shape} p={degree}", file=sys.stderr, flush=True)
    mesh = M.box_mesh(shape, degree)
    tab = M.tabulate(degree, 0)
    dof = 5 * mesh.nnodes
    ctx = PR.tgv_context(state_var=1, stabilization=2, dt=2e-3)
    rng = np.random.default_rng(1600)
    Y = PR.nodal_field(mesh, lambda x: PR.taylor_green_prim(x))
    Y[:, 1:4] += rng.uniform(-1e-3, 1e-3, (mesh.nnodes, 3))  # seeded velocity noise, amplitude 1e-3 U0
    dev = torch.device("cuda:0")
    q = torch.tensor(Y.ravel(), dtype=torch.float64, device=dev)
    qdot = torch.tensor(rng.uniform(-1e-3, 1e-3, Y.size), dtype=torch.float64, device=dev)
    dq = torch.tensor(rng.uniform(-1e-3, 1e-3, Y.size), dtype=torch.float64, device=dev)
    g = torch.zeros_like(q)
    v = torch.zeros_like(q)
    shift = 1.0 / 2e-3

    results = {}
    base = f"/gpu/hip/mi355x:scatter={args.scatter}:qdata={args.qdata}"
    for label, resource in [("fp64", base)] + ([] if args.no_fp32 else [("fp32", base + ":jacobian_precision=fp32")]):
        print(f"# setting up {label} operators", file=sys.stderr, flush=True)
        prob = NewtonianProblem(mesh, tab, ctx, 1, resource)
        for _ in range(args.warmup):
            prob.ifunction(q, qdot, g, mem=CEED_MEM_DEVICE)
            prob.ijacobian(dq, v, shift, mem=CEED_MEM_DEVICE)
        prob.synchronize()
        torch.cuda.synchronize()
        prob.set_timing(True)
        prob.reset_times()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            prob.ifunction(q, qdot, g, mem=CEED_MEM_DEVICE)
            prob.ijacobian(dq, v, shift, mem=CEED_MEM_DEVICE)
        prob.synchronize()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        times = prob.kernel_times()
        prob.set_timing(False)
        results[label] = dict(elapsed=elapsed, times=times)
        print(f"# {label}: {elapsed / args.steps * 1e3:.3f} ms/step, kernels {times}", file=sys.stderr, flush=True)
        # steady-state apply durations: `steps` back-to-back applies bracketed by two events on the backend stream
        # (events between launches slow the kernels that follow them, so the per-kernel figures above only split an
        # apply into its fused kernel and node gather)
        # (three repetitions, the fastest kept: a host stall inside one repetition leaves the GPU idle between two
        # events, which is not the apply's duration)
        apply_s, apply_reps = {}, {}
        for name, fn in (("ifunction_apply", lambda: prob.ifunction(q, qdot, g, mem=CEED_MEM_DEVICE)),
                         ("ijacobian_apply", lambda: prob.ijacobian(dq, v, shift, mem=CEED_MEM_DEVICE))):
            reps = []
            for _ in range(3):
                prob.reset_times()
                prob.region_begin()
                for _ in range(args.steps):
                    fn()
                prob.region_end(name)
                reps.append(prob.kernel_times()[name][1] / args.steps * 1e-3)
            apply_s[name], apply_reps[name] = min(reps), reps
        results[label]["apply_reps_us"] = {k: [round(x * 1e6, 2) for x in v_] for k, v_ in apply_reps.items()}
        # Jacobi and point-block diagonal assembly of the IJacobian (the preconditioners of the implicit solve)
        diag = torch.zeros_like(q)
        pb = torch.zeros(25 * mesh.nnodes, dtype=torch.float64, device=dev)
        for name, fn, n in (("ijacobian_diagonal", lambda: prob.ijacobian_diagonal(diag, shift, mem=CEED_MEM_DEVICE), args.steps),
                            ("ijacobian_point_block", lambda: prob.ijacobian_point_block_diagonal(pb, shift, mem=CEED_MEM_DEVICE),
                             max(2, args.steps // 10))):
            fn()
            prob.synchronize()
            prob.reset_times()
            prob.region_begin()
            for _ in range(n):
                fn()
            prob.region_end(name + "_assembly")
            apply_s[name] = prob.kernel_times()[name + "_assembly"][1] / n * 1e-3
        results[label]["apply_s"] = apply_s
        if label == "fp64":
            # timing-free pass for the whole-step rate (events add host work between launches)
            t0 = time.perf_counter()
            for _ in range(args.steps):
                prob.ifunction(q, qdot, g, mem=CEED_MEM_DEVICE)
                prob.ijacobian(dq, v, shift, mem=CEED_MEM_DEVICE)
            prob.synchronize()
            torch.cuda.synchronize()
            results[label]["elapsed_untimed"] = time.perf_counter() - t0
        prob.close()

    r64 = results["fp64"]
    ms_per_step = r64["elapsed_untimed"] / args.steps * 1e3
    kt = r64["times"]
    # apply durations: the fused volume kernel plus, in two-pass mode, its node-gather pass (one per apply)
    t_gather = kt["scatter_node_gather"][1] / kt["scatter_node_gather"][0] * 1e-3 if "scatter_node_gather" in kt else 0.0
    t_if_k = kt["ifunction_newtonian"][1] / kt["ifunction_newtonian"][0] * 1e-3
    t_ij_k = kt["ijacobian_newtonian"][1] / kt["ijacobian_newtonian"][0] * 1e-3
    t_if, t_ij = r64["apply_s"]["ifunction_apply"], r64["apply_s"]["ijacobian_apply"]
    b_if, b_ij = algorithmic_bytes(mesh.nelem, tab.P, tab.Q, mesh.nnodes)
    line = {
        "metric": METRIC,
        "value": dof / (ms_per_step * 1e-3) / 1e9,
        "unit": "GDoF/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Taylor-Green state, seed 1600)",
        "config": {"workload": f"TGV {shape[0]}^3 periodic hex, p={degree} Q={tab.Q}, primitive FP64, SUPG; step = IFunction + IJacobian apply",
                   "dof": dof, "elements": mesh.nelem},
        "ifunction_gdofs": dof / t_if / 1e9,
        "ijacobian_gdofs": dof / t_ij / 1e9,
        "roofline": {"bound": "hbm", "kernel": "ifunction_newtonian", "achieved": b_if / t_if / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": b_if / t_if / 1e9 / HBM_PEAK_GBS, "traffic": PMC_TRAFFIC_BYTES.get((shape, degree, args.scatter, args.qdata)),
                     "traffic_source": "stored: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this workload "
                                       "(profiles/ifunction_newtonian_tgv32_p3_traffic.md), not measured by this run",
                     "algorithmic_bytes": b_if, "avg_apply_us": t_if * 1e6,
                     "avg_launch_us": t_if_k * 1e6, "avg_node_gather_us": t_gather * 1e6, "scatter": args.scatter, "qdata": args.qdata,
                     "apply_repetitions_us": r64["apply_reps_us"]["ifunction_apply"],
                     "timing": "avg_apply_us: back-to-back applies between two HIP events on the backend stream (fused kernel + node "
                               "gather), the fastest of three repetitions; avg_launch_us / avg_node_gather_us: per-kernel events"},
    }
    f_if = flops_per_apply("ifunction", mesh.nelem, tab.P, tab.Q)
    if f_if is not None:
        ex = PMC_EXECUTED_FLOPS.get((shape, degree, "ifunction"))
        line["roofline"]["flops"] = {
            "dtype": "f64", "per_launch": f_if, "per_point": f_if / (mesh.nelem * tab.Q**3), "achieved": f_if / t_if / 1e12,
            "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": f_if / t_if / 1e12 / FP64_PEAK_TFLOPS,
            "source": "QFunction: exact count of the oracle's restatement (operation-counting scalar); basis: sum-factorized passes",
            "executed_per_launch": ex,
            "executed_frac": None if ex is None else ex / t_if_k / 1e12 / FP64_PEAK_TFLOPS,
            "executed_source": "stored: rocprofv3 --pmc SQ_INSTS_VALU_*_F64 counts of the fused kernel "
                               "(profiles/flops_newtonian_tgv32_p3.md), over its per-kernel launch time",
        }
    f_ij = flops_per_apply("ijacobian", mesh.nelem, tab.P, tab.Q)
    if f_ij is not None:
        line["ijacobian_tflops"] = f_ij / t_ij / 1e12
    if "fp32" in results:
        t32 = results["fp32"]["apply_s"]["ijacobian_apply"]
        _, b32 = algorithmic_bytes(mesh.nelem, tab.P, tab.Q, mesh.nnodes, jac_bytes=4)
        line["ijacobian_fp32_gdofs"] = dof / t32 / 1e9
        line["ijacobian_fp32_roofline_frac"] = b32 / t32 / 1e9 / HBM_PEAK_GBS
        if f_ij is not None:
            line["ijacobian_fp32_tflops"] = f_ij / t32 / 1e12
    line["ijacobian_roofline_frac"] = b_ij / t_ij / 1e9 / HBM_PEAK_GBS
    # diagonal assembly reads what the IJacobian apply reads except the input vector
    line["assembly"] = {}
    for label, r in results.items():
        a = r["apply_s"]
        b_dg = algorithmic_bytes(mesh.nelem, tab.P, tab.Q, mesh.nnodes, jac_bytes=4 if label == "fp32" else 8)[1] - 8 * dof
        line["assembly"][label] = {"jacobi_diagonal_ms": a["ijacobian_diagonal"] * 1e3,
                                   "jacobi_diagonal_hbm_frac": b_dg / a["ijacobian_diagonal"] / 1e9 / HBM_PEAK_GBS,
                                   "point_block_diagonal_ms": a["ijacobian_point_block"] * 1e3}
    if args.ts_steps > 0:
        line["time_integration"] = time_integration_c2(mesh, tab, Y, args.ts_steps)
    if not args.no_explicit:
        print("# explicit 8^3 p=2 run", file=sys.stderr, flush=True)
        line["explicit_c1"] = explicit_c1(cpu=not args.no_cpu_baseline)
    if not args.no_faces:
        print("# channel applies with faces", file=sys.stderr, flush=True)
        line["faces_c3"] = faces_c3()
    if not args.no_c5:
        line["c5_scale"] = c5_scale()
    if not args.no_les:
        line["les"] = les_study()
    if not args.no_cpu_baseline:
        print("# timing CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(shape, degree)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
