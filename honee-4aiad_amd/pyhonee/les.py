"""A-posteriori study of the learned SGS closure on decaying isotropic turbulence.

The flow is the random-phase solenoidal Fourier velocity of problems.random_fourier_modes (E(k) ~ k^4 exp(-2 (k/4)^2),
unit RMS per component) at the reference pressure and temperature (Mach 0.1 at U = 1) on the periodic box [0, 2 pi]^3.

* Fine run: the plain method (conservative FP64 state, SUPG off, SSP-RK3 with the lumped mass, acoustic CFL) on an
  n_f^3 p=2 mesh.  At every output time its state is differentially filtered (qfunctions/differential_filter.h,
  grid-based width scaled to the coarse element size) on the fine mesh.
* Coarse LES: n_c^3 p=2, primitive FP64 state, SUPG, BDF2 with Newton-GMRES and the FP32 Jacobian, started from the
  filtered fine initial state; once without a model (the fallback) and once with the data-driven closure in the
  IFunction (src/setupts.c:118-120).
* Training (on a second, independent flow): filtered fine fields sampled at the coarse nodes (the p=2 nodes of a
  uniform box form a uniform grid, and the coarse grid is the fine one with stride n_f / n_c, sampled at several
  offsets), with the velocity gradient and the grid anisotropy of the coarse mesh, as the closure sees them in the
  coarse run; targets are the filtered SGS stress tau_ij = <u_i u_j> - <u_i><u_j> (qfunctions/sgs_dd_training.h:23-60).
  The FP32 network is trained on the GPU.
* Assessment: shell-summed kinetic-energy spectra of the coarse runs against the filtered fine run sampled on the same
  coarse grid, at every output time.

Host-side orchestration and measurement only: every operator, filter, projection, training step and time step runs
on the GPU through the libraries; numpy does the FFTs of the spectra and the initial-condition synthesis.
"""
from __future__ import annotations

import time

import numpy as np

from . import problems as PR
from . import train as TR
from . import ts as TS
from .driver import NewtonianProblem, context_from_dict
from .mesh import box_mesh, tabulate

L = 2 * np.pi


def grid_velocity(n: int, modes) -> np.ndarray:
    """Velocity of random_fourier_modes on the uniform grid x_j = j 2 pi / n, as [n][n][n][3] indexed [z][y][x]
    (the node order of a periodic p=2 box mesh with n / 2 elements per direction).  The modes are placed on the FFT
    grid and transformed back: exact for integer wavevectors with |k_d| < n / 2."""
    kk, d, amp, phase = modes
    ki = np.rint(kk * (L / (2 * np.pi))).astype(int)
    if np.abs(ki).max() >= n // 2:
        raise ValueError(f"modes up to |k| = {np.abs(ki).max()} need more than {n} points per direction")
    U = np.zeros((n, n, n, 3), dtype=np.complex128)
    c = (0.5 * amp * np.exp(1j * phase))[:, None] * d
    np.add.at(U, (ki[:, 2] % n, ki[:, 1] % n, ki[:, 0] % n), c)
    np.add.at(U, (-ki[:, 2] % n, -ki[:, 1] % n, -ki[:, 0] % n), np.conj(c))
    return np.fft.ifftn(U, axes=(0, 1, 2)).real * float(n) ** 3


def energy_spectrum(u: np.ndarray) -> np.ndarray:
    """Shell-summed kinetic energy per unit mass E(k), k = 0 .. n/2, of a periodic velocity u [n][n][n][3] on the box
    [0, 2 pi]^3 (shells |k| rounded to the nearest integer); sum_k E(k) = mean(|u|^2) / 2 up to the corner shells."""
    n = u.shape[0]
    uh = np.fft.fftn(u, axes=(0, 1, 2)) / float(n) ** 3
    e = 0.5 * np.sum(uh.real ** 2 + uh.imag ** 2, axis=-1)
    k = np.fft.fftfreq(n, 1.0 / n)
    kn = np.rint(np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)).astype(int)
    return np.bincount(kn.ravel(), e.ravel(), minlength=n // 2 + 1)[: n // 2 + 1]


def spectrum_error(E: np.ndarray, E_ref: np.ndarray, kmax: int) -> float:
    """Relative l2 difference of two spectra over the shells 1 <= k <= kmax."""
    a, b = E[1:kmax + 1], E_ref[1:kmax + 1]
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _as_grid(v: np.ndarray, n: int, ncomp: int) -> np.ndarray:
    return v.reshape(n, n, n, ncomp)


def _sample(grid: np.ndarray, stride: int, offset) -> np.ndarray:
    oz, oy, ox = offset
    return np.ascontiguousarray(grid[oz::stride, oy::stride, ox::stride]).reshape(-1, grid.shape[-1])


def _gas(mu):
    g = dict(PR.TGV_GAS)
    g["mu"] = float(mu)
    return g


def _context(gas, state_var, stabilization, implicit, dt):
    d = PR.tgv_context_dict(state_var=state_var, stabilization=stabilization, implicit=implicit, dt=dt)
    d["gas"] = dict(gas)
    return d, context_from_dict(d)


def _zero_model():
    """A closure with zero weights: the coarse projections need one set up, not its values."""
    return {"b1": np.zeros(20), "b2": np.zeros(6), "W1": np.zeros((20, 6)), "W2": np.zeros((6, 20)),
            "bounds": np.tile([0.0, 1.0], (6, 1)), "alpha": 0.0}


def _initial_prim(n, modes, ref_Y=PR.TGV_REFERENCE_Y):
    Y = np.tile(np.asarray(ref_Y, dtype=np.float64), (n ** 3, 1))
    Y[:, 1:4] = grid_velocity(n, modes).reshape(-1, 3)
    return Y


def _prim_from_cons(gas, U):
    R = gas["cp"] - gas["cv"]
    rho = U[:, 0]
    u = U[:, 1:4] / rho[:, None]
    T = (U[:, 4] / rho - 0.5 * np.sum(u * u, axis=1)) / gas["cv"]
    Y = np.empty_like(U)
    Y[:, 0] = rho * R * T
    Y[:, 1:4] = u
    Y[:, 4] = T
    return Y


class FineRun:
    """The plain fine run and its differential filter (grid-based width times `width_scale`, i.e. the coarse element
    size) on one mesh."""

    def __init__(self, n_elem: int, gas: dict, width_scale: float, cfl: float = 0.3, log=print):
        self.n_elem, self.n = n_elem, 2 * n_elem
        self.gas, self.width_scale, self.log = gas, float(width_scale), log
        self.mesh = box_mesh((n_elem,) * 3, 2, lengths=(L, L, L))
        self.tab = tabulate(2, 0)
        c_ref = float(np.sqrt(gas["cp"] / gas["cv"] * (gas["cp"] - gas["cv"]) * PR.TGV_REFERENCE_Y[4]))
        self.c_ref, self.cfl = c_ref, cfl
        _, ctx = _context(gas, 1, 0, True, 1e-3)
        self.filt = NewtonianProblem(self.mesh, self.tab, ctx, 1)
        self.filt.setup_diff_filter(grid_based_width=True)
        self.filt.setup_sgs_training()

    def dt_for(self, u_max: float, interval: float):
        """dt = CFL / (sqrt(g) (|u|max + c)), g = dXdx^T dXdx / 4 for affine cube elements, rounded down so that an
        integer number of steps spans the output interval."""
        sqrt_g = 0.5 * 2.0 / (L / self.n_elem)
        dt = self.cfl / (sqrt_g * (u_max + self.c_ref))
        steps = int(np.ceil(interval / dt))
        return interval / steps, steps

    def filtered(self, Y: np.ndarray):
        """Filtered fields of a primitive fine state: training pairs [n^3][12] and filtered state [n^3][11]."""
        train, its = self.filt.sgs_training_data(np.ascontiguousarray(Y.ravel()), self.width_scale)
        f = np.zeros((self.mesh.nnodes, 11))
        self.filt._check(self.filt.lib.HoneeMiSgsTrainingFields(self.filt.handle, f.ctypes.data, None, None), "SgsTrainingFields")
        return train, f, its

    def run(self, Y0: np.ndarray, t_out, on_output):
        """Integrate from the primitive state Y0 through the output times t_out (t_out[0] = 0 is reported before any
        step) with one time step (the acoustic CFL of 1.25 times the initial maximum speed: the flow decays);
        on_output(i, t, Y_fine_prim) at each.  Returns the GPU seconds of the time steps and the step count."""
        _, ctx = _context(self.gas, 0, 0, False, 1e-3)
        prob = NewtonianProblem(self.mesh, self.tab, ctx, 0)
        u_max = float(np.linalg.norm(Y0[:, 1:4], axis=1).max())
        dt, steps = self.dt_for(1.25 * u_max, float(t_out[1] - t_out[0]))
        ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_RK, rk_type=TS.RK_SSP33, dt=dt))
        ts.set_solution(PR.cons_from_prim(self.gas, Y0).ravel(), 0.0)
        on_output(0, 0.0, Y0)
        gpu_s, nsteps = 0.0, 0
        for i, t in enumerate(t_out[1:], start=1):
            prob.synchronize()
            t0 = time.perf_counter()
            ts.step(steps)
            prob.synchronize()
            gpu_s += time.perf_counter() - t0
            nsteps += steps
            Y = _prim_from_cons(self.gas, ts.solution().reshape(-1, 5))
            if not np.all(np.isfinite(Y)):
                raise RuntimeError(f"fine run diverged before t={t}")
            self.log(f"# fine run t={t:.3f}: {nsteps} steps, dt={dt:.3e}, {gpu_s:.2f} GPU s")
            on_output(i, t, Y)
        ts.close()
        prob.close()
        return gpu_s, nsteps

    def close(self):
        self.filt.close()


class CoarseGrid:
    """Coarse p=2 mesh with the closure's projections (velocity gradient, grid anisotropy) for sampling training data."""

    def __init__(self, n_elem: int, gas: dict, model_for_projection: dict):
        self.n_elem, self.n = n_elem, 2 * n_elem
        self.mesh = box_mesh((n_elem,) * 3, 2, lengths=(L, L, L))
        self.tab = tabulate(2, 0)
        _, ctx = _context(gas, 1, 2, True, 1e-3)
        self.ctx = ctx
        self.proj = NewtonianProblem(self.mesh, self.tab, ctx, 1)
        self.proj.setup_sgs_dd(model_for_projection)
        self.A, _ = self.proj.sgs_grid_anisotropy()

    def gradient(self, Yc: np.ndarray) -> np.ndarray:
        G = np.zeros(9 * self.mesh.nnodes)
        self.proj.velocity_gradient_projection(np.ascontiguousarray(Yc.ravel()), G)
        return G.reshape(-1, 9)

    def close(self):
        self.proj.close()


def sample_offsets(stride: int, count: int, seed: int):
    """`count` distinct sampling offsets of the coarse grid within the fine grid (offset 0 first)."""
    rng = np.random.default_rng(seed)
    all_off = [(a, b, c) for a in range(stride) for b in range(stride) for c in range(stride)]
    rest = [o for o in all_off if o != (0, 0, 0)]
    return [(0, 0, 0)] + [rest[i] for i in rng.permutation(len(rest))[: max(0, count - 1)]]


def coarse_run(coarse: CoarseGrid, gas, Yc0, t_out, dt, model=None, snes_rtol=1e-3, ksp_rtol=0.1, pc_lag=50,
               resource="/gpu/hip/mi355x:jacobian_precision=fp32", mf_operator=False, linesearch="norm_slope", log=print):
    """BDF2 (Newton-GMRES(30) with point-block Jacobi and the FP32 Jacobian) on the coarse mesh from Yc0, with the
    closure in the IFunction when a model is given.  IJacobian_Newtonian leaves the closure out, so Newton converges
    only linearly with it.  The tolerances are LES-sized (relative to each step's initial residual) and the same for
    both runs; pc_lag is PETSc's -snes_lag_preconditioner.  linesearch: "bt" (PETSc's default), "basic" or
    "norm_slope" (quadratic backtracking with the slope -||F||^2, the default here).  With the closure left out of the
    Jacobian, the Newton direction need not decrease ||F||: PETSc's bt, whose slope F . (J dx) sees only the modelled
    part, fails its line search at t = 0.9 of the 24^3 study, and basic stalls on the 8^3 test case.  Returns per-output
    spectra, GPU seconds and solver counts."""
    d, ctx = _context(gas, 1, 2, True, dt)
    prob = NewtonianProblem(coarse.mesh, coarse.tab, ctx, 1, resource)
    if model is not None:
        prob.setup_sgs_dd(model)
    ts = TS.TimeStepper(prob, TS.default_options(type=TS.TS_BDF, bdf_order=2, dt=dt, ksp_pc=TS.PC_PBJACOBI, snes_rtol=snes_rtol,
                                                 ksp_rtol=ksp_rtol, pc_lag=pc_lag, snes_mf_operator=int(mf_operator),
                                                 snes_linesearch_type={"bt": 0, "basic": 1, "norm_slope": 2}[linesearch]))
    ts.set_solution(np.ascontiguousarray(Yc0.ravel()), 0.0)
    spectra = [energy_spectrum(_as_grid(Yc0[:, 1:4], coarse.n, 3))]
    gpu_s, nsteps, t_prev = 0.0, 0, 0.0
    for t in t_out[1:]:
        steps = int(round((t - t_prev) / dt))
        prob.synchronize()
        t0 = time.perf_counter()
        ts.step(steps)
        prob.synchronize()
        gpu_s += time.perf_counter() - t0
        nsteps += steps
        Y = ts.solution().reshape(-1, 5)
        if not np.all(np.isfinite(Y)):
            raise RuntimeError(f"coarse run ({'closure' if model is not None else 'no model'}) diverged before t={t}")
        spectra.append(energy_spectrum(_as_grid(Y[:, 1:4], coarse.n, 3)))
        t_prev = t
    st = ts.stats()
    log(f"# coarse run ({'closure' if model is not None else 'no model'}): {nsteps} steps, {gpu_s:.2f} GPU s, "
        f"{st['newton_its']} Newton, {st['krylov_its']} GMRES its, {st['line_search_backtracks']} backtracks")
    ts.close()
    prob.close()
    return {"spectra": spectra, "gpu_s": gpu_s, "steps": nsteps, "newton_its": st["newton_its"], "krylov_its": st["krylov_its"],
            "backtracks": st["line_search_backtracks"]}


def study(n_fine=96, n_coarse=24, mu=5e-3, t_end=1.0, n_out=4, seed_eval=2024, seed_train=2025, train_offsets=8,
          coarse_dt=0.025, snes_rtol=1e-3, ksp_rtol=0.1, pc_lag=50, epochs=30, batch=4096, lr=3e-3, kmax_frac=0.5,
          dims=(6, 20, 20, 6), closure_mf_operator=False, linesearch="norm_slope", log=print) -> dict:
    """Run the whole study; returns a JSON-ready summary.  dims: the closure network (6-20-20-6 by default; (6, 20, 6)
    is the reference model's shape)."""
    if n_fine % n_coarse:
        raise ValueError("the coarse mesh must nest in the fine one")
    stride = n_fine // n_coarse
    gas = _gas(mu)
    t_out = np.linspace(0.0, t_end, n_out + 1)
    if np.any(np.abs(np.diff(t_out) / coarse_dt - np.rint(np.diff(t_out) / coarse_dt)) > 1e-9):
        raise ValueError("the coarse time step must divide the output interval")
    fine = FineRun(n_fine, gas, width_scale=stride, log=log)
    coarse = CoarseGrid(n_coarse, gas, _zero_model())
    nf, nc = fine.n, coarse.n

    # --- training flow: filtered fine fields at the coarse nodes of several offsets, every output time after t=0
    Xs, Ss = [], []
    offsets = sample_offsets(stride, train_offsets, seed_train)

    def collect(i, t, Y):
        if i == 0:
            return
        train, f, _ = fine.filtered(Y)
        tg, fg = _as_grid(train, nf, 12), _as_grid(f, nf, 11)
        for off in offsets:
            fc = _sample(fg, stride, off)
            G = coarse.gradient(np.ascontiguousarray(fc[:, :5]))
            X, S = TR.model_samples(_sample(tg, stride, off), fc, G, coarse.A, coarse.ctx.gas)
            Xs.append(X)
            Ss.append(S)

    log(f"# training flow (seed {seed_train}) on {n_fine}^3 p=2")
    train_gpu_s, _ = fine.run(_initial_prim(nf, PR.random_fourier_modes(seed_train)), t_out, collect)
    X, S = np.concatenate(Xs), np.concatenate(Ss)
    bounds, Tn = TR.normalize_targets(S)
    mlp = TR.Mlp(alpha=0.01, seed=3, dims=dims)
    t0 = time.perf_counter()
    hist = mlp.train(X, Tn, TR.options(lr=lr, batch=batch, epochs=epochs, seed=4))
    t_train = time.perf_counter() - t0
    model = mlp.closure_model(bounds)
    mlp.close()
    log(f"# trained on {len(X)} samples: loss {hist[0]:.4e} -> {hist[-1]:.4e} in {t_train:.2f} s")

    # --- evaluation flow: the truth (filtered fine spectra on the coarse grid) and the coarse initial state
    truth, Yc0 = [], {}

    def assess(i, t, Y):
        _, f, its = fine.filtered(Y)
        fc = _sample(_as_grid(f, nf, 11), stride, (0, 0, 0))
        if i == 0:
            Yc0["Y"] = np.ascontiguousarray(fc[:, :5])
        truth.append(energy_spectrum(_as_grid(fc[:, 1:4], nc, 3)))

    log(f"# evaluation flow (seed {seed_eval})")
    fine_gpu_s, fine_steps = fine.run(_initial_prim(nf, PR.random_fourier_modes(seed_eval)), t_out, assess)
    fine.close()

    runs = {name: coarse_run(coarse, gas, Yc0["Y"], t_out, coarse_dt, m, snes_rtol, ksp_rtol, pc_lag, mf_operator=mf, linesearch=linesearch,
                             log=log)
            for name, m, mf in (("no_model", None, False), ("closure", model, closure_mf_operator))}
    coarse.close()

    kmax = max(1, int(kmax_frac * nc / 2))
    out = {"config": (f"decaying isotropic turbulence, E(k) ~ k^4 exp(-2 (k/4)^2), u_rms 1, Mach 0.1, mu {mu:g}; fine {n_fine}^3 p=2 "
                      f"conservative SSP-RK3 (CFL {fine.cfl}), filter width = coarse element; coarse {n_coarse}^3 p=2 primitive SUPG "
                      f"BDF2 dt {coarse_dt:g} FP32 Jacobian, Newton rtol {snes_rtol:g} ({linesearch} line search), GMRES rtol {ksp_rtol:g}, "
                      f"preconditioner lag {pc_lag}; "
                      f"closure network {'-'.join(str(n) for n in dims)}; outputs at t = {', '.join(f'{t:g}' for t in t_out)}; spectra over "
                      f"1 <= k <= {kmax}"),
           "fine_gpu_s": fine_gpu_s, "fine_steps": fine_steps, "training_flow_gpu_s": train_gpu_s,
           "training_samples": int(len(X)), "training_s": t_train, "training_loss": [float(hist[0]), float(hist[-1])],
           "filtered_ke": [float(E.sum()) for E in truth]}
    for name, r in runs.items():
        errs = [spectrum_error(E, Et, kmax) for E, Et in zip(r["spectra"], truth)]
        ke = [float(E.sum()) for E in r["spectra"]]
        out[name] = {"gpu_s": r["gpu_s"], "steps": r["steps"], "newton_its": r["newton_its"], "krylov_its": r["krylov_its"],
                     "line_search_backtracks": r["backtracks"],
                     "spectrum_error": errs, "ke": ke,
                     "ke_error": [abs(a - b) / b for a, b in zip(ke, out["filtered_ke"])],
                     "speedup_vs_fine": fine_gpu_s / r["gpu_s"],
                     "les_error_x_gpu_s": errs[-1] * r["gpu_s"]}
        if name == "closure":  # the one-off costs of the closure, charged to this one run
            out[name]["speedup_vs_fine_incl_training"] = fine_gpu_s / (r["gpu_s"] + t_train)
            out[name]["speedup_vs_fine_incl_training_and_data"] = fine_gpu_s / (r["gpu_s"] + t_train + train_gpu_s)
    out["spectra_t_end"] = {"k": list(range(kmax + 1)), "filtered_fine": truth[-1][: kmax + 1].tolist(),
                            "no_model": runs["no_model"]["spectra"][-1][: kmax + 1].tolist(),
                            "closure": runs["closure"]["spectra"][-1][: kmax + 1].tolist()}
    return out
