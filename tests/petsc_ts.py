"""The PETSc 3.24 solver stack behind HONEE's implicit goldens, restated in numpy for small global systems (test
infrastructure: it drives either the oracle's operators or the HIP path's, through the callbacks of `Operators`).

HONEE requires PETSc v3.24 (include/navierstokes.h:24); PETSc is not part of the reference tree, so its algorithms are
restated here from their published form and pinned by the reference's fixtures that exercise them
(tests/output/gaussianwave-IDL.bin and gaussianwave-IDL-entropy.bin, examples/navierstokes.c:230,245, with TSALPHA;
channel.bin and channel-prim.bin, examples/navierstokes.c:252-253, with TSBEULER):

* TSALPHA (src/ts/impls/implicit/alpha/alpha1.c): the generalized-alpha method for first-order systems of Jansen,
  Whiting and Hulbert (2000) with alpha_m = (3 - rho) / (2 (1 + rho)), alpha_f = 1 / (1 + rho),
  gamma = 1/2 + alpha_m - alpha_f.  The step solves R(X1) = F(t + alpha_f dt, Xa, Va) / alpha_f = 0 with
  V1 = (X1 - X0) / (gamma dt) + (1 - 1/gamma) V0, Xa = X0 + alpha_f (X1 - X0), Va = V0 + alpha_m (V1 - V0), and the
  IJacobian at shift alpha_m / (alpha_f gamma dt); V0 is the previous step's V1.  The first step starts with
  TSAlpha_Restart: two backward-Euler steps of dt/2 (alpha_m = alpha_f = gamma = 1) and the one-sided difference
  V0 = (-3 X0 + 4 X(t0 + dt/2) - X(t0 + dt)) / dt.
* TSBEULER (TSTHETA with theta 1, src/ts/impls/implicit/theta/theta.c): F(t + dt, X1, (X1 - X0) / dt) = 0 from the
  initial guess X0, with the IJacobian at shift 1/dt.
* SNESNEWTONLS with the cubic backtracking line search (SNESLINESEARCHBT, alpha 1e-4), SNESConvergedDefault
  (atol, rtol relative to the solve's first residual, stol on the step) and -snes_lag_jacobian L
  -snes_lag_jacobian_persists: the Jacobian (operator and preconditioner matrix alike) is assembled when
  (Newton iteration + iterations of earlier solves) % L == 0 and reused otherwise.
* HONEE's Jacobian is the IJacobian operator assembled into the DM's AIJ matrix (MatCeedAssembleCOO,
  src/setupts.c:129-157; the DM matrix type is AIJ on the CPU, examples/navierstokes.c:83-110), in the global vector's
  order with the constrained dofs left out.
* KSPGMRES(30) with left preconditioning, zero initial guess, classical Gram-Schmidt without refinement and Givens
  rotations; KSPConvergedDefault on the preconditioned residual, relative to the first one.
* PCILU(0), the default preconditioner of a sequential AIJ matrix, in natural order on the matrix's pattern.
"""
import numpy as np


def alpha_params(radius):
    """TSAlphaSetRadius: (alpha_m, alpha_f, gamma) of spectral radius rho at infinity."""
    am = 0.5 * (3.0 - radius) / (1.0 + radius)
    af = 1.0 / (1.0 + radius)
    return am, af, 0.5 + am - af


class Operators:
    """Callbacks of one problem, on local (L-)vectors of 5 components per node, and its global numbering.

    residual(t, dt, X_loc, V_loc) -> F_loc: the IFunction composite at state X_loc and time derivative V_loc with the
    solution-time step dt in its context; it keeps the Jacobian data of that evaluation.
    jacobian(shift, dX_loc) -> dF_loc: the IJacobian composite at that data and shift.
    free: local indices of the global vector's entries, in global order; fixed: local values of the constrained
    entries (zero time derivative).

    The Jacobian is either assembled into an AIJ matrix with ILU(0) (pattern: the matrix's nonzero structure, since COO
    preallocation keeps every element-coupled entry, zero or not), or, with point_blocks(shift) -> [nodes][5][5]
    (row: residual component, column: state component), HONEE's -amat_type shell with -pc_type vpbjacobi: the MATCEED
    itself is operator and preconditioner matrix (CreateSolveOperatorsFromMatCeed, src/petsc_ops.c:320-370).  Its
    action then uses the Jacobian data of the latest IFunction evaluation with the shift of the latest Jacobian
    evaluation (MatCeedSetContextReal in FormIJacobian_NS), and VPBJACOBI inverts each node's block of free
    components at that evaluation."""

    def __init__(self, residual, jacobian, free, fixed, pattern=None, point_blocks=None):
        self.residual, self.jacobian = residual, jacobian
        self.free = np.asarray(free)
        self.fixed = np.asarray(fixed, dtype=float)
        self.pattern = None if pattern is None else np.asarray(pattern, dtype=bool)
        self.point_blocks = point_blocks

    def to_local(self, x, fill):
        out = fill.copy()
        out[self.free] = x
        return out

    def assemble(self, shift):
        n = self.free.size
        J = np.zeros((n, n))
        e = np.zeros(self.fixed.size)
        for j, k in enumerate(self.free):
            e[k] = 1.0
            J[:, j] = self.jacobian(shift, e)[self.free]
            e[k] = 0.0
        return J

    def setup(self, shift):
        """(operator, preconditioner) of one Jacobian evaluation at the given shift, as callables on global vectors."""
        if self.point_blocks is None:
            J = self.assemble(shift)
            LU = ilu0(J, self.pattern)
            return (lambda y: J @ y), (lambda r: ilu_solve(LU, r))
        blocks = self.point_blocks(shift)
        node, comp = self.free // 5, self.free % 5
        groups = [np.nonzero(node == n)[0] for n in np.unique(node)]
        inv = [np.linalg.inv(blocks[node[g[0]]][np.ix_(comp[g], comp[g])]) for g in groups]
        zero = np.zeros(self.fixed.size)

        def pc(r):
            z = np.empty_like(r)
            for g, Bi in zip(groups, inv):
                z[g] = Bi @ r[g]
            return z

        return (lambda y: self.jacobian(shift, self.to_local(y, zero))[self.free]), pc


def ilu0(A, pattern):
    """ILU(0) of A on the given nonzero pattern (row-oriented, natural order): returns the packed factors (unit L
    below the diagonal, U on and above it)."""
    n = A.shape[0]
    pat = pattern.copy()
    np.fill_diagonal(pat, True)
    LU = A.copy()
    for i in range(1, n):
        for k in np.nonzero(pat[i, :i])[0]:
            LU[i, k] /= LU[k, k]
            cols = np.nonzero(pat[i, k + 1:])[0] + k + 1
            LU[i, cols] -= LU[i, k] * LU[k, cols]
    return LU


def ilu_solve(LU, b):
    n = b.size
    y = b.astype(float).copy()
    for i in range(n):
        y[i] -= LU[i, :i] @ y[:i]
    for i in range(n - 1, -1, -1):
        y[i] = (y[i] - LU[i, i + 1:] @ y[i + 1:]) / LU[i, i]
    return y


def gmres(A, PC, b, rtol, atol=1e-50, restart=30, max_it=10000):
    """KSPGMRES with left preconditioning (operator A and preconditioner PC as callables) and zero initial guess;
    returns (x, iterations)."""
    n = b.size
    x = np.zeros(n)
    its = 0
    ttol = None
    while True:
        r = PC(b - A(x)) if its else PC(b)
        res = np.linalg.norm(r)
        if ttol is None:
            ttol = max(rtol * res, atol)
        if res <= ttol or its >= max_it:
            return x, its
        V = [r / res]
        H = np.zeros((restart + 1, restart))
        cs, sn = np.zeros(restart), np.zeros(restart)
        g = np.zeros(restart + 1)
        g[0] = res
        k = 0
        done = False
        while k < restart and its < max_it:
            w = PC(A(V[k]))
            h = np.array([v @ w for v in V])  # classical Gram-Schmidt: all products first
            w = w - sum(hj * v for hj, v in zip(h, V))
            tt = np.linalg.norm(w)
            H[:k + 1, k] = h
            H[k + 1, k] = tt
            for j in range(k):  # previous rotations
                t0 = H[j, k]
                H[j, k] = cs[j] * t0 + sn[j] * H[j + 1, k]
                H[j + 1, k] = cs[j] * H[j + 1, k] - sn[j] * t0
            d = np.hypot(H[k, k], H[k + 1, k])
            cs[k], sn[k] = H[k, k] / d, H[k + 1, k] / d
            g[k + 1] = -sn[k] * g[k]
            g[k] = cs[k] * g[k]
            H[k, k] = cs[k] * H[k, k] + sn[k] * H[k + 1, k]
            H[k + 1, k] = 0.0
            res = abs(g[k + 1])
            k += 1
            its += 1
            if res <= ttol or tt == 0.0:
                done = True
                break
            V.append(w / tt)
        y = np.linalg.solve(np.triu(H[:k, :k]), g[:k])
        x = x + sum(yj * v for yj, v in zip(y, V[:k]))
        if done or its >= max_it:
            return x, its


class Newton:
    """SNESNEWTONLS + SNESLINESEARCHBT with a persistent lagged Jacobian."""

    def __init__(self, rtol=1e-8, atol=1e-50, stol=1e-8, max_it=50, ksp_rtol=1e-5, ksp_restart=30, lag=1, alpha=1e-4):
        self.rtol, self.atol, self.stol, self.max_it = rtol, atol, stol, max_it
        self.ksp_rtol, self.ksp_restart, self.lag, self.alpha = ksp_rtol, ksp_restart, lag, alpha
        self.jac_iter = 0
        self.A = self.PC = None
        self.newton_its = self.linear_its = self.assemblies = self.backtracks = 0

    def _linesearch(self, func, x, F, fnorm, y):
        """SNESLineSearchApply_BT (cubic): returns (x_new, F_new, fnorm_new, ynorm, xnorm)."""
        ynorm = np.linalg.norm(y)
        f = fnorm * fnorm
        initslope = F @ self.A(y)
        if initslope > 0.0:
            initslope = -initslope
        if initslope == 0.0:
            initslope = -1.0
        lam = 1.0
        W = x - lam * y
        G = func(W)
        gnorm = np.linalg.norm(G)
        g = gnorm * gnorm
        if 0.5 * g > 0.5 * f + lam * self.alpha * initslope:
            self.backtracks += 1
            lamtemp = -initslope / (g - f - 2.0 * lam * initslope)  # quadratic fit
            lamprev, gnormprev = lam, gnorm
            lam = 0.1 * lam if lamtemp <= 0.1 * lam else min(lamtemp, 0.5 * lam)
            W = x - lam * y
            G = func(W)
            gnorm = np.linalg.norm(G)
            g = gnorm * gnorm
            count = 0
            while 0.5 * g >= 0.5 * f + lam * self.alpha * initslope:  # cubic fits
                count += 1
                if lam <= 1e-12 or count > 40:
                    raise RuntimeError("line search failed")
                t1 = 0.5 * (gnorm * gnorm - f) - lam * initslope
                t2 = 0.5 * (gnormprev * gnormprev - f) - lamprev * initslope
                a = (t1 / (lam * lam) - t2 / (lamprev * lamprev)) / (lam - lamprev)
                b = (-lamprev * t1 / (lam * lam) + lam * t2 / (lamprev * lamprev)) / (lam - lamprev)
                d = max(b * b - 3.0 * a * initslope, 0.0)
                lamtemp = -initslope / (2.0 * b) if a == 0.0 else (-b + np.sqrt(d)) / (3.0 * a)
                lamprev, gnormprev = lam, gnorm
                lam = 0.1 * lam if lamtemp <= 0.1 * lam else min(lamtemp, 0.5 * lam)
                W = x - lam * y
                G = func(W)
                gnorm = np.linalg.norm(G)
                g = gnorm * gnorm
        return W, G, gnorm, lam * ynorm, np.linalg.norm(W)

    def solve(self, func, setup, x):
        """Solves func(x) = 0 from the initial guess x; setup() returns the (operator, preconditioner) of a Jacobian
        evaluation at the last func evaluation."""
        F = func(x)
        fnorm = np.linalg.norm(F)
        ttol = self.rtol * fnorm
        it = 0
        if not np.isfinite(fnorm):
            raise RuntimeError("non-finite residual")
        if fnorm >= self.atol:
            for i in range(self.max_it):
                if self.A is None or self.lag <= 1 or (i + self.jac_iter) % self.lag == 0:
                    self.A, self.PC = setup()
                    self.assemblies += 1
                y, lits = gmres(self.A, self.PC, F, self.ksp_rtol, restart=self.ksp_restart)
                self.linear_its += lits
                x, F, fnorm, ynorm, xnorm = self._linesearch(func, x, F, fnorm, y)
                it = i + 1
                if not np.isfinite(fnorm):
                    raise RuntimeError("non-finite residual")
                if fnorm < self.atol or fnorm <= ttol or ynorm < self.stol * xnorm:
                    break
            else:
                raise RuntimeError("Newton did not converge")
        self.newton_its += it
        self.jac_iter += it
        return x


def ts_alpha(ops: Operators, x0_loc, dt, nsteps, radius, newton: Newton, t0=0.0):
    """TSALPHA from the local state x0_loc (constrained entries set): returns the final global vector."""
    zero = np.zeros(ops.fixed.size)

    def stage(X0, V0, h, am, af, gam):
        shift = am / (af * gam * h)
        ta = t0_[0] + af * h

        def func(X1):
            V1 = (X1 - X0) / (gam * h) + (1.0 - 1.0 / gam) * V0
            Xa = X0 + af * (X1 - X0)
            Va = V0 + am * (V1 - V0)
            return ops.residual(ta, h, ops.to_local(Xa, ops.fixed), ops.to_local(Va, zero))[ops.free] / af

        X1 = newton.solve(func, lambda: ops.setup(shift), X0.copy())
        return X1, (X1 - X0) / (gam * h) + (1.0 - 1.0 / gam) * V0

    t0_ = [t0]
    am, af, gam = alpha_params(radius)
    X0 = np.asarray(x0_loc, dtype=float)[ops.free].copy()
    # TSAlpha_Restart: two backward-Euler half steps give the initial time derivative
    h = 0.5 * dt
    nz = np.zeros_like(X0)
    X1, _ = stage(X0, nz, h, 1.0, 1.0, 1.0)
    t0_[0] = t0 + h
    X2, _ = stage(X1, nz, h, 1.0, 1.0, 1.0)
    t0_[0] = t0
    V0 = (-3.0 * X0 + 4.0 * X1 - X2) / dt
    for n in range(nsteps):
        t0_[0] = t0 + n * dt
        X0, V0 = stage(X0, V0, dt, am, af, gam)
    return X0


def ts_beuler(ops: Operators, x0_loc, dt, nsteps, newton: Newton, t0=0.0):
    """TSBEULER from the local state x0_loc (constrained entries set): returns the final global vector."""
    zero = np.zeros(ops.fixed.size)
    X0 = np.asarray(x0_loc, dtype=float)[ops.free].copy()
    for n in range(nsteps):
        t1 = t0 + (n + 1) * dt

        def func(X1, X0=X0, t1=t1):
            return ops.residual(t1, dt, ops.to_local(X1, ops.fixed), ops.to_local((X1 - X0) / dt, zero))[ops.free]

        X0 = newton.solve(func, lambda: ops.setup(1.0 / dt), X0.copy())
    return X0
