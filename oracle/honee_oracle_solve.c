// CPU ORACLE -- test infrastructure only (see ns_pointwise.h).
//
// Time integration as HONEE configures PETSc's TS/SNES/KSP (TSSolve_NS, src/setupts.c:334-360; the "mass_" KSP,
// src/setuplibceed.c:56-95).  PETSc is a third-party dependency absent from the reference tree (v3.24.5 per
// tests/output/taylorgreen_totalke.csv:5); its published algorithms are restated here:
//   * TSRK with a Butcher tableau: Y_s = q + h sum_{j<s} a_sj K_j, K_s = M(Y_s)^{-1} R(t + c_s h, Y_s) (RHS_NS,
//     src/setupts.c:31-75), q += h sum_s b_s K_s.  The mass solve is PCJACOBI (row sum = M 1, or the diagonal) alone
//     (KSPPREONLY, HONEE's lumped default) or inside GMRES.  As in PETSc, the Jacobi vector is computed when the
//     preconditioner is first set up and then reused, because the shell matrix's object state never changes between
//     solves.  The SUPG-stabilized mass (problems/newtonian.c:212-265) depends on the stage state Y_s.
//   * KSPGMRES: left preconditioning, zero initial guess, classical Gram-Schmidt without refinement
//     (KSPGMRESCycle/KSPGMRESUpdateHessenberg/KSPGMRESBuildSoln, src/ksp/ksp/impls/gmres/gmres.c), PETSc's default
//     convergence test ||B r|| <= max(rtol ||B b||, atol).
//   * TSBDF(k): q_dot = sum_i alpha_i X_i with alpha_i the derivative at t_{n+1} of the Lagrange basis on the nodes
//     (t_{n+1}, history); the Jacobian shift is alpha_0.  An order-2 start takes a BDF1 half step, then a
//     variable-step BDF2 step to t_0 + h (TSBDF_Restart).  The Newton initial guess extrapolates the history
//     (Lagrange through up to k + 1 points).  The "timestep size" context field stays the nominal h.
//   * PCJACOBI / PCVPBJACOBI: the point diagonal (zero entries replaced by 1) or the 5x5 node blocks inverted by
//     Gauss-Jordan elimination with partial pivoting (a singular block is an error).
//   * SNESNEWTONLS with SNESLINESEARCHBT (cubic order, alpha 1e-4, minlambda 1e-12, 40 steps), restated from
//     SNESLineSearchApply_BT: initial slope F . (J dx) (made negative, -1 if zero); the full step is taken when
//     ||G||^2/2 <= ||F||^2/2 + lambda alpha slope, else a quadratic fit and then cubic fits through the last two
//     trials, each clamped to [0.1, 0.5] of the previous lambda and accepted under the strict inequality (a trial
//     with a non-finite residual cuts lambda by 10 without a fit), SNESLINESEARCHBASIC (the full step), or quadratic
//     fits only with the slope -||F||^2 of an exact Newton direction (not a PETSc type); the IJacobian
//     (src/setupts.c:121-156) uses the Jacobian data of the last IFunction evaluation and the point-Jacobi diagonal is
//     reassembled at every Newton iteration (MatGetDiagonal_Ceed, src/mat-ceed.c:1603-1630).  Convergence:
//     ||F|| <= atol, ||F|| <= max(atol, rtol ||F_0||) or ||step|| <= stol ||x||.
// Parity unpinned against PETSc itself (no PETSc here): these restatements are pinned by the reference's Taylor-Green
// kinetic-energy fixture (RK4 + GMRES mass), by the order conditions and convergence rates of the schemes, and by
// closed-form linear test problems in the tests; the line search's step (oracle_bt_next_lambda) by the exact minimizers
// of quadratic and cubic merit functions (tests/test_oracle_linesearch.py).
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "honee_oracle.h"

typedef void (*ApplyFn)(void *ctx, const double *x, double *y);

// Preconditioner z = B (x - y) (y may be NULL): B is the inverted point diagonal (nb == 1) or the inverted 5x5 node
// blocks [c_out][c_in] (nb == 5)
typedef struct {
  const double *d;
  int           nb;
} Prec;

static void prec_apply(const Prec *pc, int n, const double *x, const double *y, double *z) {
  if (pc->nb == 1) {
    for (int i = 0; i < n; i++) z[i] = pc->d[i] * (y ? x[i] - y[i] : x[i]);
    return;
  }
  for (int node = 0; node < n / 5; node++) {
    double r[5];
    for (int j = 0; j < 5; j++) r[j] = y ? x[node * 5 + j] - y[node * 5 + j] : x[node * 5 + j];
    for (int i = 0; i < 5; i++) {
      double acc = 0;
      for (int j = 0; j < 5; j++) acc += pc->d[node * 25 + i * 5 + j] * r[j];
      z[node * 5 + i] = acc;
    }
  }
}

// In-place inverse of each 5x5 block by Gauss-Jordan elimination with partial pivoting; returns the number of
// singular blocks
static int pb_invert(int nnodes, double *pb) {
  int bad = 0;
  for (int node = 0; node < nnodes; node++) {
    double A[5][5], X[5][5];
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) A[i][j] = pb[node * 25 + i * 5 + j], X[i][j] = i == j;
    for (int k = 0; k < 5; k++) {
      int piv = k;
      for (int i = k + 1; i < 5; i++)
        if (fabs(A[i][k]) > fabs(A[piv][k])) piv = i;
      for (int j = 0; j < 5; j++) {
        double ta = A[k][j], tx = X[k][j];
        A[k][j] = A[piv][j], X[k][j] = X[piv][j];
        A[piv][j] = ta, X[piv][j] = tx;
      }
      if (A[k][k] == 0) {
        bad++;
        continue;
      }
      const double inv = 1.0 / A[k][k];
      for (int j = 0; j < 5; j++) A[k][j] *= inv, X[k][j] *= inv;
      for (int i = 0; i < 5; i++) {
        if (i == k) continue;
        const double f = A[i][k];
        for (int j = 0; j < 5; j++) A[i][j] -= f * A[k][j], X[i][j] -= f * X[k][j];
      }
    }
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) pb[node * 25 + i * 5 + j] = X[i][j];
  }
  return bad;
}

static double vdot(int n, const double *a, const double *b) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

// Left-preconditioned GMRES(restart) with zero initial guess; returns the iteration count
static int gmres(int n, ApplyFn apply, void *actx, const Prec *pc, const double *b, double *x, double rtol, double atol, int restart,
                 int maxit) {
  const int K   = restart;
  double   *V   = malloc(sizeof(double) * (size_t)n * (K + 1));
  double   *w   = malloc(sizeof(double) * n);
  double   *Ax  = malloc(sizeof(double) * n);
  double   *H   = calloc((size_t)(K + 1) * (K + 1), sizeof(double));  // H[row*(K+1)+col]
  double   *cc  = calloc(K + 1, sizeof(double)), *ss = calloc(K + 1, sizeof(double)), *grs = calloc(K + 2, sizeof(double));
  double   *nrs = calloc(K + 1, sizeof(double));
  double    ttol = 0;
  int       its  = 0, converged = 0;
#define HH(r, c_) H[(r) * (K + 1) + (c_)]
  memset(x, 0, sizeof(double) * n);
  while (!converged && its < maxit) {
    // r = B (b - A x)
    apply(actx, x, Ax);
    prec_apply(pc, n, b, Ax, V);
    double res = sqrt(vdot(n, V, V));
    if (its == 0) ttol = fmax(rtol * res, atol);
    if (res <= ttol) break;
    for (int i = 0; i < n; i++) V[i] /= res;
    grs[0] = res;
    int it = 0;
    while (it < K && its < maxit) {
      double *v_new = &V[(size_t)(it + 1) * n];
      apply(actx, &V[(size_t)it * n], w);
      prec_apply(pc, n, w, NULL, v_new);
      // classical Gram-Schmidt: all dot products first, then subtract
      for (int j = 0; j <= it; j++) HH(j, it) = vdot(n, v_new, &V[(size_t)j * n]);
      for (int j = 0; j <= it; j++)
        for (int i = 0; i < n; i++) v_new[i] -= HH(j, it) * V[(size_t)j * n + i];
      HH(it + 1, it) = sqrt(vdot(n, v_new, v_new));
      if (HH(it + 1, it) != 0)
        for (int i = 0; i < n; i++) v_new[i] /= HH(it + 1, it);
      // apply previous rotations to column it, then form the new one
      for (int j = 0; j < it; j++) {
        double tt     = HH(j, it);
        HH(j, it)     = cc[j] * tt + ss[j] * HH(j + 1, it);
        HH(j + 1, it) = cc[j] * HH(j + 1, it) - ss[j] * tt;
      }
      double tt   = sqrt(HH(it, it) * HH(it, it) + HH(it + 1, it) * HH(it + 1, it));
      cc[it]      = HH(it, it) / tt;
      ss[it]      = HH(it + 1, it) / tt;
      grs[it + 1] = -(ss[it] * grs[it]);
      grs[it]     = cc[it] * grs[it];
      HH(it, it)  = cc[it] * HH(it, it) + ss[it] * HH(it + 1, it);
      res         = fabs(grs[it + 1]);
      it++;
      its++;
      if (res <= ttol) {
        converged = 1;
        break;
      }
    }
    // back substitution and update x += V y
    const int last = it - 1;
    nrs[last]      = grs[last] / HH(last, last);
    for (int k = last - 1; k >= 0; k--) {
      double tt = grs[k];
      for (int j = k + 1; j <= last; j++) tt -= HH(k, j) * nrs[j];
      nrs[k] = tt / HH(k, k);
    }
    for (int j = 0; j <= last; j++)
      for (int i = 0; i < n; i++) x[i] += nrs[j] * V[(size_t)j * n + i];
  }
#undef HH
  free(V);
  free(w);
  free(Ax);
  free(H);
  free(cc);
  free(ss);
  free(grs);
  free(nrs);
  return its;
}

static void jacobi_invert(int n, double *d) {
  for (int i = 0; i < n; i++) d[i] = d[i] != 0 ? 1.0 / d[i] : 1.0;
}

// ---- explicit ----
typedef struct {
  const OMesh  *m;
  const ONewt  *c;
  const double *qdata;
  const double *q;  // state the stabilized mass depends on
} MassCtx;

static void mass_apply(void *ctx, const double *x, double *y) {
  const MassCtx *mc = ctx;
  oracle_mass_newtonian(mc->m, mc->c, mc->qdata, mc->q, x, y);
}

static void record_row(const OMesh *m, const ONewt *c, const double *qdata, const double *q, int step, double t, double vol, double *row) {
  double sums[5];
  oracle_monitor_ke(m, c, qdata, q, sums);
  row[0] = step;
  row[1] = t;
  for (int k = 0; k < 5; k++) row[2 + k] = sums[k] / vol;
}

// The mass action with essential boundary conditions: y = M(q) (x .* mask), then the identity on the constrained
// entries (the mirror's restatement of HONEE's DMPlex-constrained mass matrix)
typedef struct {
  MassCtx       mc;
  const double *mask;
  double       *xm;
  int           n;
} MassEssCtx;

static void mass_apply_ess(void *ctx, const double *x, double *y) {
  MassEssCtx *me = ctx;
  for (int l = 0; l < me->n; l++) me->xm[l] = x[l] * me->mask[l];
  mass_apply(&me->mc, me->xm, y);
  for (int l = 0; l < me->n; l++) y[l] = y[l] * me->mask[l] + (1.0 - me->mask[l]) * x[l];
}

int oracle_rk(const OMesh *m, ONewt *c, const double *qdata, double *q, double t0, double dt, int nsteps, int s, const double *A, const double *b,
              const double *cs, const ORKMass *mo, int record_every, double *ke_rows, int *nrows, long long *mass_its) {
  return oracle_rk_ex(m, c, qdata, q, t0, dt, nsteps, s, A, b, cs, mo, NULL, record_every, ke_rows, nrows, mass_its);
}

int oracle_rk_ex(const OMesh *m, ONewt *c, const double *qdata, double *q, double t0, double dt, int nsteps, int s, const double *A, const double *b,
                 const double *cs, const ORKMass *mo, const ORKExtra *ex, int record_every, double *ke_rows, int *nrows, long long *mass_its) {
  const int     n    = 5 * m->nnodes;
  const double  vol  = oracle_volume(m, qdata);
  const double *mask = ex ? ex->mask : NULL;
  double      **K = malloc(sizeof(double *) * s), *Y = malloc(sizeof(double) * n), *R = malloc(sizeof(double) * n), *dinv = NULL;
  double       *Yin = malloc(sizeof(double) * n), *xm = malloc(sizeof(double) * n);
  for (int i = 0; i < s; i++) K[i] = malloc(sizeof(double) * n);
  double    t     = t0;
  int       rows  = 0;
  long long total = 0;
  c->dt           = dt;
  if (ke_rows) record_row(m, c, qdata, q, 0, t, vol, &ke_rows[7 * rows++]);
  for (int step = 1; step <= nsteps; step++) {
    for (int i = 0; i < s; i++) {
      memcpy(Y, q, sizeof(double) * n);
      for (int j = 0; j < i; j++) {
        const double wj = dt * A[i * s + j];
        if (wj != 0)
          for (int l = 0; l < n; l++) Y[l] += wj * K[j][l];
      }
      c->time = t + dt * cs[i];
      // the operators' input: boundary values inserted at the constrained entries (UpdateBoundaryValues)
      if (mask)
        for (int l = 0; l < n; l++) Yin[l] = Y[l] * mask[l] + ex->qbc[l];
      else memcpy(Yin, Y, sizeof(double) * n);
      oracle_rhs(m, c, qdata, Yin, R);
      if (ex && ex->faces && ex->faces->nfaces > 0) oracle_freestream(m, ex->faces, ex->freestream, ex->qdata_sur, Yin, R, NULL);
      if (mask)
        for (int l = 0; l < n; l++) R[l] *= mask[l];
      MassCtx    mc = {m, c, qdata, Yin};
      MassEssCtx me = {mc, mask, xm, n};
      if (!dinv) {  // PCSetUp at the first solve
        dinv = malloc(sizeof(double) * n);
        if (mo->jacobi_rowsum) {
          double *ones = malloc(sizeof(double) * n);
          for (int l = 0; l < n; l++) ones[l] = 1.0;
          if (mask) mass_apply_ess(&me, ones, dinv);
          else mass_apply(&mc, ones, dinv);
          free(ones);
        } else {
          oracle_mass_newtonian_diagonal(m, c, qdata, Yin, dinv);
          if (mask)
            for (int l = 0; l < n; l++) dinv[l] = dinv[l] * mask[l] + (1.0 - mask[l]);
        }
        jacobi_invert(n, dinv);
      }
      const Prec pc = {dinv, 1};
      if (mo->gmres) {
        if (mask) total += gmres(n, mass_apply_ess, &me, &pc, R, K[i], mo->rtol, 1e-50, mo->restart, mo->max_it);
        else total += gmres(n, mass_apply, &mc, &pc, R, K[i], mo->rtol, 1e-50, mo->restart, mo->max_it);
      } else {
        for (int l = 0; l < n; l++) K[i][l] = dinv[l] * R[l];
      }
    }
    for (int i = 0; i < s; i++) {
      const double ws = dt * b[i];
      if (ws != 0)
        for (int l = 0; l < n; l++) q[l] += ws * K[i][l];
    }
    t += dt;
    if (ke_rows && (step % record_every == 0 || step == nsteps)) record_row(m, c, qdata, q, step, t, vol, &ke_rows[7 * rows++]);
  }
  if (nrows) *nrows = rows;
  if (mass_its) *mass_its = total;
  for (int i = 0; i < s; i++) free(K[i]);
  free(K);
  free(Y);
  free(R);
  free(Yin);
  free(xm);
  free(dinv);
  return 0;
}

int oracle_rk4_stabilized(const OMesh *m, ONewt *c, const double *qdata, double *q, double dt, int nsteps, double ksp_rtol, int ksp_restart,
                          int record_every, double *ke_rows, int *nrows) {
  // TSRK "4" tableau, GMRES mass with the diagonal Jacobi preconditioner (examples/taylor_green_vortex.yaml:41-43)
  const double A[16] = {0, 0, 0, 0, 0.5, 0, 0, 0, 0, 0.5, 0, 0, 0, 0, 1, 0};
  const double b[4]  = {1. / 6, 1. / 3, 1. / 3, 1. / 6};
  const double cs[4] = {0, 0.5, 0.5, 1};
  ORKMass      mo    = {1, 0, ksp_rtol, ksp_restart, 10000};
  long long    its   = 0;
  oracle_rk(m, c, qdata, q, 0.0, dt, nsteps, 4, A, b, cs, &mo, record_every, ke_rows, nrows, &its);
  return (int)its;
}

// ---- implicit ----
typedef struct {
  const OMesh  *m;
  ONewt        *c;
  const double *qdata;
  const double *jac;
} JacCtx;

static void jac_apply(void *ctx, const double *x, double *y) {
  const JacCtx *jc = ctx;
  oracle_ijacobian(jc->m, jc->c, jc->qdata, jc->jac, x, y);
}

static void bdf_coefficients(const double *T, int k, double *alpha) {
  alpha[0] = 0;
  for (int m = 1; m <= k; m++) alpha[0] += 1.0 / (T[0] - T[m]);
  for (int i = 1; i <= k; i++) {
    double v = 1.0 / (T[i] - T[0]);
    for (int j = 1; j <= k; j++)
      if (j != i) v *= (T[0] - T[j]) / (T[i] - T[j]);
    alpha[i] = v;
  }
}

static void lagrange_values(const double *X, int m, double t, double *L) {
  for (int i = 0; i < m; i++) {
    double v = 1;
    for (int j = 0; j < m; j++)
      if (j != i) v *= (t - X[j]) / (X[i] - X[j]);
    L[i] = v;
  }
}

typedef struct {
  const OMesh  *m;
  ONewt        *c;
  const double *qdata;
  const OBDFOptions *o;
  int           n, nhist, cap;
  double      **hist, *hist_t;
  double       *x, *xdot, *hsum, *F, *Ftry, *dx, *xtry, *dinv, *jac, *jac_try;
  OBDFStats    *st;
} BDF;

static void residual(BDF *b, double tnew, double shift, const double *x, double *F, double *jac) {
  for (int i = 0; i < b->n; i++) b->xdot[i] = shift * x[i] + b->hsum[i];
  b->c->time = tnew;
  oracle_ifunction(b->m, b->c, b->qdata, x, b->xdot, F, jac);
  b->st->ifunction_evals++;
}

// SNESLineSearchApply_BT's step after a rejected trial with a finite residual: with f(l) = ||F(x - l dx)||^2 / 2,
// the minimizer of the quadratic through f(0), f'(0) = initslope and f(lambda) when no fit has been made yet (nfit
// == 0), else of the cubic through f(0), f'(0), f(lambda) and f(lamprev); clamped to [0.1, 0.5] lambda
double oracle_bt_next_lambda(double fnorm, double initslope, double lambda, double gnorm, int nfit, double lamprev, double gnormprev) {
  double lt;
  if (nfit == 0) {
    lt = -initslope * lambda * lambda / (gnorm * gnorm - fnorm * fnorm - 2.0 * lambda * initslope);
  } else {
    const double t1 = 0.5 * (gnorm * gnorm - fnorm * fnorm) - lambda * initslope;
    const double t2 = 0.5 * (gnormprev * gnormprev - fnorm * fnorm) - lamprev * initslope;
    const double a  = (t1 / (lambda * lambda) - t2 / (lamprev * lamprev)) / (lambda - lamprev);
    const double b  = (-lamprev * t1 / (lambda * lambda) + lambda * t2 / (lamprev * lamprev)) / (lambda - lamprev);
    double       d  = b * b - 3.0 * a * initslope;
    if (d < 0.0) d = 0.0;
    lt = a == 0.0 ? -initslope / (2.0 * b) : (-b + sqrt(d)) / (3.0 * a);
  }
  if (lt > 0.5 * lambda) lt = 0.5 * lambda;
  return lt <= 0.1 * lambda ? 0.1 * lambda : lt;
}

static int bdf_solve(BDF *b, double tnew, int k) {
  const int n = b->n;
  double    T[4], alpha[4];
  T[0] = tnew;
  for (int i = 1; i <= k; i++) T[i] = b->hist_t[i - 1];
  bdf_coefficients(T, k, alpha);
  const double shift = alpha[0];
  for (int l = 0; l < n; l++) {
    double s = 0;
    for (int i = 1; i <= k; i++) s += alpha[i] * b->hist[i - 1][l];
    b->hsum[l] = s;
  }
  const int m = k + 1 < b->nhist ? k + 1 : b->nhist;
  double    L[4];
  lagrange_values(b->hist_t, m, tnew, L);
  for (int l = 0; l < n; l++) {
    double s = 0;
    for (int i = 0; i < m; i++) s += L[i] * b->hist[i][l];
    b->x[l] = s;
  }
  const OBDFOptions *o = b->o;
  residual(b, tnew, shift, b->x, b->F, b->jac);
  double fnorm = sqrt(vdot(n, b->F, b->F)), fnorm0 = fnorm;
  b->st->last_fnorm0     = fnorm0;
  b->st->last_newton_its = 0;
  b->st->last_krylov_its = 0;
  if (!isfinite(fnorm)) return -3;
  const double ttol      = fmax(o->snes_atol, o->snes_rtol * fnorm0);
  int          converged = fnorm <= o->snes_atol;
  b->c->ijacobian_time_shift = shift;
  for (int it = 1; it <= o->snes_max_it && !converged; it++) {
    if (o->pc_pbjacobi) {
      oracle_ijacobian_point_block_diagonal(b->m, b->c, b->qdata, b->jac, b->dinv);
      if (pb_invert(n / 5, b->dinv)) return -6;
    } else {
      oracle_ijacobian_diagonal(b->m, b->c, b->qdata, b->jac, b->dinv);
      jacobi_invert(n, b->dinv);
    }
    JacCtx     jc   = {b->m, b->c, b->qdata, b->jac};
    const Prec pc   = {b->dinv, o->pc_pbjacobi ? 5 : 1};
    const int  kits = gmres(n, jac_apply, &jc, &pc, b->F, b->dx, o->ksp_rtol, o->ksp_atol, o->ksp_restart, o->ksp_max_it);
    b->st->krylov_its += kits;
    b->st->last_krylov_its += kits;
    const double ynorm = sqrt(vdot(n, b->dx, b->dx));
    const int norm_slope = o->linesearch == 2;
    double    initslope  = -fnorm * fnorm;
    if (o->linesearch == 0) {
      jac_apply(&jc, b->dx, b->Ftry);  // J dx (Ftry is free until the first trial)
      initslope = vdot(n, b->F, b->Ftry);
    }
    if (initslope > 0.0) initslope = -initslope;
    if (initslope == 0.0) initslope = -1.0;
    const double alpha_ls = 1e-4;
    double       lambda = 1.0, gnorm = 0, lamprev = 0, gnormprev = 0;
    int          nfit = 0;
    for (int bt = 0;; bt++) {
      for (int l = 0; l < n; l++) b->xtry[l] = b->x[l] - lambda * b->dx[l];
      residual(b, tnew, shift, b->xtry, b->Ftry, b->jac_try);
      gnorm = sqrt(vdot(n, b->Ftry, b->Ftry));
      if (o->linesearch == 1) {  // SNESLINESEARCHBASIC: the full step, accepted when its residual is finite
        if (!isfinite(gnorm)) return -4;
        break;
      }
      const double lhs = 0.5 * gnorm * gnorm, rhs = 0.5 * fnorm * fnorm + lambda * alpha_ls * initslope;
      if (isfinite(gnorm) && (bt == 0 || norm_slope ? lhs <= rhs : lhs < rhs)) break;
      if (bt >= 40 || (!norm_slope && lambda <= 1e-12)) return -4;
      b->st->line_search_backtracks++;
      if (!isfinite(gnorm)) {
        lambda *= 0.1;
        continue;
      }
      const double next = oracle_bt_next_lambda(fnorm, initslope, lambda, gnorm, norm_slope ? 0 : nfit, lamprev, gnormprev);
      nfit++;
      lamprev   = lambda;
      gnormprev = gnorm;
      lambda    = next;
    }
    double *sw;
    sw = b->x, b->x = b->xtry, b->xtry = sw;
    sw = b->F, b->F = b->Ftry, b->Ftry = sw;
    sw = b->jac, b->jac = b->jac_try, b->jac_try = sw;
    fnorm = gnorm;
    b->st->newton_its++;
    b->st->last_newton_its = it;
    const double xnorm = sqrt(vdot(n, b->x, b->x));
    if (fnorm <= ttol || lambda * ynorm <= o->snes_stol * xnorm) converged = 1;
  }
  b->st->last_fnorm = fnorm;
  return converged ? 0 : -4;
}

static void hist_push(BDF *b, const double *x, double tx) {
  double *last = b->hist[b->cap - 1];
  for (int i = b->cap - 1; i > 0; i--) {
    b->hist[i]   = b->hist[i - 1];
    b->hist_t[i] = b->hist_t[i - 1];
  }
  b->hist[0]   = last;
  b->hist_t[0] = tx;
  memcpy(b->hist[0], x, sizeof(double) * b->n);
  b->nhist = b->nhist + 1 < b->cap ? b->nhist + 1 : b->cap;
}

int oracle_bdf(const OMesh *m, ONewt *c, const double *qdata, double *q, double t0, int nsteps, const OBDFOptions *o, OBDFStats *st) {
  const int n  = 5 * m->nnodes;
  const int nj = 14 + (c->idl_enable ? 1 : 0);
  BDF       b  = {0};
  b.m = m, b.c = c, b.qdata = qdata, b.o = o, b.n = n, b.cap = o->order + 1, b.st = st;
  memset(st, 0, sizeof(*st));
  b.hist   = malloc(sizeof(double *) * b.cap);
  b.hist_t = calloc(b.cap, sizeof(double));
  for (int i = 0; i < b.cap; i++) b.hist[i] = malloc(sizeof(double) * n);
  double **bufs[] = {&b.x, &b.xdot, &b.hsum, &b.F, &b.Ftry, &b.dx, &b.xtry, &b.dinv};
  for (int i = 0; i < 8; i++) *bufs[i] = malloc(sizeof(double) * n * (i == 7 ? 5 : 1));
  b.jac     = malloc(sizeof(double) * (size_t)m->nelem * nj * m->Q * m->Q * m->Q);
  b.jac_try = malloc(sizeof(double) * (size_t)m->nelem * nj * m->Q * m->Q * m->Q);
  c->dt      = o->dt;
  double t   = t0;
  int    k   = 1, ierr = 0;
  for (int step = 0; step < nsteps && !ierr; step++) {
    if (b.nhist == 0) {
      hist_push(&b, q, t);
      if (o->order > 1) {
        if ((ierr = bdf_solve(&b, t + o->dt / 2, 1))) break;
        hist_push(&b, b.x, t + o->dt / 2);
        k = 2;
      } else {
        k = 1;
      }
    } else {
      k = k + 1 < o->order ? k + 1 : o->order;
    }
    if ((ierr = bdf_solve(&b, t + o->dt, k))) break;
    hist_push(&b, b.x, t + o->dt);
    memcpy(q, b.x, sizeof(double) * n);
    t += o->dt;
    st->steps++;
  }
  st->time = t;
  for (int i = 0; i < b.cap; i++) free(b.hist[i]);
  free(b.hist);
  free(b.hist_t);
  for (int i = 0; i < 8; i++) free(*bufs[i]);
  free(b.jac);
  free(b.jac_try);
  return ierr;
}
