// CPU ORACLE -- test infrastructure only (see ns_pointwise.h).
//
// Explicit time stepping used to reproduce the reference's Taylor-Green kinetic-energy fixture
// (examples/navierstokes.c:222, examples/taylor_green_vortex.yaml):
//   * PETSc TSRK type "4" (classic 4-stage RK) over RHS_NS (src/setupts.c:31-75),
//   * each stage solves M(q) G = R with the SUPG-stabilized consistent mass (problems/newtonian.c:212-265),
//   * the mass KSP is GMRES(restart) with left point-Jacobi preconditioning, zero initial guess, classical Gram-Schmidt
//     and PETSc's default convergence test ||B r|| <= rtol ||B b|| (the "mass_" KSP set up in src/setuplibceed.c:56-95).
//     As in PETSc, the Jacobi diagonal is computed when the preconditioner is first set up and then reused, because the
//     shell matrix's object state never changes between solves.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "honee_oracle.h"

typedef struct {
  const OMesh  *m;
  const ONewt  *c;
  const double *qdata;
  const double *q;  // state the stabilized mass depends on
} MassCtx;

static void mass_apply(const MassCtx *mc, const double *x, double *y) { oracle_mass_newtonian(mc->m, mc->c, mc->qdata, mc->q, x, y); }

static double vdot(int n, const double *a, const double *b) {
  double s = 0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

// Left-preconditioned GMRES(restart) with zero initial guess, mirroring KSPGMRESCycle/KSPGMRESUpdateHessenberg/
// KSPGMRESBuildSoln (PETSc src/ksp/ksp/impls/gmres/gmres.c) with classical Gram-Schmidt and no refinement.
static int gmres(int n, const MassCtx *mc, const double *dinv, const double *b, double *x, double rtol, int restart, int maxit) {
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
    mass_apply(mc, x, Ax);
    for (int i = 0; i < n; i++) V[i] = dinv[i] * (b[i] - Ax[i]);
    double res = sqrt(vdot(n, V, V));
    if (its == 0) ttol = fmax(rtol * res, 1e-50);
    if (res <= ttol) break;
    for (int i = 0; i < n; i++) V[i] /= res;
    grs[0] = res;
    int it = 0;
    while (it < K && its < maxit) {
      double *v_new = &V[(size_t)(it + 1) * n];
      mass_apply(mc, &V[(size_t)it * n], w);
      for (int i = 0; i < n; i++) v_new[i] = dinv[i] * w[i];
      // classical Gram-Schmidt: all dot products first, then subtract
      for (int j = 0; j <= it; j++) HH(j, it) = vdot(n, v_new, &V[(size_t)j * n]);
      for (int j = 0; j <= it; j++)
        for (int i = 0; i < n; i++) v_new[i] -= HH(j, it) * V[(size_t)j * n + i];
      HH(it + 1, it) = sqrt(vdot(n, v_new, v_new));
      if (HH(it + 1, it) != 0)
        for (int i = 0; i < n; i++) v_new[i] /= HH(it + 1, it);
      // apply previous rotations to column it, then form the new one
      for (int j = 0; j < it; j++) {
        double tt    = HH(j, it);
        HH(j, it)     = cc[j] * tt + ss[j] * HH(j + 1, it);
        HH(j + 1, it) = cc[j] * HH(j + 1, it) - ss[j] * tt;
      }
      double tt = sqrt(HH(it, it) * HH(it, it) + HH(it + 1, it) * HH(it + 1, it));
      cc[it]    = HH(it, it) / tt;
      ss[it]    = HH(it + 1, it) / tt;
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

static void record_row(const OMesh *m, const ONewt *c, const double *qdata, const double *q, int step, double t, double vol, double *row) {
  double sums[5];
  oracle_monitor_ke(m, c, qdata, q, sums);
  row[0] = step;
  row[1] = t;
  for (int k = 0; k < 5; k++) row[2 + k] = sums[k] / vol;
}

int oracle_rk4_stabilized(const OMesh *m, ONewt *c, const double *qdata, double *q, double dt, int nsteps, double ksp_rtol, int ksp_restart,
                          int record_every, double *ke_rows, int *nrows) {
  const int    n   = 5 * m->nnodes;
  const double vol = oracle_volume(m, qdata);
  // TSRK4 tableau
  const double A[4][4] = {{0, 0, 0, 0}, {0.5, 0, 0, 0}, {0, 0.5, 0, 0}, {0, 0, 1, 0}};
  const double bw[4]   = {1. / 6, 1. / 3, 1. / 3, 1. / 6};
  const double cst[4]  = {0, 0.5, 0.5, 1};
  double      *K[4], *Y = malloc(sizeof(double) * n), *R = malloc(sizeof(double) * n), *dinv = NULL;
  for (int i = 0; i < 4; i++) K[i] = malloc(sizeof(double) * n);
  double t     = 0;
  int    rows  = 0;
  int    total = 0;
  c->dt        = dt;
  record_row(m, c, qdata, q, 0, t, vol, &ke_rows[7 * rows++]);
  for (int step = 1; step <= nsteps; step++) {
    for (int s = 0; s < 4; s++) {
      memcpy(Y, q, sizeof(double) * n);
      for (int j = 0; j < s; j++) {
        const double wj = dt * A[s][j];
        for (int i = 0; i < n; i++) Y[i] += wj * K[j][i];
      }
      c->time = t + dt * cst[s];
      oracle_rhs(m, c, qdata, Y, R);
      MassCtx mc = {m, c, qdata, Y};
      if (!dinv) {  // PCSetUp at the first solve
        dinv = malloc(sizeof(double) * n);
        oracle_mass_newtonian_diagonal(m, c, qdata, Y, dinv);
        for (int i = 0; i < n; i++) dinv[i] = dinv[i] != 0 ? 1.0 / dinv[i] : 1.0;
      }
      total += gmres(n, &mc, dinv, R, K[s], ksp_rtol, ksp_restart, 10000);
    }
    for (int s = 0; s < 4; s++) {
      const double ws = dt * bw[s];
      for (int i = 0; i < n; i++) q[i] += ws * K[s][i];
    }
    t += dt;
    if (step % record_every == 0 || step == nsteps) record_row(m, c, qdata, q, step, t, vol, &ke_rows[7 * rows++]);
  }
  *nrows = rows;
  for (int i = 0; i < 4; i++) free(K[i]);
  free(Y);
  free(R);
  free(dinv);
  return total;
}
