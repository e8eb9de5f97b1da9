// PETSc-free GPU time integrator over the HoneeMiProblem operators (declarations and scope in honee_ts.h).
//
// Restated algorithms (PETSc is a third-party dependency of the reference, not part of it; the steps follow its
// published methods, which HONEE selects in TSSolve_NS, src/setupts.c:334-360, and the "mass_" KSP setup,
// src/setuplibceed.c:56-95):
//   * TSRK: Y_s = q + h sum_{j<s} a_sj K_j, K_s = M(Y_s)^{-1} R(t + c_s h, Y_s) (RHS_NS, src/setupts.c:31-75), then
//     q += h sum_s b_s K_s.  The mass preconditioner (PCJACOBI row sum or diagonal of M at the first stage state) is set
//     up at the first solve and reused, since the shell mass matrix never changes state.
//   * TSBDF(k): the BDF derivative at t_{n+1} is the derivative of the Lagrange interpolant of (t_{n+1}, x) and the k
//     previous solutions, q_dot = alpha_0 x + sum_i alpha_i q_{n+1-i}, with Jacobian shift alpha_0.  The start of an
//     order-2 run takes a BDF1 half step and then a variable-step BDF2 step to t_0 + h (TSBDF_Restart); the Newton
//     initial guess extrapolates the history (Lagrange, up to k+1 points).
//   * SNESNEWTONLS: F(x) = G(t, x, q_dot(x)) (IFunction_NS), J = dG/dx + alpha_0 dG/dq_dot (FormIJacobian_NS's shift),
//     Jacobi diagonal reassembled every iteration (MatGetDiagonal_Ceed), J dx = F by GMRES, backtracking on
//     ||F||^2 with the sufficient-decrease constant 1e-4, convergence by ||F|| <= max(atol, rtol ||F_0||) or
//     ||step|| <= stol ||x||.
//   * KSPGMRES: left point-Jacobi, zero initial guess, classical Gram-Schmidt without refinement, Givens rotations,
//     convergence ||B r|| <= max(rtol ||B b||, atol).
#include "honee_ts.h"

#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "krylov.hpp"

namespace {

constexpr int kMemHost = 0, kMemDevice = 1;

struct Tableau {
  int                 s = 0;
  std::vector<double> A, b, c;  // A[i * s + j]
};

// Optimal third-order SSP scheme with s = n^2 stages (Ketcheson 2008, low-storage form), converted to a Butcher
// tableau by following the scheme on coefficient vectors over (u, h K_1, ..., h K_s).
Tableau ssp_s3(int s) {
  Tableau   T;
  const int n = (int)std::lround(std::sqrt((double)s));
  T.s         = s;
  T.A.assign((size_t)s * s, 0.0);
  T.b.assign(s, 0.0);
  T.c.assign(s, 0.0);
  const double        r = (double)(s - n);
  std::vector<double> q1(s, 0.0), q2;
  int                 stage = 0;
  auto                step  = [&]() {  // q1 <- q1 + (h / r) L(q1)
    for (int j = 0; j < s; j++) T.A[(size_t)stage * s + j] = q1[j];
    q1[stage] += 1.0 / r;
    stage++;
  };
  for (int i = 1; i <= (n - 1) * (n - 2) / 2; i++) step();
  q2 = q1;
  for (int i = (n - 1) * (n - 2) / 2 + 1; i <= n * (n + 1) / 2 - 1; i++) step();
  step();
  for (int j = 0; j < s; j++) q1[j] = (n * q2[j] + (n - 1) * q1[j]) / (2 * n - 1);
  for (int i = n * (n + 1) / 2 + 1; i <= n * n; i++) step();
  T.b = q1;
  for (int i = 0; i < s; i++)
    for (int j = 0; j < s; j++) T.c[i] += T.A[(size_t)i * s + j];
  return T;
}

Tableau make_tableau(int type, int ssp_stages) {
  Tableau T;
  auto    set = [&](int s, std::initializer_list<double> A, std::initializer_list<double> b) {
    T.s = s;
    T.A.assign(A);
    T.b.assign(b);
    T.c.assign(s, 0.0);
    for (int i = 0; i < s; i++)
      for (int j = 0; j < s; j++) T.c[i] += T.A[(size_t)i * s + j];
  };
  switch (type) {
    case HONEE_MI_RK_4:
      set(4, {0, 0, 0, 0, .5, 0, 0, 0, 0, .5, 0, 0, 0, 0, 1, 0}, {1. / 6, 1. / 3, 1. / 3, 1. / 6});
      break;
    case HONEE_MI_RK_5F:
      set(6,
          {0, 0, 0, 0, 0, 0,                                                   //
           1. / 4, 0, 0, 0, 0, 0,                                              //
           3. / 32, 9. / 32, 0, 0, 0, 0,                                       //
           1932. / 2197, -7200. / 2197, 7296. / 2197, 0, 0, 0,                 //
           439. / 216, -8, 3680. / 513, -845. / 4104, 0, 0,                    //
           -8. / 27, 2, -3544. / 2565, 1859. / 4104, -11. / 40, 0},            //
          {16. / 135, 0, 6656. / 12825, 28561. / 56430, -9. / 50, 2. / 55});
      break;
    case HONEE_MI_RK_SSP_S3:
      T = ssp_s3(ssp_stages);
      break;
    case HONEE_MI_RK_SSP33:
      set(3, {0, 0, 0, 1, 0, 0, .25, .25, 0}, {1. / 6, 1. / 6, 2. / 3});
      break;
    case HONEE_MI_RK_EULER:
      set(1, {0}, {1});
      break;
    default:
      break;
  }
  return T;
}

}  // namespace

struct HoneeMiTS_ {
  HoneeMiProblem   p = nullptr;
  HoneeMiTSOptions o{};
  HoneeMiTSStats   st{};
  hipStream_t      s = nullptr;
  long long        n = 0;
  double           t = 0;
  double          *q = nullptr;
  Tableau          tab;
  // explicit
  std::vector<double *> K;
  double               *Y = nullptr, *R = nullptr, *mass_dinv = nullptr;
  bool                  mass_pc_ready = false;
  // implicit
  std::vector<double *> hist;  // most recent first
  std::vector<double>   hist_t;
  int                   nhist = 0, korder = 1;
  double               *x = nullptr, *xdot = nullptr, *hsum = nullptr, *F = nullptr, *Ftry = nullptr, *dx = nullptr, *xtry = nullptr,
         *dinv = nullptr;
  // Newton stage residual R(x) = scale F(t, a x + xc, b x + hsum) with the IFunction's "timestep size" ctx_dt
  struct Stage {
    double a = 1, b = 0, scale = 1, ctx_dt = 0;
    bool   xc = false;
  } stage;
  double   *v0 = nullptr, *xa = nullptr, *xc = nullptr;  // generalized-alpha: time derivative, stage state, its offset
  bool      alpha_started = false;
  long long jac_iter      = 0;  // Newton iterations of earlier solves counted by a persistent Jacobian lag
  double    jac_shift     = 0;  // shift of the last Jacobian evaluation
  bool      jac_ready     = false;
  // Krylov
  int     restart = 0;
  double *V = nullptr, *w = nullptr, *partial = nullptr, *d_red = nullptr;
  // Reduction results read by the host: mapped, coherent host memory (kMaxDots values, then a sequence number) that a
  // kernel stores into; h_* are the host addresses, g_* the device ones
  double             *h_red = nullptr, *g_red = nullptr;
  unsigned long long *h_seq = nullptr, *g_seq = nullptr, seq = 0;
  std::vector<void *> allocs;
  std::string         err;
};

namespace {

int fail(HoneeMiTS ts, int code, const char *fmt, ...) {
  char    buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  ts->err = buf;
  return code ? code : -1;
}

#define TS_HIP(ts, expr)                                                                   \
  do {                                                                                     \
    hipError_t e_ = (expr);                                                                \
    if (e_ != hipSuccess) return fail(ts, -2, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
  } while (0)

#define TS_OP(ts, expr)                                                                                          \
  do {                                                                                                           \
    int i_ = (expr);                                                                                             \
    if (i_) return fail(ts, i_, "%s failed (%d): %s", #expr, i_, HoneeMiProblemError(ts->p));                    \
  } while (0)

#define TS_CALL(expr)     \
  do {                    \
    int i_ = (expr);      \
    if (i_) return i_;    \
  } while (0)

int dev_alloc(HoneeMiTS ts, double **ptr, long long len) {
  TS_HIP(ts, hipMalloc((void **)ptr, sizeof(double) * (size_t)(len > 0 ? len : 1)));
  ts->allocs.push_back(*ptr);
  return 0;
}

// Host reads of reduction results d_red[0..count): the launch's final reduction kernel stores them into h_red (mapped
// host memory) followed by a new sequence number (Publish), which wait_published polls.  After 2 ms without it the host
// waits on the stream instead (which also reports a failed launch), so a long queue ahead of the read does not keep a
// core spinning.
Publish next_publish(HoneeMiTS ts, int count) {
  Publish p;
  p.h     = ts->g_red;
  p.flag  = ts->g_seq;
  p.seq   = ++ts->seq;
  p.count = count;
  return p;
}

int wait_published(HoneeMiTS ts, const Publish &p) {
  const auto t0 = std::chrono::steady_clock::now();
  while (__atomic_load_n(ts->h_seq, __ATOMIC_ACQUIRE) != p.seq) {
    if (std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
      TS_HIP(ts, hipStreamSynchronize(ts->s));
      if (__atomic_load_n(ts->h_seq, __ATOMIC_ACQUIRE) != p.seq) return fail(ts, -2, "reduction result not published");
      break;
    }
  }
  return 0;
}

int norm2(HoneeMiTS ts, const double *x, double *out) {
  const Publish pub = next_publish(ts, 1);
  TS_HIP(ts, launch_dot(x, x, ts->n, 0, ts->partial, ts->d_red, ts->s, pub));
  TS_CALL(wait_published(ts, pub));
  *out = std::sqrt(ts->h_red[0]);
  return 0;
}

int copy(HoneeMiTS ts, double *dst, const double *src) {
  TS_HIP(ts, hipMemcpyAsync(dst, src, sizeof(double) * (size_t)ts->n, hipMemcpyDeviceToDevice, ts->s));
  return 0;
}

// y = a y + sum_j c_j x_j for any number of terms (in chunks of kMaxLinComb)
int lincomb(HoneeMiTS ts, double *y, double a, const std::vector<double> &c, const std::vector<const double *> &xs) {
  size_t j = 0;
  do {
    LinComb lc{};
    lc.m = 0;
    for (; j < c.size() && lc.m < kMaxLinComb; j++) {
      if (c[j] == 0) continue;
      lc.c[lc.m]   = c[j];
      lc.x[lc.m++] = xs[j];
    }
    TS_HIP(ts, launch_lincomb(y, a, lc, ts->n, ts->s));
    a = 1.0;
  } while (j < c.size());
  return 0;
}

// Left-preconditioned GMRES(restart), zero initial guess (see the file header); apply(x, y) computes y = A x.
template <class Apply, class Prec>
int gmres(HoneeMiTS ts, Apply apply, Prec prec, const double *b, double *x, double rtol, double atol, int restart, int maxit, int *its_out) {
  const long long     n = ts->n;
  const int           K = restart;
  std::vector<double> H((size_t)(K + 1) * (K + 1), 0.0), cc(K + 1, 0.0), ss(K + 1, 0.0), grs(K + 2, 0.0), nrs(K + 1, 0.0);
  auto                HH = [&](int r, int c) -> double & { return H[(size_t)r * (K + 1) + c]; };
  double             *V  = ts->V;
  double              ttol = 0;
  int                 its = 0, converged = 0;
  TS_HIP(ts, launch_set(x, 0.0, n, ts->s));
  while (!converged && its < maxit) {
    if (its == 0) {  // r = B b (A x = 0)
      TS_CALL(prec(V, b, nullptr));
    } else {
      TS_CALL(apply(x, ts->w));
      TS_CALL(prec(V, b, ts->w));
    }
    // ||V_0||, with V_0 normalized on the device while the host reads the norm
    const Publish pub = next_publish(ts, 1);
    TS_HIP(ts, launch_dot(V, V, n, 0, ts->partial, ts->d_red, ts->s, pub));
    TS_HIP(ts, launch_normalize(V, ts->d_red, n, ts->s));
    TS_CALL(wait_published(ts, pub));
    double res = std::sqrt(ts->h_red[0]);
    if (its == 0) ttol = std::fmax(rtol * res, atol);
    if (res <= ttol) break;
    grs[0] = res;
    int it = 0;
    while (it < K && its < maxit) {
      double *v_new = V + (size_t)(it + 1) * n;
      TS_CALL(apply(V + (size_t)it * n, ts->w));
      TS_CALL(prec(v_new, ts->w, nullptr));
      // classical Gram-Schmidt: all dots first, then one projection pass that also forms ||v_new||^2; v_new is
      // normalized on the device while the host reads (h_0..h_it, ||v_new||^2)
      TS_HIP(ts, launch_mdot(v_new, V, n, it + 1, n, ts->partial, ts->d_red, ts->s));
      const Publish pub = next_publish(ts, it + 2);
      TS_HIP(ts, launch_project_out(v_new, V, n, it + 1, ts->d_red, n, it + 1, ts->partial, ts->d_red, ts->s, pub));
      TS_HIP(ts, launch_normalize(v_new, ts->d_red + it + 1, n, ts->s));
      TS_CALL(wait_published(ts, pub));
      for (int j = 0; j <= it; j++) HH(j, it) = ts->h_red[j];
      HH(it + 1, it) = std::sqrt(ts->h_red[it + 1]);
      for (int j = 0; j < it; j++) {
        const double tt = HH(j, it);
        HH(j, it)       = cc[j] * tt + ss[j] * HH(j + 1, it);
        HH(j + 1, it)   = cc[j] * HH(j + 1, it) - ss[j] * tt;
      }
      const double tt = std::sqrt(HH(it, it) * HH(it, it) + HH(it + 1, it) * HH(it + 1, it));
      cc[it]          = HH(it, it) / tt;
      ss[it]          = HH(it + 1, it) / tt;
      grs[it + 1]     = -(ss[it] * grs[it]);
      grs[it]         = cc[it] * grs[it];
      HH(it, it)      = cc[it] * HH(it, it) + ss[it] * HH(it + 1, it);
      res             = std::fabs(grs[it + 1]);
      it++;
      its++;
      if (res <= ttol) {
        converged = 1;
        break;
      }
    }
    const int last = it - 1;
    nrs[last]      = grs[last] / HH(last, last);
    for (int k = last - 1; k >= 0; k--) {
      double tt = grs[k];
      for (int j = k + 1; j <= last; j++) tt -= HH(k, j) * nrs[j];
      nrs[k] = tt / HH(k, k);
    }
    Coefs cf{};
    cf.k = last + 1;
    for (int j = 0; j <= last; j++) cf.c[j] = nrs[j];
    TS_HIP(ts, launch_basis_comb(x, V, n, cf, n, ts->s));
  }
  *its_out = its;
  return 0;
}

// Preconditioner z = B (x - y) (y may be null) with B the inverted point or point-block diagonal in d
struct Precond {
  HoneeMiTS     ts;
  const double *d;
  bool          block;
  int operator()(double *z, const double *x, const double *y) const {
    if (block) TS_HIP(ts, launch_pb_apply(z, d, x, y, ts->n / 5, ts->s));
    else TS_HIP(ts, launch_dscale_diff(z, d, x, y, ts->n, ts->s));
    return 0;
  }
};

// Invert an assembled diagonal (point: zero entries become 1) or point-block diagonal (singular blocks are an error)
int invert_diagonal(HoneeMiTS ts, double *d, bool block, const char *what) {
  if (!block) {
    TS_HIP(ts, launch_jacobi_invert(d, ts->n, ts->s));
    return 0;
  }
  const Publish pub = next_publish(ts, 1);
  TS_HIP(ts, launch_pb_invert(d, ts->n / 5, 0, ts->partial, ts->d_red, ts->s, pub));
  TS_CALL(wait_published(ts, pub));
  if (ts->h_red[0] != 0) return fail(ts, -6, "%s: %.0f singular point blocks", what, ts->h_red[0]);
  return 0;
}

int check_finite(HoneeMiTS ts, const double *x, const char *what) {
  const Publish pub = next_publish(ts, 1);
  TS_HIP(ts, launch_count_nonfinite(x, ts->n, 0, ts->partial, ts->d_red, ts->s, pub));
  TS_CALL(wait_published(ts, pub));
  if (ts->h_red[0] != 0) return fail(ts, -3, "%s has %.0f non-finite entries at step %d (t = %g)", what, ts->h_red[0], ts->st.steps, ts->t);
  return 0;
}

// ---- explicit Runge-Kutta ----
int mass_solve(HoneeMiTS ts, const double *state, const double *r, double *k) {
  HoneeMiProblem p = ts->p;
  const bool block = ts->o.mass_pc == 1;
  if (!ts->mass_pc_ready) {  // PCSetUp of the mass KSP at its first solve
    if (block) {
      TS_OP(ts, HoneeMiMassPointBlockDiagonal(p, kMemDevice, state, ts->mass_dinv));
    } else if (ts->o.mass_jacobi_rowsum) {
      TS_HIP(ts, launch_set(ts->w, 1.0, ts->n, ts->s));
      TS_OP(ts, HoneeMiApplyMass(p, kMemDevice, state, ts->w, ts->mass_dinv));
      ts->st.mass_applies++;
    } else {
      TS_OP(ts, HoneeMiMassDiagonal(p, kMemDevice, state, ts->mass_dinv));
    }
    TS_CALL(invert_diagonal(ts, ts->mass_dinv, block, "mass preconditioner"));
    ts->mass_pc_ready = true;
  }
  const Precond prec{ts, ts->mass_dinv, block};
  if (!ts->o.mass_gmres) return prec(k, r, nullptr);
  auto apply = [&](const double *xin, double *yout) -> int {
    ts->st.mass_applies++;
    TS_OP(ts, HoneeMiApplyMass(p, kMemDevice, state, xin, yout));
    return 0;
  };
  int its = 0;
  TS_CALL(gmres(ts, apply, prec, r, k, ts->o.mass_rtol, 1e-50, ts->o.mass_restart, ts->o.mass_max_it, &its));
  ts->st.mass_krylov_its += its;
  return 0;
}

int rk_step(HoneeMiTS ts) {
  const Tableau &T = ts->tab;
  const double   h = ts->o.dt;
  for (int i = 0; i < T.s; i++) {
    std::vector<double>         c{1.0};
    std::vector<const double *> xs{ts->q};
    for (int j = 0; j < i; j++) {
      c.push_back(h * T.A[(size_t)i * T.s + j]);
      xs.push_back(ts->K[j]);
    }
    TS_CALL(lincomb(ts, ts->Y, 0.0, c, xs));
    TS_OP(ts, HoneeMiApplyRHS(ts->p, kMemDevice, ts->t + T.c[i] * h, h, ts->Y, ts->R));
    ts->st.rhs_evals++;
    TS_CALL(mass_solve(ts, ts->Y, ts->R, ts->K[i]));
  }
  std::vector<double>         c;
  std::vector<const double *> xs;
  for (int i = 0; i < T.s; i++) {
    c.push_back(h * T.b[i]);
    xs.push_back(ts->K[i]);
  }
  TS_CALL(lincomb(ts, ts->q, 1.0, c, xs));
  ts->t += h;
  return 0;
}

// ---- implicit BDF with Newton-Krylov ----
// alpha_i = L_i'(T_0) of the Lagrange basis on nodes T_0..T_k
void bdf_coefficients(const double *T, int k, double *alpha) {
  alpha[0] = 0;
  for (int m = 1; m <= k; m++) alpha[0] += 1.0 / (T[0] - T[m]);
  for (int i = 1; i <= k; i++) {
    double v = 1.0 / (T[i] - T[0]);
    for (int j = 1; j <= k; j++)
      if (j != i) v *= (T[0] - T[j]) / (T[i] - T[j]);
    alpha[i] = v;
  }
}

// Lagrange basis values at t on nodes X_0..X_{m-1}
void lagrange_values(const double *X, int m, double t, double *L) {
  for (int i = 0; i < m; i++) {
    double v = 1;
    for (int j = 0; j < m; j++)
      if (j != i) v *= (t - X[j]) / (X[i] - X[j]);
    L[i] = v;
  }
}

// Stage residual R(x) = scale F(tnew, a x + xc, b x + hsum)
int ifunction_at(HoneeMiTS ts, double tnew, const double *x, double *F) {
  const HoneeMiTS_::Stage &sg = ts->stage;
  const double            *xs = x;
  if (sg.xc) {
    TS_CALL(lincomb(ts, ts->xa, 0.0, {sg.a, 1.0}, {x, ts->xc}));
    xs = ts->xa;
  }
  TS_CALL(lincomb(ts, ts->xdot, 0.0, {sg.b, 1.0}, {x, ts->hsum}));
  TS_OP(ts, HoneeMiApplyIFunction(ts->p, kMemDevice, tnew, sg.ctx_dt, xs, ts->xdot, F));
  if (sg.scale != 1.0) TS_CALL(lincomb(ts, F, sg.scale, {}, {}));
  ts->st.ifunction_evals++;
  return 0;
}

void set_stage(HoneeMiTS ts, double a, double b, double scale, double ctx_dt, bool xc) {
  ts->stage.a      = a;
  ts->stage.b      = b;
  ts->stage.scale  = scale;
  ts->stage.ctx_dt = ctx_dt;
  ts->stage.xc     = xc;
}

// Newton (SNESNEWTONLS) on the stage residual R(x) = scale F(tnew, a x + xc, b x + hsum) from the initial guess in
// ts->x.  scale a = 1 in every scheme here, so dR/dx is the IJacobian at shift b / a.
int newton_solve(HoneeMiTS ts, double tnew) {
  const double shift = ts->stage.b / ts->stage.a;
  const HoneeMiTSOptions &o = ts->o;
  double                  fnorm, fnorm0;
  TS_CALL(ifunction_at(ts, tnew, ts->x, ts->F));
  TS_CALL(norm2(ts, ts->F, &fnorm));
  fnorm0                = fnorm;
  ts->st.last_fnorm0    = fnorm0;
  ts->st.last_newton_its = 0;
  ts->st.last_krylov_its = 0;
  if (!std::isfinite(fnorm)) return fail(ts, -3, "non-finite residual at t = %g", tnew);
  const double ttol = std::fmax(o.snes_atol, o.snes_rtol * fnorm0);
  bool         converged = fnorm <= o.snes_atol;
  double       xnorm_mf  = 0;  // ||x|| of the iterate the finite-difference products are taken at
  auto         apply     = [&](const double *xin, double *yout) -> int {
    if (o.snes_mf_operator) {  // MatMFFD "wp": J v = (F(x + h v) - F(x)) / h, h = sqrt(eps) sqrt(1 + ||x||) / ||v||
      double vnorm;
      TS_CALL(norm2(ts, xin, &vnorm));
      if (vnorm == 0) return lincomb(ts, yout, 0.0, {0.0}, {xin});
      const double h = 1.4901161193847656e-08 * std::sqrt(1.0 + xnorm_mf) / vnorm;
      TS_CALL(lincomb(ts, ts->xtry, 0.0, {1.0, h}, {ts->x, xin}));  // xtry and Ftry are free until the line search
      TS_CALL(ifunction_at(ts, tnew, ts->xtry, ts->Ftry));
      TS_CALL(lincomb(ts, yout, 0.0, {1.0 / h, -1.0 / h}, {ts->Ftry, ts->F}));
      ts->st.jacobian_applies++;
      return 0;
    }
    ts->st.jacobian_applies++;
    TS_OP(ts, HoneeMiApplyIJacobian(ts->p, kMemDevice, ts->jac_shift, xin, yout));
    return 0;
  };
  const int lag = std::max(1, o.snes_lag_jacobian);
  for (int it = 1; it <= o.snes_max_it && !converged; it++) {
    const bool block = o.ksp_pc == 1;
    if (!ts->jac_ready || (it - 1 + ts->jac_iter) % lag == 0) {  // Jacobian evaluation
      ts->jac_shift = shift;
      ts->jac_ready = true;
      if ((it - 1) % std::max(1, o.pc_lag) == 0) {
        if (block) TS_OP(ts, HoneeMiIJacobianPointBlockDiagonal(ts->p, kMemDevice, shift, ts->dinv));
        else TS_OP(ts, HoneeMiIJacobianDiagonal(ts->p, kMemDevice, shift, ts->dinv));
        ts->st.diagonal_assemblies++;
        TS_CALL(invert_diagonal(ts, ts->dinv, block, "Jacobian preconditioner"));
      }
    }
    int kits = 0;
    if (o.snes_mf_operator) TS_CALL(norm2(ts, ts->x, &xnorm_mf));
    TS_CALL(gmres(ts, apply, Precond{ts, ts->dinv, block}, ts->F, ts->dx, o.ksp_rtol, o.ksp_atol, o.ksp_restart, o.ksp_max_it, &kits));
    ts->st.krylov_its += kits;
    ts->st.last_krylov_its += kits;
    // SNESLINESEARCHBT (cubic order, alpha 1e-4, minlambda 1e-12, 40 steps) along -dx, as SNESLineSearchApply_BT:
    // initial slope F . (J dx) (made negative, -1 if zero); the full step is taken when ||G||^2/2 <= ||F||^2/2 +
    // lambda alpha slope, else a quadratic fit and then cubic fits through the last two trials, each clamped to
    // [0.1, 0.5] of the previous lambda and accepted under the strict inequality.  A trial with a non-finite residual
    // cuts lambda by 10 without a fit.  Each trial reads ||F(xtry)|| and ||xtry|| (the step-length test's ||x|| if the
    // trial is accepted), the first one also ||dx|| and F . (J dx), in one host read.
    // SNESLINESEARCHBASIC takes the full step (one trial, accepted when its residual is finite); type 2 takes the slope
    // as -||F||^2 and fits quadratics only
    const bool basic = o.snes_linesearch_type == 1, norm_slope = o.snes_linesearch_type == 2;
    if (!basic && !norm_slope) {
      TS_CALL(apply(ts->dx, ts->w));  // J dx; w is free between Krylov solves
      TS_HIP(ts, launch_dot(ts->F, ts->w, ts->n, 3, ts->partial, ts->d_red, ts->s));
    }
    const double alpha_ls = 1e-4;
    double       initslope = 0, lambda = 1.0, gnorm = 0, xnorm = 0, ynorm = 0, lamprev = 0, gnormprev = 0;
    int          nfit = 0;
    for (int bt = 0;; bt++) {
      TS_CALL(lincomb(ts, ts->xtry, 0.0, {1.0, -lambda}, {ts->x, ts->dx}));
      TS_CALL(ifunction_at(ts, tnew, ts->xtry, ts->Ftry));
      const Publish pub = next_publish(ts, bt == 0 ? (basic || norm_slope ? 3 : 4) : 2);
      TS_HIP(ts, launch_dot(ts->Ftry, ts->Ftry, ts->n, 0, ts->partial, ts->d_red, ts->s));
      if (bt == 0) TS_HIP(ts, launch_dot(ts->dx, ts->dx, ts->n, 2, ts->partial, ts->d_red, ts->s));
      TS_HIP(ts, launch_dot(ts->xtry, ts->xtry, ts->n, 1, ts->partial, ts->d_red, ts->s, pub));
      TS_CALL(wait_published(ts, pub));
      gnorm = std::sqrt(ts->h_red[0]);
      xnorm = std::sqrt(ts->h_red[1]);
      if (bt == 0) ynorm = std::sqrt(ts->h_red[2]);
      if (basic) {
        if (!std::isfinite(gnorm)) return fail(ts, -4, "non-finite residual after a full Newton step at t = %g", tnew);
        break;
      }
      if (bt == 0) {
        initslope = norm_slope ? -fnorm * fnorm : ts->h_red[3];
        if (initslope > 0.0) initslope = -initslope;
        if (initslope == 0.0) initslope = -1.0;
      }
      const double lhs = 0.5 * gnorm * gnorm, rhs = 0.5 * fnorm * fnorm + lambda * alpha_ls * initslope;
      if (std::isfinite(gnorm) && (bt == 0 || norm_slope ? lhs <= rhs : lhs < rhs)) break;
      if (bt >= 40 || (!norm_slope && lambda <= 1e-12)) return fail(ts, -4, "line search failed at t = %g (||F|| = %g)", tnew, fnorm);
      ts->st.line_search_backtracks++;
      if (!std::isfinite(gnorm)) {
        lambda *= 0.1;
        continue;
      }
      double lt;
      if (nfit == 0 || norm_slope) {  // quadratic fit through f(0), f'(0) and f(lambda)
        lt = -initslope * lambda * lambda / (gnorm * gnorm - fnorm * fnorm - 2.0 * lambda * initslope);
      } else {  // cubic fit through f(0), f'(0), f(lambda) and f(lamprev)
        const double t1 = 0.5 * (gnorm * gnorm - fnorm * fnorm) - lambda * initslope;
        const double t2 = 0.5 * (gnormprev * gnormprev - fnorm * fnorm) - lamprev * initslope;
        const double a  = (t1 / (lambda * lambda) - t2 / (lamprev * lamprev)) / (lambda - lamprev);
        const double b  = (-lamprev * t1 / (lambda * lambda) + lambda * t2 / (lamprev * lamprev)) / (lambda - lamprev);
        const double d  = std::fmax(b * b - 3.0 * a * initslope, 0.0);
        lt              = a == 0.0 ? -initslope / (2.0 * b) : (-b + std::sqrt(d)) / (3.0 * a);
      }
      nfit++;
      lamprev   = lambda;
      gnormprev = gnorm;
      if (lt > 0.5 * lambda) lt = 0.5 * lambda;
      lambda = lt <= 0.1 * lambda ? 0.1 * lambda : lt;
    }
    std::swap(ts->x, ts->xtry);
    std::swap(ts->F, ts->Ftry);
    fnorm = gnorm;
    ts->st.newton_its++;
    ts->st.last_newton_its = it;
    if (fnorm <= ttol || lambda * ynorm <= o.snes_stol * xnorm) converged = true;
  }
  ts->st.last_fnorm = fnorm;
  if (!converged) return fail(ts, -4, "Newton did not converge at t = %g: ||F|| = %g from %g in %d iterations", tnew, fnorm, fnorm0, o.snes_max_it);
  ts->jac_iter = o.snes_lag_jacobian_persists ? ts->jac_iter + ts->st.last_newton_its : 0;
  return 0;
}


// Solve G(tnew, x, alpha_0 x + sum_i alpha_i hist_{i-1}) = 0 for x with the history's k most recent solutions
int bdf_solve(HoneeMiTS ts, double tnew, int k) {
  double T[4], alpha[4];
  T[0] = tnew;
  for (int i = 1; i <= k; i++) T[i] = ts->hist_t[i - 1];
  bdf_coefficients(T, k, alpha);
  {
    std::vector<double>         c;
    std::vector<const double *> xs;
    for (int i = 1; i <= k; i++) {
      c.push_back(alpha[i]);
      xs.push_back(ts->hist[i - 1]);
    }
    TS_CALL(lincomb(ts, ts->hsum, 0.0, c, xs));
  }
  {  // predictor
    const int m = std::min(k + 1, ts->nhist);
    double    L[4];
    lagrange_values(ts->hist_t.data(), m, tnew, L);
    std::vector<double>         c(L, L + m);
    std::vector<const double *> xs(ts->hist.begin(), ts->hist.begin() + m);
    TS_CALL(lincomb(ts, ts->x, 0.0, c, xs));
  }
  set_stage(ts, 1.0, alpha[0], 1.0, ts->o.dt, false);
  return newton_solve(ts, tnew);
}

// Push x (time tx) to the front of the history
int hist_push(HoneeMiTS ts, const double *x, double tx) {
  const int cap = (int)ts->hist.size();
  double   *last = ts->hist[cap - 1];
  for (int i = cap - 1; i > 0; i--) {
    ts->hist[i]   = ts->hist[i - 1];
    ts->hist_t[i] = ts->hist_t[i - 1];
  }
  ts->hist[0]   = last;
  ts->hist_t[0] = tx;
  TS_CALL(copy(ts, ts->hist[0], x));
  ts->nhist = std::min(ts->nhist + 1, cap);
  return 0;
}

int bdf_step(HoneeMiTS ts) {
  const double h = ts->o.dt;
  if (ts->nhist == 0) {  // restart
    TS_CALL(hist_push(ts, ts->q, ts->t));
    if (ts->o.bdf_order > 1) {
      TS_CALL(bdf_solve(ts, ts->t + h / 2, 1));
      TS_CALL(hist_push(ts, ts->x, ts->t + h / 2));
      ts->korder = 2;
    } else {
      ts->korder = 1;
    }
  } else {
    ts->korder = std::min(ts->korder + 1, ts->o.bdf_order);
  }
  TS_CALL(bdf_solve(ts, ts->t + h, ts->korder));
  TS_CALL(hist_push(ts, ts->x, ts->t + h));
  TS_CALL(copy(ts, ts->q, ts->x));
  ts->t += h;
  return 0;
}

// ---- backward Euler (TSBEULER: TSTHETA with theta 1) ----
// F(tprev + h, x, (x - xprev) / h) = 0 from the initial guess xprev, Jacobian shift 1 / h
int beuler_solve(HoneeMiTS ts, const double *xprev, double tprev, double h) {
  set_stage(ts, 1.0, 1.0 / h, 1.0, h, false);
  TS_CALL(lincomb(ts, ts->hsum, 0.0, {-1.0 / h}, {xprev}));
  TS_CALL(copy(ts, ts->x, xprev));
  return newton_solve(ts, tprev + h);
}

int beuler_step(HoneeMiTS ts) {
  TS_CALL(beuler_solve(ts, ts->q, ts->t, ts->o.dt));
  TS_CALL(copy(ts, ts->q, ts->x));
  ts->t += ts->o.dt;
  return 0;
}

// ---- generalized-alpha (TSALPHA, Jansen, Whiting and Hulbert 2000) ----
// With V1 = (X1 - X0) / (gamma dt) + (1 - 1/gamma) V0, Xa = X0 + alpha_f (X1 - X0) and Va = V0 + alpha_m (V1 - V0),
// the step solves F(t + alpha_f dt, Xa, Va) / alpha_f = 0 for X1 from X0, with Jacobian shift
// alpha_m / (alpha_f gamma dt); V0 is the previous step's V1.  The first step after a restart computes V0 from two
// backward-Euler half steps, V0 = (-3 X0 + 4 X(t + dt/2) - X(t + dt)) / dt (TSAlpha_Restart).
int alpha_step(HoneeMiTS ts) {
  const double dt = ts->o.dt, rho = ts->o.alpha_radius;
  const double am = 0.5 * (3.0 - rho) / (1.0 + rho), af = 1.0 / (1.0 + rho), gam = 0.5 + am - af;
  if (!ts->alpha_started) {
    const double h = 0.5 * dt;
    TS_CALL(beuler_solve(ts, ts->q, ts->t, h));
    TS_CALL(copy(ts, ts->xc, ts->x));
    TS_CALL(beuler_solve(ts, ts->xc, ts->t + h, h));
    TS_CALL(lincomb(ts, ts->v0, 0.0, {-3.0 / dt, 4.0 / dt, -1.0 / dt}, {ts->q, ts->xc, ts->x}));
    ts->alpha_started = true;
  }
  const double b = am / (gam * dt);
  set_stage(ts, af, b, 1.0 / af, dt, true);
  TS_CALL(lincomb(ts, ts->xc, 0.0, {1.0 - af}, {ts->q}));
  TS_CALL(lincomb(ts, ts->hsum, 0.0, {1.0 - am / gam, -b}, {ts->v0, ts->q}));
  TS_CALL(copy(ts, ts->x, ts->q));
  TS_CALL(newton_solve(ts, ts->t + af * dt));
  TS_CALL(lincomb(ts, ts->v0, 1.0 - 1.0 / gam, {1.0 / (gam * dt), -1.0 / (gam * dt)}, {ts->x, ts->q}));
  TS_CALL(copy(ts, ts->q, ts->x));
  ts->t += dt;
  return 0;
}

}  // namespace

extern "C" {

void HoneeMiTSOptionsDefault(HoneeMiTSOptions *o) {
  memset(o, 0, sizeof(*o));
  o->type               = HONEE_MI_TS_BDF;
  o->rk_type            = HONEE_MI_RK_5F;
  o->ssp_stages         = 4;
  o->bdf_order          = 2;
  o->dt                 = 1e-2;
  o->mass_gmres         = 0;
  o->mass_jacobi_rowsum = 1;
  o->mass_rtol          = 1e-5;
  o->mass_restart       = 30;
  o->mass_max_it        = 10000;
  o->snes_rtol          = 1e-8;
  o->snes_atol          = 1e-50;
  o->snes_stol          = 1e-8;
  o->snes_max_it        = 50;
  o->ksp_rtol           = 1e-5;
  o->ksp_atol           = 1e-50;
  o->ksp_restart        = 30;
  o->ksp_max_it         = 10000;
  o->ksp_pc             = 0;
  o->pc_lag             = 1;
  o->mass_pc            = 0;
  o->snes_mf_operator   = 0;
  o->alpha_radius       = 1.0;
  o->snes_lag_jacobian  = 1;
  o->snes_lag_jacobian_persists = 0;
  o->snes_linesearch_type       = 0;
}

int HoneeMiTSGetTableau(int rk_type, int ssp_stages, int max_stages, double *A, double *b, double *c, int *num_stages) {
  if (rk_type == HONEE_MI_RK_SSP_S3) {
    const int nn = (int)std::lround(std::sqrt((double)ssp_stages));
    if (nn < 2 || nn * nn != ssp_stages) return -5;
  }
  const Tableau T = make_tableau(rk_type, ssp_stages);
  if (T.s == 0) return -5;
  *num_stages = T.s;
  if (T.s > max_stages) return -5;
  for (int i = 0; i < T.s; i++) {
    b[i] = T.b[i];
    c[i] = T.c[i];
    for (int j = 0; j < T.s; j++) A[i * T.s + j] = T.A[(size_t)i * T.s + j];
  }
  return 0;
}

static char g_ts_error[512];

int HoneeMiTSCreate(HoneeMiProblem p, const HoneeMiTSOptions *o, HoneeMiTS *out) {
  *out          = nullptr;
  HoneeMiTS ts  = new HoneeMiTS_();
  ts->p         = p;
  ts->o         = *o;
  auto bail     = [&](int code) {
    snprintf(g_ts_error, sizeof(g_ts_error), "%s", ts->err.c_str());
    HoneeMiTSDestroy(&ts);
    return code;
  };
  const bool explicit_ts = o->type == HONEE_MI_TS_RK;
  if (explicit_ts != (bool)HoneeMiProblemIsExplicit(p))
    return bail(fail(ts, -5, "time integrator type does not match the problem (explicit RK needs is_implicit false, BDF, alpha and backward Euler need true)"));
  if (!(o->dt > 0)) return bail(fail(ts, -5, "dt must be positive"));
  if (o->ksp_pc < 0 || o->ksp_pc > 1 || o->mass_pc < 0 || o->mass_pc > 1) return bail(fail(ts, -5, "unknown preconditioner type"));
  void *stream = nullptr;
  if (CeedMiGetStream(HoneeMiProblemGetCeed(p), &stream)) return bail(fail(ts, -5, "cannot get the backend stream"));
  ts->s = (hipStream_t)stream;
  {
    CeedOperator op = HoneeMiProblemGetOperator(p, 0);
    CeedSize     lin = 0, lout = 0;
    if (CeedOperatorGetActiveVectorLengths(op, &lin, &lout) || lin <= 0) return bail(fail(ts, -5, "cannot get the L-vector length"));
    ts->n = lin;
  }
  int ierr;
  if ((ierr = dev_alloc(ts, &ts->q, ts->n))) return bail(ierr);
  if ((ierr = dev_alloc(ts, &ts->partial, (long long)kMaxReductionBlocks * kMaxDots))) return bail(ierr);
  if ((ierr = dev_alloc(ts, &ts->d_red, kMaxDots))) return bail(ierr);
  if (hipHostMalloc((void **)&ts->h_red, sizeof(double) * (kMaxDots + 1), hipHostMallocMapped | hipHostMallocCoherent) != hipSuccess)
    return bail(fail(ts, -2, "hipHostMalloc failed"));
  ts->h_seq  = reinterpret_cast<unsigned long long *>(ts->h_red + kMaxDots);
  *ts->h_seq = 0;
  if (hipHostGetDevicePointer((void **)&ts->g_red, ts->h_red, 0) != hipSuccess) return bail(fail(ts, -2, "hipHostGetDevicePointer failed"));
  ts->g_seq = reinterpret_cast<unsigned long long *>(ts->g_red + kMaxDots);
  const int restart = explicit_ts ? (o->mass_gmres ? o->mass_restart : 0) : o->ksp_restart;
  if (restart < 0 || restart + 1 > kMaxDots) return bail(fail(ts, -5, "GMRES restart must be at most %d", kMaxDots - 1));
  ts->restart = restart;
  if ((ierr = dev_alloc(ts, &ts->w, ts->n))) return bail(ierr);
  if (restart > 0 && (ierr = dev_alloc(ts, &ts->V, (long long)(restart + 1) * ts->n))) return bail(ierr);
  if (explicit_ts) {
    ts->tab = make_tableau(o->rk_type, o->ssp_stages);
    if (ts->tab.s == 0) return bail(fail(ts, -5, "unknown RK type %d", o->rk_type));
    if (o->rk_type == HONEE_MI_RK_SSP_S3) {
      const int nn = (int)std::lround(std::sqrt((double)o->ssp_stages));
      if (nn < 2 || nn * nn != o->ssp_stages) return bail(fail(ts, -5, "SSP third order needs a square number of stages >= 4"));
    }
    ts->K.assign(ts->tab.s, nullptr);
    for (auto &k : ts->K)
      if ((ierr = dev_alloc(ts, &k, ts->n))) return bail(ierr);
    if ((ierr = dev_alloc(ts, &ts->Y, ts->n)) || (ierr = dev_alloc(ts, &ts->R, ts->n)) ||
        (ierr = dev_alloc(ts, &ts->mass_dinv, o->mass_pc == 1 ? 5 * ts->n : ts->n)))
      return bail(ierr);
  } else {
    if (o->type == HONEE_MI_TS_BDF) {
      if (o->bdf_order < 1 || o->bdf_order > 2) return bail(fail(ts, -5, "BDF order must be 1 or 2"));
      ts->hist.assign(o->bdf_order + 1, nullptr);
      ts->hist_t.assign(o->bdf_order + 1, 0.0);
      for (auto &hv : ts->hist)
        if ((ierr = dev_alloc(ts, &hv, ts->n))) return bail(ierr);
    } else if (o->type == HONEE_MI_TS_ALPHA) {
      if (!(o->alpha_radius >= 0.0 && o->alpha_radius <= 1.0)) return bail(fail(ts, -5, "generalized-alpha radius must be in [0, 1]"));
      double **abufs[] = {&ts->v0, &ts->xa, &ts->xc};
      for (double **b : abufs)
        if ((ierr = dev_alloc(ts, b, ts->n))) return bail(ierr);
    } else if (o->type != HONEE_MI_TS_BEULER) {
      return bail(fail(ts, -5, "unknown time integrator type %d", o->type));
    }
    if (o->snes_lag_jacobian < 1) return bail(fail(ts, -5, "snes_lag_jacobian must be at least 1"));
    double **bufs[] = {&ts->x, &ts->xdot, &ts->hsum, &ts->F, &ts->Ftry, &ts->dx, &ts->xtry};
    for (double **b : bufs)
      if ((ierr = dev_alloc(ts, b, ts->n))) return bail(ierr);
    if ((ierr = dev_alloc(ts, &ts->dinv, o->ksp_pc == 1 ? 5 * ts->n : ts->n))) return bail(ierr);
  }
  *out = ts;
  return 0;
}

int HoneeMiTSSetSolution(HoneeMiTS ts, int mem_type, const double *q, double time) {
  TS_HIP(ts, hipMemcpyAsync(ts->q, q, sizeof(double) * (size_t)ts->n, mem_type == kMemHost ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice, ts->s));
  TS_HIP(ts, hipStreamSynchronize(ts->s));
  ts->t             = time;
  ts->nhist         = 0;
  ts->korder        = 1;
  ts->alpha_started = false;
  ts->jac_iter      = 0;
  ts->jac_ready     = false;
  ts->mass_pc_ready = false;
  ts->st            = HoneeMiTSStats{};
  ts->st.time       = time;
  return 0;
}

int HoneeMiTSGetSolution(HoneeMiTS ts, int mem_type, double *q) {
  TS_HIP(ts, hipMemcpyAsync(q, ts->q, sizeof(double) * (size_t)ts->n, mem_type == kMemHost ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice, ts->s));
  TS_HIP(ts, hipStreamSynchronize(ts->s));
  return 0;
}

const double *HoneeMiTSGetSolutionDevice(HoneeMiTS ts) { return ts->q; }

int HoneeMiTSStep(HoneeMiTS ts, int nsteps) {
  for (int i = 0; i < nsteps; i++) {
    switch (ts->o.type) {
      case HONEE_MI_TS_RK: TS_CALL(rk_step(ts)); break;
      case HONEE_MI_TS_ALPHA: TS_CALL(alpha_step(ts)); break;
      case HONEE_MI_TS_BEULER: TS_CALL(beuler_step(ts)); break;
      default: TS_CALL(bdf_step(ts));
    }
    ts->st.steps++;
    ts->st.time = ts->t;
    TS_CALL(check_finite(ts, ts->q, "solution"));
  }
  return 0;
}

int HoneeMiTSGetStats(HoneeMiTS ts, HoneeMiTSStats *stats) {
  *stats = ts->st;
  return 0;
}

const char *HoneeMiTSError(HoneeMiTS ts) { return ts ? ts->err.c_str() : g_ts_error; }

int HoneeMiTSDestroy(HoneeMiTS *pts) {
  HoneeMiTS ts = *pts;
  if (!ts) return 0;
  if (ts->s) (void)hipStreamSynchronize(ts->s);
  for (void *a : ts->allocs) (void)hipFree(a);
  if (ts->h_red) (void)hipHostFree(ts->h_red);
  delete ts;
  *pts = nullptr;
  return 0;
}

}  // extern "C"
