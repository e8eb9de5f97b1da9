// CPU ORACLE -- test infrastructure only (see ns_pointwise.h).
//
// Exact arithmetic count of the Newtonian QFunction bodies at one quadrature point.  The text of ns_qfunctions.h (and
// the pointwise helpers it uses) is compiled a second time with `double` standing for `opc::Num`, a scalar that counts
// every floating-point operation it performs, so the count is of the same code the oracle runs, on the same inputs
// (the outputs are returned too and equal the oracle's bit for bit).
//
// Counting rules:
//  * every +, -, *, / between two runtime values counts one operation of its kind;
//  * an operation with one literal operand counts too, except a product by the literal 1 or -1 (an exact identity
//    any compiler folds away); operations between literals only are constant folded and not counted;
//  * sqrt, exp and log count one each, in their own columns; negation, comparisons and copies count nothing.
// Literal operands are recognised by type: a literal is a C `double` (or int) expression, which selects the
// mixed-type overloads below; every value that has passed through a variable is an opc::Num.
#include <math.h>
#include <stdint.h>
#include <string.h>

namespace opc {

enum { kAdd = 0, kMul, kDiv, kSqrt, kExp, kLog, kNumKinds };
static thread_local long g_count[kNumKinds];

struct Num {
  double v;
  Num() = default;
  Num(double x) : v(x) {}  // NOLINT: implicit, as `double` converts from int and double literals
};
static_assert(sizeof(Num) == sizeof(double), "Num must be layout-compatible with double");

static inline Num tick(int k, double r) {
  g_count[k]++;
  return Num(r);
}
static inline bool unit(double a) { return a == 1.0 || a == -1.0; }

static inline Num operator+(Num a, Num b) { return tick(kAdd, a.v + b.v); }
static inline Num operator-(Num a, Num b) { return tick(kAdd, a.v - b.v); }
static inline Num operator*(Num a, Num b) { return tick(kMul, a.v * b.v); }
static inline Num operator/(Num a, Num b) { return tick(kDiv, a.v / b.v); }
static inline Num operator+(double a, Num b) { return tick(kAdd, a + b.v); }
static inline Num operator-(double a, Num b) { return tick(kAdd, a - b.v); }
static inline Num operator/(double a, Num b) { return tick(kDiv, a / b.v); }
static inline Num operator+(Num a, double b) { return tick(kAdd, a.v + b); }
static inline Num operator-(Num a, double b) { return tick(kAdd, a.v - b); }
static inline Num operator/(Num a, double b) { return tick(kDiv, a.v / b); }
static inline Num operator*(double a, Num b) { return unit(a) ? Num(a * b.v) : tick(kMul, a * b.v); }
static inline Num operator*(Num a, double b) { return unit(b) ? Num(a.v * b) : tick(kMul, a.v * b); }
static inline Num operator-(Num a) { return Num(-a.v); }
static inline Num &operator+=(Num &a, Num b) { return a = a + b; }
static inline Num &operator-=(Num &a, Num b) { return a = a - b; }
static inline Num &operator*=(Num &a, Num b) { return a = a * b; }
static inline Num &operator/=(Num &a, Num b) { return a = a / b; }
static inline bool operator<(Num a, Num b) { return a.v < b.v; }
static inline bool operator>(Num a, Num b) { return a.v > b.v; }
static inline bool operator<=(Num a, Num b) { return a.v <= b.v; }
static inline bool operator>=(Num a, Num b) { return a.v >= b.v; }
static inline bool operator==(Num a, Num b) { return a.v == b.v; }
static inline bool operator!=(Num a, Num b) { return a.v != b.v; }
static inline Num sqrt(Num a) { return tick(kSqrt, ::sqrt(a.v)); }
static inline Num exp(Num a) { return tick(kExp, ::exp(a.v)); }
static inline Num log(Num a) { return tick(kLog, ::log(a.v)); }
static inline Num fabs(Num a) { return Num(::fabs(a.v)); }

#define double Num
#include "ns_qfunctions.h"
#undef double

static void load(Num *dst, const double *src, int n) {
  for (int i = 0; i < n; i++) dst[i] = Num(src[i]);
}
static void store(double *dst, const Num *src, int n) {
  for (int i = 0; i < n; i++) dst[i] = src[i].v;
}
static void start() {
  for (int k = 0; k < kNumKinds; k++) g_count[k] = 0;
}
static void finish(long *counts) {
  for (int k = 0; k < kNumKinds; k++) counts[k] = g_count[k];
}

}  // namespace opc

// The conversion from conservative to primitive variables alone, for a hand count of the rules above.
extern "C" int oracle_opcount_prim_from_cons(const void *gas, const double *U, double *Y, long *counts) {
  using namespace opc;
  OGas g;
  memcpy(&g, gas, sizeof g);
  Num un[5], yn[5];
  load(un, U, 5);
  start();
  o_prim_from_cons(g, un, yn);
  finish(counts);
  store(Y, yn, 5);
  return 0;
}

// ctx points to the oracle's ONewt (honee_oracle.h); opc::ONewt has the same layout with Num for double.
// counts[6] = additions/subtractions, multiplications, divisions, sqrt, exp, log.
extern "C" int oracle_opcount_ifunction(const void *ctx, const double *q, const double *gref, const double *qdot, double wdetJ,
                                        const double *dXdx, const double *x, double *v, double *gv, double *jd, long *counts) {
  using namespace opc;
  ONewt c;
  memcpy(&c, ctx, sizeof c);
  Num qn[5], gn[3][5], qdn[5], dn[3][3], xn[3], vn[5], gvn[3][5], jdn[15] = {};
  load(qn, q, 5);
  load(&gn[0][0], gref, 15);
  load(qdn, qdot, 5);
  load(&dn[0][0], dXdx, 9);
  load(xn, x, 3);
  start();
  qf_ifunction(&c, qn, gn, qdn, Num(wdetJ), dn, xn, vn, gvn, jdn);
  finish(counts);
  store(v, vn, 5);
  store(gv, &gvn[0][0], 15);
  store(jd, jdn, njac(&c));
  return 0;
}

extern "C" int oracle_opcount_ijacobian(const void *ctx, const double *dq, const double *gref, double wdetJ, const double *dXdx,
                                        const double *jd, double *v, double *gv, long *counts) {
  using namespace opc;
  ONewt c;
  memcpy(&c, ctx, sizeof c);
  Num dqn[5], gn[3][5], dn[3][3], jdn[15] = {}, vn[5], gvn[3][5];
  load(dqn, dq, 5);
  load(&gn[0][0], gref, 15);
  load(&dn[0][0], dXdx, 9);
  load(jdn, jd, njac(&c));
  start();
  qf_ijacobian(&c, dqn, gn, Num(wdetJ), dn, jdn, vn, gvn);
  finish(counts);
  store(v, vn, 5);
  store(gv, &gvn[0][0], 15);
  return 0;
}

extern "C" int oracle_opcount_rhs(const void *ctx, const double *U, const double *gref, double wdetJ, const double *dXdx, const double *x,
                                  double *v, double *gv, long *counts) {
  using namespace opc;
  ONewt c;
  memcpy(&c, ctx, sizeof c);
  Num un[5], gn[3][5], dn[3][3], xn[3], vn[5], gvn[3][5];
  load(un, U, 5);
  load(&gn[0][0], gref, 15);
  load(&dn[0][0], dXdx, 9);
  load(xn, x, 3);
  start();
  qf_rhs(&c, un, gn, Num(wdetJ), dn, xn, vn, gvn);
  finish(counts);
  store(v, vn, 5);
  store(gv, &gvn[0][0], 15);
  return 0;
}
