/*
 * dual.h -- forward-mode derivatives of the CPU oracle (TEST INFRASTRUCTURE
 * ONLY, like the rest of oracle/).
 *
 * scale_oracle.c compiled as C++ with this header forced in first
 * (`g++ -x c++ -include dual.h`) computes, next to every value, its
 * derivative along one direction: each real becomes a pair (value, tangent)
 * and every operation applies the chain rule.  The result is the
 * tangent-linear model of the oracle, statement by statement and exactly
 * (no differencing), used to check the adjoint by the dot-product identity
 * <L dx, y> = <dx, L^T y>.
 *
 * Where the forward takes a branch on a value (sign(), max/min, abs,
 * comparisons, the upwind switch, the FCT limiter's choices), the tangent
 * follows the branch the value takes: the derivative of the branch taken,
 * as the adjoint holds such switches fixed.  The FCT ratios R+/R- themselves
 * are held constant (ORC_FROZEN): the limiter frozen at its forward values.
 *
 * Memory layout: a Dual is {value, tangent}, 16 bytes, so an array of n
 * Duals is a (2, n) array of doubles with the pair index fastest.
 */
#ifndef ORACLE_DUAL_H
#define ORACLE_DUAL_H

#include <math.h>
#include <omp.h>
#include <stddef.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define restrict __restrict__

typedef decltype(0.0) dual_base;  // the built-in double (spelled so the macro below leaves it alone)

struct Dual {
  dual_base v, t;
  constexpr Dual() : v(0), t(0) {}
  constexpr Dual(dual_base x) : v(x), t(0) {}
  constexpr Dual(dual_base x, dual_base dx) : v(x), t(dx) {}
  constexpr Dual(int x) : v(x), t(0) {}
  constexpr Dual(long x) : v((dual_base)x), t(0) {}
  constexpr Dual(unsigned long x) : v((dual_base)x), t(0) {}
  constexpr Dual(float x) : v(x), t(0) {}
  explicit operator int() const { return (int)v; }
  Dual& operator+=(const Dual& b) { v += b.v; t += b.t; return *this; }
  Dual& operator-=(const Dual& b) { v -= b.v; t -= b.t; return *this; }
  Dual& operator*=(const Dual& b) { *this = Dual(v * b.v, t * b.v + v * b.t); return *this; }
  Dual& operator/=(const Dual& b) { *this = Dual(v / b.v, (t * b.v - v * b.t) / (b.v * b.v)); return *this; }
};

static inline Dual operator+(const Dual& a) { return a; }
static inline Dual operator-(const Dual& a) { return Dual(-a.v, -a.t); }
static inline Dual operator+(const Dual& a, const Dual& b) { return Dual(a.v + b.v, a.t + b.t); }
static inline Dual operator-(const Dual& a, const Dual& b) { return Dual(a.v - b.v, a.t - b.t); }
static inline Dual operator*(const Dual& a, const Dual& b) { return Dual(a.v * b.v, a.t * b.v + a.v * b.t); }
static inline Dual operator/(const Dual& a, const Dual& b) {
  return Dual(a.v / b.v, (a.t * b.v - a.v * b.t) / (b.v * b.v));
}
#define DUAL_MIXED(op)                                                               \
  static inline Dual operator op(const Dual& a, dual_base b) { return a op Dual(b); } \
  static inline Dual operator op(dual_base a, const Dual& b) { return Dual(a) op b; } \
  static inline Dual operator op(const Dual& a, int b) { return a op Dual(b); }       \
  static inline Dual operator op(int a, const Dual& b) { return Dual(a) op b; }
DUAL_MIXED(+)
DUAL_MIXED(-)
DUAL_MIXED(*)
DUAL_MIXED(/)
#undef DUAL_MIXED
#define DUAL_CMP(op)                                                                        \
  static inline bool operator op(const Dual& a, const Dual& b) { return a.v op b.v; }       \
  static inline bool operator op(const Dual& a, dual_base b) { return a.v op b; }           \
  static inline bool operator op(dual_base a, const Dual& b) { return a op b.v; }           \
  static inline bool operator op(const Dual& a, int b) { return a.v op (dual_base)b; }      \
  static inline bool operator op(int a, const Dual& b) { return (dual_base)a op b.v; }
DUAL_CMP(<)
DUAL_CMP(>)
DUAL_CMP(<=)
DUAL_CMP(>=)
DUAL_CMP(==)
DUAL_CMP(!=)
#undef DUAL_CMP

// |a|: the tangent takes the sign of the value (+ at zero, as copysign(1, +0))
static inline Dual fabs(const Dual& a) { return a.v < 0 ? -a : a; }
// copysign(|a|, b): |a| with the sign of b's value
static inline Dual copysign(const Dual& a, const Dual& b) {
  const Dual m = fabs(a);
  return signbit(b.v) ? -m : m;
}
static inline Dual sqrt(const Dual& a) {
  const dual_base s = sqrt(a.v);
  return Dual(s, a.t / (2 * s));
}
static inline Dual exp(const Dual& a) {
  const dual_base e = exp(a.v);
  return Dual(e, a.t * e);
}
static inline Dual log(const Dual& a) { return Dual(log(a.v), a.t / a.v); }
static inline Dual cos(const Dual& a) { return Dual(cos(a.v), -sin(a.v) * a.t); }
static inline Dual sin(const Dual& a) { return Dual(sin(a.v), cos(a.v) * a.t); }
static inline Dual pow(const Dual& a, dual_base b) {
  const dual_base p = pow(a.v, b);
  return Dual(p, b * pow(a.v, b - 1) * a.t);
}
static inline Dual pow(const Dual& a, const Dual& b) {
  const dual_base p = pow(a.v, b.v);
  return Dual(p, b.v * pow(a.v, b.v - 1) * a.t + p * log(a.v) * b.t);
}
static inline Dual pow(dual_base a, const Dual& b) {
  const dual_base p = pow(a, b.v);
  return Dual(p, p * log(a) * b.t);
}
static inline Dual fmax(const Dual& a, const Dual& b) { return a.v >= b.v ? a : b; }
static inline Dual fmin(const Dual& a, const Dual& b) { return a.v <= b.v ? a : b; }
static inline int isnan(const Dual& a) { return isnan(a.v); }
// piecewise constant: no tangent
static inline Dual ceil(const Dual& a) { return Dual(ceil(a.v)); }
static inline Dual floor(const Dual& a) { return Dual(floor(a.v)); }

// the FCT limiter's ratios R+/R- are held at their values (no tangent): the
// tangent-linear model of the frozen limiter, which the tracer adjoint
// transposes
#define ORC_FROZEN(x) (Dual((x).v, 0.0))

#define double Dual

#endif
