// fdouble.hpp -- a FLOP-counting stand-in for double, force-included when
// oracle/z4c_oracle.c is compiled a second time as C++ (oracle/Makefile,
// target flops).  TEST INFRASTRUCTURE ONLY.
//
// Every +, -, *, / between floating-point operands (after constant folding
// of literal expressions, which the compiler does in the reference build
// too) increments one of four counters; each libm call (sqrt, pow, cbrt,
// exp, ...) increments the fifth; operations inside the finite-difference
// functions (marked ORACLE_FD_SCOPE in z4c_oracle.c) count separately.  Unary minus, comparisons and copies are
// not counted.  The struct holds a single double, so arrays and
// by-value arguments keep the layout and calling convention of double:
// the counting library takes the same ctypes calls as the plain one.
#pragma once
// the C headers z4c_oracle.c includes, seen before double is redefined
#include <float.h>
#include <limits.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <cmath>
#include <cstdlib>

struct fdouble;
extern "C" {
// [share][op]: share 0 = everything else, 1 = inside the finite-difference
// functions (ORACLE_FD_SCOPE); op = add, sub, mul, div, libm call
extern unsigned long long oracle_flop_counts[2][5];
extern int oracle_flop_share;
}
#define FDC(k) (++oracle_flop_counts[oracle_flop_share][k])

// Sets the stencil share for the rest of the enclosing function.
struct oracle_fd_scope_guard {
  int saved;
  oracle_fd_scope_guard() : saved(oracle_flop_share) { oracle_flop_share = 1; }
  ~oracle_fd_scope_guard() { oracle_flop_share = saved; }
};
#define ORACLE_FD_SCOPE oracle_fd_scope_guard oracle_fd_scope_

struct fdouble {
  double v;
  fdouble() = default;
  constexpr fdouble(double x) : v(x) {}
  explicit operator double() const { return v; }
  explicit operator long() const { return (long)v; }
  explicit operator int() const { return (int)v; }
  fdouble& operator+=(fdouble o) { FDC(0); v += o.v; return *this; }
  fdouble& operator-=(fdouble o) { FDC(1); v -= o.v; return *this; }
  fdouble& operator*=(fdouble o) { FDC(2); v *= o.v; return *this; }
  fdouble& operator/=(fdouble o) { FDC(3); v /= o.v; return *this; }
};
static_assert(sizeof(fdouble) == sizeof(double), "layout of double");

inline fdouble operator+(fdouble a, fdouble b) { FDC(0); return a.v + b.v; }
inline fdouble operator-(fdouble a, fdouble b) { FDC(1); return a.v - b.v; }
inline fdouble operator*(fdouble a, fdouble b) { FDC(2); return a.v * b.v; }
inline fdouble operator/(fdouble a, fdouble b) { FDC(3); return a.v / b.v; }
inline fdouble operator-(fdouble a) { return -a.v; }
inline fdouble operator+(fdouble a) { return a; }
inline bool operator<(fdouble a, fdouble b) { return a.v < b.v; }
inline bool operator>(fdouble a, fdouble b) { return a.v > b.v; }
inline bool operator<=(fdouble a, fdouble b) { return a.v <= b.v; }
inline bool operator>=(fdouble a, fdouble b) { return a.v >= b.v; }
inline bool operator==(fdouble a, fdouble b) { return a.v == b.v; }
inline bool operator!=(fdouble a, fdouble b) { return a.v != b.v; }

#define FD_MATH1(name)                                                       \
  inline fdouble name(fdouble a) { FDC(4); return std::name(a.v); }
FD_MATH1(sqrt)
FD_MATH1(cbrt)
FD_MATH1(exp)
FD_MATH1(log)
FD_MATH1(sin)
FD_MATH1(cos)
FD_MATH1(atan)
inline fdouble pow(fdouble a, fdouble b) { FDC(4); return std::pow(a.v, b.v); }
inline fdouble atan2(fdouble a, fdouble b) { FDC(4); return std::atan2(a.v, b.v); }
inline fdouble fabs(fdouble a) { return std::fabs(a.v); }
inline long lround(fdouble a) { return std::lround(a.v); }
inline bool isfinite(fdouble a) { return std::isfinite(a.v); }
inline bool isnan(fdouble a) { return std::isnan(a.v); }
inline fdouble floor(fdouble a) { return std::floor(a.v); }
inline fdouble fmax(fdouble a, fdouble b) { return std::fmax(a.v, b.v); }
inline fdouble fmin(fdouble a, fdouble b) { return std::fmin(a.v, b.v); }

#define double fdouble
