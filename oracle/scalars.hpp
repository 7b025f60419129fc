// Scalar types the oracle is instantiated with.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product library links this file.
//
//   double   - the parity path (bit-faithful restatement of the reference).
//   Counted  - a double that counts algorithmic floating-point work:
//              add/sub/mul/div/sqrt = 1 FLOP, fma = 2 FLOP,
//              abs/neg/min/max/compare = 0, acos/sin/cos = 1 T-op (separate).
//              This is the cost table the roofline numerator uses.
//   Dual     - forward-mode dual number (value, derivative along one seed
//              direction); branches follow the value, so the derivative is the
//              derivative of the oracle with its combinatorics held fixed.
#pragma once

#include <cmath>
#include <cstdint>

namespace vltr_oracle {

// ---------------------------------------------------------------- double ----
inline double vsqrt(double x) { return std::sqrt(x); }
inline double vfma(double a, double b, double c) { return std::fma(a, b, c); }
inline double vacos(double x) { return std::acos(x); }
inline double vsin(double x) { return std::sin(x); }
inline double vcos(double x) { return std::cos(x); }
inline double vabs(double x) { return std::fabs(x); }
// Rust f64::max / f64::min: a NaN argument yields the other argument.
inline double vmax(double a, double b) { return std::fmax(a, b); }
inline double vmin(double a, double b) { return std::fmin(a, b); }
inline double value_of(double x) { return x; }

// --------------------------------------------------------------- Counted ----
struct FlopCounters {
  uint64_t flops = 0;
  uint64_t tops = 0;
};
inline FlopCounters& flop_counters() {
  static thread_local FlopCounters c;
  return c;
}

struct Counted {
  double v = 0.0;
  Counted() = default;
  Counted(double x) : v(x) {}  // NOLINT: implicit from literals
};
inline Counted operator+(Counted a, Counted b) { flop_counters().flops++; return Counted(a.v + b.v); }
inline Counted operator-(Counted a, Counted b) { flop_counters().flops++; return Counted(a.v - b.v); }
inline Counted operator*(Counted a, Counted b) { flop_counters().flops++; return Counted(a.v * b.v); }
inline Counted operator/(Counted a, Counted b) { flop_counters().flops++; return Counted(a.v / b.v); }
inline Counted operator-(Counted a) { return Counted(-a.v); }
inline Counted& operator+=(Counted& a, Counted b) { a = a + b; return a; }
inline Counted& operator-=(Counted& a, Counted b) { a = a - b; return a; }
inline bool operator<(Counted a, Counted b) { return a.v < b.v; }
inline bool operator>(Counted a, Counted b) { return a.v > b.v; }
inline bool operator<=(Counted a, Counted b) { return a.v <= b.v; }
inline bool operator>=(Counted a, Counted b) { return a.v >= b.v; }
inline bool operator==(Counted a, Counted b) { return a.v == b.v; }
inline bool operator!=(Counted a, Counted b) { return a.v != b.v; }
inline Counted vsqrt(Counted x) { flop_counters().flops++; return Counted(std::sqrt(x.v)); }
inline Counted vfma(Counted a, Counted b, Counted c) { flop_counters().flops += 2; return Counted(std::fma(a.v, b.v, c.v)); }
inline Counted vacos(Counted x) { flop_counters().tops++; return Counted(std::acos(x.v)); }
inline Counted vsin(Counted x) { flop_counters().tops++; return Counted(std::sin(x.v)); }
inline Counted vcos(Counted x) { flop_counters().tops++; return Counted(std::cos(x.v)); }
inline Counted vabs(Counted x) { return Counted(std::fabs(x.v)); }
inline Counted vmax(Counted a, Counted b) { return Counted(std::fmax(a.v, b.v)); }
inline Counted vmin(Counted a, Counted b) { return Counted(std::fmin(a.v, b.v)); }
inline double value_of(Counted x) { return x.v; }

// ------------------------------------------------------------------ Dual ----
struct Dual {
  double v = 0.0;
  double d = 0.0;
  Dual() = default;
  Dual(double x) : v(x), d(0.0) {}  // NOLINT: implicit from literals
  Dual(double x, double dx) : v(x), d(dx) {}
};
inline Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
inline Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
inline Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
inline Dual operator/(Dual a, Dual b) {
  const double q = a.v / b.v;
  return Dual(q, (a.d - q * b.d) / b.v);
}
inline Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
inline Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
inline Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }
inline bool operator<(Dual a, Dual b) { return a.v < b.v; }
inline bool operator>(Dual a, Dual b) { return a.v > b.v; }
inline bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
inline bool operator>=(Dual a, Dual b) { return a.v >= b.v; }
inline bool operator==(Dual a, Dual b) { return a.v == b.v; }
inline bool operator!=(Dual a, Dual b) { return a.v != b.v; }
inline Dual vsqrt(Dual x) {
  const double s = std::sqrt(x.v);
  return Dual(s, s > 0.0 ? x.d / (2.0 * s) : 0.0);
}
inline Dual vfma(Dual a, Dual b, Dual c) {
  return Dual(std::fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d);
}
inline Dual vacos(Dual x) {
  const double one_m = 1.0 - x.v * x.v;
  return Dual(std::acos(x.v), one_m > 0.0 ? -x.d / std::sqrt(one_m) : 0.0);
}
inline Dual vsin(Dual x) { return Dual(std::sin(x.v), std::cos(x.v) * x.d); }
inline Dual vcos(Dual x) { return Dual(std::cos(x.v), -std::sin(x.v) * x.d); }
inline Dual vabs(Dual x) { return x.v < 0.0 ? -x : x; }
inline Dual vmax(Dual a, Dual b) {
  if (std::isnan(a.v)) return b;
  if (std::isnan(b.v)) return a;
  return a.v >= b.v ? a : b;
}
inline Dual vmin(Dual a, Dual b) {
  if (std::isnan(a.v)) return b;
  if (std::isnan(b.v)) return a;
  return a.v <= b.v ? a : b;
}
inline double value_of(Dual x) { return x.v; }

}  // namespace vltr_oracle
