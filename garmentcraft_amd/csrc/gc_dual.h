// gc_dual.h — forward-mode dual numbers for the adjoint of the implicit step.
//
// A Dual carries a value and one directional derivative. The element and
// per-point functions of gc_math.h are templated on their scalar type; with
// T = Dual they return the value the FP64 forward computes (the value part
// performs exactly the double operations, so every branch is decided as in
// the forward step) together with its derivative along one seeded input
// direction. The adjoint kernels evaluate each small function once per input
// direction (at most 13) and contract the derivatives with the adjoint
// vector, which gives the exact vector-Jacobian product of the element.
#pragma once
#include "gc_math.h"

struct Dual {
  double v, d;
  __host__ __device__ Dual() : v(0.0), d(0.0) {}
  __host__ __device__ Dual(double a) : v(a), d(0.0) {}
  __host__ __device__ Dual(double a, double b) : v(a), d(b) {}
  __host__ __device__ Dual& operator+=(const Dual& b) {
    v += b.v;
    d += b.d;
    return *this;
  }
  __host__ __device__ Dual& operator-=(const Dual& b) {
    v -= b.v;
    d -= b.d;
    return *this;
  }
};

GC_HD Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
GC_HD Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
GC_HD Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
GC_HD Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
GC_HD Dual operator/(Dual a, Dual b) { return Dual(a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)); }
GC_HD bool operator<(Dual a, Dual b) { return a.v < b.v; }
GC_HD bool operator>(Dual a, Dual b) { return a.v > b.v; }
GC_HD bool operator<=(Dual a, Dual b) { return a.v <= b.v; }
GC_HD bool operator>=(Dual a, Dual b) { return a.v >= b.v; }

// sqrt: the derivative at 0 is taken as 0 (psd2 and the norms only reach 0
// on branches whose result does not use it).
GC_HD Dual gsqrt(Dual a) {
  const double s = sqrt(a.v);
  return Dual(s, s > 0.0 ? a.d / (2.0 * s) : 0.0);
}
GC_HD Dual gatan2(Dual y, Dual x) {
  const double q = x.v * x.v + y.v * y.v;
  return Dual(atan2(y.v, x.v), q > 0.0 ? (x.v * y.d - y.v * x.d) / q : 0.0);
}
GC_HD Dual ground(Dual a) { return Dual(round(a.v), 0.0); }
// Rust f64::max as in gc_math.h: the larger operand (and its derivative)
GC_HD Dual gmax(Dual a, Dual b) { return (a.v >= b.v || b.v != b.v) ? a : b; }
GC_HD double wert(Dual a) { return a.v; }
GC_HD double abl(Dual a) { return a.d; }
GC_HD bool isfinite(Dual a) { return isfinite(a.v); }
