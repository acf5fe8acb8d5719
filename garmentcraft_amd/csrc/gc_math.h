// gc_math.h — per-element arithmetic of the implicit step, for the device.
//
// Templated on the element type T (double or float). With T = double and
// fp-contract off, every function performs the operations of the reference
// function it cites in the same order, so the FP64 build and operator agree
// with the reference CPU arithmetic. Paths relative to crates/klon3d-sim/src.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GC_HD __host__ __device__ __forceinline__

template <typename T>
struct V3 {
  T x, y, z;
};

template <typename T> GC_HD V3<T> v3(T a, T b, T c) { return V3<T>{a, b, c}; }
template <typename T> GC_HD V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <typename T> GC_HD V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <typename T> GC_HD V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
// koerper.rs:49-51: a0*b0 + a1*b1 + a2*b2, left to right
template <typename T> GC_HD T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <typename T> GC_HD V3<T> cross(V3<T> a, V3<T> b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
GC_HD float gsqrt(float x) { return sqrtf(x); }
GC_HD double gsqrt(double x) { return sqrt(x); }
GC_HD float gatan2(float y, float x) { return atan2f(y, x); }
GC_HD double gatan2(double y, double x) { return atan2(y, x); }
GC_HD float ground(float x) { return roundf(x); }
GC_HD double ground(double x) { return round(x); }
template <typename T> GC_HD T norm(V3<T> a) { return gsqrt(dot(a, a)); }
template <typename T> GC_HD V3<T> zero3() { return {T(0), T(0), T(0)}; }
template <typename D, typename S> GC_HD V3<D> cast3(V3<S> a) { return {D(a.x), D(a.y), D(a.z)}; }
GC_HD float gmax(float a, float b) { return fmaxf(a, b); }  // Rust f64::max: the non-NaN operand wins
GC_HD double gmax(double a, double b) { return fmax(a, b); }

// Membrane element record, the reference's `Membran` (kontinuum.rs:69-88).
template <typename T>
struct Membran {
  T flaeche;
  T r[4];       // ruhe_inv
  T kette[2];
  T schuss[2];
  T ek, es, g;  // kette_n_m, dehnung_n_m, schub_n_m
};

// `Hessen` (kontinuum.rs:116-127): F's columns and the projected geometric stress.
template <typename T>
struct Hessen {
  V3<T> f1, f2, w;
};

template <typename T> GC_HD V3<T> a3(const T w[2]) { return {w[0] * w[0], w[1] * w[1], w[0] * w[1]}; }  // kontinuum.rs:143-145
template <typename T> GC_HD V3<T> c3(const T w[2], const T v[2]) {                                       // 148-150
  return {w[0] * v[0], w[1] * v[1], T(0.5) * (w[0] * v[1] + w[1] * v[0])};
}
template <typename T> GC_HD V3<T> c1_(const Membran<T>& m) { return {-(m.r[0] + m.r[2]), m.r[0], m.r[2]}; }  // 302-305
template <typename T> GC_HD V3<T> c2_(const Membran<T>& m) { return {-(m.r[1] + m.r[3]), m.r[1], m.r[3]}; }  // 309-312
template <typename T> GC_HD T comp(V3<T> v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

// kontinuum.rs:158-177
template <typename T>
GC_HD V3<T> psd2(T a, T b, T c) {
  T m = T(0.5) * (a + b);
  T d = T(0.5) * (a - b);
  T r = gsqrt(d * d + c * c);
  if (m - r >= T(0)) return {a, b, c};
  T l_max = m + r;
  if (!(l_max > T(0))) return zero3<T>();
  T u0 = l_max - b, u1 = c;
  T n2 = u0 * u0 + u1 * u1;
  if (!(n2 > T(0))) return zero3<T>();
  T f = l_max / n2;
  return {f * u0 * u0, f * u1 * u1, f * u0 * u1};
}

// F columns (kontinuum.rs:315-322), from the two edge vectors.
template <typename T>
GC_HD void spalten(const Membran<T>& m, V3<T> e1, V3<T> e2, V3<T>* f1, V3<T>* f2) {
  *f1 = e1 * m.r[0] + e2 * m.r[2];
  *f2 = e1 * m.r[1] + e2 * m.r[3];
}

// kontinuum.rs:338-346
template <typename T>
GC_HD V3<T> dehnung_aus(const Membran<T>& m, V3<T> f1, V3<T> f2) {
  V3<T> a = a3(m.kette), b = a3(m.schuss), c = c3(m.kette, m.schuss);
  V3<T> g = {dot(f1, f1), dot(f2, f2), T(2) * dot(f1, f2)};
  return {T(0.5) * (dot(a, g) - T(1)), T(0.5) * (dot(b, g) - T(1)), T(0.5) * dot(c, g)};
}

// kontinuum.rs:376-386
template <typename T>
GC_HD V3<T> dU_dg_aus(const Membran<T>& m, V3<T> e) {
  V3<T> a = a3(m.kette), b = a3(m.schuss), c = c3(m.kette, m.schuss);
  V3<T> o;
  o.x = m.flaeche * (T(0.5) * m.ek * e.x * a.x + T(0.5) * m.es * e.y * b.x + T(2) * m.g * e.z * c.x);
  o.y = m.flaeche * (T(0.5) * m.ek * e.x * a.y + T(0.5) * m.es * e.y * b.y + T(2) * m.g * e.z * c.y);
  o.z = m.flaeche * (T(0.5) * m.ek * e.x * a.z + T(0.5) * m.es * e.y * b.z + T(2) * m.g * e.z * c.z);
  return o;
}

// kontinuum.rs:504-520, one vertex k: -J_k^T d
template <typename T>
GC_HD V3<T> kraft_k(const Membran<T>& m, V3<T> f1, V3<T> f2, V3<T> d, int k) {
  V3<T> c1 = c1_(m), c2 = c2_(m);
  T s1 = T(2) * (d.x * comp(c1, k) + d.z * comp(c2, k));
  T s2 = T(2) * (d.y * comp(c2, k) + d.z * comp(c1, k));
  return {-(s1 * f1.x + s2 * f2.x), -(s1 * f1.y + s2 * f2.y), -(s1 * f1.z + s2 * f2.z)};
}

// kontinuum.rs:475-478 (force) and 484-495 (tension-only force), all three vertices.
template <typename T>
GC_HD void krafte(const Membran<T>& m, V3<T> f1, V3<T> f2, V3<T> out[3], V3<T> zug[3]) {
  V3<T> e = dehnung_aus(m, f1, f2);
  V3<T> d = dU_dg_aus(m, e);
  V3<T> dz = dU_dg_aus(m, V3<T>{gmax(e.x, T(0)), gmax(e.y, T(0)), e.z});
  for (int k = 0; k < 3; ++k) {
    out[k] = kraft_k(m, f1, f2, d, k);
    zug[k] = kraft_k(m, f1, f2, dz, k);
  }
}

// kontinuum.rs:584-591 + 608-612
template <typename T>
GC_HD V3<T> geo_w_aus(const Membran<T>& m, V3<T> f1, V3<T> f2) {
  V3<T> e = dehnung_aus(m, f1, f2);
  V3<T> d = dU_dg_aus(m, V3<T>{gmax(e.x, T(0)), gmax(e.y, T(0)), e.z});
  return psd2(d.x, d.y, d.z);
}

// kontinuum.rs:421-435
template <typename T>
GC_HD V3<T> kg_mal(const Membran<T>& m, V3<T> q) {
  V3<T> a = a3(m.kette), b = a3(m.schuss), c = c3(m.kette, m.schuss);
  V3<T> w = zero3<T>();
  T t = m.flaeche * (T(0.25) * m.ek) * dot(a, q);
  w = w + a * t;
  t = m.flaeche * (T(0.25) * m.es) * dot(b, q);
  w = w + b * t;
  t = m.flaeche * m.g * dot(c, q);
  w = w + c * t;
  return w;
}

// kontinuum.rs:616-664 restricted to output vertex k (the other two outputs
// do not enter out[k]).
template <typename T>
GC_HD V3<T> mal_h_aus_k(const Membran<T>& m, const Hessen<T>& h, const V3<T> delta[3], int k) {
  V3<T> f1 = h.f1, f2 = h.f2;
  V3<T> c1 = c1_(m), c2 = c2_(m);
  V3<T> q = zero3<T>();
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    T d1 = dot(f1, delta[l]), d2 = dot(f2, delta[l]);
    q.x += T(2) * comp(c1, l) * d1;
    q.y += T(2) * comp(c2, l) * d2;
    q.z += T(2) * (comp(c2, l) * d1 + comp(c1, l) * d2);
  }
  V3<T> w = kg_mal(m, q);
  T s1 = T(2) * (comp(c1, k) * w.x + comp(c2, k) * w.z);
  T s2 = T(2) * (comp(c2, k) * w.y + comp(c1, k) * w.z);
  V3<T> out = {s1 * f1.x + s2 * f2.x, s1 * f1.y + s2 * f2.y, s1 * f1.z + s2 * f2.z};
  T w11 = h.w.x, w22 = h.w.y, w12 = h.w.z;
  V3<T> t1 = zero3<T>(), t2 = zero3<T>();
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    t1.x += comp(c1, b) * delta[b].x; t2.x += comp(c2, b) * delta[b].x;
    t1.y += comp(c1, b) * delta[b].y; t2.y += comp(c2, b) * delta[b].y;
    t1.z += comp(c1, b) * delta[b].z; t2.z += comp(c2, b) * delta[b].z;
  }
  T pa = T(2) * (w11 * comp(c1, k) + w12 * comp(c2, k));
  T qa = T(2) * (w12 * comp(c1, k) + w22 * comp(c2, k));
  out.x += pa * t1.x + qa * t2.x;
  out.y += pa * t1.y + qa * t2.y;
  out.z += pa * t1.z + qa * t2.z;
  return out;
}

// kontinuum.rs:616-664, all three outputs: the shared q and kg*q once, then
// per vertex the same operations as mal_h_aus_k.
template <typename T>
GC_HD void mal_h_aus_alle(const Membran<T>& m, const Hessen<T>& h, const V3<T> delta[3], V3<T> out[3]) {
  V3<T> f1 = h.f1, f2 = h.f2;
  V3<T> c1 = c1_(m), c2 = c2_(m);
  V3<T> q = zero3<T>();
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    T d1 = dot(f1, delta[l]), d2 = dot(f2, delta[l]);
    q.x += T(2) * comp(c1, l) * d1;
    q.y += T(2) * comp(c2, l) * d2;
    q.z += T(2) * (comp(c2, l) * d1 + comp(c1, l) * d2);
  }
  V3<T> w = kg_mal(m, q);
  T w11 = h.w.x, w22 = h.w.y, w12 = h.w.z;
  V3<T> t1 = zero3<T>(), t2 = zero3<T>();
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    t1.x += comp(c1, b) * delta[b].x; t2.x += comp(c2, b) * delta[b].x;
    t1.y += comp(c1, b) * delta[b].y; t2.y += comp(c2, b) * delta[b].y;
    t1.z += comp(c1, b) * delta[b].z; t2.z += comp(c2, b) * delta[b].z;
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    T s1 = T(2) * (comp(c1, k) * w.x + comp(c2, k) * w.z);
    T s2 = T(2) * (comp(c2, k) * w.y + comp(c1, k) * w.z);
    V3<T> o = {s1 * f1.x + s2 * f2.x, s1 * f1.y + s2 * f2.y, s1 * f1.z + s2 * f2.z};
    T pa = T(2) * (w11 * comp(c1, k) + w12 * comp(c2, k));
    T qa = T(2) * (w12 * comp(c1, k) + w22 * comp(c2, k));
    o.x += pa * t1.x + qa * t2.x;
    o.y += pa * t1.y + qa * t2.y;
    o.z += pa * t1.z + qa * t2.z;
    out[k] = o;
  }
}

// kontinuum.rs:675-708 for vertex k: J_k^T kg J_k + s_kk I, row-major 3x3.
template <typename T>
GC_HD void diag_block_k(const Membran<T>& m, const Hessen<T>& h, int k, T out[9]) {
  // J_k rows (kontinuum.rs:446-472)
  T J[3][3];
  const T* r = m.r;
  V3<T> f1 = h.f1, f2 = h.f2;
  auto zeile = [&](T a, T b, T* o) {
    o[0] = a * f1.x + b * f2.x; o[1] = a * f1.y + b * f2.y; o[2] = a * f1.z + b * f2.z;
  };
  if (k == 0) {
    T J1[3][3], J2[3][3];
    zeile(T(2) * r[0], T(0), J1[0]); zeile(T(0), T(2) * r[1], J1[1]); zeile(T(2) * r[1], T(2) * r[0], J1[2]);
    zeile(T(2) * r[2], T(0), J2[0]); zeile(T(0), T(2) * r[3], J2[1]); zeile(T(2) * r[3], T(2) * r[2], J2[2]);
    for (int i = 0; i < 3; ++i)
      for (int c = 0; c < 3; ++c) J[i][c] = -(J1[i][c] + J2[i][c]);
  } else if (k == 1) {
    zeile(T(2) * r[0], T(0), J[0]); zeile(T(0), T(2) * r[1], J[1]); zeile(T(2) * r[1], T(2) * r[0], J[2]);
  } else {
    zeile(T(2) * r[2], T(0), J[0]); zeile(T(0), T(2) * r[3], J[1]); zeile(T(2) * r[3], T(2) * r[2], J[2]);
  }
  // kg (kontinuum.rs:391-406)
  V3<T> a = a3(m.kette), b = a3(m.schuss), c = c3(m.kette, m.schuss);
  T kg[3][3];
  T ws[3] = {T(0.25) * m.ek, T(0.25) * m.es, m.g};
  V3<T> vs[3] = {a, b, c};
  for (int i = 0; i < 3; ++i)
    for (int q = 0; q < 3; ++q) kg[i][q] = T(0);
  for (int t = 0; t < 3; ++t)
    for (int i = 0; i < 3; ++i)
      for (int q = 0; q < 3; ++q) kg[i][q] += m.flaeche * ws[t] * comp(vs[t], i) * comp(vs[t], q);
  T tmp[3][3];
  for (int i = 0; i < 3; ++i)
    for (int cc = 0; cc < 3; ++cc) {
      T s = T(0);
      for (int mm = 0; mm < 3; ++mm) s += kg[i][mm] * J[mm][cc];
      tmp[i][cc] = s;
    }
  for (int aa = 0; aa < 3; ++aa)
    for (int cc = 0; cc < 3; ++cc) {
      T s = T(0);
      for (int i = 0; i < 3; ++i) s += J[i][aa] * tmp[i][cc];
      out[3 * aa + cc] = s;
    }
  V3<T> c1 = c1_(m), c2 = c2_(m);
  T w11 = h.w.x, w22 = h.w.y, w12 = h.w.z;
  T ck1 = comp(c1, k), ck2 = comp(c2, k);
  T s = T(2) * (w11 * ck1 * ck1 + w22 * ck2 * ck2 + T(2) * w12 * ck1 * ck2);
  out[0] += s; out[4] += s; out[8] += s;
}

// cpu.rs:279-282
template <typename T>
GC_HD T winkel_fehler(T theta, T ruhe) {
  const T TAU = T(6.283185307179586);
  T d = theta - ruhe;
  return d - TAU * ground(d / TAU);
}

// cpu.rs:288-307: signed dihedral angle about p0-p1 and its four gradients.
template <typename T>
GC_HD bool dieder(const V3<T> q[4], T* theta, V3<T> g[4]) {
  V3<T> x0 = q[0], x1 = q[1], x2 = q[2], x3 = q[3];
  V3<T> e = x1 - x0;
  T le = norm(e);
  V3<T> n1 = cross(x2 - x0, x2 - x1);
  V3<T> n2 = cross(x3 - x1, x3 - x0);
  T q1 = dot(n1, n1), q2 = dot(n2, n2);
  if (le < T(1e-12) || q1 < T(1e-30) || q2 < T(1e-30)) return false;
  V3<T> m1 = n1 * (T(1) / gsqrt(q1)), m2 = n2 * (T(1) / gsqrt(q2));
  *theta = gatan2(dot(cross(m1, m2), e * (T(1) / le)), dot(m1, m2));
  g[2] = n1 * (-le / q1);
  g[3] = n2 * (-le / q2);
  g[0] = (n1 * (dot(x2 - x1, e) / le / q1) + n2 * (dot(x3 - x1, e) / le / q2)) * T(-1);
  g[1] = n1 * (dot(x2 - x0, e) / le / q1) + n2 * (dot(x3 - x0, e) / le / q2);
  return true;
}

// 3x3 inverse by cofactors (implizit.rs:444-453). A singular or non-finite
// block gives zero, as the reference GPU backend does (gpu.rs:689-708).
template <typename T>
GC_HD void m3_inv(const T m[9], T o[9]) {
  T a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
  T det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  T s = T(1) / det;
  o[0] = (e * i - f * h) * s; o[1] = (c * h - b * i) * s; o[2] = (b * f - c * e) * s;
  o[3] = (f * g - d * i) * s; o[4] = (a * i - c * g) * s; o[5] = (c * d - a * f) * s;
  o[6] = (d * h - e * g) * s; o[7] = (b * g - a * h) * s; o[8] = (a * e - b * d) * s;
  if (!(det != T(0)) || !isfinite(s)) {
    for (int k = 0; k < 9; ++k) o[k] = T(0);
  }
}

template <typename T>
GC_HD V3<T> m3_mal(const T m[9], V3<T> v) {  // implizit.rs:441-443
  return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z, m[6] * v.x + m[7] * v.y + m[8] * v.z};
}

// ---------------------------------------------------------------------------
// Bodies. A body is a struct of plain values so the whole list rides in the
// kernel arguments; a mesh body's arrays are reached through one pointer.
// ---------------------------------------------------------------------------
#include "gc_netz.h"

enum { GC_ZYLINDER = 0, GC_SDF = 1, GC_NETZ = 2 };
struct Koerper {
  int32_t art;
  double punkt[3];          // cylinder axis point / SDF and mesh offset (versatz)
  double achse[3];          // cylinder axis (any length)
  double radius;
  double geschwindigkeit[3];
  const float* werte;       // SDF grid, x fastest
  int32_t dims[3];
  double origin[3];
  double spacing;
  const NetzDev* netz;      // mesh body (device memory)
};
#define GC_MAX_KOERPER 4

GC_HD double wert(double a) { return a; }

// Trilinear sample of the grid SDF and its gradient (per unit length).
// Templated on the scalar S of the query point (double in the step; a dual
// number in the adjoint, where the cell is chosen from the value).
template <typename S>
GC_HD bool sdf_sample(const Koerper& k, const S q[3], S* d, S grad[3]) {
  S t[3];
  int i[3];
  for (int a = 0; a < 3; ++a) {
    S u = (q[a] - S(k.origin[a])) / S(k.spacing);
    if (!(wert(u) >= 0.0) || !(wert(u) <= (double)(k.dims[a] - 1))) return false;
    int c = (int)floor(wert(u));
    if (c > k.dims[a] - 2) c = k.dims[a] - 2;
    i[a] = c;
    t[a] = u - S((double)c);
  }
  const size_t sx = 1, sy = (size_t)k.dims[0], sz = (size_t)k.dims[0] * (size_t)k.dims[1];
  const float* w0 = k.werte + (size_t)i[2] * sz + (size_t)i[1] * sy + (size_t)i[0];
  S v000 = S((double)w0[0]), v100 = S((double)w0[sx]), v010 = S((double)w0[sy]), v110 = S((double)w0[sy + sx]);
  S v001 = S((double)w0[sz]), v101 = S((double)w0[sz + sx]), v011 = S((double)w0[sz + sy]), v111 = S((double)w0[sz + sy + sx]);
  S tx = t[0], ty = t[1], tz = t[2];
  S x00 = v000 + tx * (v100 - v000), x10 = v010 + tx * (v110 - v010);
  S x01 = v001 + tx * (v101 - v001), x11 = v011 + tx * (v111 - v011);
  S y0 = x00 + ty * (x10 - x00), y1 = x01 + ty * (x11 - x01);
  *d = y0 + tz * (y1 - y0);
  S gx0 = (v100 - v000) + ty * ((v110 - v010) - (v100 - v000));
  S gx1 = (v101 - v001) + ty * ((v111 - v011) - (v101 - v001));
  grad[0] = (gx0 + tz * (gx1 - gx0)) / S(k.spacing);
  S gy0 = x10 - x00, gy1 = x11 - x01;
  grad[1] = (gy0 + tz * (gy1 - gy0)) / S(k.spacing);
  grad[2] = (y1 - y0) / S(k.spacing);
  return true;
}

// Koerper::kontakt of a mesh body (koerper.rs:82-90): the nearest surface
// point within max(suchtiefe, abstand), depth abstand - signed distance, the
// surface velocity at the nearest point. `roh` keeps a contact of depth <= 0.
GC_HD bool netz_kontakt(const Koerper& k, V3<double> x, double abstand, bool roh, V3<double>* normale, double* tiefe,
                        V3<double>* vk) {
  const NetzDev& N = *k.netz;
  const V3<double> versatz = {k.punkt[0], k.punkt[1], k.punkt[2]};
  const V3<double> starr = {k.geschwindigkeit[0], k.geschwindigkeit[1], k.geschwindigkeit[2]};
  nk::Treffer t;
  if (!nk::naechster(N, x - versatz, gmax(N.suchtiefe, abstand), &t)) return false;
  const double tf = abstand - t.abstand;
  if (!roh && !(tf > 0.0)) return false;
  *vk = nk::geschwindigkeit_an(N, t.dreieck, t.punkt, starr);
  *normale = t.normale;
  *tiefe = tf;
  return true;
}

template <typename S> struct IstDouble { static constexpr bool wert = false; };
template <> struct IstDouble<double> { static constexpr bool wert = true; };

// koerper.rs:69-91 (cylinder, mesh) and the same query on the grid SDF. The
// mesh body answers in FP64 only (the adjoint refuses it).
template <typename S>
GC_HD bool kontakt(const Koerper& k, V3<S> x, S abstand, V3<S>* normale, S* tiefe, V3<double>* vk) {
  *vk = V3<double>{k.geschwindigkeit[0], k.geschwindigkeit[1], k.geschwindigkeit[2]};
  if (k.art == GC_NETZ) {
    if constexpr (IstDouble<S>::wert) return netz_kontakt(k, x, abstand, false, normale, tiefe, vk);
    else return false;
  }
  if (k.art == GC_ZYLINDER) {
    V3<S> ach = {S(k.achse[0]), S(k.achse[1]), S(k.achse[2])};
    V3<S> a = ach * (S(1.0) / norm(ach));
    V3<S> d = x - V3<S>{S(k.punkt[0]), S(k.punkt[1]), S(k.punkt[2])};
    V3<S> quer = d - a * dot(d, a);
    S r = norm(quer);
    S t = S(k.radius) + abstand - r;
    if (wert(t) <= 0.0 || wert(r) < 1e-12) return false;
    *normale = quer * (S(1.0) / r);
    *tiefe = t;
    return true;
  }
  S q[3] = {x.x - S(k.punkt[0]), x.y - S(k.punkt[1]), x.z - S(k.punkt[2])};
  S dd, g[3];
  if (!sdf_sample(k, q, &dd, g)) return false;
  S t = abstand - dd;
  if (wert(t) <= 0.0) return false;
  V3<S> gv = {g[0], g[1], g[2]};
  S lg = norm(gv);
  if (!(wert(lg) > 1e-12)) return false;
  *normale = gv * (S(1.0) / lg);
  *tiefe = t;
  return true;
}

// The contact of kontakt() without its test for depth > 0: the normal and the
// (possibly negative) depth wherever the body's distance is defined. A step
// that replays recorded decisions (finite differences of a fixed branch)
// keeps a contact whose point has moved out by a rounding's width.
template <typename S>
GC_HD bool kontakt_roh(const Koerper& k, V3<S> x, S abstand, V3<S>* normale, S* tiefe, V3<double>* vk) {
  *vk = V3<double>{k.geschwindigkeit[0], k.geschwindigkeit[1], k.geschwindigkeit[2]};
  if (k.art == GC_NETZ) {
    if constexpr (IstDouble<S>::wert) return netz_kontakt(k, x, abstand, true, normale, tiefe, vk);
    else return false;
  }
  if (k.art == GC_ZYLINDER) {
    V3<S> ach = {S(k.achse[0]), S(k.achse[1]), S(k.achse[2])};
    V3<S> a = ach * (S(1.0) / norm(ach));
    V3<S> d = x - V3<S>{S(k.punkt[0]), S(k.punkt[1]), S(k.punkt[2])};
    V3<S> quer = d - a * dot(d, a);
    S r = norm(quer);
    if (wert(r) < 1e-12) return false;
    *normale = quer * (S(1.0) / r);
    *tiefe = S(k.radius) + abstand - r;
    return true;
  }
  S q[3] = {x.x - S(k.punkt[0]), x.y - S(k.punkt[1]), x.z - S(k.punkt[2])};
  S dd, g[3];
  if (!sdf_sample(k, q, &dd, g)) return false;
  V3<S> gv = {g[0], g[1], g[2]};
  S lg = norm(gv);
  if (!(wert(lg) > 1e-12)) return false;
  *normale = gv * (S(1.0) / lg);
  *tiefe = abstand - dd;
  return true;
}

// Entry of the path a -> b into the body: the counterpart of
// Koerpernetz::durchstoss (netzkoerper.rs:541-596, used at implizit.rs:984-1003)
// for the grid SDF that stands in for the mesh body. First sign change
// outside -> inside of the sampled distance at half-spacing samples along the
// path, refined by 60 bisections; the normal is the normalised gradient there.
// A path that starts inside does not enter; a cylinder has none (koerper.rs:101).
// t_hit: the entry's parameter along the path (p = a + (b - a) t_hit).
GC_HD bool durchstoss(const Koerper& k, V3<double> a, V3<double> b, V3<double>* p, V3<double>* n, V3<double>* vk,
                      double* t_hit = nullptr) {
  if (k.art == GC_ZYLINDER) return false;
  if (k.art == GC_NETZ) {  // koerper.rs:101-110, netzkoerper.rs:541-596
    const V3<double> versatz = {k.punkt[0], k.punkt[1], k.punkt[2]};
    const V3<double> starr = {k.geschwindigkeit[0], k.geschwindigkeit[1], k.geschwindigkeit[2]};
    V3<double> q;
    if (!nk::durchstoss(*k.netz, a - versatz, b - versatz, starr, &q, n, vk, t_hit)) return false;
    *p = q + versatz;
    return true;
  }
  const V3<double> d = b - a;
  if (dot(d, d) < 1e-24) return false;
  auto abstand = [&](V3<double> x, double* out) {
    const double q[3] = {x.x - k.punkt[0], x.y - k.punkt[1], x.z - k.punkt[2]};
    double g[3];
    return sdf_sample(k, q, out, g);
  };
  double d0;
  if (abstand(a, &d0) && !(d0 >= 0.0)) return false;
  // Only the part of the path inside the grid can enter (outside it nothing
  // is inside the body): clip to the grid box, so the samples are bounded by
  // the grid's diagonal whatever the path's length.
  double ta = 0.0, tb = 1.0;
  for (int q = 0; q < 3; ++q) {
    const double aq = q == 0 ? a.x : (q == 1 ? a.y : a.z), dq = q == 0 ? d.x : (q == 1 ? d.y : d.z);
    const double lo = k.origin[q] + k.punkt[q], hi = lo + (double)(k.dims[q] - 1) * k.spacing;
    if (dq == 0.0) {
      if (!(aq >= lo && aq <= hi)) return false;
      continue;
    }
    const double u0 = (lo - aq) / dq, u1 = (hi - aq) / dq;
    ta = fmax(ta, fmin(u0, u1));
    tb = fmin(tb, fmax(u0, u1));
  }
  if (!(ta <= tb)) return false;
  const double l = norm(d) * (tb - ta);
  if (!isfinite(l)) return false;
  int schritte = (int)ceil(l / (0.5 * k.spacing));
  if (schritte < 1) schritte = 1;
  double t0 = ta;
  for (int s = 1; s <= schritte; ++s) {
    const double t1 = ta + (tb - ta) * ((double)s / (double)schritte);
    double d1;
    if (!abstand(a + d * t1, &d1) || d1 >= 0.0) {
      t0 = t1;
      continue;
    }
    double lo = t0, hi = t1;
    for (int it = 0; it < 60; ++it) {
      const double m = 0.5 * (lo + hi);
      double dm;
      if (!abstand(a + d * m, &dm) || dm >= 0.0) lo = m;
      else hi = m;
    }
    *p = a + d * hi;
    if (t_hit) *t_hit = hi;
    const double q[3] = {p->x - k.punkt[0], p->y - k.punkt[1], p->z - k.punkt[2]};
    double dd, g[3];
    sdf_sample(k, q, &dd, g);
    const V3<double> gv = {g[0], g[1], g[2]};
    const double lg = norm(gv);
    if (!(lg > 1e-12)) return false;
    *n = gv * (1.0 / lg);
    *vk = V3<double>{k.geschwindigkeit[0], k.geschwindigkeit[1], k.geschwindigkeit[2]};
    return true;
  }
  return false;
}
