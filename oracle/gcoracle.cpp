// gcoracle.cpp — FP64 CPU restatement of klon3d-sim's implicit step.
//
// TEST INFRASTRUCTURE ONLY (the checker). Never linked into the product.
//
// Citations are "file:line" in the reference checkout, relative to
// crates/klon3d-sim/src unless stated otherwise. Loop order, summation order and
// branch order follow the cited code so that, compiled without FMA contraction
// (-ffp-contract=off), the numbers are those the reference CPU path computes.
#include "gcoracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

namespace {

using V3 = std::array<double, 3>;
using M3 = std::array<V3, 3>;

// koerper.rs:37-54
inline V3 sub(V3 a, V3 b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
inline V3 add(V3 a, V3 b) { return {a[0] + b[0], a[1] + b[1], a[2] + b[2]}; }
inline V3 mul(V3 a, double s) { return {a[0] * s, a[1] * s, a[2] * s}; }
inline double dot(V3 a, V3 b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
inline V3 cross(V3 a, V3 b) { return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }
// Rust f64::max / f64::min: the non-NaN operand wins.
inline double fmax_(double a, double b) { return std::fmax(a, b); }
inline double fmin_(double a, double b) { return std::fmin(a, b); }

// ---------------------------------------------------------------------------
// Membrane (kontinuum.rs)
// ---------------------------------------------------------------------------
struct Membran {  // kontinuum.rs:69-88
  double flaeche;
  double ruhe_inv[4];
  double kette[2];
  double schuss[2];
  double kette_n_m, dehnung_n_m, schub_n_m;
};
struct Hessen {  // kontinuum.rs:116-127
  V3 f1, f2, w;
};

Membran membran_aus(const double* e) {
  Membran m;
  m.flaeche = e[0];
  for (int i = 0; i < 4; ++i) m.ruhe_inv[i] = e[1 + i];
  m.kette[0] = e[5]; m.kette[1] = e[6];
  m.schuss[0] = e[7]; m.schuss[1] = e[8];
  m.kette_n_m = e[9]; m.dehnung_n_m = e[10]; m.schub_n_m = e[11];
  return m;
}
void membran_in(const Membran& m, double* e) {
  e[0] = m.flaeche;
  for (int i = 0; i < 4; ++i) e[1 + i] = m.ruhe_inv[i];
  e[5] = m.kette[0]; e[6] = m.kette[1]; e[7] = m.schuss[0]; e[8] = m.schuss[1];
  e[9] = m.kette_n_m; e[10] = m.dehnung_n_m; e[11] = m.schub_n_m;
}

// kontinuum.rs:143-150
inline V3 a3(const double w[2]) { return {w[0] * w[0], w[1] * w[1], w[0] * w[1]}; }
inline V3 c3(const double w[2], const double v[2]) { return {w[0] * v[0], w[1] * v[1], 0.5 * (w[0] * v[1] + w[1] * v[0])}; }

// kontinuum.rs:158-177
V3 psd2(double a, double b, double c) {
  double m = 0.5 * (a + b);
  double d = 0.5 * (a - b);
  double r = std::sqrt(d * d + c * c);
  if (m - r >= 0.0) return {a, b, c};
  double l_max = m + r;
  if (!(l_max > 0.0)) return {0.0, 0.0, 0.0};
  double u0 = l_max - b, u1 = c;
  double n2 = u0 * u0 + u1 * u1;
  if (!(n2 > 0.0)) return {0.0, 0.0, 0.0};
  double f = l_max / n2;
  return {f * u0 * u0, f * u1 * u1, f * u0 * u1};
}

// kontinuum.rs:193-216
bool rest_aus_kanten(const double kanten[3], double inv[4], double* flaeche) {
  double l01 = kanten[0], l12 = kanten[1], l20 = kanten[2];
  if (!(l01 > 0.0 && l12 > 0.0 && l20 > 0.0)) return false;
  double x = (l01 * l01 + l20 * l20 - l12 * l12) / (2.0 * l01);
  double y2 = l20 * l20 - x * x;
  if (!(y2 > 0.0)) return false;
  double y = std::sqrt(y2);
  double det = l01 * y;
  if (!(det > 0.0)) return false;
  inv[0] = 1.0 / l01; inv[1] = -x / det; inv[2] = 0.0; inv[3] = 1.0 / y;
  *flaeche = 0.5 * det;
  return true;
}

// kontinuum.rs:229-286
bool membran_bauen(const double ruhe[3][2], const double kanten[3], const double dir2[2], const double material[3], Membran* out) {
  Membran m{};
  if (!rest_aus_kanten(kanten, m.ruhe_inv, &m.flaeche)) return false;
  V3 d01 = sub(V3{ruhe[1][0], ruhe[1][1], 0.0}, V3{ruhe[0][0], ruhe[0][1], 0.0});
  double l01 = norm(d01);
  if (!(l01 > 0.0)) return false;
  double c = d01[0] / l01, s = d01[1] / l01;
  auto dreh = [&](double v0, double v1, double o[2]) { o[0] = c * v0 + s * v1; o[1] = -s * v0 + c * v1; };
  auto nach_ruhe = [&](double v0, double v1, double o[2]) -> bool {
    double l = norm(V3{v0, v1, 0.0});
    if (l > 0.0) { dreh(v0 / l, v1 / l, o); return true; }
    return false;
  };
  if (!nach_ruhe(dir2[0], dir2[1], m.kette)) return false;
  double q[2];
  if (!nach_ruhe(-dir2[1], dir2[0], q)) return false;
  double along = q[0] * m.kette[0] + q[1] * m.kette[1];
  double w0 = q[0] - along * m.kette[0], w1 = q[1] - along * m.kette[1];
  double lw = norm(V3{w0, w1, 0.0});
  if (!(lw > 1e-12)) return false;
  m.schuss[0] = w0 / lw; m.schuss[1] = w1 / lw;
  m.kette_n_m = material[0]; m.dehnung_n_m = material[1]; m.schub_n_m = material[2];
  *out = m;
  return true;
}

struct Basis { V3 a, b, c; };
inline Basis basis(const Membran& m) { return {a3(m.kette), a3(m.schuss), c3(m.kette, m.schuss)}; }  // kontinuum.rs:292-298
inline V3 c1_(const Membran& m) { const double* r = m.ruhe_inv; return {-(r[0] + r[2]), r[0], r[2]}; }  // 302-305
inline V3 c2_(const Membran& m) { const double* r = m.ruhe_inv; return {-(r[1] + r[3]), r[1], r[3]}; }  // 309-312

// kontinuum.rs:315-322
inline void f_(const Membran& m, const V3 p[3], V3* f1, V3* f2) {
  V3 e1 = sub(p[1], p[0]), e2 = sub(p[2], p[0]);
  const double* r = m.ruhe_inv;
  *f1 = add(mul(e1, r[0]), mul(e2, r[2]));
  *f2 = add(mul(e1, r[1]), mul(e2, r[3]));
}
inline V3 g_aus(V3 f1, V3 f2) { return {dot(f1, f1), dot(f2, f2), 2.0 * dot(f1, f2)}; }  // 326-328

// kontinuum.rs:338-346
V3 dehnung_aus(const Membran& m, V3 f1, V3 f2) {
  Basis B = basis(m);
  V3 g = g_aus(f1, f2);
  return {0.5 * (dot(B.a, g) - 1.0), 0.5 * (dot(B.b, g) - 1.0), 0.5 * dot(B.c, g)};
}
V3 dehnung(const Membran& m, const V3 p[3]) { V3 f1, f2; f_(m, p, &f1, &f2); return dehnung_aus(m, f1, f2); }  // 331-334

// kontinuum.rs:351-358
V3 spannung(const Membran& m, const V3 p[3]) {
  V3 e = dehnung(m, p);
  return {m.kette_n_m * e[0], m.dehnung_n_m * e[1], 2.0 * m.schub_n_m * e[2]};
}
// kontinuum.rs:361-367
double energie(const Membran& m, const V3 p[3]) {
  V3 e = dehnung(m, p);
  return m.flaeche * (0.5 * m.kette_n_m * e[0] * e[0] + 0.5 * m.dehnung_n_m * e[1] * e[1] + 2.0 * m.schub_n_m * e[2] * e[2]);
}
// kontinuum.rs:376-386
V3 dU_dg_aus(const Membran& m, V3 e) {
  Basis B = basis(m);
  V3 out{};
  for (int i = 0; i < 3; ++i)
    out[i] = m.flaeche * (0.5 * m.kette_n_m * e[0] * B.a[i] + 0.5 * m.dehnung_n_m * e[1] * B.b[i] + 2.0 * m.schub_n_m * e[2] * B.c[i]);
  return out;
}
// kontinuum.rs:391-406
std::array<V3, 3> kg_aus_basis(const Membran& m) {
  Basis B = basis(m);
  std::array<V3, 3> M{};
  const double ws[3] = {0.25 * m.kette_n_m, 0.25 * m.dehnung_n_m, m.schub_n_m};
  const V3 vs[3] = {B.a, B.b, B.c};
  for (int t = 0; t < 3; ++t)
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) M[i][k] += m.flaeche * ws[t] * vs[t][i] * vs[t][k];
  return M;
}
// kontinuum.rs:421-435
V3 kg_mal(const Membran& m, V3 q) {
  Basis B = basis(m);
  V3 w{};
  const double ss[3] = {0.25 * m.kette_n_m, 0.25 * m.dehnung_n_m, m.schub_n_m};
  const V3 vs[3] = {B.a, B.b, B.c};
  for (int t = 0; t < 3; ++t) {
    double tt = m.flaeche * ss[t] * dot(vs[t], q);
    for (int i = 0; i < 3; ++i) w[i] += tt * vs[t][i];
  }
  return w;
}
// kontinuum.rs:446-472 ; J[k][row][col]
void jacobi_aus(const Membran& m, V3 f1, V3 f2, double J[3][3][3]) {
  const double* r = m.ruhe_inv;
  auto zeile = [&](double a, double b, double* o) {
    o[0] = a * f1[0] + b * f2[0]; o[1] = a * f1[1] + b * f2[1]; o[2] = a * f1[2] + b * f2[2];
  };
  zeile(2.0 * r[0], 0.0, J[1][0]); zeile(0.0, 2.0 * r[1], J[1][1]); zeile(2.0 * r[1], 2.0 * r[0], J[1][2]);
  zeile(2.0 * r[2], 0.0, J[2][0]); zeile(0.0, 2.0 * r[3], J[2][1]); zeile(2.0 * r[3], 2.0 * r[2], J[2][2]);
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) J[0][i][k] = -(J[1][i][k] + J[2][i][k]);
}
// kontinuum.rs:504-520
void kraefte_aus(const Membran& m, const V3 p[3], V3 d, V3 out[3]) {
  V3 f1, f2; f_(m, p, &f1, &f2);
  V3 c1 = c1_(m), c2 = c2_(m);
  for (int k = 0; k < 3; ++k) {
    double s1 = 2.0 * (d[0] * c1[k] + d[2] * c2[k]);
    double s2 = 2.0 * (d[1] * c2[k] + d[2] * c1[k]);
    out[k] = {-(s1 * f1[0] + s2 * f2[0]), -(s1 * f1[1] + s2 * f2[1]), -(s1 * f1[2] + s2 * f2[2])};
  }
}
// kontinuum.rs:475-478 (with dU_dg 370-372)
void krafte(const Membran& m, const V3 p[3], V3 out[3]) { kraefte_aus(m, p, dU_dg_aus(m, dehnung(m, p)), out); }
// kontinuum.rs:484-495
void krafte_zug(const Membran& m, const V3 p[3], V3 out[3]) {
  V3 e = dehnung(m, p);
  Basis B = basis(m);
  V3 d{};
  for (int i = 0; i < 3; ++i)
    d[i] = m.flaeche * (0.5 * m.kette_n_m * fmax_(e[0], 0.0) * B.a[i] + 0.5 * m.dehnung_n_m * fmax_(e[1], 0.0) * B.b[i] + 2.0 * m.schub_n_m * e[2] * B.c[i]);
  kraefte_aus(m, p, d, out);
}
// kontinuum.rs:584-591
V3 geo_w_aus(const Membran& m, V3 f1, V3 f2) {
  V3 e = dehnung_aus(m, f1, f2);
  V3 d = dU_dg_aus(m, V3{fmax_(e[0], 0.0), fmax_(e[1], 0.0), e[2]});
  return psd2(d[0], d[1], d[2]);
}
// kontinuum.rs:608-612
Hessen hessen(const Membran& m, const V3 p[3]) {
  Hessen h; f_(m, p, &h.f1, &h.f2);
  h.w = geo_w_aus(m, h.f1, h.f2);
  return h;
}
// kontinuum.rs:616-664
void mal_h_aus(const Membran& m, const Hessen& h, const V3 delta[3], V3 out[3]) {
  V3 f1 = h.f1, f2 = h.f2;
  V3 c1 = c1_(m), c2 = c2_(m);
  V3 q{};
  for (int l = 0; l < 3; ++l) {
    double d1 = dot(f1, delta[l]), d2 = dot(f2, delta[l]);
    q[0] += 2.0 * c1[l] * d1;
    q[1] += 2.0 * c2[l] * d2;
    q[2] += 2.0 * (c2[l] * d1 + c1[l] * d2);
  }
  V3 w = kg_mal(m, q);
  for (int k = 0; k < 3; ++k) {
    double s1 = 2.0 * (c1[k] * w[0] + c2[k] * w[2]);
    double s2 = 2.0 * (c2[k] * w[1] + c1[k] * w[2]);
    out[k] = {s1 * f1[0] + s2 * f2[0], s1 * f1[1] + s2 * f2[1], s1 * f1[2] + s2 * f2[2]};
  }
  double w11 = h.w[0], w22 = h.w[1], w12 = h.w[2];
  V3 t1{}, t2{};
  for (int b = 0; b < 3; ++b)
    for (int c = 0; c < 3; ++c) { t1[c] += c1[b] * delta[b][c]; t2[c] += c2[b] * delta[b][c]; }
  for (int a = 0; a < 3; ++a) {
    double pa = 2.0 * (w11 * c1[a] + w12 * c2[a]);
    double qa = 2.0 * (w12 * c1[a] + w22 * c2[a]);
    for (int c = 0; c < 3; ++c) out[a][c] += pa * t1[c] + qa * t2[c];
  }
}
// kontinuum.rs:675-708
void diag_bloecke_aus(const Membran& m, const Hessen& h, double out[3][3][3]) {
  double J[3][3][3]; jacobi_aus(m, h.f1, h.f2, J);
  auto kg = kg_aus_basis(m);
  std::memset(out, 0, sizeof(double) * 27);
  for (int k = 0; k < 3; ++k) {
    double tmp[3][3] = {};
    for (int i = 0; i < 3; ++i)
      for (int c = 0; c < 3; ++c)
        for (int mm = 0; mm < 3; ++mm) tmp[i][c] += kg[i][mm] * J[k][mm][c];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 3; ++i) out[k][a][c] += J[k][i][a] * tmp[i][c];
  }
  double w11 = h.w[0], w22 = h.w[1], w12 = h.w[2];
  V3 c1 = c1_(m), c2 = c2_(m);
  for (int k = 0; k < 3; ++k) {
    double s = 2.0 * (w11 * c1[k] * c1[k] + w22 * c2[k] * c2[k] + 2.0 * w12 * c1[k] * c2[k]);
    for (int c = 0; c < 3; ++c) out[k][c][c] += s;
  }
}

// ---------------------------------------------------------------------------
// Dihedral bending (cpu.rs:279-307)
// ---------------------------------------------------------------------------
double winkel_fehler(double theta, double ruhe) {
  const double TAU = 6.283185307179586;
  double d = theta - ruhe;
  return d - TAU * std::round(d / TAU);
}
bool dieder(const V3 q[4], double* theta, V3 g[4]) {
  V3 x0 = q[0], x1 = q[1], x2 = q[2], x3 = q[3];
  V3 e = sub(x1, x0);
  double le = norm(e);
  V3 n1 = cross(sub(x2, x0), sub(x2, x1));
  V3 n2 = cross(sub(x3, x1), sub(x3, x0));
  double q1 = dot(n1, n1), q2 = dot(n2, n2);
  if (le < 1e-12 || q1 < 1e-30 || q2 < 1e-30) return false;
  V3 m1 = mul(n1, 1.0 / std::sqrt(q1)), m2 = mul(n2, 1.0 / std::sqrt(q2));
  *theta = std::atan2(dot(cross(m1, m2), mul(e, 1.0 / le)), dot(m1, m2));
  g[2] = mul(n1, -le / q1);
  g[3] = mul(n2, -le / q2);
  g[0] = mul(add(mul(n1, dot(sub(x2, x1), e) / le / q1), mul(n2, dot(sub(x3, x1), e) / le / q2)), -1.0);
  g[1] = add(mul(n1, dot(sub(x2, x0), e) / le / q1), mul(n2, dot(sub(x3, x0), e) / le / q2));
  return true;
}

// ---------------------------------------------------------------------------
// Directional bending stiffness (biegung.rs:92-173)
// ---------------------------------------------------------------------------
double positiv(double d, double weft, double warp, double bias) {
  if (d > 0.0) return d;
  double kleinste = std::fabs(fmin_(fmin_(weft, warp), bias));
  return 1.0e-3 * kleinste;
}
double steifigkeit(double weft, double warp, double bias, double theta) {
  if (weft == warp && warp == bias) return weft;
  double s = std::sin(theta), c = std::cos(theta);
  double s2 = s * s, c2 = c * c;
  return positiv(warp * (c2 * c2) + weft * (s2 * s2) + (4.0 * bias - warp - weft) * (c2 * s2), weft, warp, bias);
}
double steifigkeit_aus(double c2, double s2, double weft, double warp, double bias) {
  if (weft == warp && warp == bias) return weft;
  return positiv(warp * (c2 * c2) + weft * (s2 * s2) + (4.0 * bias - warp - weft) * (c2 * s2), weft, warp, bias);
}
bool cos2_sin2(const double e[2], const double warp[2], double* c2, double* s2) {
  double le = std::sqrt(e[0] * e[0] + e[1] * e[1]);
  double lw = std::sqrt(warp[0] * warp[0] + warp[1] * warp[1]);
  if (!(le > 1e-12) || !(lw > 1e-12)) return false;
  double ex = e[0] / le, ey = e[1] / le, wx = warp[0] / lw, wy = warp[1] / lw;
  double c_edge = ex * wx + ey * wy;
  double s_edge = ex * wy - ey * wx;
  double c = -s_edge, s = c_edge;
  *c2 = c * c; *s2 = s * s;
  return true;
}

// ---------------------------------------------------------------------------
// 3x3 helpers (implizit.rs:429-453)
// ---------------------------------------------------------------------------
inline void m3_add_skal(M3& m, double s) { for (int k = 0; k < 3; ++k) m[k][k] += s; }
inline void m3_add_aus(M3& m, V3 a, V3 b, double s) {
  for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m[i][j] += s * a[i] * b[j];
}
inline V3 m3_mal(const M3& m, V3 v) { return {dot(m[0], v), dot(m[1], v), dot(m[2], v)}; }
M3 m3_inv(const M3& m) {
  double a = m[0][0], b = m[0][1], c = m[0][2], d = m[1][0], e = m[1][1], f = m[1][2], g = m[2][0], h = m[2][1], i = m[2][2];
  double det = a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
  double s = 1.0 / det;
  M3 r;
  r[0] = {(e * i - f * h) * s, (c * h - b * i) * s, (b * f - c * e) * s};
  r[1] = {(f * g - d * i) * s, (a * i - c * g) * s, (c * d - a * f) * s};
  r[2] = {(d * h - e * g) * s, (b * g - a * h) * s, (a * e - b * d) * s};
  return r;
}

// ---------------------------------------------------------------------------
// Bodies (koerper.rs): the cylinder, the triangle-mesh body (netzkoerper.rs)
// and the sampled grid SDF
// ---------------------------------------------------------------------------
struct Kontakt { V3 normale; double tiefe; V3 geschwindigkeit; };

#include "netzkoerper.h"

bool sdf_sample(const float* werte, const int dims[3], const double origin[3], double spacing, const double q[3], double* d, double grad[3]) {
  // Cell and fraction per axis; outside the node box there is no field.
  double t[3]; int i[3];
  for (int a = 0; a < 3; ++a) {
    double u = (q[a] - origin[a]) / spacing;
    if (!(u >= 0.0) || !(u <= (double)(dims[a] - 1))) return false;
    int c = (int)std::floor(u);
    if (c > dims[a] - 2) c = dims[a] - 2;
    i[a] = c;
    t[a] = u - (double)c;
  }
  auto W = [&](int dx, int dy, int dz) -> double {
    size_t idx = ((size_t)(i[2] + dz) * (size_t)dims[1] + (size_t)(i[1] + dy)) * (size_t)dims[0] + (size_t)(i[0] + dx);
    return (double)werte[idx];
  };
  double v000 = W(0, 0, 0), v100 = W(1, 0, 0), v010 = W(0, 1, 0), v110 = W(1, 1, 0);
  double v001 = W(0, 0, 1), v101 = W(1, 0, 1), v011 = W(0, 1, 1), v111 = W(1, 1, 1);
  double tx = t[0], ty = t[1], tz = t[2];
  // Interpolate along x, then y, then z.
  double x00 = v000 + tx * (v100 - v000), x10 = v010 + tx * (v110 - v010);
  double x01 = v001 + tx * (v101 - v001), x11 = v011 + tx * (v111 - v011);
  double y0 = x00 + ty * (x10 - x00), y1 = x01 + ty * (x11 - x01);
  *d = y0 + tz * (y1 - y0);
  // Gradient of the trilinear interpolant, per unit length.
  double gx0 = (v100 - v000) + ty * ((v110 - v010) - (v100 - v000));
  double gx1 = (v101 - v001) + ty * ((v111 - v011) - (v101 - v001));
  grad[0] = (gx0 + tz * (gx1 - gx0)) / spacing;
  double gy0 = x10 - x00, gy1 = x11 - x01;
  grad[1] = (gy0 + tz * (gy1 - gy0)) / spacing;
  grad[2] = (y1 - y0) / spacing;
  return true;
}

struct Koerper {
  // 0 = Zylinder (koerper.rs:19), 1 = grid SDF, 2 = Netz (koerper.rs:25)
  int art = 0;
  V3 punkt{}, achse{}, geschwindigkeit{};
  double radius = 0;
  std::vector<float> werte;
  int dims[3] = {0, 0, 0};
  double origin[3] = {0, 0, 0};
  double spacing = 0;
  V3 versatz{};
  std::shared_ptr<const Koerpernetz> netz;  // shared like Arc<Koerpernetz>

  // koerper.rs:69-91. The grid SDF answers like the analytic cylinder: at any
  // depth inside its box, with the normalised field gradient as normal.
  bool kontakt(V3 x, double abstand, Kontakt* out) const {
    if (art == 0) {
      V3 a = mul(achse, 1.0 / norm(achse));
      V3 d = sub(x, punkt);
      V3 quer = sub(d, mul(a, dot(d, a)));
      double r = norm(quer);
      double tiefe = radius + abstand - r;
      if (tiefe <= 0.0 || r < 1e-12) return false;
      *out = {mul(quer, 1.0 / r), tiefe, geschwindigkeit};
      return true;
    }
    if (art == 2) {  // koerper.rs:82-90
      Naechster n;
      if (!netz->naechster(sub(x, versatz), std::fmax(netz->suchtiefe, abstand), &n)) return false;
      double tiefe = abstand - n.abstand;
      V3 vk = netz->geschwindigkeit_an(n.dreieck, n.punkt, geschwindigkeit);
      if (!(tiefe > 0.0)) return false;
      *out = {n.normale, tiefe, vk};
      return true;
    }
    V3 q = sub(x, versatz);
    double d, g[3];
    if (!sdf_sample(werte.data(), dims, origin, spacing, q.data(), &d, g)) return false;
    double tiefe = abstand - d;
    if (tiefe <= 0.0) return false;
    V3 gv{g[0], g[1], g[2]};
    double lg = norm(gv);
    if (!(lg > 1e-12)) return false;
    *out = {mul(gv, 1.0 / lg), tiefe, geschwindigkeit};
    return true;
  }
  // Entry of the path a -> b into the body, the counterpart of
  // Koerpernetz::durchstoss (netzkoerper.rs:541-596, used at implizit.rs:984-1003)
  // for the grid SDF that stands in for the mesh body: the first sign change
  // outside -> inside of the sampled distance, at half-spacing samples along
  // the path, refined by 60 bisections; the normal is the normalised gradient
  // there. A path that starts inside does not enter (exits do not count); a
  // cylinder has none (koerper.rs:101).
  bool durchstoss(V3 a, V3 b, V3* p, V3* n, V3* vk) const {
    if (art == 0) return false;
    if (art == 2) {  // koerper.rs:101-110
      V3 q;
      if (!netz->durchstoss(sub(a, versatz), sub(b, versatz), geschwindigkeit, &q, n, vk)) return false;
      *p = add(q, versatz);
      return true;
    }
    V3 d = sub(b, a);
    if (dot(d, d) < 1e-24) return false;
    auto abstand = [&](V3 x, double* out) {
      V3 q = sub(x, versatz);
      double g[3];
      return sdf_sample(werte.data(), dims, origin, spacing, q.data(), out, g);
    };
    double d0;
    if (abstand(a, &d0) && !(d0 >= 0.0)) return false;
    // clip the path to the grid box: outside the grid nothing is inside the
    // body, and the samples stay bounded by the grid's diagonal
    double ta = 0.0, tb = 1.0;
    for (int q = 0; q < 3; ++q) {
      const double lo = origin[q] + versatz[q], hi = lo + (double)(dims[q] - 1) * spacing;
      if (d[q] == 0.0) {
        if (!(a[q] >= lo && a[q] <= hi)) return false;
        continue;
      }
      const double u0 = (lo - a[q]) / d[q], u1 = (hi - a[q]) / d[q];
      ta = std::fmax(ta, std::fmin(u0, u1));
      tb = std::fmin(tb, std::fmax(u0, u1));
    }
    if (!(ta <= tb)) return false;
    const double l = norm(d) * (tb - ta);
    if (!std::isfinite(l)) return false;
    int schritte = (int)std::ceil(l / (0.5 * spacing));
    if (schritte < 1) schritte = 1;
    double t0 = ta;
    for (int k = 1; k <= schritte; ++k) {
      const double t1 = ta + (tb - ta) * ((double)k / (double)schritte);
      double d1;
      if (!abstand(add(a, mul(d, t1)), &d1) || d1 >= 0.0) {
        t0 = t1;
        continue;
      }
      double lo = t0, hi = t1;  // inside at hi, not inside at lo
      for (int it = 0; it < 60; ++it) {
        const double m = 0.5 * (lo + hi);
        double dm;
        if (!abstand(add(a, mul(d, m)), &dm) || dm >= 0.0) lo = m;
        else hi = m;
      }
      *p = add(a, mul(d, hi));
      V3 q = sub(*p, versatz);
      double dd, g[3];
      sdf_sample(werte.data(), dims, origin, spacing, q.data(), &dd, g);
      V3 gv{g[0], g[1], g[2]};
      const double lg = norm(gv);
      if (!(lg > 1e-12)) return false;
      *n = mul(gv, 1.0 / lg);
      *vk = geschwindigkeit;
      return true;
    }
    return false;
  }
  // koerper.rs:122-133
  Koerper verschoben(double t) const {
    Koerper k = *this;
    if (art == 0) k.punkt = add(punkt, mul(geschwindigkeit, t));
    else k.versatz = add(versatz, mul(geschwindigkeit, t));
    return k;
  }
};

// Exact ellipsoid distance (Eberly, "Distance from a Point to an Ellipse, an
// Ellipsoid, or a Hyperellipsoid", 2013, bisection form). e sorted descending,
// y in the first octant. Returns the unsigned distance; *inside set by the caller.
double ellipsoid_dist_sorted(const double e[3], const double y[3]) {
  // Point on an axis plane and the degenerate cases are handled by the general
  // bisection on F(t) = sum (e_i y_i/(t+e_i^2))^2 - 1 over t > -e_2^2.
  if (y[2] > 0.0) {
    if (y[1] > 0.0) {
      if (y[0] > 0.0) {
        double z[3] = {y[0] / e[0], y[1] / e[1], y[2] / e[2]};
        double g = z[0] * z[0] + z[1] * z[1] + z[2] * z[2] - 1.0;
        if (g != 0.0) {
          double r0 = (e[0] / e[2]) * (e[0] / e[2]), r1 = (e[1] / e[2]) * (e[1] / e[2]);
          double n0 = r0 * z[0], n1 = r1 * z[1];
          double s0 = z[2] - 1.0;
          double s1 = g < 0.0 ? 0.0 : std::sqrt(n0 * n0 + n1 * n1 + z[2] * z[2]) - 1.0;
          double s = 0.0;
          for (int it = 0; it < 1100; ++it) {
            s = 0.5 * (s0 + s1);
            if (s == s0 || s == s1) break;
            double ra = n0 / (s + r0), rb = n1 / (s + r1), rc = z[2] / (s + 1.0);
            double gg = ra * ra + rb * rb + rc * rc - 1.0;
            if (gg > 0.0) s0 = s; else if (gg < 0.0) s1 = s; else break;
          }
          double x0 = r0 * y[0] / (s + r0), x1 = r1 * y[1] / (s + r1), x2 = y[2] / (s + 1.0);
          return std::sqrt((x0 - y[0]) * (x0 - y[0]) + (x1 - y[1]) * (x1 - y[1]) + (x2 - y[2]) * (x2 - y[2]));
        }
        return 0.0;
      }
      // y0 == 0: ellipse in the (1,2) plane
      double ee[2] = {e[1], e[2]}, yy[2] = {y[1], y[2]};
      double z0 = yy[0] / ee[0], z1 = yy[1] / ee[1];
      double g = z0 * z0 + z1 * z1 - 1.0;
      if (g == 0.0) return 0.0;
      double r0 = (ee[0] / ee[1]) * (ee[0] / ee[1]);
      double n0 = r0 * z0, s0 = z1 - 1.0, s1 = g < 0.0 ? 0.0 : std::sqrt(n0 * n0 + z1 * z1) - 1.0, s = 0.0;
      for (int it = 0; it < 1100; ++it) {
        s = 0.5 * (s0 + s1);
        if (s == s0 || s == s1) break;
        double ra = n0 / (s + r0), rb = z1 / (s + 1.0);
        double gg = ra * ra + rb * rb - 1.0;
        if (gg > 0.0) s0 = s; else if (gg < 0.0) s1 = s; else break;
      }
      double x0 = r0 * yy[0] / (s + r0), x1 = yy[1] / (s + 1.0);
      return std::sqrt((x0 - yy[0]) * (x0 - yy[0]) + (x1 - yy[1]) * (x1 - yy[1]) + y[0] * y[0]);
    }
    if (y[0] > 0.0) {
      double ee[2] = {e[0], e[2]}, yy[2] = {y[0], y[2]};
      double z0 = yy[0] / ee[0], z1 = yy[1] / ee[1];
      double g = z0 * z0 + z1 * z1 - 1.0;
      if (g == 0.0) return 0.0;
      double r0 = (ee[0] / ee[1]) * (ee[0] / ee[1]);
      double n0 = r0 * z0, s0 = z1 - 1.0, s1 = g < 0.0 ? 0.0 : std::sqrt(n0 * n0 + z1 * z1) - 1.0, s = 0.0;
      for (int it = 0; it < 1100; ++it) {
        s = 0.5 * (s0 + s1);
        if (s == s0 || s == s1) break;
        double ra = n0 / (s + r0), rb = z1 / (s + 1.0);
        double gg = ra * ra + rb * rb - 1.0;
        if (gg > 0.0) s0 = s; else if (gg < 0.0) s1 = s; else break;
      }
      double x0 = r0 * yy[0] / (s + r0), x1 = yy[1] / (s + 1.0);
      return std::sqrt((x0 - yy[0]) * (x0 - yy[0]) + (x1 - yy[1]) * (x1 - yy[1]) + y[1] * y[1]);
    }
    return std::fabs(y[2] - e[2]);
  }
  // y2 == 0: closest point may be on the rim or interior of the (0,1) ellipse
  double d0 = e[0] * e[0] - e[2] * e[2], d1 = e[1] * e[1] - e[2] * e[2];
  double n0 = e[0] * y[0], n1 = e[1] * y[1];
  bool computed = false;
  double dist = 0.0;
  if (n0 < d0 && n1 < d1) {
    double q0 = n0 / d0, q1 = n1 / d1;
    double den = 1.0 - q0 * q0 - q1 * q1;
    if (den > 0.0) {
      double x0 = e[0] * q0, x1 = e[1] * q1, x2 = e[2] * std::sqrt(den);
      dist = std::sqrt((x0 - y[0]) * (x0 - y[0]) + (x1 - y[1]) * (x1 - y[1]) + x2 * x2);
      computed = true;
    }
  }
  if (!computed) {
    // distance to the ellipse with axes e0, e1 in the plane
    double ee[2] = {e[0], e[1]}, yy[2] = {y[0], y[1]};
    if (yy[1] > 0.0 && yy[0] > 0.0) {
      double z0 = yy[0] / ee[0], z1 = yy[1] / ee[1];
      double g = z0 * z0 + z1 * z1 - 1.0;
      if (g == 0.0) return 0.0;
      double r0 = (ee[0] / ee[1]) * (ee[0] / ee[1]);
      double m0 = r0 * z0, s0 = z1 - 1.0, s1 = g < 0.0 ? 0.0 : std::sqrt(m0 * m0 + z1 * z1) - 1.0, s = 0.0;
      for (int it = 0; it < 1100; ++it) {
        s = 0.5 * (s0 + s1);
        if (s == s0 || s == s1) break;
        double ra = m0 / (s + r0), rb = z1 / (s + 1.0);
        double gg = ra * ra + rb * rb - 1.0;
        if (gg > 0.0) s0 = s; else if (gg < 0.0) s1 = s; else break;
      }
      double x0 = r0 * yy[0] / (s + r0), x1 = yy[1] / (s + 1.0);
      dist = std::sqrt((x0 - yy[0]) * (x0 - yy[0]) + (x1 - yy[1]) * (x1 - yy[1]));
    } else if (yy[1] > 0.0) {
      dist = std::fabs(yy[1] - ee[1]);
    } else {
      double nn0 = ee[0] * yy[0], dd0 = ee[0] * ee[0] - ee[1] * ee[1];
      if (nn0 < dd0) {
        double xd0 = nn0 / dd0;
        double x0 = ee[0] * xd0, x1 = ee[1] * std::sqrt(1.0 - xd0 * xd0);
        dist = std::sqrt((x0 - yy[0]) * (x0 - yy[0]) + x1 * x1);
      } else {
        dist = std::fabs(yy[0] - ee[0]);
      }
    }
  }
  return dist;
}

double ellipsoid_sdf(const double c[3], const double r[3], const double x[3]) {
  double y[3] = {std::fabs(x[0] - c[0]), std::fabs(x[1] - c[1]), std::fabs(x[2] - c[2])};
  // Sort the axes descending, carrying the coordinates along.
  int ord[3] = {0, 1, 2};
  for (int i = 0; i < 3; ++i)
    for (int j = i + 1; j < 3; ++j)
      if (r[ord[j]] > r[ord[i]]) { int t = ord[i]; ord[i] = ord[j]; ord[j] = t; }
  double e[3] = {r[ord[0]], r[ord[1]], r[ord[2]]};
  double yy[3] = {y[ord[0]], y[ord[1]], y[ord[2]]};
  double dist = ellipsoid_dist_sorted(e, yy);
  double g = (yy[0] / e[0]) * (yy[0] / e[0]) + (yy[1] / e[1]) * (yy[1] / e[1]) + (yy[2] / e[2]) * (yy[2] / e[2]);
  return g < 1.0 ? -dist : dist;
}

double capped_cylinder_sdf(const double c[2], double rad, double z0, double z1, const double x[3]) {
  double dr = std::sqrt((x[0] - c[0]) * (x[0] - c[0]) + (x[1] - c[1]) * (x[1] - c[1])) - rad;
  double dz = fmax_(z0 - x[2], x[2] - z1);
  double a = fmax_(dr, 0.0), b = fmax_(dz, 0.0);
  return fmin_(fmax_(dr, dz), 0.0) + std::sqrt(a * a + b * b);
}

// ---------------------------------------------------------------------------
// Cloth self-collision (selbstkollision.rs)
// ---------------------------------------------------------------------------
#include "selbstkollision.h"

// ---------------------------------------------------------------------------
// The system of one substep (implizit.rs:78-280) and the solver (283-427)
// ---------------------------------------------------------------------------
struct Kante { uint32_t i, j; double k; V3 richtung; double quer; double ruhe; };  // implizit.rs:79-89
struct Biegung { uint32_t punkte[4]; V3 grad[4]; double k; double ruhe; };      // implizit.rs:94-100
struct MembranElement { uint32_t punkte[3]; Membran element; };               // kontinuum.rs:96-99
struct Fest { bool da = false; V3 normale{}; double vn_soll = 0; V3 v_koerper{}; };  // implizit.rs:594-599

struct Szene {
  std::vector<V3> x, v;
  std::vector<double> masse, w, kollisionsabstand, reibung;
  std::vector<std::array<uint32_t, 2>> abstand_punkte;
  std::vector<double> abstand_ruhe, abstand_nachgiebigkeit;
  std::vector<uint8_t> membran_kante;
  std::vector<std::array<uint32_t, 2>> naht_punkte;
  std::vector<double> naht_ruhe;
  std::vector<std::array<uint32_t, 4>> biegung_punkte;
  std::vector<double> biegung_ruhe, biegung_nachgiebigkeit;
  std::vector<MembranElement> membran_elemente;
  std::vector<std::array<uint32_t, 3>> dreiecke;
  std::vector<uint32_t> teil_von_punkt;
  // self-collision inputs (selbstkollision.rs:623-693): stretch stiffness per
  // point and per triangle, rest area per triangle
  std::vector<double> selbst_e_punkt, selbst_e_dreieck, selbst_a_dreieck;
  std::vector<Koerper> koerper;
  double bild_dt = 1.0 / 60.0;
  V3 schwerkraft{0, 0, -9.81};
  double daempfung = 0;
  double zeit = 0;
};

struct System {  // implizit.rs:123-172
  double h = 0, c = 0;
  int cg_max = 0;
  double cg_toleranz = 0;
  const Szene* s = nullptr;
  std::vector<Kante> kanten, naehte;
  std::vector<Biegung> biegungen;
  std::vector<Hessen> membran_h;
  std::vector<Feder> federn;
  std::vector<V3> rhs, f, last;
  std::vector<M3> pinv;

  size_t n() const { return s->masse.size(); }

  // implizit.rs:185-221
  void mal_h(const V3* p, V3* out) const {
    size_t n = this->n();
    for (size_t i = 0; i < n; ++i) out[i] = {0.0, 0.0, 0.0};
    auto kante = [&](const Kante& k) {
      if (k.k == 0.0) return;
      size_t i = k.i, j = k.j;
      V3 d = sub(p[i], p[j]);
      V3 bd = mul(add(mul(k.richtung, dot(k.richtung, d) * (1.0 - k.quer)), mul(d, k.quer)), k.k);
      out[i] = add(out[i], bd);
      out[j] = sub(out[j], bd);
    };
    for (const auto& k : kanten) kante(k);
    for (const auto& k : naehte) kante(k);
    for (const auto& b : biegungen) {
      double gp = 0.0;
      // Rust `(0..4).map(..).sum()` folds from 0.0, left to right.
      for (int m = 0; m < 4; ++m) gp += dot(b.grad[m], p[b.punkte[m]]);
      for (int m = 0; m < 4; ++m) out[b.punkte[m]] = add(out[b.punkte[m]], mul(b.grad[m], b.k * gp));
    }
    const auto& me = s->membran_elemente;
    for (size_t t = 0; t < me.size(); ++t) {
      V3 d3[3] = {p[me[t].punkte[0]], p[me[t].punkte[1]], p[me[t].punkte[2]]};
      V3 hu[3];
      mal_h_aus(me[t].element, membran_h[t], d3, hu);
      for (int k = 0; k < 3; ++k) out[me[t].punkte[k]] = add(out[me[t].punkte[k]], hu[k]);
    }
    for (const auto& fe : federn)
      feder_paare(fe, [&](uint32_t q, uint32_t r, double om) {
        V3 t = mul(m3_mal(fe.block, sub(p[q], p[r])), om);
        out[q] = add(out[q], t);
        out[r] = sub(out[r], t);
      });
  }
  // implizit.rs:224-229
  void mal_a(const V3* p, V3* out) const {
    mal_h(p, out);
    for (size_t i = 0; i < n(); ++i) out[i] = add(mul(out[i], h * h), mul(p[i], s->masse[i] * (1.0 + h * c)));
  }
  // implizit.rs:232-279
  std::vector<M3> diag() const {
    double h2 = h * h;
    std::vector<M3> dg(n());
    for (size_t i = 0; i < n(); ++i) {
      M3 m{};
      m3_add_skal(m, s->masse[i] * (1.0 + h * c));
      dg[i] = m;
    }
    auto kante = [&](const Kante& k) {
      if (k.k == 0.0) return;
      for (uint32_t q : {k.i, k.j}) {
        m3_add_skal(dg[q], k.k * h2 * k.quer);
        m3_add_aus(dg[q], k.richtung, k.richtung, k.k * h2 * (1.0 - k.quer));
      }
    };
    for (const auto& k : kanten) kante(k);
    for (const auto& k : naehte) kante(k);
    for (const auto& b : biegungen)
      for (int m = 0; m < 4; ++m) m3_add_aus(dg[b.punkte[m]], b.grad[m], b.grad[m], b.k * h2);
    const auto& me = s->membran_elemente;
    for (size_t t = 0; t < me.size(); ++t) {
      double bl[3][3][3];
      diag_bloecke_aus(me[t].element, membran_h[t], bl);
      for (int k = 0; k < 3; ++k)
        for (int a = 0; a < 3; ++a)
          for (int b = 0; b < 3; ++b) dg[me[t].punkte[k]][a][b] += h2 * bl[k][a][b];
    }
    for (const auto& fe : federn)
      feder_paare(fe, [&](uint32_t q, uint32_t r, double om) {
        for (uint32_t pp : {q, r})
          for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) dg[pp][a][b] += h2 * om * fe.block[a][b];
      });
    return dg;
  }
};

// implizit.rs:283-288
inline V3 filtern(const Fest& fe, V3 v) { return fe.da ? sub(v, mul(fe.normale, dot(v, fe.normale))) : v; }

struct CgLauf { int iterationen = 0; bool nicht_konvergiert = false; double rest = 0; };  // implizit.rs:294-299

// implizit.rs:370-426
CgLauf loese(const System& sys, const std::vector<Fest>& fest, std::vector<V3>& dv) {
  size_t n = sys.n();
  const auto& w = sys.s->w;
  const auto& b = sys.rhs;
  const auto& pinv = sys.pinv;
  for (size_t i = 0; i < n; ++i) {
    if (w[i] <= 0.0) { dv[i] = {0.0, 0.0, 0.0}; continue; }
    if (fest[i].da) {
      double soll = fest[i].vn_soll - dot(sys.s->v[i], fest[i].normale);
      dv[i] = add(filtern(fest[i], dv[i]), mul(fest[i].normale, soll));
    }
  }
  std::vector<V3> ap(n);
  sys.mal_a(dv.data(), ap.data());
  std::vector<V3> r(n), z(n), cdir(n);
  for (size_t i = 0; i < n; ++i) r[i] = w[i] > 0.0 ? filtern(fest[i], sub(b[i], ap[i])) : V3{0, 0, 0};
  double bb = 0.0;
  for (size_t i = 0; i < n; ++i) {
    double t = 0.0;
    if (w[i] > 0.0) { V3 fb = filtern(fest[i], b[i]); t = dot(fb, m3_mal(pinv[i], fb)); }
    bb += t;
  }
  for (size_t i = 0; i < n; ++i) z[i] = m3_mal(pinv[i], r[i]);
  for (size_t i = 0; i < n; ++i) cdir[i] = filtern(fest[i], z[i]);
  double delta = 0.0;
  for (size_t i = 0; i < n; ++i) delta += dot(r[i], cdir[i]);
  double msum = 0.0;
  for (size_t i = 0; i < n; ++i) msum += sys.s->masse[i];
  double boden = msum * 1e-18;
  double ziel = fmax_(sys.cg_toleranz * sys.cg_toleranz * fmax_(bb, delta), boden);
  int it = 0;
  while (delta > ziel && it < sys.cg_max) {
    sys.mal_a(cdir.data(), ap.data());
    for (size_t i = 0; i < n; ++i) ap[i] = w[i] > 0.0 ? filtern(fest[i], ap[i]) : V3{0, 0, 0};
    double ca = 0.0;
    for (size_t i = 0; i < n; ++i) ca += dot(cdir[i], ap[i]);
    if (ca <= std::numeric_limits<double>::min() || !std::isfinite(ca)) break;
    double alpha = delta / ca;
    for (size_t i = 0; i < n; ++i) {
      dv[i] = add(dv[i], mul(cdir[i], alpha));
      r[i] = sub(r[i], mul(ap[i], alpha));
    }
    for (size_t i = 0; i < n; ++i) z[i] = m3_mal(pinv[i], r[i]);
    double neu = 0.0;
    for (size_t i = 0; i < n; ++i) neu += dot(r[i], z[i]);
    double beta = neu / delta;
    delta = neu;
    for (size_t i = 0; i < n; ++i) cdir[i] = w[i] > 0.0 ? filtern(fest[i], add(z[i], mul(cdir[i], beta))) : V3{0, 0, 0};
    it += 1;
  }
  CgLauf l;
  l.iterationen = it;
  l.nicht_konvergiert = delta > ziel;
  l.rest = std::sqrt(delta / fmax_(bb, boden));
  return l;
}

}  // namespace

// ---------------------------------------------------------------------------
// The implicit path (implizit.rs:569-1057)
// ---------------------------------------------------------------------------
struct or_sim {
  Szene s;
  or_param param;
  or_statistik stat{};
  std::vector<V3> dv_vorher;
  System sys;
  std::vector<Fest> fest;
  // self-collision state (implizit.rs:580-584): mesh and material, built on
  // first use; springs of the step; cuts of the step; steps in a row with a cut
  bool hat_selbst = false;
  Stoffnetz netz;
  Stoffwerte werte;
  std::vector<Feder> federn;
  std::vector<Schnitt> schnitte;
  std::vector<uint32_t> folge;

  // selbstkollision.rs:623-693 (no folded allowance bands in these scenes:
  // band_von stays empty)
  void selbst_bauen() {
    size_t n = s.x.size();
    std::vector<double> r(n);
    for (size_t i = 0; i < n; ++i) r[i] = s.kollisionsabstand[i] + param.selbst_zusatzdicke / 2.0;
    netz = Stoffnetz::neu(s.dreiecke, r);
    netz.naht_partner.assign(n, {});
    for (const auto& nt : s.naht_punkte) {
      netz.naht_partner[nt[0]].push_back(nt[1]);
      netz.naht_partner[nt[1]].push_back(nt[0]);
    }
    for (size_t e = 0; e < s.abstand_punkte.size(); ++e) {
      uint32_t i = s.abstand_punkte[e][0], j = s.abstand_punkte[e][1];
      if (s.abstand_ruhe[e] != 0.0 || i == j) continue;
      auto& a = netz.naht_partner[i];
      if (std::find(a.begin(), a.end(), j) == a.end()) a.push_back(j);
      auto& b = netz.naht_partner[j];
      if (std::find(b.begin(), b.end(), i) == b.end()) b.push_back(i);
    }
    netz.teil_von = s.teil_von_punkt;
    netz.band_von.assign(n, KEIN);
    netz.dicke = s.kollisionsabstand;
    size_t nt = s.dreiecke.size();
    werte = Stoffwerte();
    werte.e_punkt = s.selbst_e_punkt;
    werte.e_punkt.resize(n, 0.0);
    werte.e_dreieck = s.selbst_e_dreieck;
    werte.e_dreieck.resize(nt, 0.0);
    werte.a_dreieck = s.selbst_a_dreieck;
    werte.a_dreieck.resize(nt, 0.0);
    werte.a_punkt.assign(n, 0.0);
    for (size_t t = 0; t < nt; ++t)
      for (uint32_t i : s.dreiecke[t]) werte.a_punkt[i] += werte.a_dreieck[t] / 3.0;
    for (const auto& kd : netz.kanten_dreiecke) {
      double ka = 0.0, ke = 0.0;
      int cnt = 0;
      for (uint32_t t : kd)
        if (t != KEIN) { ka += werte.a_dreieck[t]; ke += werte.e_dreieck[t]; ++cnt; }
      double k = (double)std::max(cnt, 1);
      werte.a_kante.push_back(ka / k);
      werte.e_kante.push_back(ke / k);
    }
    werte.mu_punkt = s.reibung;
    werte.naht_partner.assign(n, {});
    for (size_t c = 0; c < s.naht_punkte.size(); ++c) {
      uint32_t i = s.naht_punkte[c][0], j = s.naht_punkte[c][1];
      werte.naht_partner[i].push_back({j, (uint32_t)c});
      werte.naht_partner[j].push_back({i, (uint32_t)c});
    }
    hat_selbst = true;
  }

  // implizit.rs:668-707: proximity springs (damping the approach in v),
  // cuts, regions and push springs
  void selbst_federn() {
    federn.clear();
    schnitte.clear();
    if (!param.selbstkollision) return;
    if (!hat_selbst || netz.r_punkt.size() != s.x.size()) selbst_bauen();
    std::vector<Naehe> naehe = finde(netz, s.x, s.w, true, true);
    for (const auto& p : naehe) {
      if (p.art == PUNKT_DREIECK) stat.naehe_punkt_dreieck += 1;
      else stat.naehe_kante_kante += 1;
    }
    federn = ::federn(netz, werte, naehe, param.selbst_steifigkeit, s.masse, s.w, s.naht_ruhe, s.v);
    std::vector<Schnitt> sn = finde_schnitte(netz, s.x, param.selbst_schnitte_an_naehten != 0);
    std::vector<size_t> wurzeln = gebiete(netz, sn);
    stat.durchdringungen += (int64_t)sn.size();
    for (size_t i = 0; i < wurzeln.size(); ++i) stat.durchdringungsgebiete += wurzeln[i] == i ? 1 : 0;
    double rand = param.kontakt_rand;
    auto normale = [&](size_t i, V3* out) {
      for (const auto& k : s.koerper) {
        Kontakt kt;
        if (k.kontakt(s.x[i], s.kollisionsabstand[i] + rand, &kt)) { *out = kt.normale; return true; }
      }
      return false;
    };
    auto eingeklemmt = [&](const std::array<uint32_t, 2>& kp, const std::array<uint32_t, 3>& dp) {
      V3 summe{0.0, 0.0, 0.0};
      for (uint32_t i : dp) { V3 nn; if (normale(i, &nn)) summe = add(summe, nn); }
      if (!(norm(summe) > 0.0)) return false;
      for (uint32_t i : kp) { V3 nk; if (normale(i, &nk) && dot(nk, summe) < 0.0) return true; }
      return false;
    };
    std::vector<Feder> schub = schubfedern(netz, werte, sn, wurzeln, param.selbst_steifigkeit, folge, eingeklemmt);
    federn.insert(federn.end(), schub.begin(), schub.end());
    schnitte = std::move(sn);
  }

  // implizit.rs:663-896 (glue zones and self-collision springs are not restated here)
  void system_bauen(double h) {
    size_t n = s.x.size();
    V3 g = s.schwerkraft;
    double c = s.daempfung;
    selbst_federn();
    std::vector<V3> f(n), last(n);
    for (size_t i = 0; i < n; ++i) f[i] = sub(mul(g, s.masse[i]), mul(s.v[i], c * s.masse[i]));
    for (size_t i = 0; i < n; ++i) last[i] = mul(g, s.masse[i]);
    std::vector<Kante> kanten;
    kanten.reserve(s.abstand_punkte.size());
    for (size_t e = 0; e < s.abstand_punkte.size(); ++e) {
      size_t i = s.abstand_punkte[e][0], j = s.abstand_punkte[e][1];
      double k = (e < s.membran_kante.size() && s.membran_kante[e]) ? 0.0 : 1.0 / s.abstand_nachgiebigkeit[e];
      double l0 = s.abstand_ruhe[e];
      V3 d = sub(s.x[j], s.x[i]);
      double l = fmax_(norm(d), 1e-12);
      V3 nr = mul(d, 1.0 / l);
      V3 fi = mul(nr, k * (l - l0));
      f[i] = add(f[i], fi);
      f[j] = sub(f[j], fi);
      if (l > l0) {
        last[i] = add(last[i], fi);
        last[j] = sub(last[j], fi);
      }
      kanten.push_back({(uint32_t)i, (uint32_t)j, k, nr, fmax_(1.0 - l0 / l, 0.0), l0});
    }
    std::vector<Kante> naehte;
    for (size_t e = 0; e < s.naht_punkte.size(); ++e) {
      size_t i = s.naht_punkte[e][0], j = s.naht_punkte[e][1];
      double l0 = s.naht_ruhe[e];
      V3 d = sub(s.x[j], s.x[i]);
      double l = norm(d);
      l0 = std::isfinite(l0) ? l0 : l;
      V3 fi = (l0 > 0.0 && l > 1e-12) ? mul(d, param.naht_k * (l - l0) / l) : mul(d, param.naht_k);
      f[i] = add(f[i], fi);
      f[j] = sub(f[j], fi);
      V3 richtung{0, 0, 0};
      double quer = 1.0;
      if (l0 > 0.0 && l > 1e-12) { richtung = mul(d, 1.0 / l); quer = fmax_(1.0 - l0 / l, 0.0); }
      naehte.push_back({(uint32_t)i, (uint32_t)j, param.naht_k, richtung, quer, l0});
    }
    std::vector<Biegung> biegungen;
    for (size_t b = 0; b < s.biegung_punkte.size(); ++b) {
      const auto& p = s.biegung_punkte[b];
      double k = 1.0 / s.biegung_nachgiebigkeit[b];
      double ruhe = s.biegung_ruhe[b];
      V3 q[4] = {s.x[p[0]], s.x[p[1]], s.x[p[2]], s.x[p[3]]};
      double theta = 0.0;
      V3 gr[4] = {};
      if (!dieder(q, &theta, gr)) { theta = 0.0; for (auto& gg : gr) gg = {0, 0, 0}; }
      double fehler = winkel_fehler(theta, ruhe);
      for (int m = 0; m < 4; ++m) f[p[m]] = sub(f[p[m]], mul(gr[m], k * fehler));
      Biegung bg;
      for (int m = 0; m < 4; ++m) { bg.punkte[m] = p[m]; bg.grad[m] = gr[m]; }
      bg.k = k;
      bg.ruhe = ruhe;
      biegungen.push_back(bg);
    }
    for (const auto& fe : federn)
      for (int q = 0; q < 5; ++q) {
        uint32_t i = fe.punkte[q];
        f[i] = add(f[i], mul(fe.kraft, fe.gewichte[q]));
      }
    for (const auto& m : s.membran_elemente) {
      V3 p3[3] = {s.x[m.punkte[0]], s.x[m.punkte[1]], s.x[m.punkte[2]]};
      V3 fe[3];
      krafte(m.element, p3, fe);
      for (int k = 0; k < 3; ++k) f[m.punkte[k]] = add(f[m.punkte[k]], fe[k]);
      V3 fz[3];
      krafte_zug(m.element, p3, fz);
      for (int k = 0; k < 3; ++k) last[m.punkte[k]] = add(last[m.punkte[k]], fz[k]);
    }
    std::vector<Hessen> mh;
    mh.reserve(s.membran_elemente.size());
    for (const auto& m : s.membran_elemente) {
      V3 p3[3] = {s.x[m.punkte[0]], s.x[m.punkte[1]], s.x[m.punkte[2]]};
      mh.push_back(hessen(m.element, p3));
    }
    sys = System();
    sys.h = h; sys.c = c; sys.cg_max = param.cg_max; sys.cg_toleranz = param.cg_toleranz; sys.s = &s;
    sys.kanten = std::move(kanten); sys.naehte = std::move(naehte); sys.biegungen = std::move(biegungen);
    sys.membran_h = std::move(mh);
    sys.federn = federn;
    fest.assign(n, Fest());
    for (size_t i = 0; i < n; ++i) {
      if (s.w[i] <= 0.0) continue;
      for (const auto& k : s.koerper) {
        Kontakt kt;
        if (k.kontakt(s.x[i], s.kollisionsabstand[i] + param.kontakt_rand, &kt)) {
          double tiefe = kt.tiefe - param.kontakt_rand;
          fest[i].da = true;
          fest[i].normale = kt.normale;
          fest[i].vn_soll = tiefe / h + dot(kt.geschwindigkeit, kt.normale);
          fest[i].v_koerper = kt.geschwindigkeit;
        }
      }
    }
    std::vector<V3> hv(n);
    sys.mal_h(s.v.data(), hv.data());
    sys.rhs.resize(n);
    for (size_t i = 0; i < n; ++i) sys.rhs[i] = mul(sub(f[i], mul(hv[i], h)), h);
    sys.f = std::move(f);
    sys.last = std::move(last);
    auto dg = sys.diag();
    sys.pinv.resize(n);
    for (size_t i = 0; i < n; ++i) sys.pinv[i] = m3_inv(dg[i]);
  }

  // implizit.rs:899-1038 (no mesh body, so no durchstoss and no body-point pass)
  void schritt(double h) {
    size_t n = s.x.size();
    V3 g = s.schwerkraft;
    std::vector<V3> x_anfang;
    if (param.selbstkollision) x_anfang = s.x;
    system_bauen(h);
    std::vector<V3> dv;
    if (dv_vorher.size() == n) dv = dv_vorher;
    else dv.assign(n, mul(g, h));
    std::vector<double> fn_kraft(n, 0.0);
    for (int durchgang = 0; durchgang < 2; ++durchgang) {
      CgLauf lauf = loese(sys, fest, dv);
      stat.loesungen += 1;
      stat.cg_iterationen += lauf.iterationen;
      stat.rest_max = fmax_(stat.rest_max, lauf.rest);
      if (lauf.nicht_konvergiert) stat.nicht_konvergiert += 1;
      bool irgendeiner = false;
      for (size_t i = 0; i < n; ++i) irgendeiner = irgendeiner || fest[i].da;
      if (!irgendeiner) break;
      std::vector<V3> rest(n);
      sys.mal_a(dv.data(), rest.data());
      for (size_t i = 0; i < n; ++i) rest[i] = sub(rest[i], sys.rhs[i]);
      bool frei = false;
      for (size_t i = 0; i < n; ++i) {
        if (fest[i].da) {
          double fn_i = dot(rest[i], fest[i].normale) / h;
          if (fn_i < 0.0) { fest[i].da = false; frei = true; stat.freigegeben += 1; }
          else fn_kraft[i] = fn_i;
        }
      }
      if (!frei) break;
    }
    dv_vorher = dv;
    const std::vector<V3>& last = sys.last;
    for (size_t i = 0; i < n; ++i) {
      if (s.w[i] <= 0.0) continue;
      V3 v = add(s.v[i], dv[i]);
      if (fest[i].da) {
        stat.kontakte += 1;
        V3 rel = sub(v, fest[i].v_koerper);
        V3 t = sub(rel, mul(fest[i].normale, dot(rel, fest[i].normale)));
        double lt = norm(t);
        double normal = param.normalkraft_aus_last ? fmax_(-dot(last[i], fest[i].normale), 0.0) : fn_kraft[i];
        double grenze = s.reibung[i] * normal * h * s.w[i];
        if (lt <= grenze) { stat.haftend += 1; v = sub(v, t); }
        else if (lt > 0.0) v = sub(v, mul(t, grenze / lt));
      }
      s.v[i] = v;
      V3 alt = s.x[i];
      s.x[i] = add(alt, mul(v, h));
      // implizit.rs:984-1003: a path through the skin is set onto the entry;
      // the impact takes the normal velocity relative to the body, Coulomb up
      // to mu |vn| from the tangential one
      if (param.durchstoss_pruefen)
        for (const auto& k : s.koerper) {
          V3 p, nk, vk;
          if (k.durchstoss(alt, s.x[i], &p, &nk, &vk)) {
            stat.durchstoesse += 1;
            s.x[i] = add(p, mul(nk, s.kollisionsabstand[i]));
            V3 rel = sub(s.v[i], vk);
            double vn = dot(rel, nk);
            if (vn < 0.0) {
              V3 t = sub(rel, mul(nk, vn));
              double lt = norm(t);
              double grenze = s.reibung[i] * -vn;
              t = lt <= grenze ? V3{0.0, 0.0, 0.0} : mul(t, 1.0 - grenze / lt);
              s.v[i] = add(vk, t);
            }
          }
        }
      for (const auto& k : s.koerper) {
        Kontakt kt;
        if (k.kontakt(s.x[i], s.kollisionsabstand[i], &kt)) {
          s.x[i] = add(s.x[i], mul(kt.normale, kt.tiefe));
          double vn = dot(sub(s.v[i], kt.geschwindigkeit), kt.normale);
          if (vn < 0.0) s.v[i] = sub(s.v[i], mul(kt.normale, vn));
        }
      }
    }
    // implizit.rs:1017-1026: body points against cloth triangles, passes until one corrects nothing
    for (int d = 0; d < param.koerperpunkt_durchgaenge; ++d) {
      size_t zahl = 0;
      for (const auto& k : s.koerper)
        if (k.art == 2) zahl += k.netz->punkte_gegen_dreiecke(param.koerperpunkt_tiefe, k.versatz, k.geschwindigkeit, s.dreiecke, s.w, s.x, s.v);
      stat.koerperpunkte += (int64_t)zahl;
      if (zahl == 0) break;
    }
    // implizit.rs:1030-1037: impacts on the path x_anfang -> x, held points x 1e9
    if (param.selbstkollision && hat_selbst) {
      std::vector<Stoss> st = finde_stoesse(netz, x_anfang, s.x, s.w, true, true);
      st = ohne_durchdrungene(netz, st, schnitte);
      if (!st.empty()) {
        std::vector<double> m_eff(n);
        for (size_t i = 0; i < n; ++i) m_eff[i] = fest[i].da ? s.masse[i] * 1e9 : s.masse[i];
        stat.selbst_stoesse += (int64_t)stossantwort(st, s.x, m_eff, s.w, (size_t)param.selbst_stoss_iterationen);
      }
    }
  }

  // implizit.rs:1042-1057
  void bild() {
    int n = param.schritte_je_bild > 1 ? param.schritte_je_bild : 1;
    double h = s.bild_dt / (double)n;
    stat = or_statistik{};
    std::vector<Koerper> anfang = s.koerper;
    for (int k = 0; k < n; ++k) {
      double weg = (double)(k + 1) * h;
      for (size_t b = 0; b < s.koerper.size(); ++b) s.koerper[b] = anfang[b].verschoben(weg);
      schritt(h);
    }
    s.zeit += s.bild_dt;
  }
};

// ---------------------------------------------------------------------------
// The XPBD path `CpuReferenz` (cpu.rs:76-246), the reference's second frame
// loop, with the membrane's XPBD form `zuege` (kontinuum.rs:522-551)
// ---------------------------------------------------------------------------
struct Zug { double c; V3 grad[3]; double alpha; };  // kontinuum.rs:133-140

// kontinuum.rs:526-551
static void zuege(const Membran& m, const V3 p[3], Zug out[3]) {
  V3 f1, f2;
  f_(m, p, &f1, &f2);
  V3 e = dehnung_aus(m, f1, f2);
  double J[3][3][3];
  jacobi_aus(m, f1, f2, J);
  Basis B = basis(m);
  auto grad = [&](V3 v, bool halb, V3 g[3]) {
    for (int k = 0; k < 3; ++k)
      for (int col = 0; col < 3; ++col) {
        double s = 0.0;
        for (int i = 0; i < 3; ++i) s += J[k][i][col] * v[i];
        g[k][col] = halb ? 0.5 * s : s;
      }
  };
  const double fl = m.flaeche;
  auto alpha = [&](double k) { return k > 0.0 && fl > 0.0 ? 1.0 / (k * fl) : INFINITY; };
  out[0].c = e[0];
  grad(B.a, true, out[0].grad);
  out[0].alpha = alpha(m.kette_n_m);
  out[1].c = e[1];
  grad(B.b, true, out[1].grad);
  out[1].alpha = alpha(m.dehnung_n_m);
  out[2].c = 2.0 * e[2];
  grad(B.c, false, out[2].grad);
  out[2].alpha = alpha(m.schub_n_m);
}

// cpu.rs:249-268: C = |p_i - p_j| - ruhe with p = x + delta + kappa; corrects kappa.
static void xpbd_abstand(const std::vector<V3>& x, const std::vector<V3>& delta, std::vector<V3>& k, const std::vector<double>& w,
                  std::array<uint32_t, 2> p, double ruhe, double alpha_h2) {
  size_t i = p[0], j = p[1];
  V3 d = add(add(sub(x[i], x[j]), sub(delta[i], delta[j])), sub(k[i], k[j]));
  double l = norm(d);
  double wsum = w[i] + w[j] + alpha_h2;
  if (l < 1e-12 || wsum <= 0.0) return;
  double dl = -(l - ruhe) / wsum;
  V3 nrm = mul(d, 1.0 / l);
  k[i] = add(k[i], mul(nrm, w[i] * dl));
  k[j] = sub(k[j], mul(nrm, w[j] * dl));
}

// cpu.rs:310-333: C = theta - ruhe with p = x + delta + kappa; corrects kappa.
static void xpbd_biegung(const std::vector<V3>& x, const std::vector<V3>& delta, std::vector<V3>& k, const std::vector<double>& w,
                  std::array<uint32_t, 4> p, double ruhe, double alpha_h2) {
  size_t p0 = p[0];
  V3 q[4];
  for (int a = 0; a < 4; ++a) {
    size_t i = p[a];
    q[a] = add(add(sub(x[i], x[p0]), sub(delta[i], delta[p0])), sub(k[i], k[p0]));
  }
  double theta;
  V3 g[4];
  if (!dieder(q, &theta, g)) return;
  double nenner = alpha_h2;
  for (int a = 0; a < 4; ++a) nenner += w[p[a]] * dot(g[a], g[a]);
  if (nenner <= 0.0) return;
  double dl = -winkel_fehler(theta, ruhe) / nenner;
  for (int a = 0; a < 4; ++a) k[p[a]] = add(k[p[a]], mul(g[a], w[p[a]] * dl));
}

struct Xpbd {  // cpu.rs:76-86
  std::vector<V3> delta, korr, last;
  or_xpbd_statistik stat{};

  // cpu.rs:94-223
  void unterschritt(Szene& s, double h, double naht_alpha) {
    size_t n = s.x.size();
    V3 g = s.schwerkraft;
    double daempf = fmax_(1.0 - s.daempfung * h, 0.0);
    for (size_t i = 0; i < n; ++i) last[i] = mul(g, s.masse[i]);
    for (size_t c = 0; c < s.abstand_punkte.size(); ++c) {
      if (s.membran_kante[c]) continue;
      size_t i = s.abstand_punkte[c][0], j = s.abstand_punkte[c][1];
      V3 d = add(sub(s.x[j], s.x[i]), sub(delta[j], delta[i]));
      double l = norm(d);
      double zug = fmax_(l - s.abstand_ruhe[c], 0.0) / s.abstand_nachgiebigkeit[c];
      if (l > 1e-12 && zug > 0.0) {
        V3 f_ij = mul(d, zug / l);
        last[i] = add(last[i], f_ij);
        last[j] = sub(last[j], f_ij);
      }
    }
    for (const auto& m : s.membran_elemente) {
      V3 p[3];
      for (int k = 0; k < 3; ++k) {
        size_t q = m.punkte[k];
        p[k] = add(add(s.x[q], delta[q]), korr[q]);
      }
      V3 fz[3];
      krafte_zug(m.element, p, fz);
      for (int k = 0; k < 3; ++k) last[m.punkte[k]] = add(last[m.punkte[k]], fz[k]);
    }
    for (size_t i = 0; i < n; ++i) {
      korr[i] = {0.0, 0.0, 0.0};
      if (s.w[i] > 0.0) {
        s.v[i] = mul(add(s.v[i], mul(g, h)), daempf);
        delta[i] = add(delta[i], mul(s.v[i], h));
      }
    }
    double h2 = h * h;
    for (size_t c = 0; c < s.abstand_punkte.size(); ++c) {
      if (s.membran_kante[c]) continue;
      xpbd_abstand(s.x, delta, korr, s.w, s.abstand_punkte[c], s.abstand_ruhe[c], s.abstand_nachgiebigkeit[c] / h2);
    }
    for (const auto& m : s.membran_elemente) {
      V3 p[3];
      for (int k = 0; k < 3; ++k) {
        size_t q = m.punkte[k];
        p[k] = add(add(s.x[q], delta[q]), korr[q]);
      }
      Zug z[3];
      zuege(m.element, p, z);
      for (int a = 0; a < 3; ++a) {
        if (!std::isfinite(z[a].alpha)) continue;
        double wsum = z[a].alpha / h2;
        for (int k = 0; k < 3; ++k) wsum += s.w[m.punkte[k]] * dot(z[a].grad[k], z[a].grad[k]);
        if (!(wsum > 0.0)) continue;
        double dl = -z[a].c / wsum;
        for (int k = 0; k < 3; ++k) {
          size_t q = m.punkte[k];
          korr[q] = add(korr[q], mul(z[a].grad[k], s.w[q] * dl));
        }
      }
    }
    for (size_t c = 0; c < s.naht_punkte.size(); ++c)
      xpbd_abstand(s.x, delta, korr, s.w, s.naht_punkte[c], s.naht_ruhe[c], naht_alpha / h2);
    for (size_t c = 0; c < s.biegung_punkte.size(); ++c)
      xpbd_biegung(s.x, delta, korr, s.w, s.biegung_punkte[c], s.biegung_ruhe[c], s.biegung_nachgiebigkeit[c] / h2);
    for (size_t i = 0; i < n; ++i) {
      if (s.w[i] <= 0.0) continue;
      for (const auto& k : s.koerper) {
        V3 p = add(add(s.x[i], delta[i]), korr[i]);
        Kontakt kt;
        if (!k.kontakt(p, s.kollisionsabstand[i], &kt)) continue;
        korr[i] = add(korr[i], mul(kt.normale, kt.tiefe));
        V3 dx = sub(add(mul(s.v[i], h), korr[i]), mul(kt.geschwindigkeit, h));
        V3 t = sub(dx, mul(kt.normale, dot(dx, kt.normale)));
        double lt = norm(t);
        double normallast = fmax_(-dot(last[i], kt.normale), 0.0);
        double grenze = s.reibung[i] * fmax_(kt.tiefe, normallast * h2 * s.w[i]);
        stat.kontakte += 1;
        stat.tiefe += kt.tiefe;
        stat.tangential += lt;
        if (lt <= grenze) {
          stat.haftend += 1;
          korr[i] = sub(korr[i], t);
        } else if (lt > 0.0) {
          korr[i] = sub(korr[i], mul(t, grenze / lt));
        }
      }
    }
    for (size_t i = 0; i < n; ++i)
      if (s.w[i] > 0.0) {
        delta[i] = add(delta[i], korr[i]);
        s.v[i] = add(s.v[i], mul(korr[i], 1.0 / h));
      }
  }

  // cpu.rs:226-245
  void bild(Szene& s, int unterschritte, double naht_alpha) {
    int n = unterschritte > 1 ? unterschritte : 1;
    double h = s.bild_dt / (double)n;
    size_t np = s.x.size();
    stat = or_xpbd_statistik{};
    delta.assign(np, V3{0.0, 0.0, 0.0});
    korr.assign(np, V3{0.0, 0.0, 0.0});
    last.assign(np, V3{0.0, 0.0, 0.0});
    std::vector<Koerper> anfang = s.koerper;
    for (int k = 0; k < n; ++k) {
      double weg = (double)(k + 1) * h;
      for (size_t b = 0; b < s.koerper.size(); ++b) s.koerper[b] = anfang[b].verschoben(weg);
      unterschritt(s, h, naht_alpha);
    }
    for (size_t i = 0; i < np; ++i) s.x[i] = add(s.x[i], delta[i]);
    s.zeit += s.bild_dt;
  }
};

// ---------------------------------------------------------------------------
// C API
// ---------------------------------------------------------------------------
static V3 v3(const double* p) { return {p[0], p[1], p[2]}; }
static void put(double* o, V3 v) { o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; }

extern "C" {

int or_membran_bauen(const double ruhe[6], const double kanten[3], const double dir2[2], const double material[3], double out_elem[12]) {
  double r[3][2] = {{ruhe[0], ruhe[1]}, {ruhe[2], ruhe[3]}, {ruhe[4], ruhe[5]}};
  Membran m;
  if (!membran_bauen(r, kanten, dir2, material, &m)) return 0;
  membran_in(m, out_elem);
  return 1;
}
void or_membran_dehnung(const double e[12], const double p[9], double out[3]) {
  V3 q[3] = {v3(p), v3(p + 3), v3(p + 6)}; put(out, dehnung(membran_aus(e), q));
}
void or_membran_spannung(const double e[12], const double p[9], double out[3]) {
  V3 q[3] = {v3(p), v3(p + 3), v3(p + 6)}; put(out, spannung(membran_aus(e), q));
}
double or_membran_energie(const double e[12], const double p[9]) {
  V3 q[3] = {v3(p), v3(p + 3), v3(p + 6)}; return energie(membran_aus(e), q);
}
void or_membran_krafte(const double e[12], const double p[9], double out[9]) {
  V3 q[3] = {v3(p), v3(p + 3), v3(p + 6)}, o[3]; krafte(membran_aus(e), q, o);
  for (int k = 0; k < 3; ++k) put(out + 3 * k, o[k]);
}
void or_membran_krafte_zug(const double e[12], const double p[9], double out[9]) {
  V3 q[3] = {v3(p), v3(p + 3), v3(p + 6)}, o[3]; krafte_zug(membran_aus(e), q, o);
  for (int k = 0; k < 3; ++k) put(out + 3 * k, o[k]);
}
void or_membran_hessen(const double e[12], const double p[9], double out_h[9]) {
  V3 q[3] = {v3(p), v3(p + 3), v3(p + 6)}; Hessen h = hessen(membran_aus(e), q);
  put(out_h, h.f1); put(out_h + 3, h.f2); put(out_h + 6, h.w);
}
void or_membran_mal_h_aus(const double e[12], const double hh[9], const double delta[9], double out[9]) {
  Hessen h{v3(hh), v3(hh + 3), v3(hh + 6)};
  V3 d[3] = {v3(delta), v3(delta + 3), v3(delta + 6)}, o[3];
  mal_h_aus(membran_aus(e), h, d, o);
  for (int k = 0; k < 3; ++k) put(out + 3 * k, o[k]);
}
void or_membran_diag_bloecke_aus(const double e[12], const double hh[9], double out[27]) {
  Hessen h{v3(hh), v3(hh + 3), v3(hh + 6)};
  double o[3][3][3]; diag_bloecke_aus(membran_aus(e), h, o);
  std::memcpy(out, o, sizeof(o));
}
void or_psd2(double a, double b, double c, double out[3]) { put(out, psd2(a, b, c)); }
int or_dieder(const double q[12], double* theta, double grad[12]) {
  V3 qq[4] = {v3(q), v3(q + 3), v3(q + 6), v3(q + 9)}, g[4];
  if (!dieder(qq, theta, g)) return 0;
  for (int m = 0; m < 4; ++m) put(grad + 3 * m, g[m]);
  return 1;
}
double or_winkel_fehler(double t, double r) { return winkel_fehler(t, r); }
double or_steifigkeit(double weft, double warp, double bias, double theta) { return steifigkeit(weft, warp, bias, theta); }
double or_steifigkeit_aus(double c2, double s2, double weft, double warp, double bias) { return steifigkeit_aus(c2, s2, weft, warp, bias); }
int or_cos2_sin2(const double e[2], const double warp[2], double out[2]) { return cos2_sin2(e, warp, &out[0], &out[1]) ? 1 : 0; }
void or_m3_inv(const double m[9], double out[9]) {
  M3 a{V3{m[0], m[1], m[2]}, V3{m[3], m[4], m[5]}, V3{m[6], m[7], m[8]}};
  M3 r = m3_inv(a);
  for (int i = 0; i < 3; ++i) put(out + 3 * i, r[i]);
}
double or_koerper_sdf(const double ell_c[3], const double ell_r[3], const double cyl_c[2], double cyl_r, double z0, double z1, const double x[3]) {
  double a = ellipsoid_sdf(ell_c, ell_r, x);
  if (!(cyl_r > 0.0)) return a;
  return fmin_(a, capped_cylinder_sdf(cyl_c, cyl_r, z0, z1, x));
}
int or_sdf_sample(const float* werte, const int dims[3], const double origin[3], double spacing, const double q[3], double* d, double grad[3]) {
  return sdf_sample(werte, dims, origin, spacing, q, d, grad) ? 1 : 0;
}

void or_param_default(or_param* p) {  // implizit.rs:505-531
  p->schritte_je_bild = 5; p->cg_max = 50; p->cg_toleranz = 1e-5; p->naht_k = 500.0; p->kontakt_rand = 1e-4;
  p->normalkraft_aus_last = 0; p->selbstkollision = 0; p->selbst_zusatzdicke = 2.5e-3; p->selbst_steifigkeit = 1e-3;
  p->selbst_stoss_iterationen = 1; p->selbst_schnitte_an_naehten = 0; p->durchstoss_pruefen = 1;
  p->koerperpunkt_durchgaenge = 0; p->koerperpunkt_tiefe = 0.01;
}

or_sim* or_sim_neu(const or_szene* d, const or_param* p) {
  or_sim* o = new or_sim();
  Szene& s = o->s;
  size_t n = (size_t)d->n;
  s.x.resize(n); s.v.resize(n);
  for (size_t i = 0; i < n; ++i) { s.x[i] = v3(d->x + 3 * i); s.v[i] = v3(d->v + 3 * i); }
  s.masse.assign(d->masse, d->masse + n);
  s.w.assign(d->w, d->w + n);
  s.kollisionsabstand.assign(d->kollisionsabstand, d->kollisionsabstand + n);
  s.reibung.assign(d->reibung, d->reibung + n);
  for (int e = 0; e < d->n_abstand; ++e) {
    s.abstand_punkte.push_back({d->abstand_punkte[2 * e], d->abstand_punkte[2 * e + 1]});
    s.abstand_ruhe.push_back(d->abstand_ruhe[e]);
    s.abstand_nachgiebigkeit.push_back(d->abstand_nachgiebigkeit[e]);
    s.membran_kante.push_back(d->membran_kante ? d->membran_kante[e] : 0);
  }
  for (int e = 0; e < d->n_naht; ++e) {
    s.naht_punkte.push_back({d->naht_punkte[2 * e], d->naht_punkte[2 * e + 1]});
    s.naht_ruhe.push_back(d->naht_ruhe[e]);
  }
  for (int b = 0; b < d->n_biegung; ++b) {
    s.biegung_punkte.push_back({d->biegung_punkte[4 * b], d->biegung_punkte[4 * b + 1], d->biegung_punkte[4 * b + 2], d->biegung_punkte[4 * b + 3]});
    s.biegung_ruhe.push_back(d->biegung_ruhe[b]);
    s.biegung_nachgiebigkeit.push_back(d->biegung_nachgiebigkeit[b]);
  }
  for (int t = 0; t < d->n_membran; ++t) {
    MembranElement me;
    for (int k = 0; k < 3; ++k) me.punkte[k] = d->membran_punkte[3 * t + k];
    me.element = membran_aus(d->membran_daten + 12 * t);
    s.membran_elemente.push_back(me);
  }
  for (int t = 0; t < d->n_dreiecke; ++t) s.dreiecke.push_back({d->dreiecke[3 * t], d->dreiecke[3 * t + 1], d->dreiecke[3 * t + 2]});
  if (d->teil_von_punkt) s.teil_von_punkt.assign(d->teil_von_punkt, d->teil_von_punkt + n);
  if (d->selbst_e_punkt) s.selbst_e_punkt.assign(d->selbst_e_punkt, d->selbst_e_punkt + n);
  if (d->selbst_e_dreieck) s.selbst_e_dreieck.assign(d->selbst_e_dreieck, d->selbst_e_dreieck + d->n_dreiecke);
  if (d->selbst_a_dreieck) s.selbst_a_dreieck.assign(d->selbst_a_dreieck, d->selbst_a_dreieck + d->n_dreiecke);
  s.bild_dt = d->bild_dt;
  s.schwerkraft = v3(d->schwerkraft);
  s.daempfung = d->daempfung;
  if (p) o->param = *p; else or_param_default(&o->param);
  return o;
}
void or_sim_frei(or_sim* s) { delete s; }
int or_sim_koerper_zylinder(or_sim* s, const double punkt[3], const double achse[3], double radius, const double geschw[3]) {
  Koerper k; k.art = 0; k.punkt = v3(punkt); k.achse = v3(achse); k.radius = radius; k.geschwindigkeit = v3(geschw);
  s->s.koerper.push_back(k);
  return (int)s->s.koerper.size() - 1;
}
int or_sim_koerper_sdf(or_sim* s, const float* werte, const int dims[3], const double origin[3], double spacing, const double versatz[3], const double geschw[3]) {
  Koerper k; k.art = 1;
  size_t cnt = (size_t)dims[0] * (size_t)dims[1] * (size_t)dims[2];
  k.werte.assign(werte, werte + cnt);
  for (int a = 0; a < 3; ++a) { k.dims[a] = dims[a]; k.origin[a] = origin[a]; }
  k.spacing = spacing; k.versatz = v3(versatz); k.geschwindigkeit = v3(geschw);
  s->s.koerper.push_back(k);
  return (int)s->s.koerper.size() - 1;
}
int or_sim_koerper_netz(or_sim* s, const or_netz* k, const double versatz[3], const double geschw[3]);
void or_sim_param(or_sim* s, const or_param* p) { s->param = *p; }
void or_sim_lage(const or_sim* s, double* x, double* v) {
  for (size_t i = 0; i < s->s.x.size(); ++i) { if (x) put(x + 3 * i, s->s.x[i]); if (v) put(v + 3 * i, s->s.v[i]); }
}
void or_sim_setze_lage(or_sim* s, const double* x, const double* v) {
  for (size_t i = 0; i < s->s.x.size(); ++i) { if (x) s->s.x[i] = v3(x + 3 * i); if (v) s->s.v[i] = v3(v + 3 * i); }
}
void or_sim_setze_naht_ruhe(or_sim* s, const double* ruhe) { for (size_t e = 0; e < s->s.naht_ruhe.size(); ++e) s->s.naht_ruhe[e] = ruhe[e]; }
void or_sim_setze_w(or_sim* s, const double* w) { for (size_t i = 0; i < s->s.w.size(); ++i) s->s.w[i] = w[i]; }
void or_sim_setze_dv_vorher(or_sim* s, const double* dv, int32_t n) {
  s->dv_vorher.clear();
  for (int i = 0; i < n; ++i) s->dv_vorher.push_back(v3(dv + 3 * i));
}
int32_t or_sim_dv_vorher(const or_sim* s, double* out) {
  if (out)
    for (size_t i = 0; i < s->dv_vorher.size(); ++i) put(out + 3 * i, s->dv_vorher[i]);
  return (int32_t)s->dv_vorher.size();
}
void or_sim_bild(or_sim* s) { s->bild(); }
void or_sim_xpbd_bild(or_sim* s, int32_t unterschritte, double naht_nachgiebigkeit, or_xpbd_statistik* out) {
  Xpbd x;
  x.bild(s->s, unterschritte, naht_nachgiebigkeit);
  if (out) *out = x.stat;
}
void or_sim_schritt(or_sim* s, double h) { s->schritt(h); }
void or_sim_statistik(const or_sim* s, or_statistik* out) { *out = s->stat; }
void or_sim_reset_statistik(or_sim* s) { s->stat = or_statistik{}; }

void or_sys_bauen(or_sim* s, double h) { s->system_bauen(h); }
int32_t or_sys_n_kanten(const or_sim* s) { return (int32_t)(s->sys.kanten.size() + s->sys.naehte.size()); }
int32_t or_sys_n_biegungen(const or_sim* s) { return (int32_t)s->sys.biegungen.size(); }
void or_sys_kanten(const or_sim* s, uint32_t* ij, double* k, double* richtung, double* quer, double* ruhe) {
  size_t c = 0;
  for (const auto* v : {&s->sys.kanten, &s->sys.naehte})
    for (const auto& e : *v) {
      if (ij) { ij[2 * c] = e.i; ij[2 * c + 1] = e.j; }
      if (k) k[c] = e.k;
      if (richtung) put(richtung + 3 * c, e.richtung);
      if (quer) quer[c] = e.quer;
      if (ruhe) ruhe[c] = e.ruhe;
      ++c;
    }
}
void or_sys_biegungen(const or_sim* s, uint32_t* punkte, double* grad, double* k) {
  for (size_t b = 0; b < s->sys.biegungen.size(); ++b) {
    const auto& bg = s->sys.biegungen[b];
    for (int m = 0; m < 4; ++m) {
      if (punkte) punkte[4 * b + m] = bg.punkte[m];
      if (grad) put(grad + 12 * b + 3 * m, bg.grad[m]);
    }
    if (k) k[b] = bg.k;
  }
}
void or_sys_membran_h(const or_sim* s, double* out) {
  for (size_t t = 0; t < s->sys.membran_h.size(); ++t) {
    const auto& h = s->sys.membran_h[t];
    put(out + 9 * t, h.f1); put(out + 9 * t + 3, h.f2); put(out + 9 * t + 6, h.w);
  }
}
void or_sys_vektoren(const or_sim* s, double* rhs, double* f, double* last, double* pinv9) {
  for (size_t i = 0; i < s->sys.rhs.size(); ++i) {
    if (rhs) put(rhs + 3 * i, s->sys.rhs[i]);
    if (f) put(f + 3 * i, s->sys.f[i]);
    if (last) put(last + 3 * i, s->sys.last[i]);
    if (pinv9) for (int r = 0; r < 3; ++r) put(pinv9 + 9 * i + 3 * r, s->sys.pinv[i][r]);
  }
}
void or_sys_diag(const or_sim* s, double* diag9) {
  auto dg = s->sys.diag();
  for (size_t i = 0; i < dg.size(); ++i) for (int r = 0; r < 3; ++r) put(diag9 + 9 * i + 3 * r, dg[i][r]);
}
void or_sys_fest(const or_sim* s, int32_t* flag, double* normale, double* vn_soll, double* v_koerper) {
  for (size_t i = 0; i < s->fest.size(); ++i) {
    const Fest& fe = s->fest[i];
    if (flag) flag[i] = fe.da ? 1 : 0;
    if (normale) put(normale + 3 * i, fe.normale);
    if (vn_soll) vn_soll[i] = fe.vn_soll;
    if (v_koerper) put(v_koerper + 3 * i, fe.v_koerper);
  }
}
void or_sys_setze_fest(or_sim* s, const int32_t* flag, const double* normale, const double* vn_soll, const double* v_koerper) {
  s->fest.assign(s->s.x.size(), Fest());
  for (size_t i = 0; i < s->fest.size(); ++i) {
    s->fest[i].da = flag[i] != 0;
    s->fest[i].normale = v3(normale + 3 * i);
    s->fest[i].vn_soll = vn_soll[i];
    s->fest[i].v_koerper = v3(v_koerper + 3 * i);
  }
}
void or_sys_mal_h(const or_sim* s, const double* p, double* out) {
  size_t n = s->s.x.size();
  std::vector<V3> pp(n), oo(n);
  for (size_t i = 0; i < n; ++i) pp[i] = v3(p + 3 * i);
  s->sys.mal_h(pp.data(), oo.data());
  for (size_t i = 0; i < n; ++i) put(out + 3 * i, oo[i]);
}
void or_sys_mal_a(const or_sim* s, const double* p, double* out) {
  size_t n = s->s.x.size();
  std::vector<V3> pp(n), oo(n);
  for (size_t i = 0; i < n; ++i) pp[i] = v3(p + 3 * i);
  s->sys.mal_a(pp.data(), oo.data());
  for (size_t i = 0; i < n; ++i) put(out + 3 * i, oo[i]);
}
void or_sys_loese(const or_sim* s, double* dv, int32_t* iterationen, int32_t* nicht_konvergiert, double* rest) {
  size_t n = s->s.x.size();
  std::vector<V3> d(n);
  for (size_t i = 0; i < n; ++i) d[i] = v3(dv + 3 * i);
  CgLauf l = loese(s->sys, s->fest, d);
  for (size_t i = 0; i < n; ++i) put(dv + 3 * i, d[i]);
  if (iterationen) *iterationen = l.iterationen;
  if (nicht_konvergiert) *nicht_konvergiert = l.nicht_konvergiert ? 1 : 0;
  if (rest) *rest = l.rest;
}
void or_sys_rest(const or_sim* s, const double* dv, double* out) {
  or_sys_mal_a(s, dv, out);
  for (size_t i = 0; i < s->sys.rhs.size(); ++i)
    for (int a = 0; a < 3; ++a) out[3 * i + a] -= s->sys.rhs[i][a];
}

static void naehe_aus(const std::vector<Naehe>& l, int32_t max, int32_t* art, uint32_t* punkte4, double* gewichte4,
                      double* normale3, double* abstand, double* ruhe) {
  for (size_t q = 0; q < l.size() && (int32_t)q < max; ++q) {
    const Naehe& p = l[q];
    if (art) art[q] = (int32_t)p.art;
    for (int k = 0; k < 4; ++k) {
      if (punkte4) punkte4[4 * q + k] = p.punkte[k];
      if (gewichte4) gewichte4[4 * q + k] = p.gewichte[k];
    }
    if (normale3) put(normale3 + 3 * q, p.normale);
    if (abstand) abstand[q] = p.abstand;
    if (ruhe) ruhe[q] = p.ruhe;
  }
}
static void stoesse_aus(const std::vector<Stoss>& l, int32_t max, int32_t* art, uint32_t* punkte4, double* lage2,
                        double* zeit, double* seite, double* normale3) {
  for (size_t q = 0; q < l.size() && (int32_t)q < max; ++q) {
    const Stoss& p = l[q];
    if (art) art[q] = (int32_t)p.art;
    for (int k = 0; k < 4; ++k)
      if (punkte4) punkte4[4 * q + k] = p.punkte[k];
    if (lage2) { lage2[2 * q] = p.lage[0]; lage2[2 * q + 1] = p.lage[1]; }
    if (zeit) zeit[q] = p.zeit;
    if (seite) seite[q] = p.seite;
    if (normale3) put(normale3 + 3 * q, p.normale);
  }
}
int32_t or_selbst_naehe(or_sim* s, int32_t brute, int32_t max, int32_t* art, uint32_t* punkte4, double* gewichte4, double* normale3,
                        double* abstand, double* ruhe) {
  if (!s->hat_selbst || s->netz.r_punkt.size() != s->s.x.size()) s->selbst_bauen();
  std::vector<Naehe> l = finde(s->netz, s->s.x, s->s.w, true, true, brute != 0);
  naehe_aus(l, max, art, punkte4, gewichte4, normale3, abstand, ruhe);
  return (int32_t)l.size();
}
int32_t or_sys_federn(const or_sim* s, int32_t max, uint32_t* punkte5, double* gewichte5, double* kraft3, double* block9) {
  const auto& l = s->sys.federn;
  for (size_t q = 0; q < l.size() && (int32_t)q < max; ++q) {
    for (int k = 0; k < 5; ++k) {
      if (punkte5) punkte5[5 * q + k] = l[q].punkte[k];
      if (gewichte5) gewichte5[5 * q + k] = l[q].gewichte[k];
    }
    if (kraft3) put(kraft3 + 3 * q, l[q].kraft);
    if (block9)
      for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) block9[9 * q + 3 * a + b] = l[q].block[a][b];
  }
  return (int32_t)l.size();
}
void or_sim_v(const or_sim* s, double* v) {
  for (size_t i = 0; i < s->s.v.size(); ++i) put(v + 3 * i, s->s.v[i]);
}
int32_t or_selbst_schnitte(or_sim* s, int32_t max, uint32_t* kante_dreieck, double* t, double* w3, double* richtung3,
                           double* gewicht, uint32_t* wurzel_out) {
  if (!s->hat_selbst || s->netz.r_punkt.size() != s->s.x.size()) s->selbst_bauen();
  std::vector<Schnitt> l = finde_schnitte(s->netz, s->s.x, s->param.selbst_schnitte_an_naehten != 0);
  std::vector<size_t> wz = gebiete(s->netz, l);
  for (size_t q = 0; q < l.size() && (int32_t)q < max; ++q) {
    if (kante_dreieck) { kante_dreieck[2 * q] = l[q].kante; kante_dreieck[2 * q + 1] = l[q].dreieck; }
    if (t) t[q] = l[q].t;
    if (w3) for (int k = 0; k < 3; ++k) w3[3 * q + k] = l[q].w[k];
    if (richtung3) put(richtung3 + 3 * q, l[q].richtung);
    if (gewicht) gewicht[q] = l[q].gewicht;
    if (wurzel_out) wurzel_out[q] = (uint32_t)wz[q];
  }
  return (int32_t)l.size();
}
int32_t or_selbst_stoesse(or_sim* s, const double* x0, const double* x1, int32_t max, int32_t* art, uint32_t* punkte4,
                          double* lage2, double* zeit, double* seite, double* normale3) {
  if (!s->hat_selbst || s->netz.r_punkt.size() != s->s.x.size()) s->selbst_bauen();
  size_t n = s->s.x.size();
  std::vector<V3> a(n), b(n);
  for (size_t i = 0; i < n; ++i) { a[i] = v3(x0 + 3 * i); b[i] = v3(x1 + 3 * i); }
  std::vector<Stoss> l = finde_stoesse(s->netz, a, b, s->s.w, true, true);
  stoesse_aus(l, max, art, punkte4, lage2, zeit, seite, normale3);
  return (int32_t)l.size();
}
static Stoffnetz bares_netz(int32_t n, int32_t nt, const uint32_t* dreiecke, const double* r_punkt, int32_t nk, const uint32_t* kanten) {
  std::vector<std::array<uint32_t, 3>> dr;
  for (int t = 0; t < nt; ++t) dr.push_back({dreiecke[3 * t], dreiecke[3 * t + 1], dreiecke[3 * t + 2]});
  std::vector<double> r(r_punkt ? r_punkt : nullptr, r_punkt ? r_punkt + n : nullptr);
  r.resize(n, 0.0);
  if (nt > 0 || nk == 0) return Stoffnetz::neu(dr, r);
  Stoffnetz s;
  for (int e = 0; e < nk; ++e) s.kanten.push_back({kanten[2 * e], kanten[2 * e + 1]});
  s.r_punkt = r;
  return s;
}
int32_t or_netz_naehe(int32_t n, const double* x, const double* w, int32_t nt, const uint32_t* dreiecke, const double* r_punkt,
                      int32_t punkt_dreieck, int32_t kante_kante, int32_t max, int32_t* art, uint32_t* punkte4,
                      double* gewichte4, double* normale3, double* abstand, double* ruhe) {
  Stoffnetz netz = bares_netz(n, nt, dreiecke, r_punkt, 0, nullptr);
  std::vector<V3> xv(n);
  for (int i = 0; i < n; ++i) xv[i] = v3(x + 3 * i);
  std::vector<double> wv(w, w + n);
  std::vector<Naehe> l = finde(netz, xv, wv, punkt_dreieck != 0, kante_kante != 0);
  naehe_aus(l, max, art, punkte4, gewichte4, normale3, abstand, ruhe);
  return (int32_t)l.size();
}
int32_t or_netz_stoesse(int32_t n, const double* x0, double* x1, const double* w, const double* masse, int32_t nt,
                        const uint32_t* dreiecke, int32_t nk, const uint32_t* kanten, int32_t punkt_dreieck,
                        int32_t kante_kante, int32_t iterationen, int32_t max, int32_t* art, uint32_t* punkte4,
                        double* lage2, double* zeit, double* seite, int32_t* korrekturen) {
  Stoffnetz netz = bares_netz(n, nt, dreiecke, nullptr, nk, kanten);
  std::vector<V3> a(n), b(n);
  for (int i = 0; i < n; ++i) { a[i] = v3(x0 + 3 * i); b[i] = v3(x1 + 3 * i); }
  std::vector<double> wv(w, w + n), mv(masse, masse + n);
  std::vector<Stoss> l = finde_stoesse(netz, a, b, wv, punkt_dreieck != 0, kante_kante != 0);
  stoesse_aus(l, max, art, punkte4, lage2, zeit, seite, nullptr);
  size_t z = stossantwort(l, b, mv, wv, (size_t)iterationen);
  if (korrekturen) *korrekturen = (int32_t)z;
  for (int i = 0; i < n; ++i) put(x1 + 3 * i, b[i]);
  return (int32_t)l.size();
}

}  // extern "C"

// ---------------------------------------------------------------------------
// C API of the mesh body
// ---------------------------------------------------------------------------
struct or_netz { Koerpernetz k; };

static void fehler_setzen(const std::string& f, char* buf, int32_t len) {
  if (buf && len > 0) {
    size_t n = std::min<size_t>(f.size(), (size_t)len - 1);
    std::memcpy(buf, f.data(), n);
    buf[n] = 0;
  }
}
static void naechster_aus(const Naechster& n, or_naechster* o) {
  o->abstand = n.abstand;
  put(o->punkt, n.punkt);
  put(o->normale, n.normale);
  o->dreieck = n.dreieck;
  o->region = n.region;
}

extern "C" {

or_netz* or_netz_neu(int32_t n_punkte, const double* punkte, int32_t n_dreiecke, const uint32_t* dreiecke, double suchtiefe,
                     char* fehler, int32_t fehler_len) {
  std::vector<V3> p((size_t)n_punkte);
  for (int32_t i = 0; i < n_punkte; ++i) p[i] = v3(punkte + 3 * i);
  std::vector<Dreieck> d((size_t)n_dreiecke);
  for (int32_t t = 0; t < n_dreiecke; ++t) d[t] = {dreiecke[3 * t], dreiecke[3 * t + 1], dreiecke[3 * t + 2]};
  or_netz* o = new or_netz();
  std::string f = o->k.neu(std::move(p), std::move(d), suchtiefe);
  if (!f.empty()) {
    fehler_setzen(f, fehler, fehler_len);
    delete o;
    return nullptr;
  }
  return o;
}
void or_netz_frei(or_netz* k) { delete k; }
int or_netz_nachfuehren(or_netz* k, const double* punkte, char* fehler, int32_t fehler_len) {
  std::vector<V3> p(k->k.punkte.size());
  for (size_t i = 0; i < p.size(); ++i) p[i] = v3(punkte + 3 * i);
  std::string f = k->k.nachfuehren(p);
  if (!f.empty()) {
    fehler_setzen(f, fehler, fehler_len);
    return 1;
  }
  return 0;
}
int or_netz_punktgeschwindigkeit(or_netz* k, const double* v, int32_t n) {  // netzkoerper.rs:454-466
  if (n != 0 && (size_t)n != k->k.punkte.size()) return 1;
  k->k.punktgeschwindigkeit.clear();
  for (int32_t i = 0; i < n; ++i) k->k.punktgeschwindigkeit.push_back(v3(v + 3 * i));
  return 0;
}
void or_netz_info(const or_netz* k, int64_t* randkanten, double* volumen, double ursprung[3], double* zelle, int64_t zahl[3]) {
  *randkanten = (int64_t)k->k.randkanten;
  *volumen = k->k.volumen;
  put(ursprung, k->k.gitter.ursprung);
  *zelle = k->k.gitter.zelle;
  for (int a = 0; a < 3; ++a) zahl[a] = (int64_t)k->k.gitter.zahl[a];
}
void or_netz_gitter(const or_netz* k, int32_t punkte, int64_t* n_anfang, int64_t* n_eintraege, uint32_t* anfang, uint32_t* eintraege) {
  const Gitter& g = punkte ? k->k.punktgitter : k->k.gitter;
  *n_anfang = (int64_t)g.anfang.size();
  *n_eintraege = (int64_t)g.eintraege.size();
  if (anfang) std::memcpy(anfang, g.anfang.data(), g.anfang.size() * 4);
  if (eintraege) std::memcpy(eintraege, g.eintraege.data(), g.eintraege.size() * 4);
}
void or_netz_normalen(const or_netz* k, double* flaechennormale, double* punktnormale, uint32_t* nachbar, double* kugel4) {
  const Koerpernetz& n = k->k;
  for (size_t t = 0; t < n.dreiecke.size(); ++t) {
    if (flaechennormale) put(flaechennormale + 3 * t, n.flaechennormale[t]);
    if (nachbar)
      for (int e = 0; e < 3; ++e) nachbar[3 * t + e] = n.nachbar[t][e];
    if (kugel4) {
      put(kugel4 + 4 * t, n.kugel[t].m);
      kugel4[4 * t + 3] = n.kugel[t].r;
    }
  }
  if (punktnormale)
    for (size_t i = 0; i < n.punkte.size(); ++i) put(punktnormale + 3 * i, n.punktnormale[i]);
}
int or_netz_naechster(const or_netz* k, const double x[3], double radius, or_naechster* out) {
  Naechster n;
  if (!k->k.naechster(v3(x), radius, &n)) return 0;
  naechster_aus(n, out);
  return 1;
}
void or_netz_naechster_alle(const or_netz* k, const double x[3], or_naechster* out) { naechster_aus(k->k.naechster_alle(v3(x)), out); }
void or_netz_abstand_ohne_grenze(const or_netz* k, const double x[3], or_naechster* out) {
  naechster_aus(k->k.abstand_ohne_grenze(v3(x)), out);
}
int or_netz_durchstoss(const or_netz* k, const double a[3], const double b[3], const double starr[3], double p[3], double n[3],
                       double vk[3]) {
  V3 pp, nn, vv;
  if (!k->k.durchstoss(v3(a), v3(b), v3(starr), &pp, &nn, &vv)) return 0;
  put(p, pp);
  put(n, nn);
  put(vk, vv);
  return 1;
}
double or_netz_geschwindigkeit_an(const or_netz* k, int32_t t, const double p[3], const double starr[3], double out[3]) {
  put(out, k->k.geschwindigkeit_an((size_t)t, v3(p), v3(starr)));
  return 0.0;
}
int32_t or_netz_punkte_gegen_dreiecke(const or_netz* k, double tief, const double versatz[3], const double geschw[3], int32_t nd,
                                      const uint32_t* dreiecke, int32_t n, const double* w, double* x, double* v) {
  std::vector<Dreieck> d((size_t)nd);
  for (int32_t t = 0; t < nd; ++t) d[t] = {dreiecke[3 * t], dreiecke[3 * t + 1], dreiecke[3 * t + 2]};
  std::vector<double> ww(w, w + n);
  std::vector<V3> xx(n), vv(n);
  for (int32_t i = 0; i < n; ++i) {
    xx[i] = v3(x + 3 * i);
    vv[i] = v3(v + 3 * i);
  }
  size_t z = k->k.punkte_gegen_dreiecke(tief, v3(versatz), v3(geschw), d, ww, xx, vv);
  for (int32_t i = 0; i < n; ++i) {
    put(x + 3 * i, xx[i]);
    put(v + 3 * i, vv[i]);
  }
  return (int32_t)z;
}
int or_netz_kontakt(const or_netz* k, const double versatz[3], const double geschw[3], const double x[3], double abstand,
                    double normale[3], double* tiefe, double vk[3]) {
  Koerper kb;
  kb.art = 2;
  kb.netz = std::shared_ptr<const Koerpernetz>(&k->k, [](const Koerpernetz*) {});
  kb.versatz = v3(versatz);
  kb.geschwindigkeit = v3(geschw);
  Kontakt kt;
  if (!kb.kontakt(v3(x), abstand, &kt)) return 0;
  put(normale, kt.normale);
  *tiefe = kt.tiefe;
  put(vk, kt.geschwindigkeit);
  return 1;
}
int or_sim_koerper_netz(or_sim* s, const or_netz* k, const double versatz[3], const double geschw[3]) {
  Koerper kb;
  kb.art = 2;
  kb.netz = std::make_shared<const Koerpernetz>(k->k);
  kb.versatz = v3(versatz);
  kb.geschwindigkeit = v3(geschw);
  s->s.koerper.push_back(kb);
  return (int)s->s.koerper.size() - 1;
}

}  // extern "C"
