// gc_netz.h — the triangle-mesh body on the device: `Koerpernetz`
// (crates/klon3d-sim/src/netzkoerper.rs) as device arrays, its per-pose build
// (face and point normals, bounding spheres, the uniform grid of triangle
// lists) and its two queries: the nearest surface point within a radius with
// the sign from the angle-weighted pseudonormal (`naechster`, :503-531) and the
// first entry of a path from outside (`durchstoss`, :541-596).
//
// Every query is one thread per cloth point: it walks the at most 2x2x2 grid
// cells its search cube touches (cells are 2 x suchtiefe wide), tests each
// listed triangle's bounding sphere first and then takes Ericson's closest
// point. FP64 throughout, fp-contract off: a query does the reference's
// operations in the reference's order, so it returns the reference's bits
// for the same mesh arrays. The arrays themselves are built here; face normals,
// spheres and the grid are bit-identical to the reference's, the point normals
// use the device's atan2 (its last bit can differ from the host library's).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

// Regions of the nearest point (netzkoerper.rs:50-58): corner k, edge k
// (0 = p0p1, 1 = p1p2, 2 = p2p0), face.
enum { NK_ECKE = 0, NK_KANTE = 3, NK_FLAECHE = 6 };

// The device image of one mesh body. Lives in device memory; a body's
// `Koerper` record points at it.
struct NetzDev {
  const double* punkte;      // 3 per point
  const uint32_t* dreiecke;  // 3 per triangle
  const double* fn;          // face normal, 3 per triangle
  const uint32_t* nachbar;   // triangle across each side (itself at an open edge), 3 per triangle
  const double* pn;          // angle-weighted point normal, 3 per point
  const double* kugel;       // bounding sphere: centroid and largest corner distance, 4 per triangle
  const double* pv;          // velocity per point of a deforming body, or null
  const uint32_t* anfang;    // cell z lists eintraege[anfang[z] .. anfang[z + 1]]
  const uint32_t* eintraege;
  const uint32_t* p_anfang;  // the grid of the body's points on the same raster (one cell each)
  const uint32_t* p_eintraege;
  double ursprung[3];
  double zelle;
  int64_t zahl[3];
  double suchtiefe;
  int32_t n_punkte, n_dreiecke;
};

namespace nk {

GC_HD V3<double> lade3(const double* a, int64_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }

// Rust `f as i64` of a floor: saturating, NaN -> 0.
GC_HD int64_t als_i64(double f) {
  if (!(f == f)) return 0;
  if (f >= 9.2233720368547758e18) return INT64_MAX;
  if (f <= -9.2233720368547758e18) return INT64_MIN;
  return (int64_t)f;
}

// netzkoerper.rs:87-93
GC_HD void index(const NetzDev& g, V3<double> x, int64_t k[3]) {
  k[0] = als_i64(floor((x.x - g.ursprung[0]) / g.zelle));
  k[1] = als_i64(floor((x.y - g.ursprung[1]) / g.zelle));
  k[2] = als_i64(floor((x.z - g.ursprung[2]) / g.zelle));
}

// netzkoerper.rs:117-129: cell range per axis of the box [lo, hi], clipped;
// false if the box misses the grid.
GC_HD bool bereich(const NetzDev& g, V3<double> lo, V3<double> hi, int64_t r[3][2]) {
  int64_t a[3], b[3];
  index(g, lo, a);
  index(g, hi, b);
  for (int d = 0; d < 3; ++d) {
    const int64_t n = g.zahl[d];
    if (b[d] < 0 || a[d] >= n) return false;
    r[d][0] = a[d] > 0 ? a[d] : 0;
    r[d][1] = b[d] < n - 1 ? b[d] : n - 1;
  }
  return true;
}

// netzkoerper.rs:185-218 (Ericson, Real-Time Collision Detection, 5.1.5)
GC_HD V3<double> naechster_auf_dreieck(V3<double> p, V3<double> a, V3<double> b, V3<double> c, int* region) {
  auto quot = [](double z, double n) { return n > 0.0 ? z / n : 0.0; };
  const V3<double> ab = b - a, ac = c - a, ap = p - a;
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) {
    *region = NK_ECKE + 0;
    return a;
  }
  const V3<double> bp = p - b;
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) {
    *region = NK_ECKE + 1;
    return b;
  }
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) {
    *region = NK_KANTE + 0;
    return a + ab * quot(d1, d1 - d3);
  }
  const V3<double> cp = p - c;
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) {
    *region = NK_ECKE + 2;
    return c;
  }
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) {
    *region = NK_KANTE + 2;
    return a + ac * quot(d2, d2 - d6);
  }
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && d4 - d3 >= 0.0 && d5 - d6 >= 0.0) {
    const double w = quot(d4 - d3, (d4 - d3) + (d5 - d6));
    *region = NK_KANTE + 1;
    return b + (c - b) * w;
  }
  const double s = va + vb + vc;
  const double v = quot(vb, s), w = quot(vc, s);
  *region = NK_FLAECHE;
  return a + (ab * v + ac * w);
}

// netzkoerper.rs:220-227
GC_HD V3<double> normiert(V3<double> a) {
  const double l = norm(a);
  if (l > 0.0) return a * (1.0 / l);
  return zero3<double>();
}

// netzkoerper.rs:232-244
GC_HD V3<double> schwerpunktgewichte(V3<double> p, V3<double> a, V3<double> b, V3<double> c) {
  const V3<double> e1 = b - a, e2 = c - a, e3 = p - a;
  const double d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
  const double d20 = dot(e3, e1), d21 = dot(e3, e2);
  const double nenner = d00 * d11 - d01 * d01;
  if (nenner == 0.0) return {1.0, 0.0, 0.0};
  const double u = (d11 * d20 - d01 * d21) / nenner;
  const double w = (d00 * d21 - d01 * d20) / nenner;
  return {1.0 - u - w, u, w};
}

// netzkoerper.rs:474-489: the surface velocity at p on triangle t.
GC_HD V3<double> geschwindigkeit_an(const NetzDev& N, int64_t t, V3<double> p, V3<double> starr) {
  if (!N.pv) return starr;
  const uint32_t* d = N.dreiecke + 3 * t;
  const V3<double> gew = schwerpunktgewichte(p, lade3(N.punkte, d[0]), lade3(N.punkte, d[1]), lade3(N.punkte, d[2]));
  const double g[3] = {gew.x, gew.y, gew.z};
  double q[3] = {0.0, 0.0, 0.0};
  for (int j = 0; j < 3; ++j) {
    const V3<double> v = lade3(N.pv, d[j]);
    q[0] += g[j] * v.x;
    q[1] += g[j] * v.y;
    q[2] += g[j] * v.z;
  }
  return {q[0], q[1], q[2]};
}

// netzkoerper.rs:431-447: the pseudonormal of the region.
GC_HD V3<double> pseudonormale(const NetzDev& N, int64_t t, int r) {
  if (r == NK_FLAECHE) return lade3(N.fn, t);
  if (r >= NK_KANTE) {
    const int64_t o = N.nachbar[3 * t + (r - NK_KANTE)];
    if (o == t) return lade3(N.fn, t);
    return normiert(lade3(N.fn, t) + lade3(N.fn, o));
  }
  return lade3(N.pn, N.dreiecke[3 * t + (r - NK_ECKE)]);
}

struct Treffer {
  double abstand;  // outside positive
  V3<double> punkt, normale;
  uint32_t dreieck;
  int region;
};

// netzkoerper.rs:492-500
GC_HD Treffer ergebnis(const NetzDev& N, V3<double> x, int64_t t, V3<double> c, int r) {
  const V3<double> pn = pseudonormale(N, t, r);
  const V3<double> d = x - c;
  const double l = norm(d);
  const bool aussen = dot(d, pn) >= 0.0;
  Treffer e;
  e.normale = l > 1e-12 ? d * (aussen ? 1.0 / l : -1.0 / l) : pn;
  e.abstand = aussen ? l : -l;
  e.punkt = c;
  e.dreieck = (uint32_t)t;
  e.region = r;
  return e;
}

// netzkoerper.rs:503-531: the nearest surface point closer than `radius`
// (unsigned), visiting cells z-major, then y, then x, and each cell's list in
// order; a later triangle replaces the best only if strictly closer.
GC_HD bool naechster(const NetzDev& N, V3<double> x, double radius, Treffer* out) {
  int64_t r[3][2];
  const V3<double> rr = {radius, radius, radius};
  if (!bereich(N, x - rr, x + rr, r)) return false;
  double best = radius;
  int64_t bt = -1;
  V3<double> bq = zero3<double>();
  int breg = NK_FLAECHE;
  for (int64_t kz = r[2][0]; kz <= r[2][1]; ++kz)
    for (int64_t ky = r[1][0]; ky <= r[1][1]; ++ky)
      for (int64_t kx = r[0][0]; kx <= r[0][1]; ++kx) {
        const int64_t z = (kz * N.zahl[1] + ky) * N.zahl[0] + kx;
        const uint32_t e1 = N.anfang[z + 1];
        for (uint32_t e = N.anfang[z]; e < e1; ++e) {
          const int64_t t = N.eintraege[e];
          const double* kg = N.kugel + 4 * t;
          if (norm(x - V3<double>{kg[0], kg[1], kg[2]}) - kg[3] >= best) continue;
          const uint32_t* d = N.dreiecke + 3 * t;
          int reg;
          const V3<double> q = naechster_auf_dreieck(x, lade3(N.punkte, d[0]), lade3(N.punkte, d[1]), lade3(N.punkte, d[2]), &reg);
          const double l = norm(x - q);
          if (l < best) {
            best = l;
            bt = t;
            bq = q;
            breg = reg;
          }
        }
      }
  if (bt < 0) return false;
  *out = ergebnis(N, x, bt, bq, breg);
  return true;
}

// netzkoerper.rs:541-596 (Moeller & Trumbore 1997): the first entry of the
// segment a -> b through the surface from outside; *s_out is its parameter,
// clamped to [0, 1].
GC_HD bool durchstoss(const NetzDev& N, V3<double> a, V3<double> b, V3<double> starr, V3<double>* p, V3<double>* n,
                      V3<double>* vk, double* s_out) {
  const V3<double> d = b - a;
  if (dot(d, d) < 1e-24) return false;
  const V3<double> lo = {fmin(a.x, b.x), fmin(a.y, b.y), fmin(a.z, b.z)};
  const V3<double> hi = {fmax(a.x, b.x), fmax(a.y, b.y), fmax(a.z, b.z)};
  int64_t r[3][2];
  if (!bereich(N, lo, hi, r)) return false;
  const double eps = 1e-12;
  bool da = false;
  double sb = 0.0;
  int64_t tb = 0;
  for (int64_t kz = r[2][0]; kz <= r[2][1]; ++kz)
    for (int64_t ky = r[1][0]; ky <= r[1][1]; ++ky)
      for (int64_t kx = r[0][0]; kx <= r[0][1]; ++kx) {
        const int64_t z = (kz * N.zahl[1] + ky) * N.zahl[0] + kx;
        const uint32_t e1 = N.anfang[z + 1];
        for (uint32_t e = N.anfang[z]; e < e1; ++e) {
          const int64_t t = N.eintraege[e];
          if (dot(d, lade3(N.fn, t)) >= 0.0) continue;
          const uint32_t* dt = N.dreiecke + 3 * t;
          const V3<double> p0 = lade3(N.punkte, dt[0]), p1 = lade3(N.punkte, dt[1]), p2 = lade3(N.punkte, dt[2]);
          const V3<double> ea = p1 - p0, eb = p2 - p0;
          const V3<double> pv = cross(d, eb);
          const double det = dot(ea, pv);
          if (fabs(det) < 1e-30) continue;
          const V3<double> tv = a - p0;
          const double u = dot(tv, pv) / det;
          if (u < -eps || u > 1.0 + eps) continue;
          const V3<double> qv = cross(tv, ea);
          const double v = dot(d, qv) / det;
          if (v < -eps || u + v > 1.0 + eps) continue;
          const double s = dot(eb, qv) / det;
          if (s >= -eps && s <= 1.0 + eps && (!da || s < sb)) {
            da = true;
            sb = s;
            tb = t;
          }
        }
      }
  if (!da) return false;
  const double sc = sb < 0.0 ? 0.0 : (sb > 1.0 ? 1.0 : sb);
  *p = a + d * sc;
  *n = lade3(N.fn, tb);
  *vk = geschwindigkeit_an(N, tb, *p, starr);
  if (s_out) *s_out = sc;
  return true;
}

}  // namespace nk

// ---------------------------------------------------------------------------
// Per-pose build of the mesh arrays (netzkoerper.rs:257-291, 820-912)
// ---------------------------------------------------------------------------

// Per triangle: face normal and bounding sphere (netzkoerper.rs:264-287).
__global__ void k_nk_flaechen(const double* __restrict__ punkte, const uint32_t* __restrict__ dreiecke, int nt,
                              double* __restrict__ fn, double* __restrict__ kugel) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    const uint32_t* d = dreiecke + 3 * (int64_t)t;
    const V3<double> q0 = nk::lade3(punkte, d[0]), q1 = nk::lade3(punkte, d[1]), q2 = nk::lade3(punkte, d[2]);
    const V3<double> n = nk::normiert(cross(q1 - q0, q2 - q0));
    fn[3 * (int64_t)t] = n.x;
    fn[3 * (int64_t)t + 1] = n.y;
    fn[3 * (int64_t)t + 2] = n.z;
    const V3<double> m = ((q0 + q1) + q2) * (1.0 / 3.0);
    double r = 0.0;  // fold(0.0, f64::max)
    r = fmax(r, norm(q0 - m));
    r = fmax(r, norm(q1 - m));
    r = fmax(r, norm(q2 - m));
    kugel[4 * (int64_t)t] = m.x;
    kugel[4 * (int64_t)t + 1] = m.y;
    kugel[4 * (int64_t)t + 2] = m.z;
    kugel[4 * (int64_t)t + 3] = r;
  }
}

// Per point: the angle-weighted pseudonormal (netzkoerper.rs:276-290). The
// point's incidences (3t + k, ascending) are summed in the reference's
// triangle order.
__global__ void k_nk_punktnormalen(const double* __restrict__ punkte, const uint32_t* __restrict__ dreiecke,
                                   const double* __restrict__ fn, const uint32_t* __restrict__ inz_anfang,
                                   const uint32_t* __restrict__ inz, int n, double* __restrict__ pn) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    V3<double> s = zero3<double>();
    for (uint32_t e = inz_anfang[i]; e < inz_anfang[i + 1]; ++e) {
      const uint32_t t = inz[e] / 3, k = inz[e] % 3;
      const uint32_t* d = dreiecke + 3 * (int64_t)t;
      const V3<double> qk = nk::lade3(punkte, d[k]);
      const V3<double> u = nk::lade3(punkte, d[(k + 1) % 3]) - qk, v = nk::lade3(punkte, d[(k + 2) % 3]) - qk;
      const double w = atan2(norm(cross(u, v)), dot(u, v));  // netzkoerper.rs:246-248
      s = s + nk::lade3(fn, t) * w;
    }
    s = nk::normiert(s);
    pn[3 * (int64_t)i] = s.x;
    pn[3 * (int64_t)i + 1] = s.y;
    pn[3 * (int64_t)i + 2] = s.z;
  }
}

// Cell range of triangle t from its corners' cells (netzkoerper.rs:800-816),
// clipped to the grid (netzkoerper.rs:131-141).
__device__ __forceinline__ void nk_dreieck_zellen(const NetzDev& g, const double* punkte, const uint32_t* d, int64_t r[3][2]) {
  int64_t lo[3], hi[3], k[3];
  nk::index(g, nk::lade3(punkte, d[0]), lo);
  for (int a = 0; a < 3; ++a) hi[a] = lo[a];
  for (int e = 1; e < 3; ++e) {
    nk::index(g, nk::lade3(punkte, d[e]), k);
    for (int a = 0; a < 3; ++a) {
      lo[a] = lo[a] < k[a] ? lo[a] : k[a];
      hi[a] = hi[a] > k[a] ? hi[a] : k[a];
    }
  }
  for (int a = 0; a < 3; ++a) {
    r[a][0] = lo[a] > 0 ? lo[a] : 0;
    r[a][1] = hi[a] < g.zahl[a] - 1 ? hi[a] : g.zahl[a] - 1;
  }
}

// Pass 1 of the grid fill: entries per cell.
__global__ void k_nk_zaehlen(const NetzDev* __restrict__ gp, const double* __restrict__ punkte,
                             const uint32_t* __restrict__ dreiecke, int nt, uint32_t* __restrict__ zahl_je_zelle) {
  const NetzDev& g = *gp;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    int64_t r[3][2];
    nk_dreieck_zellen(g, punkte, dreiecke + 3 * (int64_t)t, r);
    for (int64_t kz = r[2][0]; kz <= r[2][1]; ++kz)
      for (int64_t ky = r[1][0]; ky <= r[1][1]; ++ky)
        for (int64_t kx = r[0][0]; kx <= r[0][1]; ++kx)
          atomicAdd(&zahl_je_zelle[(kz * g.zahl[1] + ky) * g.zahl[0] + kx], 1u);
  }
}

// Exclusive scan of the counts into the cell starts (one block; the grid has
// a few ten thousand cells). anfang[nz] is the number of entries.
__global__ void k_nk_scan(const uint32_t* __restrict__ zahl_je_zelle, int64_t nz, uint32_t* __restrict__ anfang) {
  __shared__ uint32_t teil[1024];
  const int64_t je = (nz + blockDim.x - 1) / blockDim.x;
  const int64_t a = threadIdx.x * je, b = a + je < nz ? a + je : nz;
  uint32_t s = 0;
  for (int64_t z = a; z < b; ++z) s += zahl_je_zelle[z];
  teil[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t lauf = 0;
    for (unsigned k = 0; k < blockDim.x; ++k) {
      const uint32_t v = teil[k];
      teil[k] = lauf;
      lauf += v;
    }
    anfang[nz] = lauf;
  }
  __syncthreads();
  uint32_t lauf = teil[threadIdx.x];
  for (int64_t z = a; z < b; ++z) {
    anfang[z] = lauf;
    lauf += zahl_je_zelle[z];
  }
}

// Pass 2: each triangle into every cell of its range, at a slot taken by an
// atomic cursor (the order within a cell is fixed by k_nk_ordnen).
__global__ void k_nk_fuellen(const NetzDev* __restrict__ gp, const double* __restrict__ punkte,
                             const uint32_t* __restrict__ dreiecke, int nt, uint32_t* __restrict__ cursor,
                             uint32_t* __restrict__ roh) {
  const NetzDev& g = *gp;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    int64_t r[3][2];
    nk_dreieck_zellen(g, punkte, dreiecke + 3 * (int64_t)t, r);
    for (int64_t kz = r[2][0]; kz <= r[2][1]; ++kz)
      for (int64_t ky = r[1][0]; ky <= r[1][1]; ++ky)
        for (int64_t kx = r[0][0]; kx <= r[0][1]; ++kx) {
          const uint32_t pos = atomicAdd(&cursor[(kz * g.zahl[1] + ky) * g.zahl[0] + kx], 1u);
          roh[pos] = (uint32_t)t;
        }
  }
}

// Pass 3: every entry goes to its rank among its cell's entries, so each
// cell lists its triangles in ascending order as the reference's serial fill
// does (netzkoerper.rs:899-911).
__global__ void k_nk_ordnen(const uint32_t* __restrict__ anfang, int64_t nz, const uint32_t* __restrict__ roh,
                            uint32_t* __restrict__ eintraege) {
  const uint32_t gesamt = anfang[nz];
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < gesamt; e += gridDim.x * blockDim.x) {
    // the cell: last z with anfang[z] <= e (cells without entries share a start)
    int64_t lo = 0, hi = nz - 1;
    while (lo < hi) {
      const int64_t m = (lo + hi + 1) / 2;
      if (anfang[m] <= e) lo = m;
      else hi = m - 1;
    }
    const uint32_t a = anfang[lo], b = anfang[lo + 1];
    const uint32_t mein = roh[e];
    uint32_t rang = 0;
    for (uint32_t q = a; q < b; ++q) rang += roh[q] < mein ? 1u : 0u;
    eintraege[a + rang] = mein;
  }
}

// Queries for the C ABI: the nearest point within `radius` and the path entry,
// one thread per query.
__global__ void k_nk_naechster(const NetzDev* __restrict__ gp, int n, const double* __restrict__ x, double radius,
                               int32_t* __restrict__ da, double* __restrict__ abstand, double* __restrict__ punkt,
                               double* __restrict__ normale, int32_t* __restrict__ dreieck, int32_t* __restrict__ region) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    nk::Treffer t;
    const bool ok = nk::naechster(*gp, nk::lade3(x, i), radius, &t);
    da[i] = ok ? 1 : 0;
    if (!ok) continue;
    abstand[i] = t.abstand;
    punkt[3 * (int64_t)i] = t.punkt.x;
    punkt[3 * (int64_t)i + 1] = t.punkt.y;
    punkt[3 * (int64_t)i + 2] = t.punkt.z;
    normale[3 * (int64_t)i] = t.normale.x;
    normale[3 * (int64_t)i + 1] = t.normale.y;
    normale[3 * (int64_t)i + 2] = t.normale.z;
    dreieck[i] = (int32_t)t.dreieck;
    region[i] = t.region;
  }
}

__global__ void k_nk_durchstoss(const NetzDev* __restrict__ gp, int n, const double* __restrict__ a,
                                const double* __restrict__ b, V3<double> starr, int32_t* __restrict__ da,
                                double* __restrict__ p, double* __restrict__ normale, double* __restrict__ vk) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    V3<double> q, nn, v;
    const bool ok = nk::durchstoss(*gp, nk::lade3(a, i), nk::lade3(b, i), starr, &q, &nn, &v, nullptr);
    da[i] = ok ? 1 : 0;
    if (!ok) continue;
    p[3 * (int64_t)i] = q.x;
    p[3 * (int64_t)i + 1] = q.y;
    p[3 * (int64_t)i + 2] = q.z;
    normale[3 * (int64_t)i] = nn.x;
    normale[3 * (int64_t)i + 1] = nn.y;
    normale[3 * (int64_t)i + 2] = nn.z;
    vk[3 * (int64_t)i] = v.x;
    vk[3 * (int64_t)i + 1] = v.y;
    vk[3 * (int64_t)i + 2] = v.z;
  }
}

namespace nk {

// netzkoerper.rs:770-782: the nearest surface point over all triangles.
GC_HD Treffer naechster_alle(const NetzDev& N, V3<double> x) {
  double bl = INFINITY;
  int64_t bt = 0;
  V3<double> bq = zero3<double>();
  int breg = NK_FLAECHE;
  for (int64_t t = 0; t < N.n_dreiecke; ++t) {
    const uint32_t* d = N.dreiecke + 3 * t;
    int reg;
    const V3<double> q = naechster_auf_dreieck(x, lade3(N.punkte, d[0]), lade3(N.punkte, d[1]), lade3(N.punkte, d[2]), &reg);
    const double l = norm(x - q);
    if (l < bl) {
      bl = l;
      bt = t;
      bq = q;
      breg = reg;
    }
  }
  return ergebnis(N, x, bt, bq, breg);
}

// netzkoerper.rs:741-767: the signed distance without a bound; the search
// cube doubles until it holds a triangle, the full scan answers beyond the grid.
GC_HD Treffer abstand_ohne_grenze(const NetzDev& N, V3<double> x) {
  const double xs[3] = {x.x, x.y, x.z};
  double quadrat = 0.0;
  for (int k = 0; k < 3; ++k) {
    const double e = N.ursprung[k] + N.zelle * (double)N.zahl[k];
    const double d = fmax(fabs(xs[k] - N.ursprung[k]), fabs(xs[k] - e));
    quadrat += d * d;
  }
  const double weit = sqrt(quadrat) + N.zelle;
  double r = fmax(N.suchtiefe, N.zelle);
  for (;;) {
    int64_t rr[3][2];
    const V3<double> rv = {r, r, r};
    if (bereich(N, x - rv, x + rv, rr)) {
      Treffer t;
      if (naechster(N, x, r, &t)) return t;
    }
    if (r >= weit) return naechster_alle(N, x);
    r = fmin(r * 2.0, weit);
  }
}

}  // namespace nk

// The mesh's signed distance on a regular grid (float32, x fastest): exact at
// every node closer than `band` to the surface; beyond it +-band with the
// node's side. Pass 1, one thread per node: the bounded query.
__global__ void k_nk_backen_band(const NetzDev* __restrict__ gp, V3<double> origin, double spacing, int nx, int ny,
                                 int64_t anzahl, double band, float* __restrict__ werte) {
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < anzahl; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t ix = q % nx, iy = (q / nx) % ny, iz = q / ((int64_t)nx * ny);
    const V3<double> x = {origin.x + spacing * (double)ix, origin.y + spacing * (double)iy, origin.z + spacing * (double)iz};
    nk::Treffer t;
    werte[q] = nk::naechster(*gp, x, band, &t) ? (float)t.abstand : NAN;
  }
}

// Pass 2, one thread per x-line: the side of every node beyond the band. With
// band >= spacing no surface crossing lies between two neighbouring nodes of
// which one is beyond the band (the crossing would be within one spacing of
// it), so a node beyond the band is on the side of its neighbour. A line's
// first node beyond the band takes its side from outside the mesh's box (+),
// or from the unbounded query.
__global__ void k_nk_backen_seiten(const NetzDev* __restrict__ gp, V3<double> origin, double spacing, int nx, int ny, int nz,
                                   double band, V3<double> huelle_lo, V3<double> huelle_hi, float* __restrict__ werte) {
  const int64_t linien = (int64_t)ny * nz;
  for (int64_t l = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; l < linien; l += (int64_t)gridDim.x * blockDim.x) {
    const int64_t iy = l % ny, iz = l / ny;
    float* w = werte + l * nx;
    double seite = 0.0;  // unknown
    for (int ix = 0; ix < nx; ++ix) {
      const float v = w[ix];
      if (v == v) {
        seite = v < 0.0f ? -1.0 : 1.0;
        continue;
      }
      if (seite == 0.0) {
        const V3<double> x = {origin.x + spacing * (double)ix, origin.y + spacing * (double)iy, origin.z + spacing * (double)iz};
        const bool draussen = x.x < huelle_lo.x || x.y < huelle_lo.y || x.z < huelle_lo.z || x.x > huelle_hi.x ||
                              x.y > huelle_hi.y || x.z > huelle_hi.z;
        seite = draussen ? 1.0 : (nk::abstand_ohne_grenze(*gp, x).abstand < 0.0 ? -1.0 : 1.0);
      }
      w[ix] = (float)(seite * band);
    }
  }
}

// ---------------------------------------------------------------------------
// Body points against cloth triangles (netzkoerper.rs:625-736)
// ---------------------------------------------------------------------------

// The point grid (netzkoerper.rs:861-879): each body point in its one cell.
__global__ void k_nk_punkte_zaehlen(const NetzDev* __restrict__ gp, const double* __restrict__ punkte, int n,
                                    uint32_t* __restrict__ zahl_je_zelle) {
  const NetzDev& g = *gp;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int64_t k[3];
    nk::index(g, nk::lade3(punkte, i), k);
    for (int a = 0; a < 3; ++a) k[a] = k[a] < 0 ? 0 : (k[a] > g.zahl[a] - 1 ? g.zahl[a] - 1 : k[a]);
    atomicAdd(&zahl_je_zelle[(k[2] * g.zahl[1] + k[1]) * g.zahl[0] + k[0]], 1u);
  }
}
__global__ void k_nk_punkte_fuellen(const NetzDev* __restrict__ gp, const double* __restrict__ punkte, int n,
                                    uint32_t* __restrict__ cursor, uint32_t* __restrict__ roh) {
  const NetzDev& g = *gp;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int64_t k[3];
    nk::index(g, nk::lade3(punkte, i), k);
    for (int a = 0; a < 3; ++a) k[a] = k[a] < 0 ? 0 : (k[a] > g.zahl[a] - 1 ? g.zahl[a] - 1 : k[a]);
    roh[atomicAdd(&cursor[(k[2] * g.zahl[1] + k[1]) * g.zahl[0] + k[0]], 1u)] = (uint32_t)i;
  }
}

// The deepest stab per cloth triangle (netzkoerper.rs:681-736, 641-646): every
// body point in the cells of the triangle's box widened by max(tief, band),
// cells z-major, each cell's points ascending; the line p + t n through the
// triangle (Moeller-Trumbore) at -tief < t < band; a later stab replaces the
// kept one only if deeper.
struct Stich {
  double tiefe;   // -t of the deepest stab
  V3<double> n;   // the body point's normal
  V3<double> bary;
  uint32_t kir;   // the body point
  uint32_t da;
};

__device__ __forceinline__ bool nk_stich(const NetzDev& N, V3<double> versatz, double tief, double band, V3<double> a,
                                         V3<double> b, V3<double> c, Stich* aus) {
  const double rand = fmax(tief, band);
  V3<double> lo = a, hi = a;
  for (const V3<double>& q : {b, c}) {
    lo = {fmin(lo.x, q.x), fmin(lo.y, q.y), fmin(lo.z, q.z)};
    hi = {fmax(hi.x, q.x), fmax(hi.y, q.y), fmax(hi.z, q.z)};
  }
  const V3<double> rr = {rand, rand, rand};
  lo = (lo - versatz) - rr;
  hi = (hi - versatz) + rr;
  const V3<double> e1 = b - a, e2 = c - a;
  int64_t r[3][2];
  bool da = false;
  if (!nk::bereich(N, lo, hi, r)) return false;
  for (int64_t kz = r[2][0]; kz <= r[2][1]; ++kz)
    for (int64_t ky = r[1][0]; ky <= r[1][1]; ++ky)
      for (int64_t kx = r[0][0]; kx <= r[0][1]; ++kx) {
        const int64_t z = (kz * N.zahl[1] + ky) * N.zahl[0] + kx;
        const uint32_t e1_ = N.p_anfang[z + 1];
        for (uint32_t e = N.p_anfang[z]; e < e1_; ++e) {
          const uint32_t k = N.p_eintraege[e];
          const V3<double> p = nk::lade3(N.punkte, k) + versatz;
          const V3<double> n = nk::lade3(N.pn, k);
          const V3<double> pv = cross(n, e2);
          const double det = dot(e1, pv);
          if (fabs(det) < 1e-18) continue;
          const V3<double> tv = p - a;
          const double bu = dot(tv, pv) / det;
          if (!(bu >= 0.0 && bu <= 1.0)) continue;
          const V3<double> qv = cross(tv, e1);
          const double bv = dot(n, qv) / det;
          if (bv < 0.0 || bu + bv > 1.0) continue;
          const double t = dot(e2, qv) / det;
          if (t > -tief && t < band) {
            if (!da || -t > aus->tiefe) {
              aus->tiefe = -t;
              aus->n = n;
              aus->bary = {1.0 - bu - bv, bu, bv};
              aus->kir = k;
            }
            da = true;
          }
        }
      }
  return da;
}
