// gc_selbst.h — cloth self-collision on the device (selbstkollision.rs):
// proximity pairs through a spatial hash, their springs, edge-through-triangle
// cuts with regions and push springs, and impacts on the path of a step.
//
// Every list comes out in the reference's order: the hash has the
// reference's bucket count, hash function and per-entry bucket dedup, bucket
// contents are sorted by entry index after an atomic fill, and every search
// runs twice (count, then write at the scanned offset) so that its output is
// ordered by query and, within a query, by cell and bucket position. With
// fp-contract off the FP64 arithmetic is that of the cited code, so the lists
// are bit-identical to the reference's. The few steps that depend on list
// order with shared points (approach damping of velocities, union-find of the
// cut regions) run in list order in one thread: their lists are short
// (tens to hundreds per step) next to the millions of operations around them.
//
// Paths relative to crates/klon3d-sim/src.
#pragma once
#include "gc_kernels.h"

#define SK_KEIN 0xFFFFFFFFu
#define SK_ZQ 2e-5          // ZUSCHLAG_QUADER, selbstkollision.rs:41
#define SK_ZP 4e-5          // ZUSCHLAG_PAAR, :43
#define SK_KANTE_MIN 1e-4   // :45
#define SK_AUSLENKUNG 2.5e-3  // :598
#define SK_ANNAEHERUNG (1.0 / 32.0)  // :600
#define SK_FOLGE_MAX 5u     // :806
#define SK_HEBEL 1e-3       // :809
#define SK_STOSS 1e-4       // STOSS_ABSTAND, :1092
#define SK_STOSS_KANTE 1e-4 // :1094
// Cells one entry or query may cover. A box over more cells only comes from a
// collapsed state; the step then reports an error instead of searching on.
// An entry of the hash whose box covers more cells than this has its buckets
// found by one block (k_sk_gross) instead of by one thread, whose pairwise
// dedup of the cells' buckets costs the square of their count. The reference
// accepts boxes of up to 64 cells per axis (sk_bereich_ok).
#define SK_KLEIN_ZELLEN 512

struct Quader {
  double lo[3], hi[3];
};

// Proximity pair (:60-71)
struct Naehe {
  uint32_t art;  // 0 point-triangle, 1 edge-edge
  uint32_t element;  // triangle (point-triangle) or first edge (edge-edge); second edge below
  uint32_t element2;
  uint32_t punkte[4];
  double gewichte[4];
  V3<double> normale;
  double abstand, ruhe;
};

// Cut (:812-825)
struct Schnitt {
  uint32_t kante, dreieck;
  double t;
  double w[3];
  V3<double> richtung;
  double gewicht;
};

// Impact (:1099-1112)
struct Stoss {
  uint32_t art;
  uint32_t element, element2;  // triangle or edge i; edge j
  uint32_t punkte[4];
  double lage[2];
  double zeit, seite;
  V3<double> normale;
  uint32_t gilt;  // survives ohne_durchdrungene
};

// Everything the self-collision kernels read and write; lives in device memory.
struct SelbstDev {
  int32_t n, nt, ne;
  const uint32_t* dreiecke;         // 3 nt
  const uint32_t* kanten;           // 2 ne
  const uint32_t* kanten_dreiecke;  // 2 ne, SK_KEIN = none
  const uint32_t* dreieck_kanten;   // 3 nt
  const double* r_punkt;
  const double* r_kante;
  const double* r_dreieck;
  double r_max;
  const uint32_t* partner_off;      // n + 1: seam and weld partners (gleich_oder_partner)
  const uint32_t* partner;
  const uint32_t* nahtw_off;        // n + 1: (partner, seam index) of the seam springs (Stoffwerte)
  const uint32_t* nahtw;            // 2 per entry
  const double* e_punkt;
  const double* a_punkt;
  const double* e_kante;
  const double* a_kante;
  const double* e_dreieck;
  const double* a_dreieck;
  const double* mu;                 // friction per point
  const double* naht_ruhe;          // per seam spring
  double steifigkeit;
  int32_t an_naehten;
  int32_t stoss_iterationen;
  // state
  const V3<double>* x;
  V3<double>* v;
  V3<double>* x_anfang;
  const double* masse;
  const double* w;
  const uint8_t* fflag;
  uint32_t* folge;                  // nt: steps in a row with a cut
  // scratch: normals
  V3<double>* n_dreieck;            // nt
  V3<double>* kantenebene;          // ne
  // scratch: hash
  Quader* quader;                   // max(n, ne, nt)
  uint32_t* eimer_anfang;           // eimer_max + 1
  uint32_t* eimer_fuell;            // eimer_max
  uint32_t* eintraege;              // kapazitaet_eintraege, bucket contents in entry order
  uint32_t* eintraege_roh;          // the same after the atomic fill, unordered
  uint32_t kapazitaet_eintraege;
  uint32_t* gross;                  // entries of the current hash with more than SK_KLEIN_ZELLEN cells
  uint32_t kapazitaet_gross;
  uint32_t* eimer_bits;             // eimer_max bits: the buckets one such entry has reached
  uint32_t* zahl;                   // max(n, ne, nt) + 1: per-query counts, scanned
  // lists
  Naehe* naehe;
  uint32_t kapazitaet_naehe;
  Schnitt* schnitte;
  uint32_t* schnitt_wurzel;
  uint32_t kapazitaet_schnitte;
  Stoss* stoesse;                   // nt + ne
  // springs (proximity, then push), their point lists
  uint32_t kapazitaet_federn;
  uint32_t* fe_punkte;              // 5 per spring
  double* fe_gew;                   // 5 per spring
  V3<double>* fe_kraft;
  double* fe_block;                 // 9 per spring, FP64
  uint32_t* fe_off;                 // n + 1
  uint32_t* fe_ent;                 // pair entries: spring << 5 | (q * 5 + r)
  uint32_t kapazitaet_fe_ent;
  uint32_t* fk_off;                 // n + 1
  uint32_t* fk_ent;                 // force entries: spring << 3 | q
  uint32_t kapazitaet_fk_ent;
  uint32_t* fe_roh;                 // fe_ent and fk_ent as filled (unordered per point)
  uint32_t* fk_roh;
  uint32_t* tausch;                 // scratch of the region merge (2 per pair)
  uint32_t kapazitaet_tausch;
  uint32_t* uf;                     // scratch of the region merge (3 per cut)
  V3<double>* summe;
  // stoss response scratch
  uint32_t* st_off;                 // n + 1
  uint32_t* st_ent;                 // impact << 2 | slot
  uint32_t kapazitaet_st_ent;
  V3<double>* st_dx;                // 5 per impact
  // counters (device side)
  uint32_t* zaehler;                // see SK_Z_* below
};

enum {
  SK_Z_NAEHE_PD = 0, SK_Z_NAEHE_KK, SK_Z_NAEHE, SK_Z_SCHNITTE, SK_Z_GEBIETE, SK_Z_FEDERN, SK_Z_STOESSE_PD, SK_Z_STOESSE,
  SK_Z_STOSS_ZAHL, SK_Z_UEBERLAUF, SK_Z_EINTRAEGE, SK_Z_FEDER_AUS, SK_Z_GROSS, SK_Z_ANZAHL
};

// ---------------------------------------------------------------------------
// geometry helpers (selbstkollision.rs, netzkoerper.rs)
// ---------------------------------------------------------------------------
typedef V3<double> D3;

__device__ __forceinline__ D3 sk_einheit(D3 a) {  // :196-203
  const double l = norm(a);
  return l > 1e-12 ? a * (1.0 / l) : D3{1.0, 0.0, 0.0};
}
__device__ __forceinline__ double sk_c(D3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }
__device__ __forceinline__ int64_t sk_zelle(double x, double zelle) { return (int64_t)floor(x / zelle); }  // :263-265

__device__ __forceinline__ bool sk_bereich_ok(const Quader& q, double zelle) {  // :270-272
  for (int a = 0; a < 3; ++a)
    if (!(isfinite(q.lo[a]) && isfinite(q.hi[a]) && (q.hi[a] - q.lo[a]) <= 64.0 * zelle)) return false;
  return true;
}
// u mod n by multiplication (Lemire, Kaser and Kurz, "Faster remainder by
// direct computation", 2019): with m = ceil(2^64 / n), u mod n is the high
// 64 bits of (m u mod 2^64) n, exactly, for every 32-bit u and n >= 1 (n = 1
// gives m = 0 and the remainder 0). The hash below takes its remainder this
// way, a few integer multiplies instead of the software division.
__device__ __forceinline__ uint64_t sk_rest_m(uint32_t n) { return ~0ull / (uint64_t)n + 1ull; }
__device__ __forceinline__ uint32_t sk_rest(uint32_t u, uint64_t m, uint32_t n) {
  return (uint32_t)__umul64hi(m * (uint64_t)u, (uint64_t)n);
}
__device__ __forceinline__ uint32_t sk_streue(int64_t i, int64_t j, int64_t k, uint32_t n, uint64_t m) {  // :274-278
  const uint32_t h = ((uint32_t)i * 0x8DA6B343u) ^ ((uint32_t)j * 0xD8163841u) ^ ((uint32_t)k * 0xCB1AB31Fu);
  return sk_rest(h, m, n);
}
__device__ __forceinline__ void sk_quader_rein(Quader& q, D3 p, double m) {  // one point of quader_punkte (:347-357)
  for (int a = 0; a < 3; ++a) {
    q.lo[a] = fmin(q.lo[a], sk_c(p, a) - m);
    q.hi[a] = fmax(q.hi[a], sk_c(p, a) + m);
  }
}
__device__ __forceinline__ Quader sk_quader_leer() {
  Quader q;
  for (int a = 0; a < 3; ++a) { q.lo[a] = INFINITY; q.hi[a] = -INFINITY; }
  return q;
}
__device__ __forceinline__ bool sk_ueberlappt(const Quader& a, const Quader& b) {  // :359-361
  for (int k = 0; k < 3; ++k)
    if (!(a.lo[k] <= b.hi[k] && b.lo[k] <= a.hi[k])) return false;
  return true;
}
__device__ __forceinline__ bool sk_erste_zelle(const Quader& a, const Quader& b, double zelle, const int64_t* c) {  // :365-367
  for (int k = 0; k < 3; ++k)
    if (sk_zelle(fmax(a.lo[k], b.lo[k]), zelle) != c[k]) return false;
  return true;
}
__device__ __forceinline__ Quader sk_kanten_quader(D3 a, D3 b, double m) {  // :469-484
  const D3 t = sk_einheit(b - a);
  const D3 h = fabs(t.y - 1.0) < 1e-5 ? D3{1.0, 0.0, 0.0} : D3{0.0, 1.0, 0.0};
  const D3 u = sk_einheit(cross(t, h));
  const D3 v = sk_einheit(cross(t, u));
  Quader q = sk_quader_leer();
  const D3 ps[2] = {a, b};
  const double ss[2] = {-m, m};
  for (int pi = 0; pi < 2; ++pi)
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) sk_quader_rein(q, (ps[pi] + u * ss[i]) + v * ss[j], 0.0);
  return q;
}
// netzkoerper.rs:185-218 (Ericson 2005, 5.1.5): the closest point
__device__ __forceinline__ D3 sk_naechster_auf_dreieck(D3 p, D3 a, D3 b, D3 c) {
  auto quot = [](double z, double n) { return n > 0.0 ? z / n : 0.0; };
  const D3 ab = b - a, ac = c - a, ap = p - a;
  const double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return a;
  const D3 bp = p - b;
  const double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return b;
  const double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return a + ab * quot(d1, d1 - d3);
  const D3 cp = p - c;
  const double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return c;
  const double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return a + ac * quot(d2, d2 - d6);
  const double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && d4 - d3 >= 0.0 && d5 - d6 >= 0.0) {
    const double w = quot(d4 - d3, (d4 - d3) + (d5 - d6));
    return b + (c - b) * w;
  }
  const double s = va + vb + vc;
  const double v = quot(vb, s), w = quot(vc, s);
  return a + (ab * v + ac * w);
}
__device__ __forceinline__ D3 sk_gewichte_dreieck(D3 p0, D3 p1, D3 p2, D3 x) {  // :370-383
  const D3 a = sk_naechster_auf_dreieck(x, p0, p1, p2);
  const D3 e1 = p1 - p0, e2 = p2 - p0, r = a - p0;
  const double d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
  const double det = d11 * d22 - d12 * d12;
  if (det <= 1e-30) return D3{1.0, 0.0, 0.0};
  const double r1 = dot(r, e1), r2 = dot(r, e2);
  const double w1 = (d22 * r1 - d12 * r2) / det;
  const double w2 = (d11 * r2 - d12 * r1) / det;
  return D3{1.0 - w1 - w2, w1, w2};
}
__device__ __forceinline__ D3 sk_lerp(D3 a, D3 b, double t) { return a * (1.0 - t) + b * t; }
__device__ __forceinline__ bool sk_im_dreieck(D3 a, D3 b, D3 c, D3 p, double* uv) {  // :1156-1171
  const D3 e1 = b - a, e2 = c - a, r = p - a;
  const double d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
  const double det = d11 * d22 - d12 * d12;
  if (det <= 1e-4 * d11 * d22) return false;
  const double r1 = dot(r, e1), r2 = dot(r, e2);
  const double u = (d22 * r1 - d12 * r2) / det;
  const double v = (d11 * r2 - d12 * r1) / det;
  if (u < 0.0 || v < 0.0 || u + v > 1.0) return false;
  uv[0] = u;
  uv[1] = v;
  return true;
}
__device__ __forceinline__ bool sk_kante_durch_dreieck(D3 a, D3 b, const D3* v, D3 n, double* tk, double* w) {  // :830-851
  const double d0 = dot(n, a - v[0]), d1 = dot(n, b - v[0]);
  if ((d0 > 0.0 && d1 > 0.0) || (d0 < 0.0 && d1 < 0.0) || fabs(d1 - d0) < 1e-12) return false;
  const double t = -d0 / (d1 - d0);
  if (fabs(t) < 1e-5 || fabs(t - 1.0) < 1e-5) return false;
  const D3 p = sk_lerp(a, b, t);
  const D3 e1 = v[1] - v[0], e2 = v[2] - v[0];
  if (norm(e1) < SK_STOSS_KANTE || norm(e2) < SK_STOSS_KANTE) return false;
  double uv[2];
  if (!sk_im_dreieck(v[0], v[1], v[2], p, uv)) return false;
  *tk = t;
  w[0] = 1.0 - uv[0] - uv[1];
  w[1] = uv[0];
  w[2] = uv[1];
  return true;
}
__device__ __forceinline__ double sk_ebenenabstand(D3 a, D3 b, D3 c, D3 p) {  // :1116-1119
  return dot(p - a, sk_einheit(cross(b - a, c - a)));
}
__device__ __forceinline__ bool sk_koplanar_zeit(const D3* x0, const D3* x1, double tol, double* out) {  // :1128-1153
  auto f = [&](double t) {
    D3 q[4];
    for (int k = 0; k < 4; ++k) q[k] = sk_lerp(x0[k], x1[k], t);
    return sk_ebenenabstand(q[0], q[1], q[2], q[3]);
  };
  double t0 = 0.0, t1 = 1.0;
  double d0 = f(0.0), d1 = f(1.0);
  for (int it = 0; it < 16; ++it) {
    if (d0 * d1 > 0.0) return false;
    const double tau = t0 + (t1 - t0) * fabs(d0) / fmax(fabs(d0) + fabs(d1), 2.2250738585072014e-308);
    const double d = f(tau);
    if (fabs(d) < tol || t1 - t0 < 1e-4) {
      *out = tau;
      return true;
    }
    if (d0 * d > 0.0) { t0 = tau; d0 = d; }
    else { t1 = tau; d1 = d; }
  }
  return false;
}
__device__ __forceinline__ bool sk_kanten_kreuzen(D3 a, D3 b, D3 c, D3 d, double* sw) {  // :1173-1193
  const D3 e1 = b - a, e2 = d - c;
  const double l1 = norm(e1), l2 = norm(e2);
  if (l1 < SK_STOSS_KANTE || l2 < SK_STOSS_KANTE) return false;
  const D3 u1 = e1 * (1.0 / l1), u2 = e2 * (1.0 / l2);
  const double g = dot(u1, u2);
  if (fabs(g) > 0.99999) return false;
  const D3 q = c - a;
  const double nen = 1.0 - g * g;
  const double s = (dot(q, u1) - g * dot(q, u2)) / nen;
  const double w = (g * dot(q, u1) - dot(q, u2)) / nen;
  if (s <= 0.0 || s >= l1 || w <= 0.0 || w >= l2) return false;
  sw[0] = s / l1;
  sw[1] = w / l2;
  return true;
}

__device__ __forceinline__ bool sk_partner(const SelbstDev& s, uint32_t a, uint32_t b) {  // :193-195
  if (a == b) return true;
  for (uint32_t k = s.partner_off[a]; k < s.partner_off[a + 1]; ++k)
    if (s.partner[k] == b) return true;
  return false;
}

// ---------------------------------------------------------------------------
// hash (:256-344): bucket count, per-entry dedup of buckets, sorted buckets
// ---------------------------------------------------------------------------
// The cell range of a box the hash accepts (at most 65 cells per axis, so
// at most 65^3 cells: the counts fit an int).
__device__ __forceinline__ bool sk_zellen(const Quader& q, double zelle, int64_t* lo, int64_t* hi) {
  if (!sk_bereich_ok(q, zelle)) return false;
  for (int a = 0; a < 3; ++a) {
    lo[a] = sk_zelle(q.lo[a], zelle);
    hi[a] = sk_zelle(q.hi[a], zelle);
  }
  return true;
}

// Calls f(bucket) once per distinct bucket of the entry's cells (the reference
// sorts and dedups the bucket list; the set is the same, its order does not
// matter because each bucket only receives the entry index).
// An entry with more than SK_KLEIN_ZELLEN cells is left to k_sk_gross; the
// counting pass (zaehlen) puts it on the list of such entries.
template <typename F>
__device__ __forceinline__ void sk_je_eimer(const SelbstDev& s, const Quader& q, double zelle, uint32_t eimer, uint64_t m,
                                            uint32_t e, bool zaehlen, F f) {
  int64_t lo[3], hi[3];
  if (!sk_zellen(q, zelle, lo, hi)) return;
  const int nc = (int)((hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1));
  if (nc > SK_KLEIN_ZELLEN) {
    if (zaehlen) {
      const uint32_t pos = atomicAdd(&s.zaehler[SK_Z_GROSS], 1u);
      if (pos < s.kapazitaet_gross) s.gross[pos] = e;
      else atomicAdd(&s.zaehler[SK_Z_UEBERLAUF], 1u);
    }
    return;
  }
  // the offsets are ints, stepped k fastest
  const int ny = (int)(hi[1] - lo[1] + 1), nz = (int)(hi[2] - lo[2] + 1);
  int i = 0, j = 0, k = 0;
  for (int c = 0; c < nc; ++c) {
    const uint32_t h = sk_streue(lo[0] + i, lo[1] + j, lo[2] + k, eimer, m);
    bool neu = true;
    int i2 = 0, j2 = 0, k2 = 0;
    for (int d = 0; d < c && neu; ++d) {
      neu = sk_streue(lo[0] + i2, lo[1] + j2, lo[2] + k2, eimer, m) != h;
      if (++k2 == nz) {
        k2 = 0;
        if (++j2 == ny) { j2 = 0; ++i2; }
      }
    }
    if (neu) f(h);
    if (++k == nz) {
      k = 0;
      if (++j == ny) { j = 0; ++i; }
    }
  }
}

__global__ void k_sk_eimer_zaehlen(const SelbstDev* __restrict__ sp, int anzahl, const double* zelle_p, uint32_t eimer) {
  const SelbstDev& s = *sp;
  const double zelle = *zelle_p;
  const uint64_t m = sk_rest_m(eimer);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < anzahl; e += gridDim.x * blockDim.x)
    sk_je_eimer(s, s.quader[e], zelle, eimer, m, (uint32_t)e, true, [&](uint32_t h) { atomicAdd(&s.eimer_fuell[h], 1u); });
}

__global__ void k_sk_eimer_fuellen(const SelbstDev* __restrict__ sp, int anzahl, const double* zelle_p, uint32_t eimer) {
  const SelbstDev& s = *sp;
  const double zelle = *zelle_p;
  if (s.zaehler[SK_Z_EINTRAEGE] > s.kapazitaet_eintraege) return;
  const uint64_t m = sk_rest_m(eimer);
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < anzahl; e += gridDim.x * blockDim.x)
    sk_je_eimer(s, s.quader[e], zelle, eimer, m, (uint32_t)e, false, [&](uint32_t h) {
      const uint32_t pos = atomicAdd(&s.eimer_fuell[h], 1u);
      s.eintraege_roh[pos] = (uint32_t)e;
    });
}

// The entries with more than SK_KLEIN_ZELLEN cells, one after another, by one
// block: the bucket of every cell sets its bit in a bitmap over the buckets,
// and the cell that sets it first counts the entry in that bucket (fuellen =
// 0) or puts it there (fuellen = 1). Each bucket receives the entry once, as
// with the reference's sorted and deduplicated bucket list; the order inside
// a bucket is fixed afterwards by k_sk_eimer_sortieren.
__global__ void __launch_bounds__(1024) k_sk_gross(const SelbstDev* __restrict__ sp, const double* zelle_p, uint32_t eimer,
                                                   int fuellen) {
  const SelbstDev& s = *sp;
  const double zelle = *zelle_p;
  if (fuellen && s.zaehler[SK_Z_EINTRAEGE] > s.kapazitaet_eintraege) return;
  const uint32_t ng = min(s.zaehler[SK_Z_GROSS], s.kapazitaet_gross);
  const uint32_t woerter = (eimer + 31u) / 32u;
  const uint64_t m = sk_rest_m(eimer);
  for (uint32_t g = 0; g < ng; ++g) {
    const uint32_t e = s.gross[g];
    int64_t lo[3], hi[3];
    (void)sk_zellen(s.quader[e], zelle, lo, hi);  // accepted by the counting pass
    const int ny = (int)(hi[1] - lo[1] + 1), nz = (int)(hi[2] - lo[2] + 1);
    const int nc = (int)(hi[0] - lo[0] + 1) * ny * nz;
    for (uint32_t w = threadIdx.x; w < woerter; w += blockDim.x)
      __hip_atomic_store(&s.eimer_bits[w], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
      const uint32_t h = sk_streue(lo[0] + c / (ny * nz), lo[1] + (c / nz) % ny, lo[2] + c % nz, eimer, m);
      const uint32_t bit = 1u << (h & 31u);
      if (atomicOr(&s.eimer_bits[h >> 5], bit) & bit) continue;
      const uint32_t pos = atomicAdd(&s.eimer_fuell[h], 1u);
      if (fuellen) s.eintraege_roh[pos] = e;
    }
    __syncthreads();
  }
}

// Entries of a bucket in index order, as the reference's sequential fill
// leaves them: every filled slot finds its bucket (binary search over the
// bucket starts) and its rank among the bucket's entries (entries are
// distinct within a bucket). Parallel over slots, so a bucket that grows
// large (a coarse cell) costs its size per slot, not its size squared in one
// thread.
__global__ void k_sk_eimer_sortieren(const SelbstDev* __restrict__ sp, uint32_t eimer) {
  const SelbstDev& s = *sp;
  const uint32_t gesamt = s.zaehler[SK_Z_EINTRAEGE];
  if (gesamt > s.kapazitaet_eintraege) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&s.zaehler[SK_Z_UEBERLAUF], 1u);
    return;
  }
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < gesamt; p += gridDim.x * blockDim.x) {
    // last bucket h with anfang[h] <= p (empty buckets share their start)
    uint32_t lo = 0, hi = eimer;  // anfang[lo] <= p < anfang[hi]
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (s.eimer_anfang[mid] <= p) lo = mid;
      else hi = mid;
    }
    const uint32_t a = s.eimer_anfang[lo], b = s.eimer_anfang[lo + 1];
    const uint32_t v = s.eintraege_roh[p];
    uint32_t rang = 0;
    for (uint32_t q = a; q < b; ++q) rang += s.eintraege_roh[q] < v ? 1u : 0u;
    s.eintraege[a + rang] = v;
  }
}

// Every entry in the cells of q for one query box, in the reference's order
// (cells by i, then j, then k; bucket contents in index order), walked by a
// whole wave: the bucket ranges of the cells are concatenated in that order
// and handed out 64 at a time, lane l taking entry r + l of round r. Entries
// with index <= ab are left out (buckets are sorted, so they are a prefix
// found by bisection). g(aktiv, entry, cell) runs on every lane of the wave
// once per round, aktiv false on lanes past the end, so g may use wave
// operations. q, zelle and ab must be the same on all 64 lanes, and all 64
// lanes must call.
template <typename G>
__device__ __forceinline__ void sk_welle_je_eintrag(const SelbstDev& s, const Quader& q, double zelle, uint32_t eimer, int64_t ab, G g) {
  const int lane = threadIdx.x & 63;
  if (s.zaehler[SK_Z_EINTRAEGE] > s.kapazitaet_eintraege) return;  // overflow: reported, not read
  if (!sk_bereich_ok(q, zelle)) return;
  int64_t lo[3], hi[3];
  for (int a = 0; a < 3; ++a) {
    lo[a] = sk_zelle(q.lo[a], zelle);
    hi[a] = sk_zelle(q.hi[a], zelle);
  }
  // each extent is at most 65 cells (sk_bereich_ok), so the counts fit an int
  const int nx = (int)(hi[0] - lo[0] + 1), ny = (int)(hi[1] - lo[1] + 1), nz = (int)(hi[2] - lo[2] + 1);
  const int nc = nx * ny * nz;
  const int nyz = ny * nz;
  const uint64_t mr = sk_rest_m(eimer);
  for (int c0 = 0; c0 < nc; c0 += 64) {
    const int c = c0 + lane;
    uint32_t anf = 0, n = 0;
    if (c < nc) {
      const uint32_t h = sk_streue(lo[0] + c / nyz, lo[1] + (c / nz) % ny, lo[2] + c % nz, eimer, mr);
      uint32_t u = s.eimer_anfang[h];
      const uint32_t b = s.eimer_anfang[h + 1];
      uint32_t o = b;
      while (u < o) {
        const uint32_t m = (u + o) >> 1;
        if ((int64_t)s.eintraege[m] <= ab) u = m + 1;
        else o = m;
      }
      anf = u;
      n = b - u;
    }
    uint32_t inkl = n;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const uint32_t v = __shfl_up(inkl, o, 64);
      if (lane >= o) inkl += v;
    }
    const uint32_t gesamt = __shfl(inkl, 63, 64);
    const uint32_t exkl = inkl - n;
    for (uint32_t r = 0; r < gesamt; r += 64) {
      const uint32_t k = r + (uint32_t)lane;
      // the cell of candidate k: the number of lanes whose inclusive sum is <= k
      int z = 0;
#pragma unroll
      for (int st = 32; st > 0; st >>= 1)
        if (__shfl(inkl, z + st - 1, 64) <= k) z += st;
      const uint32_t a_z = __shfl(anf, z, 64), e_z = __shfl(exkl, z, 64);
      const bool aktiv = k < gesamt;
      uint32_t eintrag = 0;
      int64_t cz[3] = {0, 0, 0};
      if (aktiv) {
        eintrag = s.eintraege[a_z + (k - e_z)];
        const int cc = c0 + z;
        cz[0] = lo[0] + cc / nyz;
        cz[1] = lo[1] + (cc / nz) % ny;
        cz[2] = lo[2] + cc % nz;
      }
      g(aktiv, eintrag, (const int64_t*)cz);
    }
  }
}

// Folds the hits of one round into a wave's running earliest hit in lane
// order, with the rule of a single thread walking the list: a hit replaces
// the best unless best.zeit <= zeit. All 64 lanes call; the result is the
// same on every lane.
__device__ __forceinline__ void sk_welle_falten(bool treffer, double zeit, uint32_t wer, bool* hat, double* best_zeit, uint32_t* best_wer) {
  unsigned long long m = __ballot(treffer);
  while (m) {
    const int l = __ffsll(m) - 1;
    const double z = __shfl(zeit, l, 64);
    const uint32_t w = __shfl(wer, l, 64);
    if (!*hat || !(*best_zeit <= z)) {
      *hat = true;
      *best_zeit = z;
      *best_wer = w;
    }
    m &= m - 1ull;
  }
}

// Exclusive scan of `anzahl` counters (werte[anzahl] = total, optionally a
// copy and the total in a counter slot) over many blocks, for long arrays
// (the buckets of the hash, one count per edge or triangle): pass 1 sums each chunk of
// SK_SCAN_CHUNK counts, pass 2 scans the chunk sums in one block and writes
// the total, pass 3 scans inside each chunk from its offset. Every load and
// store is coalesced; integer sums, so the result is the serial scan's.
#define SK_SCAN_K 4
#define SK_SCAN_CHUNK (1024 * SK_SCAN_K)

// inclusive scan of one value per thread over a block of 1024 threads; tot:
// the block's sum (every thread)
__device__ __forceinline__ uint32_t sk_block_scan(uint32_t v, uint32_t* wsum, uint32_t* tot) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const uint32_t u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  __syncthreads();
  if (lane == 63) wsum[wave] = v;
  __syncthreads();
  if (wave == 0) {
    uint32_t w = lane < 16 ? wsum[lane] : 0u;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      const uint32_t u = __shfl_up(w, o, 64);
      if (lane >= o) w += u;
    }
    if (lane < 16) wsum[lane] = w;  // inclusive over waves
  }
  __syncthreads();
  *tot = wsum[15];
  return v + (wave > 0 ? wsum[wave - 1] : 0u);
}

__global__ void __launch_bounds__(1024) k_sk_scan_teil(const uint32_t* werte, int anzahl, uint32_t* teil) {
  __shared__ uint32_t wsum[16];
  const int basis = blockIdx.x * SK_SCAN_CHUNK;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SK_SCAN_K; ++k) {
    const int i = basis + k * 1024 + (int)threadIdx.x;
    if (i < anzahl) s += werte[i];
  }
  uint32_t tot;
  sk_block_scan(s, wsum, &tot);
  if (threadIdx.x == 0) teil[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(1024) k_sk_scan_teile(uint32_t* teil, int nteile, uint32_t* werte, int anzahl, uint32_t* kopie,
                                                         uint32_t* gesamt_ziel) {
  __shared__ uint32_t wsum[16];
  uint32_t lauf = 0;
  for (int b = 0; b < nteile; b += 1024) {
    const int i = b + (int)threadIdx.x;
    const uint32_t v = i < nteile ? teil[i] : 0u;
    uint32_t tot;
    const uint32_t inkl = sk_block_scan(v, wsum, &tot);
    if (i < nteile) teil[i] = lauf + inkl - v;
    lauf += tot;
    __syncthreads();  // wsum is reused by the next round
  }
  if (threadIdx.x == 0) {
    werte[anzahl] = lauf;
    if (kopie) kopie[anzahl] = lauf;
    if (gesamt_ziel) *gesamt_ziel = lauf;
  }
}

__global__ void __launch_bounds__(1024) k_sk_scan_fertig(uint32_t* werte, int anzahl, const uint32_t* teil, uint32_t* kopie) {
  __shared__ uint32_t w[SK_SCAN_CHUNK];
  __shared__ uint32_t wsum[16];
  const int basis = blockIdx.x * SK_SCAN_CHUNK;
#pragma unroll
  for (int k = 0; k < SK_SCAN_K; ++k) {
    const int j = k * 1024 + (int)threadIdx.x, i = basis + j;
    w[j] = i < anzahl ? werte[i] : 0u;
  }
  __syncthreads();
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SK_SCAN_K; ++k) s += w[threadIdx.x * SK_SCAN_K + k];
  uint32_t tot;
  uint32_t lauf = teil[blockIdx.x] + sk_block_scan(s, wsum, &tot) - s;
#pragma unroll
  for (int k = 0; k < SK_SCAN_K; ++k) {
    const uint32_t v = w[threadIdx.x * SK_SCAN_K + k];
    w[threadIdx.x * SK_SCAN_K + k] = lauf;
    lauf += v;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SK_SCAN_K; ++k) {
    const int j = k * 1024 + (int)threadIdx.x, i = basis + j;
    if (i < anzahl) {
      werte[i] = w[j];
      if (kopie) kopie[i] = w[j];
    }
  }
}

__global__ void k_sk_null(uint32_t* p, int anzahl) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < anzahl; i += gridDim.x * blockDim.x) p[i] = 0u;
}

// ---------------------------------------------------------------------------
// normals (:209-251; the point normals are not used by the searches)
// ---------------------------------------------------------------------------
__global__ void k_sk_normalen_dreieck(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < s.nt; t += gridDim.x * blockDim.x) {
    const uint32_t* d = s.dreiecke + 3 * t;
    s.n_dreieck[t] = sk_einheit(cross(s.x[d[1]] - s.x[d[0]], s.x[d[2]] - s.x[d[0]]));
  }
}
__global__ void k_sk_normalen_kante(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < s.ne; e += gridDim.x * blockDim.x) {
    const uint32_t t0 = s.kanten_dreiecke[2 * e], t1 = s.kanten_dreiecke[2 * e + 1];
    D3 m;
    if (t0 == SK_KEIN) m = D3{1.0, 0.0, 0.0};
    else if (t1 == SK_KEIN) m = s.n_dreieck[t0];
    else {
      const D3 a = s.n_dreieck[t0], b = s.n_dreieck[t1];
      m = dot(a, b) < 0.0 ? a - b : a + b;
    }
    s.kantenebene[e] = sk_einheit(cross(m, s.x[s.kanten[2 * e]] - s.x[s.kanten[2 * e + 1]]));
  }
}

// ---------------------------------------------------------------------------
// proximity (:400-592)
// ---------------------------------------------------------------------------
__device__ __forceinline__ double sk_zelle_naehe(const SelbstDev& s) { return fmax(2.0 * s.r_max + SK_ZP, 1e-4) * 2.0; }
// the cell of the proximity hash, as the bucket kernels read it
__global__ void k_sk_zelle_naehe(const SelbstDev* __restrict__ sp, double* aus) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *aus = sk_zelle_naehe(*sp);
}

__global__ void k_sk_quader_punkte(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < s.n; v += gridDim.x * blockDim.x) {
    Quader q = sk_quader_leer();
    sk_quader_rein(q, s.x[v], s.r_punkt[v] + SK_ZQ);
    s.quader[v] = q;
  }
}

// Writes the lanes' pairs of one round at basis + *z in lane order and adds
// the round's count to *z; schreiben = false only counts. All 64 lanes call.
template <typename R>
__device__ __forceinline__ void sk_welle_schreiben(bool ok, bool schreiben, const R& o, R* ziel, uint32_t kapazitaet, uint32_t basis,
                                                   uint32_t* z) {
  const unsigned long long m = __ballot(ok);
  if (ok && schreiben) {
    const uint32_t pos = basis + *z + (uint32_t)__popcll(m & ((1ull << (threadIdx.x & 63)) - 1ull));
    if (pos < kapazitaet) ziel[pos] = o;
  }
  *z += (uint32_t)__popcll(m);
}

// One triangle's point-triangle pairs, walked by a whole wave
// (sk_welle_je_eintrag): the pairs are numbered in the order of the
// candidates, as a single thread walking them would number them. schreiben =
// false only counts. All 64 lanes call with the same t.
__device__ void sk_punkt_dreieck(const SelbstDev& s, int t, uint32_t eimer, bool schreiben, uint32_t basis, uint32_t* anzahl) {
  const uint32_t* d = s.dreiecke + 3 * t;
  const uint32_t i0 = d[0], i1 = d[1], i2 = d[2];
  uint32_t z = 0;
  if (!(i0 == i1 || i1 == i2 || i0 == i2)) {
    const D3 p[3] = {s.x[i0], s.x[i1], s.x[i2]};
    const double rt = s.r_dreieck[t];
    Quader qt = sk_quader_leer();
    for (int k = 0; k < 3; ++k) sk_quader_rein(qt, p[k], rt + SK_ZQ);
    const D3 nt = s.n_dreieck[t];
    D3 en[3];
    for (int k = 0; k < 3; ++k) en[k] = s.kantenebene[s.dreieck_kanten[3 * t + k]];
    const double sg[3] = {dot(p[2] - p[0], en[0]), dot(p[0] - p[1], en[1]), dot(p[1] - p[2], en[2])};
    const double zelle = sk_zelle_naehe(s);
    auto paar = [&](uint32_t v, const int64_t* c, Naehe& o) -> bool {
      if (sk_partner(s, v, i0) || sk_partner(s, v, i1) || sk_partner(s, v, i2)) return false;
      if (s.w[v] <= 0.0 && s.w[i0] <= 0.0 && s.w[i1] <= 0.0 && s.w[i2] <= 0.0) return false;
      const Quader qv = s.quader[v];
      if (!sk_ueberlappt(qv, qt) || !sk_erste_zelle(qt, qv, zelle, c)) return false;
      const D3 xv = s.x[v];
      const double ruhe = rt + s.r_punkt[v];  // no folded allowance bands here (band_abstand)
      const double soll = ruhe + SK_ZP;
      const double sd = dot(nt, xv - p[0]);
      if (fabs(sd) > soll) return false;
      for (int k = 0; k < 3; ++k)
        if (dot(xv - p[k], en[k]) * sg[k] < 0.0) return false;
      const D3 g = sk_gewichte_dreieck(p[0], p[1], p[2], xv);
      const D3 cp = (p[0] * g.x + p[1] * g.y) + p[2] * g.z;
      const D3 r = xv - cp;
      const double abstand = norm(r);
      if (abstand > soll) return false;
      o.art = 0;
      o.element = (uint32_t)t;
      o.element2 = SK_KEIN;
      o.punkte[0] = v; o.punkte[1] = i0; o.punkte[2] = i1; o.punkte[3] = i2;
      o.gewichte[0] = 1.0; o.gewichte[1] = -g.x; o.gewichte[2] = -g.y; o.gewichte[3] = -g.z;
      o.normale = abstand > 1e-12 ? r * (1.0 / abstand) : nt;
      o.abstand = abstand;
      o.ruhe = ruhe;
      return true;
    };
    sk_welle_je_eintrag(s, qt, zelle, eimer, -1, [&](bool aktiv, uint32_t v, const int64_t* c) {
      Naehe o;
      const bool ok = aktiv && paar(v, c, o);
      sk_welle_schreiben(ok, schreiben, o, s.naehe, s.kapazitaet_naehe, basis, &z);
    });
  }
  *anzahl = z;
}

// One wave per triangle.
__global__ void __launch_bounds__(256) k_sk_pd(const SelbstDev* __restrict__ sp, uint32_t eimer, int schreiben) {
  const SelbstDev& s = *sp;
  const int wellen = gridDim.x * (blockDim.x >> 6);
  for (int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < s.nt; t += wellen) {
    uint32_t z;
    sk_punkt_dreieck(s, t, eimer, schreiben != 0, schreiben ? s.zahl[t] : 0u, &z);
    if (!schreiben && (threadIdx.x & 63) == 0) s.zahl[t] = z;
  }
}

__global__ void k_sk_quader_kanten(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < s.ne; e += gridDim.x * blockDim.x)
    s.quader[e] = sk_kanten_quader(s.x[s.kanten[2 * e]], s.x[s.kanten[2 * e + 1]], s.r_kante[e] + SK_ZQ);
}

__device__ __forceinline__ bool sk_ist_kuerzeste(const SelbstDev& s, uint32_t e, D3 u, D3 m) {  // :489-504
  const uint32_t k0 = s.kanten[2 * e], k1 = s.kanten[2 * e + 1];
  for (int q = 0; q < 2; ++q) {
    const uint32_t t = s.kanten_dreiecke[2 * e + q];
    if (t == SK_KEIN) continue;
    const uint32_t* d = s.dreiecke + 3 * t;
    uint32_t o = SK_KEIN;
    for (int k = 0; k < 3; ++k)
      if (d[k] != k0 && d[k] != k1) { o = d[k]; break; }
    if (o == SK_KEIN) continue;
    const D3 r = s.x[o] - s.x[k0];
    const D3 wr = sk_einheit(r - u * dot(r, u));
    if (dot(m, wr) > 0.0 && fabs(dot(m, s.n_dreieck[t])) < 0.5) return false;
  }
  return true;
}

// One edge's edge-edge pairs with the edges f > e, walked by a whole wave as in
// sk_punkt_dreieck. All 64 lanes call with the same e.
__device__ void sk_kante_kante(const SelbstDev& s, int e, uint32_t eimer, bool schreiben, uint32_t basis, uint32_t* anzahl) {
  const uint32_t a0 = s.kanten[2 * e], a1 = s.kanten[2 * e + 1];
  uint32_t z = 0;
  if (a0 != a1) {
    const D3 pa0 = s.x[a0], pa1 = s.x[a1];
    const Quader qe = s.quader[e];
    const double zelle = sk_zelle_naehe(s);
    auto paar = [&](uint32_t f, const int64_t* c, Naehe& o) -> bool {
      const Quader qf = s.quader[f];
      if (!sk_ueberlappt(qe, qf) || !sk_erste_zelle(qe, qf, zelle, c)) return false;
      const uint32_t b0 = s.kanten[2 * f], b1 = s.kanten[2 * f + 1];
      if (sk_partner(s, a0, b0) || sk_partner(s, a1, b0) || sk_partner(s, a0, b1) || sk_partner(s, a1, b1)) return false;
      if (s.w[a0] <= 0.0 && s.w[a1] <= 0.0 && s.w[b0] <= 0.0 && s.w[b1] <= 0.0) return false;
      const D3 pb0 = s.x[b0], pb1 = s.x[b1];
      const D3 da = pa1 - pa0, db = pb1 - pb0;
      const double la = norm(da), lb = norm(db);
      if (la < SK_KANTE_MIN || lb < SK_KANTE_MIN) return false;
      const D3 ua = da * (1.0 / la), ub = db * (1.0 / lb);
      const double cw = dot(ua, ub);
      if (fabs(cw) > 0.99999) return false;
      const D3 r = pb0 - pa0;
      const double pp = dot(r, ua), qq = -dot(r, ub);
      const double nen = 1.0 - cw * cw;
      const double sa = (pp + cw * qq) / nen;
      const double sb = (qq + cw * pp) / nen;
      if (!(sa >= 0.0 && sa <= la) || !(sb >= 0.0 && sb <= lb)) return false;
      const D3 pa = pa0 + ua * sa;
      const D3 pb = pb0 + ub * sb;
      const D3 dv = pa - pb;
      const double abstand = norm(dv);
      const double ruhe = s.r_kante[e] + s.r_kante[f];
      if (abstand > ruhe + SK_ZP) return false;
      const D3 n = sk_einheit(dv);
      if (!sk_ist_kuerzeste(s, f, ub, n) || !sk_ist_kuerzeste(s, (uint32_t)e, ua, n * -1.0)) return false;
      const double sv = sa / la, tv = sb / lb;
      o.art = 1;
      o.element = (uint32_t)e;
      o.element2 = f;
      o.punkte[0] = a0; o.punkte[1] = a1; o.punkte[2] = b0; o.punkte[3] = b1;
      o.gewichte[0] = 1.0 - sv; o.gewichte[1] = sv; o.gewichte[2] = -(1.0 - tv); o.gewichte[3] = -tv;
      o.normale = n;
      o.abstand = abstand;
      o.ruhe = ruhe;
      return true;
    };
    // the partners f <= e, which the reference skips first, are left out of the walk
    sk_welle_je_eintrag(s, qe, zelle, eimer, (int64_t)e, [&](bool aktiv, uint32_t f, const int64_t* c) {
      Naehe o;
      const bool ok = aktiv && paar(f, c, o);
      sk_welle_schreiben(ok, schreiben, o, s.naehe, s.kapazitaet_naehe, basis, &z);
    });
  }
  *anzahl = z;
}

// One wave per edge.
__global__ void __launch_bounds__(256) k_sk_kk(const SelbstDev* __restrict__ sp, uint32_t eimer, int schreiben) {
  const SelbstDev& s = *sp;
  const uint32_t basis0 = s.zaehler[SK_Z_NAEHE_PD];
  const int wellen = gridDim.x * (blockDim.x >> 6);
  for (int e = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); e < s.ne; e += wellen) {
    uint32_t z;
    sk_kante_kante(s, e, eimer, schreiben != 0, schreiben ? basis0 + s.zahl[e] : 0u, &z);
    if (!schreiben && (threadIdx.x & 63) == 0) s.zahl[e] = z;
  }
}

// Totals of the proximity lists; a list beyond its capacity is an error of the
// step (the host reports it) and is cut to the capacity.
__global__ void k_sk_naehe_summe(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint32_t g = s.zaehler[SK_Z_NAEHE_PD] + s.zaehler[SK_Z_NAEHE_KK];
  if (g > s.kapazitaet_naehe) {
    s.zaehler[SK_Z_UEBERLAUF] += 1u;
    g = s.kapazitaet_naehe;
  }
  s.zaehler[SK_Z_NAEHE] = g;
  s.zaehler[SK_Z_FEDERN] = g;
}

// ---------------------------------------------------------------------------
// proximity springs (:713-787)
// ---------------------------------------------------------------------------
__global__ void k_sk_federn(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  const uint32_t nn = s.zaehler[SK_Z_NAEHE];
  for (uint32_t q0 = blockIdx.x * blockDim.x + threadIdx.x; q0 < nn; q0 += gridDim.x * blockDim.x) {
    const Naehe& p = s.naehe[q0];
    const uint32_t* idx = p.punkte;
    double e, a, mu;
    if (p.art == 0) {
      const uint32_t t = p.element;
      e = s.e_punkt[idx[0]] + s.e_dreieck[t];
      a = s.a_punkt[idx[0]] + s.a_dreieck[t];
      mu = 0.5 * (s.mu[idx[0]] + (s.mu[idx[1]] + s.mu[idx[2]] + s.mu[idx[3]]) / 3.0);
    } else {
      const uint32_t k1 = p.element, k2 = p.element2;
      e = s.e_kante[k1] + s.e_kante[k2];
      a = s.a_kante[k1] + s.a_kante[k2];
      mu = 0.25 * (s.mu[idx[0]] + s.mu[idx[1]] + s.mu[idx[2]] + s.mu[idx[3]]);
    }
    const double k = s.steifigkeit * e * a / 1e-6;
    double ruhe = p.ruhe;
    bool fertig = false;
    for (int q = 0; q < 4 && !fertig; ++q) {
      if (p.gewichte[q] <= 0.0) continue;
      for (uint32_t c = s.nahtw_off[idx[q]]; c < s.nahtw_off[idx[q] + 1]; ++c) {
        const uint32_t partner = s.nahtw[2 * c], naht = s.nahtw[2 * c + 1];
        bool auf_zweiter = false;
        for (int r = 0; r < 4; ++r) auf_zweiter = auf_zweiter || (p.gewichte[r] < 0.0 && idx[r] == partner);
        const double l0 = s.naht_ruhe[naht];
        if (auf_zweiter && isfinite(l0) && l0 < ruhe) {
          ruhe = l0;
          fertig = true;
          break;
        }
      }
    }
    const double cc = fmin(ruhe - p.abstand, SK_AUSLENKUNG);
    const D3 n = p.normale;
    s.fe_kraft[q0] = n * (k * cc);
    const double nv[3] = {n.x, n.y, n.z};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) s.fe_block[9 * q0 + 3 * i + j] = k * ((1.0 - mu) * nv[i] * nv[j] + (i == j ? mu : 0.0));
    for (int q = 0; q < 4; ++q) {
      s.fe_punkte[5 * q0 + q] = idx[q];
      s.fe_gew[5 * q0 + q] = p.gewichte[q];
    }
    s.fe_punkte[5 * q0 + 4] = idx[0];
    s.fe_gew[5 * q0 + 4] = 0.0;
  }
}

// Approach damping in list order (:762-781): each pair reads the velocities
// the pairs before it left, so the pairs are applied one after another by one
// thread. The block stages them in chunks: all threads load a chunk's pair
// records (mu, the points' mass shares, weights, normal, held flags) and put the velocities
// of the points they touch into an LDS hash table keyed by point; thread 0
// then applies the chunk's pairs in order on the table (the arithmetic and
// its order are the sequential loop's), and the block writes the touched
// velocities back before the next chunk reads them.
#define SK_DP_CHUNK 128
#define SK_DP_TAB 1024  // power of two, at least twice the chunk's 4 * SK_DP_CHUNK points
__global__ void __launch_bounds__(1024) k_sk_daempfung(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  __shared__ uint32_t tab_key[SK_DP_TAB];
  __shared__ D3 tab_v[SK_DP_TAB];
  __shared__ D3 p_n[SK_DP_CHUNK];
  __shared__ double p_g[SK_DP_CHUNK][4];
  __shared__ double p_mu[SK_DP_CHUNK], p_anteil[SK_DP_CHUNK][4];
  __shared__ uint8_t p_massiv[SK_DP_CHUNK];
  __shared__ uint16_t p_pos[SK_DP_CHUNK][4];
  __shared__ uint8_t p_fest[SK_DP_CHUNK];
  const uint32_t nn = s.zaehler[SK_Z_NAEHE];
  const uint32_t leer = 0xFFFFFFFFu;
  for (uint32_t q0 = 0; q0 < nn; q0 += SK_DP_CHUNK) {
    const uint32_t anzahl = nn - q0 < SK_DP_CHUNK ? nn - q0 : SK_DP_CHUNK;
    for (uint32_t e = threadIdx.x; e < SK_DP_TAB; e += blockDim.x) tab_key[e] = leer;
    __syncthreads();
    // the chunk's points into the table (linear probing; at most 4 * SK_DP_CHUNK
    // distinct keys, so the table is at most half full)
    for (uint32_t t = threadIdx.x; t < 4 * anzahl; t += blockDim.x) {
      const uint32_t pr = t >> 2, q = t & 3u;
      const uint32_t i = s.naehe[q0 + pr].punkte[q];
      uint32_t e = (i * 2654435761u) & (SK_DP_TAB - 1);
      for (;;) {
        const uint32_t alt = atomicCAS(&tab_key[e], leer, i);
        if (alt == leer || alt == i) break;
        e = (e + 1) & (SK_DP_TAB - 1);
      }
      p_pos[pr][q] = (uint16_t)e;
    }
    // the per-pair values that do not depend on velocities (same expressions
    // and order as the sequential loop)
    for (uint32_t pr = threadIdx.x; pr < anzahl; pr += blockDim.x) {
      const Naehe& p = s.naehe[q0 + pr];
      double mu;
      if (p.art == 0) mu = 0.5 * (s.mu[p.punkte[0]] + (s.mu[p.punkte[1]] + s.mu[p.punkte[2]] + s.mu[p.punkte[3]]) / 3.0);
      else mu = 0.25 * (s.mu[p.punkte[0]] + s.mu[p.punkte[1]] + s.mu[p.punkte[2]] + s.mu[p.punkte[3]]);
      double m1 = 0.0, m2 = 0.0;
      uint8_t fest = 0;
      for (int q = 0; q < 4; ++q) {
        const double g = p.gewichte[q];
        p_g[pr][q] = g;
        if (g > 0.0) m1 += g * s.masse[p.punkte[q]];
        else m2 -= g * s.masse[p.punkte[q]];
        if (s.w[p.punkte[q]] <= 0.0) fest |= (uint8_t)(1u << q);
      }
      // each point's share of the correction: the quotients of the sequential
      // loop, taken here so that thread 0's loop holds no division
      for (int q = 0; q < 4; ++q) p_anteil[pr][q] = p_g[pr][q] > 0.0 ? m2 / (m1 + m2) : m1 / (m1 + m2);
      p_mu[pr] = mu;
      p_massiv[pr] = m1 + m2 > 0.0 ? 1 : 0;
      p_n[pr] = p.normale;
      p_fest[pr] = fest;
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < SK_DP_TAB; e += blockDim.x)
      if (tab_key[e] != leer) tab_v[e] = s.v[tab_key[e]];
    __syncthreads();
    if (threadIdx.x == 0)
      for (uint32_t pr = 0; pr < anzahl; ++pr) {
        const D3 n = p_n[pr];
        D3 dv = zero3<double>();
        for (int q = 0; q < 4; ++q) dv = dv + tab_v[p_pos[pr][q]] * p_g[pr][q];
        const double an = dot(dv, n) * SK_ANNAEHERUNG;
        if (an < 0.0 && p_massiv[pr]) {
          const D3 u = n * an + (dv - n * an) * p_mu[pr];
          for (int q = 0; q < 4; ++q) {
            if (p_fest[pr] & (1u << q)) continue;
            const uint16_t e = p_pos[pr][q];
            tab_v[e] = tab_v[e] - u * (p_g[pr][q] * p_anteil[pr][q]);
          }
        }
      }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < SK_DP_TAB; e += blockDim.x)
      if (tab_key[e] != leer) s.v[tab_key[e]] = tab_v[e];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// cuts, regions, push springs (:858-1059)
// ---------------------------------------------------------------------------
__global__ void k_sk_quader_dreiecke(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < s.nt; t += gridDim.x * blockDim.x) {
    Quader q = sk_quader_leer();
    for (int k = 0; k < 3; ++k) sk_quader_rein(q, s.x[s.dreiecke[3 * t + k]], SK_ZQ);
    s.quader[t] = q;
  }
}

// Longest edge at positions xs (one block; `aus` gets max(l, untergrenze)).
// Longest edge, at least `untergrenze`: *aus is set to untergrenze first
// (k_sk_laengste_start), then every block folds its maximum in with an atomic
// max on the bits. The values are non-negative and never NaN (fmax drops a NaN
// length, as the serial fold does), and for such doubles the order of the bit
// patterns is the order of the values, so the result does not depend on the
// order of the blocks.
__global__ void k_sk_laengste_start(double untergrenze, double* aus) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *aus = untergrenze;
}
__global__ void __launch_bounds__(256) k_sk_laengste(const SelbstDev* __restrict__ sp, const V3<double>* xs, double* aus) {
  const SelbstDev& s = *sp;
  __shared__ double m[256];
  double l = 0.0;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < s.ne; e += gridDim.x * blockDim.x)
    l = fmax(l, norm(xs[s.kanten[2 * e]] - xs[s.kanten[2 * e + 1]]));
  m[threadIdx.x] = l;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) m[threadIdx.x] = fmax(m[threadIdx.x], m[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax((unsigned long long*)aus, (unsigned long long)__double_as_longlong(m[0]));
}

// One edge's cuts through triangles, walked by a whole wave as in
// sk_punkt_dreieck. All 64 lanes call with the same e.
__device__ void sk_schnitte_kante(const SelbstDev& s, int e, double zelle, uint32_t eimer, bool schreiben, uint32_t basis,
                                  uint32_t* anzahl) {
  const uint32_t a0 = s.kanten[2 * e], a1 = s.kanten[2 * e + 1];
  const uint32_t kd0 = s.kanten_dreiecke[2 * e], kd1 = s.kanten_dreiecke[2 * e + 1];
  uint32_t z = 0;
  if (!(a0 == a1 || kd0 == SK_KEIN)) {
    const D3 pa = s.x[a0], pb = s.x[a1];
    Quader qe = sk_quader_leer();
    sk_quader_rein(qe, pa, SK_ZQ);
    sk_quader_rein(qe, pb, SK_ZQ);
    const D3 ehat = sk_einheit(pb - pa);
    auto schnitt = [&](uint32_t t, const int64_t* c, Schnitt& o) -> bool {
      const uint32_t* d = s.dreiecke + 3 * t;
      bool nachbar;
      if (s.an_naehten) {
        nachbar = d[0] == a0 || d[1] == a0 || d[2] == a0 || d[0] == a1 || d[1] == a1 || d[2] == a1;
      } else {
        nachbar = false;
        for (int k = 0; k < 3; ++k) nachbar = nachbar || sk_partner(s, a0, d[k]) || sk_partner(s, a1, d[k]);
      }
      const Quader qt = s.quader[t];
      if (nachbar || !sk_ueberlappt(qt, qe) || !sk_erste_zelle(qe, qt, zelle, c)) return false;
      const D3 n = s.n_dreieck[t];
      const D3 v[3] = {s.x[d[0]], s.x[d[1]], s.x[d[2]]};
      double tk, w[3];
      if (!sk_kante_durch_dreieck(pa, pb, v, n, &tk, w)) return false;
      const double cn = dot(n, ehat);
      if (fabs(cn) < 1e-6) return false;
      if (schreiben) {
        const D3 p = sk_lerp(pa, pb, tk);
        D3 g = zero3<double>();
        const uint32_t kd[2] = {kd0, kd1};
        for (int q = 0; q < 2; ++q) {
          const uint32_t k = kd[q];
          if (k == SK_KEIN) continue;
          const uint32_t* dk = s.dreiecke + 3 * k;
          uint32_t o = SK_KEIN;
          for (int j = 0; j < 3; ++j)
            if (dk[j] != a0 && dk[j] != a1) { o = dk[j]; break; }
          if (o == SK_KEIN) continue;
          D3 m = sk_einheit(cross(n, s.n_dreieck[k]));
          const D3 r = s.x[o] - p;
          const D3 r_quer = r - ehat * dot(ehat, r);
          if (dot(m, r_quer) < 0.0) m = m * -1.0;
          g = g + (m - n * (dot(ehat, m) / cn));
        }
        const double lg = norm(g);
        o.kante = (uint32_t)e;
        o.dreieck = t;
        o.t = tk;
        o.w[0] = w[0]; o.w[1] = w[1]; o.w[2] = w[2];
        o.richtung = sk_einheit(g);
        o.gewicht = lg / sqrt(lg * lg + 2.25);
      }
      return true;
    };
    sk_welle_je_eintrag(s, qe, zelle, eimer, -1, [&](bool aktiv, uint32_t t, const int64_t* c) {
      Schnitt o;
      const bool ok = aktiv && schnitt(t, c, o);
      sk_welle_schreiben(ok, schreiben, o, s.schnitte, s.kapazitaet_schnitte, basis, &z);
    });
  }
  *anzahl = z;
}

// One wave per edge.
__global__ void __launch_bounds__(256) k_sk_schnitte(const SelbstDev* __restrict__ sp, const double* zelle_p, uint32_t eimer,
                                                     int schreiben) {
  const SelbstDev& s = *sp;
  const double zelle = *zelle_p;
  const int wellen = gridDim.x * (blockDim.x >> 6);
  for (int e = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); e < s.ne; e += wellen) {
    uint32_t z;
    sk_schnitte_kante(s, e, zelle, eimer, schreiben != 0, schreiben ? s.zahl[e] : 0u, &z);
    if (!schreiben && (threadIdx.x & 63) == 0) s.zahl[e] = z;
  }
}

// Index of the cut (e, t), or SK_KEIN; cuts of one edge are contiguous at
// zahl[e] .. zahl[e + 1] (the scanned per-edge counts).
__device__ __forceinline__ uint32_t sk_schnitt_von(const SelbstDev& s, uint32_t e, uint32_t t) {
  const uint32_t b = min(s.zahl[e + 1], s.kapazitaet_schnitte);
  for (uint32_t i = s.zahl[e]; i < b; ++i)
    if (s.schnitte[i].dreieck == t) return i;
  return SK_KEIN;
}
__global__ void k_sk_schnitte_summe(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  if (s.zaehler[SK_Z_SCHNITTE] > s.kapazitaet_schnitte) {
    s.zaehler[SK_Z_UEBERLAUF] += 1u;
    s.zaehler[SK_Z_SCHNITTE] = s.kapazitaet_schnitte;
  }
}
__device__ __forceinline__ uint32_t sk_wurzel(uint32_t* eltern, uint32_t i) {  // :910-916
  while (eltern[i] != i) {
    eltern[i] = eltern[eltern[i]];
    i = eltern[i];
  }
  return i;
}
__device__ __forceinline__ void sk_vereinige(uint32_t* eltern, uint32_t a, uint32_t b) {  // :918-924
  const uint32_t ra = sk_wurzel(eltern, a), rb = sk_wurzel(eltern, b);
  if (ra != rb) eltern[max(ra, rb)] = min(ra, rb);
}

// gebiete (:931-1007) in list order, one thread; the region direction sum in
// list order. Leaves the root of every cut in schnitt_wurzel.
__global__ void k_sk_gebiete(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const uint32_t n = s.zaehler[SK_Z_SCHNITTE];  // clamped to the capacity
  uint32_t* eltern = s.uf;
  uint32_t* partner = s.uf + n;
  for (uint32_t i = 0; i < n; ++i) eltern[i] = i;
  uint32_t nt = 0;
  bool voll = false;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t e = s.schnitte[i].kante, t = s.schnitte[i].dreieck;
    uint32_t kdr[2];
    int nkd = 0;
    for (int q = 0; q < 2; ++q)
      if (s.kanten_dreiecke[2 * e + q] != SK_KEIN) kdr[nkd++] = s.kanten_dreiecke[2 * e + q];
    uint32_t andere[6];
    int na = 0;
    for (int q = 0; q < nkd; ++q)
      for (int j = 0; j < 3; ++j) {
        const uint32_t f = s.dreieck_kanten[3 * kdr[q] + j];
        if (f != e) andere[na++] = f;
      }
    uint32_t tn[6];
    int ntn = 0;
    for (int j = 0; j < 3; ++j) {
      const uint32_t k = s.dreieck_kanten[3 * t + j];
      for (int q = 0; q < 2; ++q) {
        const uint32_t u = s.kanten_dreiecke[2 * k + q];
        if (u != SK_KEIN && u != t) tn[ntn++] = u;
      }
    }
    // "gleich": other edges of e's triangles through t, e through t's
    // neighbours, other edges through t's neighbours (in that order)
    for (int a = 0; a < na; ++a) {
      const uint32_t j = sk_schnitt_von(s, andere[a], t);
      if (j != SK_KEIN) sk_vereinige(eltern, i, j);
    }
    for (int u = 0; u < ntn; ++u) {
      uint32_t j = sk_schnitt_von(s, e, tn[u]);
      if (j != SK_KEIN) sk_vereinige(eltern, i, j);
      for (int a = 0; a < na; ++a) {
        j = sk_schnitt_von(s, andere[a], tn[u]);
        if (j != SK_KEIN) sk_vereinige(eltern, i, j);
      }
    }
    // role swap: an edge of t through a triangle of e
    for (int j3 = 0; j3 < 3; ++j3)
      for (int q = 0; q < nkd; ++q) {
        const uint32_t j = sk_schnitt_von(s, s.dreieck_kanten[3 * t + j3], kdr[q]);
        if (j != SK_KEIN) {
          if (nt < s.kapazitaet_tausch) {
            s.tausch[2 * nt] = i;
            s.tausch[2 * nt + 1] = j;
            ++nt;
          } else {
            voll = true;
          }
        }
      }
  }
  if (voll) atomicAdd(&s.zaehler[SK_Z_UEBERLAUF], 1u);
  for (int durchgang = 0; durchgang < 2; ++durchgang) {
    for (uint32_t i = 0; i < n; ++i) partner[i] = SK_KEIN;
    for (uint32_t q = 0; q < nt; ++q) {
      const uint32_t ri = sk_wurzel(eltern, s.tausch[2 * q]), rj = sk_wurzel(eltern, s.tausch[2 * q + 1]);
      partner[ri] = min(partner[ri], rj);
      partner[rj] = min(partner[rj], ri);
    }
    for (uint32_t q = 0; q < nt; ++q) {
      const uint32_t rj = sk_wurzel(eltern, s.tausch[2 * q + 1]);
      if (partner[rj] != SK_KEIN) sk_vereinige(eltern, s.tausch[2 * q], partner[rj]);
    }
  }
  uint32_t gebiete = 0;
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t r = sk_wurzel(eltern, i);
    s.schnitt_wurzel[i] = r;
    gebiete += r == i ? 1u : 0u;
    s.summe[i] = zero3<double>();
  }
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t r = s.schnitt_wurzel[i];
    s.summe[r] = s.summe[r] + s.schnitte[i].richtung * s.schnitte[i].gewicht;
  }
  for (uint32_t i = 0; i < n; ++i) s.schnitte[i].richtung = sk_einheit(s.summe[s.schnitt_wurzel[i]]);
  s.zaehler[SK_Z_GEBIETE] = gebiete;
}

// Body normal at a point, first body in contact (implizit.rs:699).
template <typename T>
__device__ __forceinline__ bool sk_koerper_normale(const SysDev<T>& sy, uint32_t i, D3* out) {
  for (int b = 0; b < sy.n_koerper; ++b) {
    D3 nor, vk;
    double tiefe;
    if (kontakt(sy.koerper[b], sy.x[i], sy.koll[i] + sy.kontakt_rand, &nor, &tiefe, &vk)) {
      *out = nor;
      return true;
    }
  }
  return false;
}

// schubfedern (:1014-1059), one thread in list order: clamped regions,
// the per-triangle run counter, push springs after the proximity springs.
template <typename T>
__global__ void k_sk_schub(const SelbstDev* __restrict__ sp, const SysDev<T>* __restrict__ syp) {
  const SelbstDev& s = *sp;
  const SysDev<T>& sy = *syp;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const uint32_t n = s.zaehler[SK_Z_SCHNITTE];
  uint32_t* klemm = s.uf + n;  // reuse: one flag per root
  for (uint32_t i = 0; i < n; ++i) klemm[i] = 0u;
  for (uint32_t i = 0; i < n; ++i) {
    const Schnitt& c = s.schnitte[i];
    const uint32_t kp[2] = {s.kanten[2 * c.kante], s.kanten[2 * c.kante + 1]};
    const uint32_t* dp = s.dreiecke + 3 * c.dreieck;
    D3 summe = zero3<double>();
    for (int q = 0; q < 3; ++q) {
      D3 nn;
      if (sk_koerper_normale(sy, dp[q], &nn)) summe = summe + nn;
    }
    bool eingeklemmt = false;
    if (norm(summe) > 0.0)
      for (int q = 0; q < 2 && !eingeklemmt; ++q) {
        D3 nk;
        if (sk_koerper_normale(sy, kp[q], &nk) && dot(nk, summe) < 0.0) eingeklemmt = true;
      }
    if (eingeklemmt) klemm[s.schnitt_wurzel[i]] = 1u;
  }
  uint32_t f = s.zaehler[SK_Z_NAEHE];
  for (uint32_t i = 0; i < n; ++i) {
    if (klemm[s.schnitt_wurzel[i]]) continue;
    const Schnitt& c = s.schnitte[i];
    const uint32_t e = c.kante, t = c.dreieck;
    if (f >= s.kapazitaet_federn) {
      atomicAdd(&s.zaehler[SK_Z_UEBERLAUF], 1u);
      break;
    }
    const double k = 2.0 * s.steifigkeit * (s.e_kante[e] + s.e_dreieck[t]) * (s.a_kante[e] + s.a_dreieck[t]) / 1e-6;
    const D3 g = c.richtung;
    // the run counter is read after this step's update (below, per triangle)
    const uint32_t folge = s.folge[t] + 1u;
    s.fe_kraft[f] = g * (k * (double)min(folge, SK_FOLGE_MAX) * SK_HEBEL);
    const double gv[3] = {g.x, g.y, g.z};
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) s.fe_block[9 * f + 3 * a + b] = k * gv[a] * gv[b];
    s.fe_punkte[5 * f + 0] = s.kanten[2 * e];
    s.fe_punkte[5 * f + 1] = s.kanten[2 * e + 1];
    for (int q = 0; q < 3; ++q) s.fe_punkte[5 * f + 2 + q] = s.dreiecke[3 * t + q];
    s.fe_gew[5 * f + 0] = 1.0 - c.t;
    s.fe_gew[5 * f + 1] = c.t;
    for (int q = 0; q < 3; ++q) s.fe_gew[5 * f + 2 + q] = -c.w[q];
    ++f;
  }
  s.zaehler[SK_Z_FEDERN] = f;
}

// folge (:1030-1033): triangles with a cut count up, the others start at 1.
__global__ void k_sk_folge(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  const uint32_t n = s.zaehler[SK_Z_SCHNITTE];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < s.nt; t += gridDim.x * blockDim.x) {
    bool hat = false;
    // cuts are few; a linear scan is cheaper than a per-triangle index
    for (uint32_t i = 0; i < n && !hat; ++i) hat = s.schnitte[i].dreieck == (uint32_t)t;
    s.folge[t] = hat ? s.folge[t] + 1u : 1u;
  }
}

// ---------------------------------------------------------------------------
// per-point spring lists (implizit.rs:63-70 pairs; force slots :800-806)
// ---------------------------------------------------------------------------
__global__ void k_sk_feder_zaehlen(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  const uint32_t nf = min(s.zaehler[SK_Z_FEDERN], s.kapazitaet_federn);
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
    const uint32_t* p = s.fe_punkte + 5 * f;
    const double* g = s.fe_gew + 5 * f;
    for (int q = 0; q < 5; ++q) {
      if (g[q] != 0.0) atomicAdd(&s.fk_off[p[q]], 1u);
      for (int r = 0; r < 5; ++r)
        if (g[q] > 0.0 && g[r] < 0.0) {
          atomicAdd(&s.fe_off[p[q]], 1u);
          atomicAdd(&s.fe_off[p[r]], 1u);
        }
    }
  }
}
// Lists beyond their capacity are not filled; their offsets are cleared
// afterwards, so no row reads them (the step reports the error).
__global__ void k_sk_feder_pruefen(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const bool aus = s.fe_off[s.n] > s.kapazitaet_fe_ent || s.fk_off[s.n] > s.kapazitaet_fk_ent;
  s.zaehler[SK_Z_FEDER_AUS] = aus ? 1u : 0u;
  if (aus) s.zaehler[SK_Z_UEBERLAUF] += 1u;
}
__global__ void k_sk_feder_leeren(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  if (!s.zaehler[SK_Z_FEDER_AUS]) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i <= s.n; i += gridDim.x * blockDim.x) {
    s.fe_off[i] = 0u;
    s.fk_off[i] = 0u;
  }
}
__global__ void k_sk_feder_fuellen(const SelbstDev* __restrict__ sp, uint32_t* fe_fuell, uint32_t* fk_fuell) {
  const SelbstDev& s = *sp;
  if (s.zaehler[SK_Z_FEDER_AUS]) return;
  const uint32_t nf = min(s.zaehler[SK_Z_FEDERN], s.kapazitaet_federn);
  for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < nf; f += gridDim.x * blockDim.x) {
    const uint32_t* p = s.fe_punkte + 5 * f;
    const double* g = s.fe_gew + 5 * f;
    for (int q = 0; q < 5; ++q) {
      if (g[q] != 0.0) s.fk_roh[atomicAdd(&fk_fuell[p[q]], 1u)] = (f << 3) | (uint32_t)q;
      for (int r = 0; r < 5; ++r)
        if (g[q] > 0.0 && g[r] < 0.0) {
          const uint32_t code = (f << 5) | (uint32_t)(5 * q + r);
          s.fe_roh[atomicAdd(&fe_fuell[p[q]], 1u)] = code;
          s.fe_roh[atomicAdd(&fe_fuell[p[r]], 1u)] = code;
        }
    }
  }
}
// Insertion sort of a[lo, hi) (short lists: the impacts of one point).
__device__ __forceinline__ void sk_sortiere(uint32_t* a, uint32_t lo, uint32_t hi) {
  for (uint32_t i = lo + 1; i < hi; ++i) {
    const uint32_t v = a[i];
    uint32_t j = i;
    while (j > lo && a[j - 1] > v) {
      a[j] = a[j - 1];
      --j;
    }
    a[j] = v;
  }
}

// Every entry of a per-point list to its place in ascending order: its
// rank among the entries of its list (the smaller ones, and the equal ones
// filled before it). Parallel over entries, so a point with a long list costs
// its length per entry, not its length squared in one thread. Equal entries
// are equal codes, so the lists are those of a sequential sort.
__device__ __forceinline__ void sk_rang_platz(const uint32_t* roh, uint32_t* ent, const uint32_t* off, int n, uint32_t p) {
  uint32_t u = 0, o = (uint32_t)n;  // the list: the last i with off[i] <= p
  while (o - u > 1) {
    const uint32_t m = (u + o) >> 1;
    if (off[m] <= p) u = m;
    else o = m;
  }
  const uint32_t a = off[u], b = off[u + 1], v = roh[p];
  uint32_t rang = 0;
  for (uint32_t q = a; q < b; ++q) rang += (roh[q] < v || (roh[q] == v && q < p)) ? 1u : 0u;
  ent[a + rang] = v;
}
__global__ void k_sk_feder_sortieren(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  if (s.zaehler[SK_Z_FEDER_AUS]) return;
  const uint32_t ne = s.fe_off[s.n], nk = s.fk_off[s.n];
  for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < ne + nk; p += gridDim.x * blockDim.x) {
    if (p < ne) sk_rang_platz(s.fe_roh, s.fe_ent, s.fe_off, s.n, p);
    else sk_rang_platz(s.fk_roh, s.fk_ent, s.fk_off, s.n, p - ne);
  }
}
// Operator records of the springs in the bulk type.
template <typename T>
__global__ void k_sk_feder_bloecke(const SelbstDev* __restrict__ sp, T* block_t) {
  const SelbstDev& s = *sp;
  const uint32_t nf = min(s.zaehler[SK_Z_FEDERN], s.kapazitaet_federn);
  for (uint32_t q = blockIdx.x * blockDim.x + threadIdx.x; q < 9 * nf; q += gridDim.x * blockDim.x) block_t[q] = T(s.fe_block[q]);
}

// ---------------------------------------------------------------------------
// impacts (:1196-1343)
// ---------------------------------------------------------------------------
__global__ void k_sk_kopie_x(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) s.x_anfang[i] = s.x[i];
}
__global__ void k_sk_quader_weg_punkte(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < s.n; v += gridDim.x * blockDim.x) {
    Quader q = sk_quader_leer();
    sk_quader_rein(q, s.x_anfang[v], 1e-8);
    sk_quader_rein(q, s.x[v], 1e-8);
    s.quader[v] = q;
  }
}
__global__ void k_sk_quader_weg_kanten(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < s.ne; e += gridDim.x * blockDim.x) {
    const uint32_t a = s.kanten[2 * e], b = s.kanten[2 * e + 1];
    Quader q = sk_quader_leer();
    sk_quader_rein(q, s.x_anfang[a], 1e-8);
    sk_quader_rein(q, s.x_anfang[b], 1e-8);
    sk_quader_rein(q, s.x[a], 1e-8);
    sk_quader_rein(q, s.x[b], 1e-8);
    s.quader[e] = q;
  }
}

// Point v through triangle d on the path of the step (:1201-1231): the time
// the four points are coplanar and the barycentric position at that time.
// With hat, a time not before `grenze` is rejected before the position test.
__device__ __forceinline__ bool sk_pd_zeit(const SelbstDev& s, const uint32_t* d, uint32_t v, bool hat, double grenze, double* zeit,
                                           double* uv) {
  const uint32_t idx[4] = {d[0], d[1], d[2], v};
  if (s.w[idx[0]] <= 0.0 && s.w[idx[1]] <= 0.0 && s.w[idx[2]] <= 0.0 && s.w[idx[3]] <= 0.0) return false;
  D3 a0[4], a1[4];
  for (int k = 0; k < 4; ++k) { a0[k] = s.x_anfang[idx[k]]; a1[k] = s.x[idx[k]]; }
  if (!sk_koplanar_zeit(a0, a1, SK_STOSS * 0.1, zeit)) return false;
  if (hat && grenze <= *zeit) return false;
  D3 q[4];
  for (int k = 0; k < 4; ++k) q[k] = sk_lerp(a0[k], a1[k], *zeit);
  return sk_im_dreieck(q[0], q[1], q[2], q[3], uv);
}
// The impact record of the hit of point v through triangle t.
__device__ __forceinline__ void sk_pd_satz(const SelbstDev& s, int t, const uint32_t* d, uint32_t v, Stoss* out) {
  double zeit, uv[2];
  sk_pd_zeit(s, d, v, false, 0.0, &zeit, uv);
  const uint32_t idx[4] = {d[0], d[1], d[2], v};
  D3 a0[4], a1[4], q[4];
  for (int k = 0; k < 4; ++k) {
    a0[k] = s.x_anfang[idx[k]];
    a1[k] = s.x[idx[k]];
    q[k] = sk_lerp(a0[k], a1[k], zeit);
  }
  const D3 n = sk_einheit(cross(q[1] - q[0], q[2] - q[0]));
  const double g[3] = {1.0 - uv[0] - uv[1], uv[0], uv[1]};
  const D3 rel = (a1[3] - a0[3]) - (((a1[0] - a0[0]) * g[0] + (a1[1] - a0[1]) * g[1]) + (a1[2] - a0[2]) * g[2]);
  const double seite = dot(rel, n) <= 0.0 ? 1.0 : -1.0;
  Stoss best;
  best.art = 0;
  best.element = (uint32_t)t;
  best.element2 = SK_KEIN;
  best.punkte[0] = v; best.punkte[1] = d[0]; best.punkte[2] = d[1]; best.punkte[3] = d[2];
  best.lage[0] = uv[0];
  best.lage[1] = uv[1];
  best.zeit = zeit;
  best.seite = seite;
  best.normale = n * seite;
  *out = best;
}

// Earliest point through each triangle; slot t of the impact array. One wave
// per triangle walks the point hash (sk_welle_je_eintrag) and folds its hits
// in candidate order, so the impact kept is the one a single thread walking
// the candidates in the reference's order keeps.
__global__ void __launch_bounds__(256) k_sk_stoss_pd(const SelbstDev* __restrict__ sp, const double* zelle_p, uint32_t eimer) {
  const SelbstDev& s = *sp;
  const double zelle = *zelle_p;
  const int lane = threadIdx.x & 63;
  const int wellen = gridDim.x * (blockDim.x >> 6);
  for (int t = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); t < s.nt; t += wellen) {
    const uint32_t* d = s.dreiecke + 3 * t;
    bool hat = false;
    double best_zeit = 0.0;
    uint32_t best_v = 0;
    if (!(d[0] == d[1] || d[1] == d[2] || d[0] == d[2])) {
      Quader qt = sk_quader_leer();
      for (int k = 0; k < 3; ++k) sk_quader_rein(qt, s.x_anfang[d[k]], 1e-8);
      for (int k = 0; k < 3; ++k) sk_quader_rein(qt, s.x[d[k]], 1e-8);
      sk_welle_je_eintrag(s, qt, zelle, eimer, -1, [&](bool aktiv, uint32_t v, const int64_t* c) {
        bool ok = false;
        double zeit = 0.0;
        if (aktiv && !(sk_partner(s, v, d[0]) || sk_partner(s, v, d[1]) || sk_partner(s, v, d[2]))) {
          const Quader qv = s.quader[v];
          if (sk_ueberlappt(qv, qt) && sk_erste_zelle(qt, qv, zelle, c)) {
            double uv[2];
            ok = sk_pd_zeit(s, d, v, hat, best_zeit, &zeit, uv);
          }
        }
        sk_welle_falten(ok, zeit, v, &hat, &best_zeit, &best_v);
      });
    }
    if (lane == 0) {
      if (hat) sk_pd_satz(s, t, d, best_v, &s.stoesse[t]);
      s.zahl[t] = hat ? 1u : 0u;
    }
  }
}

// Edges (i0, i1) and (j0, j1) crossing on the path of the step (:1233-1262):
// the coplanarity time and the two edge parameters at that time. With hat, a
// time not before `grenze` is rejected before the crossing test.
__device__ __forceinline__ bool sk_kk_zeit(const SelbstDev& s, uint32_t i0, uint32_t i1, uint32_t j0, uint32_t j1, bool hat,
                                           double grenze, double* zeit, double* sw) {
  const uint32_t idx[4] = {i1, j0, j1, i0};
  if (s.w[idx[0]] <= 0.0 && s.w[idx[1]] <= 0.0 && s.w[idx[2]] <= 0.0 && s.w[idx[3]] <= 0.0) return false;
  D3 a0[4], a1[4];
  for (int k = 0; k < 4; ++k) { a0[k] = s.x_anfang[idx[k]]; a1[k] = s.x[idx[k]]; }
  if (!sk_koplanar_zeit(a0, a1, SK_STOSS * 0.1, zeit)) return false;
  if (hat && grenze <= *zeit) return false;
  D3 q[4];
  for (int k = 0; k < 4; ++k) q[k] = sk_lerp(a0[k], a1[k], *zeit);
  return sk_kanten_kreuzen(q[3], q[0], q[1], q[2], sw);
}
// The impact record of the hit of edge i with edge j.
__device__ __forceinline__ void sk_kk_satz(const SelbstDev& s, int i, uint32_t j, Stoss* out) {
  const uint32_t i0 = s.kanten[2 * i], i1 = s.kanten[2 * i + 1];
  const uint32_t j0 = s.kanten[2 * j], j1 = s.kanten[2 * j + 1];
  double zeit, sw[2];
  sk_kk_zeit(s, i0, i1, j0, j1, false, 0.0, &zeit, sw);
  const uint32_t idx[4] = {i1, j0, j1, i0};
  D3 a0[4], a1[4], q[4];
  for (int k = 0; k < 4; ++k) {
    a0[k] = s.x_anfang[idx[k]];
    a1[k] = s.x[idx[k]];
    q[k] = sk_lerp(a0[k], a1[k], zeit);
  }
  const D3 n = sk_einheit(cross(q[0] - q[3], q[2] - q[1]));
  const D3 auf_i = (a1[3] - a0[3]) * (1.0 - sw[0]) + (a1[0] - a0[0]) * sw[0];
  const D3 auf_j = (a1[1] - a0[1]) * (1.0 - sw[1]) + (a1[2] - a0[2]) * sw[1];
  const double seite = dot(auf_i - auf_j, n) <= 0.0 ? 1.0 : -1.0;
  Stoss best;
  best.art = 1;
  best.element = (uint32_t)i;
  best.element2 = j;
  best.punkte[0] = i0; best.punkte[1] = i1; best.punkte[2] = j0; best.punkte[3] = j1;
  best.lage[0] = sw[0];
  best.lage[1] = sw[1];
  best.zeit = zeit;
  best.seite = seite;
  best.normale = n * seite;
  *out = best;
}

// Earliest partner edge j > i of each edge i; slot nt + i of the impact array.
// One wave per edge, as in k_sk_stoss_pd; the partners j <= i, which the
// reference skips first, are left out of the walk.
__global__ void __launch_bounds__(256) k_sk_stoss_kk(const SelbstDev* __restrict__ sp, const double* zelle_p, uint32_t eimer) {
  const SelbstDev& s = *sp;
  const double zelle = *zelle_p;
  const int lane = threadIdx.x & 63;
  const int wellen = gridDim.x * (blockDim.x >> 6);
  for (int i = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); i < s.ne; i += wellen) {
    const uint32_t i0 = s.kanten[2 * i], i1 = s.kanten[2 * i + 1];
    bool hat = false;
    double best_zeit = 0.0;
    uint32_t best_j = 0;
    if (i0 != i1) {
      const Quader qi = s.quader[i];
      sk_welle_je_eintrag(s, qi, zelle, eimer, (int64_t)i, [&](bool aktiv, uint32_t j, const int64_t* c) {
        bool ok = false;
        double zeit = 0.0;
        if (aktiv) {
          const Quader qj = s.quader[j];
          if (sk_ueberlappt(qi, qj) && sk_erste_zelle(qi, qj, zelle, c)) {
            const uint32_t j0 = s.kanten[2 * j], j1 = s.kanten[2 * j + 1];
            if (!(sk_partner(s, j0, i0) || sk_partner(s, j0, i1) || sk_partner(s, j1, i0) || sk_partner(s, j1, i1))) {
              double sw[2];
              ok = sk_kk_zeit(s, i0, i1, j0, j1, hat, best_zeit, &zeit, sw);
            }
          }
        }
        sk_welle_falten(ok, zeit, j, &hat, &best_zeit, &best_j);
      });
    }
    if (lane == 0) {
      if (hat) sk_kk_satz(s, i, best_j, &s.stoesse[s.nt + i]);
      s.zahl[s.nt + i] = hat ? 1u : 0u;
    }
  }
}

// ohne_durchdrungene (:1063-1086) per impact slot; cuts are few.
__global__ void k_sk_stoss_filter(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  const uint32_t nc = s.zaehler[SK_Z_SCHNITTE];
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < s.nt + s.ne; q += gridDim.x * blockDim.x) {
    if (!s.zahl[q]) continue;
    Stoss& st = s.stoesse[q];
    bool bleibt = true;
    if (st.art == 0) {
      for (uint32_t i = 0; i < nc && bleibt; ++i) {
        const Schnitt& c = s.schnitte[i];
        if (c.dreieck == st.element && (s.kanten[2 * c.kante] == st.punkte[0] || s.kanten[2 * c.kante + 1] == st.punkte[0])) bleibt = false;
      }
    } else {
      const uint32_t ki = st.element, kj = st.element2;
      auto trifft = [&](uint32_t a, uint32_t b) {
        for (int q2 = 0; q2 < 2; ++q2) {
          const uint32_t t = s.kanten_dreiecke[2 * b + q2];
          if (t == SK_KEIN) continue;
          const uint32_t* dk = s.dreieck_kanten + 3 * t;
          if (!(dk[0] == b || dk[1] == b || dk[2] == b)) continue;
          for (uint32_t i = 0; i < nc; ++i)
            if (s.schnitte[i].kante == a && s.schnitte[i].dreieck == t) return true;
        }
        return false;
      };
      bleibt = !(trifft(ki, kj) || trifft(kj, ki));
    }
    st.gilt = bleibt ? 1u : 0u;
  }
}

// stossantwort (:1288-1343): one iteration. Per impact the displacement of
// each of its points, from the positions the iteration starts with.
template <typename T>
__global__ void k_sk_stoss_beitrag(const SelbstDev* __restrict__ sp, const SysDev<T>* __restrict__ syp) {
  const SelbstDev& s = *sp;
  const SysDev<T>& sy = *syp;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < s.nt + s.ne; q += gridDim.x * blockDim.x) {
    if (!s.zahl[q] || !(s.stoesse[q].gilt & 1u)) continue;
    const Stoss& st = s.stoesse[q];
    const uint32_t p0 = st.punkte[0], p1 = st.punkte[1], p2 = st.punkte[2], p3 = st.punkte[3];
    uint32_t erste[2], zweite[3];
    double ge[2], gz[3];
    int ne, nz;
    D3 n;
    double d;
    if (st.art == 0) {
      n = sk_einheit(cross(sy.x[p2] - sy.x[p1], sy.x[p3] - sy.x[p1])) * st.seite;
      erste[0] = p0; ge[0] = 1.0; ne = 1;
      zweite[0] = p1; zweite[1] = p2; zweite[2] = p3;
      gz[0] = 1.0 - st.lage[0] - st.lage[1]; gz[1] = st.lage[0]; gz[2] = st.lage[1];
      nz = 3;
      d = dot(n, sy.x[p0] - sy.x[p1]);
    } else {
      n = st.normale;
      const D3 a = sk_lerp(sy.x[p0], sy.x[p1], st.lage[0]), b = sk_lerp(sy.x[p2], sy.x[p3], st.lage[1]);
      erste[0] = p0; erste[1] = p1; ge[0] = 1.0 - st.lage[0]; ge[1] = st.lage[0]; ne = 2;
      zweite[0] = p2; zweite[1] = p3; gz[0] = 1.0 - st.lage[1]; gz[1] = st.lage[1]; nz = 2;
      d = dot(n, a - b);
    }
    const double tief = SK_STOSS - d;
    D3* dx = s.st_dx + 5 * q;
    for (int k = 0; k < 5; ++k) dx[k] = zero3<double>();
    s.stoesse[q].gilt = tief > 0.0 ? 3u : 1u;  // bit 1: pushes in this iteration
    if (tief <= 0.0) continue;
    auto m_eff = [&](uint32_t i) { return sy.fflag[i] == 1 ? s.masse[i] * 1e9 : s.masse[i]; };
    double m1 = 0.0, m2 = 0.0;
    for (int k = 0; k < ne; ++k) m1 += ge[k] * m_eff(erste[k]);
    for (int k = 0; k < nz; ++k) m2 += gz[k] * m_eff(zweite[k]);
    bool fest1 = true, fest2 = true;
    for (int k = 0; k < ne; ++k) fest1 = fest1 && s.w[erste[k]] <= 0.0;
    for (int k = 0; k < nz; ++k) fest2 = fest2 && s.w[zweite[k]] <= 0.0;
    const double alpha = fest2 ? 1.0 : (fest1 ? 0.0 : m2 / (m1 + m2));
    for (int k = 0; k < ne; ++k) dx[k] = n * (tief * alpha);
    for (int k = 0; k < nz; ++k) dx[ne + k] = n * (tief * (1.0 - alpha)) * -1.0;
    atomicAdd(&s.zaehler[SK_Z_STOSS_ZAHL], 1u);
  }
}

// Per point: its impact contributions in list order (slot order within an
// impact), averaged.
__global__ void k_sk_stoss_zaehlen(const SelbstDev* __restrict__ sp) {
  const SelbstDev& s = *sp;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < s.nt + s.ne; q += gridDim.x * blockDim.x) {
    if (!s.zahl[q] || !(s.stoesse[q].gilt & 1u)) continue;
    const Stoss& st = s.stoesse[q];
    for (int k = 0; k < 4; ++k)
      if (s.w[st.punkte[k]] > 0.0) atomicAdd(&s.st_off[st.punkte[k]], 1u);
  }
}
__global__ void k_sk_stoss_fuellen(const SelbstDev* __restrict__ sp, uint32_t* fuell) {
  const SelbstDev& s = *sp;
  if (s.st_off[s.n] > s.kapazitaet_st_ent) return;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < s.nt + s.ne; q += gridDim.x * blockDim.x) {
    if (!s.zahl[q] || !(s.stoesse[q].gilt & 1u)) continue;
    const Stoss& st = s.stoesse[q];
    for (int k = 0; k < 4; ++k)
      if (s.w[st.punkte[k]] > 0.0) s.st_ent[atomicAdd(&fuell[st.punkte[k]], 1u)] = ((uint32_t)q << 2) | (uint32_t)k;
  }
}
template <typename T>
__global__ void k_sk_stoss_anwenden(const SelbstDev* __restrict__ sp, const SysDev<T>* __restrict__ syp) {
  const SelbstDev& s = *sp;
  const SysDev<T>& sy = *syp;
  if (s.st_off[s.n] > s.kapazitaet_st_ent) {
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&s.zaehler[SK_Z_UEBERLAUF], 1u);
    return;
  }
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    const uint32_t a = s.st_off[i], b = s.st_off[i + 1];
    if (a == b) continue;
    sk_sortiere(s.st_ent, a, b);  // impact order, as the reference's loop visits them
    D3 dx = zero3<double>();
    uint32_t cnt = 0;
    for (uint32_t e = a; e < b; ++e) {
      const uint32_t q = s.st_ent[e] >> 2, k = s.st_ent[e] & 3u;
      if (!(s.stoesse[q].gilt & 2u)) continue;
      dx = dx + s.st_dx[5 * q + k];
      ++cnt;
    }
    if (cnt > 0) sy.x[i] = sy.x[i] + dx * (1.0 / (double)cnt);
  }
}

// Step statistics (ImplizitStatistik, implizit.rs:550-561). phase 0: after the
// build (pairs, cuts, regions); phase 1: after the impacts.
__global__ void k_sk_statistik(const SelbstDev* __restrict__ sp, GridState* gs, int phase) {
  const SelbstDev& s = *sp;
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  if (phase == 0) {
    gs->naehe_pd += s.zaehler[SK_Z_NAEHE_PD];
    gs->naehe_kk += s.zaehler[SK_Z_NAEHE_KK];
    gs->durchdringungen += s.zaehler[SK_Z_SCHNITTE];
    gs->gebiete += s.zaehler[SK_Z_GEBIETE];
  } else {
    gs->selbst_stoesse += s.zaehler[SK_Z_STOSS_ZAHL];
  }
  if (s.zaehler[SK_Z_UEBERLAUF]) gs->selbst_ueberlauf = 1u;
}
