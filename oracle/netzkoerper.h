// netzkoerper.h — FP64 restatement of the triangle-mesh body `Koerpernetz`
// (crates/klon3d-sim/src/netzkoerper.rs).
//
// TEST INFRASTRUCTURE ONLY (the checker). Included by gcoracle.cpp inside its
// anonymous namespace, after the vector helpers sub/add/mul/dot/cross/norm
// (koerper.rs:37-54). Loop order, summation order and branch order follow the
// cited reference lines, so that compiled without FMA contraction the numbers
// are those the reference computes.

// Region of the nearest point on a triangle (netzkoerper.rs:50-58):
// 0..2 corner k, 3..5 edge k (0 = p0p1, 1 = p1p2, 2 = p2p0), 6 face.
enum { NK_ECKE = 0, NK_KANTE = 3, NK_FLAECHE = 6 };

struct Naechster {  // netzkoerper.rs:60-72
  double abstand = 0;
  V3 punkt{}, normale{};
  uint32_t dreieck = 0;
  int region = NK_FLAECHE;
};

// Rust `f as i64`: saturating, NaN -> 0.
inline int64_t als_i64(double f) {
  if (!(f == f)) return 0;
  if (f >= 9.2233720368547758e18) return INT64_MAX;
  if (f <= -9.2233720368547758e18) return INT64_MIN;
  return (int64_t)f;
}
// Rust `f as usize`: saturating at 0 and at the top, NaN -> 0.
inline uint64_t als_usize(double f) {
  if (!(f > 0.0)) return 0;
  if (f >= 1.8446744073709552e19) return UINT64_MAX;
  return (uint64_t)f;
}

struct Gitter {  // netzkoerper.rs:75-83
  V3 ursprung{};
  double zelle = 0;
  size_t zahl[3] = {0, 0, 0};
  std::vector<uint32_t> anfang, eintraege;

  // netzkoerper.rs:87-93
  std::array<int64_t, 3> index(V3 x) const {
    std::array<int64_t, 3> k{};
    for (int a = 0; a < 3; ++a) k[a] = als_i64(std::floor((x[a] - ursprung[a]) / zelle));
    return k;
  }
  // netzkoerper.rs:95-97
  size_t zellnummer(size_t kx, size_t ky, size_t kz) const { return (kz * zahl[1] + ky) * zahl[0] + kx; }
  // netzkoerper.rs:117-129
  bool bereich(V3 lo, V3 hi, size_t r[3][2]) const {
    auto a = index(lo), b = index(hi);
    for (int d = 0; d < 3; ++d) {
      int64_t n = (int64_t)zahl[d];
      if (b[d] < 0 || a[d] >= n) return false;
      r[d][0] = (size_t)std::max<int64_t>(a[d], 0);
      r[d][1] = (size_t)std::min<int64_t>(b[d], n - 1);
    }
    return true;
  }
  // netzkoerper.rs:131-141
  void klammer(std::array<int64_t, 3> a, std::array<int64_t, 3> b, size_t r[3][2]) const {
    for (int d = 0; d < 3; ++d) {
      int64_t n = (int64_t)zahl[d];
      r[d][0] = (size_t)std::max<int64_t>(a[d], 0);
      r[d][1] = (size_t)std::min<int64_t>(b[d], n - 1);
    }
  }
  // netzkoerper.rs:101-113
  template <typename F>
  void je_eintrag(V3 lo, V3 hi, F f) const {
    size_t r[3][2];
    if (!bereich(lo, hi, r)) return;
    for (size_t kz = r[2][0]; kz <= r[2][1]; ++kz)
      for (size_t ky = r[1][0]; ky <= r[1][1]; ++ky)
        for (size_t kx = r[0][0]; kx <= r[0][1]; ++kx) {
          size_t z = zellnummer(kx, ky, kz);
          for (uint32_t e = anfang[z]; e < anfang[z + 1]; ++e) f(eintraege[e]);
        }
  }
};

// netzkoerper.rs:185-218 (Ericson 2005, 5.1.5)
inline V3 naechster_auf_dreieck(V3 p, V3 a, V3 b, V3 c, int* region) {
  auto quot = [](double z, double n) { return n > 0.0 ? z / n : 0.0; };
  V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) { *region = NK_ECKE + 0; return a; }
  V3 bp = sub(p, b);
  double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) { *region = NK_ECKE + 1; return b; }
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) { *region = NK_KANTE + 0; return add(a, mul(ab, quot(d1, d1 - d3))); }
  V3 cp = sub(p, c);
  double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) { *region = NK_ECKE + 2; return c; }
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) { *region = NK_KANTE + 2; return add(a, mul(ac, quot(d2, d2 - d6))); }
  double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && d4 - d3 >= 0.0 && d5 - d6 >= 0.0) {
    double w = quot(d4 - d3, (d4 - d3) + (d5 - d6));
    *region = NK_KANTE + 1;
    return add(b, mul(sub(c, b), w));
  }
  double s = va + vb + vc;
  double v = quot(vb, s), w = quot(vc, s);
  *region = NK_FLAECHE;
  return add(a, add(mul(ab, v), mul(ac, w)));
}

// netzkoerper.rs:220-227
inline V3 normiert(V3 a) {
  double l = norm(a);
  if (l > 0.0) return mul(a, 1.0 / l);
  return {0.0, 0.0, 0.0};
}

// netzkoerper.rs:232-244
inline V3 schwerpunktgewichte(V3 p, V3 a, V3 b, V3 c) {
  V3 e1 = sub(b, a), e2 = sub(c, a), e3 = sub(p, a);
  double d00 = dot(e1, e1), d01 = dot(e1, e2), d11 = dot(e2, e2);
  double d20 = dot(e3, e1), d21 = dot(e3, e2);
  double nenner = d00 * d11 - d01 * d01;
  if (nenner == 0.0) return {1.0, 0.0, 0.0};
  double u = (d11 * d20 - d01 * d21) / nenner;
  double w = (d00 * d21 - d01 * d20) / nenner;
  return {1.0 - u - w, u, w};
}

// netzkoerper.rs:246-248
inline double winkel(V3 u, V3 v) { return std::atan2(norm(cross(u, v)), dot(u, v)); }

struct Kugel { V3 m; double r; };
using Dreieck = std::array<uint32_t, 3>;

// netzkoerper.rs:257-291: everything that depends on the points, in one pass.
inline double punktabhaengig(const std::vector<V3>& punkte, const std::vector<Dreieck>& dreiecke, std::vector<V3>& fn,
                             std::vector<V3>& pn, std::vector<Kugel>& kugel) {
  fn.clear();
  for (const auto& d : dreiecke) {
    V3 a = punkte[d[0]], b = punkte[d[1]], c = punkte[d[2]];
    fn.push_back(normiert(cross(sub(b, a), sub(c, a))));
  }
  pn.assign(punkte.size(), V3{0.0, 0.0, 0.0});
  double volumen = 0.0;
  kugel.clear();
  for (size_t t = 0; t < dreiecke.size(); ++t) {
    const auto& d = dreiecke[t];
    V3 q[3] = {punkte[d[0]], punkte[d[1]], punkte[d[2]]};
    for (int k = 0; k < 3; ++k) {
      double w = winkel(sub(q[(k + 1) % 3], q[k]), sub(q[(k + 2) % 3], q[k]));
      size_t p = d[k];
      pn[p] = add(pn[p], mul(fn[t], w));
    }
    volumen += dot(q[0], cross(q[1], q[2])) / 6.0;
    V3 m = mul(add(add(q[0], q[1]), q[2]), 1.0 / 3.0);
    double r = 0.0;
    for (int k = 0; k < 3; ++k) r = std::fmax(r, norm(sub(q[k], m)));  // fold(0.0, f64::max)
    kugel.push_back({m, r});
  }
  for (auto& p : pn) p = normiert(p);
  return volumen;
}

// netzkoerper.rs:800-816
inline void dreieck_klammer(const std::vector<std::array<int64_t, 3>>& zellen, const Dreieck& d, std::array<int64_t, 3>& lo,
                            std::array<int64_t, 3>& hi) {
  lo = zellen[d[0]];
  hi = zellen[d[0]];
  for (int e = 1; e < 3; ++e) {
    const auto& k = zellen[d[e]];
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::min(lo[a], k[a]);
      hi[a] = std::max(hi[a], k[a]);
    }
  }
}

// netzkoerper.rs:915-921
inline void leere_tabelle(Gitter& g, size_t nz) { g.anfang.assign(nz + 1, 0u); }
// netzkoerper.rs:925-931
inline void summiere(Gitter& g, size_t nz) {
  for (size_t z = 0; z < nz; ++z) g.anfang[z + 1] += g.anfang[z];
}

// netzkoerper.rs:861-879
inline void gitter_punkte(Gitter& g, const std::vector<std::array<int64_t, 3>>& zellen, size_t nz) {
  leere_tabelle(g, nz);
  size_t r[3][2];
  for (const auto& k : zellen) {
    g.klammer(k, k, r);
    g.anfang[g.zellnummer(r[0][0], r[1][0], r[2][0]) + 1] += 1;
  }
  summiere(g, nz);
  g.eintraege.assign(g.anfang[nz], 0u);
  std::vector<uint32_t> frei = g.anfang;
  for (size_t e = 0; e < zellen.size(); ++e) {
    g.klammer(zellen[e], zellen[e], r);
    size_t z = g.zellnummer(r[0][0], r[1][0], r[2][0]);
    g.eintraege[frei[z]] = (uint32_t)e;
    frei[z] += 1;
  }
}

// netzkoerper.rs:881-912
inline void gitter_dreiecke(Gitter& g, const std::vector<std::array<int64_t, 3>>& zellen, const std::vector<Dreieck>& dreiecke,
                            size_t nz) {
  leere_tabelle(g, nz);
  size_t r[3][2];
  std::array<int64_t, 3> a, b;
  for (const auto& d : dreiecke) {
    dreieck_klammer(zellen, d, a, b);
    g.klammer(a, b, r);
    for (size_t kz = r[2][0]; kz <= r[2][1]; ++kz)
      for (size_t ky = r[1][0]; ky <= r[1][1]; ++ky)
        for (size_t kx = r[0][0]; kx <= r[0][1]; ++kx) g.anfang[g.zellnummer(kx, ky, kz) + 1] += 1;
  }
  summiere(g, nz);
  g.eintraege.assign(g.anfang[nz], 0u);
  std::vector<uint32_t> frei = g.anfang;
  for (size_t e = 0; e < dreiecke.size(); ++e) {
    dreieck_klammer(zellen, dreiecke[e], a, b);
    g.klammer(a, b, r);
    for (size_t kz = r[2][0]; kz <= r[2][1]; ++kz)
      for (size_t ky = r[1][0]; ky <= r[1][1]; ++ky)
        for (size_t kx = r[0][0]; kx <= r[0][1]; ++kx) {
          size_t z = g.zellnummer(kx, ky, kz);
          g.eintraege[frei[z]] = (uint32_t)e;
          frei[z] += 1;
        }
  }
}

// netzkoerper.rs:785-797
inline void huelle_punkte(const std::vector<V3>& punkte, V3& lo, V3& hi) {
  lo = {INFINITY, INFINITY, INFINITY};
  hi = {-INFINITY, -INFINITY, -INFINITY};
  for (const auto& p : punkte)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::fmin(lo[a], p[a]);
      hi[a] = std::fmax(hi[a], p[a]);
    }
}

// netzkoerper.rs:820-858: one raster for both grids, cells 2*suchtiefe over the
// points' box widened by suchtiefe, never shrinking below the old box.
inline void gitter_beide(Gitter& g, Gitter& pg, double suchtiefe, const std::vector<V3>& punkte,
                         const std::vector<Dreieck>& dreiecke) {
  V3 lo, hi;
  huelle_punkte(punkte, lo, hi);
  double zelle = 2.0 * suchtiefe;
  V3 bedarf_lo = sub(lo, V3{suchtiefe, suchtiefe, suchtiefe});
  V3 bedarf_hi = hi;
  for (int a = 0; a < 3; ++a) bedarf_hi[a] += suchtiefe;
  V3 ursprung = bedarf_lo, ende = bedarf_hi;
  for (const Gitter* alt : {&g, &pg}) {
    if (alt->zahl[0] == 0 && alt->zahl[1] == 0 && alt->zahl[2] == 0) continue;
    for (int a = 0; a < 3; ++a) {
      ursprung[a] = std::fmin(ursprung[a], alt->ursprung[a]);
      ende[a] = std::fmax(ende[a], alt->ursprung[a] + alt->zelle * (double)alt->zahl[a]);
    }
  }
  size_t zahl[3];
  for (int a = 0; a < 3; ++a) zahl[a] = std::max<uint64_t>(als_usize(std::ceil((ende[a] - ursprung[a]) / zelle)), 1);
  for (Gitter* gg : {&g, &pg}) {
    gg->ursprung = ursprung;
    gg->zelle = zelle;
    for (int a = 0; a < 3; ++a) gg->zahl[a] = zahl[a];
  }
  std::vector<std::array<int64_t, 3>> zellen;
  zellen.reserve(punkte.size());
  for (const auto& p : punkte) zellen.push_back(g.index(p));
  size_t nz = zahl[0] * zahl[1] * zahl[2];
  gitter_punkte(pg, zellen, nz);
  gitter_dreiecke(g, zellen, dreiecke, nz);
}

struct Koerpernetz {  // netzkoerper.rs:145-177
  std::vector<V3> punkte;
  std::vector<Dreieck> dreiecke;
  std::vector<V3> flaechennormale;
  std::vector<std::array<uint32_t, 3>> nachbar;
  std::vector<V3> punktnormale;
  std::vector<Kugel> kugel;
  double suchtiefe = 0;
  Gitter gitter, punktgitter;
  size_t randkanten = 0;
  double volumen = 0;
  std::vector<V3> punktgeschwindigkeit;

  // netzkoerper.rs:297-388. Returns an empty string on success, else the reason.
  std::string neu(std::vector<V3> p, std::vector<Dreieck> d, double tiefe) {
    if (d.empty()) return "Koerpernetz ohne Dreiecke";
    if (!(tiefe > 0.0)) return "Suchtiefe muss positiv sein";
    size_t n = p.size();
    for (const auto& t : d)
      for (uint32_t i : t)
        if ((size_t)i >= n) return "Dreieck verweist auf einen Punkt ausserhalb";
    // per undirected edge: the triangles with edge number and direction
    struct Seite { uint32_t t; uint8_t k; bool r; };
    std::map<std::pair<uint32_t, uint32_t>, std::vector<Seite>> kanten;
    for (size_t t = 0; t < d.size(); ++t)
      for (uint8_t k = 0; k < 3; ++k) {
        uint32_t i = d[t][k], j = d[t][(k + 1) % 3];
        kanten[{std::min(i, j), std::max(i, j)}].push_back({(uint32_t)t, k, i < j});
      }
    std::vector<std::array<uint32_t, 3>> nb(d.size(), std::array<uint32_t, 3>{0, 0, 0});
    size_t rand = 0;
    for (const auto& [ij, liste] : kanten) {
      if (liste.size() == 1) {
        rand += 1;
        nb[liste[0].t][liste[0].k] = liste[0].t;
      } else if (liste.size() == 2) {
        if (liste[0].r == liste[1].r) return "Netz uneinheitlich orientiert";
        nb[liste[0].t][liste[0].k] = liste[1].t;
        nb[liste[1].t][liste[1].k] = liste[0].t;
      } else {
        return "Kante mit mehr als zwei Dreiecken";
      }
    }
    std::vector<V3> fn, pn;
    std::vector<Kugel> kg;
    double vol = punktabhaengig(p, d, fn, pn, kg);
    if (vol <= 0.0) return "Volumen <= 0: die Normalen zeigen nach innen";
    Gitter g, pg;
    gitter_beide(g, pg, tiefe, p, d);
    punkte = std::move(p);
    dreiecke = std::move(d);
    flaechennormale = std::move(fn);
    nachbar = std::move(nb);
    punktnormale = std::move(pn);
    kugel = std::move(kg);
    suchtiefe = tiefe;
    gitter = std::move(g);
    punktgitter = std::move(pg);
    randkanten = rand;
    volumen = vol;
    punktgeschwindigkeit.clear();
    return "";
  }

  // netzkoerper.rs:390-418
  std::string nachfuehren(const std::vector<V3>& p) {
    if (p.size() != punkte.size()) return "nachfuehren: falsche Punktzahl";
    double vol = punktabhaengig(p, dreiecke, flaechennormale, punktnormale, kugel);
    if (vol <= 0.0) return "Volumen <= 0 nach dem Verformen";
    punkte = p;
    volumen = vol;
    punktgeschwindigkeit.clear();
    gitter_beide(gitter, punktgitter, suchtiefe, punkte, dreiecke);
    return "";
  }

  // netzkoerper.rs:431-438
  V3 kantennormale(size_t t, size_t k) const {
    size_t o = nachbar[t][k];
    if (o == t) return flaechennormale[t];
    return normiert(add(flaechennormale[t], flaechennormale[o]));
  }
  // netzkoerper.rs:441-447
  V3 pseudonormale(size_t t, int r) const {
    if (r == NK_FLAECHE) return flaechennormale[t];
    if (r >= NK_KANTE) return kantennormale(t, (size_t)(r - NK_KANTE));
    return punktnormale[dreiecke[t][r - NK_ECKE]];
  }
  // netzkoerper.rs:474-489
  V3 geschwindigkeit_an(size_t t, V3 p, V3 starr) const {
    if (punktgeschwindigkeit.empty()) return starr;
    const auto& d = dreiecke[t];
    V3 gew = schwerpunktgewichte(p, punkte[d[0]], punkte[d[1]], punkte[d[2]]);
    V3 q{0.0, 0.0, 0.0};
    for (int j = 0; j < 3; ++j) {
      V3 v = punktgeschwindigkeit[d[j]];
      for (int k = 0; k < 3; ++k) q[k] += gew[j] * v[k];
    }
    return q;
  }
  // netzkoerper.rs:492-500
  Naechster ergebnis(V3 x, size_t t, V3 c, int r) const {
    V3 pn = pseudonormale(t, r);
    V3 d = sub(x, c);
    double l = norm(d);
    bool aussen = dot(d, pn) >= 0.0;
    V3 normale = l > 1e-12 ? mul(d, aussen ? 1.0 / l : -1.0 / l) : pn;
    Naechster e;
    e.abstand = aussen ? l : -l;
    e.punkt = c;
    e.normale = normale;
    e.dreieck = (uint32_t)t;
    e.region = r;
    return e;
  }
  // netzkoerper.rs:503-531
  bool naechster(V3 x, double radius, Naechster* out) const {
    const Gitter& g = gitter;
    size_t r[3][2];
    if (!g.bereich(sub(x, V3{radius, radius, radius}), add(x, V3{radius, radius, radius}), r)) return false;
    double best = radius;
    bool da = false;
    size_t bt = 0;
    V3 bq{};
    int breg = NK_FLAECHE;
    for (size_t kz = r[2][0]; kz <= r[2][1]; ++kz)
      for (size_t ky = r[1][0]; ky <= r[1][1]; ++ky)
        for (size_t kx = r[0][0]; kx <= r[0][1]; ++kx) {
          size_t z = g.zellnummer(kx, ky, kz);
          for (uint32_t e = g.anfang[z]; e < g.anfang[z + 1]; ++e) {
            size_t t = g.eintraege[e];
            const Kugel& kg = kugel[t];
            if (norm(sub(x, kg.m)) - kg.r >= best) continue;
            const auto& d = dreiecke[t];
            int reg;
            V3 q = naechster_auf_dreieck(x, punkte[d[0]], punkte[d[1]], punkte[d[2]], &reg);
            double l = norm(sub(x, q));
            if (l < best) {
              best = l;
              da = true;
              bt = t;
              bq = q;
              breg = reg;
            }
          }
        }
    if (!da) return false;
    *out = ergebnis(x, bt, bq, breg);
    return true;
  }
  // netzkoerper.rs:541-596 (Moeller & Trumbore 1997)
  bool durchstoss(V3 a, V3 b, V3 starr, V3* p, V3* n, V3* vk) const {
    V3 d = sub(b, a);
    if (dot(d, d) < 1e-24) return false;
    V3 lo = a, hi = a;
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::fmin(lo[k], b[k]);
      hi[k] = std::fmax(hi[k], b[k]);
    }
    const Gitter& g = gitter;
    size_t r[3][2];
    if (!g.bereich(lo, hi, r)) return false;
    bool da = false;
    double sb = 0;
    size_t tb = 0;
    const double eps = 1e-12;
    for (size_t kz = r[2][0]; kz <= r[2][1]; ++kz)
      for (size_t ky = r[1][0]; ky <= r[1][1]; ++ky)
        for (size_t kx = r[0][0]; kx <= r[0][1]; ++kx) {
          size_t z = g.zellnummer(kx, ky, kz);
          for (uint32_t e = g.anfang[z]; e < g.anfang[z + 1]; ++e) {
            size_t t = g.eintraege[e];
            if (dot(d, flaechennormale[t]) >= 0.0) continue;
            const auto& dt = dreiecke[t];
            V3 p0 = punkte[dt[0]], p1 = punkte[dt[1]], p2 = punkte[dt[2]];
            V3 e1 = sub(p1, p0), e2 = sub(p2, p0);
            V3 pv = cross(d, e2);
            double det = dot(e1, pv);
            if (std::fabs(det) < 1e-30) continue;
            V3 tv = sub(a, p0);
            double u = dot(tv, pv) / det;
            if (u < -eps || u > 1.0 + eps) continue;
            V3 qv = cross(tv, e1);
            double v = dot(d, qv) / det;
            if (v < -eps || u + v > 1.0 + eps) continue;
            double s = dot(e2, qv) / det;
            if (s >= -eps && s <= 1.0 + eps && (!da || s < sb)) {
              da = true;
              sb = s;
              tb = t;
            }
          }
        }
    if (!da) return false;
    // s.clamp(0.0, 1.0)
    double sc = sb < 0.0 ? 0.0 : (sb > 1.0 ? 1.0 : sb);
    *p = add(a, mul(d, sc));
    *n = flaechennormale[tb];
    *vk = geschwindigkeit_an(tb, *p, starr);
    return true;
  }
  // netzkoerper.rs:668-675
  V3 geschwindigkeit_je_punkt(size_t k, V3 starr) const {
    return punktgeschwindigkeit.empty() ? starr : punktgeschwindigkeit[k];
  }
  // netzkoerper.rs:681-736
  template <typename F>
  void stiche(double tief, double band, V3 versatz, const std::vector<Dreieck>& stoff, const std::vector<V3>& x, F f) const {
    double rand = std::fmax(tief, band);
    for (size_t di = 0; di < stoff.size(); ++di) {
      const auto& d = stoff[di];
      V3 a = x[d[0]], b = x[d[1]], c = x[d[2]];
      V3 lo = a, hi = a;
      for (V3 q : {b, c})
        for (int k = 0; k < 3; ++k) {
          lo[k] = std::fmin(lo[k], q[k]);
          hi[k] = std::fmax(hi[k], q[k]);
        }
      lo = sub(sub(lo, versatz), V3{rand, rand, rand});
      hi = add(sub(hi, versatz), V3{rand, rand, rand});
      V3 e1 = sub(b, a), e2 = sub(c, a);
      punktgitter.je_eintrag(lo, hi, [&](uint32_t kk) {
        size_t k = kk;
        V3 p = add(punkte[k], versatz);
        V3 n = punktnormale[k];
        V3 pv = cross(n, e2);
        double det = dot(e1, pv);
        if (std::fabs(det) < 1e-18) return;
        V3 tv = sub(p, a);
        double bu = dot(tv, pv) / det;
        if (!(bu >= 0.0 && bu <= 1.0)) return;
        V3 qv = cross(tv, e1);
        double bv = dot(n, qv) / det;
        if (bv < 0.0 || bu + bv > 1.0) return;
        double t = dot(e2, qv) / det;
        if (t > -tief && t < band) f(di, kk, V3{1.0 - bu - bv, bu, bv}, n, t);
      });
    }
  }
  // netzkoerper.rs:625-665
  size_t punkte_gegen_dreiecke(double tief, V3 versatz, V3 geschw, const std::vector<Dreieck>& stoff, const std::vector<double>& w,
                               std::vector<V3>& x, std::vector<V3>& v) const {
    struct Stich { bool da = false; double tiefe = 0; V3 n{}, bary{}; uint32_t kir = 0; };
    std::vector<Stich> tiefster(stoff.size());
    stiche(tief, -1e-9, versatz, stoff, x, [&](size_t di, uint32_t kir, V3 bary, V3 n, double t) {
      if (!tiefster[di].da || -t > tiefster[di].tiefe) tiefster[di] = {true, -t, n, bary, kir};
    });
    size_t zahl = 0;
    for (size_t di = 0; di < stoff.size(); ++di) {
      const Stich& st = tiefster[di];
      if (!st.da) continue;
      const auto& d = stoff[di];
      double gew[3] = {w[d[0]], w[d[1]], w[d[2]]};
      double nenner = 0.0;
      for (int j = 0; j < 3; ++j) nenner += st.bary[j] * st.bary[j] * gew[j];
      if (nenner <= 0.0) continue;
      V3 vk = geschwindigkeit_je_punkt(st.kir, geschw);
      V3 s{0.0, 0.0, 0.0};
      for (int j = 0; j < 3; ++j) s = add(s, mul(v[d[j]], st.bary[j]));
      V3 vq = sub(s, vk);
      double vn = std::fmin(dot(vq, st.n), 0.0);
      for (int j = 0; j < 3; ++j) {
        double f = st.bary[j] * gew[j] / nenner;
        x[d[j]] = add(x[d[j]], mul(st.n, f * st.tiefe));
        v[d[j]] = sub(v[d[j]], mul(st.n, f * vn));
      }
      zahl += 1;
    }
    return zahl;
  }
  // netzkoerper.rs:770-782
  Naechster naechster_alle(V3 x) const {
    double bl = INFINITY;
    size_t bt = 0;
    V3 bq{0.0, 0.0, 0.0};
    int breg = NK_FLAECHE;
    for (size_t t = 0; t < dreiecke.size(); ++t) {
      const auto& d = dreiecke[t];
      int reg;
      V3 q = naechster_auf_dreieck(x, punkte[d[0]], punkte[d[1]], punkte[d[2]], &reg);
      double l = norm(sub(x, q));
      if (l < bl) {
        bl = l;
        bt = t;
        bq = q;
        breg = reg;
      }
    }
    return ergebnis(x, bt, bq, breg);
  }
  // netzkoerper.rs:741-767
  Naechster abstand_ohne_grenze(V3 x) const {
    const Gitter& g = gitter;
    double quadrat = 0.0;
    for (int k = 0; k < 3; ++k) {
      double e = g.ursprung[k] + g.zelle * (double)g.zahl[k];
      double d = std::fmax(std::fabs(x[k] - g.ursprung[k]), std::fabs(x[k] - e));
      quadrat += d * d;
    }
    double weit = std::sqrt(quadrat) + g.zelle;
    double r = std::fmax(suchtiefe, g.zelle);
    for (;;) {
      size_t rr[3][2];
      if (g.bereich(sub(x, V3{r, r, r}), add(x, V3{r, r, r}), rr)) {
        Naechster n;
        if (naechster(x, r, &n)) return n;
      }
      if (r >= weit) return naechster_alle(x);
      r = std::fmin(r * 2.0, weit);
    }
  }
};
