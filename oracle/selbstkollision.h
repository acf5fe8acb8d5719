// selbstkollision.h — FP64 restatement of klon3d-sim's cloth self-collision
// (crates/klon3d-sim/src/selbstkollision.rs), included into gcoracle.cpp's
// anonymous namespace after the vector helpers.
//
// TEST INFRASTRUCTURE ONLY (the checker). Loop, traversal and summation order
// follow the cited code: candidate lists come out in the reference's order,
// with the same hash (Ericson's primes, bucket count, per-entry bucket dedup).
#pragma once

constexpr double ZUSCHLAG_QUADER = 2e-5;     // selbstkollision.rs:41
constexpr double ZUSCHLAG_PAAR = 4e-5;       // :43
constexpr double KANTE_MIN = 1e-4;           // :45
constexpr double AUSLENKUNG_MAX = 2.5e-3;    // :598
constexpr double ANNAEHERUNG_ANTEIL = 1.0 / 32.0;  // :600
constexpr uint32_t FOLGE_MAX = 5;            // :806
constexpr double SCHUB_HEBEL = 1e-3;         // :809
constexpr double STOSS_ABSTAND = 1e-4;       // :1092
constexpr double STOSS_KANTE_MIN = 1e-4;     // :1094
constexpr uint32_t KEIN = 0xFFFFFFFFu;

enum Art { PUNKT_DREIECK = 0, KANTE_KANTE = 1 };

// :60-71
struct Naehe {
  Art art;
  uint32_t punkte[4];
  double gewichte[4];
  V3 normale;
  double abstand;
  double ruhe;
};

// :700-707
struct Feder {
  uint32_t punkte[5];
  double gewichte[5];
  V3 kraft;
  M3 block;
};

// :812-825
struct Schnitt {
  uint32_t kante, dreieck;
  double t;
  double w[3];
  V3 richtung;
  double gewicht;
};

// :1099-1112
struct Stoss {
  Art art;
  uint32_t punkte[4];
  double lage[2];
  double zeit;
  double seite;
  V3 normale;
};

struct Quader { V3 lo, hi; };

// :75-123, built by Stoffnetz::neu (:157-190)
struct Stoffnetz {
  std::vector<std::array<uint32_t, 3>> dreiecke;
  std::vector<std::array<uint32_t, 2>> kanten;
  std::vector<std::array<uint32_t, 2>> kanten_dreiecke;
  std::vector<std::array<uint32_t, 3>> dreieck_kanten;
  std::vector<double> r_punkt, r_kante, r_dreieck;
  std::map<std::pair<uint32_t, uint32_t>, uint32_t> kanten_index;
  std::vector<std::vector<uint32_t>> naht_partner;
  std::vector<uint32_t> teil_von, band_von;
  std::vector<double> dicke;

  static Stoffnetz neu(const std::vector<std::array<uint32_t, 3>>& dr, const std::vector<double>& r) {
    Stoffnetz s;
    s.dreiecke = dr;
    s.r_punkt = r;
    for (size_t t = 0; t < dr.size(); ++t) {
      std::array<uint32_t, 3> dk{};
      for (int k = 0; k < 3; ++k) {
        uint32_t a = dr[t][k], b = dr[t][(k + 1) % 3];
        std::pair<uint32_t, uint32_t> key{std::min(a, b), std::max(a, b)};
        auto it = s.kanten_index.find(key);
        uint32_t e;
        if (it == s.kanten_index.end()) {
          s.kanten.push_back({key.first, key.second});
          s.kanten_dreiecke.push_back({KEIN, KEIN});
          e = (uint32_t)(s.kanten.size() - 1);
          s.kanten_index[key] = e;
        } else {
          e = it->second;
        }
        auto& kd = s.kanten_dreiecke[e];
        if (kd[0] == KEIN) kd[0] = (uint32_t)t;
        else if (kd[1] == KEIN) kd[1] = (uint32_t)t;
        dk[k] = e;
      }
      s.dreieck_kanten.push_back(dk);
    }
    for (const auto& k : s.kanten) s.r_kante.push_back(fmax_(r[k[0]], r[k[1]]));
    for (const auto& d : dr) {
      double m = 0.0;
      for (uint32_t i : d) m = fmax_(m, r[i]);
      s.r_dreieck.push_back(m);
    }
    return s;
  }
  // :128-140
  bool band_und_traeger(uint32_t a, uint32_t b) const {
    auto traeger = [&](uint32_t p, uint32_t* t) {
      if (p >= band_von.size()) return false;
      *t = band_von[p];
      return *t != KEIN;
    };
    auto teil = [&](uint32_t p, uint32_t* t) {
      if (p >= teil_von.size()) return false;
      *t = teil_von[p];
      return true;
    };
    uint32_t t, tb;
    if (traeger(a, &t) && teil(b, &tb) && tb == t) return true;
    if (traeger(b, &t) && teil(a, &tb) && tb == t) return true;
    return false;
  }
  // :147-156
  bool band_abstand(uint32_t a, const uint32_t* andere, int na, double* out) const {
    bool irgend = false;
    for (int q = 0; q < na; ++q) irgend = irgend || band_und_traeger(a, andere[q]);
    if (!irgend) return false;
    auto dick = [&](uint32_t p) { return p < dicke.size() ? dicke[p] : 0.0; };
    double m = 0.0;
    for (int q = 0; q < na; ++q) m = fmax_(m, dick(andere[q]));
    *out = dick(a) + m;
    return true;
  }
  // :193-195
  bool gleich_oder_partner(uint32_t a, uint32_t b) const {
    if (a == b) return true;
    if (a >= naht_partner.size()) return false;
    for (uint32_t q : naht_partner[a])
      if (q == b) return true;
    return false;
  }
};

// :196-203
inline V3 einheit(V3 a) {
  double l = norm(a);
  return l > 1e-12 ? mul(a, 1.0 / l) : V3{1.0, 0.0, 0.0};
}

// :209-251
struct Normalen {
  std::vector<V3> dreieck, punkt, kantenebene;
  static Normalen neu(const Stoffnetz& netz, const std::vector<V3>& x) {
    Normalen n;
    for (const auto& d : netz.dreiecke) n.dreieck.push_back(einheit(cross(sub(x[d[1]], x[d[0]]), sub(x[d[2]], x[d[0]]))));
    n.punkt.assign(x.size(), V3{0.0, 0.0, 0.0});
    for (size_t t = 0; t < netz.dreiecke.size(); ++t)
      for (uint32_t i : netz.dreiecke[t]) n.punkt[i] = add(n.punkt[i], n.dreieck[t]);
    for (auto& p : n.punkt) p = einheit(p);
    for (size_t e = 0; e < netz.kanten.size(); ++e) {
      const auto& k = netz.kanten[e];
      const auto& kd = netz.kanten_dreiecke[e];
      V3 m;
      if (kd[0] == KEIN) m = {1.0, 0.0, 0.0};
      else if (kd[1] == KEIN) m = n.dreieck[kd[0]];
      else {
        V3 a = n.dreieck[kd[0]], b = n.dreieck[kd[1]];
        m = dot(a, b) < 0.0 ? sub(a, b) : add(a, b);
      }
      n.kantenebene.push_back(einheit(cross(m, sub(x[k[0]], x[k[1]]))));
    }
    return n;
  }
};

// :263-265 (`as i64` of a finite value: truncation toward zero of floor)
inline int64_t zellindex(double x, double zelle) { return (int64_t)std::floor(x / zelle); }

// :270-272
inline bool zellbereich_ok(const Quader& q, double zelle) {
  for (int a = 0; a < 3; ++a)
    if (!(std::isfinite(q.lo[a]) && std::isfinite(q.hi[a]) && (q.hi[a] - q.lo[a]) <= 64.0 * zelle)) return false;
  return true;
}

// :274-278
inline size_t streue(int64_t i, int64_t j, int64_t k, size_t n) {
  uint32_t h = ((uint32_t)i * 0x8DA6B343u) ^ ((uint32_t)j * 0xD8163841u) ^ ((uint32_t)k * 0xCB1AB31Fu);
  return (size_t)h % n;
}

// :256-344
struct Hashgitter {
  double zelle = 0;
  size_t eimer = 0;
  std::vector<uint32_t> anfang, eintraege;

  template <typename F>
  static void je_zelle(const Quader& q, double zelle, size_t eimer, F f) {
    if (!zellbereich_ok(q, zelle)) return;
    int64_t lo[3], hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = zellindex(q.lo[a], zelle); hi[a] = zellindex(q.hi[a], zelle); }
    std::vector<size_t> hs;
    for (int64_t i = lo[0]; i <= hi[0]; ++i)
      for (int64_t j = lo[1]; j <= hi[1]; ++j)
        for (int64_t k = lo[2]; k <= hi[2]; ++k) hs.push_back(streue(i, j, k, eimer));
    std::sort(hs.begin(), hs.end());
    hs.erase(std::unique(hs.begin(), hs.end()), hs.end());
    for (size_t h : hs) f(h);
  }

  static Hashgitter neu(const std::vector<Quader>& quader, double zelle) {
    Hashgitter g;
    g.zelle = zelle;
    g.eimer = std::max<size_t>(2 * quader.size(), 64);
    std::vector<uint32_t> zahl(g.eimer + 1, 0);
    for (const auto& q : quader) je_zelle(q, zelle, g.eimer, [&](size_t h) { zahl[h + 1] += 1; });
    for (size_t h = 0; h < g.eimer; ++h) zahl[h + 1] += zahl[h];
    g.anfang = zahl;
    std::vector<uint32_t> fuell = zahl;
    g.eintraege.assign(g.anfang[g.eimer], 0);
    for (size_t e = 0; e < quader.size(); ++e)
      je_zelle(quader[e], zelle, g.eimer, [&](size_t h) { g.eintraege[fuell[h]++] = (uint32_t)e; });
    return g;
  }

  template <typename F>
  void je_eintrag(const Quader& q, F f) const {
    if (!zellbereich_ok(q, zelle)) return;
    int64_t lo[3], hi[3];
    for (int a = 0; a < 3; ++a) { lo[a] = zellindex(q.lo[a], zelle); hi[a] = zellindex(q.hi[a], zelle); }
    for (int64_t i = lo[0]; i <= hi[0]; ++i)
      for (int64_t j = lo[1]; j <= hi[1]; ++j)
        for (int64_t k = lo[2]; k <= hi[2]; ++k) {
          size_t h = streue(i, j, k, eimer);
          for (uint32_t a = anfang[h]; a < anfang[h + 1]; ++a) {
            int64_t c[3] = {i, j, k};
            f(eintraege[a], c);
          }
        }
  }
};

// :347-357
inline Quader quader_punkte(const V3* p, int n, double m) {
  double inf = std::numeric_limits<double>::infinity();
  Quader q{{inf, inf, inf}, {-inf, -inf, -inf}};
  for (int s = 0; s < n; ++s)
    for (int a = 0; a < 3; ++a) {
      q.lo[a] = fmin_(q.lo[a], p[s][a] - m);
      q.hi[a] = fmax_(q.hi[a], p[s][a] + m);
    }
  return q;
}

// :359-361
inline bool ueberlappt(const Quader& a, const Quader& b) {
  for (int k = 0; k < 3; ++k)
    if (!(a.lo[k] <= b.hi[k] && b.lo[k] <= a.hi[k])) return false;
  return true;
}

// :365-367
inline bool erste_zelle(const Quader& a, const Quader& b, double zelle, const int64_t* c) {
  for (int k = 0; k < 3; ++k)
    if (zellindex(fmax_(a.lo[k], b.lo[k]), zelle) != c[k]) return false;
  return true;
}

// netzkoerper.rs:185-218 (Ericson 2005, 5.1.5): the closest point only
inline V3 naechster_auf_dreieck(V3 p, V3 a, V3 b, V3 c) {
  auto quot = [](double z, double n) { return n > 0.0 ? z / n : 0.0; };
  V3 ab = sub(b, a), ac = sub(c, a), ap = sub(p, a);
  double d1 = dot(ab, ap), d2 = dot(ac, ap);
  if (d1 <= 0.0 && d2 <= 0.0) return a;
  V3 bp = sub(p, b);
  double d3 = dot(ab, bp), d4 = dot(ac, bp);
  if (d3 >= 0.0 && d4 <= d3) return b;
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0.0 && d1 >= 0.0 && d3 <= 0.0) return add(a, mul(ab, quot(d1, d1 - d3)));
  V3 cp = sub(p, c);
  double d5 = dot(ab, cp), d6 = dot(ac, cp);
  if (d6 >= 0.0 && d5 <= d6) return c;
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0.0 && d2 >= 0.0 && d6 <= 0.0) return add(a, mul(ac, quot(d2, d2 - d6)));
  double va = d3 * d6 - d5 * d4;
  if (va <= 0.0 && d4 - d3 >= 0.0 && d5 - d6 >= 0.0) {
    double w = quot(d4 - d3, (d4 - d3) + (d5 - d6));
    return add(b, mul(sub(c, b), w));
  }
  double s = va + vb + vc;
  double v = quot(vb, s), w = quot(vc, s);
  return add(a, add(mul(ab, v), mul(ac, w)));
}

// :370-383
inline V3 naechster_dreieck(V3 p0, V3 p1, V3 p2, V3 x) {
  V3 a = naechster_auf_dreieck(x, p0, p1, p2);
  V3 e1 = sub(p1, p0), e2 = sub(p2, p0), r = sub(a, p0);
  double d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
  double det = d11 * d22 - d12 * d12;
  if (det <= 1e-30) return {1.0, 0.0, 0.0};
  double r1 = dot(r, e1), r2 = dot(r, e2);
  double w1 = (d22 * r1 - d12 * r2) / det;
  double w2 = (d11 * r2 - d12 * r1) / det;
  return {1.0 - w1 - w2, w1, w2};
}

inline double r_max_von(const Stoffnetz& netz) {
  double m = 0.0;
  for (double r : netz.r_punkt) m = fmax_(m, r);
  return m;
}

// Brute-force stand-in for Hashgitter::je_eintrag: every entry in index
// order, each once (checks that the hash finds every pair exactly once).
template <typename F>
inline void alle_eintraege(size_t n, F f) {
  const int64_t c[3] = {0, 0, 0};
  for (size_t e = 0; e < n; ++e) f((uint32_t)e, c);
}

// :400-467 (brute: search without the hash, for the hash's own check)
inline void finde_punkt_dreieck(const Stoffnetz& netz, const std::vector<V3>& x, const std::vector<double>& w,
                                const Normalen& nrm, double r_max, std::vector<Naehe>& aus, bool brute = false) {
  std::vector<Quader> qp(x.size());
  for (size_t v = 0; v < x.size(); ++v) qp[v] = quader_punkte(&x[v], 1, netz.r_punkt[v] + ZUSCHLAG_QUADER);
  double zelle = fmax_(2.0 * r_max + ZUSCHLAG_PAAR, 1e-4) * 2.0;
  Hashgitter gitter = Hashgitter::neu(qp, zelle);
  for (size_t t = 0; t < netz.dreiecke.size(); ++t) {
    const auto& d = netz.dreiecke[t];
    uint32_t i0 = d[0], i1 = d[1], i2 = d[2];
    if (i0 == i1 || i1 == i2 || i0 == i2) continue;
    V3 p[3] = {x[i0], x[i1], x[i2]};
    double rt = netz.r_dreieck[t];
    Quader qt = quader_punkte(p, 3, rt + ZUSCHLAG_QUADER);
    V3 nt = nrm.dreieck[t];
    V3 en[3];
    for (int k = 0; k < 3; ++k) en[k] = nrm.kantenebene[netz.dreieck_kanten[t][k]];
    double s[3] = {dot(sub(p[2], p[0]), en[0]), dot(sub(p[0], p[1]), en[1]), dot(sub(p[1], p[2]), en[2])};
    auto paar = [&](uint32_t v, const int64_t* c) {
      for (uint32_t i : d)
        if (netz.gleich_oder_partner(v, i)) return;
      size_t vu = v;
      if (w[vu] <= 0.0 && w[i0] <= 0.0 && w[i1] <= 0.0 && w[i2] <= 0.0) return;
      if (!ueberlappt(qp[vu], qt) || (!brute && !erste_zelle(qt, qp[vu], gitter.zelle, c))) return;
      V3 xv = x[vu];
      double ruhe = rt + netz.r_punkt[vu];
      double eine;
      if (netz.band_abstand(v, d.data(), 3, &eine)) ruhe = eine;
      double soll = ruhe + ZUSCHLAG_PAAR;
      double sd = dot(nt, sub(xv, p[0]));
      if (std::fabs(sd) > soll) return;
      for (int k = 0; k < 3; ++k)
        if (dot(sub(xv, p[k]), en[k]) * s[k] < 0.0) return;
      V3 g = naechster_dreieck(p[0], p[1], p[2], xv);
      V3 cp = add(add(mul(p[0], g[0]), mul(p[1], g[1])), mul(p[2], g[2]));
      V3 r = sub(xv, cp);
      double abstand = norm(r);
      if (abstand > soll) return;
      Naehe n;
      n.art = PUNKT_DREIECK;
      n.punkte[0] = v; n.punkte[1] = i0; n.punkte[2] = i1; n.punkte[3] = i2;
      n.gewichte[0] = 1.0; n.gewichte[1] = -g[0]; n.gewichte[2] = -g[1]; n.gewichte[3] = -g[2];
      n.normale = abstand > 1e-12 ? mul(r, 1.0 / abstand) : nt;
      n.abstand = abstand;
      n.ruhe = ruhe;
      aus.push_back(n);
    };
    if (brute) alle_eintraege(x.size(), paar);
    else gitter.je_eintrag(qt, paar);
  }
}

// :469-484
inline Quader kanten_quader(V3 a, V3 b, double m) {
  V3 t = einheit(sub(b, a));
  V3 h = std::fabs(t[1] - 1.0) < 1e-5 ? V3{1.0, 0.0, 0.0} : V3{0.0, 1.0, 0.0};
  V3 u = einheit(cross(t, h));
  V3 v = einheit(cross(t, u));
  V3 ecken[8];
  int z = 0;
  for (V3 p : {a, b})
    for (double su : {-m, m})
      for (double sv : {-m, m}) ecken[z++] = add(add(p, mul(u, su)), mul(v, sv));
  return quader_punkte(ecken, 8, 0.0);
}

// :489-504
inline bool ist_kuerzeste(const Stoffnetz& netz, const Normalen& nrm, const std::vector<V3>& x, size_t e, V3 u, V3 m) {
  uint32_t k0 = netz.kanten[e][0], k1 = netz.kanten[e][1];
  for (uint32_t t : netz.kanten_dreiecke[e]) {
    if (t == KEIN) continue;
    const auto& d = netz.dreiecke[t];
    uint32_t o = KEIN;
    for (uint32_t i : d)
      if (i != k0 && i != k1) { o = i; break; }
    if (o == KEIN) continue;
    V3 r = sub(x[o], x[k0]);
    V3 wr = einheit(sub(r, mul(u, dot(r, u))));
    if (dot(m, wr) > 0.0 && std::fabs(dot(m, nrm.dreieck[t])) < 0.5) return false;
  }
  return true;
}

// :506-592
inline void finde_kante_kante(const Stoffnetz& netz, const std::vector<V3>& x, const std::vector<double>& w,
                              const Normalen& nrm, double r_max, std::vector<Naehe>& aus, bool brute = false) {
  std::vector<Quader> qk;
  for (size_t e = 0; e < netz.kanten.size(); ++e)
    qk.push_back(kanten_quader(x[netz.kanten[e][0]], x[netz.kanten[e][1]], netz.r_kante[e] + ZUSCHLAG_QUADER));
  double zelle = fmax_(2.0 * r_max + ZUSCHLAG_PAAR, 1e-4) * 2.0;
  Hashgitter gitter = Hashgitter::neu(qk, zelle);
  for (size_t e = 0; e < netz.kanten.size(); ++e) {
    uint32_t a0 = netz.kanten[e][0], a1 = netz.kanten[e][1];
    if (a0 == a1) continue;
    V3 pa0 = x[a0], pa1 = x[a1];
    auto paar = [&](uint32_t fi, const int64_t* c) {
      size_t f = fi;
      if (f <= e || !ueberlappt(qk[e], qk[f]) || (!brute && !erste_zelle(qk[e], qk[f], gitter.zelle, c))) return;
      uint32_t b0 = netz.kanten[f][0], b1 = netz.kanten[f][1];
      for (uint32_t b : {b0, b1})
        if (netz.gleich_oder_partner(a0, b) || netz.gleich_oder_partner(a1, b)) return;
      uint32_t bb[2] = {b0, b1}, aa[2] = {a0, a1};
      double band_dick = 0.0;
      bool band = netz.band_abstand(a0, bb, 2, &band_dick) || netz.band_abstand(a1, bb, 2, &band_dick) ||
                  netz.band_abstand(b0, aa, 2, &band_dick) || netz.band_abstand(b1, aa, 2, &band_dick);
      if (w[a0] <= 0.0 && w[a1] <= 0.0 && w[b0] <= 0.0 && w[b1] <= 0.0) return;
      V3 pb0 = x[b0], pb1 = x[b1];
      V3 da = sub(pa1, pa0), db = sub(pb1, pb0);
      double la = norm(da), lb = norm(db);
      if (la < KANTE_MIN || lb < KANTE_MIN) return;
      V3 ua = mul(da, 1.0 / la), ub = mul(db, 1.0 / lb);
      double cw = dot(ua, ub);
      if (std::fabs(cw) > 0.99999) return;
      V3 r = sub(pb0, pa0);
      double p = dot(r, ua), q = -dot(r, ub);
      double nen = 1.0 - cw * cw;
      double sa = (p + cw * q) / nen;
      double sb = (q + cw * p) / nen;
      if (!(sa >= 0.0 && sa <= la) || !(sb >= 0.0 && sb <= lb)) return;
      V3 pa = add(pa0, mul(ua, sa));
      V3 pb = add(pb0, mul(ub, sb));
      V3 dv = sub(pa, pb);
      double abstand = norm(dv);
      double ruhe = band ? band_dick : netz.r_kante[e] + netz.r_kante[f];
      if (abstand > ruhe + ZUSCHLAG_PAAR) return;
      V3 n = einheit(dv);
      if (!ist_kuerzeste(netz, nrm, x, f, ub, n) || !ist_kuerzeste(netz, nrm, x, e, ua, mul(n, -1.0))) return;
      double s = sa / la, t = sb / lb;
      Naehe nh;
      nh.art = KANTE_KANTE;
      nh.punkte[0] = a0; nh.punkte[1] = a1; nh.punkte[2] = b0; nh.punkte[3] = b1;
      nh.gewichte[0] = 1.0 - s; nh.gewichte[1] = s; nh.gewichte[2] = -(1.0 - t); nh.gewichte[3] = -t;
      nh.normale = n;
      nh.abstand = abstand;
      nh.ruhe = ruhe;
      aus.push_back(nh);
    };
    if (brute) alle_eintraege(netz.kanten.size(), paar);
    else gitter.je_eintrag(qk[e], paar);
  }
}

// :387-398
inline std::vector<Naehe> finde(const Stoffnetz& netz, const std::vector<V3>& x, const std::vector<double>& w, bool pd, bool kk,
                                bool brute = false) {
  Normalen nrm = Normalen::neu(netz, x);
  std::vector<Naehe> aus;
  double r_max = r_max_von(netz);
  if (pd) finde_punkt_dreieck(netz, x, w, nrm, r_max, aus, brute);
  if (kk) finde_kante_kante(netz, x, w, nrm, r_max, aus, brute);
  return aus;
}

// :609-621
struct Stoffwerte {
  std::vector<double> e_punkt, a_punkt, e_kante, a_kante, e_dreieck, a_dreieck, mu_punkt;
  std::vector<std::vector<std::pair<uint32_t, uint32_t>>> naht_partner;
};

inline size_t kante_von(const Stoffnetz& netz, size_t a, size_t b) {  // :796-799
  std::pair<uint32_t, uint32_t> key{(uint32_t)std::min(a, b), (uint32_t)std::max(a, b)};
  auto it = netz.kanten_index.find(key);
  return it == netz.kanten_index.end() ? 0 : it->second;
}

inline size_t dreieck_von(const Stoffnetz& netz, const uint32_t* p) {  // :790-794
  size_t e = kante_von(netz, p[0], p[1]);
  for (uint32_t t : netz.kanten_dreiecke[e]) {
    if (t == KEIN) continue;
    const auto& d = netz.dreiecke[t];
    if (d[0] == p[2] || d[1] == p[2] || d[2] == p[2]) return t;
  }
  return 0;
}

// :713-787
inline std::vector<Feder> federn(const Stoffnetz& netz, const Stoffwerte& werte, const std::vector<Naehe>& naehe, double steifigkeit,
                                 const std::vector<double>& masse, const std::vector<double>& w, const std::vector<double>& naht_ruhe,
                                 std::vector<V3>& v) {
  std::vector<Feder> aus;
  aus.reserve(naehe.size());
  for (const auto& p : naehe) {
    size_t idx[4] = {p.punkte[0], p.punkte[1], p.punkte[2], p.punkte[3]};
    double e, a, mu;
    if (p.art == PUNKT_DREIECK) {
      size_t t = dreieck_von(netz, &p.punkte[1]);
      e = werte.e_punkt[idx[0]] + werte.e_dreieck[t];
      a = werte.a_punkt[idx[0]] + werte.a_dreieck[t];
      mu = 0.5 * (werte.mu_punkt[idx[0]] + (werte.mu_punkt[idx[1]] + werte.mu_punkt[idx[2]] + werte.mu_punkt[idx[3]]) / 3.0);
    } else {
      size_t k1 = kante_von(netz, idx[0], idx[1]), k2 = kante_von(netz, idx[2], idx[3]);
      e = werte.e_kante[k1] + werte.e_kante[k2];
      a = werte.a_kante[k1] + werte.a_kante[k2];
      mu = 0.25 * (werte.mu_punkt[idx[0]] + werte.mu_punkt[idx[1]] + werte.mu_punkt[idx[2]] + werte.mu_punkt[idx[3]]);
    }
    double k = steifigkeit * e * a / 1e-6;
    double ruhe = p.ruhe;
    bool fertig = false;
    for (int q = 0; q < 4 && !fertig; ++q) {
      if (p.gewichte[q] <= 0.0) continue;
      for (const auto& pc : werte.naht_partner[idx[q]]) {
        bool auf_zweiter = false;
        for (int r = 0; r < 4; ++r) auf_zweiter = auf_zweiter || (p.gewichte[r] < 0.0 && p.punkte[r] == pc.first);
        double l0 = naht_ruhe[pc.second];
        if (auf_zweiter && std::isfinite(l0) && l0 < ruhe) { ruhe = l0; fertig = true; break; }
      }
    }
    double c = fmin_(ruhe - p.abstand, AUSLENKUNG_MAX);
    V3 n = p.normale;
    Feder fe;
    fe.kraft = mul(n, k * c);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) fe.block[i][j] = k * ((1.0 - mu) * n[i] * n[j] + (i == j ? mu : 0.0));
    double m1 = 0.0, m2 = 0.0;
    V3 dv{0.0, 0.0, 0.0};
    for (int q = 0; q < 4; ++q) {
      double g = p.gewichte[q];
      dv = add(dv, mul(v[idx[q]], g));
      if (g > 0.0) m1 += g * masse[idx[q]];
      else m2 -= g * masse[idx[q]];
    }
    double an = dot(dv, n) * ANNAEHERUNG_ANTEIL;
    if (an < 0.0 && m1 + m2 > 0.0) {
      V3 u = add(mul(n, an), mul(sub(dv, mul(n, an)), mu));
      for (int q = 0; q < 4; ++q) {
        double g = p.gewichte[q];
        if (w[idx[q]] <= 0.0) continue;
        double anteil = g > 0.0 ? m2 / (m1 + m2) : m1 / (m1 + m2);
        v[idx[q]] = sub(v[idx[q]], mul(u, g * anteil));
      }
    }
    for (int q = 0; q < 4; ++q) { fe.punkte[q] = p.punkte[q]; fe.gewichte[q] = p.gewichte[q]; }
    fe.punkte[4] = p.punkte[0];
    fe.gewichte[4] = 0.0;
    aus.push_back(fe);
  }
  return aus;
}

// implizit.rs:63-70: the pairs (first side, second side, weight) of a spring
template <typename F>
inline void feder_paare(const Feder& fe, F f) {
  for (int q = 0; q < 5; ++q)
    for (int r = 0; r < 5; ++r) {
      double gq = fe.gewichte[q], gr = fe.gewichte[r];
      if (gq > 0.0 && gr < 0.0) f(fe.punkte[q], fe.punkte[r], -gq * gr);
    }
}

inline V3 lerp(V3 a, V3 b, double t) { return add(mul(a, 1.0 - t), mul(b, t)); }

// :1156-1171
inline bool im_dreieck(V3 a, V3 b, V3 c, V3 p, double* uv) {
  V3 e1 = sub(b, a), e2 = sub(c, a), r = sub(p, a);
  double d11 = dot(e1, e1), d12 = dot(e1, e2), d22 = dot(e2, e2);
  double det = d11 * d22 - d12 * d12;
  if (det <= 1e-4 * d11 * d22) return false;
  double r1 = dot(r, e1), r2 = dot(r, e2);
  double u = (d22 * r1 - d12 * r2) / det;
  double v = (d11 * r2 - d12 * r1) / det;
  if (u < 0.0 || v < 0.0 || u + v > 1.0) return false;
  uv[0] = u;
  uv[1] = v;
  return true;
}

// :830-851
inline bool kante_durch_dreieck(V3 a, V3 b, const V3* v, V3 n, double* t_out, double* w) {
  double d0 = dot(n, sub(a, v[0])), d1 = dot(n, sub(b, v[0]));
  if ((d0 > 0.0 && d1 > 0.0) || (d0 < 0.0 && d1 < 0.0) || std::fabs(d1 - d0) < 1e-12) return false;
  double t = -d0 / (d1 - d0);
  if (std::fabs(t) < 1e-5 || std::fabs(t - 1.0) < 1e-5) return false;
  V3 p = lerp(a, b, t);
  V3 e1 = sub(v[1], v[0]), e2 = sub(v[2], v[0]);
  if (norm(e1) < STOSS_KANTE_MIN || norm(e2) < STOSS_KANTE_MIN) return false;
  double uv[2];
  if (!im_dreieck(v[0], v[1], v[2], p, uv)) return false;
  *t_out = t;
  w[0] = 1.0 - uv[0] - uv[1];
  w[1] = uv[0];
  w[2] = uv[1];
  return true;
}

// :858-907
inline std::vector<Schnitt> finde_schnitte(const Stoffnetz& netz, const std::vector<V3>& x, bool an_naehten) {
  Normalen nrm = Normalen::neu(netz, x);
  std::vector<Quader> qt;
  for (const auto& d : netz.dreiecke) {
    V3 p[3] = {x[d[0]], x[d[1]], x[d[2]]};
    qt.push_back(quader_punkte(p, 3, ZUSCHLAG_QUADER));
  }
  double laenge = 0.0;
  for (const auto& k : netz.kanten) laenge = fmax_(laenge, norm(sub(x[k[0]], x[k[1]])));
  Hashgitter gitter = Hashgitter::neu(qt, fmax_(laenge, 1e-3));
  std::vector<Schnitt> aus;
  for (size_t e = 0; e < netz.kanten.size(); ++e) {
    uint32_t a0 = netz.kanten[e][0], a1 = netz.kanten[e][1];
    const auto& kd = netz.kanten_dreiecke[e];
    if (a0 == a1 || kd[0] == KEIN) continue;
    V3 pa = x[a0], pb = x[a1];
    V3 pe[2] = {pa, pb};
    Quader qe = quader_punkte(pe, 2, ZUSCHLAG_QUADER);
    V3 ehat = einheit(sub(pb, pa));
    gitter.je_eintrag(qe, [&](uint32_t t, const int64_t* c) {
      const auto& d = netz.dreiecke[t];
      bool nachbar;
      if (an_naehten) {
        nachbar = d[0] == a0 || d[1] == a0 || d[2] == a0 || d[0] == a1 || d[1] == a1 || d[2] == a1;
      } else {
        nachbar = false;
        for (uint32_t i : d) nachbar = nachbar || netz.gleich_oder_partner(a0, i) || netz.gleich_oder_partner(a1, i);
      }
      if (nachbar || !ueberlappt(qt[t], qe) || !erste_zelle(qe, qt[t], gitter.zelle, c)) return;
      V3 n = nrm.dreieck[t];
      V3 v[3] = {x[d[0]], x[d[1]], x[d[2]]};
      double tk, w[3];
      if (!kante_durch_dreieck(pa, pb, v, n, &tk, w)) return;
      double cn = dot(n, ehat);
      if (std::fabs(cn) < 1e-6) return;
      V3 p = lerp(pa, pb, tk);
      V3 g{0.0, 0.0, 0.0};
      for (uint32_t k : kd) {
        if (k == KEIN) continue;
        const auto& dk = netz.dreiecke[k];
        uint32_t o = KEIN;
        for (uint32_t i : dk)
          if (i != a0 && i != a1) { o = i; break; }
        if (o == KEIN) continue;
        V3 m = einheit(cross(n, nrm.dreieck[k]));
        V3 r = sub(x[o], p);
        V3 r_quer = sub(r, mul(ehat, dot(ehat, r)));
        if (dot(m, r_quer) < 0.0) m = mul(m, -1.0);
        g = add(g, sub(m, mul(n, dot(ehat, m) / cn)));
      }
      double lg = norm(g);
      Schnitt s;
      s.kante = (uint32_t)e;
      s.dreieck = t;
      s.t = tk;
      for (int q = 0; q < 3; ++q) s.w[q] = w[q];
      s.richtung = einheit(g);
      s.gewicht = lg / std::sqrt(lg * lg + 2.25);
      aus.push_back(s);
    });
  }
  return aus;
}

// :910-924
inline size_t wurzel(std::vector<size_t>& eltern, size_t i) {
  while (eltern[i] != i) {
    eltern[i] = eltern[eltern[i]];
    i = eltern[i];
  }
  return i;
}
inline void vereinige(std::vector<size_t>& eltern, size_t a, size_t b) {
  size_t ra = wurzel(eltern, a), rb = wurzel(eltern, b);
  if (ra != rb) eltern[std::max(ra, rb)] = std::min(ra, rb);
}

// :931-1007
inline std::vector<size_t> gebiete(const Stoffnetz& netz, std::vector<Schnitt>& schnitte) {
  size_t n = schnitte.size();
  std::map<std::pair<uint32_t, uint32_t>, size_t> index;
  for (size_t i = 0; i < n; ++i) index.emplace(std::make_pair(schnitte[i].kante, schnitte[i].dreieck), i);
  auto finde_i = [&](uint32_t e, uint32_t t, std::vector<size_t>& out) {
    auto it = index.find({e, t});
    if (it != index.end()) out.push_back(it->second);
  };
  auto nachbardreiecke = [&](uint32_t t) {
    std::vector<uint32_t> v;
    for (uint32_t k : netz.dreieck_kanten[t])
      for (uint32_t u : netz.kanten_dreiecke[k])
        if (u != KEIN && u != t) v.push_back(u);
    return v;
  };
  std::vector<size_t> eltern(n);
  for (size_t i = 0; i < n; ++i) eltern[i] = i;
  std::vector<std::pair<size_t, size_t>> tausch;
  for (size_t i = 0; i < n; ++i) {
    uint32_t e = schnitte[i].kante, t = schnitte[i].dreieck;
    std::vector<uint32_t> kanten_dreiecke;
    for (uint32_t k : netz.kanten_dreiecke[e])
      if (k != KEIN) kanten_dreiecke.push_back(k);
    std::vector<uint32_t> andere_kanten;
    for (uint32_t k : kanten_dreiecke)
      for (uint32_t f : netz.dreieck_kanten[k])
        if (f != e) andere_kanten.push_back(f);
    std::vector<uint32_t> tn = nachbardreiecke(t);
    std::vector<size_t> gleich;
    for (uint32_t f : andere_kanten) finde_i(f, t, gleich);
    for (uint32_t u : tn) {
      finde_i(e, u, gleich);
      for (uint32_t f : andere_kanten) finde_i(f, u, gleich);
    }
    for (size_t j : gleich) vereinige(eltern, i, j);
    for (uint32_t f : netz.dreieck_kanten[t])
      for (uint32_t k : kanten_dreiecke) {
        std::vector<size_t> j;
        finde_i(f, k, j);
        for (size_t jj : j) tausch.push_back({i, jj});
      }
  }
  for (int durchgang = 0; durchgang < 2; ++durchgang) {
    std::vector<size_t> partner(n, std::numeric_limits<size_t>::max());
    for (const auto& ij : tausch) {
      size_t ri = wurzel(eltern, ij.first), rj = wurzel(eltern, ij.second);
      partner[ri] = std::min(partner[ri], rj);
      partner[rj] = std::min(partner[rj], ri);
    }
    for (const auto& ij : tausch) {
      size_t rj = wurzel(eltern, ij.second);
      if (partner[rj] != std::numeric_limits<size_t>::max()) vereinige(eltern, ij.first, partner[rj]);
    }
  }
  std::vector<size_t> wurzeln(n);
  for (size_t i = 0; i < n; ++i) wurzeln[i] = wurzel(eltern, i);
  std::vector<V3> summe(n, V3{0.0, 0.0, 0.0});
  for (size_t i = 0; i < n; ++i) summe[wurzeln[i]] = add(summe[wurzeln[i]], mul(schnitte[i].richtung, schnitte[i].gewicht));
  for (size_t i = 0; i < n; ++i) schnitte[i].richtung = einheit(summe[wurzeln[i]]);
  return wurzeln;
}

// :1014-1059
template <typename K>
inline std::vector<Feder> schubfedern(const Stoffnetz& netz, const Stoffwerte& werte, const std::vector<Schnitt>& schnitte,
                                      const std::vector<size_t>& wurzeln, double steifigkeit, std::vector<uint32_t>& folge,
                                      K eingeklemmt) {
  if (folge.size() != netz.dreiecke.size()) folge.assign(netz.dreiecke.size(), 1);
  std::vector<bool> hat(netz.dreiecke.size(), false);
  std::vector<bool> klemm(schnitte.size(), false);
  for (size_t i = 0; i < schnitte.size(); ++i) {
    const auto& s = schnitte[i];
    hat[s.dreieck] = true;
    if (eingeklemmt(netz.kanten[s.kante], netz.dreiecke[s.dreieck])) klemm[wurzeln[i]] = true;
  }
  for (size_t t = 0; t < folge.size(); ++t) folge[t] = hat[t] ? folge[t] + 1 : 1;
  std::vector<Feder> aus;
  for (size_t i = 0; i < schnitte.size(); ++i) {
    if (klemm[wurzeln[i]]) continue;
    const auto& s = schnitte[i];
    size_t e = s.kante, t = s.dreieck;
    double k = 2.0 * steifigkeit * (werte.e_kante[e] + werte.e_dreieck[t]) * (werte.a_kante[e] + werte.a_dreieck[t]) / 1e-6;
    V3 g = s.richtung;
    Feder fe;
    fe.kraft = mul(g, k * (double)std::min(folge[t], FOLGE_MAX) * SCHUB_HEBEL);
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) fe.block[a][b] = k * g[a] * g[b];
    fe.punkte[0] = netz.kanten[e][0]; fe.punkte[1] = netz.kanten[e][1];
    fe.punkte[2] = netz.dreiecke[t][0]; fe.punkte[3] = netz.dreiecke[t][1]; fe.punkte[4] = netz.dreiecke[t][2];
    fe.gewichte[0] = 1.0 - s.t; fe.gewichte[1] = s.t;
    fe.gewichte[2] = -s.w[0]; fe.gewichte[3] = -s.w[1]; fe.gewichte[4] = -s.w[2];
    aus.push_back(fe);
  }
  return aus;
}

// :1063-1086
inline std::vector<Stoss> ohne_durchdrungene(const Stoffnetz& netz, const std::vector<Stoss>& stoesse, const std::vector<Schnitt>& schnitte) {
  if (schnitte.empty()) return stoesse;
  std::set<std::pair<uint32_t, uint32_t>> paare;
  for (const auto& s : schnitte) paare.insert({s.kante, s.dreieck});
  std::vector<Stoss> aus;
  for (const auto& st : stoesse) {
    bool bleibt;
    if (st.art == PUNKT_DREIECK) {
      uint32_t t = (uint32_t)dreieck_von(netz, &st.punkte[1]);
      bool sticht = false;
      for (const auto& s : schnitte)
        sticht = sticht || (s.dreieck == t && (netz.kanten[s.kante][0] == st.punkte[0] || netz.kanten[s.kante][1] == st.punkte[0]));
      bleibt = !sticht;
    } else {
      size_t i = kante_von(netz, st.punkte[0], st.punkte[1]), j = kante_von(netz, st.punkte[2], st.punkte[3]);
      auto trifft = [&](size_t a, size_t b) {
        for (uint32_t t : netz.kanten_dreiecke[b]) {
          if (t == KEIN) continue;
          const auto& dk = netz.dreieck_kanten[t];
          if (paare.count({(uint32_t)a, t}) && (dk[0] == b || dk[1] == b || dk[2] == b)) return true;
        }
        return false;
      };
      bleibt = !(trifft(i, j) || trifft(j, i));
    }
    if (bleibt) aus.push_back(st);
  }
  return aus;
}

// :1116-1119
inline double ebenenabstand(V3 a, V3 b, V3 c, V3 p) { return dot(sub(p, a), einheit(cross(sub(b, a), sub(c, a)))); }

// :1128-1153
inline bool koplanar_zeit(const V3* x0, const V3* x1, double tol, double* out) {
  auto f = [&](double t) {
    V3 q[4];
    for (int k = 0; k < 4; ++k) q[k] = lerp(x0[k], x1[k], t);
    return ebenenabstand(q[0], q[1], q[2], q[3]);
  };
  double t0 = 0.0, t1 = 1.0;
  double d0 = f(0.0), d1 = f(1.0);
  for (int it = 0; it < 16; ++it) {
    if (d0 * d1 > 0.0) return false;
    double tau = t0 + (t1 - t0) * std::fabs(d0) / fmax_(std::fabs(d0) + std::fabs(d1), std::numeric_limits<double>::min());
    double d = f(tau);
    if (std::fabs(d) < tol || t1 - t0 < 1e-4) { *out = tau; return true; }
    if (d0 * d > 0.0) { t0 = tau; d0 = d; }
    else { t1 = tau; d1 = d; }
  }
  return false;
}

// :1173-1193
inline bool kanten_kreuzen(V3 a, V3 b, V3 c, V3 d, double* sw) {
  V3 e1 = sub(b, a), e2 = sub(d, c);
  double l1 = norm(e1), l2 = norm(e2);
  if (l1 < STOSS_KANTE_MIN || l2 < STOSS_KANTE_MIN) return false;
  V3 u1 = mul(e1, 1.0 / l1), u2 = mul(e2, 1.0 / l2);
  double g = dot(u1, u2);
  if (std::fabs(g) > 0.99999) return false;
  V3 q = sub(c, a);
  double nen = 1.0 - g * g;
  double s = (dot(q, u1) - g * dot(q, u2)) / nen;
  double w = (g * dot(q, u1) - dot(q, u2)) / nen;
  if (s <= 0.0 || s >= l1 || w <= 0.0 || w >= l2) return false;
  sw[0] = s / l1;
  sw[1] = w / l2;
  return true;
}

// :1196-1282
inline std::vector<Stoss> finde_stoesse(const Stoffnetz& netz, const std::vector<V3>& x0, const std::vector<V3>& x1,
                                        const std::vector<double>& w, bool pd, bool kk) {
  std::vector<Stoss> aus;
  auto weg = [&](const uint32_t* idx, int n) {
    V3 p[8];
    for (int q = 0; q < n; ++q) p[q] = x0[idx[q]];
    for (int q = 0; q < n; ++q) p[n + q] = x1[idx[q]];
    return quader_punkte(p, 2 * n, 1e-8);
  };
  double laenge = 0.0;
  for (const auto& k : netz.kanten) laenge = fmax_(laenge, norm(sub(x1[k[0]], x1[k[1]])));
  double zelle = fmax_(laenge, 1e-3);
  if (pd) {
    std::vector<Quader> qp;
    for (uint32_t v = 0; v < (uint32_t)x0.size(); ++v) qp.push_back(weg(&v, 1));
    Hashgitter gitter = Hashgitter::neu(qp, zelle);
    for (const auto& d : netz.dreiecke) {
      if (d[0] == d[1] || d[1] == d[2] || d[0] == d[2]) continue;
      Quader qt = weg(d.data(), 3);
      bool hat = false;
      Stoss best{};
      gitter.je_eintrag(qt, [&](uint32_t v, const int64_t* c) {
        for (uint32_t i : d)
          if (netz.gleich_oder_partner(v, i)) return;
        if (!ueberlappt(qp[v], qt) || !erste_zelle(qt, qp[v], zelle, c)) return;
        uint32_t idx[4] = {d[0], d[1], d[2], v};
        if (w[idx[0]] <= 0.0 && w[idx[1]] <= 0.0 && w[idx[2]] <= 0.0 && w[idx[3]] <= 0.0) return;
        V3 a0[4], a1[4];
        for (int k = 0; k < 4; ++k) { a0[k] = x0[idx[k]]; a1[k] = x1[idx[k]]; }
        double zeit;
        if (!koplanar_zeit(a0, a1, STOSS_ABSTAND * 0.1, &zeit)) return;
        if (hat && best.zeit <= zeit) return;
        V3 q[4];
        for (int k = 0; k < 4; ++k) q[k] = lerp(a0[k], a1[k], zeit);
        double uv[2];
        if (!im_dreieck(q[0], q[1], q[2], q[3], uv)) return;
        V3 n = einheit(cross(sub(q[1], q[0]), sub(q[2], q[0])));
        double g[3] = {1.0 - uv[0] - uv[1], uv[0], uv[1]};
        auto bew = [&](int k) { return sub(a1[k], a0[k]); };
        V3 rel = sub(bew(3), add(add(mul(bew(0), g[0]), mul(bew(1), g[1])), mul(bew(2), g[2])));
        double seite = dot(rel, n) <= 0.0 ? 1.0 : -1.0;
        best.art = PUNKT_DREIECK;
        best.punkte[0] = v; best.punkte[1] = d[0]; best.punkte[2] = d[1]; best.punkte[3] = d[2];
        best.lage[0] = uv[0]; best.lage[1] = uv[1];
        best.zeit = zeit;
        best.seite = seite;
        best.normale = mul(n, seite);
        hat = true;
      });
      if (hat) aus.push_back(best);
    }
  }
  if (kk) {
    std::vector<Quader> qk;
    for (const auto& k : netz.kanten) qk.push_back(weg(k.data(), 2));
    Hashgitter gitter = Hashgitter::neu(qk, zelle);
    for (size_t i = 0; i < netz.kanten.size(); ++i) {
      uint32_t i0 = netz.kanten[i][0], i1 = netz.kanten[i][1];
      if (i0 == i1) continue;
      bool hat = false;
      Stoss best{};
      gitter.je_eintrag(qk[i], [&](uint32_t jj, const int64_t* c) {
        size_t j = jj;
        if (j <= i || !ueberlappt(qk[i], qk[j]) || !erste_zelle(qk[i], qk[j], zelle, c)) return;
        uint32_t j0 = netz.kanten[j][0], j1 = netz.kanten[j][1];
        for (uint32_t q : {j0, j1})
          if (netz.gleich_oder_partner(q, i0) || netz.gleich_oder_partner(q, i1)) return;
        uint32_t idx[4] = {i1, j0, j1, i0};
        if (w[idx[0]] <= 0.0 && w[idx[1]] <= 0.0 && w[idx[2]] <= 0.0 && w[idx[3]] <= 0.0) return;
        V3 a0[4], a1[4];
        for (int k = 0; k < 4; ++k) { a0[k] = x0[idx[k]]; a1[k] = x1[idx[k]]; }
        double zeit;
        if (!koplanar_zeit(a0, a1, STOSS_ABSTAND * 0.1, &zeit)) return;
        if (hat && best.zeit <= zeit) return;
        V3 q[4];
        for (int k = 0; k < 4; ++k) q[k] = lerp(a0[k], a1[k], zeit);
        double sw[2];
        if (!kanten_kreuzen(q[3], q[0], q[1], q[2], sw)) return;
        V3 n = einheit(cross(sub(q[0], q[3]), sub(q[2], q[1])));
        auto bew = [&](int k) { return sub(a1[k], a0[k]); };
        V3 auf_i = add(mul(bew(3), 1.0 - sw[0]), mul(bew(0), sw[0]));
        V3 auf_j = add(mul(bew(1), 1.0 - sw[1]), mul(bew(2), sw[1]));
        double seite = dot(sub(auf_i, auf_j), n) <= 0.0 ? 1.0 : -1.0;
        best.art = KANTE_KANTE;
        best.punkte[0] = i0; best.punkte[1] = i1; best.punkte[2] = j0; best.punkte[3] = j1;
        best.lage[0] = sw[0]; best.lage[1] = sw[1];
        best.zeit = zeit;
        best.seite = seite;
        best.normale = mul(n, seite);
        hat = true;
      });
      if (hat) aus.push_back(best);
    }
  }
  return aus;
}

// :1288-1343
inline size_t stossantwort(const std::vector<Stoss>& stoesse, std::vector<V3>& x, const std::vector<double>& masse,
                           const std::vector<double>& w, size_t iterationen) {
  size_t zahl = 0;
  for (size_t it = 0; it < iterationen; ++it) {
    std::vector<V3> dx(x.size(), V3{0.0, 0.0, 0.0});
    std::vector<uint32_t> cnt(x.size(), 0);
    for (const auto& s : stoesse) {
      size_t p0 = s.punkte[0], p1 = s.punkte[1], p2 = s.punkte[2], p3 = s.punkte[3];
      std::pair<size_t, double> erste[2], zweite[3];
      int ne, nz;
      V3 n;
      double d;
      if (s.art == PUNKT_DREIECK) {
        n = mul(einheit(cross(sub(x[p2], x[p1]), sub(x[p3], x[p1]))), s.seite);
        double g[3] = {1.0 - s.lage[0] - s.lage[1], s.lage[0], s.lage[1]};
        erste[0] = {p0, 1.0}; ne = 1;
        zweite[0] = {p1, g[0]}; zweite[1] = {p2, g[1]}; zweite[2] = {p3, g[2]}; nz = 3;
        d = dot(n, sub(x[p0], x[p1]));
      } else {
        n = s.normale;
        V3 a = lerp(x[p0], x[p1], s.lage[0]), b = lerp(x[p2], x[p3], s.lage[1]);
        erste[0] = {p0, 1.0 - s.lage[0]}; erste[1] = {p1, s.lage[0]}; ne = 2;
        zweite[0] = {p2, 1.0 - s.lage[1]}; zweite[1] = {p3, s.lage[1]}; nz = 2;
        d = dot(n, sub(a, b));
      }
      double tief = STOSS_ABSTAND - d;
      if (tief <= 0.0) continue;
      double m1 = 0.0, m2 = 0.0;
      for (int q = 0; q < ne; ++q) m1 += erste[q].second * masse[erste[q].first];
      for (int q = 0; q < nz; ++q) m2 += zweite[q].second * masse[zweite[q].first];
      bool fest1 = true, fest2 = true;
      for (int q = 0; q < ne; ++q) fest1 = fest1 && w[erste[q].first] <= 0.0;
      for (int q = 0; q < nz; ++q) fest2 = fest2 && w[zweite[q].first] <= 0.0;
      double alpha = fest2 ? 1.0 : (fest1 ? 0.0 : m2 / (m1 + m2));
      for (int q = 0; q < ne; ++q) {
        size_t i = erste[q].first;
        if (w[i] > 0.0) { dx[i] = add(dx[i], mul(n, tief * alpha)); cnt[i] += 1; }
      }
      for (int q = 0; q < nz; ++q) {
        size_t i = zweite[q].first;
        if (w[i] > 0.0) { dx[i] = sub(dx[i], mul(n, tief * (1.0 - alpha))); cnt[i] += 1; }
      }
      zahl += 1;
    }
    bool bewegt = false;
    for (size_t i = 0; i < x.size(); ++i)
      if (cnt[i] > 0) { x[i] = add(x[i], mul(dx[i], 1.0 / (double)cnt[i])); bewegt = true; }
    if (!bewegt) break;
  }
  return zahl;
}
