// Seam pairing by length fraction: the boundary plan of klon3d-mesh
// (crates/klon3d-mesh/src/rand.rs) and the stitched point pairs it yields
// (lib.rs:351-381). Host code, run once per pattern; every decision (merging
// stations within fang_mm, splitting a boundary piece) is taken from FP64
// values computed in the reference's order, so the plan, and the integer
// pairs that come out of it, are the reference's to the bit.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <stdexcept>
#include <set>
#include <string>
#include <utility>
#include <vector>

namespace naht {

static double abstand(double ax, double ay, double bx, double by) { return std::hypot(ax - bx, ay - by); }

// Distance from p to the segment a-b (rand.rs:46-54).
static double punkt_strecke(double px, double py, double ax, double ay, double bx, double by) {
  const double dx = bx - ax, dy = by - ay;
  const double l2 = dx * dx + dy * dy;
  if (l2 <= 0.0) return abstand(px, py, ax, ay);
  double t = ((px - ax) * dx + (py - ay) * dy) / l2;
  t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
  return abstand(px, py, ax + t * dx, ay + t * dy);
}

// A boundary edge as a polyline with arc length (Kantenzug, rand.rs:56-130):
// points closer than 1e-9 mm to the previous one are dropped.
struct Zug {
  std::vector<double> x, y, kum;
  double laenge = 0.0;

  void aus(const double* p, int n) {
    for (int i = 0; i < n; ++i) {
      const double px = p[2 * i], py = p[2 * i + 1];
      if (x.empty()) {
        kum.push_back(0.0);
      } else {
        const double d = abstand(x.back(), y.back(), px, py);
        if (d <= 1e-9) continue;
        kum.push_back(kum.back() + d);
      }
      x.push_back(px);
      y.push_back(py);
    }
    laenge = kum.empty() ? 0.0 : kum.back();
  }
  // the first index whose arc length exceeds s (partition_point of x <= s)
  size_t nach(double s) const { return (size_t)(std::upper_bound(kum.begin(), kum.end(), s) - kum.begin()); }
  void ort(double s, double* ox, double* oy) const {
    if (s <= 0.0) {
      *ox = x[0];
      *oy = y[0];
      return;
    }
    if (s >= laenge) {
      *ox = x.back();
      *oy = y.back();
      return;
    }
    const size_t i = nach(s) - 1;
    const double t = (s - kum[i]) / (kum[i + 1] - kum[i]);
    *ox = x[i] + t * (x[i + 1] - x[i]);
    *oy = y[i] + t * (y[i + 1] - y[i]);
  }
  // chord from s0 to s1: the largest distance of the polyline's points from
  // it, its length loss, and the arc length of the farthest point (or -1)
  void sehne(double s0, double s1, double* hoehe, double* verlust, double* ecke, bool* hat_ecke) const {
    const double a0 = std::fmin(s0, s1), a1 = std::fmax(s0, s1);
    double ax, ay, bx, by;
    ort(a0, &ax, &ay);
    ort(a1, &bx, &by);
    double h_max = 0.0;
    bool gefunden = false;
    double e = 0.0;
    for (size_t i = nach(a0); i < kum.size(); ++i) {
      if (kum[i] >= a1) break;
      const double h = punkt_strecke(x[i], y[i], ax, ay, bx, by);
      if (!gefunden || h > h_max) {
        h_max = h;
        e = kum[i];
        gefunden = true;
      }
    }
    *hoehe = h_max;
    *verlust = (a1 - a0) - abstand(ax, ay, bx, by);
    *ecke = e;
    *hat_ecke = gefunden;
  }
};

struct Teil {
  std::vector<Zug> kanten;
};

// A piece of a boundary edge between two stations, s0 -> s1 in seam
// direction (Glied, rand.rs:194-218).
struct Glied {
  int teil, kante;
  double s0, s1;
  double s(double f) const {
    if (f <= 0.0) return s0;
    if (f >= 1.0) return s1;
    return s0 + f * (s1 - s0);
  }
  double laenge() const { return std::fabs(s1 - s0); }
};

// Pieces split together: a free edge (one piece) or a seam section (two);
// fractions 0 = f0 < ... < fm = 1 (Gruppe, rand.rs:220-227).
struct Gruppe {
  std::vector<Glied> glieder;
  std::vector<double> f;
};

struct Param {
  double kante_mm, sehne_mm, laenge_mm, fang_mm, min_rand_mm;
};

// Even split to at most kante_mm, then every piece whose chord strays or
// loses too much on any of the group's pieces is split at the polyline point
// farthest from the chord, else in the middle (unterteile, rand.rs:329-351).
static void unterteile(Gruppe& g, const std::vector<Teil>& teile, const Param& p) {
  double lmax = 0.0;
  for (const Glied& gl : g.glieder) lmax = std::fmax(lmax, gl.laenge());
  size_t m = (size_t)std::ceil(lmax / p.kante_mm);
  if (m < 1) m = 1;
  g.f.clear();
  for (size_t j = 0; j <= m; ++j) g.f.push_back((double)j / (double)m);
  size_t j = 0;
  while (j + 1 < g.f.size()) {
    const double f0 = g.f[j], f1 = g.f[j + 1];
    bool grob = false, hat = false;
    double ecke = 0.0;
    for (const Glied& gl : g.glieder) {
      const Zug& kz = teile[gl.teil].kanten[gl.kante];
      const double s0 = gl.s(f0), s1 = gl.s(f1);
      double hoehe, verlust, e;
      bool he;
      kz.sehne(s0, s1, &hoehe, &verlust, &e, &he);
      if (hoehe > p.sehne_mm || verlust > p.laenge_mm * std::fabs(s1 - s0) / kz.laenge) {
        grob = true;
        if (!hat && he) {
          ecke = (e - gl.s0) / (gl.s1 - gl.s0);
          hat = true;
        }
      }
    }
    bool teilen = false;
    double ft = 0.0;
    if (grob) {
      if (hat && (ecke - f0) * lmax >= p.min_rand_mm && (f1 - ecke) * lmax >= p.min_rand_mm) {
        teilen = true;
        ft = ecke;
      } else if ((f1 - f0) * lmax >= 2.0 * p.min_rand_mm) {
        teilen = true;
        ft = 0.5 * (f0 + f1);
      }
    }
    if (teilen) g.f.insert(g.f.begin() + (long)(j + 1), ft);
    else ++j;
  }
}

// Linear map over the anchor pairs (rho_A, rho_B); rueck: from B to A
// (abbilden, rand.rs:353-359).
static double abbilden(const std::vector<std::pair<double, double>>& anker, double x, bool rueck) {
  auto paar = [&](size_t k) { return rueck ? std::make_pair(anker[k].second, anker[k].first) : anker[k]; };
  long k = -1;
  for (long i = (long)anker.size() - 1; i >= 0; --i)
    if (paar((size_t)i).first <= x) {
      k = i;
      break;
    }
  if (k < 0) k = 0;
  if (k > (long)anker.size() - 2) k = (long)anker.size() - 2;
  const auto a = paar((size_t)k), b = paar((size_t)k + 1);
  return a.second + (x - a.first) / (b.first - a.first) * (b.second - a.second);
}

// One edge of a seam side in seam direction (Stueck, rand.rs:278-295).
struct Stueck {
  int teil, kante;
  bool rueck;
  double laenge;
  int teilstueck;
  double anfang() const { return rueck ? laenge : 0.0; }
  double ende() const { return rueck ? 0.0 : laenge; }
};

enum { ENDE = 0, GEMEINSAM = 1, NUR_A = 2, NUR_B = 3 };

struct Roh {
  double ra, rb;
  int ga, gb;  // the edge boundary before run piece i, or -1 (a new point)
  int art;
};

struct Abschnitt {
  int gruppe;
  double sigma_a[2], sigma_b[2];
};

struct Teilnahtplan {
  int teil_a, teil_b;
  std::vector<Abschnitt> abschnitte;
};

struct Stelle {
  int art;
  int knoten[4];  // piece, edge (start node) on A; on B; -1 for a new point
  double rho_a, rho_b, versatz_mm;
};

struct Nahtplan {
  double la, lb;
  std::vector<Stelle> stellen;
  std::vector<Teilnahtplan> teilnaehte;
};

// Where a station lies on a side: (piece before, s) and (piece after, s)
// (orte, rand.rs:373-384).
static void orte(const std::vector<Stueck>& s, const std::vector<double>& kum, double l, double rho, int grenze,
                 int* vor_k, double* vor_s, int* nach_k, double* nach_s) {
  *vor_k = *nach_k = -1;
  if (grenze >= 0) {
    if (grenze > 0) {
      *vor_k = grenze - 1;
      *vor_s = s[(size_t)grenze - 1].ende();
    }
    if (grenze < (int)s.size()) {
      *nach_k = grenze;
      *nach_s = s[(size_t)grenze].anfang();
    }
    return;
  }
  const double sigma = rho * l;
  size_t k = (size_t)(std::upper_bound(kum.begin(), kum.end(), sigma) - kum.begin());
  k = k == 0 ? 0 : k - 1;
  if (k > s.size() - 1) k = s.size() - 1;
  double d = sigma - kum[k];
  d = d < 0.0 ? 0.0 : (d > s[k].laenge ? s[k].laenge : d);
  const double pos = s[k].rueck ? s[k].laenge - d : d;
  *vor_k = *nach_k = (int)k;
  *vor_s = *nach_s = pos;
}

// The node a station at edge boundary i of a side sits on (stellenname,
// rand.rs:386-398): piece and the edge whose start node it is.
static void knoten_bei(const std::vector<Stueck>& s, const std::vector<Teil>& teile, int i, int* teil, int* kante) {
  auto anfang = [&](const Stueck& x) { return x.rueck ? x.kante + 1 : x.kante; };
  auto ende = [&](const Stueck& x) { return x.rueck ? x.kante : x.kante + 1; };
  int t, k;
  if (i == (int)s.size()) {
    t = s[(size_t)i - 1].teil;
    k = ende(s[(size_t)i - 1]);
  } else {
    t = s[(size_t)i].teil;
    k = anfang(s[(size_t)i]);
  }
  *teil = t;
  *kante = k % (int)teile[(size_t)t].kanten.size();
}

static bool naht_planen(const std::vector<Stueck>& a, const std::vector<Stueck>& b, const std::vector<Teil>& teile,
                        const Param& p, std::vector<Gruppe>& gruppen, Nahtplan& out, std::string& fehler) {
  auto kum = [](const std::vector<Stueck>& s) {
    std::vector<double> v{0.0};
    for (const Stueck& x : s) v.push_back(v.back() + x.laenge);
    return v;
  };
  const std::vector<double> ka = kum(a), kb = kum(b);
  const double la = ka.back(), lb = kb.back();
  const double lm = 0.5 * (la + lb);
  std::vector<double> ra, rb;
  for (size_t i = 1; i < a.size(); ++i) ra.push_back(ka[i] / la);
  for (size_t i = 1; i < b.size(); ++i) rb.push_back(kb[i] / lb);
  auto naechst = [](double x, const std::vector<double>& v) -> long {
    long best = -1;
    double bd = 0.0;
    for (size_t i = 0; i < v.size(); ++i) {
      const double d = std::fabs(v[i] - x);
      if (best < 0 || d < bd) {
        best = (long)i;
        bd = d;
      }
    }
    return best;
  };
  const double fang = std::fmax(p.fang_mm, 1e-3);
  std::vector<std::pair<double, double>> anker{{0.0, 0.0}};
  std::vector<std::pair<int, int>> gemeinsam;
  for (size_t i = 0; i < ra.size(); ++i) {
    const double x = ra[i];
    const long j = naechst(x, rb);
    if (j < 0) continue;
    const auto letzt = anker.back();
    if (std::fabs(rb[(size_t)j] - x) * lm <= fang && naechst(rb[(size_t)j], ra) == (long)i && x > letzt.first &&
        rb[(size_t)j] > letzt.second) {
      anker.push_back({x, rb[(size_t)j]});
      gemeinsam.push_back({(int)i, (int)j});
    }
  }
  anker.push_back({1.0, 1.0});

  std::vector<Roh> st{{0.0, 0.0, 0, 0, ENDE}};
  std::set<int> ia, jb;
  for (const auto& g : gemeinsam) {
    ia.insert(g.first);
    jb.insert(g.second);
    st.push_back({ra[(size_t)g.first], rb[(size_t)g.second], g.first + 1, g.second + 1, GEMEINSAM});
  }
  for (size_t i = 0; i < ra.size(); ++i)
    if (!ia.count((int)i)) st.push_back({ra[i], abbilden(anker, ra[i], false), (int)i + 1, -1, NUR_A});
  for (size_t j = 0; j < rb.size(); ++j)
    if (!jb.count((int)j)) st.push_back({abbilden(anker, rb[j], true), rb[j], -1, (int)j + 1, NUR_B});
  st.push_back({1.0, 1.0, (int)a.size(), (int)b.size(), ENDE});
  std::stable_sort(st.begin(), st.end(), [](const Roh& x, const Roh& y) { return x.ra < y.ra; });
  for (size_t w = 0; w + 1 < st.size(); ++w) {
    if ((st[w + 1].ra - st[w].ra) * la < 1e-3 || (st[w + 1].rb - st[w].rb) * lb < 1e-3) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "seam stations at %.3f mm (side A) coincide", st[w].ra * la);
      fehler = buf;
      return false;
    }
  }

  struct Roh_ab {
    int ts_a, ts_b, teil_a, teil_b;
    Abschnitt ab;
  };
  std::vector<Roh_ab> abschnitte;
  for (size_t w = 0; w + 1 < st.size(); ++w) {
    int sa, sa1, sb, sb1, x0, x1;
    double s0a, s1a, s0b, s1b, y0, y1;
    orte(a, ka, la, st[w].ra, st[w].ga, &x0, &y0, &sa, &s0a);
    orte(a, ka, la, st[w + 1].ra, st[w + 1].ga, &sa1, &s1a, &x1, &y1);
    orte(b, kb, lb, st[w].rb, st[w].gb, &x0, &y0, &sb, &s0b);
    orte(b, kb, lb, st[w + 1].rb, st[w + 1].gb, &sb1, &s1b, &x1, &y1);
    if (sa < 0 || sa1 < 0 || sb < 0 || sb1 < 0 || sa != sa1 || sb != sb1) {
      char buf[160];
      std::snprintf(buf, sizeof buf, "the section from %.3f mm runs across an edge boundary", st[w].ra * la);
      fehler = buf;
      return false;
    }
    Gruppe g;
    g.glieder.push_back({a[(size_t)sa].teil, a[(size_t)sa].kante, s0a, s1a});
    g.glieder.push_back({b[(size_t)sb].teil, b[(size_t)sb].kante, s0b, s1b});
    unterteile(g, teile, p);
    gruppen.push_back(g);
    Abschnitt ab{(int)gruppen.size() - 1, {st[w].ra * la, st[w + 1].ra * la}, {st[w].rb * lb, st[w + 1].rb * lb}};
    abschnitte.push_back({a[(size_t)sa].teilstueck, b[(size_t)sb].teilstueck, a[(size_t)sa].teil, b[(size_t)sb].teil, ab});
  }
  out.teilnaehte.clear();
  bool hat_letzt = false;
  std::pair<int, int> letzt{0, 0};
  for (const Roh_ab& x : abschnitte) {
    const std::pair<int, int> schl{x.ts_a, x.ts_b};
    if (!hat_letzt || letzt != schl) {
      out.teilnaehte.push_back({x.teil_a, x.teil_b, {}});
      letzt = schl;
      hat_letzt = true;
    }
    out.teilnaehte.back().abschnitte.push_back(x.ab);
  }
  out.la = la;
  out.lb = lb;
  out.stellen.clear();
  for (const Roh& r : st) {
    Stelle s{r.art, {-1, -1, -1, -1}, r.ra, r.rb, (r.ra - r.rb) * lm};
    if (r.ga >= 0) knoten_bei(a, teile, r.ga, &s.knoten[0], &s.knoten[1]);
    if (r.gb >= 0) knoten_bei(b, teile, r.gb, &s.knoten[2], &s.knoten[3]);
    out.stellen.push_back(s);
  }
  return true;
}

// The boundary points of one piece, keyed by (edge, bits of s): the end of
// an edge is the start of the next (schluessel, rand.rs:536-545).
struct Randindex {
  std::map<std::pair<int, uint64_t>, std::pair<int, double>> menge;
  std::map<std::pair<int, uint64_t>, int> index;
};

static std::pair<int, uint64_t> schluessel(const Teil& t, int kante, double s) {
  if (s >= t.kanten[(size_t)kante].laenge) return {(kante + 1) % (int)t.kanten.size(), 0ull};
  if (s <= 0.0) return {kante, 0ull};
  uint64_t b;
  std::memcpy(&b, &s, sizeof b);
  return {kante, b};
}

static void randindex(const std::vector<Teil>& teile, const std::vector<Gruppe>& gruppen, int teil, Randindex& ri) {
  const Teil& t = teile[(size_t)teil];
  for (const Gruppe& g : gruppen)
    for (const Glied& gl : g.glieder) {
      if (gl.teil != teil) continue;
      for (double f : g.f) {
        const double s = gl.s(f);
        const auto k = schluessel(t, gl.kante, s);
        if (!ri.menge.count(k)) ri.menge[k] = k.second == 0 ? std::make_pair(k.first, 0.0) : std::make_pair(gl.kante, s);
      }
    }
  int i = 0;
  for (const auto& e : ri.menge) ri.index[e.first] = i++;
}

}  // namespace naht

static int naht_planen_c(const gc_naht_eingabe* ein, gc_naht_ausgabe* aus) {
  using namespace naht;
  aus->fehler[0] = '\0';
  auto fail = [&](const std::string& m) {
    std::snprintf(aus->fehler, sizeof aus->fehler, "%s", m.c_str());
    return 1;
  };
  const Param p{ein->kante_mm, ein->sehne_mm, ein->laenge_mm, ein->fang_mm, ein->min_rand_mm};
  std::vector<Teil> teile((size_t)ein->n_teile);
  for (int t = 0; t < ein->n_teile; ++t) {
    for (int e = ein->teil_kanten[t]; e < ein->teil_kanten[t + 1]; ++e) {
      Zug z;
      z.aus(ein->punkte + 2 * (size_t)ein->kante_punkte[e], ein->kante_punkte[e + 1] - ein->kante_punkte[e]);
      if (z.x.size() < 2 || z.laenge <= 1e-6) return fail("piece " + std::to_string(t) + ": boundary edge " +
                                                             std::to_string(e - ein->teil_kanten[t]) + " has no length");
      teile[(size_t)t].kanten.push_back(std::move(z));
    }
    if (teile[(size_t)t].kanten.empty()) return fail("piece " + std::to_string(t) + " has no boundary edge");
  }
  std::vector<Gruppe> gruppen;
  std::set<std::pair<int, int>> belegt;
  std::vector<Nahtplan> plaene((size_t)ein->n_naehte);
  for (int n = 0; n < ein->n_naehte; ++n) {
    std::vector<Stueck> seiten[2];
    for (int sd = 0; sd < 2; ++sd) {
      const int s = 2 * n + sd;
      for (int r = ein->seite_stuecke[s]; r < ein->seite_stuecke[s + 1]; ++r) {
        const int t = ein->stueck_teil[r];
        if (t < 0 || t >= ein->n_teile) return fail("seam " + std::to_string(n) + ": no piece " + std::to_string(t));
        for (int q = ein->stueck_kanten[r]; q < ein->stueck_kanten[r + 1]; ++q) {
          const int k = ein->kanten[q];
          if (k < 0 || k >= (int)teile[(size_t)t].kanten.size())
            return fail("seam " + std::to_string(n) + ": piece " + std::to_string(t) + " has no edge " + std::to_string(k));
          seiten[sd].push_back({t, k, ein->stueck_rueck[r] != 0, teile[(size_t)t].kanten[(size_t)k].laenge,
                                r - ein->seite_stuecke[s]});
        }
      }
      if (seiten[sd].empty()) return fail("seam " + std::to_string(n) + ": a side without edges");
    }
    for (int sd = 0; sd < 2; ++sd)
      for (const Stueck& x : seiten[sd])
        if (!belegt.insert({x.teil, x.kante}).second)
          return fail("seam " + std::to_string(n) + ": piece " + std::to_string(x.teil) + " edge " +
                      std::to_string(x.kante) + " is in two seams");
    std::string f;
    if (!naht_planen(seiten[0], seiten[1], teile, p, gruppen, plaene[(size_t)n], f))
      return fail("seam " + std::to_string(n) + ": " + f);
  }
  for (int t = 0; t < ein->n_teile; ++t)
    for (int k = 0; k < (int)teile[(size_t)t].kanten.size(); ++k)
      if (!belegt.count({t, k})) {
        Gruppe g;
        g.glieder.push_back({t, k, 0.0, teile[(size_t)t].kanten[(size_t)k].laenge});
        unterteile(g, teile, p);
        gruppen.push_back(g);
      }
  std::vector<Randindex> ri((size_t)ein->n_teile);
  int64_t n_rand = 0, n_stellen = 0, n_tn = 0, n_paare = 0;
  for (int t = 0; t < ein->n_teile; ++t) {
    randindex(teile, gruppen, t, ri[(size_t)t]);
    n_rand += (int64_t)ri[(size_t)t].menge.size();
  }
  for (const Nahtplan& np : plaene) {
    n_stellen += (int64_t)np.stellen.size();
    n_tn += (int64_t)np.teilnaehte.size();
    for (const Teilnahtplan& tn : np.teilnaehte)
      for (size_t k = 0; k < tn.abschnitte.size(); ++k)
        n_paare += (int64_t)gruppen[(size_t)tn.abschnitte[k].gruppe].f.size() - (k > 0 ? 1 : 0);
  }
  aus->n_rand = n_rand;
  aus->n_stellen = n_stellen;
  aus->n_teilnaehte = n_tn;
  aus->n_paare = n_paare;
  if (n_rand > aus->kap_rand || n_stellen > aus->kap_stellen || n_tn > aus->kap_teilnaehte || n_paare > aus->kap_paare)
    return 2;
  int64_t r = 0;
  int kante0 = 0;
  for (int t = 0; t < ein->n_teile; ++t) {
    aus->rand_teil[t] = (int32_t)r;
    const Teil& tt = teile[(size_t)t];
    for (const auto& e : ri[(size_t)t].menge) {
      aus->rand_kante[r] = e.second.first;
      aus->rand_s[r] = e.second.second;
      tt.kanten[(size_t)e.second.first].ort(e.second.second, &aus->rand_xy[2 * r], &aus->rand_xy[2 * r + 1]);
      ++r;
    }
    for (int k = 0; k < (int)tt.kanten.size(); ++k) aus->rand_knoten[kante0 + k] = ri[(size_t)t].index.at({k, 0ull});
    kante0 += (int)tt.kanten.size();
  }
  aus->rand_teil[ein->n_teile] = (int32_t)r;
  int64_t si = 0, ti = 0, pi = 0;
  for (int n = 0; n < ein->n_naehte; ++n) {
    const Nahtplan& np = plaene[(size_t)n];
    aus->naht_laengen_mm[2 * n] = np.la;
    aus->naht_laengen_mm[2 * n + 1] = np.lb;
    aus->naht_stellen[n] = (int32_t)si;
    for (const Stelle& s : np.stellen) {
      aus->stelle_art[si] = s.art;
      for (int q = 0; q < 4; ++q) aus->stelle_knoten[4 * si + q] = s.knoten[q];
      aus->stelle_rho[2 * si] = s.rho_a;
      aus->stelle_rho[2 * si + 1] = s.rho_b;
      aus->stelle_versatz_mm[si] = s.versatz_mm;
      ++si;
    }
    aus->naht_teilnaehte[n] = (int32_t)ti;
    for (const Teilnahtplan& tn : np.teilnaehte) {
      aus->teilnaht_teile[2 * ti] = tn.teil_a;
      aus->teilnaht_teile[2 * ti + 1] = tn.teil_b;
      aus->teilnaht_paare[ti] = (int32_t)pi;
      const Randindex &ria = ri[(size_t)tn.teil_a], &rib = ri[(size_t)tn.teil_b];
      for (size_t k = 0; k < tn.abschnitte.size(); ++k) {
        const Abschnitt& ab = tn.abschnitte[k];
        const Gruppe& g = gruppen[(size_t)ab.gruppe];
        for (size_t j = 0; j < g.f.size(); ++j) {
          if (k > 0 && j == 0) continue;
          const double f = g.f[j];
          const Glied &ga = g.glieder[0], &gb = g.glieder[1];
          aus->paar_punkte[2 * pi] = ria.index.at(schluessel(teile[(size_t)ga.teil], ga.kante, ga.s(f)));
          aus->paar_punkte[2 * pi + 1] = rib.index.at(schluessel(teile[(size_t)gb.teil], gb.kante, gb.s(f)));
          aus->paar_rho[2 * pi] = (ab.sigma_a[0] + f * (ab.sigma_a[1] - ab.sigma_a[0])) / np.la;
          aus->paar_rho[2 * pi + 1] = (ab.sigma_b[0] + f * (ab.sigma_b[1] - ab.sigma_b[0])) / np.lb;
          ++pi;
        }
      }
      ++ti;
    }
  }
  aus->naht_stellen[ein->n_naehte] = (int32_t)si;
  aus->naht_teilnaehte[ein->n_naehte] = (int32_t)ti;
  aus->teilnaht_paare[ti] = (int32_t)pi;
  return 0;
}

extern "C" int gc_naht_planen(const gc_naht_eingabe* ein, gc_naht_ausgabe* aus) {
  if (!ein || !aus || ein->n_teile < 0 || ein->n_naehte < 0 || !(ein->kante_mm > 0.0)) return -1;
  try {
    return naht_planen_c(ein, aus);
  } catch (const std::exception& e) {
    std::snprintf(aus->fehler, sizeof aus->fehler, "seam plan: %s", e.what());
    return 1;
  }
}
