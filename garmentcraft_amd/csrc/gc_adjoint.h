// gc_adjoint.h — reverse mode of one implicit substep (the discrete adjoint).
//
// The reference has no adjoint; this is the gradient of the step that
// ImplizitCg::schritt computes (implizit.rs:899-1038), with every discrete
// decision of the step held as the forward pass took it: the contact set after
// the release pass, stick or slip, path entries, push-outs. Within those
// decisions the step is smooth, and its exact vector-Jacobian product is:
//
//   forward:  b = h (f(x) - h H(x) v),  A = M(1 + h c) + h^2 H(x),
//             [A N; N^T 0] (dv, -lambda) = (b, c(x, v)),   lambda = h fn,
//             (x', v') = update(x, v, dv, fn, mu)     (friction, x += h v, entry, push-out)
//   reverse:  (gx', gv') -> per point the update's VJP gives gx, gv, g_dv, g_fn, g_mu;
//             the adjoint solve [A N; N^T 0] (mu, nu) = (g_dv, -g_fn / h) is the
//             step's own modified PCG with right-hand side g_dv and normal
//             targets -g_fn / h (A is symmetric); nu = n.(g_dv - A mu);
//             dL += mu.db - mu.dA dv + lambda mu.dn - nu dv.dn + nu dc,
//             i.e. per element the gradient of
//               psi_e(x_e, theta_e) = h mu_e.f_e - h^2 mu_e.H_e (v + dv)_e,
//             per held point of chi_i(x, v) = n(x).(lambda mu - nu dv) + nu c(x, v),
//             and gv += -h c m mu - h^2 H mu.
//
// Every small function (an element, a point's update, a contact) is the
// forward's own templated code evaluated in dual numbers, once per input
// direction; its value part repeats the forward's FP64 operations, so the
// branches are the forward's. Self-collision springs and impacts are not
// differentiated (the adjoint step refuses a context with them switched on).
#pragma once
#include "gc_dual.h"
#include "gc_kernels.h"

// Device buffers of the adjoint, FP64.
struct AdjDev {
  V3<double>* gx;          // dL/dx: in, of the state after the step; out, of the state before
  V3<double>* gv;          // dL/dv, the same
  V3<double>* x0;          // state before the step
  V3<double>* v0;
  V3<double>* gx_p;        // the point update's part of dL/dx, dL/dv
  V3<double>* gv_p;
  V3<double>* gdv;         // dL/d(dv): right-hand side of the adjoint solve
  double* ziel;            // normal target of mu at held points: -dL/dfn / h
  double* nu;              // per held point: n.(g_dv - A mu)
  V3<double>* e_kante;     // per constraint and corner: d psi / dx
  V3<double>* e_biegung;
  V3<double>* e_membran;
  double* d_kante;         // accumulated parameter gradients: per distance constraint dL/dk,
  double* d_biegung;       // per bending constraint dL/dk,
  double* d_membran;       // per membrane dL/d(E_k, E_s, G),
  double* d_reibung;       // per point dL/d(friction coefficient)
  unsigned long long* zaehler;  // [0] points whose recomputed update differs from the step's, [1] path entries
};

GC_HD V3<Dual> saat(V3<double> a, int k, int basis) {
  return {Dual(a.x, k == basis ? 1.0 : 0.0), Dual(a.y, k == basis + 1 ? 1.0 : 0.0), Dual(a.z, k == basis + 2 ? 1.0 : 0.0)};
}
GC_HD V3<Dual> konst(V3<double> a) { return {Dual(a.x), Dual(a.y), Dual(a.z)}; }
GC_HD V3<double> wert3(V3<Dual> a) { return {a.x.v, a.y.v, a.z.v}; }
GC_HD V3<double> abl3(V3<Dual> a) { return {a.x.d, a.y.d, a.z.d}; }
template <typename S> GC_HD V3<S> konst_s(V3<double> a) { return {S(a.x), S(a.y), S(a.z)}; }
GC_HD V3<double> wert3s(V3<double> a) { return a; }
GC_HD V3<double> wert3s(V3<Dual> a) { return wert3(a); }

// Membrane element in dual numbers; q = 0, 1, 2 seeds E_k, E_s, G.
GC_HD Membran<Dual> dual_el(const Membran<double>& m, int q) {
  Membran<Dual> e;
  e.flaeche = Dual(m.flaeche);
  for (int k = 0; k < 4; ++k) e.r[k] = Dual(m.r[k]);
  e.kette[0] = Dual(m.kette[0]); e.kette[1] = Dual(m.kette[1]);
  e.schuss[0] = Dual(m.schuss[0]); e.schuss[1] = Dual(m.schuss[1]);
  e.ek = Dual(m.ek, q == 0 ? 1.0 : 0.0);
  e.es = Dual(m.es, q == 1 ? 1.0 : 0.0);
  e.g = Dual(m.g, q == 2 ? 1.0 : 0.0);
  return e;
}

// One point's friction, velocity and position update, path entry and push-out
// (k_update, implizit.rs:945-1013) as a function of the state before the
// step, the solution dv, the normal force fn and the friction coefficient mu.
// The contact normal is the contact query at x0 (k_fest, implizit.rs:872-883);
// a path entry is the point where the sampled distance along the path is 0,
// so its parameter moves by -d(phi) / (grad phi . path).
template <typename T, typename S>
__device__ __forceinline__ void adj_punkt(const SysDev<T>& s, int i, uint8_t fl, V3<S> x0, V3<S> v0, V3<S> dv, S fn, S mu,
                                          V3<S>* xo, V3<S>* vo, int* eintritte) {
  V3<S> v = v0 + dv;
  if (fl == 1) {
    V3<S> n = {S(0.0), S(0.0), S(0.0)};
    V3<double> vk = zero3<double>();
    for (int b = 0; b < s.n_koerper; ++b) {
      V3<S> nor;
      S tiefe;
      V3<double> vkk;
      if (kontakt(s.koerper[b], x0, S(s.koll[i] + s.kontakt_rand), &nor, &tiefe, &vkk)) {
        n = nor;
        vk = vkk;
      }
    }
    const V3<S> rel = v - konst_s<S>(vk);
    const V3<S> t = rel - n * dot(rel, n);
    const S lt = norm(t);
    const S grenze = mu * fn * S(s.h) * S(s.w[i]);
    if (wert(lt) <= wert(grenze)) {
      v = v - t;
    } else if (wert(lt) > 0.0) {
      v = v - t * (grenze / lt);
    }
  }
  V3<S> vv = v;
  const V3<S> alt = x0;
  V3<S> x = alt + v * S(s.h);
  if (s.durchstoss_pruefen)
    for (int b = 0; b < s.n_koerper; ++b) {
      const Koerper& k = s.koerper[b];
      V3<double> p, nk, vk;
      double th = 0.0;
      if (durchstoss(k, wert3s(alt), wert3s(x), &p, &nk, &vk, &th)) {
        ++*eintritte;
        const V3<S> seg = x - alt;
        const V3<S> q = alt + seg * S(th);
        S qq[3] = {q.x - S(k.punkt[0]), q.y - S(k.punkt[1]), q.z - S(k.punkt[2])};
        S phi, gq[3];
        sdf_sample(k, qq, &phi, gq);
        const double nenner = wert(dot(V3<S>{gq[0], gq[1], gq[2]}, seg));
        const S tt = nenner != 0.0 ? S(th) - (phi - S(wert(phi))) / S(nenner) : S(th);
        const V3<S> ps = alt + seg * tt;
        S qp[3] = {ps.x - S(k.punkt[0]), ps.y - S(k.punkt[1]), ps.z - S(k.punkt[2])};
        S dd, g[3];
        sdf_sample(k, qp, &dd, g);
        const V3<S> gv = {g[0], g[1], g[2]};
        const V3<S> nks = gv * (S(1.0) / norm(gv));
        x = ps + nks * S(s.koll[i]);
        const V3<S> rel = vv - konst_s<S>(vk);
        const S vn = dot(rel, nks);
        if (wert(vn) < 0.0) {
          V3<S> t = rel - nks * vn;
          const S lt = norm(t);
          const S grenze = mu * -vn;
          t = wert(lt) <= wert(grenze) ? V3<S>{S(0.0), S(0.0), S(0.0)} : t * (S(1.0) - grenze / lt);
          vv = konst_s<S>(vk) + t;
        }
      }
    }
  for (int b = 0; b < s.n_koerper; ++b) {
    V3<S> nor;
    S tiefe;
    V3<double> vk;
    if (kontakt(s.koerper[b], x, S(s.koll[i]), &nor, &tiefe, &vk)) {
      x = x + nor * tiefe;
      const S vn = dot(vv - konst_s<S>(vk), nor);
      if (wert(vn) < 0.0) vv = vv - nor * vn;
    }
  }
  *xo = x;
  *vo = vv;
}

// Reverse of the point update: from (dL/dx', dL/dv') the parts of dL/dx,
// dL/dv, dL/d(dv) (the adjoint right-hand side, also in the bulk type),
// the normal target -dL/dfn / h and dL/dmu. Also checks that the recomputed
// update is the step's, bit for bit.
template <typename T>
__global__ void k_adj_update(const SysDev<T>* __restrict__ sp, AdjDev a, V3<T>* b_adj) {
  const SysDev<T>& s = *sp;
  unsigned long long abw = 0, ein = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    const V3<double> gxn = a.gx[i], gvn = a.gv[i];
    if (s.w[i] <= 0.0) {  // held points keep x and v (k_update)
      a.gx_p[i] = gxn;
      a.gv_p[i] = gvn;
      a.gdv[i] = zero3<double>();
      b_adj[i] = zero3<T>();
      a.ziel[i] = 0.0;
      continue;
    }
    const uint8_t fl = s.fflag[i];
    const V3<double> x0 = a.x0[i], v0 = a.v0[i], dv = dv_loesung(s, (uint32_t)i);
    const double fn = s.fn_kraft[i], mu = s.reib[i];
    double g[11];
    int eintritte = 0;
    for (int k = 0; k < 11; ++k) {
      V3<Dual> xo, vo;
      adj_punkt<T, Dual>(s, i, fl, saat(x0, k, 0), saat(v0, k, 3), saat(dv, k, 6), Dual(fn, k == 9 ? 1.0 : 0.0),
                         Dual(mu, k == 10 ? 1.0 : 0.0), &xo, &vo, &eintritte);
      g[k] = dot(gxn, abl3(xo)) + dot(gvn, abl3(vo));
      if (k == 0) {
        const V3<double> xw = wert3(xo), vw = wert3(vo), xs = s.x[i], vs = s.v[i];
        if (xw.x != xs.x || xw.y != xs.y || xw.z != xs.z || vw.x != vs.x || vw.y != vs.y || vw.z != vs.z) ++abw;
        ein += (unsigned long long)eintritte;
      }
    }
    a.gx_p[i] = V3<double>{g[0], g[1], g[2]};
    a.gv_p[i] = V3<double>{g[3], g[4], g[5]};
    const V3<double> gdv = {g[6], g[7], g[8]};
    a.gdv[i] = gdv;
    b_adj[i] = cast3<T>(gdv);
    a.ziel[i] = fl == 1 ? -g[9] / s.h : 0.0;
    a.d_reibung[i] += g[10];
  }
  if (abw) atomicAdd(&a.zaehler[0], abw);
  if (ein) atomicAdd(&a.zaehler[1], ein);
}

// nu_i = n_i.(g_dv - A mu)_i at held points.
template <typename T>
__global__ void k_adj_nu(const SysDev<T>* __restrict__ sp, AdjDev a, const V3<T>* __restrict__ mu) {
  const SysDev<T>& s = *sp;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    if (s.fflag[i] != 1 && s.fflag[i] != 3) {  // held in the last solve
      a.nu[i] = 0.0;
      continue;
    }
    const V3<T> am = mal_a_zeile(s, (uint32_t)i, [&](uint32_t j) { return mu[j]; });
    a.nu[i] = dot(a.gdv[i] - cast3<double>(am), s.fnorm64[i]);
  }
}

// psi of a distance constraint (k_bau_kanten and the edge rows of mal_h):
// 6 coordinates and k.
template <typename T>
__global__ void k_adj_kanten(const SysDev<T>* __restrict__ sp, AdjDev a, const V3<T>* __restrict__ mu) {
  const SysDev<T>& s = *sp;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < s.nk; c += gridDim.x * blockDim.x) {
    const uint32_t i = s.k_ij[2 * c], j = s.k_ij[2 * c + 1];
    const double k0 = s.k_k[c], l0in = s.k_ruhe[c];
    const bool naht = s.k_naht[c] != 0;
    const V3<double> mi = cast3<double>(mu[i]), mj = cast3<double>(mu[j]);
    const V3<double> ud = (a.v0[i] + dv_loesung(s, (uint32_t)i)) - (a.v0[j] + dv_loesung(s, (uint32_t)j));
    double g[7];
    for (int r = 0; r < 7; ++r) {
      const V3<Dual> xi = saat(a.x0[i], r, 0), xj = saat(a.x0[j], r, 3);
      const Dual k(k0, r == 6 ? 1.0 : 0.0);
      const V3<Dual> d = xj - xi;
      V3<Dual> nr, fi;
      Dual quer;
      if (!naht) {
        const Dual l = gmax(norm(d), Dual(1e-12));
        nr = d * (Dual(1.0) / l);
        fi = nr * (k * (l - Dual(l0in)));
        quer = gmax(Dual(1.0) - Dual(l0in) / l, Dual(0.0));
      } else {
        const Dual l = norm(d);
        const Dual l0 = isfinite(l0in) ? Dual(l0in) : l;
        const bool lang = wert(l0) > 0.0 && wert(l) > 1e-12;
        fi = lang ? d * (k * (l - l0) / l) : d * k;
        nr = lang ? d * (Dual(1.0) / l) : konst(zero3<double>());
        quer = lang ? gmax(Dual(1.0) - l0 / l, Dual(0.0)) : Dual(1.0);
      }
      const V3<Dual> u = konst(ud);
      const V3<Dual> bd = (nr * (dot(nr, u) * (Dual(1.0) - quer)) + u * quer) * k;
      const Dual psi = (dot(konst(mi), fi) - dot(konst(mj), fi)) * Dual(s.h) - (dot(konst(mi), bd) - dot(konst(mj), bd)) * Dual(s.h2);
      g[r] = psi.d;
    }
    a.e_kante[2 * c] = V3<double>{g[0], g[1], g[2]};
    a.e_kante[2 * c + 1] = V3<double>{g[3], g[4], g[5]};
    a.d_kante[c] += g[6];
  }
}

// psi of a bending constraint (k_bau_biegung, the force of k_sammeln and the
// bending rows of mal_h): 12 coordinates and k.
template <typename T>
__global__ void k_adj_biegung(const SysDev<T>* __restrict__ sp, AdjDev a, const V3<T>* __restrict__ mu) {
  const SysDev<T>& s = *sp;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < s.nb; c += gridDim.x * blockDim.x) {
    const uint32_t* bp = s.b_p + 4 * c;
    V3<double> xs[4], ms[4], us[4];
    for (int m = 0; m < 4; ++m) {
      xs[m] = a.x0[bp[m]];
      ms[m] = cast3<double>(mu[bp[m]]);
      us[m] = a.v0[bp[m]] + dv_loesung(s, bp[m]);
    }
    const double k0 = s.b_k[c], ruhe = s.b_ruhe[c];
    double g[13];
    bool entartet = false;
    for (int r = 0; r < 13 && !entartet; ++r) {
      V3<Dual> q[4];
      for (int m = 0; m < 4; ++m) q[m] = saat(xs[m], r, 3 * m);
      Dual theta;
      V3<Dual> gr[4];
      if (!dieder(q, &theta, gr)) {  // a degenerate constraint has no stiffness (implizit.rs:780-788)
        entartet = true;
        break;
      }
      const Dual k(k0, r == 12 ? 1.0 : 0.0);
      const Dual kf = k * winkel_fehler(theta, Dual(ruhe));
      Dual gp(0.0);
      for (int l = 0; l < 4; ++l) gp += dot(gr[l], konst(us[l]));
      const Dual kg = k * gp;
      Dual psi(0.0);
      for (int m = 0; m < 4; ++m)
        psi += dot(konst(ms[m]), gr[m] * kf) * Dual(-s.h) - dot(konst(ms[m]), gr[m] * kg) * Dual(s.h2);
      g[r] = psi.d;
    }
    if (entartet)
      for (int r = 0; r < 13; ++r) g[r] = 0.0;
    for (int m = 0; m < 4; ++m) a.e_biegung[4 * c + m] = V3<double>{g[3 * m], g[3 * m + 1], g[3 * m + 2]};
    a.d_biegung[c] += g[12];
  }
}

// psi of a membrane (k_bau_membran, its force and the membrane rows of
// mal_h): 9 coordinates and E_k, E_s, G.
template <typename T>
__global__ void k_adj_membran(const SysDev<T>* __restrict__ sp, AdjDev a, const V3<T>* __restrict__ mu) {
  const SysDev<T>& s = *sp;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < s.nm; c += gridDim.x * blockDim.x) {
    const uint32_t* mp = s.m_p + 3 * c;
    V3<double> xs[3], ms[3], us[3];
    for (int k = 0; k < 3; ++k) {
      xs[k] = a.x0[mp[k]];
      ms[k] = cast3<double>(mu[mp[k]]);
      us[k] = a.v0[mp[k]] + dv_loesung(s, mp[k]);
    }
    const Membran<double> el = s.m_el64[c];
    double g[12];
    for (int r = 0; r < 12; ++r) {
      const Membran<Dual> e = dual_el(el, r - 9);
      const V3<Dual> x0 = saat(xs[0], r, 0), x1 = saat(xs[1], r, 3), x2 = saat(xs[2], r, 6);
      V3<Dual> f1, f2;
      spalten(e, x1 - x0, x2 - x0, &f1, &f2);
      V3<Dual> fk[3], zk[3];
      krafte(e, f1, f2, fk, zk);
      Hessen<Dual> hh;
      hh.f1 = f1;
      hh.f2 = f2;
      hh.w = geo_w_aus(e, f1, f2);
      const V3<Dual> u3[3] = {konst(us[0]), konst(us[1]), konst(us[2])};
      V3<Dual> o[3];
      mal_h_aus_alle(e, hh, u3, o);
      Dual psi(0.0);
      for (int k = 0; k < 3; ++k) psi += dot(konst(ms[k]), fk[k]) * Dual(s.h) - dot(konst(ms[k]), o[k]) * Dual(s.h2);
      g[r] = psi.d;
    }
    for (int k = 0; k < 3; ++k) a.e_membran[3 * c + k] = V3<double>{g[3 * k], g[3 * k + 1], g[3 * k + 2]};
    for (int q = 0; q < 3; ++q) a.d_membran[3 * c + q] += g[9 + q];
  }
}

// Per point: dL/dx and dL/dv of the state before the step, gathered over the
// point's constraints in CSR order, plus the held-point contact term and the
// right-hand side's dependence on v.
template <typename T>
__global__ void k_adj_sammeln(const SysDev<T>* __restrict__ sp, AdjDev a, const V3<T>* __restrict__ mu) {
  const SysDev<T>& s = *sp;
  for (int ii = blockIdx.x * blockDim.x + threadIdx.x; ii < s.n; ii += gridDim.x * blockDim.x) {
    const uint32_t i = (uint32_t)ii;
    V3<double> gx = a.gx_p[i], gv = a.gv_p[i];
    const uint32_t* o = s.off + GC_KAT * i;
    for (uint32_t e = o[0]; e < o[1]; ++e) {
      const uint32_t ent = s.csr[e];
      gx = gx + a.e_kante[2 * (ent >> 2) + (ent & 3u)];
    }
    for (uint32_t e = o[1]; e < o[2]; ++e) {
      const uint32_t ent = s.csr[e];
      gx = gx + a.e_biegung[4 * (ent >> 2) + (ent & 3u)];
    }
    for (uint32_t e = o[2]; e < o[3]; ++e) {
      const uint32_t ent = s.csr[e];
      gx = gx + a.e_membran[3 * (ent >> 2) + (ent & 3u)];
    }
    const V3<double> mui = cast3<double>(mu[i]);
    const V3<T> hm = mal_h_zeile(s, i, [&](uint32_t j) { return mu[j]; });
    gv = gv - mui * (s.h * s.c * s.masse[i]) - cast3<double>(hm) * s.h2;
    if (s.fflag[i] == 1 || s.fflag[i] == 3) {  // held in the last solve
      const double lam = s.h * s.fn_kraft[i], nu = a.nu[i];
      const V3<double> w = mui * lam - dv_loesung(s, (uint32_t)i) * nu;
      double g[6];
      for (int r = 0; r < 6; ++r) {
        const V3<Dual> X = saat(a.x0[i], r, 0), V = saat(a.v0[i], r, 3);
        V3<Dual> n = konst(zero3<double>());
        Dual t(0.0);
        V3<double> vk = zero3<double>();
        for (int b = 0; b < s.n_koerper; ++b) {
          V3<Dual> nor;
          Dual tiefe;
          V3<double> vkk;
          if (kontakt(s.koerper[b], X, Dual(s.koll[i] + s.kontakt_rand), &nor, &tiefe, &vkk)) {
            n = nor;
            t = tiefe;
            vk = vkk;
          }
        }
        const Dual c = (t - Dual(s.kontakt_rand)) / Dual(s.h) + dot(konst(vk), n) - dot(V, n);
        const Dual chi = dot(n, konst(w)) + Dual(nu) * c;
        g[r] = chi.d;
      }
      gx = gx + V3<double>{g[0], g[1], g[2]};
      gv = gv + V3<double>{g[3], g[4], g[5]};
    }
    a.gx[i] = gx;
    a.gv[i] = gv;
  }
}
