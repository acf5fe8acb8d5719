// gc_xpbd.h — the reference's second frame loop, XPBD with small substeps
// (`CpuReferenz`, crates/klon3d-sim/src/cpu.rs:76-246), on the device.
//
// Per substep: the load per point (gravity, edge tension, membrane tension),
// the prediction, one Gauss-Seidel round of constraints (distance, membrane
// strains, seams, bending), body contact with Coulomb friction as a position
// correction, and the velocity update, in the reference's FP64 arithmetic.
//
// Gauss-Seidel order on the device. The reference walks each constraint group
// in its stored order; two constraints that share a point must see each
// other's corrections in that order, two that do not commute exactly. The
// host cuts each group into consecutive runs whose constraints share no point
// (colour-ordered groups give their colours), and orders the membrane
// elements by dependency level: an element's level is one more than the
// level of the last earlier element that touched any of its corners. Within a
// run or a level every constraint is one thread; the runs and levels follow
// one another. Every point therefore receives its corrections, and its load
// terms, in the reference's sequence, and the FP64 results are the
// reference's up to the device's atan2 in the dihedral angle.
#pragma once
#include "gc_kernels.h"

namespace xp {

// kontinuum.rs:446-472: dg/dx per corner k as J[k][row][col].
GC_HD void jacobi_aus(const Membran<double>& m, V3<double> f1, V3<double> f2, double J[3][3][3]) {
  const double* r = m.r;
  auto zeile = [&](double a, double b, double* o) {
    o[0] = a * f1.x + b * f2.x;
    o[1] = a * f1.y + b * f2.y;
    o[2] = a * f1.z + b * f2.z;
  };
  zeile(2.0 * r[0], 0.0, J[1][0]);
  zeile(0.0, 2.0 * r[1], J[1][1]);
  zeile(2.0 * r[1], 2.0 * r[0], J[1][2]);
  zeile(2.0 * r[2], 0.0, J[2][0]);
  zeile(0.0, 2.0 * r[3], J[2][1]);
  zeile(2.0 * r[3], 2.0 * r[2], J[2][2]);
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) J[0][i][k] = -(J[1][i][k] + J[2][i][k]);
}

struct Zug {  // kontinuum.rs:133-140
  double c;
  V3<double> grad[3];
  double alpha;
};

// kontinuum.rs:526-551: warp strain, weft strain and engineering shear as
// three scalar constraints with their gradients and compliances.
GC_HD void zuege(const Membran<double>& m, const V3<double> p[3], Zug out[3]) {
  V3<double> f1, f2;
  spalten(m, p[1] - p[0], p[2] - p[0], &f1, &f2);
  const V3<double> e = dehnung_aus(m, f1, f2);
  double J[3][3][3];
  jacobi_aus(m, f1, f2, J);
  const V3<double> a = a3(m.kette), b = a3(m.schuss), c = c3(m.kette, m.schuss);
  auto grad = [&](V3<double> v, bool halb, V3<double> g[3]) {
    const double vv[3] = {v.x, v.y, v.z};
    for (int k = 0; k < 3; ++k) {
      double gc[3];
      for (int col = 0; col < 3; ++col) {
        double s = 0.0;
        for (int i = 0; i < 3; ++i) s += J[k][i][col] * vv[i];
        gc[col] = halb ? 0.5 * s : s;
      }
      g[k] = {gc[0], gc[1], gc[2]};
    }
  };
  const double fl = m.flaeche;
  auto alpha = [&](double k) { return k > 0.0 && fl > 0.0 ? 1.0 / (k * fl) : INFINITY; };
  out[0].c = e.x;
  grad(a, true, out[0].grad);
  out[0].alpha = alpha(m.ek);
  out[1].c = e.y;
  grad(b, true, out[1].grad);
  out[1].alpha = alpha(m.es);
  out[2].c = 2.0 * e.z;
  grad(c, false, out[2].grad);
  out[2].alpha = alpha(m.g);
}

}  // namespace xp

// The load at the start of the substep (cpu.rs:97-135): gravity, then edge
// tension in constraint order, then membrane tension in element order. The
// terms depend only on positions, so each is computed by its own thread and
// every point gathers its terms in the reference's order (its distance
// constraints ascending, then its elements ascending): the sum per point is
// the reference's, without one launch per run or level.

// Edge tension k max(0, L - L0) per distance constraint (cpu.rs:102-121);
// da[c] = 0 where the reference adds nothing.
__global__ void k_xp_zug_abstand(int na, const uint32_t* __restrict__ ap, const double* __restrict__ ruhe,
                                 const double* __restrict__ alpha, const V3<double>* __restrict__ x,
                                 const V3<double>* __restrict__ delta, V3<double>* __restrict__ fa, uint8_t* __restrict__ da) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < na; c += gridDim.x * blockDim.x) {
    const uint32_t i = ap[2 * c], j = ap[2 * c + 1];
    const V3<double> d = (x[j] - x[i]) + (delta[j] - delta[i]);
    const double l = norm(d);
    const double zug = fmax(l - ruhe[c], 0.0) / alpha[c];
    const bool an = l > 1e-12 && zug > 0.0;
    da[c] = an ? 1 : 0;
    if (an) fa[c] = d * (zug / l);
  }
}

// Membrane tension per element and corner (cpu.rs:123-135).
__global__ void k_xp_zug_membran(int nm, const uint32_t* __restrict__ mp, const Membran<double>* __restrict__ mel,
                                 const V3<double>* __restrict__ x, const V3<double>* __restrict__ delta,
                                 const V3<double>* __restrict__ korr, V3<double>* __restrict__ fm) {
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nm; t += gridDim.x * blockDim.x) {
    const uint32_t* d = mp + 3 * (size_t)t;
    V3<double> p[3];
    for (int k = 0; k < 3; ++k) p[k] = (x[d[k]] + delta[d[k]]) + korr[d[k]];
    V3<double> f1, f2, kr[3], zug[3];
    spalten(mel[t], p[1] - p[0], p[2] - p[0], &f1, &f2);
    krafte(mel[t], f1, f2, kr, zug);
    for (int k = 0; k < 3; ++k) fm[3 * (size_t)t + k] = zug[k];
  }
}

// Per point: gravity, then its distance terms (entry 2c + side: side 0 adds,
// side 1 subtracts), then its membrane terms (entry 3t + corner).
__global__ void k_xp_last(int n, const double* __restrict__ masse, V3<double> g, const uint32_t* __restrict__ a_off,
                          const uint32_t* __restrict__ a_liste, const V3<double>* __restrict__ fa,
                          const uint8_t* __restrict__ da, const uint32_t* __restrict__ m_off,
                          const uint32_t* __restrict__ m_liste, const V3<double>* __restrict__ fm,
                          V3<double>* __restrict__ last) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    V3<double> l = g * masse[i];
    for (uint32_t e = a_off[i]; e < a_off[i + 1]; ++e) {
      const uint32_t c = a_liste[e] >> 1;
      if (!da[c]) continue;
      l = (a_liste[e] & 1u) ? l - fa[c] : l + fa[c];
    }
    for (uint32_t e = m_off[i]; e < m_off[i + 1]; ++e) l = l + fm[m_liste[e]];
    last[i] = l;
  }
}

// Prediction (cpu.rs:137-143).
__global__ void k_xp_vorhersage(int n, const double* __restrict__ w, V3<double> g, double h, double daempf,
                                V3<double>* __restrict__ v, V3<double>* __restrict__ delta, V3<double>* __restrict__ korr) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    korr[i] = zero3<double>();
    if (w[i] > 0.0) {
      v[i] = (v[i] + g * h) * daempf;
      delta[i] = delta[i] + v[i] * h;
    }
  }
}

// cpu.rs:249-268 for one run of distance constraints (alpha per constraint,
// or one alpha for all when alpha_je is null); alpha / h2 as the reference
// divides it.
__global__ void k_xp_abstand(int lo, int hi, const uint32_t* __restrict__ ap, const double* __restrict__ ruhe,
                             const double* __restrict__ alpha_je, double alpha_alle, double h2,
                             const V3<double>* __restrict__ x, const V3<double>* __restrict__ delta,
                             V3<double>* __restrict__ korr, const double* __restrict__ w) {
  for (int c = lo + blockIdx.x * blockDim.x + threadIdx.x; c < hi; c += gridDim.x * blockDim.x) {
    const uint32_t i = ap[2 * c], j = ap[2 * c + 1];
    const double alpha_h2 = (alpha_je ? alpha_je[c] : alpha_alle) / h2;
    const V3<double> d = ((x[i] - x[j]) + (delta[i] - delta[j])) + (korr[i] - korr[j]);
    const double l = norm(d);
    const double wsum = w[i] + w[j] + alpha_h2;
    if (l < 1e-12 || wsum <= 0.0) continue;
    const double dl = -(l - ruhe[c]) / wsum;
    const V3<double> nrm = d * (1.0 / l);
    korr[i] = korr[i] + nrm * (w[i] * dl);
    korr[j] = korr[j] - nrm * (w[j] * dl);
  }
}

// Membrane strains of one dependency level (cpu.rs:160-189).
__global__ void k_xp_membran(int lo, int hi, const uint32_t* __restrict__ reihe, const uint32_t* __restrict__ mp,
                             const Membran<double>* __restrict__ mel, double h2, const V3<double>* __restrict__ x,
                             const V3<double>* __restrict__ delta, V3<double>* __restrict__ korr, const double* __restrict__ w) {
  for (int q = lo + blockIdx.x * blockDim.x + threadIdx.x; q < hi; q += gridDim.x * blockDim.x) {
    const uint32_t t = reihe[q];
    const uint32_t* d = mp + 3 * (size_t)t;
    V3<double> p[3];
    for (int k = 0; k < 3; ++k) p[k] = (x[d[k]] + delta[d[k]]) + korr[d[k]];
    xp::Zug z[3];
    xp::zuege(mel[t], p, z);
    for (int a = 0; a < 3; ++a) {
      if (!isfinite(z[a].alpha)) continue;
      double wsum = z[a].alpha / h2;
      for (int k = 0; k < 3; ++k) wsum += w[d[k]] * dot(z[a].grad[k], z[a].grad[k]);
      if (!(wsum > 0.0)) continue;
      const double dl = -z[a].c / wsum;
      for (int k = 0; k < 3; ++k) korr[d[k]] = korr[d[k]] + z[a].grad[k] * (w[d[k]] * dl);
    }
  }
}

// All membrane levels in one launch of one block: the levels follow one
// another behind a workgroup barrier, which orders the block's global stores
// before the next level's loads. For levels of at most a few hundred elements
// this replaces a launch per level by a barrier per level; larger levels need
// more than one CU.
__global__ void __launch_bounds__(1024) k_xp_membran_stufen(int nstufen, const int* __restrict__ stufen,
                                                            const uint32_t* __restrict__ reihe, const uint32_t* __restrict__ mp,
                                                            const Membran<double>* __restrict__ mel, double h2,
                                                            const V3<double>* __restrict__ x, const V3<double>* __restrict__ delta,
                                                            V3<double>* __restrict__ korr, const double* __restrict__ w) {
  for (int l = 0; l < nstufen; ++l) {
    const int lo = stufen[l], hi = stufen[l + 1];
    for (int q = lo + (int)threadIdx.x; q < hi; q += (int)blockDim.x) {
      const uint32_t t = reihe[q];
      const uint32_t* d = mp + 3 * (size_t)t;
      V3<double> p[3];
      for (int k = 0; k < 3; ++k) p[k] = (x[d[k]] + delta[d[k]]) + korr[d[k]];
      xp::Zug z[3];
      xp::zuege(mel[t], p, z);
      for (int a = 0; a < 3; ++a) {
        if (!isfinite(z[a].alpha)) continue;
        double wsum = z[a].alpha / h2;
        for (int k = 0; k < 3; ++k) wsum += w[d[k]] * dot(z[a].grad[k], z[a].grad[k]);
        if (!(wsum > 0.0)) continue;
        const double dl = -z[a].c / wsum;
        for (int k = 0; k < 3; ++k) korr[d[k]] = korr[d[k]] + z[a].grad[k] * (w[d[k]] * dl);
      }
    }
    __syncthreads();
  }
}

// cpu.rs:310-333 for one run of bending constraints.
__global__ void k_xp_biegung(int lo, int hi, const uint32_t* __restrict__ bp, const double* __restrict__ ruhe,
                             const double* __restrict__ alpha, double h2, const V3<double>* __restrict__ x,
                             const V3<double>* __restrict__ delta, V3<double>* __restrict__ korr, const double* __restrict__ w) {
  for (int c = lo + blockIdx.x * blockDim.x + threadIdx.x; c < hi; c += gridDim.x * blockDim.x) {
    const uint32_t* p = bp + 4 * (size_t)c;
    const uint32_t p0 = p[0];
    V3<double> q[4];
    for (int a = 0; a < 4; ++a) q[a] = ((x[p[a]] - x[p0]) + (delta[p[a]] - delta[p0])) + (korr[p[a]] - korr[p0]);
    double theta;
    V3<double> g[4];
    if (!dieder(q, &theta, g)) continue;
    double nenner = alpha[c] / h2;
    for (int a = 0; a < 4; ++a) nenner += w[p[a]] * dot(g[a], g[a]);
    if (nenner <= 0.0) continue;
    const double dl = -winkel_fehler(theta, ruhe[c]) / nenner;
    for (int a = 0; a < 4; ++a) korr[p[a]] = korr[p[a]] + g[a] * (w[p[a]] * dl);
  }
}

// Contacts of the substep statistics (cpu.rs:63-73), summed per frame.
struct XpStatDev {
  unsigned long long kontakte, haftend;
  double tiefe, tangential;
};

// Body contact with Coulomb friction as a position correction (cpu.rs:203-222),
// bodies in list order.
template <typename T>
__global__ void k_xp_kontakt(const SysDev<T>* __restrict__ sp, double h, double h2, const V3<double>* __restrict__ delta,
                             V3<double>* __restrict__ korr, const V3<double>* __restrict__ last, XpStatDev* __restrict__ st) {
  const SysDev<T>& s = *sp;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    if (s.w[i] <= 0.0) continue;
    for (int b = 0; b < s.n_koerper; ++b) {
      const V3<double> p = (s.x[i] + delta[i]) + korr[i];
      V3<double> n, vk;
      double tiefe;
      if (!kontakt(s.koerper[b], p, s.koll[i], &n, &tiefe, &vk)) continue;
      korr[i] = korr[i] + n * tiefe;
      const V3<double> dx = (s.v[i] * h + korr[i]) - vk * h;
      const V3<double> t = dx - n * dot(dx, n);
      const double lt = norm(t);
      const double normallast = fmax(-dot(last[i], n), 0.0);
      const double grenze = s.reib[i] * fmax(tiefe, normallast * h2 * s.w[i]);
      atomicAdd(&st->kontakte, 1ull);
      atomicAdd(&st->tiefe, tiefe);
      atomicAdd(&st->tangential, lt);
      if (lt <= grenze) {
        atomicAdd(&st->haftend, 1ull);
        korr[i] = korr[i] - t;
      } else if (lt > 0.0) {
        korr[i] = korr[i] - t * (grenze / lt);
      }
    }
  }
}

// End of the substep: delta += kappa, v += kappa / h (cpu.rs:215-222).
__global__ void k_xp_ende(int n, const double* __restrict__ w, double h, V3<double>* __restrict__ v,
                          V3<double>* __restrict__ delta, const V3<double>* __restrict__ korr) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    if (w[i] > 0.0) {
      delta[i] = delta[i] + korr[i];
      v[i] = v[i] + korr[i] * (1.0 / h);
    }
}

// End of the frame: x += delta (cpu.rs:240-242).
__global__ void k_xp_lage(int n, V3<double>* __restrict__ x, const V3<double>* __restrict__ delta) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] = x[i] + delta[i];
}
