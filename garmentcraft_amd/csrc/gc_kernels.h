// gc_kernels.h — device kernels of the implicit step for CDNA4 (gfx950).
//
// Layout: structure of arrays in HBM, one writer per constraint, one gather per
// point over a CSR list of the constraints that touch it (no atomics on data).
// The CSR list of a point holds its constraints by category — distance
// constraints (edges with k != 0, then seams), bending, membranes — each in
// constraint order, which is the order in which the reference's sequential
// loops add into that point (implizit.rs:185-221, 232-279, 725-827). The FP64
// build therefore reproduces the reference's per-point sums.
#pragma once
#include <type_traits>
#include "gc_math.h"

// CSR entry: constraint index << 2 | slot. Kante: slot 0 = first point (+), 1 = second (-).
#define GC_KAT 4  // categories per point in the offset table: kanten, biege, membran, (end)

// Operator records of one tile entry, in tile order (see SysDev::kachel_kante).
template <typename T>
struct KanteK {
  V3<T> n;
  T q, kk;
};
template <typename T>
struct BiegeK {
  V3<T> g[4];
  T kt;
};
template <typename T>
struct MembranK {
  Membran<T> el;
  Hessen<T> h;
};

// Global-memory view of a pointer the compiler only knows as generic (every
// pointer read from SysDev): loads and stores through it are global_*
// instead of flat_* instructions, so they do not count against the LDS wait
// counter that the stages' LDS reads wait on.
#define GC_GM __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const GC_GM T* gm(const T* p) {
  return (const GC_GM T*)p;
}
template <typename T>
__device__ __forceinline__ GC_GM T* gm(T* p) {
  return (GC_GM T*)p;
}
// One element read or written through a global view. Class types (vectors,
// records) have no copy through a global lvalue, so they go word by word; the
// compiler merges the words into wide loads and stores.
template <typename T>
__device__ __forceinline__ T lg(const GC_GM T* p) {
  if constexpr (std::is_class<T>::value) {
    static_assert(sizeof(T) % 4 == 0, "records are whole words");
    const GC_GM uint32_t* q = (const GC_GM uint32_t*)p;
    uint32_t w[sizeof(T) / 4];
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) w[k] = q[k];
    T out;
    __builtin_memcpy(&out, w, sizeof(T));
    return out;
  } else {
    return *p;
  }
}
template <typename T>
__device__ __forceinline__ T lg(const T* p) {
  return lg(gm(p));
}
template <typename T>
__device__ __forceinline__ void sg(GC_GM T* p, const T& v) {
  if constexpr (std::is_class<T>::value) {
    static_assert(sizeof(T) % 4 == 0, "records are whole words");
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &v, sizeof(T));
    GC_GM uint32_t* q = (GC_GM uint32_t*)p;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); ++k) q[k] = w[k];
  } else {
    *p = v;
  }
}
// m3_mal with the matrix read through a global view
template <typename T>
__device__ __forceinline__ V3<T> m3_mal_g(const GC_GM T* m, V3<T> v) {
  T a[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) a[q] = m[q];
  return m3_mal(a, v);
}

// Data of one entry of a point's spring list (see SysDev::fe_satz).
template <typename T>
struct FederK {
  uint32_t pq, pr;  // the points on sides q and r of the entry's pair
  T om;             // -g_q g_r, negated when the entry's row is not p_q
  T B[9];           // the spring's block
};

// Decisions of a step held for replay (one body): contact of the step,
// released after solve 1 / 2, sticking, path entry and its impact (normal
// velocity removed, tangential one stopped), push-out and its normal velocity
// removal.
enum : uint16_t {
  GC_H_KONTAKT = 1, GC_H_FREI0 = 2, GC_H_FREI1 = 4, GC_H_HAFT = 8, GC_H_EINTRITT = 16, GC_H_EIN_VN = 32,
  GC_H_EIN_HAFT = 64, GC_H_RAUS = 128, GC_H_RAUS_VN = 256
};

template <typename T>
struct SysDev {
  int32_t n, nk, nb, nm;
  double h, h2, c, naht_k, kontakt_rand;
  double g[3];
  int32_t cg_max;
  double cg_tol2, boden;
  int32_t aus_last;
  int32_t durchstoss_pruefen;
  // CSR
  const uint32_t* off;   // 4 per point + 1
  const uint32_t* csr;
  // distance constraints (edges with k != 0 in input order, then seams)
  const uint32_t* k_ij;
  const double* k_k;
  const double* k_ruhe;
  const uint8_t* k_naht;
  V3<T>* k_n;
  T* k_q;
  T* k_kk;
  V3<double>* k_f;
  uint8_t* k_zug;
  // bending
  const uint32_t* b_p;
  const double* b_k;
  const double* b_ruhe;
  T* b_kt;
  V3<T>* b_grad;  // 4 per constraint
  double* b_kf;   // k * wrapped angle error
  // membranes
  const uint32_t* m_p;
  const Membran<T>* m_el;         // operator records (per iteration)
  const Membran<double>* m_el64;  // the same elements for the per-step build
  Hessen<T>* m_h;
  V3<double>* m_f;  // 3 per element: force
  V3<double>* m_z;  // 3 per element: tension-only force
  // points
  const double* masse;
  const double* w;
  const double* koll;
  const double* reib;
  V3<double>* x;
  V3<double>* v;
  T* mdiag;
  V3<double>* f;
  V3<double>* last;
  V3<T>* b;
  T* pinv;  // 9 per point
  // contacts: 0 free, 1 held along a normal, 2 held completely (w <= 0)
  uint8_t* fflag;
  V3<T>* fnorm;
  V3<double>* fnorm64;
  double* vn_soll;
  V3<double>* vk;
  double* fn_kraft;
  // PCG
  V3<T>* dv;
  V3<T>* r;
  V3<T>* z;
  V3<T>* p[2];
  V3<T>* ap;
  V3<double>* dv_vorher;
  // FP32 bulk only (null in FP64): the solve's solution in FP64, and the
  // device address of an FP64 copy of this system (SysDev<double>: the same
  // topology with FP64 operator records, right-hand side and mass diagonal).
  // The FP32 PCG then solves for corrections of dv64 whose residual is taken
  // in FP64 from that copy (iterative refinement), so its result converges to
  // the FP64 system's solution rather than to the FP32-rounded one's.
  V3<double>* dv64;
  const void* schatten;
  // The step's discrete decisions per point (bits GC_H_*): recorded
  // (halten_modus 1) or replayed instead of decided (2), for finite
  // differences on the branch the adjoint differentiates; 0: off.
  uint16_t* halten;
  int32_t halten_modus;
  // normal target of held points for the solve; null: vn_soll - v.n of the
  // step (implizit.rs:375-384); set: the adjoint solve's own targets
  const double* soll_direkt;
  // tiles of the PCG operator: row ranges whose CSR entries (and the points
  // their constraints touch) fit in LDS
  int32_t n_kacheln;
  const int32_t* kachel_zeilen;   // lo, hi per tile
  const uint32_t* kachel_el_off;  // element-list offsets (n_kacheln + 1)
  const uint32_t* kachel_kat;     // per tile: first bending entry, first membrane entry (edges come first)
  const uint32_t* kachel_el;      // category << 30 | constraint
  const uint16_t* kachel_lv;      // 4 tile-local point indices per listed element
  const uint16_t* kachel_pos;     // 4 LDS slots per listed element, 0xFFFF = point not in the tile
  const uint32_t* kachel_v_off;   // point-list offsets (n_kacheln + 1)
  const uint32_t* kachel_v;       // points touched by the tile's constraints, global index
  uint32_t kachel_slots_max;      // LDS slot area (entries) of the largest tile
  uint32_t kachel_v_max;          // LDS staging area (points) of the largest tile
  uint32_t punkte_lds;            // 1: a one-tile block keeps its point list, flags and normals in LDS
  // the records of every tile entry at its entry index, copied from the
  // per-constraint arrays after each build (k_kachel_saetze), so stage B
  // loads them directly instead of through the constraint index; an entry
  // uses the array of its category only
  KanteK<T>* kachel_kante;
  BiegeK<T>* kachel_biege;
  MembranK<T>* kachel_membran;

  // springs of cloth self-collision, rebuilt every step (counts on the
  // device; empty lists when it is off): per point its spring pairs
  // (spring << 5 | q * 5 + r, in spring then pair order) and its force slots
  // (spring << 3 | q)
  int32_t selbst;
  const uint32_t* fe_off;
  const uint32_t* fe_ent;
  const uint32_t* fk_off;
  const uint32_t* fk_ent;
  const uint32_t* fe_punkte;    // 5 per spring
  const double* fe_gew;         // 5 per spring
  const V3<double>* fe_kraft;
  const T* fe_block;            // 9 per spring, bulk type
  V3<T>* fe_t;                  // per entry of fe_ent: its term K (p_q - p_r) (-g_q g_r)
  FederK<T>* fe_satz;           // per entry: its points, signed weight product and block (k_feder_saetze)
  uint32_t fe_lds_kap;          // spring terms of a tile kept in LDS (after the staging); the rest in fe_t

  // bodies
  int32_t n_koerper;
  Koerper koerper[GC_MAX_KOERPER];
};

// Shared device state of the grid-wide reductions and the step's counters.
struct GridState {
  unsigned int zaehler;      // epoch of the last grid barrier, carried from launch to launch within a
                             // refined FP32 solve (k_verf, k_pcg rounds, k_pcg64: same grid, so the
                             // barrier counters stay monotonic); 0 whenever the counters are cleared
  unsigned int fehler;       // 1: spin limit hit; 2: an all-reduce's partials were not all of its barrier
  unsigned int frei[2];      // a point was released in pass k
  unsigned int n_fest;       // points held along a normal
  // FP32 bulk with the FP64 copy (k_verf rounds): the solve has converged or
  // ended; it is handed over to FP64 iterations (k_pcg64)
  unsigned int fertig;
  unsigned int uebergabe;
  unsigned int pad0;
  // statistics of the last step (device side)
  long long cg_iterationen, loesungen, nicht_konvergiert, kontakte, haftend, freigegeben;
  double rest_max;
  unsigned int spins_max;    // longest barrier wait (polls) since the last read
  unsigned int selbst_ueberlauf;  // a self-collision list exceeded its capacity
  long long naehe_pd, naehe_kk, selbst_stoesse, durchdringungen, gebiete, durchstoesse;
  long long halten_abweichungen;  // replayed decisions that the state no longer allowed (taken as decided)
  long long rundungsgrenze;  // FP32 solves stopped at their rounding floor
  long long neustarts;       // FP32 true-residual restarts
  long long koerperpunkte;   // cloth triangles lifted over body points
  unsigned int kp_pass[2];   // lifts of the last two passes of body points against cloth triangles (pass d at d & 1)
  // result of the last solve
  int iterationen;
  int nicht_konv;
  double rest;
  double delta, ziel, bb;
  double delta_best;         // k_verf: FP64 residual at the FP64 solution dv64
#ifdef GC_PHASEN
  // build flag GC_PHASEN: block 0's wall-clock ticks per phase of a PCG
  // iteration (stage A, stage B, stage C, all-reduce p.Ap, update,
  // all-reduce r.z) and the number of iterations they cover
  unsigned long long phasen[6];
  unsigned long long phasen_it;
  // block 0 in the all-reduces: ticks from its own partial to the summed
  // result (its wait for the other blocks included), and the count
  unsigned long long warte0[2];
  // per block (rank r at 6 r): ticks of stages A, B and C, of the update, and
  // of the waits in the two all-reduces (null: not recorded)
  unsigned long long* pro_block;
#endif
};

// One block's partial of a grid all-reduce: value k as two 64-bit words
// (2k: low half, 2k + 1: high half), each tagged with the barrier's epoch in
// its upper 32 bits. One 64-byte line per block.
#define GC_MAX_BLOECKE 1024
struct alignas(64) Slot {
  unsigned long long wort[8];
};

// A system's slot array (GC_SLOTS entries): one slot per block for each
// epoch parity.
#define GC_SLOTS (2 * GC_MAX_BLOECKE)
// Clears the partials (one block; a fill kernel's work, no memset): an
// all-reduce reads a slot as this barrier's once it carries the barrier's
// epoch, and the epoch starts anew at 1 after each clear.
__device__ __forceinline__ void barriere_leeren(Slot* slots) {
  for (unsigned k = threadIdx.x; k < 16 * GC_SLOTS; k += blockDim.x) ((unsigned*)slots)[k] = 0u;
}

// ---------------------------------------------------------------------------
// Grid-wide all-reduce. Each block writes its partial into its slot, then
// polls the partials of all nb blocks and sums them itself, in one fixed
// order, so every block holds the identical value (deterministic, equal to
// the bit across blocks and across a batch's systems of the same launch
// shape). Slots are double-buffered by epoch parity: a block writes the
// partial of barrier e + 2 into the slot of barrier e only after barrier
// e + 1, which needs every block's partial of e + 1, so no block still reads
// the slot for barrier e then.
//
// Block sums are deterministic: an xor butterfly inside each wave (every lane
// ends with the same value, IEEE addition being commutative), then the wave
// sums in wave order. rang, nb: this block's rank among the nb blocks that
// share gs and slots (the whole grid for one system, a group of blocks in a
// batched launch). Wave 0 alone arrives, waits and hands the result to the
// block through LDS (two workgroup barriers per grid barrier).
//
// A 64-bit access is single-copy atomic, so a reader that finds this
// barrier's epoch in every word of a slot holds that block's partial for
// this barrier, with no counter or flag to order against it: a block's
// arrival IS its partial's store. The critical path is the last partial's
// store and the polls that see it. (Where one leader block sums and
// publishes the result for the others to poll, the result's store and a
// second poll follow: one more trip through the level that is coherent
// across the XCDs.) A slot holding a later epoch than the barrier's is stale
// (left from another launch): the all-reduce then fails and the block says
// why in gs. Slots are cleared before each launch that starts the epoch
// anew.
//
// Release: each block writes back its stores (the L2 write-back to the
// coherent level) before its partial. There is no acquire (no L2
// invalidation): everything a block reads that another block wrote in this
// launch (partials, dv / z / p of points it does not own) is read with
// agent-scope loads, which the program order after the observed partials
// puts after every writer's write-back.
// VOLL = false: the caller's stores before this all-reduce are not read by
// other blocks until after the next full one, so the block's partial alone
// is published; the L2 write-back of the release is left to that next
// all-reduce.
__device__ __forceinline__ unsigned long long barr_wort(double x, int haelfte, unsigned long long marke) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  return marke | (haelfte ? (b >> 32) : (b & 0xffffffffull));
}

// A workgroup barrier that orders LDS only. On gfx950 (no threadgroup split)
// a __syncthreads emits the same s_barrier with no wait on global memory
// either; this form states that nothing global is ordered here, so loads
// issued before it stay in flight until their first use.
__device__ __forceinline__ void lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <int NV, bool VOLL = true>
__device__ __forceinline__ bool grid_allreduce(double (&v)[NV], double* sh, Slot* slots, GridState* gs, unsigned& epoch,
                                               unsigned rang, unsigned nb) {
  static_assert(NV <= 4, "a slot holds four values");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v[k] += __shfl_xor(v[k], o, 64);
  if (lane == 0)
    for (int k = 0; k < NV; ++k) sh[wave * NV + k] = v[k];
  // every wave's stores and wave sums are in; sh's previous readers (wave 0
  // below, the broadcast readers at the end) finished before this barrier.
  // VOLL = false publishes no global stores here, so its barriers order LDS
  // only; with no release fence either, wave 0 stores its partial without
  // waiting for loads the caller issued before it.
  if constexpr (VOLL) __syncthreads();
  else lds_sync();
  ++epoch;
  Slot* buf = slots + (epoch & 1u) * GC_MAX_BLOECKE;
  double* aus = sh + 16 * NV;
  __shared__ int s_fehler;
  if (wave == 0) {
    int fehler = 0;
    const unsigned long long marke = (unsigned long long)epoch << 32;
    // the block's partial: word 2k + h holds half h of value k (the waves'
    // sums in wave order); lane l stores word l mod 2 NV, so that no branch
    // splits the wave here (a release fence inside a divergent branch is
    // duplicated into both sides, and the L2 write-back runs twice)
    const int kw = lane % (2 * NV);
    double a = 0.0;
    for (int w = 0; w < nw; ++w) a += sh[w * NV + (kw >> 1)];
    if constexpr (VOLL) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
#ifdef GC_PHASEN
    const unsigned long long t_an = wall_clock64();
#endif
    __hip_atomic_store(&buf[rang].wort[kw], barr_wort(a, kw & 1, marke), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the nb partials, summed in the fixed order (lane l takes l, l + 64,
    // ..., four slots per round, then the wave's butterfly); each lane waits
    // for its own slots
    double acc[NV];
    for (int k = 0; k < NV; ++k) acc[k] = 0.0;
    unsigned spins = 0;
    bool spaet = false;  // a slot of a later barrier: stale
    for (unsigned b0 = lane; b0 < nb && !fehler; b0 += 256) {
      unsigned long long x[4][2 * NV];
      for (;;) {
        bool da = true;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const bool gibt = b0 + 64u * u < nb;
#pragma unroll
          for (int k = 0; k < 2 * NV; ++k) {
            x[u][k] = gibt ? __hip_atomic_load(&buf[b0 + 64u * u].wort[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                           : marke;
            const unsigned e = (unsigned)(x[u][k] >> 32);
            da = da && e == epoch;
            spaet = spaet || (e > epoch && e != 0u);
          }
        }
        if (da || spaet) break;
        if (++spins > (1u << 22)) { fehler = 1; break; }
        if ((spins & 63u) == 0u && __hip_atomic_load(&gs->fehler, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
          fehler = 1;
          break;
        }
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (b0 + 64u * u < nb)
#pragma unroll
          for (int k = 0; k < NV; ++k)
            acc[k] += __longlong_as_double((long long)((x[u][2 * k + 1] << 32) | (x[u][2 * k] & 0xffffffffull)));
    }
    if (__ballot(fehler == 1) != 0ull) fehler = 1;
    else if (__ballot(spaet) != 0ull) fehler = 2;
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) acc[k] += __shfl_xor(acc[k], o, 64);
#ifdef GC_PHASEN
    if (lane == 0 && rang == 0) {
      gs->warte0[0] += wall_clock64() - t_an;
      gs->warte0[1] += 1ull;
    }
#endif
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) spins = max(spins, (unsigned)__shfl_xor((int)spins, o, 64));
    if (lane == 0) {
      for (int k = 0; k < NV; ++k) aus[k] = acc[k];
      if (fehler)
        __hip_atomic_store(&gs->fehler, (unsigned)fehler, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_fetch_max(&gs->spins_max, spins, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      s_fehler = fehler;
    }
  }
  if constexpr (VOLL) __syncthreads();
  else lds_sync();
  if (s_fehler) return false;
  for (int k = 0; k < NV; ++k) v[k] = aus[k];
  return true;
}

// ---------------------------------------------------------------------------
// Operator rows
// ---------------------------------------------------------------------------
// H*p at row i (implizit.rs:185-221), p given by a functor j -> V3<T>.
template <typename T, typename PF>
__device__ __forceinline__ V3<T> mal_h_zeile(const SysDev<T>& s, uint32_t i, PF p) {
  V3<T> out = zero3<T>();
  const uint32_t* o = s.off + GC_KAT * i;
  const uint32_t e0 = o[0], e1 = o[1], e2 = o[2], e3 = o[3];
  for (uint32_t e = e0; e < e1; ++e) {
    const uint32_t ent = s.csr[e], c = ent >> 2;
    const V3<T> d = p(s.k_ij[2 * c]) - p(s.k_ij[2 * c + 1]);
    const V3<T> n = s.k_n[c];
    const T q = s.k_q[c];
    const V3<T> bd = (n * (dot(n, d) * (T(1) - q)) + d * q) * s.k_kk[c];
    out = (ent & 3u) == 0u ? out + bd : out - bd;
  }
  for (uint32_t e = e1; e < e2; ++e) {
    const uint32_t ent = s.csr[e], c = ent >> 2, m = ent & 3u;
    const uint32_t* bp = s.b_p + 4 * c;
    const V3<T>* gr = s.b_grad + 4 * c;
    T gp = T(0);
#pragma unroll
    for (int l = 0; l < 4; ++l) gp += dot(gr[l], p(bp[l]));
    out = out + gr[m] * (s.b_kt[c] * gp);
  }
  for (uint32_t e = e2; e < e3; ++e) {
    const uint32_t ent = s.csr[e], c = ent >> 2, k = ent & 3u;
    const uint32_t* mp = s.m_p + 3 * c;
    V3<T> d3[3] = {p(mp[0]), p(mp[1]), p(mp[2])};
    out = out + mal_h_aus_k(s.m_el[c], s.m_h[c], d3, (int)k);
  }
  return feder_zeile(s, i, out, p);
}

// Springs of row i, added to acc in the order of System::mal_h
// (implizit.rs:214-220): per spring its pairs (first side q, second side r)
// with the block K times (p_q - p_r) times -g_q g_r.
template <typename T, typename PF>
__device__ __forceinline__ V3<T> feder_zeile(const SysDev<T>& s, uint32_t i, V3<T> acc, PF p) {
  if (!s.selbst) return acc;
  for (uint32_t e = s.fe_off[i]; e < s.fe_off[i + 1]; ++e) {
    const uint32_t code = s.fe_ent[e], f = code >> 5, pr = code & 31u, q = pr / 5u, r = pr % 5u;
    const uint32_t pq = s.fe_punkte[5 * f + q], prr = s.fe_punkte[5 * f + r];
    const T om = T(-s.fe_gew[5 * f + q] * s.fe_gew[5 * f + r]);
    const T* B = s.fe_block + 9 * f;
    const V3<T> d = p(pq) - p(prr);
    const V3<T> t = V3<T>{B[0] * d.x + B[1] * d.y + B[2] * d.z, B[3] * d.x + B[4] * d.y + B[5] * d.z,
                          B[6] * d.x + B[7] * d.y + B[8] * d.z} * om;
    acc = pq == i ? acc + t : acc - t;
  }
  return acc;
}

// The spring terms of rows lo..hi-1 (the entries fe_off[lo]..fe_off[hi]),
// computed by the whole block, one entry per thread: each entry gathers two
// directions and a block of its own, so a row with many springs no longer
// walks them in sequence while the grid waits on it. Same expression as
// feder_zeile, with the row's sign in the weight product (c * -om is exactly
// -(c * om)). Entry e - e0 < fe_lds_kap goes to lds, the rest to fe_t; the
// block's __syncthreads() publishes them to feder_summe.
template <typename T, typename PF>
__device__ __forceinline__ void feder_terme(const SysDev<T>& s, int lo, int hi, PF p, V3<T>* lds) {
  const uint32_t e0 = gm(s.fe_off)[lo], e1 = gm(s.fe_off)[hi];
  for (uint32_t e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const FederK<T> k = lg(gm(s.fe_satz) + (e));
    const T* B = k.B;
    const V3<T> d = p(k.pq) - p(k.pr);
    const V3<T> t = V3<T>{B[0] * d.x + B[1] * d.y + B[2] * d.z, B[3] * d.x + B[4] * d.y + B[5] * d.z,
                          B[6] * d.x + B[7] * d.y + B[8] * d.z} * k.om;
    if (e - e0 < s.fe_lds_kap) lds[e - e0] = t;
    else sg(gm(s.fe_t) + (e), t);
  }
}

// The data of every spring list entry for feder_terme, after each rebuild of
// the lists and blocks: its pair's two points, the block and -g_q g_r,
// negated when the list the entry sits in is not p_q's (the row adds the term
// then, feder_zeile subtracts it), so a row just adds its terms. The row of
// entry e is found by bisection over the list offsets.
template <typename T>
__global__ void k_feder_saetze(const SysDev<T>* __restrict__ sp) {
  const SysDev<T>& s = *sp;
  const uint32_t ne = gm(s.fe_off)[s.n];
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
    const uint32_t code = s.fe_ent[e], f = code >> 5, pr = code & 31u, q = pr / 5u, r = pr % 5u;
    // the row: the last i with fe_off[i] <= e (empty rows share their offset)
    uint32_t u = 0, o = (uint32_t)s.n;
    while (o - u > 1) {
      const uint32_t m = (u + o) >> 1;
      if (s.fe_off[m] <= e) u = m;
      else o = m;
    }
    FederK<T> k;
    k.pq = s.fe_punkte[5 * f + q];
    k.pr = s.fe_punkte[5 * f + r];
    const T om = T(-s.fe_gew[5 * f + q] * s.fe_gew[5 * f + r]);
    k.om = k.pq == u ? om : -om;
    for (int a = 0; a < 9; ++a) k.B[a] = s.fe_block[9 * (size_t)f + a];
    sg(gm(s.fe_satz) + (e), k);
  }
}

// Row i's springs from the terms of feder_terme (in LDS from entry e0 of the
// tile on, up to fe_lds_kap of them), added in the order of feder_zeile; eight
// (FP32) or four terms are loaded before their sums.
template <typename T>
__device__ __forceinline__ V3<T> feder_summe_ab(const SysDev<T>& s, uint32_t e, uint32_t b, V3<T> acc, const V3<T>* lds, uint32_t e0) {
  auto term = [&](uint32_t q) { return q - e0 < s.fe_lds_kap ? lds[q - e0] : lg(gm(s.fe_t) + (q)); };
  // the order of the additions is the entries' order either way; the FP64
  // kernel has no registers to spare for eight
  if constexpr (sizeof(T) == 4) {
    for (; e + 8 <= b; e += 8) {
      V3<T> t[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) t[k] = term(e + k);
#pragma unroll
      for (int k = 0; k < 8; ++k) acc = acc + t[k];
    }
  }
  for (; e + 4 <= b; e += 4) {
    V3<T> t[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) t[k] = term(e + k);
#pragma unroll
    for (int k = 0; k < 4; ++k) acc = acc + t[k];
  }
  for (; e < b; ++e) acc = acc + term(e);
  return acc;
}
template <typename T>
__device__ __forceinline__ V3<T> feder_summe(const SysDev<T>& s, uint32_t i, V3<T> acc, const V3<T>* lds, uint32_t e0) {
  return feder_summe_ab(s, gm(s.fe_off)[i], gm(s.fe_off)[i + 1], acc, lds, e0);
}

// A*p at row i (implizit.rs:224-229).
template <typename T, typename PF>
__device__ __forceinline__ V3<T> mal_a_zeile(const SysDev<T>& s, uint32_t i, PF p) {
  V3<T> hp = mal_h_zeile(s, i, p);
  return hp * T(s.h2) + p(i) * s.mdiag[i];
}

// implizit.rs:283-288 with the held-point rule of CpuLoeser::loese.
template <typename T>
__device__ __forceinline__ V3<T> filtern(const SysDev<T>& s, uint32_t i, V3<T> v) {
  const uint8_t fl = s.fflag[i];
  if (fl == 0) return v;
  if (fl == 2) return zero3<T>();
  const V3<T> n = s.fnorm[i];
  return v - n * dot(v, n);
}

// ---------------------------------------------------------------------------
// Build kernels (implizit.rs:725-844; device form of gpu.rs aufbau_k/aufbau_b)
// ---------------------------------------------------------------------------
template <typename T>
__global__ void k_bau_kanten(const SysDev<T>* __restrict__ sp) {
  const SysDev<T>& s = *sp;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < s.nk; c += gridDim.x * blockDim.x) {
    const uint32_t i = s.k_ij[2 * c], j = s.k_ij[2 * c + 1];
    const double k = s.k_k[c];
    double l0 = s.k_ruhe[c];
    const V3<double> d = s.x[j] - s.x[i];
    V3<double> nr, fi;
    double quer;
    uint8_t zug = 0;
    if (!s.k_naht[c]) {  // implizit.rs:742-752
      const double l = gmax(norm(d), 1e-12);
      nr = d * (1.0 / l);
      fi = nr * (k * (l - l0));
      zug = l > l0 ? 1 : 0;
      quer = gmax(1.0 - l0 / l, 0.0);
    } else {  // implizit.rs:763-771
      const double l = norm(d);
      l0 = isfinite(l0) ? l0 : l;
      const bool lang = l0 > 0.0 && l > 1e-12;
      fi = lang ? d * (k * (l - l0) / l) : d * k;
      nr = lang ? d * (1.0 / l) : zero3<double>();
      quer = lang ? gmax(1.0 - l0 / l, 0.0) : 1.0;
    }
    s.k_n[c] = cast3<T>(nr);
    s.k_q[c] = T(quer);
    s.k_kk[c] = T(k);
    s.k_f[c] = fi;
    s.k_zug[c] = zug;
  }
}

// Bending geometry in FP64 in both modes (the angle and its gradients are a
// per-step cost); only the gradients the operator streams are stored as T.
template <typename T>
__global__ void k_bau_biegung(const SysDev<T>* __restrict__ sp) {
  const SysDev<T>& s = *sp;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < s.nb; c += gridDim.x * blockDim.x) {
    const uint32_t* bp = s.b_p + 4 * c;
    V3<double> q[4];
    for (int m = 0; m < 4; ++m) q[m] = s.x[bp[m]];
    double theta = 0.0;
    V3<double> g[4];
    if (!dieder(q, &theta, g)) {  // implizit.rs:780-788: degenerate keeps its slot with no stiffness
      theta = 0.0;
      for (int m = 0; m < 4; ++m) g[m] = zero3<double>();
    }
    const double k = s.b_k[c];
    const double fehler = winkel_fehler(theta, s.b_ruhe[c]);
    for (int m = 0; m < 4; ++m) s.b_grad[4 * c + m] = cast3<T>(g[m]);
    s.b_kf[c] = k * fehler;
    s.b_kt[c] = T(k);
  }
}

// Membrane strain, forces and tangent in FP64 in both modes: the Green strain
// 0.5*(F^T F - 1) cancels at small stretch, which in FP32 costs ~1e-4 relative
// in the force. Only the tangent the operator streams is stored as T.
template <typename T>
__global__ void k_bau_membran(const SysDev<T>* __restrict__ sp) {
  const SysDev<T>& s = *sp;
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < s.nm; c += gridDim.x * blockDim.x) {
    const uint32_t* mp = s.m_p + 3 * c;
    const V3<double> x0 = s.x[mp[0]];
    const V3<double> e1 = s.x[mp[1]] - x0, e2 = s.x[mp[2]] - x0;
    const Membran<double> el = s.m_el64[c];
    V3<double> f1, f2;
    spalten(el, e1, e2, &f1, &f2);
    V3<double> fk[3], zk[3];
    krafte(el, f1, f2, fk, zk);
    for (int k = 0; k < 3; ++k) {
      s.m_f[3 * c + k] = fk[k];
      s.m_z[3 * c + k] = zk[k];
    }
    Hessen<T> h;
    h.f1 = cast3<T>(f1);
    h.f2 = cast3<T>(f2);
    h.w = cast3<T>(geo_w_aus(el, f1, f2));
    s.m_h[c] = h;
  }
}

// Contacts of the substep (implizit.rs:872-883) and the held-point flags.
template <typename T>
__global__ void k_fest(const SysDev<T>* __restrict__ sp, GridState* gs) {
  const SysDev<T>& s = *sp;
  unsigned lokal = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    uint8_t fl = 0;
    V3<double> nn = zero3<double>(), vkk = zero3<double>();
    double soll = 0.0;
    if (s.w[i] <= 0.0) {
      fl = 2;
    } else if (s.halten_modus == 2) {  // the recorded contact (one body)
      V3<double> nor, vk;
      double tiefe;
      if (s.halten[i] & GC_H_KONTAKT) {
        if (kontakt_roh(s.koerper[0], s.x[i], s.koll[i] + s.kontakt_rand, &nor, &tiefe, &vk)) {
          fl = 1;
          nn = nor;
          soll = (tiefe - s.kontakt_rand) / s.h + dot(vk, nor);
          vkk = vk;
        } else {
          atomicAdd((unsigned long long*)&gs->halten_abweichungen, 1ull);
        }
      }
    } else {
      for (int b = 0; b < s.n_koerper; ++b) {
        V3<double> nor, vk;
        double tiefe;
        if (kontakt(s.koerper[b], s.x[i], s.koll[i] + s.kontakt_rand, &nor, &tiefe, &vk)) {
          const double t = tiefe - s.kontakt_rand;
          fl = 1;
          nn = nor;
          soll = t / s.h + dot(vk, nor);
          vkk = vk;
        }
      }
    }
    if (s.halten_modus == 1) s.halten[i] = fl == 1 ? GC_H_KONTAKT : 0;
    s.fflag[i] = fl;
    s.fnorm[i] = cast3<T>(nn);
    s.fnorm64[i] = nn;
    s.vn_soll[i] = soll;
    s.vk[i] = vkk;
    s.fn_kraft[i] = 0.0;
    if (fl == 1) ++lokal;
  }
  if (lokal) atomicAdd(&gs->n_fest, lokal);
}

// Per point: force, load, H*v, right-hand side, diagonal block and its inverse
// (implizit.rs:722-723, 745-751, 767-768, 793-795, 819-825, 887-894; the device
// form of gpu.rs sammeln).
template <typename T>
__global__ void k_sammeln(const SysDev<T>* __restrict__ sp) {
  const SysDev<T>& s = *sp;
  for (int ii = blockIdx.x * blockDim.x + threadIdx.x; ii < s.n; ii += gridDim.x * blockDim.x) {
    const uint32_t i = (uint32_t)ii;
    const double m = s.masse[i];
    const V3<double> g = {s.g[0], s.g[1], s.g[2]};
    V3<double> f = g * m - s.v[i] * (s.c * m);
    V3<double> last = g * m;
    const uint32_t* o = s.off + GC_KAT * i;
    for (uint32_t e = o[0]; e < o[1]; ++e) {
      const uint32_t ent = s.csr[e], c = ent >> 2;
      const V3<double> fi = s.k_f[c];
      if ((ent & 3u) == 0u) {
        f = f + fi;
        if (s.k_zug[c]) last = last + fi;
      } else {
        f = f - fi;
        if (s.k_zug[c]) last = last - fi;
      }
    }
    for (uint32_t e = o[1]; e < o[2]; ++e) {
      const uint32_t ent = s.csr[e], c = ent >> 2, mm = ent & 3u;
      f = f - cast3<double>(s.b_grad[4 * c + mm]) * s.b_kf[c];
    }
    if (s.selbst)
      for (uint32_t e = s.fk_off[i]; e < s.fk_off[i + 1]; ++e) {
        const uint32_t code = s.fk_ent[e], fe = code >> 3, q = code & 7u;
        f = f + s.fe_kraft[fe] * s.fe_gew[5 * fe + q];
      }
    for (uint32_t e = o[2]; e < o[3]; ++e) {
      const uint32_t ent = s.csr[e], c = ent >> 2, k = ent & 3u;
      f = f + s.m_f[3 * c + k];
      last = last + s.m_z[3 * c + k];
    }
    const V3<T> hv = mal_h_zeile(s, i, [&](uint32_t j) { return cast3<T>(s.v[j]); });
    const V3<double> rhs = (f - cast3<double>(hv) * s.h) * s.h;
    s.f[i] = f;
    s.last[i] = last;
    s.b[i] = cast3<T>(rhs);
    const T md = T(m * (1.0 + s.h * s.c));
    s.mdiag[i] = md;
    // diagonal block (implizit.rs:232-279)
    T d[9] = {md, T(0), T(0), T(0), md, T(0), T(0), T(0), md};
    const T h2 = T(s.h2);
    for (uint32_t e = o[0]; e < o[1]; ++e) {
      const uint32_t c = s.csr[e] >> 2;
      const T kk = s.k_kk[c], q = s.k_q[c];
      const V3<T> n = s.k_n[c];
      const T sk = kk * h2 * q;
      d[0] += sk; d[4] += sk; d[8] += sk;
      const T sa = kk * h2 * (T(1) - q);
      const T nv[3] = {n.x, n.y, n.z};
      for (int a = 0; a < 3; ++a)
        for (int bb = 0; bb < 3; ++bb) d[3 * a + bb] += sa * nv[a] * nv[bb];
    }
    for (uint32_t e = o[1]; e < o[2]; ++e) {
      const uint32_t ent = s.csr[e], c = ent >> 2, mm = ent & 3u;
      const V3<T> gr = s.b_grad[4 * c + mm];
      const T sb = s.b_kt[c] * h2;
      const T gv[3] = {gr.x, gr.y, gr.z};
      for (int a = 0; a < 3; ++a)
        for (int bb = 0; bb < 3; ++bb) d[3 * a + bb] += sb * gv[a] * gv[bb];
    }
    for (uint32_t e = o[2]; e < o[3]; ++e) {
      const uint32_t ent = s.csr[e], c = ent >> 2, k = ent & 3u;
      T bl[9];
      diag_block_k(s.m_el[c], s.m_h[c], (int)k, bl);
      for (int q = 0; q < 9; ++q) d[q] += h2 * bl[q];
    }
    if (s.selbst)
      for (uint32_t e = s.fe_off[i]; e < s.fe_off[i + 1]; ++e) {  // implizit.rs:267-277
        const uint32_t code = s.fe_ent[e], fe = code >> 5, pr = code & 31u, q = pr / 5u, r = pr % 5u;
        const T hom = T(s.h2 * (-s.fe_gew[5 * fe + q] * s.fe_gew[5 * fe + r]));
        for (int a = 0; a < 9; ++a) d[a] += hom * s.fe_block[9 * fe + a];
      }
    T inv[9];
    m3_inv(d, inv);
    for (int q = 0; q < 9; ++q) s.pinv[9 * i + q] = inv[q];
  }
}

// Per-step reset of the shared counters (a fill kernel, so a captured step
// records no memset node). The barrier error flag is sticky: the host reads
// and clears it after every call, so a failure in any substep is reported.
__global__ void k_reset(GridState* gs, Slot* slots) {
  if (threadIdx.x == 0) {
    gs->zaehler = 0;
    gs->frei[0] = 0;
    gs->frei[1] = 0;
    gs->n_fest = 0;
  }
  if (threadIdx.x < 2) gs->kp_pass[threadIdx.x] = 0;
  barriere_leeren(slots);
}

// Clears a system's barrier counters before a PCG launch that no k_reset or
// release pass precedes.
__global__ void k_barriere_leeren(GridState* gs, Slot* slots) {
  if (threadIdx.x == 0) gs->zaehler = 0;
  barriere_leeren(slots);
}

// Copies every tile entry's operator record from the per-constraint arrays to
// its entry index (after each build; the records are constant during a solve).
template <typename T>
__global__ void k_kachel_saetze(const SysDev<T>* __restrict__ sp) {
  const SysDev<T>& s = *sp;
  const uint32_t ne = s.kachel_el_off[s.n_kacheln];
  for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < ne; e += gridDim.x * blockDim.x) {
    const uint32_t code = s.kachel_el[e], c = code & 0x3FFFFFFFu, kat = code >> 30;
    if (kat == 0) {
      KanteK<T> r;
      r.n = s.k_n[c];
      r.q = s.k_q[c];
      r.kk = s.k_kk[c];
      s.kachel_kante[e] = r;
    } else if (kat == 1) {
      BiegeK<T> r;
      for (int l = 0; l < 4; ++l) r.g[l] = s.b_grad[4 * (size_t)c + l];
      r.kt = s.b_kt[c];
      s.kachel_biege[e] = r;
    } else {
      MembranK<T> r;
      r.el = s.m_el[c];
      r.h = s.m_h[c];
      s.kachel_membran[e] = r;
    }
  }
}

// Warm start (implizit.rs:906): dv of the previous step, h*g on the first.
template <typename T>
__global__ void k_warm(const SysDev<T>* __restrict__ sp, int hat_vorher) {
  const SysDev<T>& s = *sp;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    const V3<double> gh = V3<double>{s.g[0], s.g[1], s.g[2]} * s.h;
    const V3<double> w = hat_vorher ? s.dv_vorher[i] : gh;
    s.dv[i] = cast3<T>(w);
    if (s.dv64) s.dv64[i] = w;
  }
}

// The step's solution dv at point i in FP64: dv64 where the FP32 solve
// refines in FP64, the bulk-type dv otherwise.
template <typename T>
__device__ __forceinline__ V3<double> dv_loesung(const SysDev<T>& s, uint32_t i) {
  return s.dv64 ? s.dv64[i] : cast3<double>(s.dv[i]);
}

// ---------------------------------------------------------------------------
// Tiled operator for the PCG. Stage A puts the operand of every point the
// tile's constraints touch into LDS (computed once per point); stage B
// computes every listed constraint's contribution to the tile's own points,
// in parallel over constraints, into one LDS slot per CSR entry; stage C sums
// each own point's slots in CSR order — the reference's summation order
// (implizit.rs:185-221).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ V3<T> filter_fl(uint8_t fl, V3<T> n, V3<T> v) {
  if (fl == 0) return v;
  if (fl == 2) return zero3<T>();
  return v - n * dot(v, n);
}

// Coherent load of a vector that another block of the grid wrote during this
// launch (the PCG's dv, z, p). Agent-scope loads are served by the level that
// is coherent across the XCDs, where the writer's release at the grid barrier
// (its L2 write-back) has put the data; so the barrier needs no acquire (no L2
// invalidation), and the read-only operator records stay cached in each XCD's
// L2 from one iteration to the next. Every read of such data across blocks
// goes through this function; a block's own rows are read with plain loads.
template <typename T>
__device__ __forceinline__ V3<T> lade_k(const GC_GM V3<T>* a, uint32_t j) {
  const GC_GM T* q = (const GC_GM T*)(a + j);
  return {__hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
          __hip_atomic_load(q + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
          __hip_atomic_load(q + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)};
}
template <typename T>
__device__ __forceinline__ V3<T> lade_k(const V3<T>* a, uint32_t j) {
  return lade_k(gm(a), j);
}

// One constraint's contributions to the tile's own points, each into its LDS
// slot (pos = 0xFFFF: that corner is not in the tile).
template <typename T>
__device__ __forceinline__ void membran_beitrag(const Membran<T>& el, const Hessen<T>& hh, uint2 lw, uint2 pw, V3<T>* slot,
                                                const V3<T>* u) {
  const V3<T> d3[3] = {u[lw.x & 0xFFFFu], u[lw.x >> 16], u[lw.y & 0xFFFFu]};
  V3<T> o[3];
  mal_h_aus_alle(el, hh, d3, o);
  const uint32_t pos[3] = {pw.x & 0xFFFFu, pw.x >> 16, pw.y & 0xFFFFu};
#pragma unroll
  for (int k = 0; k < 3; ++k)
    if (pos[k] != 0xFFFFu) slot[pos[k]] = o[k];
}

template <typename T>
__device__ __forceinline__ void biege_beitrag(const V3<T> (&g)[4], T kt, uint2 lw, uint2 pw, V3<T>* slot, const V3<T>* u) {
  const uint32_t lv[4] = {lw.x & 0xFFFFu, lw.x >> 16, lw.y & 0xFFFFu, lw.y >> 16};
  const uint32_t pos[4] = {pw.x & 0xFFFFu, pw.x >> 16, pw.y & 0xFFFFu, pw.y >> 16};
  T gp = T(0);
#pragma unroll
  for (int l = 0; l < 4; ++l) gp += dot(g[l], u[lv[l]]);
  const T kg = kt * gp;
#pragma unroll
  for (int m = 0; m < 4; ++m)
    if (pos[m] != 0xFFFFu) slot[pos[m]] = g[m] * kg;
}

// Stage B of a tile: every listed constraint's contributions. The tile's list
// is sorted by category (edges, bending, membranes; kachel_kat gives the
// boundaries), so each loop runs one kind of constraint without a branch on
// the kind. (Two constraints per thread and pass, loads first, was measured
// slower: 7.7 -> 10.0 us per iteration for this stage on the drape.)
// The per-tile constants of a solve (row range, CSR base, point-list and
// element-list bounds, first spring entry).
struct KachelK {
  int lo, hi;
  uint32_t basis, v0, v1, e0, e1, e2, e3, fe0;
};
template <typename T>
__device__ __forceinline__ KachelK kachel_konst(const SysDev<T>& s, int t) {
  KachelK k;
  k.lo = gm(s.kachel_zeilen)[2 * t];
  k.hi = gm(s.kachel_zeilen)[2 * t + 1];
  k.basis = gm(s.off)[GC_KAT * k.lo];
  k.v0 = gm(s.kachel_v_off)[t];
  k.v1 = gm(s.kachel_v_off)[t + 1];
  k.e0 = gm(s.kachel_el_off)[t];
  k.e3 = gm(s.kachel_el_off)[t + 1];
  k.e1 = gm(s.kachel_kat)[2 * t];
  k.e2 = gm(s.kachel_kat)[2 * t + 1];
  k.fe0 = s.selbst ? gm(s.fe_off)[k.lo] : 0u;
  return k;
}

template <typename T>
__device__ __forceinline__ void kachel_beitraege(const SysDev<T>& s, const KachelK& kk, V3<T>* slot, const V3<T>* u) {
  const uint32_t e0 = kk.e0, e1 = kk.e1, e2 = kk.e2, e3 = kk.e3;
  const uint32_t bd = blockDim.x;
  for (uint32_t e = e0 + threadIdx.x; e < e1; e += bd) {
    const uint2 lw = lg((const GC_GM uint2*)(s.kachel_lv + 4 * (size_t)e));
    const uint2 pw = lg((const GC_GM uint2*)(s.kachel_pos + 4 * (size_t)e));
    const KanteK<T> rk = lg(gm(s.kachel_kante) + (e));
    const V3<T> n = rk.n;
    const T q = rk.q, kk = rk.kk;
    const V3<T> d = u[lw.x & 0xFFFFu] - u[lw.x >> 16];
    const V3<T> bd3 = (n * (dot(n, d) * (T(1) - q)) + d * q) * kk;
    if ((pw.x & 0xFFFFu) != 0xFFFFu) slot[pw.x & 0xFFFFu] = bd3;
    if ((pw.x >> 16) != 0xFFFFu) slot[pw.x >> 16] = bd3 * T(-1);
  }
  for (uint32_t e = e1 + threadIdx.x; e < e2; e += bd) {
    const uint2 lw = lg((const GC_GM uint2*)(s.kachel_lv + 4 * (size_t)e));
    const uint2 pw = lg((const GC_GM uint2*)(s.kachel_pos + 4 * (size_t)e));
    const BiegeK<T> rb = lg(gm(s.kachel_biege) + (e));
    biege_beitrag(rb.g, rb.kt, lw, pw, slot, u);
  }
  for (uint32_t e = e2 + threadIdx.x; e < e3; e += bd) {
    const uint2 lw = lg((const GC_GM uint2*)(s.kachel_lv + 4 * (size_t)e));
    const uint2 pw = lg((const GC_GM uint2*)(s.kachel_pos + 4 * (size_t)e));
    const MembranK<T> rm = lg(gm(s.kachel_membran) + (e));
    membran_beitrag(rm.el, rm.h, lw, pw, slot, u);
  }
}

// H*u of point i from its LDS slots, in CSR order.
template <typename T>
__device__ __forceinline__ V3<T> kachel_zeile_ab(uint32_t a, uint32_t b, const V3<T>* slot) {
  V3<T> hp = zero3<T>();
  uint32_t q = a;
  // four slots are loaded before their sums; the sums keep CSR order
  for (; q + 4 <= b; q += 4) {
    const V3<T> v0 = slot[q], v1 = slot[q + 1], v2 = slot[q + 2], v3 = slot[q + 3];
    hp = hp + v0;
    hp = hp + v1;
    hp = hp + v2;
    hp = hp + v3;
  }
  for (; q < b; ++q) hp = hp + slot[q];
  return hp;
}
template <typename T>
__device__ __forceinline__ V3<T> kachel_zeile(const SysDev<T>& s, int i, uint32_t basis, const V3<T>* slot) {
  return kachel_zeile_ab(gm(s.off)[GC_KAT * i] - basis, gm(s.off)[GC_KAT * i + GC_KAT] - basis, slot);
}

// The search direction of point j, p = S(z + beta p_old) (implizit.rs:420-422);
// the first direction is S(z) (implizit.rs:390). The owner and every tile that
// reads the point evaluate this same expression, so they agree to the bit.
// K = true for a point another block owns: its z and p are read with
// coherent loads (see lade_k).
// richtung with the point's flag and normal already at hand (staged in LDS)
template <typename T>
__device__ __forceinline__ V3<T> richtung_fn(const SysDev<T>& s, uint32_t j, uint8_t fl, V3<T> n, const GC_GM V3<T>* po, T tb,
                                              bool erste) {
  V3<T> q = lade_k(s.z, j);
  if (!erste) q = q + lade_k(po, j) * tb;
  return filter_fl(fl, n, q);
}
template <bool K, typename T>
__device__ __forceinline__ V3<T> richtung(const SysDev<T>& s, uint32_t j, const GC_GM V3<T>* po, T tb, bool erste) {
  const uint8_t fl = gm(s.fflag)[j];
  V3<T> q = K ? lade_k(s.z, j) : lg(s.z + j);
  if (!erste) q = q + (K ? lade_k(po, j) : lg(po + j)) * tb;
  return filter_fl(fl, lg(gm(s.fnorm) + (j)), q);
}

// ---------------------------------------------------------------------------
// Persistent modified PCG (CpuLoeser::loese, implizit.rs:370-426), the
// reference's recurrences. One launch per solve, one block per CU; a block
// owns tiles t = b, b+nb, ... Two grid all-reduces per iteration (p.Ap and
// r.z); the new direction needs no barrier of its own because every tile
// rebuilds the directions of the points it reads (stage A).
// durchgang 1 is the second solve after a release; it returns at once unless
// a point was released.
// ---------------------------------------------------------------------------
// The solve of one system by the nb blocks of rank 0..nb-1 that share gs and
// slots; lds_roh is the block's dynamic LDS.
// ZR: rows may be held in registers (below); off for the batched kernel,
// whose blocks carry more state and would spill.
// Iterative refinement of the FP32 solve (SysDev::schatten), over the rows of
// the block's tiles.
//
// dv64 += dv, dv = 0: the correction of the last iterations folded in.
template <typename T>
__device__ __forceinline__ void falten_fp64(const SysDev<T>& s, int rang, int nb) {
  for (int t = rang; t < s.n_kacheln; t += nb) {
    const int lo = gm(s.kachel_zeilen)[2 * t], hi = gm(s.kachel_zeilen)[2 * t + 1];
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      s.dv64[i] = s.dv64[i] + cast3<double>(lg(gm(s.dv) + (i)));
      sg(gm(s.dv) + (i), zero3<T>());
    }
  }
}

// dv = 0: the correction of the last iterations dropped (it did not lower
// the FP64 residual).
template <typename T>
__device__ __forceinline__ void verwerfen_fp64(const SysDev<T>& s, int rang, int nb) {
  for (int t = rang; t < s.n_kacheln; t += nb) {
    const int lo = gm(s.kachel_zeilen)[2 * t], hi = gm(s.kachel_zeilen)[2 * t + 1];
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) sg(gm(s.dv) + (i), zero3<T>());
  }
}

// r = S(b - A x) of the FP64 system at x = dv64 (+ dv where mit_dv);
// part[0] += sum S(b) P S(b), part[1] += r . S(P r), in FP64 (P, the
// bulk-type preconditioner, applied in FP64); r and z = P r go to the bulk
// type for the iterations. dv64 and dv of other blocks' points are read with
// agent-scope loads (after the grid barrier that follows their last writes).
template <typename T>
__device__ __forceinline__ void rest_fp64(const SysDev<T>& s, int rang, int nb, bool mit_dv, double* part) {
  const SysDev<double>& sd = *(const SysDev<double>*)s.schatten;
  for (int t = rang; t < s.n_kacheln; t += nb) {
    const int lo = gm(s.kachel_zeilen)[2 * t], hi = gm(s.kachel_zeilen)[2 * t + 1];
    for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
      const V3<double> ap = mal_a_zeile(sd, (uint32_t)i, [&](uint32_t j) {
        const V3<double> x = lade_k(s.dv64, j);
        return mit_dv ? x + cast3<double>(lade_k(s.dv, j)) : x;
      });
      const uint8_t fl = gm(s.fflag)[i];
      const V3<double> n = lg(gm(s.fnorm64) + (i));
      const V3<double> bi = sd.b[i];
      const V3<double> ri = filter_fl(fl, n, bi - ap);
      const GC_GM T* P = gm(s.pinv) + 9 * (size_t)i;
      auto pmal = [&](V3<double> v) {
        return V3<double>{(double)P[0] * v.x + (double)P[1] * v.y + (double)P[2] * v.z,
                          (double)P[3] * v.x + (double)P[4] * v.y + (double)P[5] * v.z,
                          (double)P[6] * v.x + (double)P[7] * v.y + (double)P[8] * v.z};
      };
      if (fl != 2) {
        const V3<double> fb = filter_fl(fl, n, bi);
        part[0] += dot(fb, pmal(fb));
      }
      part[1] += dot(ri, filter_fl(fl, n, pmal(ri)));
      const V3<T> rt = cast3<T>(ri);
      sg(gm(s.r) + (i), rt);
      sg(gm(s.z) + (i), m3_mal_g(P, rt));
    }
  }
}

// The outcome of a refined FP32 solve (after its k_verf rounds and, if
// handed over, its FP64 iterations), as k_pcg reports an unrefined one.
template <typename T>
__device__ __forceinline__ void loesung_zaehlen(const SysDev<T>& s, GridState* gs, bool uebergeben) {
  const double delta = gs->delta, ziel = gs->ziel;
  gs->nicht_konv = delta <= ziel ? 0 : 1;  // a NaN delta counts as not converged
  gs->rest = sqrt(delta / gmax(gs->bb, s.boden));
  gs->loesungen += 1;
  if (!(delta <= ziel)) gs->nicht_konvergiert += 1;
  if (uebergeben) gs->rundungsgrenze += 1;
  gs->rest_max = gmax(gs->rest_max, gs->rest);
}

// k_verf: the FP64 side of a refined FP32 solve (SysDev::schatten), between
// the rounds of FP32 iterations (k_pcg), one launch per step of the
// sequence start, round, check, round, check, ..., k_pcg64:
// - modus 0 (start): held points get their target normal velocity in the
//   FP64 solution dv64 (implizit.rs:375-384), the correction dv is cleared;
//   r = S(b - A dv64) of the FP64 system, bb, delta and the stopping target
//   (implizit.rs:385-400), r and z = P r in the bulk type for the round.
// - modus 1 (check): the residual at dv64 + dv. A correction that lowered
//   the FP64 residual is folded into dv64; one that did not (a round that
//   diverged in FP32) is dropped and the residual taken at dv64 again. The
//   solve is done when that residual meets the stopping test or the
//   iteration budget is spent; otherwise the next round starts from it, and
//   after the last round (letzte) the solve is handed over to FP64
//   iterations (k_pcg64).
template <typename T>
__device__ __forceinline__ void verf_schritt(const SysDev<T>& s, GridState* gs, Slot* slots, int durchgang, int modus, int letzte,
                                             int rang, int nb, double* sh) {
  if (durchgang == 1 && __hip_atomic_load(&gs->frei[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
  if (modus == 1 && __hip_atomic_load(&gs->fertig, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  // every block reads the epoch the solve's last launch ended at before its
  // first arrival, so rank 0 (which passes the last barrier only after every
  // block arrived) can store this launch's last epoch for the next one
  unsigned epoch = __hip_atomic_load(&gs->zaehler, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (modus == 0) {
    for (int t = rang; t < s.n_kacheln; t += nb) {
      const int lo = gm(s.kachel_zeilen)[2 * t], hi = gm(s.kachel_zeilen)[2 * t + 1];
      for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const uint8_t fl = gm(s.fflag)[i];
        sg(gm(s.dv) + (i), zero3<T>());
        if (fl == 2) {
          s.dv64[i] = zero3<double>();
        } else if (fl != 0) {
          const V3<double> n = lg(gm(s.fnorm64) + (i));
          const double soll = s.soll_direkt ? gm(s.soll_direkt)[i] : gm(s.vn_soll)[i] - dot(lg(gm(s.v) + (i)), n);
          s.dv64[i] = filter_fl(fl, n, s.dv64[i]) + n * soll;
        }
      }
    }
    double nix[1] = {0.0};
    if (!grid_allreduce<1>(nix, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
    double part[2] = {0.0, 0.0};
    rest_fp64(s, rang, nb, false, part);
    if (!grid_allreduce<2>(part, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
    if (rang == 0 && threadIdx.x == 0) {
      const double bb = part[0], delta = part[1];
      const double ziel = gmax(s.cg_tol2 * gmax(bb, delta), s.boden);
      gs->bb = bb;
      gs->ziel = ziel;
      gs->delta = delta;
      gs->delta_best = delta;
      gs->iterationen = 0;
      gs->uebergabe = 0u;
      gs->fertig = delta <= ziel ? 1u : 0u;
      gs->zaehler = epoch;
    }
    return;
  }
  const double best = gs->delta_best;
  double delta = 0.0, neu_best = best;
  for (bool mit_dv = true;;) {
    double part[2] = {0.0, 0.0};
    rest_fp64(s, rang, nb, mit_dv, part);
    if (!grid_allreduce<2>(part, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
    delta = part[1];
    if (!mit_dv) break;  // the residual at dv64 alone, after a dropped correction
    if (delta <= best) {
      // the correction lowered the FP64 residual: dv64 += dv (every block
      // has read dv and dv64 before the all-reduce above)
      falten_fp64(s, rang, nb);
      neu_best = delta;
      break;
    }
    verwerfen_fp64(s, rang, nb);  // dropped
    mit_dv = false;
  }
  if (rang == 0 && threadIdx.x == 0) {
    gs->zaehler = epoch;
    gs->delta = delta;
    gs->delta_best = neu_best;
    if (delta <= gs->ziel || gs->iterationen >= s.cg_max) {
      gs->fertig = 1u;
    } else {
      gs->neustarts += 1;
      if (letzte) {
        gs->uebergabe = 1u;
        gs->fertig = 1u;
      }
    }
  }
}

// The FP32 solve's last resort: when refinement makes no headway (the
// system is too ill-conditioned for FP32 iterations to lower the FP64
// residual), k_pcg hands the solve over (GridState::uebergabe) and
// k_pcg64, launched after every FP32 PCG launch, goes on from dv64 as the
// reference's modified PCG (implizit.rs:385-417) entirely in FP64, on the
// FP64 copy of the system: its records through the per-point CSR lists, its
// FP64 preconditioner, and FP64 vectors r, z, p, A p of its own. Three grid
// barriers per iteration (the direction p is gathered by other blocks'
// rows). A launch in which nothing was handed over returns at once; its own
// kernel keeps this loop's registers out of k_pcg's.
template <typename T>
__device__ __forceinline__ void pcg64_loesen(const SysDev<T>& s, GridState* gs, Slot* slots, int durchgang, int rang, int nb,
                                             double* sh) {
  if (durchgang == 1 && __hip_atomic_load(&gs->frei[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
  if (__hip_atomic_load(&gs->uebergabe, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
    if (rang == 0 && threadIdx.x == 0) loesung_zaehlen(s, gs, false);
    return;
  }
  const SysDev<double>& sd = *(const SysDev<double>*)s.schatten;
  V3<double>* const r = sd.r;
  V3<double>* const z = sd.z;
  V3<double>* const p = sd.p[0];
  V3<double>* const ap = sd.ap;
  const double ziel = gs->ziel;
  int it = gs->iterationen;
  unsigned epoch = __hip_atomic_load(&gs->zaehler, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  auto zeilen = [&](auto f) {
    for (int t = rang; t < s.n_kacheln; t += nb) {
      const int lo = gm(s.kachel_zeilen)[2 * t], hi = gm(s.kachel_zeilen)[2 * t + 1];
      for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) f(i);
    }
  };
  // r = S(b - A dv64), z = P r, delta = r . S(z)
  double part[1] = {0.0};
  zeilen([&](int i) {
    const uint8_t fl = gm(s.fflag)[i];
    const V3<double> n = lg(gm(s.fnorm64) + (i));
    const V3<double> ri = filter_fl(fl, n, sd.b[i] - mal_a_zeile(sd, (uint32_t)i, [&](uint32_t j) { return lade_k(s.dv64, j); }));
    const V3<double> zi = m3_mal(sd.pinv + 9 * (size_t)i, ri);
    r[i] = ri;
    z[i] = zi;
    part[0] += dot(ri, filter_fl(fl, n, zi));
  });
  if (!grid_allreduce<1>(part, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
  double delta = part[0];
  const int it0 = it;
  double beta = 0.0;
  for (bool erste = true; delta > ziel && it < s.cg_max; erste = false) {
    // p = S(z + beta p)
    zeilen([&](int i) {
      const V3<double> q = erste ? z[i] : z[i] + p[i] * beta;
      p[i] = filter_fl(gm(s.fflag)[i], lg(gm(s.fnorm64) + (i)), q);
    });
    double nix[1] = {0.0};
    if (!grid_allreduce<1>(nix, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
    // ap = S(A p), ca = p . ap
    double ca[1] = {0.0};
    zeilen([&](int i) {
      const V3<double> a = filter_fl(gm(s.fflag)[i], lg(gm(s.fnorm64) + (i)),
                                     mal_a_zeile(sd, (uint32_t)i, [&](uint32_t j) { return lade_k(p, j); }));
      ap[i] = a;
      ca[0] += dot(p[i], a);
    });
    if (!grid_allreduce<1, false>(ca, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
    if (ca[0] <= 2.2250738585072014e-308 || !isfinite(ca[0])) break;
    const double alpha = delta / ca[0];
    double neu[1] = {0.0};
    zeilen([&](int i) {
      s.dv64[i] = s.dv64[i] + p[i] * alpha;
      const V3<double> ri = r[i] - ap[i] * alpha;
      r[i] = ri;
      const V3<double> zi = m3_mal(sd.pinv + 9 * (size_t)i, ri);
      z[i] = zi;
      neu[0] += dot(ri, zi);
    });
    if (!grid_allreduce<1>(neu, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
    beta = neu[0] / delta;
    delta = neu[0];
    ++it;
  }
  if (rang == 0 && threadIdx.x == 0) {
    gs->zaehler = epoch;
    gs->uebergabe = 0u;
    gs->iterationen = it;
    gs->delta = delta;
    gs->cg_iterationen += it - it0;
    loesung_zaehlen(s, gs, true);
  }
}

template <typename T, bool ZR>
__device__ __forceinline__ void pcg_loesen(const SysDev<T>& s, GridState* gs, Slot* slots, int durchgang, int rang, int nb,
                                           unsigned char* lds_roh, double* sh) {
  // the second solve runs only if the release pass freed a point (the same
  // answer for every block of the launch)
  if (durchgang == 1) {
    if (__hip_atomic_load(&gs->frei[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
  }
  // a refined FP32 solve that k_verf has finished takes no further round
  if (sizeof(T) == 4 && s.schatten && __hip_atomic_load(&gs->fertig, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return;
  V3<T>* slot = (V3<T>*)lds_roh;
  V3<T>* ul = slot + s.kachel_slots_max;
  V3<T>* fterm = ul + s.kachel_v_max;
  // a block that owns one tile (the default shape) keeps the tile's constants
  // in registers and its point list with flags and normals in LDS for the
  // whole solve (flags and normals change only between solves)
  V3<T>* kn = fterm + s.fe_lds_kap;
  uint32_t* kj = (uint32_t*)(kn + s.kachel_v_max);
  uint8_t* kf = (uint8_t*)(kj + s.kachel_v_max);
  const bool eine = rang < s.n_kacheln && rang + nb >= s.n_kacheln;
  const bool eine_lds = eine && s.punkte_lds;
  KachelK k1{};
  if (eine) k1 = kachel_konst(s, rang);
  if (eine_lds) {
    for (uint32_t v = k1.v0 + threadIdx.x; v < k1.v1; v += blockDim.x) {
      const uint32_t j = gm(s.kachel_v)[v];
      kj[v - k1.v0] = j;
      kf[v - k1.v0] = gm(s.fflag)[j];
      kn[v - k1.v0] = lg(gm(s.fnorm) + (j));
    }
    __syncthreads();
  }
  auto konst = [&](int t) { return eine ? k1 : kachel_konst(s, t); };
  // FP32 one-tile blocks with at most one row per thread also hold that row's
  // solve constants (slot and spring ranges, diagonal, flag, normal) in
  // registers and load its z and previous direction at the start of each
  // operator pass, beside the staging loads, so stage C reads only LDS.
  const bool zeile_reg = ZR && sizeof(T) == 4 && eine && k1.hi - k1.lo <= (int)blockDim.x;
  const int zr = eine ? k1.lo + (int)threadIdx.x : -1;
  const bool hat_zeile = zeile_reg && zr < k1.hi;
  uint32_t r_a = 0, r_b = 0, r_fa = 0, r_fb = 0;
  T r_md = T(0);
  uint8_t r_fl = 0;
  V3<T> r_n = zero3<T>();
  if (hat_zeile) {
    r_a = gm(s.off)[GC_KAT * zr] - k1.basis;
    r_b = gm(s.off)[GC_KAT * zr + GC_KAT] - k1.basis;
    if (s.selbst) {
      r_fa = gm(s.fe_off)[zr];
      r_fb = gm(s.fe_off)[zr + 1];
    }
    r_md = gm(s.mdiag)[zr];
    r_fl = gm(s.fflag)[zr];
    r_n = lg(gm(s.fnorm) + (zr));
  }
  unsigned epoch = 0;
  auto zeilen = [&](int t, int* lo, int* hi) { *lo = gm(s.kachel_zeilen)[2 * t]; *hi = gm(s.kachel_zeilen)[2 * t + 1]; };
  const T th2 = T(s.h2);

  // FP32 bulk with an FP64 copy of the system (verf): the launch is one round
  // of FP32 iterations on the correction dv. k_verf, launched before and
  // after every round, holds the FP64 solution dv64, takes the FP64
  // residual, keeps or drops the round's correction and decides whether the
  // solve has converged (SysDev::schatten); this kernel keeps no FP64 code.
  const bool verf = sizeof(T) == 4 && s.schatten != nullptr;
  double bb = 0.0, delta = 0.0, ziel = 0.0;
  int it = 0;
  bool grenze = false;
  int neustarts = 0;
  if (verf) {
    bb = gs->bb;
    ziel = gs->ziel;
    delta = gs->delta;
    it = gs->iterationen;
    epoch = __hip_atomic_load(&gs->zaehler, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    // start: held along the normal get the target normal velocity (implizit.rs:375-384)
    for (int t = rang; t < s.n_kacheln; t += nb) {
      int lo, hi;
      zeilen(t, &lo, &hi);
      for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const uint8_t fl = gm(s.fflag)[i];
        if (fl == 2) {
          sg(gm(s.dv) + (i), zero3<T>());
        } else if (fl != 0) {  // 1; 3 only in the adjoint's solve (see k_freigabe)
          const V3<double> n = lg(gm(s.fnorm64) + (i));
          const double soll = s.soll_direkt ? gm(s.soll_direkt)[i] : gm(s.vn_soll)[i] - dot(lg(gm(s.v) + (i)), n);
          const V3<T> dvf = filter_fl(fl, lg(gm(s.fnorm) + (i)), lg(gm(s.dv) + (i)));
          sg(gm(s.dv) + (i), dvf + lg(gm(s.fnorm) + (i)) * T(soll));
        }
      }
    }
    double nix[1] = {0.0};
    if (!grid_allreduce<1>(nix, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
  }
  const int it_anfang = it;
  // FP32 restarts without the FP64 copy (the adjoint's FP32 solve). The FP32
  // product A p is exact only to a few ulps, so with a stiff system the
  // recursively updated residual drifts from b - A dv: it stalls near what
  // FP32 can reach and can then grow without bound. When delta stops
  // improving (400 iterations without a new best) or grows 1e4x above its
  // best, the solve starts again from the current dv with the true FP32
  // residual (residual replacement), at most 3 times; after that it ends as
  // not converged and is counted at its rounding floor. With the FP64 copy
  // such a round ends instead and k_verf restarts from the FP64 residual.
  // The FP64 build runs the reference's loop unchanged.
  for (bool erster_start = true;; erster_start = false) {
    if (!verf) {
    // r = S(b - A dv), bb = sum S(b) P S(b), z = P r, delta = r . S(z) (implizit.rs:385-391)
    double part[2] = {0.0, 0.0};
    for (int t = rang; t < s.n_kacheln; t += nb) {
      const KachelK kk = konst(t);
      const int lo = kk.lo, hi = kk.hi;
      const uint32_t basis = kk.basis, v0 = kk.v0, v1 = kk.v1;
      for (uint32_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) ul[v - v0] = lade_k(s.dv, gm(s.kachel_v)[v]);
      __syncthreads();
      kachel_beitraege(s, kk, slot, ul);
      __syncthreads();
      for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
        const V3<T> hp = feder_zeile(s, (uint32_t)i, kachel_zeile(s, i, basis, slot), [&](uint32_t j) { return lade_k(s.dv, j); });
        const V3<T> ap = hp * th2 + lg(gm(s.dv) + (i)) * gm(s.mdiag)[i];
        const uint8_t fl = gm(s.fflag)[i];
        const V3<T> n = lg(gm(s.fnorm) + (i));
        const V3<T> bi = lg(gm(s.b) + (i));
        const V3<T> ri = filter_fl(fl, n, bi - ap);
        const GC_GM T* P = gm(s.pinv) + 9 * (size_t)i;
        if (fl != 2) {
          const V3<T> fb = filter_fl(fl, n, bi);
          part[0] += (double)dot(fb, m3_mal_g(P, fb));
        }
        const V3<T> zi = m3_mal_g(P, ri);
        part[1] += (double)dot(ri, filter_fl(fl, n, zi));
        sg(gm(s.r) + (i), ri);
        sg(gm(s.z) + (i), zi);
      }
      __syncthreads();
    }
    if (!grid_allreduce<2>(part, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
    delta = part[1];
    if (erster_start) {
      bb = part[0];
      ziel = gmax(s.cg_tol2 * gmax(bb, delta), s.boden);
    }
    }
    double beta = 0.0;
    int cur = 0;
    const int it_start = it;
    // one row per thread (FP32): the row's dv and r stay in registers from
    // here to the end of the loop (no other block reads them), A p is never
    // stored, and an iteration stores only p and z, which other blocks read
    V3<T> u_dv = zero3<T>(), u_r = zero3<T>();
    if (hat_zeile) {
      u_dv = lg(gm(s.dv) + (zr));
      u_r = lg(gm(s.r) + (zr));
    }
    double dmin = delta;
    int it_min = it;
    bool neu_starten = false;
    while (delta > ziel && it < s.cg_max) {
#ifdef GC_PHASEN
      // thread 0's stamps, in LDS: registers would be held by every lane
      const bool uhr = threadIdx.x == 0;
      __shared__ unsigned long long tp[7];
      if (uhr) tp[0] = wall_clock64();
#define GC_PHASE(k) \
  if (uhr) tp[k] = wall_clock64();
#else
#define GC_PHASE(k)
#endif
      GC_GM V3<T>* pn = gm(s.p[cur]);
      const GC_GM V3<T>* po = gm(s.p[cur ^ 1]);
      const T tb = T(beta);
      const bool erste = it == it_start;
      // ap = S(A p), ca = p . ap (implizit.rs:401-405)
      double ca[1] = {0.0};
      // one row per thread (FP32): the row's p and A p stay in registers for
      // the update, whose preconditioner block is loaded at the end of stage
      // C, so that the load runs during the p.Ap all-reduce
      V3<T> u_p = zero3<T>(), u_ap = zero3<T>();
      T u_pinv[9];
      for (int t = rang; t < s.n_kacheln; t += nb) {
        const KachelK kk = konst(t);
        const int lo = kk.lo, hi = kk.hi;
        const uint32_t fe0 = kk.fe0, basis = kk.basis, v0 = kk.v0, v1 = kk.v1;
        V3<T> vor_z = zero3<T>(), vor_p = zero3<T>();
        if (hat_zeile) {
          vor_z = lg(gm(s.z) + (zr));
          if (!erste) vor_p = lg(po + zr);
        }
        if (eine_lds) {
          for (uint32_t v = v0 + threadIdx.x; v < v1; v += blockDim.x)
            ul[v - v0] = richtung_fn(s, kj[v - v0], kf[v - v0], kn[v - v0], po, tb, erste);
        } else {
          for (uint32_t v = v0 + threadIdx.x; v < v1; v += blockDim.x) ul[v - v0] = richtung<true>(s, gm(s.kachel_v)[v], po, tb, erste);
        }
        __syncthreads();
        GC_PHASE(1)
        kachel_beitraege(s, kk, slot, ul);
        if (s.selbst) feder_terme(s, lo, hi, [&](uint32_t j) { return richtung<true>(s, j, po, tb, erste); }, fterm);
        __syncthreads();
        GC_PHASE(2)
        if (zeile_reg) {
          if (hat_zeile) {
            V3<T> q = vor_z;
            if (!erste) q = q + vor_p * tb;
            const V3<T> pi = filter_fl(r_fl, r_n, q);
            const V3<T> hk = kachel_zeile_ab(r_a, r_b, slot);
            const V3<T> hp = s.selbst ? feder_summe_ab(s, r_fa, r_fb, hk, fterm, fe0) : hk;
            const V3<T> ap = filter_fl(r_fl, r_n, hp * th2 + pi * r_md);
            sg(pn + zr, pi);
            ca[0] += (double)dot(pi, ap);
            u_p = pi;
            u_ap = ap;
#pragma unroll
            for (int q = 0; q < 9; ++q) u_pinv[q] = gm(s.pinv)[9 * (size_t)zr + q];
          }
        } else {
          for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
            const uint8_t fl = gm(s.fflag)[i];
            const V3<T> pi = richtung<false>(s, (uint32_t)i, po, tb, erste);
            const V3<T> hk = kachel_zeile(s, i, basis, slot);
            const V3<T> hp = s.selbst ? feder_summe(s, (uint32_t)i, hk, fterm, fe0) : hk;
            const V3<T> ap = filter_fl(fl, lg(gm(s.fnorm) + (i)), hp * th2 + pi * gm(s.mdiag)[i]);
            sg(pn + i, pi);
            sg(gm(s.ap) + (i), ap);
            ca[0] += (double)dot(pi, ap);
          }
        }
        __syncthreads();
      }
      GC_PHASE(3)
      // p and A p are read by other blocks only in the next operator pass,
      // after the r.z all-reduce, whose release writes them back
      if (!grid_allreduce<1, false>(ca, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
      GC_PHASE(4)
      if (ca[0] <= 2.2250738585072014e-308 || !isfinite(ca[0])) break;
      const double alpha = delta / ca[0];
      // dv += p alpha, r -= ap alpha, z = P r, neu = r . z (implizit.rs:410-417)
      double neu[1] = {0.0};
      const T ta = T(alpha);
      if (zeile_reg) {
        if (hat_zeile) {
          u_dv = u_dv + u_p * ta;
          const V3<T> ri = u_r - u_ap * ta;
          u_r = ri;
          const V3<T> zi = m3_mal(u_pinv, ri);
          sg(gm(s.z) + (zr), zi);
          neu[0] += (double)dot(ri, zi);
        }
      } else
      for (int t = rang; t < s.n_kacheln; t += nb) {
        int lo, hi;
        if (eine) {
          lo = k1.lo;
          hi = k1.hi;
        } else {
          zeilen(t, &lo, &hi);
        }
        for (int i = lo + threadIdx.x; i < hi; i += blockDim.x) {
          sg(gm(s.dv) + (i), lg(gm(s.dv) + (i)) + lg(pn + i) * ta);
          const V3<T> ri = lg(gm(s.r) + (i)) - lg(gm(s.ap) + (i)) * ta;
          sg(gm(s.r) + (i), ri);
          const V3<T> zi = m3_mal_g(gm(s.pinv) + 9 * (size_t)i, ri);
          sg(gm(s.z) + (i), zi);
          neu[0] += (double)dot(ri, zi);
        }
      }
      GC_PHASE(5)
      if (!grid_allreduce<1>(neu, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
      GC_PHASE(6)
#ifdef GC_PHASEN
      if (uhr && s.n_kacheln <= nb) {  // one tile per block: the phases are one pass each
        if (rang == 0) {
          for (int k = 0; k < 6; ++k) gs->phasen[k] += tp[k + 1] - tp[k];
          gs->phasen_it += 1;
        }
        unsigned long long* pb = gs->pro_block;
        if (pb) {
          pb[6 * rang] += tp[1] - tp[0];
          pb[6 * rang + 1] += tp[2] - tp[1];
          pb[6 * rang + 2] += tp[3] - tp[2];
          pb[6 * rang + 3] += tp[5] - tp[4];
          pb[6 * rang + 4] += tp[4] - tp[3];
          pb[6 * rang + 5] += tp[6] - tp[5];
        }
      }
#endif
#undef GC_PHASE
      beta = neu[0] / delta;
      delta = neu[0];
      cur ^= 1;
      ++it;
      if constexpr (sizeof(T) == 4) {
        if (delta < dmin) {
          dmin = delta;
          it_min = it;
        } else if (delta > 1e4 * dmin || it - it_min >= 400 || !isfinite(delta)) {
          neu_starten = true;
          break;
        }
      }
    }
    if (hat_zeile) {
      sg(gm(s.dv) + (zr), u_dv);
      sg(gm(s.r) + (zr), u_r);
    }
    // a restart reads the dv of other blocks' points: publish it first
    if (!verf && neu_starten && neustarts < 3) {
      double nix[1] = {0.0};
      if (!grid_allreduce<1>(nix, sh, slots, gs, epoch, (unsigned)rang, (unsigned)nb)) return;
    }
    if (verf) break;  // the round ends; k_verf checks it in FP64
    if (!neu_starten) break;
    if (neustarts == 3) {
      grenze = true;
      break;
    }
    ++neustarts;
  }
  if (verf) {  // the round's outcome for k_verf
    if (rang == 0 && threadIdx.x == 0) {
      gs->zaehler = epoch;
      gs->iterationen = it;
      gs->delta = delta;
      gs->cg_iterationen += it - it_anfang;
    }
    return;
  }
  if (rang == 0 && threadIdx.x == 0) {
    gs->iterationen = it;
    gs->nicht_konv = delta <= ziel ? 0 : 1;  // a NaN delta counts as not converged
    gs->rest = sqrt(delta / gmax(bb, s.boden));
    gs->delta = delta;
    gs->ziel = ziel;
    gs->bb = bb;
    gs->loesungen += 1;
    gs->cg_iterationen += it;
    if (!(delta <= ziel)) gs->nicht_konvergiert += 1;
    if (grenze) gs->rundungsgrenze += 1;
    gs->neustarts += neustarts;
    gs->rest_max = gmax(gs->rest_max, gs->rest);
  }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_pcg(const SysDev<T>* __restrict__ sp, GridState* gs, Slot* slots, int durchgang) {
  extern __shared__ __align__(16) unsigned char lds_roh[];
  __shared__ double sh[16 * 2 + 2];
  pcg_loesen<T, true>(*sp, gs, slots, durchgang, (int)blockIdx.x, (int)gridDim.x, lds_roh, sh);
}

// The FP64 steps of a refined FP32 solve (verf_schritt), with the PCG's
// launch shape and no dynamic LDS.
__global__ void __launch_bounds__(1024) k_verf(const SysDev<float>* __restrict__ sp, GridState* gs, Slot* slots, int durchgang,
                                               int modus, int letzte) {
  __shared__ double sh[16 * 2 + 2];
  verf_schritt(*sp, gs, slots, durchgang, modus, letzte, (int)blockIdx.x, (int)gridDim.x, sh);
}

// The FP64 iterations of a handed-over FP32 solve and the statistics of
// every refined one (pcg64_loesen); the last launch of the sequence.
__global__ void __launch_bounds__(1024) k_pcg64(const SysDev<float>* __restrict__ sp, GridState* gs, Slot* slots, int durchgang) {
  __shared__ double sh[16 * 2 + 2];
  if (!sp->schatten) return;
  pcg64_loesen(*sp, gs, slots, durchgang, (int)blockIdx.x, (int)gridDim.x, sh);
}

// Batched solves: block b works for system tab[b].sys as rank tab[b].rang of
// its tab[b].nb blocks; every system runs its own iterations, reductions and
// stopping test, all in one persistent launch (all blocks co-resident).
template <typename T>
struct StapelBlock {
  const SysDev<T>* sys;
  GridState* gs;
  Slot* slots;
  int32_t rang, nb;
};

template <typename T>
__global__ void __launch_bounds__(1024) k_pcg_stapel(const StapelBlock<T>* __restrict__ tab, int durchgang) {
  extern __shared__ __align__(16) unsigned char lds_roh[];
  __shared__ double sh[16 * 2 + 2];
  const StapelBlock<T> e = tab[blockIdx.x];
  pcg_loesen<T, false>(*e.sys, e.gs, e.slots, durchgang, e.rang, e.nb, lds_roh, sh);
}

// k_verf for a batch: each system's blocks take its FP64 step.
__global__ void __launch_bounds__(1024) k_verf_stapel(const StapelBlock<float>* __restrict__ tab, int durchgang, int modus,
                                                      int letzte) {
  __shared__ double sh[16 * 2 + 2];
  const StapelBlock<float> e = tab[blockIdx.x];
  if (!e.sys->schatten) return;
  verf_schritt(*e.sys, e.gs, e.slots, durchgang, modus, letzte, e.rang, e.nb, sh);
}

// k_pcg64 for a batch: each system's blocks finish its handed-over solve.
__global__ void __launch_bounds__(1024) k_pcg64_stapel(const StapelBlock<float>* __restrict__ tab, int durchgang) {
  __shared__ double sh[16 * 2 + 2];
  const StapelBlock<float> e = tab[blockIdx.x];
  if (!e.sys->schatten) return;
  pcg64_loesen(*e.sys, e.gs, e.slots, durchgang, e.rang, e.nb, sh);
}

// Release pass (implizit.rs:916-942): normal force from the residual
// A*dv - b; a held point that pulls is released.
template <typename T>
__global__ void k_freigabe(const SysDev<T>* __restrict__ sp, GridState* gs, Slot* slots, int durchgang) {
  const SysDev<T>& s = *sp;
  // the second PCG launch starts with clear barrier counters. Block 0 clears
  // them before any early return: the held count it reads below can already
  // be 0 because other blocks released every held point, and the second solve
  // still runs then (frei[0] > 0)
  if (durchgang == 0 && blockIdx.x == 0) {
    if (threadIdx.x == 0) gs->zaehler = 0;
    barriere_leeren(slots);
  }
  if (durchgang == 1 && __hip_atomic_load(&gs->frei[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
  if (__hip_atomic_load(&gs->n_fest, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) return;
  unsigned frei = 0, weg = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    if (s.fflag[i] != 1) continue;
    V3<double> rest;
    if (s.schatten) {  // FP32 bulk: the residual of the FP64 system at the FP64 solution
      const SysDev<double>& sd = *(const SysDev<double>*)s.schatten;
      rest = mal_a_zeile(sd, (uint32_t)i, [&](uint32_t j) { return s.dv64[j]; }) - sd.b[i];
    } else {
      const V3<T> ap = mal_a_zeile(s, (uint32_t)i, [&](uint32_t j) { return s.dv[j]; });
      rest = cast3<double>(ap) - cast3<double>(s.b[i]);
    }
    const double fn = dot(rest, s.fnorm64[i]) / s.h;
    const uint16_t bit = durchgang == 1 ? GC_H_FREI1 : GC_H_FREI0;
    const bool los = s.halten_modus == 2 ? (s.halten[i] & bit) != 0 : fn < 0.0;
    if (s.halten_modus == 1 && los) s.halten[i] |= bit;
    if (los) {
      // released; after the second solve (no third follows) the point keeps
      // flag 3 = "held in the last solve, free in the update", with the
      // normal force of that solve for the step's adjoint. Every kernel of
      // the step treats 3 as free (the PCG never runs after this pass).
      if (durchgang == 1) {
        s.fflag[i] = 3;
        s.fn_kraft[i] = fn;
      } else {
        s.fflag[i] = 0;
      }
      ++frei;
      ++weg;
    } else {
      s.fn_kraft[i] = fn;
    }
  }
  if (frei) {
    atomicAdd(&gs->frei[durchgang], frei);
    atomicAdd((unsigned long long*)&gs->freigegeben, (unsigned long long)frei);
    atomicSub(&gs->n_fest, weg);
  }
}

// Friction, velocity and position update, push-out (implizit.rs:945-1013).
template <typename T>
__global__ void k_update(const SysDev<T>* __restrict__ sp, GridState* gs) {
  const SysDev<T>& s = *sp;
  unsigned kont = 0, haft = 0, stoss = 0;
  unsigned long long abweichung = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    const V3<double> dvi = dv_loesung(s, (uint32_t)i);
    s.dv_vorher[i] = dvi;
    if (s.w[i] <= 0.0) continue;
    // a decision as taken (recorded when halten_modus is 1) or as recorded (2)
    uint16_t bits = s.halten_modus == 2 ? s.halten[i] : 0;
    auto wahl = [&](uint16_t bit, bool frei) {
      const bool j = s.halten_modus == 2 ? (bits & bit) != 0 : frei;
      if (s.halten_modus == 1 && j) bits |= bit;
      return j;
    };
    if (s.halten_modus == 1) bits = s.halten[i];
    V3<double> v = s.v[i] + dvi;
    if (s.fflag[i] == 1) {
      ++kont;
      const V3<double> n = s.fnorm64[i];
      const V3<double> rel = v - s.vk[i];
      const V3<double> t = rel - n * dot(rel, n);
      const double lt = norm(t);
      const double normal = s.aus_last ? gmax(-dot(s.last[i], n), 0.0) : s.fn_kraft[i];
      const double grenze = s.reib[i] * normal * s.h * s.w[i];
      if (wahl(GC_H_HAFT, lt <= grenze)) {
        ++haft;
        v = v - t;
      } else if (lt > 0.0) {
        v = v - t * (grenze / lt);
      }
    }
    s.v[i] = v;
    const V3<double> alt = s.x[i];
    V3<double> x = alt + v * s.h;
    // implizit.rs:984-1003: a path through the skin is set onto the entry
    if (s.durchstoss_pruefen)
      for (int b = 0; b < s.n_koerper; ++b) {
        V3<double> p, nk, vk;
        bool ein = false;
        if (s.halten_modus == 2) {
          if (bits & GC_H_EINTRITT) {
            ein = durchstoss(s.koerper[b], alt, x, &p, &nk, &vk);
            if (!ein) ++abweichung;
          }
        } else {
          ein = wahl(GC_H_EINTRITT, durchstoss(s.koerper[b], alt, x, &p, &nk, &vk));
        }
        if (ein) {
          ++stoss;
          x = p + nk * s.koll[i];
          const V3<double> rel = s.v[i] - vk;
          const double vn = dot(rel, nk);
          if (wahl(GC_H_EIN_VN, vn < 0.0)) {
            V3<double> t = rel - nk * vn;
            const double lt = norm(t);
            const double grenze = s.reib[i] * -vn;
            t = wahl(GC_H_EIN_HAFT, lt <= grenze) ? zero3<double>() : t * (1.0 - grenze / lt);
            s.v[i] = vk + t;
          }
        }
      }
    for (int b = 0; b < s.n_koerper; ++b) {
      V3<double> nor, vk;
      double tiefe;
      bool raus = false;
      if (s.halten_modus == 2) {
        if (bits & GC_H_RAUS) {
          raus = kontakt_roh(s.koerper[b], x, s.koll[i], &nor, &tiefe, &vk);
          if (!raus) ++abweichung;
        }
      } else {
        raus = wahl(GC_H_RAUS, kontakt(s.koerper[b], x, s.koll[i], &nor, &tiefe, &vk));
      }
      if (raus) {
        x = x + nor * tiefe;
        const double vn = dot(s.v[i] - vk, nor);
        if (wahl(GC_H_RAUS_VN, vn < 0.0)) s.v[i] = s.v[i] - nor * vn;
      }
    }
    s.x[i] = x;
    if (s.halten_modus == 1) s.halten[i] = bits;
  }
  if (kont) atomicAdd((unsigned long long*)&gs->kontakte, (unsigned long long)kont);
  if (haft) atomicAdd((unsigned long long*)&gs->haftend, (unsigned long long)haft);
  if (stoss) atomicAdd((unsigned long long*)&gs->durchstoesse, (unsigned long long)stoss);
  if (abweichung) atomicAdd((unsigned long long*)&gs->halten_abweichungen, abweichung);
}

// Utility kernels for the Loeser-level entry points.
template <typename T>
__global__ void k_mal(const SysDev<T>* __restrict__ sp, const V3<T>* p, V3<T>* out, int nur_h) {
  const SysDev<T>& s = *sp;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < s.n; i += gridDim.x * blockDim.x) {
    auto pf = [&](uint32_t j) { return p[j]; };
    out[i] = nur_h ? mal_h_zeile(s, (uint32_t)i, pf) : mal_a_zeile(s, (uint32_t)i, pf);
  }
}

__global__ void k_fest_zaehlen(const uint8_t* fl, int n, GridState* gs) {
  unsigned lokal = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) lokal += fl[i] == 1 ? 1u : 0u;
  if (lokal) atomicAdd(&gs->n_fest, lokal);
}

// ---------------------------------------------------------------------------
// Body points against cloth triangles in the step (implizit.rs:1017-1026,
// netzkoerper.rs:625-665): pass d of body b. A pass after one that lifted
// nothing returns at once (the reference leaves the loop), so the launch
// sequence stays fixed. Pass d counts its lifts in kp_pass[d & 1], which the
// pass's first body clears while the other slot holds pass d - 1's count.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool kp_aus(const GridState* gs, int d) {
  return d > 0 && __hip_atomic_load(&gs->kp_pass[(d - 1) & 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u;
}

// Per cloth triangle: the deepest stab of the body's points. A stabbed
// triangle keeps its record at its index and joins the list (any order).
template <typename T>
__global__ void k_kp_stiche(const SysDev<T>* __restrict__ sp, int b, GridState* gs, int d, int erster, double tief,
                            const uint32_t* __restrict__ dreiecke, int nt, Stich* __restrict__ st,
                            uint32_t* __restrict__ liste, uint32_t* __restrict__ zahl) {
  if (erster && blockIdx.x == 0 && threadIdx.x == 0) gs->kp_pass[d & 1] = 0u;
  if (kp_aus(gs, d)) return;
  const SysDev<T>& s = *sp;
  const Koerper& k = s.koerper[b];
  const V3<double> versatz = {k.punkt[0], k.punkt[1], k.punkt[2]};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nt; t += gridDim.x * blockDim.x) {
    const uint32_t* q = dreiecke + 3 * (size_t)t;
    Stich h;
    h.tiefe = 0.0;
    h.kir = 0;
    h.da = 1u;
    if (!nk_stich(*k.netz, versatz, tief, -1e-9, s.x[q[0]], s.x[q[1]], s.x[q[2]], &h)) continue;
    st[t] = h;
    liste[atomicAdd(zahl, 1u)] = (uint32_t)t;
  }
}

// The lifts in cloth-triangle order (netzkoerper.rs:647-665): one block walks
// the dependency levels (a level's triangles share no point, and each point
// meets its triangles level by level in ascending order), each level over the
// list of stabbed triangles, so every point receives its lifts in the
// reference's sequence. Empties the list for the next pass.
template <typename T>
__global__ void __launch_bounds__(1024) k_kp_heben(const SysDev<T>* __restrict__ sp, int b, GridState* gs, int d, int nstufen,
                                                   const int* __restrict__ stufe, const uint32_t* __restrict__ dreiecke,
                                                   const Stich* __restrict__ st, const uint32_t* __restrict__ liste,
                                                   uint32_t* __restrict__ zahl) {
  if (kp_aus(gs, d)) return;
  const SysDev<T>& s = *sp;
  const Koerper& k = s.koerper[b];
  const NetzDev& N = *k.netz;
  const V3<double> starr = {k.geschwindigkeit[0], k.geschwindigkeit[1], k.geschwindigkeit[2]};
  const uint32_t m = *zahl;
  unsigned gehoben = 0;
  for (int l = 0; l < nstufen && m > 0; ++l) {
    for (uint32_t q = threadIdx.x; q < m; q += blockDim.x) {
      const uint32_t t = liste[q];
      if (stufe[t] != l) continue;
      const Stich h = st[t];
      const uint32_t* idx = dreiecke + 3 * (size_t)t;
      const double bary[3] = {h.bary.x, h.bary.y, h.bary.z};
      double gew[3], nenner = 0.0;
      for (int j = 0; j < 3; ++j) {
        gew[j] = s.w[idx[j]];
        nenner += bary[j] * bary[j] * gew[j];
      }
      if (nenner <= 0.0) continue;
      const V3<double> vk = N.pv ? nk::lade3(N.pv, h.kir) : starr;
      V3<double> summe = zero3<double>();
      for (int j = 0; j < 3; ++j) summe = summe + s.v[idx[j]] * bary[j];
      const double vn = fmin(dot(summe - vk, h.n), 0.0);
      for (int j = 0; j < 3; ++j) {
        const double f = bary[j] * gew[j] / nenner;
        s.x[idx[j]] = s.x[idx[j]] + h.n * (f * h.tiefe);
        s.v[idx[j]] = s.v[idx[j]] - h.n * (f * vn);
      }
      ++gehoben;
    }
    __syncthreads();
  }
  __shared__ unsigned summe_block;
  if (threadIdx.x == 0) summe_block = 0;
  __syncthreads();
  if (gehoben) atomicAdd(&summe_block, gehoben);
  __syncthreads();
  if (threadIdx.x == 0) {
    *zahl = 0u;
    if (summe_block) {
      atomicAdd(&gs->kp_pass[d & 1], summe_block);
      atomicAdd((unsigned long long*)&gs->koerperpunkte, (unsigned long long)summe_block);
    }
  }
}
