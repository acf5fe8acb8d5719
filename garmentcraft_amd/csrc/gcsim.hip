// gcsim.hip — host runtime and C ABI of the MI355X implicit step.
//
// One context = one garment system resident in HBM. The host builds the
// per-point CSR adjacency once per topology, then every substep is a fixed
// sequence of kernel launches on one stream (capturable as a HIP graph):
//   reset, build edges / bending / membranes, contacts, per-point gather,
//   warm start, PCG, release, PCG (only after a release), release, update.
#include "../../include/gcsim.h"
#include "gc_adjoint.h"
#include "gc_selbst.h"
#include "gc_naht.h"
#include "gc_xpbd.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <memory>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

const int BUILD_THREADS = 256;

// A mesh body: its topology on the host, its arrays on the device.
struct Netzhost {
  int n = 0, nt = 0;
  double suchtiefe = 0.0;
  std::vector<uint32_t> dreiecke;
  std::vector<double> punkte;  // the current pose (host copy: volume, hull)
  int64_t randkanten = 0;
  double volumen = 0.0;
  bool hat_gitter = false;
  double huelle_lo[3] = {0, 0, 0}, huelle_hi[3] = {0, 0, 0};  // the pose's point box
  NetzDev bild{};  // host image of the device struct
  double *punkte_d = nullptr, *fn_d = nullptr, *pn_d = nullptr, *kugel_d = nullptr, *pv_d = nullptr;
  uint32_t *dreiecke_d = nullptr, *nachbar_d = nullptr, *inz_anfang_d = nullptr, *inz_d = nullptr;
  uint32_t *anfang_d = nullptr, *zaehler_d = nullptr, *cursor_d = nullptr, *roh_d = nullptr, *eintraege_d = nullptr;
  // the grid of body points (same raster), for body points against cloth triangles
  uint32_t *p_anfang_d = nullptr, *p_roh_d = nullptr, *p_eintraege_d = nullptr;
  size_t nz_kap = 0, e_kap = 0;
  NetzDev* dev = nullptr;
  void frei() {
    for (void* q : {(void*)punkte_d, (void*)fn_d, (void*)pn_d, (void*)kugel_d, (void*)pv_d, (void*)dreiecke_d,
                    (void*)nachbar_d, (void*)inz_anfang_d, (void*)inz_d, (void*)anfang_d, (void*)zaehler_d,
                    (void*)cursor_d, (void*)roh_d, (void*)eintraege_d, (void*)p_anfang_d, (void*)p_roh_d,
                    (void*)p_eintraege_d, (void*)dev})
      if (q) (void)hipFree(q);
    punkte_d = fn_d = pn_d = kugel_d = pv_d = nullptr;
    dreiecke_d = nachbar_d = inz_anfang_d = inz_d = anfang_d = zaehler_d = cursor_d = roh_d = eintraege_d = nullptr;
    p_anfang_d = p_roh_d = p_eintraege_d = nullptr;
    dev = nullptr;
  }
};

// The XPBD path's own constraint lists (cpu.rs), cut into runs and levels
// whose constraints share no point.
struct XpbdHost {
  bool bereit = false;
  int na = 0, nn = 0, nb = 0, nm = 0;
  std::vector<int> a_laeufe, n_laeufe, b_laeufe, m_stufen;  // run / level starts, last entry = count
  uint32_t *a_p = nullptr, *n_p = nullptr, *b_p = nullptr, *m_reihe = nullptr;
  // per point: its distance constraints (2c + side) and membrane corners (3t + k), ascending
  uint32_t *a_off = nullptr, *a_liste = nullptr, *m_off = nullptr, *m_liste = nullptr;
  V3<double>*fa = nullptr, *fm = nullptr;
  uint8_t* da = nullptr;
  // membrane levels in one single-block launch (levels of a few hundred
  // elements), or one launch per level (large levels)
  bool stufen_einzeln = false;
  int* stufen_d = nullptr;
  double *a_ruhe = nullptr, *a_alpha = nullptr, *n_ruhe = nullptr, *b_ruhe = nullptr, *b_alpha = nullptr;
  double n_alpha = 0.0;
  V3<double> *delta = nullptr, *korr = nullptr, *last = nullptr;
  XpStatDev* st = nullptr;
};

struct Koerperhost {
  Koerper k;
  float* werte_dev = nullptr;
  std::shared_ptr<Netzhost> netz;
};

}  // namespace

struct gc_ctx {
  gc_optionen opt{};
  std::string err;
  std::string name;
  hipStream_t stream = nullptr;
  int cus = 0;
  // host copy of the topology
  int n = 0;
  std::vector<double> masse, w, koll, reib;
  std::vector<uint32_t> k_ij;
  std::vector<double> k_k, k_ruhe;
  std::vector<uint8_t> k_naht;
  std::vector<int32_t> k_quelle;  // device kante -> input edge index, or -(1 + seam index)
  int n_kanten_in = 0, n_naehte = 0, nk = 0;
  std::vector<uint32_t> b_p;
  std::vector<double> b_k, b_ruhe;
  std::vector<uint32_t> m_p;
  std::vector<double> m_d;
  double msum = 0.0;
  gc_param param{};
  std::vector<Koerperhost> koerper;
  XpbdHost xp;
  // body points against cloth triangles (implizit.rs:1017-1026): the cloth
  // triangles' dependency levels (a level's triangles share no point, each
  // point meets its triangles in ascending order), the deepest stab per
  // triangle and the list of stabbed triangles; built once passes are asked for
  bool kp_bereit = false;
  int kp_nt = 0, kp_nstufen = 0;
  uint32_t* kp_dreiecke = nullptr;
  int* kp_stufe = nullptr;
  Stich* kp_stich = nullptr;
  uint32_t* kp_liste = nullptr;
  uint32_t* kp_zahl = nullptr;
  // device
  std::vector<void*> allocs;
  SysDev<double> sd{};
  SysDev<float> sf{};
  // device copy of the active system description (kernels read it through a
  // pointer) and the bytes of its last upload
  void* sys_dev = nullptr;
  std::vector<unsigned char> sys_alt;
  // FP32 bulk: the FP64 copy of the system that the PCG's refinement and the
  // release pass take their residual from (its records, right-hand side and
  // mass diagonal are its own; topology, state and forces are shared)
  SysDev<double> schatten{};
  void* schatten_dev = nullptr;
  std::vector<unsigned char> schatten_alt;
  GridState* gs = nullptr;
  Slot* slots = nullptr;
  bool topo = false, hat_param = false, built = false;
  bool hat_vorher = false;
  int pcg_blocks = 0, pcg_threads = 1024, pcg_max_blocks = 0;
  size_t pcg_lds = 0;       // dynamic LDS of the PCG kernel (largest tile, spring terms)
  size_t pcg_lds_kachel = 0;  // its part for the largest tile (slots and staging)
  uint32_t fe_kap_allein = 0;  // spring terms in LDS when the system runs alone
  size_t lds_cu = 65536;    // LDS of one CU
  size_t lds_budget = 0;    // bytes of LDS a tile may fill
  int n_kacheln = 0;
  long kachel_eintraege = 0;  // entries of all tiles' element lists
  // timing
  std::vector<hipEvent_t> ev;
  double ms_schritt = 0.0, ms_pcg = 0.0;
  // cloth self-collision: inputs, the mesh tables built from them on first
  // use (for the extra thickness they were built with), device buffers
  std::vector<uint32_t> sk_dreiecke_in;
  std::vector<double> sk_e_punkt, sk_e_dreieck, sk_a_dreieck;
  std::vector<uint32_t> sk_abstand_in;        // all distance constraints (weld partners)
  std::vector<double> sk_abstand_ruhe_in;
  std::vector<uint32_t> sk_naht_in;
  std::vector<double> sk_naht_ruhe;
  bool sk_bereit = false;
  double sk_dicke = -1.0;
  // capacity factor of the self-collision lists: a step whose lists overflow
  // is taken again from its start with the capacities doubled (the reference
  // keeps them in growable vectors)
  int sk_mal = 1;
  std::vector<void*> sk_allocs;              // buffers of the self-collision tables (freed when they grow)
  SelbstDev sk{};
  SelbstDev* sk_dev = nullptr;
  uint32_t sk_eimer_p = 0, sk_eimer_k = 0, sk_eimer_d = 0, sk_eimer_max = 0;
  double* sk_zellen = nullptr;               // device: proximity, cut, impact cell
  uint32_t* sk_fe_fuell = nullptr;
  uint32_t* sk_fk_fuell = nullptr;
  uint32_t* sk_st_fuell = nullptr;
  uint32_t* sk_scan_teil = nullptr;          // chunk sums of the multi-block scan
  uint32_t sk_scan_max = 0;                  // longest array it scans
  double* sk_naht_ruhe_dev = nullptr;
  void* sk_block_t = nullptr;                // spring blocks in the bulk type
  void* sk_fe_t = nullptr;                   // per spring entry: its term in the PCG (bulk type)
  void* sk_fe_satz = nullptr;                // per spring entry: its points, weights and block (bulk type)
  // graph of one substep
  hipGraphExec_t graph = nullptr;
  // generation of the launch sequence: grows whenever a setting that the
  // host turns into launches changes (self-collision on or off, impact
  // iterations, a body moved), so that captured graphs are captured anew
  uint64_t folge = 0;
  // adjoint of the step (allocated on first use): FP64 buffers, the adjoint
  // right-hand side and solution in the bulk type, a copy of the system
  // description whose b, dv and normal targets are the adjoint's, and a grid
  // state of its own (the step's statistics stay the step's)
  bool adj_bereit = false;
  AdjDev adj{};
  void* adj_b = nullptr;
  void* adj_mu = nullptr;
  void* adj_sys_dev = nullptr;
  std::vector<unsigned char> adj_sys_host;
  GridState* gs_adj = nullptr;
  long long adj_schritte = 0;
  // saved states (x, v, warm start) by slot number, for checkpointed rollouts
  struct Zustand {
    V3<double>* x = nullptr;
    V3<double>* v = nullptr;
    V3<double>* dvv = nullptr;
    uint32_t* folge = nullptr;   // self-collision push-spring sequence per triangle (selbstkollision.rs:1024-1026)
    bool hat_vorher = false;
    // the solutions of the step taken from this state (gc_step_aufnehmen):
    // dv after solve 1 and solve 2 in the bulk type, and for a refined FP32
    // solve its FP64 solution after each
    void* dv_loes[2] = {nullptr, nullptr};
    V3<double>* dv64_loes[2] = {nullptr, nullptr};
    bool hat_loesung = false;
  };
  std::map<int, Zustand> zustaende;
  // while a step records its solutions into a slot, or takes them from one
  // instead of solving (set only for the duration of that step)
  Zustand* aufnahme = nullptr;
  const Zustand* wiedergabe = nullptr;
  // recorded decisions of steps by slot (gc_entscheidungen)
  std::map<int, uint16_t*> entscheidungen;
  // the state at the start of the current gc_step / gc_step_graph call, with
  // the step counters, for taking the call again after a list overflow
  Zustand anfang;
  GridState anfang_gs{};
};

#define HIPCK(ctx, call)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);                       \
      return 1;                                                                             \
    }                                                                                       \
  } while (0)

static int fehler(gc_ctx* c, const std::string& m) {
  c->err = m;
  return 1;
}

// Device buffer, zeroed in order with the uploads that follow on the
// context's stream (a plain hipMemset runs on the null stream, which a
// non-blocking stream does not wait for, and could clear a buffer after its
// upload).
template <typename X>
static X* dalloc(gc_ctx* c, size_t count) {
  void* p = nullptr;
  size_t bytes = std::max<size_t>(1, count) * sizeof(X);
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  if (hipMemsetAsync(p, 0, bytes, c->stream) != hipSuccess) return nullptr;
  c->allocs.push_back(p);
  return (X*)p;
}

template <typename X>
static int hoch(gc_ctx* c, X* dst, const X* src, size_t count) {
  if (count == 0) return 0;
  HIPCK(c, hipMemcpyAsync(dst, src, count * sizeof(X), hipMemcpyHostToDevice, c->stream));
  return 0;
}

// A setting that the launch sequence depends on changed: the context's graph
// is dropped, and a batch's graph is captured again at its next replay.
static void folge_neu(gc_ctx* c) {
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  ++c->folge;
}

// The tables of body points against cloth triangles, once the parameters ask
// for passes and the topology holds the cloth triangles. Outside any step, so
// that no allocation falls into a captured substep.
static int kp_bereiten(gc_ctx* c) {
  if (c->kp_bereit || !c->topo || c->param.koerperpunkt_durchgaenge <= 0 || c->sk_dreiecke_in.empty()) return 0;
  const int nt = (int)(c->sk_dreiecke_in.size() / 3);
  std::vector<int> stufe((size_t)nt), zuletzt((size_t)c->n, -1);
  int stufen = 0;
  for (int t = 0; t < nt; ++t) {
    int l = 0;
    for (int k = 0; k < 3; ++k) l = std::max(l, zuletzt[c->sk_dreiecke_in[3 * (size_t)t + k]] + 1);
    stufe[t] = l;
    for (int k = 0; k < 3; ++k) zuletzt[c->sk_dreiecke_in[3 * (size_t)t + k]] = l;
    stufen = std::max(stufen, l + 1);
  }
  c->kp_dreiecke = dalloc<uint32_t>(c, c->sk_dreiecke_in.size());
  c->kp_stufe = dalloc<int>(c, (size_t)nt);
  c->kp_stich = dalloc<Stich>(c, (size_t)nt);
  c->kp_liste = dalloc<uint32_t>(c, (size_t)nt);
  c->kp_zahl = dalloc<uint32_t>(c, 1);  // cleared by dalloc, emptied by each lift pass
  if (!c->kp_dreiecke || !c->kp_stufe || !c->kp_stich || !c->kp_liste || !c->kp_zahl)
    return fehler(c, "hipMalloc failed (body points against cloth triangles)");
  if (hoch(c, c->kp_dreiecke, c->sk_dreiecke_in.data(), c->sk_dreiecke_in.size()) || hoch(c, c->kp_stufe, stufe.data(), stufe.size()))
    return 1;
  HIPCK(c, hipStreamSynchronize(c->stream));
  c->kp_nt = nt;
  c->kp_nstufen = stufen;
  c->kp_bereit = true;
  return 0;
}

static void frei_alles(gc_ctx* c) {
  if (c->graph) {
    (void)hipGraphExecDestroy(c->graph);
    c->graph = nullptr;
  }
  for (void* p : c->allocs) (void)hipFree(p);
  c->allocs.clear();
  for (void* p : c->sk_allocs) (void)hipFree(p);
  c->sk_allocs.clear();
  for (auto& k : c->koerper) {
    if (k.werte_dev) (void)hipFree(k.werte_dev);
    if (k.netz) k.netz->frei();
  }
}

template <typename T> static SysDev<T>& sys(gc_ctx* c);
template <> SysDev<double>& sys<double>(gc_ctx* c) { return c->sd; }
template <> SysDev<float>& sys<float>(gc_ctx* c) { return c->sf; }

template <typename T>
static const SysDev<T>* dsys(gc_ctx* c) { return (const SysDev<T>*)c->sys_dev; }

// Uploads the host copy of the system description when it differs from the
// last upload. Called by every entry point before it launches kernels, never
// inside a graph capture: it waits for the stream first, so no kernel still
// reads the old copy. A captured graph reads the description at replay time,
// so parameter and body updates reach it without a new capture.
// The FP64 copy of an FP32 system: the scalars and the shared arrays of the
// bulk system, its own FP64 records (allocated with the topology), uploaded
// when it changed.
static int schatten_hoch(gc_ctx* c) {
  if (!c->schatten_dev) return 0;
  const SysDev<float>& f = c->sf;
  SysDev<double>& d = c->schatten;
  d.n = f.n; d.nk = f.nk; d.nb = f.nb; d.nm = f.nm;
  d.h = f.h; d.h2 = f.h2; d.c = f.c; d.naht_k = f.naht_k; d.kontakt_rand = f.kontakt_rand;
  for (int q = 0; q < 3; ++q) d.g[q] = f.g[q];
  d.cg_max = f.cg_max; d.cg_tol2 = f.cg_tol2; d.boden = f.boden;
  d.aus_last = f.aus_last; d.durchstoss_pruefen = f.durchstoss_pruefen;
  d.off = f.off; d.csr = f.csr;
  d.k_ij = f.k_ij; d.k_k = f.k_k; d.k_ruhe = f.k_ruhe; d.k_naht = f.k_naht; d.k_f = f.k_f; d.k_zug = f.k_zug;
  d.b_p = f.b_p; d.b_k = f.b_k; d.b_ruhe = f.b_ruhe; d.b_kf = f.b_kf;
  d.m_p = f.m_p; d.m_el = f.m_el64; d.m_el64 = f.m_el64; d.m_f = f.m_f; d.m_z = f.m_z;
  d.masse = f.masse; d.w = f.w; d.koll = f.koll; d.reib = f.reib;
  d.x = f.x; d.v = f.v; d.f = f.f; d.last = f.last;
  d.fflag = f.fflag; d.fnorm = f.fnorm64; d.fnorm64 = f.fnorm64;
  d.vn_soll = f.vn_soll; d.vk = f.vk; d.fn_kraft = f.fn_kraft;
  d.dv = f.dv64; d.dv_vorher = f.dv_vorher;
  d.selbst = f.selbst;
  d.fe_off = f.fe_off; d.fe_ent = f.fe_ent; d.fk_off = f.fk_off; d.fk_ent = f.fk_ent;
  d.fe_punkte = f.fe_punkte; d.fe_gew = f.fe_gew; d.fe_kraft = f.fe_kraft;
  d.fe_block = f.selbst ? c->sk.fe_block : nullptr;  // the FP64 blocks of the springs
  d.n_koerper = f.n_koerper;
  for (int b = 0; b < GC_MAX_KOERPER; ++b) d.koerper[b] = f.koerper[b];
  const unsigned char* roh = (const unsigned char*)&d;
  if (c->schatten_alt.size() == sizeof(SysDev<double>) && std::memcmp(c->schatten_alt.data(), roh, sizeof(SysDev<double>)) == 0)
    return 0;
  HIPCK(c, hipStreamSynchronize(c->stream));
  HIPCK(c, hipMemcpyAsync(c->schatten_dev, roh, sizeof(SysDev<double>), hipMemcpyHostToDevice, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  c->schatten_alt.assign(roh, roh + sizeof(SysDev<double>));
  return 0;
}

template <typename T>
static int sys_hoch(gc_ctx* c) {
  if (sizeof(T) == 4 && schatten_hoch(c)) return 1;
  const SysDev<T>& s = sys<T>(c);
  const unsigned char* roh = (const unsigned char*)&s;
  if (c->sys_alt.size() == sizeof(SysDev<T>) && std::memcmp(c->sys_alt.data(), roh, sizeof(SysDev<T>)) == 0) return 0;
  HIPCK(c, hipStreamSynchronize(c->stream));
  HIPCK(c, hipMemcpyAsync(c->sys_dev, roh, sizeof(SysDev<T>), hipMemcpyHostToDevice, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  c->sys_alt.assign(roh, roh + sizeof(SysDev<T>));
  return 0;
}

static unsigned gitter(long count) {
  long b = (count + BUILD_THREADS - 1) / BUILD_THREADS;
  return (unsigned)std::max(1L, std::min(b, 8192L));
}

// ---------------------------------------------------------------------------
// topology upload
// ---------------------------------------------------------------------------
template <typename T>
static int topologie_laden(gc_ctx* c, const std::vector<uint32_t>& off, const std::vector<uint32_t>& csr) {
  SysDev<T>& s = sys<T>(c);
  std::memset((void*)&s, 0, sizeof(SysDev<T>));  // padding too: uploads compare raw bytes
  c->sys_alt.clear();
  c->sys_dev = dalloc<SysDev<T>>(c, 1);
  if (!c->sys_dev) return fehler(c, "hipMalloc failed (system description)");
  const int n = c->n, nk = c->nk, nb = (int)c->b_k.size(), nm = (int)(c->m_p.size() / 3);
  s.n = n; s.nk = nk; s.nb = nb; s.nm = nm;
  uint32_t* d_off = dalloc<uint32_t>(c, off.size());
  uint32_t* d_csr = dalloc<uint32_t>(c, csr.size());
  uint32_t* d_kij = dalloc<uint32_t>(c, 2 * (size_t)nk);
  double* d_kk = dalloc<double>(c, nk);
  double* d_kr = dalloc<double>(c, nk);
  uint8_t* d_kn = dalloc<uint8_t>(c, nk);
  uint32_t* d_bp = dalloc<uint32_t>(c, 4 * (size_t)nb);
  double* d_bk = dalloc<double>(c, nb);
  double* d_br = dalloc<double>(c, nb);
  uint32_t* d_mp = dalloc<uint32_t>(c, 3 * (size_t)nm);
  Membran<T>* d_mel = dalloc<Membran<T>>(c, nm);
  Membran<double>* d_mel64 = sizeof(T) == sizeof(double) ? (Membran<double>*)d_mel : dalloc<Membran<double>>(c, nm);
  double* d_masse = dalloc<double>(c, n);
  double* d_w = dalloc<double>(c, n);
  double* d_koll = dalloc<double>(c, n);
  double* d_reib = dalloc<double>(c, n);
  if (!d_off || !d_csr || !d_kij || !d_kk || !d_kr || !d_kn || !d_bp || !d_bk || !d_br || !d_mp || !d_mel || !d_mel64 || !d_masse ||
      !d_w || !d_koll || !d_reib)
    return fehler(c, "hipMalloc failed (topology)");
  std::vector<Membran<T>> mel(nm);
  std::vector<Membran<double>> mel64(nm);
  for (int e = 0; e < nm; ++e) {
    const double* d = &c->m_d[12 * (size_t)e];
    Membran<double>& m = mel64[e];
    m.flaeche = d[0];
    for (int q = 0; q < 4; ++q) m.r[q] = d[1 + q];
    m.kette[0] = d[5]; m.kette[1] = d[6];
    m.schuss[0] = d[7]; m.schuss[1] = d[8];
    m.ek = d[9]; m.es = d[10]; m.g = d[11];
    Membran<T>& t = mel[e];
    t.flaeche = T(m.flaeche);
    for (int q = 0; q < 4; ++q) t.r[q] = T(m.r[q]);
    t.kette[0] = T(m.kette[0]); t.kette[1] = T(m.kette[1]);
    t.schuss[0] = T(m.schuss[0]); t.schuss[1] = T(m.schuss[1]);
    t.ek = T(m.ek); t.es = T(m.es); t.g = T(m.g);
  }
  if (sizeof(T) != sizeof(double) && hoch(c, d_mel64, mel64.data(), mel64.size())) return 1;
  if (hoch(c, d_off, off.data(), off.size()) || hoch(c, d_csr, csr.data(), csr.size()) ||
      hoch(c, d_kij, c->k_ij.data(), c->k_ij.size()) || hoch(c, d_kk, c->k_k.data(), c->k_k.size()) ||
      hoch(c, d_kr, c->k_ruhe.data(), c->k_ruhe.size()) || hoch(c, d_kn, c->k_naht.data(), c->k_naht.size()) ||
      hoch(c, d_bp, c->b_p.data(), c->b_p.size()) || hoch(c, d_bk, c->b_k.data(), c->b_k.size()) ||
      hoch(c, d_br, c->b_ruhe.data(), c->b_ruhe.size()) || hoch(c, d_mp, c->m_p.data(), c->m_p.size()) ||
      hoch(c, d_mel, mel.data(), mel.size()) || hoch(c, d_masse, c->masse.data(), c->masse.size()) ||
      hoch(c, d_w, c->w.data(), c->w.size()) || hoch(c, d_koll, c->koll.data(), c->koll.size()) ||
      hoch(c, d_reib, c->reib.data(), c->reib.size()))
    return 1;
  s.off = d_off; s.csr = d_csr;
  s.k_ij = d_kij; s.k_k = d_kk; s.k_ruhe = d_kr; s.k_naht = d_kn;
  s.b_p = d_bp; s.b_k = d_bk; s.b_ruhe = d_br;
  s.m_p = d_mp; s.m_el = d_mel; s.m_el64 = d_mel64;
  s.masse = d_masse; s.w = d_w; s.koll = d_koll; s.reib = d_reib;
  // per-step and state buffers
  s.k_n = dalloc<V3<T>>(c, nk); s.k_q = dalloc<T>(c, nk); s.k_kk = dalloc<T>(c, nk);
  s.k_f = dalloc<V3<double>>(c, nk); s.k_zug = dalloc<uint8_t>(c, nk);
  s.b_kt = dalloc<T>(c, nb); s.b_grad = dalloc<V3<T>>(c, 4 * (size_t)nb); s.b_kf = dalloc<double>(c, nb);
  s.m_h = dalloc<Hessen<T>>(c, nm); s.m_f = dalloc<V3<double>>(c, 3 * (size_t)nm); s.m_z = dalloc<V3<double>>(c, 3 * (size_t)nm);
  s.x = dalloc<V3<double>>(c, n); s.v = dalloc<V3<double>>(c, n);
  s.mdiag = dalloc<T>(c, n); s.f = dalloc<V3<double>>(c, n); s.last = dalloc<V3<double>>(c, n);
  s.b = dalloc<V3<T>>(c, n); s.pinv = dalloc<T>(c, 9 * (size_t)n);
  s.fflag = dalloc<uint8_t>(c, n); s.fnorm = dalloc<V3<T>>(c, n); s.fnorm64 = dalloc<V3<double>>(c, n);
  s.vn_soll = dalloc<double>(c, n); s.vk = dalloc<V3<double>>(c, n); s.fn_kraft = dalloc<double>(c, n);
  s.dv = dalloc<V3<T>>(c, n); s.r = dalloc<V3<T>>(c, n); s.z = dalloc<V3<T>>(c, n);
  s.p[0] = dalloc<V3<T>>(c, n); s.p[1] = dalloc<V3<T>>(c, n); s.ap = dalloc<V3<T>>(c, n);
  s.dv_vorher = dalloc<V3<double>>(c, n);
  if (!s.k_n || !s.k_q || !s.k_kk || !s.k_f || !s.k_zug || !s.b_kt || !s.b_grad || !s.b_kf || !s.m_h || !s.m_f || !s.m_z ||
      !s.x || !s.v || !s.mdiag || !s.f || !s.last || !s.b || !s.pinv || !s.fflag || !s.fnorm || !s.fnorm64 || !s.vn_soll ||
      !s.vk || !s.fn_kraft || !s.dv || !s.r || !s.z || !s.p[0] || !s.p[1] || !s.ap || !s.dv_vorher)
    return fehler(c, "hipMalloc failed (vectors)");
  if (sizeof(T) == 4) {
    // the FP64 copy's own arrays: operator records, right-hand side, mass
    // diagonal, its (unused) preconditioner, and the FP64 solution
    SysDev<double>& d = c->schatten;
    std::memset((void*)&d, 0, sizeof(SysDev<double>));
    c->schatten_alt.clear();
    d.k_n = dalloc<V3<double>>(c, nk); d.k_q = dalloc<double>(c, nk); d.k_kk = dalloc<double>(c, nk);
    d.b_kt = dalloc<double>(c, nb); d.b_grad = dalloc<V3<double>>(c, 4 * (size_t)nb);
    d.m_h = dalloc<Hessen<double>>(c, nm);
    d.mdiag = dalloc<double>(c, n); d.b = dalloc<V3<double>>(c, n); d.pinv = dalloc<double>(c, 9 * (size_t)n);
    // the FP64 iterations a stalled FP32 solve hands over to
    d.r = dalloc<V3<double>>(c, n); d.z = dalloc<V3<double>>(c, n); d.p[0] = dalloc<V3<double>>(c, n);
    d.ap = dalloc<V3<double>>(c, n);
    s.dv64 = dalloc<V3<double>>(c, n);
    c->schatten_dev = dalloc<SysDev<double>>(c, 1);
    if (!d.k_n || !d.k_q || !d.k_kk || !d.b_kt || !d.b_grad || !d.m_h || !d.mdiag || !d.b || !d.pinv || !d.r || !d.z ||
        !d.p[0] || !d.ap || !s.dv64 || !c->schatten_dev)
      return fehler(c, "hipMalloc failed (FP64 copy of the system)");
    s.schatten = c->schatten_dev;
  }
  // Tiles: consecutive rows whose CSR entries plus the points their
  // constraints touch fit the LDS budget. Per tile: the constraints touching
  // its rows, each with the tile-local index of its points and the LDS slot of
  // every (constraint, own point).
  const size_t budget = c->lds_budget / sizeof(V3<T>);
  std::vector<int32_t> kz;
  std::vector<uint32_t> k_off{0}, k_el, v_off{0}, k_v, k_kat;
  std::vector<uint16_t> k_pos, k_lv;
  size_t max_slots = 0, max_v = 0;
  {
    auto punkte_von = [&](uint32_t kat, uint32_t cc, uint32_t* q) -> int {
      if (kat == 0) { q[0] = c->k_ij[2 * cc]; q[1] = c->k_ij[2 * cc + 1]; return 2; }
      if (kat == 1) { for (int m = 0; m < 4; ++m) q[m] = c->b_p[4 * (size_t)cc + m]; return 4; }
      for (int m = 0; m < 3; ++m) q[m] = c->m_p[3 * (size_t)cc + m];
      return 3;
    };
    std::vector<std::pair<uint64_t, std::array<uint16_t, 4>>> liste;
    std::vector<uint32_t> pkt;
    int lo = 0;
    size_t ziel_slots = budget * 3 / 4;
    while (lo < n) {
      int hi = lo;
      while (hi < n && (size_t)(off[GC_KAT * (size_t)(hi + 1)] - off[GC_KAT * (size_t)lo]) <= ziel_slots) ++hi;
      if (hi == lo) return fehler(c, "a single point has more constraints than fit in LDS");
      const uint32_t basis = off[GC_KAT * (size_t)lo];
      liste.clear();
      for (int i = lo; i < hi; ++i)
        for (int kat = 0; kat < 3; ++kat)
          for (uint32_t e = off[GC_KAT * (size_t)i + kat]; e < off[GC_KAT * (size_t)i + kat + 1]; ++e) {
            const uint64_t key = ((uint64_t)kat << 32) | (csr[e] >> 2);
            std::array<uint16_t, 4> pos = {0xFFFF, 0xFFFF, 0xFFFF, 0xFFFF};
            pos[csr[e] & 3u] = (uint16_t)(e - basis);
            liste.push_back({key, pos});
          }
      std::sort(liste.begin(), liste.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
      // unique points of the listed constraints
      pkt.clear();
      for (const auto& l : liste) {
        uint32_t q[4];
        int np = punkte_von((uint32_t)(l.first >> 32), (uint32_t)(l.first & 0xFFFFFFFFu), q);
        for (int m = 0; m < np; ++m) pkt.push_back(q[m]);
      }
      std::sort(pkt.begin(), pkt.end());
      pkt.erase(std::unique(pkt.begin(), pkt.end()), pkt.end());
      const size_t slots_hier = off[GC_KAT * (size_t)hi] - basis;
      if (slots_hier + pkt.size() > budget || pkt.size() > 0xFFFEu) {
        if (ziel_slots < 64) return fehler(c, "LDS budget too small for one tile");
        ziel_slots = ziel_slots * 7 / 8;  // fewer rows, try again
        continue;
      }
      max_slots = std::max(max_slots, slots_hier);
      max_v = std::max(max_v, pkt.size());
      uint32_t kat_start[3] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu};
      for (size_t q = 0; q < liste.size();) {
        std::array<uint16_t, 4> pos = liste[q].second;
        const uint32_t kat_q = (uint32_t)(liste[q].first >> 32);
        if (kat_start[kat_q] == 0xFFFFFFFFu) kat_start[kat_q] = (uint32_t)k_el.size();
        size_t r = q + 1;
        for (; r < liste.size() && liste[r].first == liste[q].first; ++r)
          for (int m = 0; m < 4; ++m)
            if (liste[r].second[m] != 0xFFFF) pos[m] = liste[r].second[m];
        const uint32_t kat = (uint32_t)(liste[q].first >> 32), cc = (uint32_t)(liste[q].first & 0xFFFFFFFFu);
        k_el.push_back((kat << 30) | cc);
        uint32_t qp[4] = {0, 0, 0, 0};
        int np = punkte_von(kat, cc, qp);
        for (int m = 0; m < 4; ++m) {
          k_pos.push_back(pos[m]);
          uint16_t lvv = 0;
          if (m < np) lvv = (uint16_t)(std::lower_bound(pkt.begin(), pkt.end(), qp[m]) - pkt.begin());
          k_lv.push_back(lvv);
        }
        q = r;
      }
      // the list is sorted by category: edges, bending, membranes; per tile
      // the first bending and the first membrane entry
      {
        const uint32_t ende = (uint32_t)k_el.size();
        uint32_t b1 = kat_start[1] != 0xFFFFFFFFu ? kat_start[1] : (kat_start[2] != 0xFFFFFFFFu ? kat_start[2] : ende);
        uint32_t b2 = kat_start[2] != 0xFFFFFFFFu ? kat_start[2] : ende;
        k_kat.push_back(b1);
        k_kat.push_back(b2);
      }
      for (uint32_t v : pkt) k_v.push_back(v);
      kz.push_back(lo);
      kz.push_back(hi);
      k_off.push_back((uint32_t)k_el.size());
      v_off.push_back((uint32_t)k_v.size());
      lo = hi;
      ziel_slots = budget * 3 / 4;
    }
  }
  c->n_kacheln = (int)(kz.size() / 2);
  int32_t* d_kz = dalloc<int32_t>(c, kz.size());
  uint32_t* d_koff = dalloc<uint32_t>(c, k_off.size());
  uint32_t* d_kkat = dalloc<uint32_t>(c, k_kat.size());
  uint32_t* d_kel = dalloc<uint32_t>(c, k_el.size());
  uint16_t* d_kpos = dalloc<uint16_t>(c, k_pos.size());
  uint16_t* d_klv = dalloc<uint16_t>(c, k_lv.size());
  uint32_t* d_voff = dalloc<uint32_t>(c, v_off.size());
  uint32_t* d_kv = dalloc<uint32_t>(c, k_v.size());
  const size_t n_ein = std::max<size_t>(1, k_el.size());
  s.kachel_kante = dalloc<KanteK<T>>(c, n_ein);
  s.kachel_biege = dalloc<BiegeK<T>>(c, n_ein);
  s.kachel_membran = dalloc<MembranK<T>>(c, n_ein);
  c->kachel_eintraege = (long)k_el.size();
  if (!d_kz || !d_koff || !d_kkat || !d_kel || !d_kpos || !d_klv || !d_voff || !d_kv || !s.kachel_kante || !s.kachel_biege ||
      !s.kachel_membran)
    return fehler(c, "hipMalloc failed (tiles)");
  if (hoch(c, d_kz, kz.data(), kz.size()) || hoch(c, d_koff, k_off.data(), k_off.size()) ||
      hoch(c, d_kkat, k_kat.data(), k_kat.size()) ||
      hoch(c, d_kel, k_el.data(), k_el.size()) || hoch(c, d_kpos, k_pos.data(), k_pos.size()) ||
      hoch(c, d_klv, k_lv.data(), k_lv.size()) || hoch(c, d_voff, v_off.data(), v_off.size()) ||
      hoch(c, d_kv, k_v.data(), k_v.size()))
    return 1;
  s.n_kacheln = c->n_kacheln;
  s.kachel_zeilen = d_kz;
  s.kachel_el_off = d_koff;
  s.kachel_kat = d_kkat;
  s.kachel_el = d_kel;
  s.kachel_pos = d_kpos;
  s.kachel_lv = d_klv;
  s.kachel_v_off = d_voff;
  s.kachel_v = d_kv;
  s.kachel_slots_max = (uint32_t)max_slots;
  s.kachel_v_max = (uint32_t)max_v;
  const size_t groesste = max_slots + max_v;
  // the CU's LDS left after the largest tile's slots and staging keeps a
  // tile's self-collision spring terms (up to 3072; more go to global memory)
  // the tile's part: slots and staging, plus the point list with flags and
  // normals that a one-tile block keeps for the whole solve
  const size_t lds_punkte = (max_v * (sizeof(V3<T>) + sizeof(uint32_t) + 1) + 15) / 16 * 16;
  const size_t lds_tile = groesste * sizeof(V3<T>) + lds_punkte, lds_rest = 2048;
  size_t fe_kap = lds_tile + lds_rest < c->lds_cu ? (c->lds_cu - lds_tile - lds_rest) / sizeof(V3<T>) : 0;
  fe_kap = std::min<size_t>(fe_kap, 3072);
  s.fe_lds_kap = (uint32_t)fe_kap;
  s.punkte_lds = 1;
  c->fe_kap_allein = (uint32_t)fe_kap;
  // a batch keeps no point lists in LDS (so two 512-thread blocks still fit a CU)
  c->pcg_lds_kachel = std::max<size_t>(16, groesste * sizeof(V3<T>));
  // PCG launch shape: co-resident blocks only (the kernel waits on all of them)
  c->pcg_lds = std::max<size_t>(16, lds_tile + fe_kap * sizeof(V3<T>));
  {
    // the kernel's static LDS plus the tile's dynamic LDS must fit one CU;
    // the occupancy query below then confirms one block per CU
    hipFuncAttributes fa{};
    HIPCK(c, hipFuncGetAttributes(&fa, (const void*)k_pcg<T>));
    if (fa.sharedSizeBytes + c->pcg_lds > c->lds_cu)
      return fehler(c, "PCG tile LDS (" + std::to_string(c->pcg_lds) + " B) plus the kernel's static LDS (" +
                           std::to_string(fa.sharedSizeBytes) + " B) exceed the CU's LDS");
  }
  HIPCK(c, hipFuncSetAttribute((const void*)k_pcg<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c->pcg_lds));
  int occ = 0;
  HIPCK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_pcg<T>, c->pcg_threads, c->pcg_lds));
  if (occ < 1) return fehler(c, "PCG kernel does not fit on a CU with this LDS budget");
  c->pcg_max_blocks = std::min(GC_MAX_BLOECKE, occ * c->cus);
  // default: one block per tile up to the co-resident maximum (one block of
  // 1024 threads per CU). A block's tiles run one after another, so a large
  // system is fastest with every CU; a small one keeps its tile count.
  int want = c->opt.bloecke > 0 ? c->opt.bloecke : c->n_kacheln;
  want = std::min(want, c->n_kacheln);
  c->pcg_blocks = std::max(1, std::min(want, c->pcg_max_blocks));
  return 0;
}

template <typename T>
static void param_setzen(gc_ctx* c) {
  SysDev<T>& s = sys<T>(c);
  for (int q = 0; q < 3; ++q) s.g[q] = c->param.schwerkraft[q];
  s.c = c->param.daempfung;
  s.cg_max = c->param.cg_max;
  s.cg_tol2 = c->param.cg_toleranz * c->param.cg_toleranz;
  s.naht_k = c->param.naht_k;
  s.kontakt_rand = c->param.kontakt_rand;
  s.aus_last = c->param.normalkraft_aus_last;
  s.durchstoss_pruefen = c->param.durchstoss_pruefen;
  s.boden = c->msum * 1e-18;
  s.n_koerper = (int)c->koerper.size();
  for (size_t b = 0; b < c->koerper.size() && b < GC_MAX_KOERPER; ++b) s.koerper[b] = c->koerper[b].k;
}

static void alle_param(gc_ctx* c) {
  if (c->opt.praezision == GC_FP32) param_setzen<float>(c);
  else param_setzen<double>(c);
}

// ---------------------------------------------------------------------------
// cloth self-collision: mesh tables and buffers (selbstkollision.rs:157-190,
// 623-693), built on the host once per garment and extra thickness
// ---------------------------------------------------------------------------
static bool selbst_an(const gc_ctx* c) { return c->param.selbstkollision != 0 && !c->sk_dreiecke_in.empty(); }

template <typename X>
static X* hoch_neu(gc_ctx* c, const std::vector<X>& v) {
  X* d = dalloc<X>(c, v.size());
  if (!d) return nullptr;
  if (!v.empty() && hipMemcpyAsync(d, v.data(), v.size() * sizeof(X), hipMemcpyHostToDevice, c->stream) != hipSuccess) return nullptr;
  return d;
}

template <typename T>
static int selbst_vorbereiten(gc_ctx* c) {
  SysDev<T>& sy = sys<T>(c);
  if (!selbst_an(c)) {
    sy.selbst = 0;
    return 0;
  }
  const double dicke = c->param.selbst_zusatzdicke;
  if (c->sk_bereit && c->sk_dicke == dicke) {
    sy.selbst = 1;
    return 0;
  }
  if (c->sk_bereit) return fehler(c, "selbst_zusatzdicke changed after the first step (create a new context)");
  HIPCK(c, hipStreamSynchronize(c->stream));
  const int n = c->n;
  const int nt = (int)(c->sk_dreiecke_in.size() / 3);
  // Stoffnetz::neu (:157-190): edges in first-encounter order
  std::map<std::pair<uint32_t, uint32_t>, uint32_t> index;
  std::vector<uint32_t> kanten, kd, dk(3 * (size_t)nt);
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k) {
      const uint32_t a = c->sk_dreiecke_in[3 * t + k], b = c->sk_dreiecke_in[3 * t + (k + 1) % 3];
      const std::pair<uint32_t, uint32_t> key{std::min(a, b), std::max(a, b)};
      auto it = index.find(key);
      uint32_t e;
      if (it == index.end()) {
        e = (uint32_t)(kanten.size() / 2);
        kanten.push_back(key.first);
        kanten.push_back(key.second);
        kd.push_back(SK_KEIN);
        kd.push_back(SK_KEIN);
        index[key] = e;
      } else {
        e = it->second;
      }
      if (kd[2 * e] == SK_KEIN) kd[2 * e] = (uint32_t)t;
      else if (kd[2 * e + 1] == SK_KEIN) kd[2 * e + 1] = (uint32_t)t;
      dk[3 * (size_t)t + k] = e;
    }
  const int ne = (int)(kanten.size() / 2);
  std::vector<double> rp(n), rk(ne), rd(nt);
  double r_max = 0.0;
  for (int i = 0; i < n; ++i) {
    rp[i] = c->koll[i] + dicke / 2.0;  // aus_szene (:625)
    r_max = std::fmax(r_max, rp[i]);
  }
  for (int e = 0; e < ne; ++e) rk[e] = std::fmax(rp[kanten[2 * e]], rp[kanten[2 * e + 1]]);
  for (int t = 0; t < nt; ++t) {
    double m = 0.0;
    for (int k = 0; k < 3; ++k) m = std::fmax(m, rp[c->sk_dreiecke_in[3 * t + k]]);
    rd[t] = m;
  }
  // seam and weld partners (:629-650)
  std::vector<std::vector<uint32_t>> partner(n);
  const size_t ns = c->sk_naht_in.size() / 2;
  for (size_t q = 0; q < ns; ++q) {
    const uint32_t i = c->sk_naht_in[2 * q], j = c->sk_naht_in[2 * q + 1];
    partner[i].push_back(j);
    partner[j].push_back(i);
  }
  for (size_t e = 0; e < c->sk_abstand_ruhe_in.size(); ++e) {
    const uint32_t i = c->sk_abstand_in[2 * e], j = c->sk_abstand_in[2 * e + 1];
    if (c->sk_abstand_ruhe_in[e] != 0.0 || i == j) continue;
    if (std::find(partner[i].begin(), partner[i].end(), j) == partner[i].end()) partner[i].push_back(j);
    if (std::find(partner[j].begin(), partner[j].end(), i) == partner[j].end()) partner[j].push_back(i);
  }
  std::vector<uint32_t> p_off(n + 1, 0), p_list;
  for (int i = 0; i < n; ++i) {
    p_off[i] = (uint32_t)p_list.size();
    p_list.insert(p_list.end(), partner[i].begin(), partner[i].end());
  }
  p_off[n] = (uint32_t)p_list.size();
  // Stoffwerte (:655-692)
  std::vector<double> e_p = c->sk_e_punkt, e_d = c->sk_e_dreieck, a_d = c->sk_a_dreieck;
  e_p.resize(n, 0.0);
  e_d.resize(nt, 0.0);
  a_d.resize(nt, 0.0);
  std::vector<double> a_p(n, 0.0), e_k(ne), a_k(ne);
  for (int t = 0; t < nt; ++t)
    for (int k = 0; k < 3; ++k) a_p[c->sk_dreiecke_in[3 * t + k]] += a_d[t] / 3.0;
  for (int e = 0; e < ne; ++e) {
    double sa = 0.0, se = 0.0;
    int cnt = 0;
    for (int q = 0; q < 2; ++q)
      if (kd[2 * e + q] != SK_KEIN) {
        sa += a_d[kd[2 * e + q]];
        se += e_d[kd[2 * e + q]];
        ++cnt;
      }
    const double k = (double)std::max(cnt, 1);
    a_k[e] = sa / k;
    e_k[e] = se / k;
  }
  std::vector<std::vector<uint32_t>> nw(n);
  for (size_t q = 0; q < ns; ++q) {
    const uint32_t i = c->sk_naht_in[2 * q], j = c->sk_naht_in[2 * q + 1];
    nw[i].push_back(j);
    nw[i].push_back((uint32_t)q);
    nw[j].push_back(i);
    nw[j].push_back((uint32_t)q);
  }
  std::vector<uint32_t> nw_off(n + 1, 0), nw_list;
  for (int i = 0; i < n; ++i) {
    nw_off[i] = (uint32_t)(nw_list.size() / 2);
    nw_list.insert(nw_list.end(), nw[i].begin(), nw[i].end());
  }
  nw_off[n] = (uint32_t)(nw_list.size() / 2);

  SelbstDev& k = c->sk;
  k = SelbstDev{};
  k.n = n;
  k.nt = nt;
  k.ne = ne;
  k.r_max = r_max;
  k.steifigkeit = c->param.selbst_steifigkeit;
  k.an_naehten = c->param.selbst_schnitte_an_naehten;
  k.stoss_iterationen = c->param.selbst_stoss_iterationen;
  const size_t mx = (size_t)std::max(n, std::max(ne, nt));
  c->sk_eimer_p = (uint32_t)std::max<size_t>(2 * (size_t)n, 64);
  c->sk_eimer_k = (uint32_t)std::max<size_t>(2 * (size_t)ne, 64);
  c->sk_eimer_d = (uint32_t)std::max<size_t>(2 * (size_t)nt, 64);
  c->sk_eimer_max = std::max(c->sk_eimer_p, std::max(c->sk_eimer_k, c->sk_eimer_d));
  const size_t mal = (size_t)c->sk_mal;
  if (mal * (32 * mx + 12 * (4 * ((size_t)n + ne) + 2 * (size_t)ne + 1024)) >= (size_t)UINT32_MAX)
    return fehler(c, "self-collision lists would exceed 32-bit indexing");
  k.kapazitaet_eintraege = (uint32_t)(mal * 32 * mx);
  k.kapazitaet_naehe = (uint32_t)(mal * 4 * ((size_t)n + ne));
  k.kapazitaet_schnitte = (uint32_t)(mal * (2 * (size_t)ne + 1024));
  k.kapazitaet_tausch = 6 * k.kapazitaet_schnitte;
  k.kapazitaet_federn = k.kapazitaet_naehe + k.kapazitaet_schnitte;
  k.kapazitaet_fe_ent = 12 * k.kapazitaet_federn;
  k.kapazitaet_fk_ent = 5 * k.kapazitaet_federn;
  k.kapazitaet_st_ent = (uint32_t)(mal * 4 * ((size_t)nt + ne));
  // every buffer allocated below belongs to the tables and is freed when they grow
  const size_t a0 = c->allocs.size();
  bool ok = true;
  auto chk = [&](const void* q) { ok = ok && q != nullptr; };
  k.dreiecke = hoch_neu(c, c->sk_dreiecke_in); chk(k.dreiecke);
  k.kanten = hoch_neu(c, kanten); chk(k.kanten);
  k.kanten_dreiecke = hoch_neu(c, kd); chk(k.kanten_dreiecke);
  k.dreieck_kanten = hoch_neu(c, dk); chk(k.dreieck_kanten);
  k.r_punkt = hoch_neu(c, rp); chk(k.r_punkt);
  k.r_kante = hoch_neu(c, rk); chk(k.r_kante);
  k.r_dreieck = hoch_neu(c, rd); chk(k.r_dreieck);
  k.partner_off = hoch_neu(c, p_off); chk(k.partner_off);
  k.partner = hoch_neu(c, p_list); chk(k.partner);
  k.nahtw_off = hoch_neu(c, nw_off); chk(k.nahtw_off);
  k.nahtw = hoch_neu(c, nw_list); chk(k.nahtw);
  k.e_punkt = hoch_neu(c, e_p); chk(k.e_punkt);
  k.a_punkt = hoch_neu(c, a_p); chk(k.a_punkt);
  k.e_kante = hoch_neu(c, e_k); chk(k.e_kante);
  k.a_kante = hoch_neu(c, a_k); chk(k.a_kante);
  k.e_dreieck = hoch_neu(c, e_d); chk(k.e_dreieck);
  k.a_dreieck = hoch_neu(c, a_d); chk(k.a_dreieck);
  k.mu = sy.reib;
  c->sk_naht_ruhe_dev = hoch_neu(c, c->sk_naht_ruhe); chk(c->sk_naht_ruhe_dev);
  k.naht_ruhe = c->sk_naht_ruhe_dev;
  k.x = sy.x;
  k.v = sy.v;
  k.masse = sy.masse;
  k.w = sy.w;
  k.fflag = sy.fflag;
  k.x_anfang = dalloc<V3<double>>(c, n); chk(k.x_anfang);
  const std::vector<uint32_t> eins(nt, 1u);  // folge starts at 1 (:1024-1026)
  k.folge = hoch_neu(c, eins); chk(k.folge);
  k.n_dreieck = dalloc<V3<double>>(c, nt); chk(k.n_dreieck);
  k.kantenebene = dalloc<V3<double>>(c, ne); chk(k.kantenebene);
  k.quader = dalloc<Quader>(c, mx); chk(k.quader);
  k.eimer_anfang = dalloc<uint32_t>(c, c->sk_eimer_max + 1); chk(k.eimer_anfang);
  k.eimer_fuell = dalloc<uint32_t>(c, c->sk_eimer_max + 1); chk(k.eimer_fuell);
  k.eintraege = dalloc<uint32_t>(c, k.kapazitaet_eintraege); chk(k.eintraege);
  k.eintraege_roh = dalloc<uint32_t>(c, k.kapazitaet_eintraege); chk(k.eintraege_roh);
  k.kapazitaet_gross = (uint32_t)(1024 * mal);
  k.gross = dalloc<uint32_t>(c, k.kapazitaet_gross); chk(k.gross);
  k.eimer_bits = dalloc<uint32_t>(c, c->sk_eimer_max / 32 + 1); chk(k.eimer_bits);
  k.zahl = dalloc<uint32_t>(c, (size_t)n + ne + nt + 1); chk(k.zahl);
  k.naehe = dalloc<Naehe>(c, k.kapazitaet_naehe); chk(k.naehe);
  k.schnitte = dalloc<Schnitt>(c, k.kapazitaet_schnitte); chk(k.schnitte);
  k.schnitt_wurzel = dalloc<uint32_t>(c, k.kapazitaet_schnitte); chk(k.schnitt_wurzel);
  k.uf = dalloc<uint32_t>(c, 2 * (size_t)k.kapazitaet_schnitte); chk(k.uf);
  k.summe = dalloc<V3<double>>(c, k.kapazitaet_schnitte); chk(k.summe);
  k.tausch = dalloc<uint32_t>(c, 2 * (size_t)k.kapazitaet_tausch); chk(k.tausch);
  k.stoesse = dalloc<Stoss>(c, (size_t)nt + ne); chk(k.stoesse);
  k.fe_punkte = dalloc<uint32_t>(c, 5 * (size_t)k.kapazitaet_federn); chk(k.fe_punkte);
  k.fe_gew = dalloc<double>(c, 5 * (size_t)k.kapazitaet_federn); chk(k.fe_gew);
  k.fe_kraft = dalloc<V3<double>>(c, k.kapazitaet_federn); chk(k.fe_kraft);
  k.fe_block = dalloc<double>(c, 9 * (size_t)k.kapazitaet_federn); chk(k.fe_block);
  k.fe_off = dalloc<uint32_t>(c, (size_t)n + 1); chk(k.fe_off);
  k.fe_ent = dalloc<uint32_t>(c, k.kapazitaet_fe_ent); chk(k.fe_ent);
  k.fk_off = dalloc<uint32_t>(c, (size_t)n + 1); chk(k.fk_off);
  k.fk_ent = dalloc<uint32_t>(c, k.kapazitaet_fk_ent); chk(k.fk_ent);
  k.fe_roh = dalloc<uint32_t>(c, k.kapazitaet_fe_ent); chk(k.fe_roh);
  k.fk_roh = dalloc<uint32_t>(c, k.kapazitaet_fk_ent); chk(k.fk_roh);
  k.st_off = dalloc<uint32_t>(c, (size_t)n + 1); chk(k.st_off);
  k.st_ent = dalloc<uint32_t>(c, k.kapazitaet_st_ent); chk(k.st_ent);
  k.st_dx = dalloc<V3<double>>(c, 5 * ((size_t)nt + ne)); chk(k.st_dx);
  k.zaehler = dalloc<uint32_t>(c, SK_Z_ANZAHL); chk(k.zaehler);
  c->sk_zellen = dalloc<double>(c, 3); chk(c->sk_zellen);
  c->sk_fe_fuell = dalloc<uint32_t>(c, (size_t)n + 1); chk(c->sk_fe_fuell);
  c->sk_fk_fuell = dalloc<uint32_t>(c, (size_t)n + 1); chk(c->sk_fk_fuell);
  c->sk_st_fuell = dalloc<uint32_t>(c, (size_t)n + 1); chk(c->sk_st_fuell);
  c->sk_scan_max = std::max<uint32_t>(c->sk_eimer_max, std::max<uint32_t>((uint32_t)n, std::max<uint32_t>(nt, ne)));
  c->sk_scan_teil = dalloc<uint32_t>(c, c->sk_scan_max / SK_SCAN_CHUNK + 1); chk(c->sk_scan_teil);
  c->sk_block_t = dalloc<T>(c, 9 * (size_t)k.kapazitaet_federn); chk(c->sk_block_t);
  c->sk_fe_t = dalloc<V3<T>>(c, k.kapazitaet_fe_ent); chk(c->sk_fe_t);
  c->sk_fe_satz = dalloc<FederK<T>>(c, k.kapazitaet_fe_ent); chk(c->sk_fe_satz);
  c->sk_dev = dalloc<SelbstDev>(c, 1); chk(c->sk_dev);
  c->sk_allocs.insert(c->sk_allocs.end(), c->allocs.begin() + (long)a0, c->allocs.end());
  c->allocs.resize(a0);
  if (!ok) return fehler(c, "hipMalloc/hipMemcpy failed (self-collision buffers)");
  HIPCK(c, hipMemcpyAsync(c->sk_dev, &k, sizeof(SelbstDev), hipMemcpyHostToDevice, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  sy.selbst = 1;
  sy.fe_off = k.fe_off;
  sy.fe_ent = k.fe_ent;
  sy.fk_off = k.fk_off;
  sy.fk_ent = k.fk_ent;
  sy.fe_punkte = k.fe_punkte;
  sy.fe_gew = k.fe_gew;
  sy.fe_kraft = k.fe_kraft;
  sy.fe_block = (const T*)c->sk_block_t;
  sy.fe_t = (V3<T>*)c->sk_fe_t;
  sy.fe_satz = (FederK<T>*)c->sk_fe_satz;
  c->sk_bereit = true;
  c->sk_dicke = dicke;
  return 0;
}

static unsigned sk_gitter(long count) {
  long b = (count + BUILD_THREADS - 1) / BUILD_THREADS;
  return (unsigned)std::max(1L, std::min(b, 4096L));
}
// blocks of BUILD_THREADS for kernels that give each item a whole wave
static unsigned sk_gitter_wellen(long count) {
  long b = (count * 64 + BUILD_THREADS - 1) / BUILD_THREADS;
  return (unsigned)std::max(1L, std::min(b, 4096L));
}

// Exclusive scan of werte[0..anzahl) in place, total into werte[anzahl]
// (and into kopie[0..anzahl] and *gesamt_ziel when given), over
// ceil(anzahl / SK_SCAN_CHUNK) blocks.
static void sk_scan(gc_ctx* c, uint32_t* werte, int anzahl, uint32_t* kopie, uint32_t* gesamt_ziel) {
  hipStream_t st = c->stream;
  const int teile = std::max(1, (anzahl + SK_SCAN_CHUNK - 1) / SK_SCAN_CHUNK);
  hipLaunchKernelGGL(k_sk_scan_teil, dim3(teile), dim3(1024), 0, st, (const uint32_t*)werte, anzahl, c->sk_scan_teil);
  hipLaunchKernelGGL(k_sk_scan_teile, dim3(1), dim3(1024), 0, st, c->sk_scan_teil, teile, werte, anzahl, kopie, gesamt_ziel);
  hipLaunchKernelGGL(k_sk_scan_fertig, dim3(teile), dim3(1024), 0, st, werte, anzahl, (const uint32_t*)c->sk_scan_teil, kopie);
}

// One spatial hash over c->sk.quader[0..anzahl) with the cell at *zelle_p.
static void sk_hash(gc_ctx* c, int anzahl, const double* zelle_p, uint32_t eimer) {
  hipStream_t st = c->stream;
  SelbstDev* sp = c->sk_dev;
  hipLaunchKernelGGL(k_sk_null, dim3(sk_gitter(eimer + 1)), dim3(BUILD_THREADS), 0, st, c->sk.eimer_fuell, (int)eimer + 1);
  hipLaunchKernelGGL(k_sk_null, dim3(1), dim3(64), 0, st, c->sk.zaehler + SK_Z_GROSS, 1);
  hipLaunchKernelGGL(k_sk_eimer_zaehlen, dim3(sk_gitter(anzahl)), dim3(BUILD_THREADS), 0, st, sp, anzahl, zelle_p, eimer);
  hipLaunchKernelGGL(k_sk_gross, dim3(1), dim3(1024), 0, st, sp, zelle_p, eimer, 0);
  sk_scan(c, c->sk.eimer_fuell, (int)eimer, c->sk.eimer_anfang, c->sk.zaehler + SK_Z_EINTRAEGE);
  hipLaunchKernelGGL(k_sk_eimer_fuellen, dim3(sk_gitter(anzahl)), dim3(BUILD_THREADS), 0, st, sp, anzahl, zelle_p, eimer);
  hipLaunchKernelGGL(k_sk_gross, dim3(1), dim3(1024), 0, st, sp, zelle_p, eimer, 1);
  hipLaunchKernelGGL(k_sk_eimer_sortieren, dim3(sk_gitter(c->sk.kapazitaet_eintraege)), dim3(BUILD_THREADS), 0, st, sp, eimer);
}

// Before the build (implizit.rs:668-707): pairs, springs and damping, cuts,
// regions, push springs, the per-point spring lists.
template <typename T>
static int launch_selbst_bau(gc_ctx* c) {
  hipStream_t st = c->stream;
  SelbstDev* sp = c->sk_dev;
  const SelbstDev& k = c->sk;
  const SysDev<T>* syp = dsys<T>(c);
  hipLaunchKernelGGL(k_sk_null, dim3(1), dim3(64), 0, st, k.zaehler, (int)SK_Z_ANZAHL);
  hipLaunchKernelGGL(k_sk_zelle_naehe, dim3(1), dim3(64), 0, st, sp, c->sk_zellen);
  hipLaunchKernelGGL(k_sk_normalen_dreieck, dim3(sk_gitter(k.nt)), dim3(BUILD_THREADS), 0, st, sp);
  hipLaunchKernelGGL(k_sk_normalen_kante, dim3(sk_gitter(k.ne)), dim3(BUILD_THREADS), 0, st, sp);
  // point-triangle (:400-467)
  hipLaunchKernelGGL(k_sk_quader_punkte, dim3(sk_gitter(k.n)), dim3(BUILD_THREADS), 0, st, sp);
  sk_hash(c, k.n, c->sk_zellen, c->sk_eimer_p);
  hipLaunchKernelGGL(k_sk_pd, dim3(sk_gitter_wellen(k.nt)), dim3(BUILD_THREADS), 0, st, sp, c->sk_eimer_p, 0);
  sk_scan(c, k.zahl, k.nt, nullptr, k.zaehler + SK_Z_NAEHE_PD);
  hipLaunchKernelGGL(k_sk_pd, dim3(sk_gitter_wellen(k.nt)), dim3(BUILD_THREADS), 0, st, sp, c->sk_eimer_p, 1);
  // edge-edge (:506-592)
  hipLaunchKernelGGL(k_sk_quader_kanten, dim3(sk_gitter(k.ne)), dim3(BUILD_THREADS), 0, st, sp);
  sk_hash(c, k.ne, c->sk_zellen, c->sk_eimer_k);
  hipLaunchKernelGGL(k_sk_kk, dim3(sk_gitter_wellen(k.ne)), dim3(BUILD_THREADS), 0, st, sp, c->sk_eimer_k, 0);
  sk_scan(c, k.zahl, k.ne, nullptr, k.zaehler + SK_Z_NAEHE_KK);
  hipLaunchKernelGGL(k_sk_kk, dim3(sk_gitter_wellen(k.ne)), dim3(BUILD_THREADS), 0, st, sp, c->sk_eimer_k, 1);
  hipLaunchKernelGGL(k_sk_naehe_summe, dim3(1), dim3(64), 0, st, sp);
  // springs and the damping of the approach (:713-787)
  hipLaunchKernelGGL(k_sk_federn, dim3(sk_gitter(k.kapazitaet_naehe)), dim3(BUILD_THREADS), 0, st, sp);
  hipLaunchKernelGGL(k_sk_daempfung, dim3(1), dim3(1024), 0, st, sp);
  // cuts, regions, push springs (:858-1059)
  hipLaunchKernelGGL(k_sk_quader_dreiecke, dim3(sk_gitter(k.nt)), dim3(BUILD_THREADS), 0, st, sp);
  hipLaunchKernelGGL(k_sk_laengste_start, dim3(1), dim3(64), 0, st, 1e-3, c->sk_zellen + 1);
  hipLaunchKernelGGL(k_sk_laengste, dim3(sk_gitter(k.ne)), dim3(BUILD_THREADS), 0, st, sp, (const V3<double>*)k.x, c->sk_zellen + 1);
  sk_hash(c, k.nt, c->sk_zellen + 1, c->sk_eimer_d);
  hipLaunchKernelGGL(k_sk_schnitte, dim3(sk_gitter_wellen(k.ne)), dim3(BUILD_THREADS), 0, st, sp, (const double*)(c->sk_zellen + 1),
                     c->sk_eimer_d, 0);
  sk_scan(c, k.zahl, k.ne, nullptr, k.zaehler + SK_Z_SCHNITTE);
  hipLaunchKernelGGL(k_sk_schnitte, dim3(sk_gitter_wellen(k.ne)), dim3(BUILD_THREADS), 0, st, sp, (const double*)(c->sk_zellen + 1),
                     c->sk_eimer_d, 1);
  hipLaunchKernelGGL(k_sk_schnitte_summe, dim3(1), dim3(64), 0, st, sp);
  hipLaunchKernelGGL(k_sk_gebiete, dim3(1), dim3(64), 0, st, sp);
  hipLaunchKernelGGL(k_sk_schub<T>, dim3(1), dim3(64), 0, st, sp, syp);
  hipLaunchKernelGGL(k_sk_folge, dim3(sk_gitter(k.nt)), dim3(BUILD_THREADS), 0, st, sp);
  // per-point spring lists
  hipLaunchKernelGGL(k_sk_null, dim3(sk_gitter(k.n + 1)), dim3(BUILD_THREADS), 0, st, k.fe_off, k.n + 1);
  hipLaunchKernelGGL(k_sk_null, dim3(sk_gitter(k.n + 1)), dim3(BUILD_THREADS), 0, st, k.fk_off, k.n + 1);
  hipLaunchKernelGGL(k_sk_feder_zaehlen, dim3(sk_gitter(k.kapazitaet_federn)), dim3(BUILD_THREADS), 0, st, sp);
  sk_scan(c, k.fe_off, k.n, c->sk_fe_fuell, nullptr);
  sk_scan(c, k.fk_off, k.n, c->sk_fk_fuell, nullptr);
  hipLaunchKernelGGL(k_sk_feder_pruefen, dim3(1), dim3(64), 0, st, sp);
  hipLaunchKernelGGL(k_sk_feder_fuellen, dim3(sk_gitter(k.kapazitaet_federn)), dim3(BUILD_THREADS), 0, st, sp, c->sk_fe_fuell,
                     c->sk_fk_fuell);
  hipLaunchKernelGGL(k_sk_feder_sortieren, dim3(sk_gitter(k.n)), dim3(BUILD_THREADS), 0, st, sp);
  hipLaunchKernelGGL(k_sk_feder_leeren, dim3(sk_gitter(k.n + 1)), dim3(BUILD_THREADS), 0, st, sp);
  hipLaunchKernelGGL(k_sk_feder_bloecke<T>, dim3(sk_gitter(9L * k.kapazitaet_federn)), dim3(BUILD_THREADS), 0, st, sp,
                     (T*)c->sk_block_t);
  hipLaunchKernelGGL(k_feder_saetze<T>, dim3(sk_gitter(k.kapazitaet_fe_ent)), dim3(BUILD_THREADS), 0, st, dsys<T>(c));
  hipLaunchKernelGGL(k_sk_statistik, dim3(1), dim3(64), 0, st, sp, c->gs, 0);
  HIPCK(c, hipGetLastError());
  return 0;
}

// After the update (implizit.rs:1030-1037): impacts on the path of the step.
template <typename T>
static int launch_selbst_stoss(gc_ctx* c) {
  hipStream_t st = c->stream;
  SelbstDev* sp = c->sk_dev;
  const SelbstDev& k = c->sk;
  const SysDev<T>* syp = dsys<T>(c);
  hipLaunchKernelGGL(k_sk_null, dim3(1), dim3(64), 0, st, k.zaehler + SK_Z_STOSS_ZAHL, 1);
  hipLaunchKernelGGL(k_sk_laengste_start, dim3(1), dim3(64), 0, st, 1e-3, c->sk_zellen + 2);
  hipLaunchKernelGGL(k_sk_laengste, dim3(sk_gitter(k.ne)), dim3(BUILD_THREADS), 0, st, sp, (const V3<double>*)k.x, c->sk_zellen + 2);
  hipLaunchKernelGGL(k_sk_quader_weg_punkte, dim3(sk_gitter(k.n)), dim3(BUILD_THREADS), 0, st, sp);
  sk_hash(c, k.n, c->sk_zellen + 2, c->sk_eimer_p);
  hipLaunchKernelGGL(k_sk_stoss_pd, dim3(sk_gitter_wellen(k.nt)), dim3(BUILD_THREADS), 0, st, sp, (const double*)(c->sk_zellen + 2),
                     c->sk_eimer_p);
  hipLaunchKernelGGL(k_sk_quader_weg_kanten, dim3(sk_gitter(k.ne)), dim3(BUILD_THREADS), 0, st, sp);
  sk_hash(c, k.ne, c->sk_zellen + 2, c->sk_eimer_k);
  hipLaunchKernelGGL(k_sk_stoss_kk, dim3(sk_gitter_wellen(k.ne)), dim3(BUILD_THREADS), 0, st, sp, (const double*)(c->sk_zellen + 2),
                     c->sk_eimer_k);
  hipLaunchKernelGGL(k_sk_stoss_filter, dim3(sk_gitter(k.nt + k.ne)), dim3(BUILD_THREADS), 0, st, sp);
  hipLaunchKernelGGL(k_sk_null, dim3(sk_gitter(k.n + 1)), dim3(BUILD_THREADS), 0, st, k.st_off, k.n + 1);
  hipLaunchKernelGGL(k_sk_stoss_zaehlen, dim3(sk_gitter(k.nt + k.ne)), dim3(BUILD_THREADS), 0, st, sp);
  sk_scan(c, k.st_off, k.n, c->sk_st_fuell, nullptr);
  hipLaunchKernelGGL(k_sk_stoss_fuellen, dim3(sk_gitter(k.nt + k.ne)), dim3(BUILD_THREADS), 0, st, sp, c->sk_st_fuell);
  for (int it = 0; it < k.stoss_iterationen; ++it) {
    hipLaunchKernelGGL(k_sk_stoss_beitrag<T>, dim3(sk_gitter(k.nt + k.ne)), dim3(BUILD_THREADS), 0, st, sp, syp);
    hipLaunchKernelGGL(k_sk_stoss_anwenden<T>, dim3(sk_gitter(k.n)), dim3(BUILD_THREADS), 0, st, sp, syp);
  }
  hipLaunchKernelGGL(k_sk_statistik, dim3(1), dim3(64), 0, st, sp, c->gs, 1);
  HIPCK(c, hipGetLastError());
  return 0;
}

// ---------------------------------------------------------------------------
// launch sequences
// ---------------------------------------------------------------------------
template <typename T>
static int launch_bau(gc_ctx* c) {
  const SysDev<T>& s = sys<T>(c);
  const SysDev<T>* sp = dsys<T>(c);
  hipStream_t st = c->stream;
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(256), 0, st, c->gs, c->slots);
  if (s.selbst && launch_selbst_bau<T>(c)) return 1;
  // FP32 bulk: the FP64 copy's records, right-hand side and mass diagonal
  // first (the shared FP64 outputs, forces and loads, are the same in both
  // builds, and the bulk build below writes them last)
  const SysDev<double>* sdp = (const SysDev<double>*)c->schatten_dev;
  if (sizeof(T) == 4 && sdp) {
    if (s.nk) hipLaunchKernelGGL(k_bau_kanten<double>, dim3(gitter(s.nk)), dim3(BUILD_THREADS), 0, st, sdp);
    if (s.nb) hipLaunchKernelGGL(k_bau_biegung<double>, dim3(gitter(s.nb)), dim3(BUILD_THREADS), 0, st, sdp);
    if (s.nm) hipLaunchKernelGGL(k_bau_membran<double>, dim3(gitter(s.nm)), dim3(BUILD_THREADS), 0, st, sdp);
    hipLaunchKernelGGL(k_sammeln<double>, dim3(gitter(s.n)), dim3(BUILD_THREADS), 0, st, sdp);
  }
  if (s.nk) hipLaunchKernelGGL(k_bau_kanten<T>, dim3(gitter(s.nk)), dim3(BUILD_THREADS), 0, st, sp);
  if (s.nb) hipLaunchKernelGGL(k_bau_biegung<T>, dim3(gitter(s.nb)), dim3(BUILD_THREADS), 0, st, sp);
  if (s.nm) hipLaunchKernelGGL(k_bau_membran<T>, dim3(gitter(s.nm)), dim3(BUILD_THREADS), 0, st, sp);
  hipLaunchKernelGGL(k_fest<T>, dim3(gitter(s.n)), dim3(BUILD_THREADS), 0, st, sp, c->gs);
  hipLaunchKernelGGL(k_sammeln<T>, dim3(gitter(s.n)), dim3(BUILD_THREADS), 0, st, sp);
  hipLaunchKernelGGL(k_kachel_saetze<T>, dim3(gitter(c->kachel_eintraege)), dim3(BUILD_THREADS), 0, st, sp);
  HIPCK(c, hipGetLastError());
  return 0;
}

// Rounds of FP32 iterations of a refined solve: its FP64 start, then up to
// GC_RUNDEN rounds each checked in FP64, then the FP64 iterations of a solve
// handed over and the solve's statistics. A launch after the solve is done
// returns at once, so the sequence is fixed (and capturable). The launches
// share one grid and carry the barrier epoch from one to the next
// (GridState::zaehler), so the counters, cleared before the first, are not
// cleared between them.
#define GC_RUNDEN 4
static void launch_verfeinert(hipStream_t st, int bloecke, int threads, size_t lds, const SysDev<float>* sp, GridState* gs,
                              Slot* slots, int durchgang) {
  hipLaunchKernelGGL(k_verf, dim3(bloecke), dim3(threads), 0, st, sp, gs, slots, durchgang, 0, 0);
  for (int r = 0; r < GC_RUNDEN; ++r) {
    hipLaunchKernelGGL(k_pcg<float>, dim3(bloecke), dim3(threads), lds, st, sp, gs, slots, durchgang);
    hipLaunchKernelGGL(k_verf, dim3(bloecke), dim3(threads), 0, st, sp, gs, slots, durchgang, 1, r == GC_RUNDEN - 1 ? 1 : 0);
  }
  hipLaunchKernelGGL(k_pcg64, dim3(bloecke), dim3(threads), 0, st, sp, gs, slots, durchgang);
}

template <typename T>
static int launch_pcg(gc_ctx* c, int durchgang) {
  if (sizeof(T) == 4 && c->schatten_dev) {
    launch_verfeinert(c->stream, c->pcg_blocks, c->pcg_threads, c->pcg_lds, (const SysDev<float>*)dsys<T>(c), c->gs, c->slots,
                      durchgang);
  } else {
    hipLaunchKernelGGL(k_pcg<T>, dim3(c->pcg_blocks), dim3(c->pcg_threads), c->pcg_lds, c->stream, dsys<T>(c), c->gs,
                       c->slots, durchgang);
  }
  HIPCK(c, hipGetLastError());
  return 0;
}

// A substep in three parts around its two PCG launches, so that a batch can
// put the solves of many systems into one launch: before (positions at the
// start, self-collision, build, warm start), between (release pass 0), after
// (release pass 1, friction and update, impacts).
template <typename T>
static int launch_vor_pcg(gc_ctx* c, int hat_vorher) {
  const SysDev<T>& s = sys<T>(c);
  if (s.selbst) hipLaunchKernelGGL(k_sk_kopie_x, dim3(sk_gitter(s.n)), dim3(BUILD_THREADS), 0, c->stream, c->sk_dev);
  if (launch_bau<T>(c)) return 1;
  hipLaunchKernelGGL(k_warm<T>, dim3(gitter(s.n)), dim3(BUILD_THREADS), 0, c->stream, dsys<T>(c), hat_vorher);
  HIPCK(c, hipGetLastError());
  return 0;
}

template <typename T>
static int launch_zwischen_pcg(gc_ctx* c) {
  hipLaunchKernelGGL(k_freigabe<T>, dim3(gitter(sys<T>(c).n)), dim3(BUILD_THREADS), 0, c->stream, dsys<T>(c), c->gs, c->slots, 0);
  HIPCK(c, hipGetLastError());
  return 0;
}

// Body points against cloth triangles after the update (implizit.rs:1017-1026):
// per pass and mesh body in order, the deepest stab per triangle, then the
// lifts level by level in one block over the stabbed triangles.
template <typename T>
static int launch_koerperpunkte(gc_ctx* c) {
  if (!c->kp_bereit || c->param.koerperpunkt_durchgaenge <= 0) return 0;
  std::vector<int> netze;
  for (size_t b = 0; b < c->koerper.size(); ++b)
    if (c->koerper[b].netz) netze.push_back((int)b);
  if (netze.empty()) return 0;
  const SysDev<T>* sp = dsys<T>(c);
  for (int d = 0; d < c->param.koerperpunkt_durchgaenge; ++d)
    for (size_t q = 0; q < netze.size(); ++q) {
      hipLaunchKernelGGL(k_kp_stiche<T>, dim3(gitter(c->kp_nt)), dim3(BUILD_THREADS), 0, c->stream, sp, netze[q], c->gs, d,
                         q == 0 ? 1 : 0, c->param.koerperpunkt_tiefe, (const uint32_t*)c->kp_dreiecke, c->kp_nt, c->kp_stich,
                         c->kp_liste, c->kp_zahl);
      hipLaunchKernelGGL(k_kp_heben<T>, dim3(1), dim3(1024), 0, c->stream, sp, netze[q], c->gs, d, c->kp_nstufen,
                         (const int*)c->kp_stufe, (const uint32_t*)c->kp_dreiecke, (const Stich*)c->kp_stich,
                         (const uint32_t*)c->kp_liste, c->kp_zahl);
    }
  HIPCK(c, hipGetLastError());
  return 0;
}

template <typename T>
static int launch_nach_pcg(gc_ctx* c) {
  const SysDev<T>& s = sys<T>(c);
  const SysDev<T>* sp = dsys<T>(c);
  hipLaunchKernelGGL(k_freigabe<T>, dim3(gitter(s.n)), dim3(BUILD_THREADS), 0, c->stream, sp, c->gs, c->slots, 1);
  hipLaunchKernelGGL(k_update<T>, dim3(gitter(s.n)), dim3(BUILD_THREADS), 0, c->stream, sp, c->gs);
  if (launch_koerperpunkte<T>(c)) return 1;
  if (s.selbst && launch_selbst_stoss<T>(c)) return 1;
  HIPCK(c, hipGetLastError());
  return 0;
}

// Solve k (0, 1) of a substep, or its recorded result instead (c->wiedergabe);
// with c->aufnahme the result is kept. The solution is dv, and for a refined
// FP32 solve also its FP64 solution dv64, which the release pass and the
// update read.
template <typename T>
static int loesung(gc_ctx* c, int k) {
  const SysDev<T>& s = sys<T>(c);
  const size_t n = (size_t)s.n;
  if (c->wiedergabe) {
    HIPCK(c, hipMemcpyAsync(s.dv, c->wiedergabe->dv_loes[k], n * sizeof(V3<T>), hipMemcpyDeviceToDevice, c->stream));
    if (s.dv64)
      HIPCK(c, hipMemcpyAsync(s.dv64, c->wiedergabe->dv64_loes[k], n * sizeof(V3<double>), hipMemcpyDeviceToDevice, c->stream));
  } else if (launch_pcg<T>(c, k)) {
    return 1;
  }
  if (c->aufnahme) {
    HIPCK(c, hipMemcpyAsync(c->aufnahme->dv_loes[k], s.dv, n * sizeof(V3<T>), hipMemcpyDeviceToDevice, c->stream));
    if (s.dv64)
      HIPCK(c, hipMemcpyAsync(c->aufnahme->dv64_loes[k], s.dv64, n * sizeof(V3<double>), hipMemcpyDeviceToDevice, c->stream));
  }
  return 0;
}

template <typename T>
// ev: four events (before/after each PCG launch) or nullptr
static int launch_schritt(gc_ctx* c, int hat_vorher, const hipEvent_t* ev) {
  hipStream_t st = c->stream;
  if (launch_vor_pcg<T>(c, hat_vorher)) return 1;
  if (ev) HIPCK(c, hipEventRecord(ev[0], st));
  if (loesung<T>(c, 0)) return 1;
  if (ev) HIPCK(c, hipEventRecord(ev[1], st));
  if (launch_zwischen_pcg<T>(c)) return 1;
  if (ev) HIPCK(c, hipEventRecord(ev[2], st));
  if (loesung<T>(c, 1)) return 1;
  if (ev) HIPCK(c, hipEventRecord(ev[3], st));
  return launch_nach_pcg<T>(c);
}

static V3<double>* zustand_x(gc_ctx* c) { return c->opt.praezision == GC_FP32 ? c->sf.x : c->sd.x; }
static V3<double>* zustand_v(gc_ctx* c) { return c->opt.praezision == GC_FP32 ? c->sf.v : c->sd.v; }
static V3<double>* zustand_dvv(gc_ctx* c) { return c->opt.praezision == GC_FP32 ? c->sf.dv_vorher : c->sd.dv_vorher; }

// Saves x, v, the warm start and the self-collision push-spring sequence of
// the context's state into z (device copies, allocated on first use).
static int zustand_sichern_in(gc_ctx* c, gc_ctx::Zustand& z) {
  const size_t n = (size_t)c->n;
  if (!z.x) {
    z.x = dalloc<V3<double>>(c, n);
    z.v = dalloc<V3<double>>(c, n);
    z.dvv = dalloc<V3<double>>(c, n);
    if (!z.x || !z.v || !z.dvv) return fehler(c, "hipMalloc failed (saved state)");
  }
  const size_t bytes = n * sizeof(V3<double>);
  HIPCK(c, hipMemcpyAsync(z.x, zustand_x(c), bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCK(c, hipMemcpyAsync(z.v, zustand_v(c), bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCK(c, hipMemcpyAsync(z.dvv, zustand_dvv(c), bytes, hipMemcpyDeviceToDevice, c->stream));
  if (c->sk_bereit) {
    if (!z.folge) z.folge = dalloc<uint32_t>(c, (size_t)c->sk.nt);
    if (!z.folge) return fehler(c, "hipMalloc failed (saved state)");
    HIPCK(c, hipMemcpyAsync(z.folge, c->sk.folge, (size_t)c->sk.nt * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
  }
  z.hat_vorher = c->hat_vorher;
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

static int zustand_laden_aus(gc_ctx* c, const gc_ctx::Zustand& z) {
  const size_t bytes = (size_t)c->n * sizeof(V3<double>);
  HIPCK(c, hipMemcpyAsync(zustand_x(c), z.x, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCK(c, hipMemcpyAsync(zustand_v(c), z.v, bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCK(c, hipMemcpyAsync(zustand_dvv(c), z.dvv, bytes, hipMemcpyDeviceToDevice, c->stream));
  if (c->sk_bereit && z.folge)
    HIPCK(c, hipMemcpyAsync(c->sk.folge, z.folge, (size_t)c->sk.nt * sizeof(uint32_t), hipMemcpyDeviceToDevice, c->stream));
  c->hat_vorher = z.hat_vorher;
  c->built = false;
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

// The state and the step counters at the start of a stepping call.
static int anfang_sichern(gc_ctx* c) {
  if (zustand_sichern_in(c, c->anfang)) return 1;
  HIPCK(c, hipMemcpy(&c->anfang_gs, c->gs, sizeof(GridState), hipMemcpyDeviceToHost));
  return 0;
}

// After a self-collision list overflowed: free the tables so that the next
// selbst_vorbereiten builds them with twice the capacities, and put the
// state and the counters back to the start of the call (the push-spring
// sequence is restored by anfang_zurueck once the tables exist again).
static int selbst_wachsen(gc_ctx* c) {
  if (c->sk_mal >= 16)
    return fehler(c, "a self-collision list exceeded its capacity at 16 times the default sizes (state of the step is "
                     "not the reference's)");
  HIPCK(c, hipStreamSynchronize(c->stream));
  for (void* p : c->sk_allocs) (void)hipFree(p);
  c->sk_allocs.clear();
  c->sk_bereit = false;
  c->sk_dev = nullptr;
  c->sk_mal *= 2;
  folge_neu(c);
  return 0;
}

static int anfang_zurueck(gc_ctx* c) {
  if (zustand_laden_aus(c, c->anfang)) return 1;
  GridState g = c->anfang_gs;
  g.fehler = 0;
  g.selbst_ueberlauf = 0;
  HIPCK(c, hipMemcpy(c->gs, &g, sizeof(GridState), hipMemcpyHostToDevice));
  return 0;
}

// The message of a grid all-reduce's error flag (GridState::fehler).
static std::string barriere_meldung(unsigned code) {
  if (code == 2)
    return "grid all-reduce found a partial of a later barrier (slots left from another launch); the solve's "
           "result is not valid";
  return "grid barrier wait exceeded its spin limit (blocks not co-resident?)";
}

// ueberlauf: where given, a self-collision list overflow is reported there
// (and cleared) instead of as an error.
static int pruefe_gitter(gc_ctx* c, bool* ueberlauf = nullptr) {
  GridState g;
  HIPCK(c, hipMemcpyAsync(&g, c->gs, sizeof(GridState), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  if (g.fehler) {
    unsigned null = 0;
    HIPCK(c, hipMemcpy(&c->gs->fehler, &null, sizeof(unsigned), hipMemcpyHostToDevice));
    return fehler(c, barriere_meldung(g.fehler));
  }
  if (g.selbst_ueberlauf && ueberlauf) {
    unsigned null = 0;
    HIPCK(c, hipMemcpy(&c->gs->selbst_ueberlauf, &null, sizeof(unsigned), hipMemcpyHostToDevice));
    *ueberlauf = true;
    return 0;
  }
  if (g.selbst_ueberlauf) {
    unsigned null = 0;
    HIPCK(c, hipMemcpy(&c->gs->selbst_ueberlauf, &null, sizeof(unsigned), hipMemcpyHostToDevice));
    return fehler(c, "a self-collision list exceeded its capacity (state of the step is not the reference's)");
  }
  return 0;
}

template <typename T>
static int schritte(gc_ctx* c, double h, int n) {
  if (!c->topo || !c->hat_param) return fehler(c, "gc_step: topology and parameters must be set first");
  if ((int)c->ev.size() < 4 * n + 2) {
    while ((int)c->ev.size() < 4 * n + 2) {
      hipEvent_t e;
      HIPCK(c, hipEventCreate(&e));
      c->ev.push_back(e);
    }
  }
  sys<T>(c).h = h;
  sys<T>(c).h2 = h * h;
  if (selbst_vorbereiten<T>(c)) return 1;
  if (sys_hoch<T>(c)) return 1;
  // with self-collision on, a call whose lists overflow is taken again from
  // its start with larger lists
  const bool sichern = sys<T>(c).selbst != 0;
  if (sichern && anfang_sichern(c)) return 1;
  for (;;) {
    HIPCK(c, hipEventRecord(c->ev[0], c->stream));
    for (int k = 0; k < n; ++k) {
      if (launch_schritt<T>(c, c->hat_vorher ? 1 : 0, &c->ev[2 + 4 * k])) return 1;
      c->hat_vorher = true;
    }
    HIPCK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCK(c, hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->ms_schritt = ms;
    double pcg = 0.0;
    for (int k = 0; k < n; ++k) {
      float m0 = 0.f, m1 = 0.f;
      HIPCK(c, hipEventElapsedTime(&m0, c->ev[2 + 4 * k], c->ev[3 + 4 * k]));
      HIPCK(c, hipEventElapsedTime(&m1, c->ev[4 + 4 * k], c->ev[5 + 4 * k]));
      pcg += (double)m0 + (double)m1;
    }
    c->ms_pcg = pcg;
    c->built = false;
    bool ueberlauf = false;
    if (pruefe_gitter(c, sichern ? &ueberlauf : nullptr)) return 1;
    if (!ueberlauf) return 0;
    if (selbst_wachsen(c) || selbst_vorbereiten<T>(c) || sys_hoch<T>(c) || anfang_zurueck(c)) return 1;
  }
}

template <typename T>
static int schritte_graph(gc_ctx* c, double h, int n) {
  if (!c->topo || !c->hat_param) return fehler(c, "gc_step_graph: topology and parameters must be set first");
  if (n <= 0) return 0;
  int done = 0;
  if (!c->hat_vorher) {  // the first substep starts from h*g; the captured one from dv_vorher
    if (schritte<T>(c, h, 1)) return 1;
    done = 1;
    if (n == 1) return 0;
  }
  sys<T>(c).h = h;
  sys<T>(c).h2 = h * h;
  if (selbst_vorbereiten<T>(c)) return 1;
  if (sys_hoch<T>(c)) return 1;
  const bool sichern = sys<T>(c).selbst != 0;
  if (sichern && anfang_sichern(c)) return 1;
  if (c->ev.size() < 2) {
    while (c->ev.size() < 2) {
      hipEvent_t e;
      HIPCK(c, hipEventCreate(&e));
      c->ev.push_back(e);
    }
  }
  for (;;) {
    if (!c->graph) {
      hipGraph_t g;
      HIPCK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
      int rc = launch_schritt<T>(c, 1, nullptr);
      hipError_t e2 = hipStreamEndCapture(c->stream, &g);
      if (rc) return 1;
      if (e2 != hipSuccess) return fehler(c, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
      hipError_t e3 = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e3 != hipSuccess) return fehler(c, std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
    }
    HIPCK(c, hipEventRecord(c->ev[0], c->stream));
    for (int k = done; k < n; ++k) HIPCK(c, hipGraphLaunch(c->graph, c->stream));
    HIPCK(c, hipEventRecord(c->ev[1], c->stream));
    HIPCK(c, hipEventSynchronize(c->ev[1]));
    float ms = 0.f;
    HIPCK(c, hipEventElapsedTime(&ms, c->ev[0], c->ev[1]));
    c->ms_schritt = ms;
    c->ms_pcg = std::nan("");
    c->built = false;
    bool ueberlauf = false;
    if (pruefe_gitter(c, sichern ? &ueberlauf : nullptr)) return 1;
    if (!ueberlauf) return 0;
    // the tables grow (the graph is dropped with them) and the replays are
    // taken again from the state after the eager first substep, if any
    if (selbst_wachsen(c) || selbst_vorbereiten<T>(c) || sys_hoch<T>(c) || anfang_zurueck(c)) return 1;
  }
}

// ---------------------------------------------------------------------------
// Batches: many systems, one persistent PCG launch per solve pass
// ---------------------------------------------------------------------------
struct gc_batch {
  std::vector<gc_ctx*> ctx;
  std::string err;
  hipStream_t stream = nullptr;
  int praezision = GC_FP32;
  int cus = 0;
  int threads = 1024;
  size_t lds = 0;
  int max_bloecke = 0;             // co-resident blocks of the batched kernel
  std::vector<double> gewicht;     // per system: tiles x iterations per solve (last measured)
  std::vector<int> bloecke;        // per system: blocks in the batched launch
  void* tab = nullptr;             // StapelBlock<T>[sum of bloecke]
  int n_tab = 0;
  hipGraphExec_t graph = nullptr;
  std::vector<uint64_t> folge;     // each system's launch-sequence generation at the graph's capture
  std::vector<hipEvent_t> ev;
  double ms_schritt = 0.0, ms_pcg = 0.0;
  // the systems' own kernels (build, release pass, update) run on several
  // branch streams, forked from and joined back into the batch's stream
  std::vector<hipStream_t> zweige;
  std::vector<hipEvent_t> gabel;   // per phase: fork event, then one join event per branch
};

static int bfehler(gc_batch* b, const std::string& m) {
  b->err = m;
  return 1;
}

#define BHIPCK(b, call)                                                                     \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      (b)->err = std::string(#call) + ": " + hipGetErrorString(e_);                         \
      return 1;                                                                             \
    }                                                                                       \
  } while (0)

// While a batch call runs, system k works on branch stream k mod branches.
struct StromTausch {
  gc_batch* b;
  std::vector<hipStream_t> alt;
  explicit StromTausch(gc_batch* bb) : b(bb) {
    for (size_t k = 0; k < b->ctx.size(); ++k) {
      alt.push_back(b->ctx[k]->stream);
      b->ctx[k]->stream = b->zweige[k % b->zweige.size()];
    }
  }
  ~StromTausch() {
    for (size_t k = 0; k < b->ctx.size(); ++k) b->ctx[k]->stream = alt[k];
  }
};

// Blocks per system in proportion to its weight (tiles x iterations), at least
// one and at most one per tile, summing to at most `gesamt`.
static void bloecke_verteilen(gc_batch* b, int gesamt) {
  const size_t n = b->ctx.size();
  double sw = 0.0;
  for (size_t k = 0; k < n; ++k) sw += b->gewicht[k];
  std::vector<int> nb(n);
  int summe = 0;
  for (size_t k = 0; k < n; ++k) {
    const int kacheln = std::max(1, b->ctx[k]->n_kacheln);
    int q = (int)std::floor((double)gesamt * b->gewicht[k] / std::max(sw, 1e-300));
    q = std::min(std::max(q, 1), std::min(kacheln, GC_MAX_BLOECKE));
    nb[k] = q;
    summe += q;
  }
  // hand out what is left to the systems with the most weight per block
  while (summe < gesamt) {
    int best = -1;
    double bw = 0.0;
    for (size_t k = 0; k < n; ++k) {
      if (nb[k] >= std::min(std::max(1, b->ctx[k]->n_kacheln), GC_MAX_BLOECKE)) continue;
      const double w = b->gewicht[k] / nb[k];
      if (w > bw) { bw = w; best = (int)k; }
    }
    if (best < 0) break;
    ++nb[best];
    ++summe;
  }
  while (summe > gesamt) {  // only when there are more systems than blocks allow
    int best = -1;
    double bw = 1e300;
    for (size_t k = 0; k < n; ++k) {
      if (nb[k] <= 1) continue;
      const double w = b->gewicht[k] / nb[k];
      if (w < bw) { bw = w; best = (int)k; }
    }
    if (best < 0) break;
    --nb[best];
    --summe;
  }
  b->bloecke = nb;
}

template <typename T>
static int tabelle_laden(gc_batch* b) {
  std::vector<StapelBlock<T>> tab;
  for (size_t k = 0; k < b->ctx.size(); ++k) {
    gc_ctx* c = b->ctx[k];
    for (int r = 0; r < b->bloecke[k]; ++r)
      tab.push_back(StapelBlock<T>{dsys<T>(c), c->gs, c->slots, (int32_t)r, (int32_t)b->bloecke[k]});
  }
  if ((int)tab.size() > b->max_bloecke) return bfehler(b, "batch needs more blocks than can be co-resident");
  if (b->tab) (void)hipFree(b->tab);
  b->tab = nullptr;
  BHIPCK(b, hipMalloc(&b->tab, tab.size() * sizeof(StapelBlock<T>)));
  BHIPCK(b, hipMemcpyAsync(b->tab, tab.data(), tab.size() * sizeof(StapelBlock<T>), hipMemcpyHostToDevice, b->stream));
  BHIPCK(b, hipStreamSynchronize(b->stream));
  b->n_tab = (int)tab.size();
  if (b->graph) {
    (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
  }
  return 0;
}

// LDS of a batched launch: the largest tile of its systems (slots and
// staging; no point lists). Each system keeps as many spring terms in LDS as
// fit beside its own tile in that size (never more than when it runs alone),
// so a batch gets the blocks per CU of its tiles. The settings reach the
// device with the next upload of each system.
template <typename T>
static size_t batch_lds(gc_batch* b) {
  size_t lds = 16;
  for (gc_ctx* c : b->ctx) lds = std::max(lds, c->pcg_lds_kachel);
  for (gc_ctx* c : b->ctx) {
    const size_t frei = (lds - c->pcg_lds_kachel) / sizeof(V3<T>);
    sys<T>(c).fe_lds_kap = (uint32_t)std::min<size_t>(frei, c->fe_kap_allein);
    sys<T>(c).punkte_lds = 0;
  }
  return lds;
}

template <typename T>
static int batch_form(gc_batch* b, int threads, int gesamt) {
  const size_t lds = batch_lds<T>(b);
  // the systems' spring counts must be on the device before any launch of
  // this shape (the autotuner launches right after)
  BHIPCK(b, hipStreamSynchronize(b->stream));
  for (gc_ctx* c : b->ctx)
    if (sys_hoch<T>(c)) return bfehler(b, c->err);
  BHIPCK(b, hipFuncSetAttribute((const void*)k_pcg_stapel<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int occ = 0;
  BHIPCK(b, hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_pcg_stapel<T>, threads, lds));
  if (occ < 1) return bfehler(b, "batched PCG kernel does not fit on a CU");
  b->threads = threads;
  b->lds = lds;
  b->max_bloecke = occ * b->cus;
  const int n = (int)b->ctx.size();
  if (n > b->max_bloecke) return bfehler(b, "more systems than co-resident blocks");
  bloecke_verteilen(b, std::max(n, std::min(gesamt > 0 ? gesamt : b->max_bloecke, b->max_bloecke)));
  return tabelle_laden<T>(b);
}

template <typename T>
static int batch_pcg(gc_batch* b, int durchgang) {
  if (sizeof(T) == 4) {  // refined solves in rounds, as launch_verfeinert
    const StapelBlock<float>* tab = (const StapelBlock<float>*)b->tab;
    hipLaunchKernelGGL(k_verf_stapel, dim3(b->n_tab), dim3(b->threads), 0, b->stream, tab, durchgang, 0, 0);
    for (int r = 0; r < GC_RUNDEN; ++r) {
      hipLaunchKernelGGL(k_pcg_stapel<T>, dim3(b->n_tab), dim3(b->threads), b->lds, b->stream, (const StapelBlock<T>*)b->tab,
                         durchgang);
      hipLaunchKernelGGL(k_verf_stapel, dim3(b->n_tab), dim3(b->threads), 0, b->stream, tab, durchgang, 1,
                         r == GC_RUNDEN - 1 ? 1 : 0);
    }
    hipLaunchKernelGGL(k_pcg64_stapel, dim3(b->n_tab), dim3(b->threads), 0, b->stream, tab, durchgang);
  } else {
    hipLaunchKernelGGL(k_pcg_stapel<T>, dim3(b->n_tab), dim3(b->threads), b->lds, b->stream, (const StapelBlock<T>*)b->tab,
                       durchgang);
  }
  BHIPCK(b, hipGetLastError());
  return 0;
}

// Branch streams wait for the batch stream (phase ph's fork event) ...
static int gabeln(gc_batch* b, int ph) {
  const size_t z = b->zweige.size();
  hipEvent_t* e = &b->gabel[(size_t)ph * (z + 1)];
  BHIPCK(b, hipEventRecord(e[0], b->stream));
  for (size_t q = 0; q < z; ++q) BHIPCK(b, hipStreamWaitEvent(b->zweige[q], e[0], 0));
  return 0;
}
// ... and the batch stream waits for every branch.
static int vereinen(gc_batch* b, int ph) {
  const size_t z = b->zweige.size();
  hipEvent_t* e = &b->gabel[(size_t)ph * (z + 1)];
  for (size_t q = 0; q < z; ++q) {
    BHIPCK(b, hipEventRecord(e[1 + q], b->zweige[q]));
    BHIPCK(b, hipStreamWaitEvent(b->stream, e[1 + q], 0));
  }
  return 0;
}

template <typename T>
static int batch_schritt(gc_batch* b, const hipEvent_t* ev) {
  if (gabeln(b, 0)) return 1;
  for (gc_ctx* c : b->ctx)
    if (launch_vor_pcg<T>(c, c->hat_vorher ? 1 : 0)) return bfehler(b, c->err);
  if (vereinen(b, 0)) return 1;
  if (ev) BHIPCK(b, hipEventRecord(ev[0], b->stream));
  if (batch_pcg<T>(b, 0)) return 1;
  if (ev) BHIPCK(b, hipEventRecord(ev[1], b->stream));
  if (gabeln(b, 1)) return 1;
  for (gc_ctx* c : b->ctx)
    if (launch_zwischen_pcg<T>(c)) return bfehler(b, c->err);
  if (vereinen(b, 1)) return 1;
  if (ev) BHIPCK(b, hipEventRecord(ev[2], b->stream));
  if (batch_pcg<T>(b, 1)) return 1;
  if (ev) BHIPCK(b, hipEventRecord(ev[3], b->stream));
  if (gabeln(b, 2)) return 1;
  for (gc_ctx* c : b->ctx)
    if (launch_nach_pcg<T>(c)) return bfehler(b, c->err);
  return vereinen(b, 2);
}

static int batch_sync(gc_batch* b) {
  for (hipStream_t z : b->zweige) BHIPCK(b, hipStreamSynchronize(z));
  BHIPCK(b, hipStreamSynchronize(b->stream));
  return 0;
}

template <typename T>
static int batch_vorbereiten(gc_batch* b, double h) {
  for (gc_ctx* c : b->ctx) {
    if (!c->topo || !c->hat_param) return bfehler(b, "every system of a batch needs topology and parameters");
    sys<T>(c).h = h;
    sys<T>(c).h2 = h * h;
    if (selbst_vorbereiten<T>(c) || sys_hoch<T>(c)) return bfehler(b, c->err);
  }
  return 0;
}

static int batch_ev(gc_batch* b, size_t anzahl) {
  while (b->ev.size() < anzahl) {
    hipEvent_t e;
    BHIPCK(b, hipEventCreate(&e));
    b->ev.push_back(e);
  }
  return 0;
}

// nochmal: where given, the systems whose self-collision lists overflowed get
// larger tables and *nochmal says that the call is to be taken again.
template <typename T>
static int batch_pruefen(gc_batch* b, bool* nochmal = nullptr) {
  if (batch_sync(b)) return 1;
  for (gc_ctx* c : b->ctx) {
    c->built = false;
    bool ueberlauf = false;
    if (pruefe_gitter(c, nochmal ? &ueberlauf : nullptr)) return bfehler(b, c->err);
    if (ueberlauf) {
      if (selbst_wachsen(c)) return bfehler(b, c->err);
      *nochmal = true;
    }
  }
  return 0;
}

// With any system's self-collision on, every system's state and counters at
// the start of a batch call are saved, so that the call can be taken again.
template <typename T>
static int batch_sichern(gc_batch* b, bool* sichern) {
  *sichern = false;
  for (gc_ctx* c : b->ctx) *sichern = *sichern || sys<T>(c).selbst != 0;
  if (*sichern)
    for (gc_ctx* c : b->ctx)
      if (anfang_sichern(c)) return bfehler(b, c->err);
  return 0;
}

template <typename T>
static int batch_zurueck(gc_batch* b, double h) {
  if (batch_vorbereiten<T>(b, h)) return 1;
  for (gc_ctx* c : b->ctx)
    if (anfang_zurueck(c)) return bfehler(b, c->err);
  return 0;
}

template <typename T>
static int batch_schritte(gc_batch* b, double h, int n) {
  StromTausch tausch(b);
  if (batch_vorbereiten<T>(b, h)) return 1;
  if (batch_ev(b, 4 * (size_t)n + 2)) return 1;
  bool sichern = false;
  if (batch_sichern<T>(b, &sichern)) return 1;
  for (;;) {
    BHIPCK(b, hipEventRecord(b->ev[0], b->stream));
    for (int k = 0; k < n; ++k) {
      if (batch_schritt<T>(b, &b->ev[2 + 4 * k])) return 1;
      for (gc_ctx* c : b->ctx) c->hat_vorher = true;
    }
    BHIPCK(b, hipEventRecord(b->ev[1], b->stream));
    BHIPCK(b, hipEventSynchronize(b->ev[1]));
    float ms = 0.f;
    BHIPCK(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->ms_schritt = ms;
    double pcg = 0.0;
    for (int k = 0; k < n; ++k) {
      float m0 = 0.f, m1 = 0.f;
      BHIPCK(b, hipEventElapsedTime(&m0, b->ev[2 + 4 * k], b->ev[3 + 4 * k]));
      BHIPCK(b, hipEventElapsedTime(&m1, b->ev[4 + 4 * k], b->ev[5 + 4 * k]));
      pcg += (double)m0 + (double)m1;
    }
    b->ms_pcg = pcg;
    bool nochmal = false;
    if (batch_pruefen<T>(b, sichern ? &nochmal : nullptr)) return 1;
    if (!nochmal) return 0;
    if (batch_zurueck<T>(b, h)) return 1;
  }
}

template <typename T>
static int batch_schritte_graph(gc_batch* b, double h, int n) {
  if (n <= 0) return 0;
  int done = 0;
  bool alle = true;
  for (gc_ctx* c : b->ctx) alle = alle && c->hat_vorher;
  if (!alle) {  // the first substep starts from h*g, the captured one from dv_vorher
    if (batch_schritte<T>(b, h, 1)) return 1;
    done = 1;
    if (n == 1) return 0;
  }
  StromTausch tausch(b);
  if (batch_vorbereiten<T>(b, h)) return 1;
  if (batch_ev(b, 2)) return 1;
  bool sichern = false;
  if (batch_sichern<T>(b, &sichern)) return 1;
  for (;;) {
    // a system's launch sequence changed since the capture (or its
    // self-collision tables grew): capture again
    bool gleich = b->folge.size() == b->ctx.size();
    for (size_t q = 0; gleich && q < b->ctx.size(); ++q) gleich = b->folge[q] == b->ctx[q]->folge;
    if (!gleich && b->graph) {
      (void)hipGraphExecDestroy(b->graph);
      b->graph = nullptr;
    }
    if (!b->graph) {
      b->folge.clear();
      for (gc_ctx* c : b->ctx) b->folge.push_back(c->folge);
      hipGraph_t g;
      BHIPCK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
      int rc = batch_schritt<T>(b, nullptr);
      hipError_t e2 = hipStreamEndCapture(b->stream, &g);
      if (rc) return 1;
      if (e2 != hipSuccess) return bfehler(b, std::string("hipStreamEndCapture: ") + hipGetErrorString(e2));
      hipError_t e3 = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e3 != hipSuccess) return bfehler(b, std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
    }
    BHIPCK(b, hipEventRecord(b->ev[0], b->stream));
    for (int k = done; k < n; ++k) BHIPCK(b, hipGraphLaunch(b->graph, b->stream));
    BHIPCK(b, hipEventRecord(b->ev[1], b->stream));
    BHIPCK(b, hipEventSynchronize(b->ev[1]));
    float ms = 0.f;
    BHIPCK(b, hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    b->ms_schritt = ms;
    b->ms_pcg = std::nan("");
    bool nochmal = false;
    if (batch_pruefen<T>(b, sichern ? &nochmal : nullptr)) return 1;
    if (!nochmal) return 0;
    if (batch_zurueck<T>(b, h)) return 1;
  }
}

// Weights from the last measured iterations per solve of every system.
static int batch_gewichte_messen(gc_batch* b) {
  for (size_t k = 0; k < b->ctx.size(); ++k) {
    GridState g;
    BHIPCK(b, hipMemcpy(&g, b->ctx[k]->gs, sizeof(GridState), hipMemcpyDeviceToHost));
    const double it = g.loesungen > 0 ? (double)g.cg_iterationen / (double)g.loesungen : 1.0;
    b->gewicht[k] = (double)std::max(1, b->ctx[k]->n_kacheln) * std::max(it, 1.0);
  }
  return 0;
}

// Autotuner: builds every system once, then times the batched first solve from
// the same warm start for each candidate launch shape (threads per block, total
// blocks, blocks by tiles or by tiles x measured iterations) and keeps the
// fastest. Positions, velocities, the warm start and the counters are restored;
// the run counters of self-collision cuts advance by the one build.
template <typename T>
static int batch_autotune(gc_batch* b, double h, int32_t* threads_out, int32_t* bloecke_out, double* ms_out) {
  StromTausch tausch(b);
  if (batch_vorbereiten<T>(b, h)) return 1;
  const size_t n = b->ctx.size();
  std::vector<void*> tmp;
  auto neu = [&](size_t bytes) -> void* {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(bytes, 1)) != hipSuccess) return nullptr;
    tmp.push_back(q);
    return q;
  };
  struct Gesichert { void *x, *v, *dv, *dv64; GridState alt, start; };
  std::vector<Gesichert> g(n);
  int rc = batch_sync(b);
  auto ende = [&]() {
    for (void* q : tmp) (void)hipFree(q);
    return rc;
  };
  for (size_t k = 0; k < n && !rc; ++k) {
    gc_ctx* c = b->ctx[k];
    const SysDev<T>& s = sys<T>(c);
    g[k].x = neu(s.n * sizeof(V3<double>));
    g[k].v = neu(s.n * sizeof(V3<double>));
    g[k].dv = neu(s.n * sizeof(V3<T>));
    g[k].dv64 = s.dv64 ? neu(s.n * sizeof(V3<double>)) : nullptr;  // the FP64 warm start of a refined FP32 solve
    if (!g[k].x || !g[k].v || !g[k].dv || (s.dv64 && !g[k].dv64)) { rc = bfehler(b, "hipMalloc failed (autotune)"); break; }
    if (hipMemcpyAsync(g[k].x, s.x, s.n * sizeof(V3<double>), hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(g[k].v, s.v, s.n * sizeof(V3<double>), hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(&g[k].alt, c->gs, sizeof(GridState), hipMemcpyDeviceToHost, b->stream) != hipSuccess)
      rc = bfehler(b, "hipMemcpy failed (autotune)");
  }
  if (rc) return ende();
  if (gabeln(b, 0)) { rc = 1; return ende(); }
  for (gc_ctx* c : b->ctx)
    if (launch_vor_pcg<T>(c, c->hat_vorher ? 1 : 0)) { rc = bfehler(b, c->err); return ende(); }
  if (vereinen(b, 0) || batch_sync(b)) { rc = 1; return ende(); }
  for (size_t k = 0; k < n && !rc; ++k) {
    const SysDev<T>& s = sys<T>(b->ctx[k]);
    if (hipMemcpyAsync(g[k].dv, s.dv, s.n * sizeof(V3<T>), hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
        (s.dv64 && hipMemcpyAsync(g[k].dv64, s.dv64, s.n * sizeof(V3<double>), hipMemcpyDeviceToDevice, b->stream) !=
                       hipSuccess) ||
        hipMemcpyAsync(&g[k].start, b->ctx[k]->gs, sizeof(GridState), hipMemcpyDeviceToHost, b->stream) != hipSuccess)
      rc = bfehler(b, "hipMemcpy failed (autotune)");
  }
  if (rc || hipStreamSynchronize(b->stream) != hipSuccess) { rc = rc ? rc : bfehler(b, "autotune copy failed"); return ende(); }
  if (batch_ev(b, 2)) { rc = 1; return ende(); }
  auto versuch = [&](int threads, int gesamt, double* ms) -> int {
    if (batch_form<T>(b, threads, gesamt)) return 1;
    for (size_t k = 0; k < n; ++k) {
      const SysDev<T>& s = sys<T>(b->ctx[k]);
      BHIPCK(b, hipMemcpyAsync(s.dv, g[k].dv, s.n * sizeof(V3<T>), hipMemcpyDeviceToDevice, b->stream));
      if (s.dv64) BHIPCK(b, hipMemcpyAsync(s.dv64, g[k].dv64, s.n * sizeof(V3<double>), hipMemcpyDeviceToDevice, b->stream));
      BHIPCK(b, hipMemcpyAsync(b->ctx[k]->gs, &g[k].start, sizeof(GridState), hipMemcpyHostToDevice, b->stream));
      hipLaunchKernelGGL(k_barriere_leeren, dim3(1), dim3(256), 0, b->stream, b->ctx[k]->gs, b->ctx[k]->slots);
    }
    BHIPCK(b, hipEventRecord(b->ev[0], b->stream));
    if (batch_pcg<T>(b, 0)) return 1;
    BHIPCK(b, hipEventRecord(b->ev[1], b->stream));
    BHIPCK(b, hipEventSynchronize(b->ev[1]));
    float m = 0.f;
    BHIPCK(b, hipEventElapsedTime(&m, b->ev[0], b->ev[1]));
    *ms = m;
    for (gc_ctx* c : b->ctx)
      if (pruefe_gitter(c)) return bfehler(b, c->err);
    return 0;
  };
  // weights by tiles; one solve measures every system's iterations
  std::vector<double> nach_kacheln(n), nach_arbeit(n);
  for (size_t k = 0; k < n; ++k) nach_kacheln[k] = (double)std::max(1, b->ctx[k]->n_kacheln);
  b->gewicht = nach_kacheln;
  double ms = 0.0;
  if (versuch(1024, 0, &ms)) { rc = 1; return ende(); }
  for (size_t k = 0; k < n; ++k) {
    GridState q;
    if (hipMemcpy(&q, b->ctx[k]->gs, sizeof(GridState), hipMemcpyDeviceToHost) != hipSuccess) { rc = bfehler(b, "hipMemcpy"); return ende(); }
    nach_arbeit[k] = nach_kacheln[k] * std::max(1.0, (double)q.iterationen);
  }
  int best_t = 1024, best_b = 0;
  int best_g = 0;
  double best_ms = 1e300;
  const int threads_kand[3] = {256, 512, 1024};
  const double anteil[4] = {0.25, 0.5, 0.75, 1.0};
  for (int gw = 0; gw < 2; ++gw)
    for (int ti = 0; ti < 3; ++ti)
      for (int ai = 0; ai < 4; ++ai) {
        b->gewicht = gw ? nach_arbeit : nach_kacheln;
        // blocks available at this thread count
        int occ = 0;
        const size_t lds = batch_lds<T>(b);
        if (hipFuncSetAttribute((const void*)k_pcg_stapel<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess ||
            hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, (const void*)k_pcg_stapel<T>, threads_kand[ti], lds) != hipSuccess ||
            occ < 1)
          continue;
        const int gesamt = std::max((int)n, (int)std::lround(anteil[ai] * occ * b->cus));
        double m = 0.0;
        if (versuch(threads_kand[ti], gesamt, &m)) { rc = 1; return ende(); }
        if (m < best_ms) { best_ms = m; best_t = threads_kand[ti]; best_b = gesamt; best_g = gw; }
      }
  b->gewicht = best_g ? nach_arbeit : nach_kacheln;
  if (batch_form<T>(b, best_t, best_b)) { rc = 1; return ende(); }
  for (size_t k = 0; k < n && !rc; ++k) {
    gc_ctx* c = b->ctx[k];
    const SysDev<T>& s = sys<T>(c);
    if (hipMemcpyAsync(s.x, g[k].x, s.n * sizeof(V3<double>), hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(s.v, g[k].v, s.n * sizeof(V3<double>), hipMemcpyDeviceToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(c->gs, &g[k].alt, sizeof(GridState), hipMemcpyHostToDevice, b->stream) != hipSuccess)
      rc = bfehler(b, "hipMemcpy failed (autotune restore)");
    c->built = false;
  }
  if (!rc && hipStreamSynchronize(b->stream) != hipSuccess) rc = bfehler(b, "autotune restore failed");
  if (threads_out) *threads_out = best_t;
  if (bloecke_out) *bloecke_out = b->n_tab;
  if (ms_out) *ms_out = best_ms;
  return ende();
}

// ---------------------------------------------------------------------------
// Loeser-level helpers
// ---------------------------------------------------------------------------
template <typename T>
static int hoch_vec(gc_ctx* c, V3<T>* dst, const double* src) {
  std::vector<V3<T>> tmp(c->n);
  for (int i = 0; i < c->n; ++i) tmp[i] = V3<T>{T(src[3 * i]), T(src[3 * i + 1]), T(src[3 * i + 2])};
  if (hoch(c, dst, tmp.data(), tmp.size())) return 1;
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

template <typename T>
static int runter_vec(gc_ctx* c, double* dst, const V3<T>* src, size_t count) {
  std::vector<V3<T>> tmp(count);
  if (count) HIPCK(c, hipMemcpyAsync(tmp.data(), src, count * sizeof(V3<T>), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < count; ++i) {
    dst[3 * i] = (double)tmp[i].x;
    dst[3 * i + 1] = (double)tmp[i].y;
    dst[3 * i + 2] = (double)tmp[i].z;
  }
  return 0;
}

template <typename T>
static int runter_skalar(gc_ctx* c, double* dst, const T* src, size_t count) {
  std::vector<T> tmp(count);
  if (count) HIPCK(c, hipMemcpyAsync(tmp.data(), src, count * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < count; ++i) dst[i] = (double)tmp[i];
  return 0;
}

template <typename T>
static int mal(gc_ctx* c, const double* p, double* out, int nur_h) {
  if (!c->built) return fehler(c, "no system built (call gc_build first)");
  SysDev<T>& s = sys<T>(c);
  if (hoch_vec<T>(c, s.ap, p)) return 1;
  if (sys_hoch<T>(c)) return 1;
  hipLaunchKernelGGL(k_mal<T>, dim3(gitter(s.n)), dim3(BUILD_THREADS), 0, c->stream, dsys<T>(c), (const V3<T>*)s.ap, s.r, nur_h);
  HIPCK(c, hipGetLastError());
  return runter_vec<T>(c, out, s.r, c->n);
}

template <typename T>
static int loese(gc_ctx* c, double* dv, gc_cg_lauf* lauf) {
  if (!c->built) return fehler(c, "no system built (call gc_build first)");
  SysDev<T>& s = sys<T>(c);
  if (hoch_vec<T>(c, s.dv, dv)) return 1;
  // FP32 bulk: the warm start and the solution are the FP64 dv64
  if (s.dv64 && hoch_vec<double>(c, s.dv64, dv)) return 1;
  if (sys_hoch<T>(c)) return 1;
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(256), 0, c->stream, c->gs, c->slots);
  if (launch_pcg<T>(c, 0)) return 1;
  GridState g;
  HIPCK(c, hipMemcpyAsync(&g, c->gs, sizeof(GridState), hipMemcpyDeviceToHost, c->stream));
  if (s.dv64 ? runter_vec<double>(c, dv, s.dv64, c->n) : runter_vec<T>(c, dv, s.dv, c->n)) return 1;
  if (g.fehler) {
    unsigned null = 0;
    HIPCK(c, hipMemcpy(&c->gs->fehler, &null, sizeof(unsigned), hipMemcpyHostToDevice));
    return fehler(c, barriere_meldung(g.fehler));
  }
  if (lauf) {
    lauf->iterationen = g.iterationen;
    lauf->nicht_konvergiert = g.nicht_konv;
    lauf->rest = g.rest;
  }
  return 0;
}

// ---------------------------------------------------------------------------
// The solve without allocations (gc_solve_ohne_speicher): the grid state and
// the reduction/barrier slots of its PCG launch live in this array instead
// of the context's buffers. Its layout: the grid state at byte 0, the
// GC_SLOTS slots from byte GC_OS_SLOTS on. Each device holds its own copy;
// host code uses the per-device address from hipGetSymbolAddress, never the
// array's name. Every call on a device shares that device's copy, so calls
// that overlap on one device race.
// ---------------------------------------------------------------------------
#ifdef GC_PHASEN
#define GC_OS_SLOTS 512
#else
#define GC_OS_SLOTS 256
#endif
#define GC_OS_BYTES (GC_OS_SLOTS + GC_SLOTS * sizeof(Slot))
static_assert(sizeof(GridState) <= GC_OS_SLOTS, "the grid state fits before the slots");
static __device__ __align__(256) unsigned char scratch[GC_OS_BYTES];

#define GC_OS_GERAETE 64
static std::mutex os_mutex;
static void* os_adresse[GC_OS_GERAETE] = {};

// The device address of `scratch` on device g (current), fetched once per
// device. Called by gc_create, so the entry point itself only looks it up.
static int os_adresse_holen(int g, std::string* err) {
  if (g < 0 || g >= GC_OS_GERAETE) {
    *err = "device ordinal beyond the scratch table";
    return 1;
  }
  std::lock_guard<std::mutex> lock(os_mutex);
  if (os_adresse[g]) return 0;
  void* ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) {
    *err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    return 1;
  }
  os_adresse[g] = ptr;
  return 0;
}

// Clears the whole grid state (counters, error flag, results) and the
// barrier counters of a slot array (a fill kernel: no memset node when a
// caller captures the call).
__global__ void k_os_leeren(GridState* gs, Slot* slots) {
  unsigned* w = (unsigned*)gs;
  for (unsigned k = threadIdx.x; k < sizeof(GridState) / sizeof(unsigned); k += blockDim.x) w[k] = 0u;
  barriere_leeren(slots);
}

template <typename T>
static int loese_ohne_speicher(gc_ctx* c, double* dv, gc_cg_lauf* lauf) {
  if (!c->built) return fehler(c, "no system built (call gc_build first)");
  if (c->pcg_blocks > GC_MAX_BLOECKE)
    return fehler(c, "gc_solve_ohne_speicher: " + std::to_string(c->pcg_blocks) + " PCG blocks exceed the " +
                         std::to_string(GC_MAX_BLOECKE) + " the scratch array holds reductions for");
  void* basis = nullptr;
  {
    std::lock_guard<std::mutex> lock(os_mutex);
    basis = os_adresse[c->opt.geraet];
  }
  if (!basis) return fehler(c, "gc_solve_ohne_speicher: no scratch address for this device");
  GridState* gs = (GridState*)basis;
  Slot* slots = (Slot*)((unsigned char*)basis + GC_OS_SLOTS);
  SysDev<T>& s = sys<T>(c);
  // the warm start straight into the solution vector: FP64 dv64 in the
  // FP32-bulk form (the correction dv is cleared by the solve), dv in FP64
  const size_t bytes = (size_t)c->n * sizeof(V3<double>);
  V3<double>* ziel = sizeof(T) == 4 ? s.dv64 : (V3<double>*)s.dv;
  if (!ziel) return fehler(c, "gc_solve_ohne_speicher: the FP32 form needs its FP64 solution vector");
  if (sys_hoch<T>(c)) return 1;
  HIPCK(c, hipMemcpyAsync(ziel, dv, bytes, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_os_leeren, dim3(1), dim3(256), 0, c->stream, gs, slots);
  if (sizeof(T) == 4)
    launch_verfeinert(c->stream, c->pcg_blocks, c->pcg_threads, c->pcg_lds, (const SysDev<float>*)dsys<T>(c), gs, slots, 0);
  else
    hipLaunchKernelGGL(k_pcg<T>, dim3(c->pcg_blocks), dim3(c->pcg_threads), c->pcg_lds, c->stream, dsys<T>(c), gs, slots, 0);
  HIPCK(c, hipGetLastError());
  GridState g;
  HIPCK(c, hipMemcpyAsync(&g, gs, sizeof(GridState), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipMemcpyAsync(dv, ziel, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  if (g.fehler) return fehler(c, barriere_meldung(g.fehler));
  if (lauf) {
    lauf->iterationen = g.iterationen;
    lauf->nicht_konvergiert = g.nicht_konv;
    lauf->rest = g.rest;
  }
  return 0;
}

template <typename T>
static int fest_setzen(gc_ctx* c, const int32_t* flag, const double* normale, const double* vn_soll, const double* vk) {
  SysDev<T>& s = sys<T>(c);
  const int n = c->n;
  std::vector<uint8_t> fl(n);
  std::vector<V3<T>> nt(n);
  std::vector<V3<double>> nd(n), vv(n);
  for (int i = 0; i < n; ++i) {
    fl[i] = c->w[i] <= 0.0 ? 2 : (flag[i] ? 1 : 0);
    nd[i] = V3<double>{normale[3 * i], normale[3 * i + 1], normale[3 * i + 2]};
    nt[i] = cast3<T>(nd[i]);
    vv[i] = V3<double>{vk[3 * i], vk[3 * i + 1], vk[3 * i + 2]};
  }
  if (hoch(c, s.fflag, fl.data(), n) || hoch(c, s.fnorm, nt.data(), n) || hoch(c, s.fnorm64, nd.data(), n) ||
      hoch(c, s.vn_soll, vn_soll, n) || hoch(c, s.vk, vv.data(), n))
    return 1;
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

template <typename T>
static int fest_lesen(gc_ctx* c, int32_t* flag, double* normale, double* vn_soll, double* vk) {
  SysDev<T>& s = sys<T>(c);
  const int n = c->n;
  std::vector<uint8_t> fl(n);
  std::vector<V3<double>> nd(n), vv(n);
  HIPCK(c, hipMemcpyAsync(fl.data(), s.fflag, n, hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipMemcpyAsync(nd.data(), s.fnorm64, n * sizeof(V3<double>), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipMemcpyAsync(vv.data(), s.vk, n * sizeof(V3<double>), hipMemcpyDeviceToHost, c->stream));
  if (vn_soll) HIPCK(c, hipMemcpyAsync(vn_soll, s.vn_soll, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) {
    if (flag) flag[i] = fl[i] == 1 ? 1 : 0;
    if (normale) { normale[3 * i] = nd[i].x; normale[3 * i + 1] = nd[i].y; normale[3 * i + 2] = nd[i].z; }
    if (vk) { vk[3 * i] = vv[i].x; vk[3 * i + 1] = vv[i].y; vk[3 * i + 2] = vv[i].z; }
  }
  return 0;
}

template <typename T>
static int system_lesen(gc_ctx* c, double* rhs, double* f, double* last, double* pinv9) {
  if (!c->built) return fehler(c, "no system built (call gc_build first)");
  SysDev<T>& s = sys<T>(c);
  const size_t n = c->n;
  if (rhs && runter_vec<T>(c, rhs, s.b, n)) return 1;
  if (f) HIPCK(c, hipMemcpyAsync(f, s.f, n * sizeof(V3<double>), hipMemcpyDeviceToHost, c->stream));
  if (last) HIPCK(c, hipMemcpyAsync(last, s.last, n * sizeof(V3<double>), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  if (pinv9 && runter_skalar<T>(c, pinv9, s.pinv, 9 * n)) return 1;
  return 0;
}

template <typename T>
static int records_lesen(gc_ctx* c, double* richtung, double* quer, double* grad) {
  if (!c->built) return fehler(c, "no system built (call gc_build first)");
  SysDev<T>& s = sys<T>(c);
  if (richtung && runter_vec<T>(c, richtung, s.k_n, c->nk)) return 1;
  if (quer && runter_skalar<T>(c, quer, s.k_q, c->nk)) return 1;
  if (grad && runter_vec<T>(c, grad, s.b_grad, 4 * c->b_k.size())) return 1;
  return 0;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// adjoint of the step (reverse mode over one substep, gc_adjoint.h)
// ---------------------------------------------------------------------------
template <typename T>
static int adj_vorbereiten(gc_ctx* c) {
  if (c->adj_bereit) return 0;
  const SysDev<T>& s = sys<T>(c);
  const size_t n = (size_t)c->n;
  AdjDev& a = c->adj;
  a.gx = dalloc<V3<double>>(c, n);
  a.gv = dalloc<V3<double>>(c, n);
  a.x0 = dalloc<V3<double>>(c, n);
  a.v0 = dalloc<V3<double>>(c, n);
  a.gx_p = dalloc<V3<double>>(c, n);
  a.gv_p = dalloc<V3<double>>(c, n);
  a.gdv = dalloc<V3<double>>(c, n);
  a.ziel = dalloc<double>(c, n);
  a.nu = dalloc<double>(c, n);
  a.e_kante = dalloc<V3<double>>(c, 2 * (size_t)s.nk);
  a.e_biegung = dalloc<V3<double>>(c, 4 * (size_t)s.nb);
  a.e_membran = dalloc<V3<double>>(c, 3 * (size_t)s.nm);
  a.d_kante = dalloc<double>(c, s.nk);
  a.d_biegung = dalloc<double>(c, s.nb);
  a.d_membran = dalloc<double>(c, 3 * (size_t)s.nm);
  a.d_reibung = dalloc<double>(c, n);
  a.zaehler = dalloc<unsigned long long>(c, 2);
  c->adj_b = dalloc<V3<T>>(c, n);
  c->adj_mu = dalloc<V3<T>>(c, n);
  c->adj_sys_dev = dalloc<SysDev<T>>(c, 1);
  c->gs_adj = dalloc<GridState>(c, 1);
  if (!a.gx || !a.gv || !a.x0 || !a.v0 || !a.gx_p || !a.gv_p || !a.gdv || !a.ziel || !a.nu || !a.e_kante || !a.e_biegung ||
      !a.e_membran || !a.d_kante || !a.d_biegung || !a.d_membran || !a.d_reibung || !a.zaehler || !c->adj_b || !c->adj_mu ||
      !c->adj_sys_dev || !c->gs_adj)
    return fehler(c, "hipMalloc failed (adjoint)");
  HIPCK(c, hipStreamSynchronize(c->stream));
  c->adj_bereit = true;
  return 0;
}

// One step from the context's state (as gc_step(h, 1) computes it), then its
// reverse: the adjoint seed (dL/dx, dL/dv of the state after the step)
// becomes dL/dx, dL/dv of the state before it, and the step's parameter
// gradients are added to the accumulators.
template <typename T>
static int adj_schritt(gc_ctx* c, double h) {
  if (!c->topo || !c->hat_param) return fehler(c, "gc_adjoint_step: topology and parameters must be set first");
  if (!(h > 0.0)) return fehler(c, "h must be > 0");
  if (selbst_an(c))
    return fehler(c, "the adjoint does not differentiate cloth self-collision: switch it off (gc_param.selbstkollision = 0)");
  for (const auto& k : c->koerper)
    if (k.k.art == GC_NETZ) return fehler(c, "the adjoint does not differentiate the mesh body's contact; use a grid SDF body");
  if (c->param.normalkraft_aus_last)
    return fehler(c, "the adjoint takes the normal force from the residual (gc_param.normalkraft_aus_last = 0)");
  if (adj_vorbereiten<T>(c)) return 1;
  SysDev<T>& s = sys<T>(c);
  s.h = h;
  s.h2 = h * h;
  if (sys_hoch<T>(c)) return 1;
  AdjDev& a = c->adj;
  hipStream_t st = c->stream;
  const size_t n = (size_t)c->n;
  HIPCK(c, hipMemcpyAsync(a.x0, s.x, n * sizeof(V3<double>), hipMemcpyDeviceToDevice, st));
  HIPCK(c, hipMemcpyAsync(a.v0, s.v, n * sizeof(V3<double>), hipMemcpyDeviceToDevice, st));
  if (launch_schritt<T>(c, c->hat_vorher ? 1 : 0, nullptr)) return 1;
  c->hat_vorher = true;
  c->built = false;
  // the adjoint solve: the step's system with right-hand side dL/d(dv),
  // solution mu and the normal targets of the adjoint; a purely relative
  // stopping test (the seed's scale is arbitrary)
  SysDev<T> sa = s;
  sa.b = (V3<T>*)c->adj_b;
  sa.dv = (V3<T>*)c->adj_mu;
  sa.soll_direkt = a.ziel;
  // the adjoint solve runs on the bulk system alone (its right-hand side has
  // no FP64 copy); the step's own solution stays readable through sp's dv64
  sa.schatten = nullptr;
  sa.dv64 = nullptr;
  sa.boden = 0.0;
  c->adj_sys_host.assign((const unsigned char*)&sa, (const unsigned char*)&sa + sizeof(sa));
  HIPCK(c, hipMemcpyAsync(c->adj_sys_dev, c->adj_sys_host.data(), sizeof(sa), hipMemcpyHostToDevice, st));
  const SysDev<T>* sp = dsys<T>(c);
  const SysDev<T>* sap = (const SysDev<T>*)c->adj_sys_dev;
  const V3<T>* mu = (const V3<T>*)c->adj_mu;
  hipLaunchKernelGGL(k_adj_update<T>, dim3(gitter(s.n)), dim3(BUILD_THREADS), 0, st, sp, a, (V3<T>*)c->adj_b);
  // the adjoint solve starts from the previous (later) step's mu, cleared by
  // gc_adjoint_set: the adjoint state changes little from step to step
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(256), 0, st, c->gs_adj, c->slots);
  hipLaunchKernelGGL(k_pcg<T>, dim3(c->pcg_blocks), dim3(c->pcg_threads), c->pcg_lds, st, sap, c->gs_adj, c->slots, 0);
  hipLaunchKernelGGL(k_adj_nu<T>, dim3(gitter(s.n)), dim3(BUILD_THREADS), 0, st, sp, a, mu);
  if (s.nk) hipLaunchKernelGGL(k_adj_kanten<T>, dim3(gitter(s.nk)), dim3(BUILD_THREADS), 0, st, sp, a, mu);
  if (s.nb) hipLaunchKernelGGL(k_adj_biegung<T>, dim3(gitter(s.nb)), dim3(BUILD_THREADS), 0, st, sp, a, mu);
  if (s.nm) hipLaunchKernelGGL(k_adj_membran<T>, dim3(gitter(s.nm)), dim3(BUILD_THREADS), 0, st, sp, a, mu);
  hipLaunchKernelGGL(k_adj_sammeln<T>, dim3(gitter(s.n)), dim3(BUILD_THREADS), 0, st, sp, a, mu);
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipStreamSynchronize(st));
  ++c->adj_schritte;
  if (pruefe_gitter(c)) return 1;
  unsigned fe = 0;
  HIPCK(c, hipMemcpy(&fe, &c->gs_adj->fehler, sizeof(unsigned), hipMemcpyDeviceToHost));
  if (fe) {
    unsigned null = 0;
    HIPCK(c, hipMemcpy(&c->gs_adj->fehler, &null, sizeof(unsigned), hipMemcpyHostToDevice));
    return fehler(c, barriere_meldung(fe) + " (adjoint solve)");
  }
  return 0;
}

// New stiffness and friction values on an uploaded topology (the zero
// pattern of the edge stiffnesses stays: a k = 0 edge is carried by a membrane
// and not on the device).
template <typename T>
static int steifigkeit_setzen(gc_ctx* c, const double* kk, const double* bk, const double* md, const double* rb) {
  SysDev<T>& s = sys<T>(c);
  if (kk) {
    int nz = 0, nz_dev = 0;
    for (int e = 0; e < c->n_kanten_in; ++e) nz += kk[e] != 0.0 ? 1 : 0;
    for (int q = 0; q < c->nk; ++q) nz_dev += c->k_quelle[q] >= 0 ? 1 : 0;
    for (int q = 0; q < c->nk; ++q)
      if (c->k_quelle[q] >= 0 && kk[c->k_quelle[q]] == 0.0) nz = -1;
    if (nz != nz_dev) return fehler(c, "gc_set_steifigkeit: the edges with k = 0 must stay the topology's");
    for (int q = 0; q < c->nk; ++q)
      if (c->k_quelle[q] >= 0) c->k_k[q] = kk[c->k_quelle[q]];
    if (hoch(c, const_cast<double*>(s.k_k), c->k_k.data(), c->k_k.size())) return 1;
  }
  if (bk) {
    c->b_k.assign(bk, bk + c->b_k.size());
    if (hoch(c, const_cast<double*>(s.b_k), c->b_k.data(), c->b_k.size())) return 1;
  }
  if (md) {
    const size_t nm = c->m_p.size() / 3;
    std::vector<Membran<T>> mel(nm);
    std::vector<Membran<double>> mel64(nm);
    for (size_t e = 0; e < nm; ++e) {
      for (int q = 0; q < 3; ++q) c->m_d[12 * e + 9 + q] = md[3 * e + q];
      const double* d = &c->m_d[12 * e];
      Membran<double>& m = mel64[e];
      m.flaeche = d[0];
      for (int q = 0; q < 4; ++q) m.r[q] = d[1 + q];
      m.kette[0] = d[5]; m.kette[1] = d[6];
      m.schuss[0] = d[7]; m.schuss[1] = d[8];
      m.ek = d[9]; m.es = d[10]; m.g = d[11];
      Membran<T>& t = mel[e];
      t.flaeche = T(m.flaeche);
      for (int q = 0; q < 4; ++q) t.r[q] = T(m.r[q]);
      t.kette[0] = T(m.kette[0]); t.kette[1] = T(m.kette[1]);
      t.schuss[0] = T(m.schuss[0]); t.schuss[1] = T(m.schuss[1]);
      t.ek = T(m.ek); t.es = T(m.es); t.g = T(m.g);
    }
    if (hoch(c, const_cast<Membran<T>*>(s.m_el), mel.data(), nm)) return 1;
    if ((const void*)s.m_el64 != (const void*)s.m_el && hoch(c, const_cast<Membran<double>*>(s.m_el64), mel64.data(), nm)) return 1;
  }
  if (rb) {
    c->reib.assign(rb, rb + c->n);
    if (hoch(c, const_cast<double*>(s.reib), c->reib.data(), c->reib.size())) return 1;
  }
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

template <typename T>
static int adj_gradient_lesen(gc_ctx* c, double* d_kanten, double* d_naehte, double* d_biegung, double* d_membran,
                              double* d_reibung, int reset) {
  if (adj_vorbereiten<T>(c)) return 1;
  const SysDev<T>& s = sys<T>(c);
  const AdjDev& a = c->adj;
  std::vector<double> dk(s.nk), db(s.nb), dm(3 * (size_t)s.nm), dr(c->n);
  hipStream_t st = c->stream;
  if (s.nk) HIPCK(c, hipMemcpyAsync(dk.data(), a.d_kante, dk.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  if (s.nb) HIPCK(c, hipMemcpyAsync(db.data(), a.d_biegung, db.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  if (s.nm) HIPCK(c, hipMemcpyAsync(dm.data(), a.d_membran, dm.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCK(c, hipMemcpyAsync(dr.data(), a.d_reibung, dr.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCK(c, hipStreamSynchronize(st));
  if (d_kanten) std::fill(d_kanten, d_kanten + c->n_kanten_in, 0.0);
  if (d_naehte) std::fill(d_naehte, d_naehte + c->n_naehte, 0.0);
  for (int q = 0; q < s.nk; ++q) {
    const int e = c->k_quelle[q];
    if (e >= 0 && d_kanten) d_kanten[e] = dk[q];
    if (e < 0 && d_naehte) d_naehte[-(e + 1)] = dk[q];
  }
  if (d_biegung) std::copy(db.begin(), db.end(), d_biegung);
  if (d_membran) std::copy(dm.begin(), dm.end(), d_membran);
  if (d_reibung) std::copy(dr.begin(), dr.end(), d_reibung);
  if (reset) {
    if (s.nk) HIPCK(c, hipMemsetAsync(a.d_kante, 0, dk.size() * sizeof(double), st));
    if (s.nb) HIPCK(c, hipMemsetAsync(a.d_biegung, 0, db.size() * sizeof(double), st));
    if (s.nm) HIPCK(c, hipMemsetAsync(a.d_membran, 0, dm.size() * sizeof(double), st));
    HIPCK(c, hipMemsetAsync(a.d_reibung, 0, dr.size() * sizeof(double), st));
    HIPCK(c, hipStreamSynchronize(st));
  }
  return 0;
}


extern "C" {

gc_ctx* gc_create(const gc_optionen* opt, char* err, int32_t err_len) {
  auto melde = [&](const std::string& m) {
    if (err && err_len > 0) {
      std::snprintf(err, (size_t)err_len, "%s", m.c_str());
    }
  };
  gc_optionen o{};
  if (opt) o = *opt;
  if (o.praezision != GC_FP64 && o.praezision != GC_FP32) {
    melde("praezision must be GC_FP64 or GC_FP32");
    return nullptr;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    melde("no HIP device");
    return nullptr;
  }
  if (o.geraet < 0 || o.geraet >= ndev) {
    melde("device ordinal out of range");
    return nullptr;
  }
  if (hipSetDevice(o.geraet) != hipSuccess) {
    melde("hipSetDevice failed");
    return nullptr;
  }
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, o.geraet) != hipSuccess) {
    melde("hipGetDeviceProperties failed");
    return nullptr;
  }
  if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) {
    melde(std::string("built for gfx950, device is ") + p.gcnArchName);
    return nullptr;
  }
  {
    // the device address of the no-allocation solve's scratch array, once
    // per device, with the device current
    std::string e;
    if (os_adresse_holen(o.geraet, &e)) {
      melde(e);
      return nullptr;
    }
  }
  gc_ctx* c = new gc_ctx();
  c->opt = o;
  c->cus = p.multiProcessorCount;
  c->lds_cu = p.maxSharedMemoryPerMultiProcessor > 0 ? (size_t)p.maxSharedMemoryPerMultiProcessor : 65536;
  // default tile budget 80 KiB: measured best on the 40k drape (80-112 KiB
  // equal within noise, 64 and 128 slower), the 250k drape and the batch
  c->lds_budget = (size_t)(o.lds_kb > 0 ? o.lds_kb : 80) * 1024;
  if (c->lds_budget + 1024 > (size_t)p.sharedMemPerBlock) {
    melde("lds_kb exceeds the LDS of a workgroup");
    delete c;
    return nullptr;
  }
  c->pcg_threads = o.threads > 0 ? o.threads : 1024;
  if (c->pcg_threads > 1024 || c->pcg_threads < 64 || (c->pcg_threads & (c->pcg_threads - 1))) {
    melde("threads must be a power of two in [64, 1024]");
    delete c;
    return nullptr;
  }
  c->name = o.praezision == GC_FP32 ? "MI355X HIP (f32, f64 reductions)" : "MI355X HIP (f64)";
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    melde("hipStreamCreate failed");
    delete c;
    return nullptr;
  }
  c->gs = dalloc<GridState>(c, 1);
  c->slots = dalloc<Slot>(c, GC_SLOTS);
  if (!c->gs || !c->slots) {
    melde("hipMalloc failed");
    frei_alles(c);
    delete c;
    return nullptr;
  }
  c->param.schwerkraft[2] = -9.81;
  c->param.cg_max = 50;
  c->param.cg_toleranz = 1e-5;
  c->param.naht_k = 500.0;
  c->param.kontakt_rand = 1e-4;
  return c;
}

void gc_destroy(gc_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->opt.geraet);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  frei_alles(c);
  for (auto e : c->ev) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* gc_last_error(const gc_ctx* c) { return c ? c->err.c_str() : "null context"; }
const char* gc_name(const gc_ctx* c) { return c ? c->name.c_str() : ""; }

int gc_set_topology(gc_ctx* c, const gc_topologie* t) {
  if (!c || !t) return 1;
  if (c->topo) return fehler(c, "topology already set (create a new context for a new garment)");
  const int n = t->n;
  if (n <= 0) return fehler(c, "n must be > 0");
  auto pruefe = [&](const uint32_t* idx, size_t cnt) {
    for (size_t q = 0; q < cnt; ++q)
      if (idx[q] >= (uint32_t)n) return false;
    return true;
  };
  if (!pruefe(t->kanten_punkte, 2 * (size_t)t->n_kanten) || !pruefe(t->naht_punkte, 2 * (size_t)t->n_naehte) ||
      !pruefe(t->biegung_punkte, 4 * (size_t)t->n_biegungen) || !pruefe(t->membran_punkte, 3 * (size_t)t->n_membran))
    return fehler(c, "point index out of range in topology");
  c->n = n;
  c->masse.assign(t->masse, t->masse + n);
  c->w.assign(t->w, t->w + n);
  c->koll.assign(t->kollisionsabstand, t->kollisionsabstand + n);
  c->reib.assign(t->reibung, t->reibung + n);
  c->msum = 0.0;
  for (int i = 0; i < n; ++i) c->msum += c->masse[i];  // implizit.rs:397 (sequential sum)
  // device distance list: edges with k != 0 (implizit.rs:190-192 skips k = 0), then seams
  c->k_ij.clear(); c->k_k.clear(); c->k_ruhe.clear(); c->k_naht.clear(); c->k_quelle.clear();
  c->n_kanten_in = t->n_kanten;
  c->n_naehte = t->n_naehte;
  for (int e = 0; e < t->n_kanten; ++e) {
    if (t->kanten_k[e] == 0.0) continue;
    c->k_ij.push_back(t->kanten_punkte[2 * e]);
    c->k_ij.push_back(t->kanten_punkte[2 * e + 1]);
    c->k_k.push_back(t->kanten_k[e]);
    c->k_ruhe.push_back(t->kanten_ruhe[e]);
    c->k_naht.push_back(0);
    c->k_quelle.push_back(e);
  }
  for (int e = 0; e < t->n_naehte; ++e) {
    c->k_ij.push_back(t->naht_punkte[2 * e]);
    c->k_ij.push_back(t->naht_punkte[2 * e + 1]);
    c->k_k.push_back(c->param.naht_k);
    c->k_ruhe.push_back(t->naht_ruhe[e]);
    c->k_naht.push_back(1);
    c->k_quelle.push_back(-(1 + e));
  }
  c->nk = (int)c->k_k.size();
  c->b_p.assign(t->biegung_punkte, t->biegung_punkte + 4 * (size_t)t->n_biegungen);
  c->b_k.assign(t->biegung_k, t->biegung_k + t->n_biegungen);
  c->b_ruhe.assign(t->biegung_ruhe, t->biegung_ruhe + t->n_biegungen);
  c->m_p.assign(t->membran_punkte, t->membran_punkte + 3 * (size_t)t->n_membran);
  c->m_d.assign(t->membran_daten, t->membran_daten + 12 * (size_t)t->n_membran);
  if (t->n_dreiecke > 0 && t->dreiecke) {
    if (!pruefe(t->dreiecke, 3 * (size_t)t->n_dreiecke)) return fehler(c, "point index out of range in triangles");
    c->sk_dreiecke_in.assign(t->dreiecke, t->dreiecke + 3 * (size_t)t->n_dreiecke);
    if (t->selbst_e_punkt) c->sk_e_punkt.assign(t->selbst_e_punkt, t->selbst_e_punkt + n);
    if (t->selbst_e_dreieck) c->sk_e_dreieck.assign(t->selbst_e_dreieck, t->selbst_e_dreieck + t->n_dreiecke);
    if (t->selbst_a_dreieck) c->sk_a_dreieck.assign(t->selbst_a_dreieck, t->selbst_a_dreieck + t->n_dreiecke);
  }
  c->sk_abstand_in.assign(t->kanten_punkte, t->kanten_punkte + 2 * (size_t)t->n_kanten);
  c->sk_abstand_ruhe_in.assign(t->kanten_ruhe, t->kanten_ruhe + t->n_kanten);
  c->sk_naht_in.assign(t->naht_punkte, t->naht_punkte + 2 * (size_t)t->n_naehte);
  c->sk_naht_ruhe.assign(t->naht_ruhe, t->naht_ruhe + t->n_naehte);
  // CSR by point and category, each category in constraint order
  std::vector<uint32_t> cnt((size_t)GC_KAT * n, 0);
  for (int k = 0; k < c->nk; ++k) {
    cnt[GC_KAT * (size_t)c->k_ij[2 * k] + 0]++;
    cnt[GC_KAT * (size_t)c->k_ij[2 * k + 1] + 0]++;
  }
  for (size_t q = 0; q < c->b_p.size(); ++q) cnt[GC_KAT * (size_t)c->b_p[q] + 1]++;
  for (size_t q = 0; q < c->m_p.size(); ++q) cnt[GC_KAT * (size_t)c->m_p[q] + 2]++;
  std::vector<uint32_t> off((size_t)GC_KAT * n + 1);
  uint32_t run = 0;
  for (size_t q = 0; q < (size_t)GC_KAT * n; ++q) {
    off[q] = run;
    run += cnt[q];
  }
  off[(size_t)GC_KAT * n] = run;
  std::vector<uint32_t> csr(run);
  std::vector<uint32_t> cur(off.begin(), off.end() - 1);
  for (int k = 0; k < c->nk; ++k) {
    csr[cur[GC_KAT * (size_t)c->k_ij[2 * k] + 0]++] = ((uint32_t)k << 2) | 0u;
    csr[cur[GC_KAT * (size_t)c->k_ij[2 * k + 1] + 0]++] = ((uint32_t)k << 2) | 1u;
  }
  for (size_t b = 0; b < c->b_k.size(); ++b)
    for (uint32_t m = 0; m < 4; ++m) csr[cur[GC_KAT * (size_t)c->b_p[4 * b + m] + 1]++] = ((uint32_t)b << 2) | m;
  for (size_t e = 0; e < c->m_p.size() / 3; ++e)
    for (uint32_t k = 0; k < 3; ++k) csr[cur[GC_KAT * (size_t)c->m_p[3 * e + k] + 2]++] = ((uint32_t)e << 2) | k;
  HIPCK(c, hipSetDevice(c->opt.geraet));
  int rc = c->opt.praezision == GC_FP32 ? topologie_laden<float>(c, off, csr) : topologie_laden<double>(c, off, csr);
  if (rc) return rc;
  HIPCK(c, hipStreamSynchronize(c->stream));
  c->topo = true;
  if (kp_bereiten(c)) return 1;
  alle_param(c);
  return 0;
}

int gc_set_param(gc_ctx* c, const gc_param* p) {
  if (!c || !p) return 1;
  if (p->cg_max < 0) return fehler(c, "cg_max must be >= 0");
  if (p->koerperpunkt_durchgaenge < 0) return fehler(c, "koerperpunkt_durchgaenge must be >= 0");
  const bool selbst_vorher = selbst_an(c);
  const int stoss_vorher = c->param.selbst_stoss_iterationen;
  const int kp_vorher = c->param.koerperpunkt_durchgaenge;
  const double kp_tiefe_vorher = c->param.koerperpunkt_tiefe;
  c->param = *p;
  c->hat_param = true;
  // self-collision on or off, the impact iterations and the passes of body
  // points decide which kernels a substep launches (and with which depth): a
  // captured graph of the old sequence is dropped
  if (selbst_an(c) != selbst_vorher || (selbst_an(c) && p->selbst_stoss_iterationen != stoss_vorher) ||
      p->koerperpunkt_durchgaenge != kp_vorher || !(p->koerperpunkt_tiefe == kp_tiefe_vorher))
    folge_neu(c);
  if (kp_bereiten(c)) return 1;
  if (c->sk_bereit) {
    // the kernels read the spring stiffness factor, the seam-cut switch and
    // the impact iterations from the device copy of the self-collision tables
    c->sk.steifigkeit = p->selbst_steifigkeit;
    c->sk.an_naehten = p->selbst_schnitte_an_naehten;
    c->sk.stoss_iterationen = p->selbst_stoss_iterationen;
    HIPCK(c, hipMemcpyAsync(c->sk_dev, &c->sk, sizeof(SelbstDev), hipMemcpyHostToDevice, c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
  }
  if (c->topo) {
    // seam stiffness is the parameter naht_k (implizit.rs:760, 771)
    for (int k = 0; k < c->nk; ++k)
      if (c->k_naht[k]) c->k_k[k] = p->naht_k;
    double* dk = const_cast<double*>(c->opt.praezision == GC_FP32 ? c->sf.k_k : c->sd.k_k);
    if (hoch(c, dk, c->k_k.data(), c->k_k.size())) return 1;
    HIPCK(c, hipStreamSynchronize(c->stream));
  }
  alle_param(c);
  return 0;
}

int gc_set_naht_ruhe(gc_ctx* c, const double* ruhe) {
  if (!c || !ruhe || !c->topo) return 1;
  for (int k = 0; k < c->nk; ++k)
    if (c->k_naht[k]) c->k_ruhe[k] = ruhe[-(c->k_quelle[k] + 1)];
  double* dr = const_cast<double*>(c->opt.praezision == GC_FP32 ? c->sf.k_ruhe : c->sd.k_ruhe);
  if (hoch(c, dr, c->k_ruhe.data(), c->k_ruhe.size())) return 1;
  c->sk_naht_ruhe.assign(ruhe, ruhe + c->n_naehte);
  if (c->sk_naht_ruhe_dev && hoch(c, c->sk_naht_ruhe_dev, c->sk_naht_ruhe.data(), c->sk_naht_ruhe.size())) return 1;
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gc_set_w(gc_ctx* c, const double* w) {
  if (!c || !w || !c->topo) return 1;
  c->w.assign(w, w + c->n);
  double* dw = const_cast<double*>(c->opt.praezision == GC_FP32 ? c->sf.w : c->sd.w);
  if (hoch(c, dw, c->w.data(), c->w.size())) return 1;
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gc_add_zylinder(gc_ctx* c, const double punkt[3], const double achse[3], double radius, const double v[3]) {
  if (!c) return -1;
  if (c->koerper.size() >= GC_MAX_KOERPER) {
    c->err = "too many bodies";
    return -1;
  }
  Koerperhost kh;
  std::memset(&kh.k, 0, sizeof(Koerper));
  kh.k.art = GC_ZYLINDER;
  for (int q = 0; q < 3; ++q) {
    kh.k.punkt[q] = punkt[q];
    kh.k.achse[q] = achse[q];
    kh.k.geschwindigkeit[q] = v[q];
  }
  kh.k.radius = radius;
  c->koerper.push_back(kh);
  alle_param(c);
  return (int)c->koerper.size() - 1;
}

int gc_add_sdf(gc_ctx* c, const float* werte, const int32_t dims[3], const double origin[3], double spacing,
               const double versatz[3], const double v[3]) {
  if (!c) return -1;
  if (c->koerper.size() >= GC_MAX_KOERPER) {
    c->err = "too many bodies";
    return -1;
  }
  if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2 || !(spacing > 0.0)) {
    c->err = "SDF grid needs at least 2 nodes per axis and a positive spacing";
    return -1;
  }
  size_t cnt = (size_t)dims[0] * dims[1] * dims[2];
  Koerperhost kh;
  std::memset(&kh.k, 0, sizeof(Koerper));
  if (hipMalloc(&kh.werte_dev, cnt * sizeof(float)) != hipSuccess) {
    c->err = "hipMalloc failed (SDF)";
    return -1;
  }
  if (hipMemcpy(kh.werte_dev, werte, cnt * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    c->err = "hipMemcpy failed (SDF)";
    (void)hipFree(kh.werte_dev);
    return -1;
  }
  kh.k.art = GC_SDF;
  kh.k.werte = kh.werte_dev;
  for (int q = 0; q < 3; ++q) {
    kh.k.dims[q] = dims[q];
    kh.k.origin[q] = origin[q];
    kh.k.punkt[q] = versatz[q];
    kh.k.geschwindigkeit[q] = v[q];
  }
  kh.k.spacing = spacing;
  c->koerper.push_back(kh);
  alle_param(c);
  return (int)c->koerper.size() - 1;
}

// --- mesh body (netzkoerper.rs) ----------------------------------------------
}  // extern "C"

// Koerpernetz::neu's checks and topology (netzkoerper.rs:297-348): every
// undirected edge has one or two triangles, two of them run it in opposite
// directions; the neighbour across each side (the triangle itself at an open
// edge). Also each point's incidences (3t + k, ascending) for its normal.
static std::string netz_topologie(Netzhost& h, std::vector<uint32_t>& nachbar, std::vector<uint32_t>& inz_anfang,
                                  std::vector<uint32_t>& inz) {
  if (h.nt == 0) return "mesh body without triangles";
  if (!(h.suchtiefe > 0.0)) return "search depth " + std::to_string(h.suchtiefe) + " m must be positive";
  for (int t = 0; t < h.nt; ++t)
    for (int k = 0; k < 3; ++k)
      if (h.dreiecke[3 * t + k] >= (uint32_t)h.n)
        return "triangle " + std::to_string(t) + " refers to a point >= " + std::to_string(h.n);
  struct Seite {
    uint64_t schluessel;
    uint32_t t;
    uint8_t k;
    bool r;
  };
  std::vector<Seite> seiten;
  seiten.reserve(3 * (size_t)h.nt);
  for (int t = 0; t < h.nt; ++t)
    for (uint8_t k = 0; k < 3; ++k) {
      const uint32_t i = h.dreiecke[3 * t + k], j = h.dreiecke[3 * t + (k + 1) % 3];
      const uint64_t lo = std::min(i, j), hi = std::max(i, j);
      seiten.push_back({(lo << 32) | hi, (uint32_t)t, k, i < j});
    }
  std::stable_sort(seiten.begin(), seiten.end(), [](const Seite& a, const Seite& b) { return a.schluessel < b.schluessel; });
  nachbar.assign(3 * (size_t)h.nt, 0u);
  h.randkanten = 0;
  for (size_t a = 0; a < seiten.size();) {
    size_t b = a;
    while (b < seiten.size() && seiten[b].schluessel == seiten[a].schluessel) ++b;
    const uint32_t i = (uint32_t)(seiten[a].schluessel >> 32), j = (uint32_t)seiten[a].schluessel;
    if (b - a == 1) {
      h.randkanten += 1;
      nachbar[3 * seiten[a].t + seiten[a].k] = seiten[a].t;
    } else if (b - a == 2) {
      if (seiten[a].r == seiten[a + 1].r)
        return "mesh not consistently oriented at the edge " + std::to_string(i) + "-" + std::to_string(j);
      nachbar[3 * seiten[a].t + seiten[a].k] = seiten[a + 1].t;
      nachbar[3 * seiten[a + 1].t + seiten[a + 1].k] = seiten[a].t;
    } else {
      return "edge " + std::to_string(i) + "-" + std::to_string(j) + " has " + std::to_string(b - a) + " triangles";
    }
    a = b;
  }
  inz_anfang.assign((size_t)h.n + 1, 0u);
  for (int t = 0; t < h.nt; ++t)
    for (int k = 0; k < 3; ++k) inz_anfang[h.dreiecke[3 * t + k] + 1] += 1;
  for (int i = 0; i < h.n; ++i) inz_anfang[i + 1] += inz_anfang[i];
  inz.assign(inz_anfang[h.n], 0u);
  std::vector<uint32_t> frei(inz_anfang.begin(), inz_anfang.end() - 1);
  for (int t = 0; t < h.nt; ++t)
    for (int k = 0; k < 3; ++k) inz[frei[h.dreiecke[3 * t + k]]++] = (uint32_t)(3 * t + k);
  return "";
}

// Enclosed volume of the pose, summed in triangle order (netzkoerper.rs:282).
static double netz_volumen(const Netzhost& h) {
  double vol = 0.0;
  for (int t = 0; t < h.nt; ++t) {
    const double* q0 = &h.punkte[3 * h.dreiecke[3 * t]];
    const double* q1 = &h.punkte[3 * h.dreiecke[3 * t + 1]];
    const double* q2 = &h.punkte[3 * h.dreiecke[3 * t + 2]];
    const double cx = q1[1] * q2[2] - q1[2] * q2[1], cy = q1[2] * q2[0] - q1[0] * q2[2], cz = q1[0] * q2[1] - q1[1] * q2[0];
    vol += (q0[0] * cx + q0[1] * cy + q0[2] * cz) / 6.0;
  }
  return vol;
}

// The grid raster of the pose (netzkoerper.rs:820-850): cells 2 x suchtiefe
// over the points' box widened by suchtiefe, joined to the previous raster
// (it never shrinks).
static void netz_raster(Netzhost& h) {
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = 0; i < h.n; ++i)
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::fmin(lo[a], h.punkte[3 * i + a]);
      hi[a] = std::fmax(hi[a], h.punkte[3 * i + a]);
    }
  for (int a = 0; a < 3; ++a) {
    h.huelle_lo[a] = lo[a];
    h.huelle_hi[a] = hi[a];
  }
  const double zelle = 2.0 * h.suchtiefe;
  double ursprung[3], ende[3];
  for (int a = 0; a < 3; ++a) {
    ursprung[a] = lo[a] - h.suchtiefe;
    ende[a] = hi[a] + h.suchtiefe;
  }
  if (h.hat_gitter)
    for (int a = 0; a < 3; ++a) {
      ursprung[a] = std::fmin(ursprung[a], h.bild.ursprung[a]);
      ende[a] = std::fmax(ende[a], h.bild.ursprung[a] + h.bild.zelle * (double)h.bild.zahl[a]);
    }
  for (int a = 0; a < 3; ++a) {
    const double z = std::ceil((ende[a] - ursprung[a]) / zelle);
    const int64_t zahl = z > 0.0 ? (int64_t)z : 0;  // `as usize`, then max(1)
    h.bild.zahl[a] = std::max<int64_t>(zahl, 1);
    h.bild.ursprung[a] = ursprung[a];
  }
  h.bild.zelle = zelle;
  h.hat_gitter = true;
}

template <typename X>
static int netz_alloc(gc_ctx* c, X** p, size_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  if (hipMalloc((void**)p, std::max<size_t>(1, count) * sizeof(X)) != hipSuccess) return fehler(c, "hipMalloc failed (mesh body)");
  return 0;
}

// Everything that depends on the points (netzkoerper.rs:257-291, 820-912):
// face normals, spheres, point normals, the grid of triangle lists and the
// grid of body points (each cell's entries ascending), built on the device
// from the pose in h.punkte.
static int netz_pose(gc_ctx* c, Netzhost& h) {
  const double vol = netz_volumen(h);
  if (!(vol > 0.0))
    return fehler(c, "volume " + std::to_string(vol) + " m^3 <= 0: the normals point inwards");
  h.volumen = vol;
  netz_raster(h);
  const int64_t nz = h.bild.zahl[0] * h.bild.zahl[1] * h.bild.zahl[2];
  if (nz <= 0 || nz > (int64_t)1 << 31) return fehler(c, "mesh body grid of " + std::to_string(nz) + " cells");
  hipStream_t st = c->stream;
  HIPCK(c, hipMemcpyAsync(h.punkte_d, h.punkte.data(), h.punkte.size() * sizeof(double), hipMemcpyHostToDevice, st));
  if ((size_t)nz + 1 > h.nz_kap) {
    if (netz_alloc(c, &h.anfang_d, (size_t)nz + 1) || netz_alloc(c, &h.zaehler_d, (size_t)nz) ||
        netz_alloc(c, &h.cursor_d, (size_t)nz) || netz_alloc(c, &h.p_anfang_d, (size_t)nz + 1))
      return 1;
    h.nz_kap = (size_t)nz + 1;
  }
  h.bild.anfang = h.anfang_d;
  h.bild.p_anfang = h.p_anfang_d;
  h.bild.p_eintraege = h.p_eintraege_d;
  h.bild.pv = h.pv_d;
  HIPCK(c, hipMemcpyAsync(h.dev, &h.bild, sizeof(NetzDev), hipMemcpyHostToDevice, st));
  const unsigned gt = (unsigned)std::min<int64_t>(4096, (h.nt + 255) / 256), gp = (unsigned)std::min<int64_t>(4096, (h.n + 255) / 256);
  hipLaunchKernelGGL(k_nk_flaechen, dim3(gt), dim3(256), 0, st, h.punkte_d, h.dreiecke_d, h.nt, h.fn_d, h.kugel_d);
  hipLaunchKernelGGL(k_nk_punktnormalen, dim3(gp), dim3(256), 0, st, h.punkte_d, h.dreiecke_d, h.fn_d, h.inz_anfang_d, h.inz_d,
                     h.n, h.pn_d);
  HIPCK(c, hipMemsetAsync(h.zaehler_d, 0, (size_t)nz * sizeof(uint32_t), st));
  hipLaunchKernelGGL(k_nk_zaehlen, dim3(gt), dim3(256), 0, st, (const NetzDev*)h.dev, h.punkte_d, h.dreiecke_d, h.nt, h.zaehler_d);
  hipLaunchKernelGGL(k_nk_scan, dim3(1), dim3(1024), 0, st, h.zaehler_d, nz, h.anfang_d);
  HIPCK(c, hipGetLastError());
  uint32_t gesamt = 0;
  HIPCK(c, hipMemcpyAsync(&gesamt, h.anfang_d + nz, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  HIPCK(c, hipStreamSynchronize(st));
  if (gesamt > h.e_kap || !h.eintraege_d) {
    if (netz_alloc(c, &h.roh_d, gesamt) || netz_alloc(c, &h.eintraege_d, gesamt)) return 1;
    h.e_kap = gesamt;
  }
  h.bild.eintraege = h.eintraege_d;
  HIPCK(c, hipMemcpyAsync(h.dev, &h.bild, sizeof(NetzDev), hipMemcpyHostToDevice, st));
  HIPCK(c, hipMemcpyAsync(h.cursor_d, h.anfang_d, (size_t)nz * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_nk_fuellen, dim3(gt), dim3(256), 0, st, (const NetzDev*)h.dev, h.punkte_d, h.dreiecke_d, h.nt, h.cursor_d,
                     h.roh_d);
  if (gesamt > 0)
    hipLaunchKernelGGL(k_nk_ordnen, dim3((unsigned)std::min<uint32_t>(4096, (gesamt + 255) / 256)), dim3(256), 0, st, h.anfang_d,
                       nz, h.roh_d, h.eintraege_d);
  // the point grid: count, scan, fill, order (netzkoerper.rs:862-879)
  HIPCK(c, hipMemsetAsync(h.zaehler_d, 0, (size_t)nz * sizeof(uint32_t), st));
  hipLaunchKernelGGL(k_nk_punkte_zaehlen, dim3(gp), dim3(256), 0, st, (const NetzDev*)h.dev, h.punkte_d, h.n, h.zaehler_d);
  hipLaunchKernelGGL(k_nk_scan, dim3(1), dim3(1024), 0, st, h.zaehler_d, nz, h.p_anfang_d);
  HIPCK(c, hipMemcpyAsync(h.cursor_d, h.p_anfang_d, (size_t)nz * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_nk_punkte_fuellen, dim3(gp), dim3(256), 0, st, (const NetzDev*)h.dev, h.punkte_d, h.n, h.cursor_d,
                     h.p_roh_d);
  hipLaunchKernelGGL(k_nk_ordnen, dim3(gp), dim3(256), 0, st, h.p_anfang_d, nz, h.p_roh_d, h.p_eintraege_d);
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipStreamSynchronize(st));
  return 0;
}

extern "C" {

int gc_add_netz(gc_ctx* c, const double* punkte, int32_t n_punkte, const uint32_t* dreiecke, int32_t n_dreiecke, double suchtiefe,
                const double versatz[3], const double v[3]) {
  if (!c) return -1;
  if (c->koerper.size() >= GC_MAX_KOERPER) {
    c->err = "too many bodies";
    return -1;
  }
  if (n_punkte < 0 || n_dreiecke < 0 || (n_punkte > 0 && !punkte) || (n_dreiecke > 0 && !dreiecke)) {
    c->err = "gc_add_netz: bad arrays";
    return -1;
  }
  auto h = std::make_shared<Netzhost>();
  h->n = n_punkte;
  h->nt = n_dreiecke;
  h->suchtiefe = suchtiefe;
  h->dreiecke.assign(dreiecke, dreiecke + 3 * (size_t)n_dreiecke);
  h->punkte.assign(punkte, punkte + 3 * (size_t)n_punkte);
  std::vector<uint32_t> nachbar, inz_anfang, inz;
  const std::string f = netz_topologie(*h, nachbar, inz_anfang, inz);
  if (!f.empty()) {
    c->err = f;
    return -1;
  }
  const size_t n = (size_t)h->n, nt = (size_t)h->nt;
  if (netz_alloc(c, &h->punkte_d, 3 * n) || netz_alloc(c, &h->fn_d, 3 * nt) || netz_alloc(c, &h->pn_d, 3 * n) ||
      netz_alloc(c, &h->kugel_d, 4 * nt) || netz_alloc(c, &h->dreiecke_d, 3 * nt) || netz_alloc(c, &h->nachbar_d, 3 * nt) ||
      netz_alloc(c, &h->inz_anfang_d, n + 1) || netz_alloc(c, &h->inz_d, inz.size()) || netz_alloc(c, &h->dev, 1) ||
      netz_alloc(c, &h->p_roh_d, n) || netz_alloc(c, &h->p_eintraege_d, n)) {
    h->frei();
    return -1;
  }
  hipStream_t st = c->stream;
  if (hipMemcpyAsync(h->dreiecke_d, h->dreiecke.data(), 3 * nt * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(h->nachbar_d, nachbar.data(), 3 * nt * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      hipMemcpyAsync(h->inz_anfang_d, inz_anfang.data(), (n + 1) * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
      (!inz.empty() && hipMemcpyAsync(h->inz_d, inz.data(), inz.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess)) {
    c->err = "hipMemcpy failed (mesh body)";
    h->frei();
    return -1;
  }
  NetzDev& b = h->bild;
  b.punkte = h->punkte_d;
  b.dreiecke = h->dreiecke_d;
  b.fn = h->fn_d;
  b.nachbar = h->nachbar_d;
  b.pn = h->pn_d;
  b.kugel = h->kugel_d;
  b.pv = nullptr;
  b.suchtiefe = suchtiefe;
  b.n_punkte = n_punkte;
  b.n_dreiecke = n_dreiecke;
  if (netz_pose(c, *h)) {
    h->frei();
    return -1;
  }
  Koerperhost kh;
  std::memset(&kh.k, 0, sizeof(Koerper));
  kh.k.art = GC_NETZ;
  kh.k.netz = h->dev;
  for (int q = 0; q < 3; ++q) {
    kh.k.punkt[q] = versatz[q];
    kh.k.geschwindigkeit[q] = v[q];
  }
  kh.netz = h;
  c->koerper.push_back(kh);
  folge_neu(c);  // a mesh body adds the passes of body points to the substep
  alle_param(c);
  return (int)c->koerper.size() - 1;
}

static Netzhost* netz_von(gc_ctx* c, int32_t k) {
  if (!c || k < 0 || k >= (int32_t)c->koerper.size() || !c->koerper[(size_t)k].netz) {
    if (c) c->err = "body " + std::to_string(k) + " is not a mesh body";
    return nullptr;
  }
  return c->koerper[(size_t)k].netz.get();
}

int gc_netz_nachfuehren(gc_ctx* c, int32_t k, const double* punkte) {
  Netzhost* h = netz_von(c, k);
  if (!h) return 1;
  std::vector<double> alt = std::move(h->punkte);
  h->punkte.assign(punkte, punkte + 3 * (size_t)h->n);
  if (netz_volumen(*h) <= 0.0 || !(netz_volumen(*h) > 0.0)) {
    h->punkte = std::move(alt);
    return fehler(c, "volume <= 0 after the deformation: the mesh has turned inside out");
  }
  if (h->pv_d) {  // the field belonged to the old pose (netzkoerper.rs:409)
    (void)hipFree(h->pv_d);
    h->pv_d = nullptr;
  }
  return netz_pose(c, *h);
}

int gc_netz_punktgeschwindigkeit(gc_ctx* c, int32_t k, const double* v, int32_t n) {
  Netzhost* h = netz_von(c, k);
  if (!h) return 1;
  if (n != 0 && n != h->n)
    return fehler(c, "velocity for " + std::to_string(n) + " points, the mesh has " + std::to_string(h->n));
  if (n == 0) {
    if (h->pv_d) (void)hipFree(h->pv_d);
    h->pv_d = nullptr;
  } else {
    if (!h->pv_d && netz_alloc(c, &h->pv_d, 3 * (size_t)n)) return 1;
    HIPCK(c, hipMemcpyAsync(h->pv_d, v, 3 * (size_t)n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  h->bild.pv = h->pv_d;
  HIPCK(c, hipMemcpyAsync(h->dev, &h->bild, sizeof(NetzDev), hipMemcpyHostToDevice, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gc_netz_info(gc_ctx* c, int32_t k, gc_netz_angaben* out) {
  Netzhost* h = netz_von(c, k);
  if (!h || !out) return 1;
  out->n_punkte = h->n;
  out->n_dreiecke = h->nt;
  out->randkanten = h->randkanten;
  out->volumen = h->volumen;
  for (int a = 0; a < 3; ++a) {
    out->ursprung[a] = h->bild.ursprung[a];
    out->zahl[a] = h->bild.zahl[a];
  }
  out->zelle = h->bild.zelle;
  uint32_t gesamt = 0;
  const int64_t nz = h->bild.zahl[0] * h->bild.zahl[1] * h->bild.zahl[2];
  HIPCK(c, hipMemcpy(&gesamt, h->anfang_d + nz, sizeof(uint32_t), hipMemcpyDeviceToHost));
  out->n_eintraege = gesamt;
  return 0;
}

int gc_netz_lesen(gc_ctx* c, int32_t k, double* flaechennormale, double* punktnormale, uint32_t* nachbar, double* kugel,
                  uint32_t* anfang, uint32_t* eintraege) {
  Netzhost* h = netz_von(c, k);
  if (!h) return 1;
  HIPCK(c, hipStreamSynchronize(c->stream));
  const size_t n = (size_t)h->n, nt = (size_t)h->nt;
  const int64_t nz = h->bild.zahl[0] * h->bild.zahl[1] * h->bild.zahl[2];
  uint32_t gesamt = 0;
  HIPCK(c, hipMemcpy(&gesamt, h->anfang_d + nz, sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (flaechennormale) HIPCK(c, hipMemcpy(flaechennormale, h->fn_d, 3 * nt * sizeof(double), hipMemcpyDeviceToHost));
  if (punktnormale) HIPCK(c, hipMemcpy(punktnormale, h->pn_d, 3 * n * sizeof(double), hipMemcpyDeviceToHost));
  if (nachbar) HIPCK(c, hipMemcpy(nachbar, h->nachbar_d, 3 * nt * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (kugel) HIPCK(c, hipMemcpy(kugel, h->kugel_d, 4 * nt * sizeof(double), hipMemcpyDeviceToHost));
  if (anfang) HIPCK(c, hipMemcpy(anfang, h->anfang_d, ((size_t)nz + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  if (eintraege && gesamt) HIPCK(c, hipMemcpy(eintraege, h->eintraege_d, (size_t)gesamt * sizeof(uint32_t), hipMemcpyDeviceToHost));
  return 0;
}

}  // extern "C"

// Device scratch for one query call (host arrays in and out).
struct NetzAbfrage {
  std::vector<void*> p;
  ~NetzAbfrage() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename X>
  X* neu(size_t count) {
    void* q = nullptr;
    if (hipMalloc(&q, std::max<size_t>(1, count) * sizeof(X)) != hipSuccess) return nullptr;
    p.push_back(q);
    return (X*)q;
  }
};

extern "C" {

int gc_netz_naechster(gc_ctx* c, int32_t k, int32_t n, const double* x, double radius, int32_t* da, double* abstand,
                      double* punkt, double* normale, int32_t* dreieck, int32_t* region) {
  Netzhost* h = netz_von(c, k);
  if (!h) return 1;
  if (n <= 0) return 0;
  const size_t m = (size_t)n;
  NetzAbfrage q;
  double *xd = q.neu<double>(3 * m), *ad = q.neu<double>(m), *pd = q.neu<double>(3 * m), *nd = q.neu<double>(3 * m);
  int32_t *dd = q.neu<int32_t>(m), *td = q.neu<int32_t>(m), *rd = q.neu<int32_t>(m);
  if (!xd || !ad || !pd || !nd || !dd || !td || !rd) return fehler(c, "hipMalloc failed (mesh query)");
  hipStream_t st = c->stream;
  HIPCK(c, hipMemcpyAsync(xd, x, 3 * m * sizeof(double), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_nk_naechster, dim3((unsigned)std::min<size_t>(4096, (m + 255) / 256)), dim3(256), 0, st,
                     (const NetzDev*)h->dev, n, xd, radius, dd, ad, pd, nd, td, rd);
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipMemcpyAsync(da, dd, m * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (abstand) HIPCK(c, hipMemcpyAsync(abstand, ad, m * sizeof(double), hipMemcpyDeviceToHost, st));
  if (punkt) HIPCK(c, hipMemcpyAsync(punkt, pd, 3 * m * sizeof(double), hipMemcpyDeviceToHost, st));
  if (normale) HIPCK(c, hipMemcpyAsync(normale, nd, 3 * m * sizeof(double), hipMemcpyDeviceToHost, st));
  if (dreieck) HIPCK(c, hipMemcpyAsync(dreieck, td, m * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (region) HIPCK(c, hipMemcpyAsync(region, rd, m * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCK(c, hipStreamSynchronize(st));
  return 0;
}

int gc_netz_durchstoss(gc_ctx* c, int32_t k, int32_t n, const double* a, const double* b, const double starr[3], int32_t* da,
                       double* p, double* normale, double* vk) {
  Netzhost* h = netz_von(c, k);
  if (!h) return 1;
  if (n <= 0) return 0;
  const size_t m = (size_t)n;
  NetzAbfrage q;
  double *ad = q.neu<double>(3 * m), *bd = q.neu<double>(3 * m), *pd = q.neu<double>(3 * m), *nd = q.neu<double>(3 * m),
         *vd = q.neu<double>(3 * m);
  int32_t* dd = q.neu<int32_t>(m);
  if (!ad || !bd || !pd || !nd || !vd || !dd) return fehler(c, "hipMalloc failed (mesh query)");
  hipStream_t st = c->stream;
  HIPCK(c, hipMemcpyAsync(ad, a, 3 * m * sizeof(double), hipMemcpyHostToDevice, st));
  HIPCK(c, hipMemcpyAsync(bd, b, 3 * m * sizeof(double), hipMemcpyHostToDevice, st));
  const V3<double> sv = {starr[0], starr[1], starr[2]};
  hipLaunchKernelGGL(k_nk_durchstoss, dim3((unsigned)std::min<size_t>(4096, (m + 255) / 256)), dim3(256), 0, st,
                     (const NetzDev*)h->dev, n, ad, bd, sv, dd, pd, nd, vd);
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipMemcpyAsync(da, dd, m * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (p) HIPCK(c, hipMemcpyAsync(p, pd, 3 * m * sizeof(double), hipMemcpyDeviceToHost, st));
  if (normale) HIPCK(c, hipMemcpyAsync(normale, nd, 3 * m * sizeof(double), hipMemcpyDeviceToHost, st));
  if (vk) HIPCK(c, hipMemcpyAsync(vk, vd, 3 * m * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCK(c, hipStreamSynchronize(st));
  return 0;
}

int gc_netz_sdf_backen(gc_ctx* c, int32_t k, const int32_t dims[3], const double origin[3], double spacing, double band,
                       float* werte) {
  Netzhost* h = netz_von(c, k);
  if (!h) return 1;
  if (dims[0] < 2 || dims[1] < 2 || dims[2] < 2 || !(spacing > 0.0))
    return fehler(c, "gc_netz_sdf_backen: at least 2 nodes per axis and a positive spacing");
  if (!(band >= spacing)) return fehler(c, "gc_netz_sdf_backen: the band must be at least one spacing wide");
  const int64_t anzahl = (int64_t)dims[0] * dims[1] * dims[2];
  NetzAbfrage q;
  float* wd = q.neu<float>((size_t)anzahl);
  if (!wd) return fehler(c, "hipMalloc failed (SDF bake)");
  hipStream_t st = c->stream;
  const V3<double> o = {origin[0], origin[1], origin[2]};
  hipLaunchKernelGGL(k_nk_backen_band, dim3((unsigned)std::min<int64_t>(65536, (anzahl + 255) / 256)), dim3(256), 0, st,
                     (const NetzDev*)h->dev, o, spacing, dims[0], dims[1], anzahl, band, wd);
  const int64_t linien = (int64_t)dims[1] * dims[2];
  const V3<double> lo = {h->huelle_lo[0], h->huelle_lo[1], h->huelle_lo[2]}, hi = {h->huelle_hi[0], h->huelle_hi[1], h->huelle_hi[2]};
  hipLaunchKernelGGL(k_nk_backen_seiten, dim3((unsigned)std::min<int64_t>(65536, (linien + 127) / 128)), dim3(128), 0, st,
                     (const NetzDev*)h->dev, o, spacing, dims[0], dims[1], dims[2], band, lo, hi, wd);
  HIPCK(c, hipGetLastError());
  HIPCK(c, hipMemcpyAsync(werte, wd, (size_t)anzahl * sizeof(float), hipMemcpyDeviceToHost, st));
  HIPCK(c, hipStreamSynchronize(st));
  return 0;
}

// --- XPBD path (cpu.rs:76-246) ------------------------------------------------
}  // extern "C"

// Consecutive runs of a constraint list in which no two constraints share a
// point: a run closes before the first constraint that touches a point of it.
template <int N>
static std::vector<int> laeufe(const uint32_t* p, int anzahl, int n) {
  std::vector<int> start{0};
  std::vector<int> marke((size_t)n, -1);
  int lauf = 0;
  for (int c = 0; c < anzahl; ++c) {
    bool teilt = false;
    for (int k = 0; k < N; ++k) teilt = teilt || marke[p[N * (size_t)c + k]] == lauf;
    if (teilt) {
      start.push_back(c);
      ++lauf;
    }
    for (int k = 0; k < N; ++k) marke[p[N * (size_t)c + k]] = lauf;
  }
  start.push_back(anzahl);
  return start;
}

template <typename T>
static int xpbd_bild(gc_ctx* c, double bild_dt, int unterschritte, gc_xpbd_statistik* out) {
  XpbdHost& x = c->xp;
  SysDev<T>& s = sys<T>(c);
  const int n = unterschritte > 1 ? unterschritte : 1;
  const double h = bild_dt / (double)n, h2 = h * h;
  const V3<double> g = {c->param.schwerkraft[0], c->param.schwerkraft[1], c->param.schwerkraft[2]};
  const double daempf = fmax(1.0 - c->param.daempfung * h, 0.0);
  hipStream_t st = c->stream;
  const size_t np = (size_t)c->n;
  HIPCK(c, hipMemsetAsync(x.delta, 0, np * sizeof(V3<double>), st));
  HIPCK(c, hipMemsetAsync(x.korr, 0, np * sizeof(V3<double>), st));
  HIPCK(c, hipMemsetAsync(x.last, 0, np * sizeof(V3<double>), st));
  HIPCK(c, hipMemsetAsync(x.st, 0, sizeof(XpStatDev), st));
  const unsigned gp = gitter(c->n);
  auto zahl = [](int lo, int hi) { return dim3((unsigned)std::max(1, std::min(4096, (hi - lo + 255) / 256))); };
  std::vector<Koerper> anfang;
  for (auto& k : c->koerper) anfang.push_back(k.k);
  for (int k = 0; k < n; ++k) {
    const double weg = (double)(k + 1) * h;  // koerper.rs:122-133 from the frame start
    for (size_t b = 0; b < c->koerper.size(); ++b)
      for (int q = 0; q < 3; ++q) c->koerper[b].k.punkt[q] = anfang[b].punkt[q] + anfang[b].geschwindigkeit[q] * weg;
    alle_param(c);
    if (sys_hoch<T>(c)) return 1;
    if (x.na > 0)
      hipLaunchKernelGGL(k_xp_zug_abstand, zahl(0, x.na), dim3(256), 0, st, x.na, x.a_p, x.a_ruhe, x.a_alpha, s.x, x.delta, x.fa,
                         x.da);
    if (x.nm > 0)
      hipLaunchKernelGGL(k_xp_zug_membran, zahl(0, x.nm), dim3(256), 0, st, x.nm, s.m_p, s.m_el64, s.x, x.delta, x.korr, x.fm);
    hipLaunchKernelGGL(k_xp_last, dim3(gp), dim3(256), 0, st, c->n, s.masse, g, x.a_off, x.a_liste, x.fa, x.da, x.m_off,
                       x.m_liste, x.fm, x.last);
    hipLaunchKernelGGL(k_xp_vorhersage, dim3(gp), dim3(256), 0, st, c->n, s.w, g, h, daempf, s.v, x.delta, x.korr);
    for (size_t r = 0; r + 1 < x.a_laeufe.size(); ++r) {
      const int lo = x.a_laeufe[r], hi = x.a_laeufe[r + 1];
      hipLaunchKernelGGL(k_xp_abstand, zahl(lo, hi), dim3(256), 0, st, lo, hi, x.a_p, x.a_ruhe, (const double*)x.a_alpha, 0.0, h2,
                         s.x, x.delta, x.korr, s.w);
    }
    if (x.stufen_einzeln) {
      hipLaunchKernelGGL(k_xp_membran_stufen, dim3(1), dim3(1024), 0, st, (int)x.m_stufen.size() - 1, x.stufen_d, x.m_reihe, s.m_p,
                         s.m_el64, h2, s.x, x.delta, x.korr, s.w);
    } else {
      for (size_t r = 0; r + 1 < x.m_stufen.size(); ++r) {
        const int lo = x.m_stufen[r], hi = x.m_stufen[r + 1];
        hipLaunchKernelGGL(k_xp_membran, zahl(lo, hi), dim3(256), 0, st, lo, hi, x.m_reihe, s.m_p, s.m_el64, h2, s.x, x.delta,
                           x.korr, s.w);
      }
    }
    for (size_t r = 0; r + 1 < x.n_laeufe.size(); ++r) {
      const int lo = x.n_laeufe[r], hi = x.n_laeufe[r + 1];
      hipLaunchKernelGGL(k_xp_abstand, zahl(lo, hi), dim3(256), 0, st, lo, hi, x.n_p, x.n_ruhe, (const double*)nullptr, x.n_alpha,
                         h2, s.x, x.delta, x.korr, s.w);
    }
    for (size_t r = 0; r + 1 < x.b_laeufe.size(); ++r) {
      const int lo = x.b_laeufe[r], hi = x.b_laeufe[r + 1];
      hipLaunchKernelGGL(k_xp_biegung, zahl(lo, hi), dim3(256), 0, st, lo, hi, x.b_p, x.b_ruhe, x.b_alpha, h2, s.x, x.delta, x.korr,
                         s.w);
    }
    hipLaunchKernelGGL(k_xp_kontakt<T>, dim3(gp), dim3(256), 0, st, dsys<T>(c), h, h2, x.delta, x.korr, x.last, x.st);
    hipLaunchKernelGGL(k_xp_ende, dim3(gp), dim3(256), 0, st, c->n, s.w, h, s.v, x.delta, x.korr);
    HIPCK(c, hipGetLastError());
  }
  hipLaunchKernelGGL(k_xp_lage, dim3(gp), dim3(256), 0, st, c->n, s.x, x.delta);
  HIPCK(c, hipGetLastError());
  XpStatDev stat{};
  HIPCK(c, hipMemcpyAsync(&stat, x.st, sizeof(XpStatDev), hipMemcpyDeviceToHost, st));
  HIPCK(c, hipStreamSynchronize(st));
  if (out) {
    out->kontakte = (int64_t)stat.kontakte;
    out->haftend = (int64_t)stat.haftend;
    out->tiefe = stat.tiefe;
    out->tangential = stat.tangential;
  }
  return 0;
}

extern "C" {

int gc_xpbd_topology(gc_ctx* c, const gc_xpbd_topologie* t) {
  if (!c || !t) return 1;
  if (!c->topo) return fehler(c, "gc_xpbd_topology: the context's topology must be set first");
  if (c->xp.bereit) return fehler(c, "gc_xpbd_topology: already set");
  const int n = c->n;
  auto pruefe = [&](const uint32_t* idx, int anzahl, int je) {
    for (size_t q = 0; q < (size_t)anzahl * je; ++q)
      if (idx[q] >= (uint32_t)n) return false;
    return true;
  };
  if (t->n_abstand < 0 || t->n_naehte < 0 || t->n_biegungen < 0 || !pruefe(t->abstand_punkte, t->n_abstand, 2) ||
      !pruefe(t->naht_punkte, t->n_naehte, 2) || !pruefe(t->biegung_punkte, t->n_biegungen, 4))
    return fehler(c, "gc_xpbd_topology: a constraint refers to a point out of range");
  XpbdHost& x = c->xp;
  // distance constraints without the ones a membrane carries (cpu.rs:105-109, 148-152)
  std::vector<uint32_t> ap;
  std::vector<double> ar, aa;
  for (int e = 0; e < t->n_abstand; ++e) {
    if (t->membran_kante && t->membran_kante[e]) continue;
    ap.push_back(t->abstand_punkte[2 * e]);
    ap.push_back(t->abstand_punkte[2 * e + 1]);
    ar.push_back(t->abstand_ruhe[e]);
    aa.push_back(t->abstand_nachgiebigkeit[e]);
  }
  x.na = (int)ar.size();
  x.nn = t->n_naehte;
  x.nb = t->n_biegungen;
  x.nm = (int)(c->m_p.size() / 3);
  x.a_laeufe = laeufe<2>(ap.data(), x.na, n);
  x.n_laeufe = laeufe<2>(t->naht_punkte, x.nn, n);
  x.b_laeufe = laeufe<4>(t->biegung_punkte, x.nb, n);
  // membrane elements by dependency level (cpu.rs:123-135, 160-189)
  std::vector<int> stufe((size_t)x.nm), zuletzt((size_t)n, -1);
  int stufen = 0;
  for (int e = 0; e < x.nm; ++e) {
    int l = 0;
    for (int k = 0; k < 3; ++k) l = std::max(l, zuletzt[c->m_p[3 * (size_t)e + k]] + 1);
    stufe[e] = l;
    for (int k = 0; k < 3; ++k) zuletzt[c->m_p[3 * (size_t)e + k]] = l;
    stufen = std::max(stufen, l + 1);
  }
  x.m_stufen.assign((size_t)stufen + 1, 0);
  for (int e = 0; e < x.nm; ++e) x.m_stufen[stufe[e] + 1] += 1;
  for (int l = 0; l < stufen; ++l) x.m_stufen[l + 1] += x.m_stufen[l];
  std::vector<uint32_t> reihe((size_t)x.nm);
  std::vector<int> frei(x.m_stufen.begin(), x.m_stufen.end() - 1);
  for (int e = 0; e < x.nm; ++e) reihe[frei[stufe[e]]++] = (uint32_t)e;
  // per-point term lists for the load (cpu.rs:97-135)
  std::vector<uint32_t> a_off((size_t)n + 1, 0u), m_off((size_t)n + 1, 0u);
  for (size_t q = 0; q < ap.size(); ++q) a_off[ap[q] + 1] += 1;
  for (size_t q = 0; q < 3 * (size_t)x.nm; ++q) m_off[c->m_p[q] + 1] += 1;
  for (int i = 0; i < n; ++i) {
    a_off[i + 1] += a_off[i];
    m_off[i + 1] += m_off[i];
  }
  std::vector<uint32_t> a_liste(a_off[n]), m_liste(m_off[n]);
  {
    std::vector<uint32_t> fa_(a_off.begin(), a_off.end() - 1), fm_(m_off.begin(), m_off.end() - 1);
    for (size_t q = 0; q < ap.size(); ++q) a_liste[fa_[ap[q]]++] = (uint32_t)q;  // q = 2c + side
    for (size_t q = 0; q < 3 * (size_t)x.nm; ++q) m_liste[fm_[c->m_p[q]]++] = (uint32_t)q;  // q = 3t + corner
  }
  x.n_alpha = t->naht_nachgiebigkeit;
  x.a_p = dalloc<uint32_t>(c, ap.size());
  x.a_ruhe = dalloc<double>(c, ar.size());
  x.a_alpha = dalloc<double>(c, aa.size());
  x.n_p = dalloc<uint32_t>(c, 2 * (size_t)x.nn);
  x.n_ruhe = dalloc<double>(c, (size_t)x.nn);
  x.b_p = dalloc<uint32_t>(c, 4 * (size_t)x.nb);
  x.b_ruhe = dalloc<double>(c, (size_t)x.nb);
  x.b_alpha = dalloc<double>(c, (size_t)x.nb);
  x.m_reihe = dalloc<uint32_t>(c, reihe.size());
  x.delta = dalloc<V3<double>>(c, (size_t)n);
  x.korr = dalloc<V3<double>>(c, (size_t)n);
  x.last = dalloc<V3<double>>(c, (size_t)n);
  x.st = dalloc<XpStatDev>(c, 1);
  x.a_off = dalloc<uint32_t>(c, a_off.size());
  x.a_liste = dalloc<uint32_t>(c, a_liste.size());
  x.m_off = dalloc<uint32_t>(c, m_off.size());
  x.m_liste = dalloc<uint32_t>(c, m_liste.size());
  x.fa = dalloc<V3<double>>(c, (size_t)x.na);
  x.da = dalloc<uint8_t>(c, (size_t)x.na);
  x.fm = dalloc<V3<double>>(c, 3 * (size_t)x.nm);
  if (!x.a_off || !x.a_liste || !x.m_off || !x.m_liste || !x.fa || !x.da || !x.fm) return fehler(c, "hipMalloc failed (XPBD)");
  // one block walks all levels when they average at most 128 elements: its
  // barrier then costs less than a launch per level. Larger levels need the
  // whole chip (a substep of the 40k drape, 90 levels of ~900 elements,
  // takes 3.5 ms with them in one block, 1.1 ms with a launch per level).
  const int nstufen = (int)x.m_stufen.size() - 1;
  x.stufen_einzeln = nstufen > 0 && x.nm <= 128 * nstufen;
  x.stufen_d = dalloc<int>(c, x.m_stufen.size());
  if (!x.stufen_d || hoch(c, x.stufen_d, x.m_stufen.data(), x.m_stufen.size())) return fehler(c, "hipMalloc failed (XPBD)");
  if (hoch(c, x.a_off, a_off.data(), a_off.size()) || hoch(c, x.a_liste, a_liste.data(), a_liste.size()) ||
      hoch(c, x.m_off, m_off.data(), m_off.size()) || hoch(c, x.m_liste, m_liste.data(), m_liste.size()))
    return 1;
  if (!x.a_p || !x.a_ruhe || !x.a_alpha || !x.n_p || !x.n_ruhe || !x.b_p || !x.b_ruhe || !x.b_alpha || !x.m_reihe || !x.delta ||
      !x.korr || !x.last || !x.st)
    return fehler(c, "hipMalloc failed (XPBD)");
  if (hoch(c, x.a_p, ap.data(), ap.size()) || hoch(c, x.a_ruhe, ar.data(), ar.size()) || hoch(c, x.a_alpha, aa.data(), aa.size()) ||
      hoch(c, x.n_p, t->naht_punkte, 2 * (size_t)x.nn) || hoch(c, x.n_ruhe, t->naht_ruhe, (size_t)x.nn) ||
      hoch(c, x.b_p, t->biegung_punkte, 4 * (size_t)x.nb) || hoch(c, x.b_ruhe, t->biegung_ruhe, (size_t)x.nb) ||
      hoch(c, x.b_alpha, t->biegung_nachgiebigkeit, (size_t)x.nb) || hoch(c, x.m_reihe, reihe.data(), reihe.size()))
    return 1;
  HIPCK(c, hipStreamSynchronize(c->stream));
  x.bereit = true;
  return 0;
}

int gc_xpbd_bild(gc_ctx* c, double bild_dt, int32_t unterschritte, gc_xpbd_statistik* out) {
  if (!c) return 1;
  if (!c->xp.bereit || !c->hat_param) return fehler(c, "gc_xpbd_bild: XPBD topology and parameters must be set first");
  if (!(bild_dt > 0.0)) return fehler(c, "gc_xpbd_bild: bild_dt must be > 0");
  return c->opt.praezision == GC_FP32 ? xpbd_bild<float>(c, bild_dt, unterschritte, out)
                                      : xpbd_bild<double>(c, bild_dt, unterschritte, out);
}

int32_t gc_xpbd_laeufe(gc_ctx* c, int32_t* abstand, int32_t* naehte, int32_t* biegungen, int32_t* membran_stufen) {
  if (!c || !c->xp.bereit) return 1;
  if (abstand) *abstand = (int32_t)c->xp.a_laeufe.size() - 1;
  if (naehte) *naehte = (int32_t)c->xp.n_laeufe.size() - 1;
  if (biegungen) *biegungen = (int32_t)c->xp.b_laeufe.size() - 1;
  if (membran_stufen) *membran_stufen = (int32_t)c->xp.m_stufen.size() - 1;
  return 0;
}

int gc_set_koerper_lage(gc_ctx* c, int32_t k, const double lage[3]) {
  if (!c || k < 0 || k >= (int)c->koerper.size()) return 1;
  for (int q = 0; q < 3; ++q) c->koerper[k].k.punkt[q] = lage[q];
  alle_param(c);
  folge_neu(c);
  return 0;
}

int gc_set_state(gc_ctx* c, const double* x, const double* v) {
  if (!c || !c->topo) return 1;
  V3<double>* dx = c->opt.praezision == GC_FP32 ? c->sf.x : c->sd.x;
  V3<double>* dv = c->opt.praezision == GC_FP32 ? c->sf.v : c->sd.v;
  if (x && hoch(c, dx, (const V3<double>*)x, c->n)) return 1;
  if (v && hoch(c, dv, (const V3<double>*)v, c->n)) return 1;
  HIPCK(c, hipStreamSynchronize(c->stream));
  c->built = false;
  return 0;
}

int gc_get_state(gc_ctx* c, double* x, double* v) {
  if (!c || !c->topo) return 1;
  V3<double>* dx = c->opt.praezision == GC_FP32 ? c->sf.x : c->sd.x;
  V3<double>* dv = c->opt.praezision == GC_FP32 ? c->sf.v : c->sd.v;
  if (x) HIPCK(c, hipMemcpyAsync(x, dx, c->n * sizeof(V3<double>), hipMemcpyDeviceToHost, c->stream));
  if (v) HIPCK(c, hipMemcpyAsync(v, dv, c->n * sizeof(V3<double>), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gc_set_dv_vorher(gc_ctx* c, const double* dv, int32_t n) {
  if (!c || !c->topo) return 1;
  if (n == 0) {
    c->hat_vorher = false;
    alle_param(c);
    return 0;
  }
  if (n != c->n) return fehler(c, "gc_set_dv_vorher: length must be n or 0");
  V3<double>* d = c->opt.praezision == GC_FP32 ? c->sf.dv_vorher : c->sd.dv_vorher;
  if (hoch(c, d, (const V3<double>*)dv, c->n)) return 1;
  HIPCK(c, hipStreamSynchronize(c->stream));
  c->hat_vorher = true;
  alle_param(c);
  return 0;
}

int gc_build(gc_ctx* c, double h) {
  if (!c || !c->topo || !c->hat_param) return c ? fehler(c, "gc_build: topology and parameters must be set first") : 1;
  if (!(h > 0.0)) return fehler(c, "h must be > 0");
  int rc;
  if (c->opt.praezision == GC_FP32) {
    c->sf.h = h;
    c->sf.h2 = h * h;
    rc = selbst_vorbereiten<float>(c) || sys_hoch<float>(c) || launch_bau<float>(c);
  } else {
    c->sd.h = h;
    c->sd.h2 = h * h;
    rc = selbst_vorbereiten<double>(c) || sys_hoch<double>(c) || launch_bau<double>(c);
  }
  if (rc) return rc;
  HIPCK(c, hipStreamSynchronize(c->stream));
  c->built = true;
  return pruefe_gitter(c);
}

int gc_set_fest(gc_ctx* c, const int32_t* flag, const double* normale, const double* vn_soll, const double* vk) {
  if (!c || !c->topo || !flag || !normale || !vn_soll || !vk) return 1;
  return c->opt.praezision == GC_FP32 ? fest_setzen<float>(c, flag, normale, vn_soll, vk) : fest_setzen<double>(c, flag, normale, vn_soll, vk);
}

int gc_get_fest(gc_ctx* c, int32_t* flag, double* normale, double* vn_soll, double* vk) {
  if (!c || !c->topo) return 1;
  return c->opt.praezision == GC_FP32 ? fest_lesen<float>(c, flag, normale, vn_soll, vk) : fest_lesen<double>(c, flag, normale, vn_soll, vk);
}

int gc_solve(gc_ctx* c, double* dv, gc_cg_lauf* lauf) {
  if (!c || !dv) return 1;
  return c->opt.praezision == GC_FP32 ? loese<float>(c, dv, lauf) : loese<double>(c, dv, lauf);
}

int gc_solve_ohne_speicher(gc_ctx* c, double* dv, gc_cg_lauf* lauf) {
  if (!c || !dv) return 1;
  return c->opt.praezision == GC_FP32 ? loese_ohne_speicher<float>(c, dv, lauf) : loese_ohne_speicher<double>(c, dv, lauf);
}

int gc_apply_A(gc_ctx* c, const double* p, double* out) {
  if (!c || !p || !out) return 1;
  return c->opt.praezision == GC_FP32 ? mal<float>(c, p, out, 0) : mal<double>(c, p, out, 0);
}

int gc_apply_H(gc_ctx* c, const double* p, double* out) {
  if (!c || !p || !out) return 1;
  return c->opt.praezision == GC_FP32 ? mal<float>(c, p, out, 1) : mal<double>(c, p, out, 1);
}

int gc_residual(gc_ctx* c, const double* dv, double* out) {
  if (!c || !dv || !out) return 1;
  if (gc_apply_A(c, dv, out)) return 1;
  std::vector<double> b(3 * (size_t)c->n);
  if (gc_read_system(c, b.data(), nullptr, nullptr, nullptr)) return 1;
  for (size_t q = 0; q < b.size(); ++q) out[q] = out[q] - b[q];
  return 0;
}

int gc_read_load(gc_ctx* c, double* out) {
  if (!c || !out) return 1;
  return gc_read_system(c, nullptr, nullptr, out, nullptr);
}

int gc_read_system(gc_ctx* c, double* rhs, double* f, double* last, double* pinv9) {
  if (!c) return 1;
  return c->opt.praezision == GC_FP32 ? system_lesen<float>(c, rhs, f, last, pinv9) : system_lesen<double>(c, rhs, f, last, pinv9);
}

int gc_read_records(gc_ctx* c, double* richtung, double* quer, double* grad) {
  if (!c) return 1;
  return c->opt.praezision == GC_FP32 ? records_lesen<float>(c, richtung, quer, grad) : records_lesen<double>(c, richtung, quer, grad);
}

int gc_step(gc_ctx* c, double h, int32_t n) {
  if (!c) return 1;
  if (!(h > 0.0) || n < 0) return fehler(c, "gc_step: h must be > 0 and n >= 0");
  return c->opt.praezision == GC_FP32 ? schritte<float>(c, h, n) : schritte<double>(c, h, n);
}

int gc_step_graph(gc_ctx* c, double h, int32_t n) {
  if (!c) return 1;
  if (!(h > 0.0) || n < 0) return fehler(c, "gc_step_graph: h must be > 0 and n >= 0");
  return c->opt.praezision == GC_FP32 ? schritte_graph<float>(c, h, n) : schritte_graph<double>(c, h, n);
}

int gc_bild(gc_ctx* c, double bild_dt, int32_t schritte_je_bild) {
  if (!c) return 1;
  const int n = std::max(1, schritte_je_bild);
  const double h = bild_dt / (double)n;
  std::vector<Koerper> anfang;
  for (auto& k : c->koerper) anfang.push_back(k.k);
  for (int k = 0; k < n; ++k) {
    const double weg = (double)(k + 1) * h;
    for (size_t b = 0; b < c->koerper.size(); ++b)
      for (int q = 0; q < 3; ++q) c->koerper[b].k.punkt[q] = anfang[b].punkt[q] + anfang[b].geschwindigkeit[q] * weg;
    alle_param(c);
    if (gc_step(c, h, 1)) return 1;
  }
  return 0;
}

int gc_read_statistik(gc_ctx* c, gc_statistik* out, int32_t reset) {
  if (!c) return 1;
  GridState g;
  HIPCK(c, hipMemcpy(&g, c->gs, sizeof(GridState), hipMemcpyDeviceToHost));
  if (out) {
    out->cg_iterationen = g.cg_iterationen;
    out->loesungen = g.loesungen;
    out->nicht_konvergiert = g.nicht_konvergiert;
    out->kontakte = g.kontakte;
    out->haftend = g.haftend;
    out->freigegeben = g.freigegeben;
    out->rest_max = g.rest_max;
    out->barriere_spins_max = g.spins_max;
    out->naehe_punkt_dreieck = g.naehe_pd;
    out->naehe_kante_kante = g.naehe_kk;
    out->selbst_stoesse = g.selbst_stoesse;
    out->durchdringungen = g.durchdringungen;
    out->durchdringungsgebiete = g.gebiete;
    out->durchstoesse = g.durchstoesse;
    out->cg_rundungsgrenze = g.rundungsgrenze;
    out->cg_neustarts = g.neustarts;
    out->selbst_listen_faktor = c->sk_mal;
    out->koerperpunkte = g.koerperpunkte;
  }
  if (reset) {
    g.cg_iterationen = g.loesungen = g.nicht_konvergiert = g.kontakte = g.haftend = g.freigegeben = 0;
    g.naehe_pd = g.naehe_kk = g.selbst_stoesse = g.durchdringungen = g.gebiete = g.durchstoesse = 0;
    g.rundungsgrenze = 0;
    g.neustarts = 0;
    g.koerperpunkte = 0;
    g.rest_max = 0.0;
    g.spins_max = 0;
    HIPCK(c, hipMemcpy(c->gs, &g, sizeof(GridState), hipMemcpyHostToDevice));
  }
  return 0;
}

static int sk_zaehler_lesen(gc_ctx* c, uint32_t* z) {
  HIPCK(c, hipMemcpyAsync(z, c->sk.zaehler, SK_Z_ANZAHL * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gc_read_naehe(gc_ctx* c, int32_t max, int32_t* art, uint32_t* punkte4, double* gewichte4, double* normale3, double* abstand,
                  double* ruhe) {
  if (!c) return -1;
  if (!c->sk_bereit) return 0;
  uint32_t z[SK_Z_ANZAHL];
  if (sk_zaehler_lesen(c, z)) return -1;
  const uint32_t m = z[SK_Z_NAEHE];
  std::vector<Naehe> l(std::min<uint32_t>(m, (uint32_t)std::max(max, 0)));
  if (!l.empty()) {
    if (hipMemcpy(l.data(), c->sk.naehe, l.size() * sizeof(Naehe), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  }
  for (size_t q = 0; q < l.size(); ++q) {
    if (art) art[q] = (int32_t)l[q].art;
    for (int k = 0; k < 4; ++k) {
      if (punkte4) punkte4[4 * q + k] = l[q].punkte[k];
      if (gewichte4) gewichte4[4 * q + k] = l[q].gewichte[k];
    }
    if (normale3) { normale3[3 * q] = l[q].normale.x; normale3[3 * q + 1] = l[q].normale.y; normale3[3 * q + 2] = l[q].normale.z; }
    if (abstand) abstand[q] = l[q].abstand;
    if (ruhe) ruhe[q] = l[q].ruhe;
  }
  return (int)m;
}

int gc_read_federn(gc_ctx* c, int32_t max, uint32_t* punkte5, double* gewichte5, double* kraft3, double* block9) {
  if (!c) return -1;
  if (!c->sk_bereit) return 0;
  uint32_t z[SK_Z_ANZAHL];
  if (sk_zaehler_lesen(c, z)) return -1;
  const uint32_t m = std::min(z[SK_Z_FEDERN], c->sk.kapazitaet_federn);
  const size_t k = std::min<uint32_t>(m, (uint32_t)std::max(max, 0));
  if (k) {
    if (punkte5 && hipMemcpy(punkte5, c->sk.fe_punkte, 5 * k * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (gewichte5 && hipMemcpy(gewichte5, c->sk.fe_gew, 5 * k * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (kraft3 && hipMemcpy(kraft3, c->sk.fe_kraft, k * sizeof(V3<double>), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (block9 && hipMemcpy(block9, c->sk.fe_block, 9 * k * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  }
  return (int)m;
}

int gc_read_schnitte(gc_ctx* c, int32_t max, uint32_t* kante_dreieck, double* t, double* w3, double* richtung3, double* gewicht,
                     uint32_t* wurzel) {
  if (!c) return -1;
  if (!c->sk_bereit) return 0;
  uint32_t z[SK_Z_ANZAHL];
  if (sk_zaehler_lesen(c, z)) return -1;
  const uint32_t m = z[SK_Z_SCHNITTE];
  const size_t k = std::min<uint32_t>(m, (uint32_t)std::max(max, 0));
  std::vector<Schnitt> l(k);
  std::vector<uint32_t> wz(k);
  if (k) {
    if (hipMemcpy(l.data(), c->sk.schnitte, k * sizeof(Schnitt), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (hipMemcpy(wz.data(), c->sk.schnitt_wurzel, k * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  }
  for (size_t q = 0; q < k; ++q) {
    if (kante_dreieck) { kante_dreieck[2 * q] = l[q].kante; kante_dreieck[2 * q + 1] = l[q].dreieck; }
    if (t) t[q] = l[q].t;
    if (w3) for (int j = 0; j < 3; ++j) w3[3 * q + j] = l[q].w[j];
    if (richtung3) { richtung3[3 * q] = l[q].richtung.x; richtung3[3 * q + 1] = l[q].richtung.y; richtung3[3 * q + 2] = l[q].richtung.z; }
    if (gewicht) gewicht[q] = l[q].gewicht;
    if (wurzel) wurzel[q] = wz[q];
  }
  return (int)m;
}

int gc_read_stoesse(gc_ctx* c, int32_t max, int32_t* art, uint32_t* punkte4, double* lage2, double* zeit, double* seite,
                    double* normale3, int32_t* gilt) {
  if (!c) return -1;
  if (!c->sk_bereit) return 0;
  const size_t slots = (size_t)c->sk.nt + c->sk.ne;
  std::vector<uint32_t> da(slots);
  std::vector<Stoss> l(slots);
  HIPCK(c, hipStreamSynchronize(c->stream));
  if (slots) {
    if (hipMemcpy(da.data(), c->sk.zahl, slots * sizeof(uint32_t), hipMemcpyDeviceToHost) != hipSuccess) return -1;
    if (hipMemcpy(l.data(), c->sk.stoesse, slots * sizeof(Stoss), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  }
  int m = 0;
  for (size_t q = 0; q < slots; ++q) {
    if (!da[q]) continue;
    if (m < max) {
      const Stoss& st = l[q];
      if (art) art[m] = (int32_t)st.art;
      for (int k = 0; k < 4; ++k)
        if (punkte4) punkte4[4 * m + k] = st.punkte[k];
      if (lage2) { lage2[2 * m] = st.lage[0]; lage2[2 * m + 1] = st.lage[1]; }
      if (zeit) zeit[m] = st.zeit;
      if (seite) seite[m] = st.seite;
      if (normale3) { normale3[3 * m] = st.normale.x; normale3[3 * m + 1] = st.normale.y; normale3[3 * m + 2] = st.normale.z; }
      if (gilt) gilt[m] = (st.gilt & 1u) ? 1 : 0;
    }
    ++m;
  }
  return m;
}

int gc_get_v(gc_ctx* c, double* v) {
  if (!c || !c->topo || !v) return 1;
  V3<double>* dv = c->opt.praezision == GC_FP32 ? c->sf.v : c->sd.v;
  HIPCK(c, hipMemcpyAsync(v, dv, c->n * sizeof(V3<double>), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gc_zeiten(gc_ctx* c, double* ms_schritt, double* ms_pcg) {
  if (!c) return 1;
  if (ms_schritt) *ms_schritt = c->ms_schritt;
  if (ms_pcg) *ms_pcg = c->ms_pcg;
  return 0;
}

int gc_add_sdf_von(gc_ctx* c, const gc_ctx* quelle, int32_t k, const double versatz[3], const double v[3]) {
  if (!c) return -1;
  if (!quelle || k < 0 || k >= (int32_t)quelle->koerper.size() || quelle->koerper[(size_t)k].k.art != GC_SDF) {
    c->err = "gc_add_sdf_von: no such grid body in the source context";
    return -1;
  }
  if (c->koerper.size() >= GC_MAX_KOERPER) {
    c->err = "too many bodies";
    return -1;
  }
  Koerperhost kh;
  kh.k = quelle->koerper[(size_t)k].k;  // same device grid; not owned here
  for (int q = 0; q < 3; ++q) {
    kh.k.punkt[q] = versatz[q];
    kh.k.geschwindigkeit[q] = v[q];
  }
  c->koerper.push_back(kh);
  alle_param(c);
  return (int)c->koerper.size() - 1;
}

gc_batch* gc_batch_create(gc_ctx* const* ctx, int32_t n, int32_t threads, int32_t bloecke, char* err, int32_t err_len) {
  auto melde = [&](const std::string& m) {
    if (err && err_len > 0) std::snprintf(err, (size_t)err_len, "%s", m.c_str());
  };
  if (!ctx || n <= 0) {
    melde("gc_batch_create: no systems");
    return nullptr;
  }
  for (int32_t k = 0; k < n; ++k) {
    if (!ctx[k] || !ctx[k]->topo || !ctx[k]->hat_param) {
      melde("gc_batch_create: every system needs topology and parameters");
      return nullptr;
    }
    if (ctx[k]->opt.praezision != ctx[0]->opt.praezision || ctx[k]->opt.geraet != ctx[0]->opt.geraet) {
      melde("gc_batch_create: systems differ in precision or device");
      return nullptr;
    }
    for (int32_t q = 0; q < k; ++q)
      if (ctx[q] == ctx[k]) {
        melde("gc_batch_create: a system is listed twice");
        return nullptr;
      }
  }
  if (threads <= 0) threads = 1024;
  if (threads > 1024 || threads < 64 || (threads & (threads - 1))) {
    melde("threads must be a power of two in [64, 1024]");
    return nullptr;
  }
  gc_batch* b = new gc_batch();
  b->ctx.assign(ctx, ctx + n);
  b->praezision = ctx[0]->opt.praezision;
  b->cus = ctx[0]->cus;
  b->gewicht.resize((size_t)n);
  for (int32_t k = 0; k < n; ++k) b->gewicht[(size_t)k] = (double)std::max(1, ctx[k]->n_kacheln);
  for (int32_t k = 0; k < n; ++k)
    if (hipStreamSynchronize(ctx[k]->stream) != hipSuccess) {
      melde("gc_batch_create: a system's stream failed");
      delete b;
      return nullptr;
    }
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    melde("hipStreamCreate failed");
    delete b;
    return nullptr;
  }
  const int n_zweige = std::min<int>(n, 8);
  for (int q = 0; q < n_zweige; ++q) {
    hipStream_t z = nullptr;
    if (hipStreamCreateWithFlags(&z, hipStreamNonBlocking) != hipSuccess) {
      melde("hipStreamCreate failed");
      gc_batch_destroy(b);
      return nullptr;
    }
    b->zweige.push_back(z);
  }
  for (int q = 0; q < 3 * (n_zweige + 1); ++q) {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      melde("hipEventCreate failed");
      gc_batch_destroy(b);
      return nullptr;
    }
    b->gabel.push_back(e);
  }
  int rc = b->praezision == GC_FP32 ? batch_form<float>(b, threads, bloecke) : batch_form<double>(b, threads, bloecke);
  if (rc) {
    melde(b->err);
    gc_batch_destroy(b);
    return nullptr;
  }
  return b;
}

void gc_batch_destroy(gc_batch* b) {
  if (!b) return;
  for (hipStream_t z : b->zweige) (void)hipStreamSynchronize(z);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  if (b->graph) (void)hipGraphExecDestroy(b->graph);
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->gabel) (void)hipEventDestroy(e);
  if (b->tab) (void)hipFree(b->tab);
  for (hipStream_t z : b->zweige) (void)hipStreamDestroy(z);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

const char* gc_batch_last_error(const gc_batch* b) { return b ? b->err.c_str() : "no batch"; }

int gc_batch_step(gc_batch* b, double h, int32_t n) {
  if (!b) return 1;
  return b->praezision == GC_FP32 ? batch_schritte<float>(b, h, n) : batch_schritte<double>(b, h, n);
}

int gc_batch_step_graph(gc_batch* b, double h, int32_t n) {
  if (!b) return 1;
  return b->praezision == GC_FP32 ? batch_schritte_graph<float>(b, h, n) : batch_schritte_graph<double>(b, h, n);
}

int gc_batch_autotune(gc_batch* b, double h, int32_t* threads, int32_t* bloecke, double* ms) {
  if (!b) return 1;
  return b->praezision == GC_FP32 ? batch_autotune<float>(b, h, threads, bloecke, ms)
                                  : batch_autotune<double>(b, h, threads, bloecke, ms);
}

int gc_batch_bloecke(gc_batch* b, int32_t* bloecke, int32_t* threads) {
  if (!b) return 1;
  if (bloecke)
    for (size_t k = 0; k < b->bloecke.size(); ++k) bloecke[k] = b->bloecke[k];
  if (threads) *threads = b->threads;
  return 0;
}

int gc_batch_zeiten(gc_batch* b, double* ms_schritt, double* ms_pcg) {
  if (!b) return 1;
  if (ms_schritt) *ms_schritt = b->ms_schritt;
  if (ms_pcg) *ms_pcg = b->ms_pcg;
  return 0;
}

// --- adjoint and saved states ----------------------------------------------

int gc_adjoint_step(gc_ctx* c, double h) {
  if (!c) return 1;
  return c->opt.praezision == GC_FP32 ? adj_schritt<float>(c, h) : adj_schritt<double>(c, h);
}

int gc_step_aufnehmen(gc_ctx* c, double h, int32_t slot) {
  if (!c || !c->topo) return 1;
  const bool neu = c->zustaende.find(slot) == c->zustaende.end();
  gc_ctx::Zustand& z = c->zustaende[slot];
  z.hat_loesung = false;
  if (zustand_sichern_in(c, z)) {
    if (neu && !z.x) c->zustaende.erase(slot);
    return 1;
  }
  const size_t n = (size_t)c->n;
  const bool fp32 = c->opt.praezision == GC_FP32;
  for (int k = 0; k < 2; ++k) {
    if (!z.dv_loes[k]) z.dv_loes[k] = fp32 ? (void*)dalloc<V3<float>>(c, n) : (void*)dalloc<V3<double>>(c, n);
    if (fp32 && c->sf.dv64 && !z.dv64_loes[k]) z.dv64_loes[k] = dalloc<V3<double>>(c, n);
    if (!z.dv_loes[k] || (fp32 && c->sf.dv64 && !z.dv64_loes[k])) return fehler(c, "hipMalloc failed (recorded solutions)");
  }
  c->aufnahme = &z;
  const int rc = gc_step(c, h, 1);
  c->aufnahme = nullptr;
  z.hat_loesung = rc == 0;
  return rc;
}

int gc_adjoint_step_aus(gc_ctx* c, double h, int32_t slot) {
  if (!c || !c->topo) return 1;
  auto it = c->zustaende.find(slot);
  if (it == c->zustaende.end() || !it->second.hat_loesung)
    return fehler(c, "gc_adjoint_step_aus: no step recorded in this slot (gc_step_aufnehmen)");
  if (zustand_laden_aus(c, it->second)) return 1;
  alle_param(c);
  c->wiedergabe = &it->second;
  const int rc = c->opt.praezision == GC_FP32 ? adj_schritt<float>(c, h) : adj_schritt<double>(c, h);
  c->wiedergabe = nullptr;
  return rc;
}

int gc_adjoint_set(gc_ctx* c, const double* gx, const double* gv) {
  if (!c || !c->topo) return 1;
  int rc = c->opt.praezision == GC_FP32 ? adj_vorbereiten<float>(c) : adj_vorbereiten<double>(c);
  if (rc) return rc;
  const size_t bytes = (size_t)c->n * sizeof(V3<double>);
  if (gx) HIPCK(c, hipMemcpyAsync(c->adj.gx, gx, bytes, hipMemcpyHostToDevice, c->stream));
  else HIPCK(c, hipMemsetAsync(c->adj.gx, 0, bytes, c->stream));
  if (gv) HIPCK(c, hipMemcpyAsync(c->adj.gv, gv, bytes, hipMemcpyHostToDevice, c->stream));
  else HIPCK(c, hipMemsetAsync(c->adj.gv, 0, bytes, c->stream));
  // a new seed: the next adjoint solve starts from mu = 0
  const size_t mu_bytes = (size_t)c->n * (c->opt.praezision == GC_FP32 ? sizeof(V3<float>) : sizeof(V3<double>));
  HIPCK(c, hipMemsetAsync(c->adj_mu, 0, mu_bytes, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gc_adjoint_get(gc_ctx* c, double* gx, double* gv) {
  if (!c || !c->topo) return 1;
  if (!c->adj_bereit) return fehler(c, "no adjoint state yet (gc_adjoint_set)");
  const size_t bytes = (size_t)c->n * sizeof(V3<double>);
  if (gx) HIPCK(c, hipMemcpyAsync(gx, c->adj.gx, bytes, hipMemcpyDeviceToHost, c->stream));
  if (gv) HIPCK(c, hipMemcpyAsync(gv, c->adj.gv, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  return 0;
}

int gc_adjoint_gradient(gc_ctx* c, double* d_kanten, double* d_naehte, double* d_biegung, double* d_membran,
                        double* d_reibung, int32_t reset) {
  if (!c || !c->topo) return 1;
  return c->opt.praezision == GC_FP32 ? adj_gradient_lesen<float>(c, d_kanten, d_naehte, d_biegung, d_membran, d_reibung, reset)
                                      : adj_gradient_lesen<double>(c, d_kanten, d_naehte, d_biegung, d_membran, d_reibung, reset);
}

int gc_adjoint_statistik(gc_ctx* c, gc_adjoint_stat* out, int32_t reset) {
  if (!c || !c->topo) return 1;
  if (!c->adj_bereit) {
    if (out) std::memset(out, 0, sizeof(*out));
    return 0;
  }
  GridState g;
  unsigned long long z[2];
  HIPCK(c, hipMemcpyAsync(&g, c->gs_adj, sizeof(GridState), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipMemcpyAsync(z, c->adj.zaehler, sizeof(z), hipMemcpyDeviceToHost, c->stream));
  HIPCK(c, hipStreamSynchronize(c->stream));
  if (out) {
    out->schritte = c->adj_schritte;
    out->cg_iterationen = g.cg_iterationen;
    out->nicht_konvergiert = g.nicht_konvergiert;
    out->abweichungen = (int64_t)z[0];
    out->eintritte = (int64_t)z[1];
    out->rest_max = g.rest_max;
  }
  if (reset) {
    c->adj_schritte = 0;
    HIPCK(c, hipMemsetAsync(c->gs_adj, 0, sizeof(GridState), c->stream));
    HIPCK(c, hipMemsetAsync(c->adj.zaehler, 0, sizeof(z), c->stream));
    HIPCK(c, hipStreamSynchronize(c->stream));
  }
  return 0;
}

int gc_set_steifigkeit(gc_ctx* c, const double* kanten_k, const double* biegung_k, const double* membran_ek_es_g,
                       const double* reibung) {
  if (!c || !c->topo) return 1;
  return c->opt.praezision == GC_FP32 ? steifigkeit_setzen<float>(c, kanten_k, biegung_k, membran_ek_es_g, reibung)
                                      : steifigkeit_setzen<double>(c, kanten_k, biegung_k, membran_ek_es_g, reibung);
}

int gc_zustand_sichern(gc_ctx* c, int32_t slot) {
  if (!c || !c->topo) return 1;
  const bool neu = c->zustaende.find(slot) == c->zustaende.end();
  gc_ctx::Zustand& z = c->zustaende[slot];
  if (zustand_sichern_in(c, z)) {
    if (neu && !z.x) c->zustaende.erase(slot);
    return 1;
  }
  return 0;
}

int gc_entscheidungen(gc_ctx* c, int32_t modus, int32_t slot) {
  if (!c || !c->topo) return 1;
  if (modus < 0 || modus > 2) return fehler(c, "gc_entscheidungen: modus must be 0 (off), 1 (record) or 2 (replay)");
  uint16_t* puffer = nullptr;
  if (modus != 0) {
    if (c->koerper.size() > 1) return fehler(c, "gc_entscheidungen: decisions are held for at most one body");
    if (selbst_an(c)) return fehler(c, "gc_entscheidungen: self-collision decisions are not held (switch it off)");
    auto it = c->entscheidungen.find(slot);
    if (it == c->entscheidungen.end()) {
      if (modus == 2) return fehler(c, "gc_entscheidungen: nothing recorded in this slot");
      puffer = dalloc<uint16_t>(c, (size_t)c->n);
      if (!puffer) return fehler(c, "hipMalloc failed (decisions)");
      c->entscheidungen[slot] = puffer;
    } else {
      puffer = it->second;
    }
  }
  if (c->opt.praezision == GC_FP32) {
    c->sf.halten = puffer;
    c->sf.halten_modus = modus;
  } else {
    c->sd.halten = puffer;
    c->sd.halten_modus = modus;
  }
  return 0;
}

#ifdef GC_PHASEN
// Build flag GC_PHASEN only (a diagnostic, not part of the C ABI's header):
// block 0's mean microseconds per PCG iteration in stage A, stage B, stage C,
// the p.Ap all-reduce, the update and the r.z all-reduce, over the
// iterations since the last reset; returns the iteration count.
// us7: the six phases of a PCG iteration seen by block 0 (stage A, stage B,
// stage C, all-reduce p.Ap, update, all-reduce r.z), then block 0's time
// from its partial to the summed result per all-reduce, in microseconds;
// returns the iterations they cover
extern "C" int gc_phasen_lesen(gc_ctx* c, double* us7, int32_t reset) {
  if (!c || !us7) return -1;
  double* us6 = us7;
  GridState g;
  if (hipMemcpy(&g, c->gs, sizeof(GridState), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->opt.geraet) != hipSuccess || khz <= 0) return -1;
  const double it = (double)std::max<unsigned long long>(g.phasen_it, 1ull);
  for (int k = 0; k < 6; ++k) us6[k] = (double)g.phasen[k] / ((double)khz * 1e-3) / it;
  us7[6] = (double)g.warte0[0] / ((double)khz * 1e-3) / (double)std::max<unsigned long long>(g.warte0[1], 1ull);
  if (reset) {
    for (int k = 0; k < 6; ++k) g.phasen[k] = 0;
    for (int k = 0; k < 2; ++k) g.warte0[k] = 0;
    const unsigned long long n = g.phasen_it;
    g.phasen_it = 0;
    if (hipMemcpy(c->gs, &g, sizeof(GridState), hipMemcpyHostToDevice) != hipSuccess) return -1;
    return (int)n;
  }
  return (int)g.phasen_it;
}
#endif

#ifdef GC_PHASEN
// out (6 x n, microseconds summed over the recorded iterations): per block,
// stages A, B and C, update, wait in the p.Ap all-reduce, wait in the r.z
// one. The first call (out null) turns the record on.
extern "C" int gc_phasen_bloecke(gc_ctx* c, double* out, int32_t n, int32_t reset) {
  if (!c) return -1;
  GridState g;
  if (hipMemcpy(&g, c->gs, sizeof(GridState), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  if (!g.pro_block) {
    if (hipMalloc((void**)&g.pro_block, 6 * GC_MAX_BLOECKE * sizeof(unsigned long long)) != hipSuccess) return -1;
    if (hipMemset(g.pro_block, 0, 6 * GC_MAX_BLOECKE * sizeof(unsigned long long)) != hipSuccess) return -1;
    if (hipMemcpy(c->gs, &g, sizeof(GridState), hipMemcpyHostToDevice) != hipSuccess) return -1;
  }
  if (!out) return 0;
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->opt.geraet) != hipSuccess || khz <= 0) return -1;
  std::vector<unsigned long long> h(6 * GC_MAX_BLOECKE);
  if (hipMemcpy(h.data(), g.pro_block, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost) != hipSuccess) return -1;
  for (int k = 0; k < 6 * std::min<int>(n, GC_MAX_BLOECKE); ++k) out[k] = (double)h[k] / ((double)khz * 1e-3);
  if (reset && hipMemset(g.pro_block, 0, 6 * GC_MAX_BLOECKE * sizeof(unsigned long long)) != hipSuccess) return -1;
  return 0;
}
#endif

#ifdef GC_PHASEN
}  // extern "C"
// out (4 x n): per tile its first row, its end row, the spring list entries
// of its rows after the last rebuild and its listed elements, so that the
// per-block times of one-tile blocks can be set against the tile's work.
template <typename T>
static int phasen_kacheln(gc_ctx* c, int32_t* out, int32_t n) {
  const SysDev<T>& s = sys<T>(c);
  const int nt = std::min<int>(n, s.n_kacheln);
  std::vector<int32_t> kz(2 * (size_t)s.n_kacheln);
  std::vector<uint32_t> el((size_t)s.n_kacheln + 1), fe((size_t)s.n + 1, 0u);
  if (hipMemcpy(kz.data(), s.kachel_zeilen, kz.size() * 4, hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(el.data(), s.kachel_el_off, el.size() * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return -1;
  if (s.selbst && s.fe_off && hipMemcpy(fe.data(), s.fe_off, fe.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) return -1;
  for (int t = 0; t < nt; ++t) {
    out[4 * t] = kz[2 * t];
    out[4 * t + 1] = kz[2 * t + 1];
    out[4 * t + 2] = (int32_t)(fe[kz[2 * t + 1]] - fe[kz[2 * t]]);
    out[4 * t + 3] = (int32_t)(el[t + 1] - el[t]);
  }
  return nt;
}
extern "C" {
int gc_phasen_kacheln(gc_ctx* c, int32_t* out, int32_t n) {
  if (!c || !out) return -1;
  return c->opt.praezision == GC_FP32 ? phasen_kacheln<float>(c, out, n) : phasen_kacheln<double>(c, out, n);
}
#endif

int gc_entscheidungen_abweichungen(gc_ctx* c, int64_t* anzahl, int32_t reset) {
  if (!c || !anzahl) return 1;
  GridState g;
  HIPCK(c, hipMemcpy(&g, c->gs, sizeof(GridState), hipMemcpyDeviceToHost));
  *anzahl = g.halten_abweichungen;
  if (reset) {
    long long null = 0;
    HIPCK(c, hipMemcpy(&c->gs->halten_abweichungen, &null, sizeof(null), hipMemcpyHostToDevice));
  }
  return 0;
}

int gc_zustand_laden(gc_ctx* c, int32_t slot) {
  if (!c || !c->topo) return 1;
  auto it = c->zustaende.find(slot);
  if (it == c->zustaende.end()) return fehler(c, "gc_zustand_laden: no state saved in this slot");
  if (zustand_laden_aus(c, it->second)) return 1;
  alle_param(c);
  return 0;
}

}  // extern "C"
