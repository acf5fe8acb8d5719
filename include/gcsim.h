/*
 * gcsim — MI355X-native implicit backward-Euler cloth step (C ABI).
 *
 * Drop-in for klon3d-sim's implicit path. The reference plugs solvers into
 * `ImplizitCg` through the trait `Loeser` (crates/klon3d-sim/src/implizit.rs:304-351)
 * and steps frames through `Rechenweg::bild` (crates/klon3d-sim/src/cpu.rs:58-60,
 * implemented at implizit.rs:1042-1057). Each entry point below names the
 * reference item it replaces. All paths are relative to the reference checkout.
 *
 * Conventions
 *   - Every function returns int status: 0 = ok, nonzero = error; the message is
 *     available from gc_last_error(ctx). Non-convergence is not an error: it is
 *     reported in gc_cg_lauf, as the reference reports it in `CgLauf`
 *     (implizit.rs:293-299).
 *   - Point arrays are caller-owned, interleaved double[3*n] like Vec<[f64;3]>.
 *   - One context is not thread-safe (the reference takes &mut self); several
 *     contexts may live side by side. A context owns its device buffers and
 *     keeps the topology resident between calls (the reference GPU backend does
 *     the same, gpu.rs:1356-1456).
 *   - No torch or HIP types cross this boundary.
 */
#ifndef GCSIM_H
#define GCSIM_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct gc_ctx gc_ctx;

/* Arithmetic of the bulk kernels. Reductions, residual norms and the
 * convergence test are FP64 in both modes. */
enum { GC_FP64 = 0, GC_FP32 = 1 };

typedef struct gc_optionen {
  int32_t praezision;     /* GC_FP64 or GC_FP32 */
  int32_t geraet;         /* HIP device ordinal */
  int32_t bloecke;        /* persistent PCG blocks; 0 = one per tile, up to the
                           * co-resident maximum (one 1024-thread block per CU) */
  int32_t threads;        /* threads per block; 0 = choose */
  int32_t lds_kb;         /* LDS a PCG tile may fill (KiB); 0 = 80 */
} gc_optionen;

/* Topology of the system, as `System` carries it when the backend builds the
 * geometry itself (`System::ohne_geometrie`, implizit.rs:123-172, filled at
 * implizit.rs:718-797): per distance constraint the two points, the
 * stiffness (0 = carried by a membrane) and the rest length; per seam spring
 * the two points and the rest length (stiffness = gc_param.naht_k, rest +inf =
 * open, implizit.rs:757-765); per bending constraint four points, stiffness and
 * rest angle; per membrane triangle three points and the 12-value element
 * {A0, ruhe_inv[4], kette[2], schuss[2], E_k, E_s, G} (kontinuum.rs:69-99). */
typedef struct gc_topologie {
  int32_t n;
  const double* masse;               /* n */
  const double* w;                   /* n, inverse mass; <= 0 means held */
  const double* kollisionsabstand;   /* n */
  const double* reibung;             /* n */
  int32_t n_kanten;
  const uint32_t* kanten_punkte;     /* 2 per edge */
  const double* kanten_k;
  const double* kanten_ruhe;
  int32_t n_naehte;
  const uint32_t* naht_punkte;       /* 2 per seam spring */
  const double* naht_ruhe;
  int32_t n_biegungen;
  const uint32_t* biegung_punkte;    /* 4 per constraint */
  const double* biegung_k;
  const double* biegung_ruhe;
  int32_t n_membran;
  const uint32_t* membran_punkte;    /* 3 per element */
  const double* membran_daten;       /* 12 per element */
  /* Cloth self-collision (selbstkollision.rs:623-693): the cloth triangles
   * (Szene::dreiecke), the stretch stiffness (N/m) of each point's and each
   * triangle's piece, and each triangle's rest area. May be 0/NULL when
   * self-collision stays off. */
  int32_t n_dreiecke;
  const uint32_t* dreiecke;          /* 3 per triangle */
  const double* selbst_e_punkt;      /* n */
  const double* selbst_e_dreieck;    /* n_dreiecke */
  const double* selbst_a_dreieck;    /* n_dreiecke */
} gc_topologie;

/* ImplizitParameter (implizit.rs:456-531) plus the scene's Simparameter
 * (szene.rs:95-126) that the step reads. */
typedef struct gc_param {
  double schwerkraft[3];
  double daempfung;          /* air damping c (1/s) */
  int32_t cg_max;
  double cg_toleranz;
  double naht_k;
  double kontakt_rand;
  int32_t normalkraft_aus_last;
  /* ImplizitParameter's self-collision fields (implizit.rs:482-494, defaults
   * 522-526): on/off, extra collision thickness (m), spring stiffness factor,
   * impact response iterations, search cuts across seams too. */
  int32_t selbstkollision;
  double selbst_zusatzdicke;
  double selbst_steifigkeit;
  int32_t selbst_stoss_iterationen;
  int32_t selbst_schnitte_an_naehten;
  /* implizit.rs:476 (default on): a path through a body's skin is set onto
   * its entry (Koerpernetz::durchstoss, netzkoerper.rs:541-596; for the grid
   * SDF, the first outside -> inside crossing of the sampled distance). */
  int32_t durchstoss_pruefen;
  /* Passes of body points against cloth triangles per step after the update
   * (implizit.rs:477-481, 1017-1026; Koerpernetz::punkte_gegen_dreiecke,
   * netzkoerper.rs:625-665): every cloth triangle under which a point of a
   * mesh body lies at most koerperpunkt_tiefe (m) deep is lifted over the
   * deepest such point, split over its corners by inverse mass; the passes
   * end early after one that lifts nothing. 0 switches it off (default). */
  int32_t koerperpunkt_durchgaenge;
  double koerperpunkt_tiefe;  /* default 0.01 */
} gc_param;

/* CgLauf (implizit.rs:293-299). */
typedef struct gc_cg_lauf {
  int32_t iterationen;
  int32_t nicht_konvergiert;
  double rest;
} gc_cg_lauf;

/* ImplizitStatistik (implizit.rs:534-561), the counters this path keeps. */
typedef struct gc_statistik {
  int64_t cg_iterationen;
  int64_t loesungen;
  int64_t nicht_konvergiert;
  int64_t kontakte;
  int64_t haftend;
  int64_t freigegeben;
  double rest_max;
  int64_t barriere_spins_max;  /* longest grid-barrier wait in polls (diagnostic) */
  int64_t naehe_punkt_dreieck;
  int64_t naehe_kante_kante;
  int64_t selbst_stoesse;
  int64_t durchdringungen;
  int64_t durchdringungsgebiete;
  int64_t durchstoesse;
  int64_t cg_rundungsgrenze;   /* FP32 solves whose iterations stopped lowering the FP64 residual (the
                                * system too ill-conditioned for FP32) and that FP64 iterations finished
                                * (no FP64 counterpart) */
  int64_t cg_neustarts;        /* FP32 restarts of the PCG from the FP64 residual (no FP64 counterpart) */
  int64_t selbst_listen_faktor; /* capacity of the self-collision lists, in multiples of the default (a call
                                 * whose lists overflow is taken again with them doubled, up to 16; the
                                 * reference keeps them in growable vectors). Not reset. */
  int64_t koerperpunkte;       /* cloth triangles lifted over body points (implizit.rs:550, 1022) */
} gc_statistik;

/* Construction. Replaces `GpuLoeser::neu() -> Result<Self, String>` (gpu.rs:1136):
 * NULL on failure with the reason in err (if given). */
gc_ctx* gc_create(const gc_optionen* opt, char* err, int32_t err_len);
void gc_destroy(gc_ctx* ctx);
const char* gc_last_error(const gc_ctx* ctx);
/* `Loeser::name` (implizit.rs:306). */
const char* gc_name(const gc_ctx* ctx);

/* Upload topology, stiffnesses and rest lengths; kept on the device until the
 * next call (gpu.rs:1362-1456 re-uploads only on change). */
int gc_set_topology(gc_ctx* ctx, const gc_topologie* topo);
int gc_set_param(gc_ctx* ctx, const gc_param* p);
/* Seam rest lengths change while a garment is sewn (`Szene::naht_kuerzen`,
 * szene.rs:768-774). */
int gc_set_naht_ruhe(gc_ctx* ctx, const double* ruhe);
/* Held points (w = 0) are set by callers that clamp (tests/kontinuum.rs:382-386). */
int gc_set_w(gc_ctx* ctx, const double* w);

/* Bodies (koerper.rs:16-26): the cylinder, the triangle-mesh body
 * (gc_add_netz below) and a sampled grid SDF, which answers
 * `Koerper::kontakt` (koerper.rs:69-91) like the mesh does.
 * werte: float[nx*ny*nz], x fastest. Returns the body index or -1. */
int gc_add_zylinder(gc_ctx* ctx, const double punkt[3], const double achse[3], double radius, const double geschwindigkeit[3]);
int gc_add_sdf(gc_ctx* ctx, const float* werte, const int32_t dims[3], const double origin[3], double spacing,
               const double versatz[3], const double geschwindigkeit[3]);
/* Move a body to an absolute offset (cylinder: axis point; SDF: versatz), as
 * `Koerper::verschoben` does per substep (koerper.rs:122-133). */
int gc_set_koerper_lage(gc_ctx* ctx, int32_t koerper, const double lage[3]);
/* A body that uses the device grid of body k of another context (a batch of
 * garments on one body keeps one copy of its SDF). The source context must
 * outlive this one. No counterpart in the reference, which has one scene per
 * solver. Returns the body index or -1. */
int gc_add_sdf_von(gc_ctx* ctx, const gc_ctx* quelle, int32_t k, const double versatz[3], const double geschwindigkeit[3]);

/* A closed, consistently oriented triangle mesh as body: `Koerper::Netz`
 * (koerper.rs:25-26) over `Koerpernetz::neu` (netzkoerper.rs:297-388).
 * punkte double[3n] (m), dreiecke uint32[3m] with outward normals, suchtiefe
 * (netzkoerper.rs:46, 10 mm in the reference) how deep inside a point is still
 * found. Rejected like `neu` rejects: no triangles, suchtiefe <= 0, an index
 * out of range, an edge of three or more triangles, an inconsistently oriented
 * edge, volume <= 0. Contacts are `Koerper::kontakt` (koerper.rs:82-90), path
 * entries `Koerpernetz::durchstoss` (netzkoerper.rs:541-596). The body point
 * pass (`punkte_gegen_dreiecke`, netzkoerper.rs:625-665; off by default,
 * implizit.rs:523) is not built. Returns the body index or -1. */
int gc_add_netz(gc_ctx* ctx, const double* punkte, int32_t n_punkte, const uint32_t* dreiecke, int32_t n_dreiecke,
                double suchtiefe, const double versatz[3], const double geschwindigkeit[3]);
/* `Koerpernetz::nachfuehren` (netzkoerper.rs:390-418): a new pose of the same
 * points; normals, spheres and grid are rebuilt on the device, the grid never
 * shrinks, a velocity field is cleared. */
int gc_netz_nachfuehren(gc_ctx* ctx, int32_t koerper, const double* punkte);
/* `Koerpernetz::setze_punktgeschwindigkeit` (netzkoerper.rs:454-466): one
 * velocity per point (double[3n]); n = 0 clears it (one velocity for the body). */
int gc_netz_punktgeschwindigkeit(gc_ctx* ctx, int32_t koerper, const double* v, int32_t n);
typedef struct gc_netz_angaben {
  int32_t n_punkte, n_dreiecke;
  int64_t randkanten;   /* edges with one triangle (open borders) */
  double volumen;       /* enclosed volume, m^3 */
  double ursprung[3];   /* grid origin */
  double zelle;         /* cell edge, 2 x suchtiefe */
  int64_t zahl[3];      /* cells per axis */
  int64_t n_eintraege;  /* triangle entries over all cells */
} gc_netz_angaben;
int gc_netz_info(gc_ctx* ctx, int32_t koerper, gc_netz_angaben* out);
/* The device arrays (each may be NULL): face normals double[3m], point normals
 * double[3n], neighbour per side uint32[3m], spheres double[4m] (centroid,
 * radius), grid starts uint32[cells + 1], grid entries uint32[n_eintraege]. */
int gc_netz_lesen(gc_ctx* ctx, int32_t koerper, double* flaechennormale, double* punktnormale, uint32_t* nachbar,
                  double* kugel, uint32_t* anfang, uint32_t* eintraege);
/* `Koerpernetz::naechster` (netzkoerper.rs:503-531) for n points in the
 * mesh's frame: da[i] = 1 if a surface point lies closer than radius; then the
 * signed distance, nearest point, outward normal, triangle and region
 * (0..2 corner k, 3..5 edge k, 6 face). Output arrays other than da may be NULL. */
int gc_netz_naechster(gc_ctx* ctx, int32_t koerper, int32_t n, const double* x, double radius, int32_t* da, double* abstand,
                      double* punkt, double* normale, int32_t* dreieck, int32_t* region);
/* `Koerpernetz::durchstoss` (netzkoerper.rs:541-596) for n segments a -> b in
 * the mesh's frame: entry point, face normal, surface velocity there. */
int gc_netz_durchstoss(gc_ctx* ctx, int32_t koerper, int32_t n, const double* a, const double* b, const double starr[3],
                       int32_t* da, double* p, double* normale, double* vk);
/* The mesh's signed distance on a node grid (werte float[nx*ny*nz], x fastest,
 * node (i,j,k) at origin + spacing*(i,j,k), in the mesh's frame), for a
 * gc_add_sdf body. Exact (`naechster`, sign by the pseudonormal) at nodes closer
 * than `band` to the surface, +-band beyond it with the node's side; band must
 * be at least one spacing. No counterpart in the reference, whose mesh body is
 * queried directly; it feeds the grid-SDF body from a mesh. */
int gc_netz_sdf_backen(gc_ctx* ctx, int32_t koerper, const int32_t dims[3], const double origin[3], double spacing, double band,
                       float* werte);

/* The XPBD path, the reference's other `Rechenweg` (`CpuReferenz`,
 * cpu.rs:76-246), on the device. Its constraint lists as `Szene` holds them
 * (szene.rs:127-160, colour-ordered): distance constraints with rest length,
 * compliance (m/N) and the flag of those a membrane carries (skipped);
 * seams with their rest and one compliance (Simparameter::naht_nachgiebigkeit,
 * szene.rs:545); bending with rest angle and compliance. Masses, inverse
 * masses, collision distances, friction, membranes and bodies are the
 * context's. Call once, after gc_set_topology. */
typedef struct gc_xpbd_topologie {
  int32_t n_abstand;
  const uint32_t* abstand_punkte;        /* 2 per constraint */
  const double* abstand_ruhe;
  const double* abstand_nachgiebigkeit;
  const uint8_t* membran_kante;          /* may be NULL */
  int32_t n_naehte;
  const uint32_t* naht_punkte;           /* 2 per seam */
  const double* naht_ruhe;
  double naht_nachgiebigkeit;
  int32_t n_biegungen;
  const uint32_t* biegung_punkte;        /* 4 per constraint */
  const double* biegung_ruhe;
  const double* biegung_nachgiebigkeit;
} gc_xpbd_topologie;
int gc_xpbd_topology(gc_ctx* ctx, const gc_xpbd_topologie* t);
/* Contacts of the last XPBD frame (`Kontaktstatistik`, cpu.rs:63-73). */
typedef struct gc_xpbd_statistik {
  int64_t kontakte, haftend;
  double tiefe, tangential;
} gc_xpbd_statistik;
/* One frame of `CpuReferenz::bild` (cpu.rs:226-245): `unterschritte`
 * substeps of bild_dt / unterschritte, gravity and damping from gc_param,
 * bodies moved from their frame-start offsets per substep. */
int gc_xpbd_bild(gc_ctx* ctx, double bild_dt, int32_t unterschritte, gc_xpbd_statistik* out);
/* Runs (distance, seams, bending) and membrane dependency levels the frame
 * launches one after another (a diagnostic). No counterpart in the reference. */
int32_t gc_xpbd_laeufe(gc_ctx* ctx, int32_t* abstand, int32_t* naehte, int32_t* biegungen, int32_t* membran_stufen);

/* State: positions and velocities (Szene::x, Szene::v). */
int gc_set_state(gc_ctx* ctx, const double* x, const double* v);
int gc_get_state(gc_ctx* ctx, double* x, double* v);
/* Warm start of the next step (`ImplizitCg::dv_vorher`, implizit.rs:578, 906, 945).
 * n = 0 clears it (the next step then starts from h*g). */
int gc_set_dv_vorher(gc_ctx* ctx, const double* dv, int32_t n);

/* --- the Loeser plug-in (implizit.rs:304-351) ------------------------------ */

/* Build the system of a substep of length h at the current state: geometry,
 * forces, load, right-hand side, preconditioner (`baut_selbst` = true:
 * the device does what ImplizitCg::system_bauen does at implizit.rs:725-894),
 * and the contacts against the bodies (implizit.rs:872-883). */
int gc_build(gc_ctx* ctx, double h);
/* Replace the contact set of the built system (`fest`, implizit.rs:593-599):
 * flag per point (0/1), normal, target normal velocity, body velocity. */
int gc_set_fest(gc_ctx* ctx, const int32_t* flag, const double* normale, const double* vn_soll, const double* v_koerper);
int gc_get_fest(gc_ctx* ctx, int32_t* flag, double* normale, double* vn_soll, double* v_koerper);
/* `Loeser::loese` (implizit.rs:308, CpuLoeser at 370-426): modified PCG;
 * dv in = warm start, out = solution. */
int gc_solve(gc_ctx* ctx, double* dv, gc_cg_lauf* lauf);
/* No counterpart in the reference. For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid it): gc_solve without any allocation. The grid state and the
 * reduction and barrier slots of its PCG launch (up to 1024 blocks, about
 * 128.3 KiB) live in a static device array, one per device, instead of the
 * context's buffers; everything else is the context's, set up before. The
 * warm start and the solution are copied straight between dv and the
 * device's solution vector. Calls that overlap on one device race on that
 * array; the result equals gc_solve's to the bit. */
int gc_solve_ohne_speicher(gc_ctx* ctx, double* dv, gc_cg_lauf* lauf);
/* `Loeser::mal_a` (implizit.rs:310): unfiltered A*p. */
int gc_apply_A(gc_ctx* ctx, const double* p, double* out);
/* `System::mal_h` (implizit.rs:185-221): H*p without mass and h^2. */
int gc_apply_H(gc_ctx* ctx, const double* p, double* out);
/* `Loeser::rest_geloest` (implizit.rs:337-342): A*dv - b. */
int gc_residual(gc_ctx* ctx, const double* dv, double* out);
/* `Loeser::last_lesen` (implizit.rs:348-350): load per point (tension + weight). */
int gc_read_load(gc_ctx* ctx, double* out);
/* The built system for record-by-record checks (tests/gpu.rs:356-436):
 * rhs, f, last (3n each), pinv (9n, row-major), any may be NULL. */
int gc_read_system(gc_ctx* ctx, double* rhs, double* f, double* last, double* pinv9);
/* Per distance constraint (edges with k != 0 in input order, then seams):
 * direction (3), cross factor; per bending constraint: 4 gradients (12). */
int gc_read_records(gc_ctx* ctx, double* richtung, double* quer, double* biege_grad);

/* --- self-collision lists of the last build / step -------------------------- */

/* Proximity pairs (selbstkollision::finde, :387-398) of the last build, in the
 * reference's order. art: 0 point-triangle, 1 edge-edge. Writes up to max,
 * returns the count. */
int gc_read_naehe(gc_ctx* ctx, int32_t max, int32_t* art, uint32_t* punkte4, double* gewichte4, double* normale3,
                  double* abstand, double* ruhe);
/* Springs of the last build (`System::federn`, implizit.rs:154): proximity
 * springs then push springs; block row-major. */
int gc_read_federn(gc_ctx* ctx, int32_t max, uint32_t* punkte5, double* gewichte5, double* kraft3, double* block9);
/* Cuts of the last build with their region roots (finde_schnitte, gebiete). */
int gc_read_schnitte(gc_ctx* ctx, int32_t max, uint32_t* kante_dreieck, double* t, double* w3, double* richtung3,
                     double* gewicht, uint32_t* wurzel);
/* Impacts found on the path of the last step (finde_stoesse, :1196-1282), in
 * the reference's order; gilt = 1 when ohne_durchdrungene kept it. */
int gc_read_stoesse(gc_ctx* ctx, int32_t max, int32_t* art, uint32_t* punkte4, double* lage2, double* zeit,
                    double* seite, double* normale3, int32_t* gilt);
/* Velocities (after a build they carry the proximity damping). */
int gc_get_v(gc_ctx* ctx, double* v);

/* --- the whole step on the device ----------------------------------------- */

/* `ImplizitCg::schritt` (implizit.rs:899-1038) for n substeps of length h:
 * build, warm start, solve, release pass and second solve, friction, position
 * update, push-out. Bodies stay where they are. */
int gc_step(gc_ctx* ctx, double h, int32_t n);
/* `Rechenweg::bild` for ImplizitCg (implizit.rs:1042-1057): schritte substeps of
 * bild_dt/schritte, bodies moved from their frame-start position by
 * v*(k+1)*h before substep k. */
int gc_bild(gc_ctx* ctx, double bild_dt, int32_t schritte);
/* Capture one substep of length h into a HIP graph and replay it n times
 * (same semantics as gc_step). */
int gc_step_graph(gc_ctx* ctx, double h, int32_t n);
int gc_read_statistik(gc_ctx* ctx, gc_statistik* out, int32_t reset);
/* Per-kernel device timings of the last gc_step/gc_step_graph call (HIP events
 * on the stream the kernels run on): total ms and PCG-kernel ms. */
int gc_zeiten(gc_ctx* ctx, double* ms_schritt, double* ms_pcg);

/* Batches. The reference steps one `ImplizitCg` per scene (implizit.rs:899-1038);
 * a batch steps many independent systems (contexts with topology and
 * parameters, same precision and device) as one: each substep runs every
 * system's build, then ONE persistent PCG launch in which each system owns a
 * group of blocks with its own reductions, iterations and stopping test, the
 * release passes, a second such launch, and every system's update. Per system
 * the substep is `ImplizitCg::schritt`, as gc_step computes it; only the
 * grouping of the reductions' block partials follows the batch's launch shape.
 * threads: power of two in [64, 1024] (0: 1024); bloecke: total blocks (0: all
 * co-resident ones). The contexts must outlive the batch and are not to be
 * stepped on their own while a batch call runs. NULL on failure with the
 * reason in err. */
typedef struct gc_batch gc_batch;
gc_batch* gc_batch_create(gc_ctx* const* ctx, int32_t n, int32_t threads, int32_t bloecke, char* err, int32_t err_len);
void gc_batch_destroy(gc_batch* b);
const char* gc_batch_last_error(const gc_batch* b);
/* n substeps of length h for every system of the batch, eager launches. */
int gc_batch_step(gc_batch* b, double h, int32_t n);
/* The same batched substep captured once into a HIP graph and replayed n times. */
int gc_batch_step_graph(gc_batch* b, double h, int32_t n);
/* Chooses the launch shape (threads per block; total blocks; blocks per system
 * in proportion to its tiles or to tiles x measured iterations) by timing the
 * batched solve of the current substep from one warm start per candidate; the
 * state is left as it was. Returns the chosen threads, total blocks and the
 * solve time in ms. */
int gc_batch_autotune(gc_batch* b, double h, int32_t* threads, int32_t* bloecke, double* ms);
/* Blocks per system (n values) and threads per block of the current shape. */
int gc_batch_bloecke(gc_batch* b, int32_t* bloecke, int32_t* threads);
/* Device time of the last gc_batch_step(_graph) call: total ms and the batched
 * PCG launches' ms (NaN for graph replays). */
int gc_batch_zeiten(gc_batch* b, double* ms_schritt, double* ms_pcg);

/* --- adjoint of the step and saved states --------------------------------- */

/* No counterpart in the reference, which has no gradient: its step
 * (ImplizitCg::schritt, implizit.rs:899-1038) run in reverse so that material
 * stiffnesses and friction can be fitted to target drapes.
 *
 * gc_adjoint_step runs one substep of length h from the context's state,
 * exactly as gc_step(ctx, h, 1), and then its reverse: the context's adjoint
 * seed, dL/dx and dL/dv of the state AFTER the step, is replaced by dL/dx and
 * dL/dv of the state BEFORE it, and the step's parameter gradients are added
 * to the accumulators. The discrete decisions of the step (contacts after the
 * release pass, stick or slip, path entries, push-outs) are held as the step
 * took them; within them the gradient is exact (per element and per point by
 * dual numbers over the step's own functions; one adjoint solve with the
 * step's modified PCG, reductions in FP64). Refused while cloth
 * self-collision or normalkraft_aus_last is on. A rollout runs its steps
 * forward saving states (gc_zustand_sichern), then for each step from the
 * last: gc_zustand_laden of the state before it, gc_adjoint_step. */
int gc_adjoint_step(gc_ctx* ctx, double h);
/* A rollout's reverse pass re-runs each segment from its checkpoint to get the
 * states before its steps; gc_step_aufnehmen takes such a step (as
 * gc_step(ctx, h, 1)) after saving the state into slot, and keeps the step's
 * two solutions with it. gc_adjoint_step_aus loads that state and runs
 * gc_adjoint_step with the recorded solutions in place of the forward's two
 * PCG solves (the same dv, to the bit), so the step's forward is not solved
 * a second time. The slot is valid for the same h, parameters and bodies as
 * the recording. */
int gc_step_aufnehmen(gc_ctx* ctx, double h, int32_t slot);
int gc_adjoint_step_aus(gc_ctx* ctx, double h, int32_t slot);
/* Seed / read the adjoint state: dL/dx and dL/dv (double[3n] each; NULL sets zero). */
int gc_adjoint_set(gc_ctx* ctx, const double* gx, const double* gv);
int gc_adjoint_get(gc_ctx* ctx, double* gx, double* gv);
/* Accumulated parameter gradients since the last reset: per input edge dL/dk
 * (0 for the k = 0 edges a membrane carries), per seam dL/dk (its stiffness is
 * gc_param.naht_k), per bending constraint dL/dk, per membrane
 * dL/d(E_k, E_s, G) (3 per element, the membran_daten fields 9-11), per point
 * dL/d(friction coefficient). Any pointer may be NULL. */
int gc_adjoint_gradient(gc_ctx* ctx, double* d_kanten, double* d_naehte, double* d_biegung, double* d_membran,
                        double* d_reibung, int32_t reset);
typedef struct gc_adjoint_stat {
  int64_t schritte;           /* adjoint steps since the last reset */
  int64_t cg_iterationen;     /* iterations of their adjoint solves */
  int64_t nicht_konvergiert;  /* adjoint solves that stopped above the tolerance */
  int64_t abweichungen;       /* points whose recomputed update differed from the step's (0 expected) */
  int64_t eintritte;          /* path entries differentiated */
  double rest_max;            /* largest relative residual of an adjoint solve */
} gc_adjoint_stat;
int gc_adjoint_statistik(gc_ctx* ctx, gc_adjoint_stat* out, int32_t reset);
/* New material values on the uploaded topology: per input edge k (the k = 0
 * pattern must stay), per bending constraint k, per membrane (E_k, E_s, G),
 * per point friction coefficient; NULL leaves a group as it is. The builder
 * writes these values when it builds the scene (szene.rs:425-461, 602-647). */
int gc_set_steifigkeit(gc_ctx* ctx, const double* kanten_k, const double* biegung_k, const double* membran_ek_es_g,
                       const double* reibung);
/* The discrete decisions of a step held fixed, for finite differences of the
 * branch the adjoint differentiates (no counterpart in the reference): modus 1
 * records, per point, the step's contact, its releases after solve 1 and 2,
 * stick or slip, path entry and push-out (with their velocity branches) into
 * slot `slot`; modus 2 replays that slot's record instead of deciding (a
 * contact whose point moved out by a rounding's width keeps its normal and
 * a depth <= 0); modus 0 decides as the reference does. One body at most,
 * self-collision off. The setting holds for the following steps until
 * changed. gc_entscheidungen_abweichungen counts replayed decisions the state
 * no longer allowed (taken as decided instead). */
int gc_entscheidungen(gc_ctx* ctx, int32_t modus, int32_t slot);
int gc_entscheidungen_abweichungen(gc_ctx* ctx, int64_t* anzahl, int32_t reset);
/* Save / restore the state (x, v and the warm start of the next step) in
 * device memory under a slot number, e.g. every k steps of a rollout. */
int gc_zustand_sichern(gc_ctx* ctx, int32_t slot);
int gc_zustand_laden(gc_ctx* ctx, int32_t slot);

/* ------------------------------------------------------------------------
 * Seam pairing by length fraction (klon3d-mesh rand.rs:297-501, Randplan
 * rand.rs:503-589, pairs emitted as in lib.rs:351-381). Host code: no device
 * is touched. Input: pattern pieces whose boundary edges are polylines in
 * mm, and seams whose two sides are runs of whole edges (Nahtseite). Output:
 * every piece's boundary points (the first points of its mesh, in key order
 * of edge and arc length, rand.rs:536-583), every seam's stations
 * (Nahtstelle: end, shared, only A, only B) and its sub-seams (Teilnaht) with
 * the stitched point pairs (Punktpaar). Boundary splits the mesher's
 * refinement asks for (Randplan::teile_auf, rand.rs:591-604) are not taken.
 * ---------------------------------------------------------------------- */
typedef struct gc_naht_eingabe {
  int32_t n_teile;
  const int32_t* teil_kanten;   /* n_teile + 1: edges of piece t are teil_kanten[t] .. teil_kanten[t+1] - 1 */
  const int32_t* kante_punkte;  /* n_kanten + 1: polyline points of edge e */
  const double* punkte;         /* x, y (mm) per polyline point */
  int32_t n_naehte;
  const int32_t* seite_stuecke; /* 2 n_naehte + 1: side A of seam n is 2n, side B 2n + 1; its runs */
  const int32_t* stueck_teil;   /* per run: the piece */
  const int32_t* stueck_rueck;  /* per run: 1 = the seam runs the edges backwards */
  const int32_t* stueck_kanten; /* runs + 1: the run's edges (indices within the piece) */
  const int32_t* kanten;
  double kante_mm;              /* target edge length (10) */
  double sehne_mm;              /* largest chord height (0.1) */
  double laenge_mm;             /* largest length loss per boundary edge (0.1) */
  double fang_mm;               /* stations at most this far apart on A and B are merged (2) */
  double min_rand_mm;           /* shortest boundary piece a split may make (0.5) */
} gc_naht_eingabe;

typedef struct gc_naht_ausgabe {
  /* capacities in (entries), counts out */
  int64_t kap_rand, n_rand;
  int32_t* rand_teil;           /* n_teile + 1: boundary points of piece t (offsets) */
  int32_t* rand_kante;          /* per boundary point: its edge */
  double* rand_s;               /* per boundary point: arc length from the edge's start (mm) */
  double* rand_xy;              /* per boundary point: x, y (mm) */
  int32_t* rand_knoten;         /* per edge (n_kanten): the boundary point of the node it starts at */
  int64_t kap_stellen, n_stellen;
  int32_t* naht_stellen;        /* n_naehte + 1 */
  int32_t* stelle_art;          /* 0 end, 1 shared, 2 only A, 3 only B (Stellenart) */
  int32_t* stelle_knoten;       /* piece and edge whose start node the station is on A, then on B; -1 -1 for a new point */
  double* stelle_rho;           /* length fraction on A and B */
  double* stelle_versatz_mm;    /* (rho_A - rho_B) x mean side length */
  int64_t kap_teilnaehte, n_teilnaehte;
  int32_t* naht_teilnaehte;     /* n_naehte + 1 */
  int32_t* teilnaht_teile;      /* piece on A, piece on B */
  int32_t* teilnaht_paare;      /* n_teilnaehte + 1 */
  int64_t kap_paare, n_paare;
  int32_t* paar_punkte;         /* boundary point on A, on B (indices within the piece's points) */
  double* paar_rho;             /* length fraction on A, on B */
  double* naht_laengen_mm;      /* 2 per seam: length of side A, of side B (mm) */
  char fehler[256];             /* the reason of a failed plan */
} gc_naht_ausgabe;
/* 0: planned; 1: the pattern cannot be planned (reason in fehler, as the
 * reference's Err: an edge in two seams, stations that coincide, a section
 * across an edge boundary); 2: a capacity is too small (all counts are set,
 * nothing else written); -1: bad arguments. */
int gc_naht_planen(const gc_naht_eingabe* ein, gc_naht_ausgabe* aus);

#ifdef __cplusplus
}
#endif
#endif
