/*
 * gcoracle — FP64 CPU restatement of klon3d-sim's implicit backward-Euler step.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker that the HIP path is
 * compared against; it is never linked into, loaded by, or called from the
 * product library. Only tests/, the smoke check and bench.py's cpu_baseline leg
 * may load it.
 *
 * Every function restates a reference function with the same loop order,
 * summation order and branch order; the citation is given at each definition in
 * gcoracle.cpp (paths relative to the reference checkout, crates/klon3d-sim/src).
 *
 * Layout conventions (same as the product C ABI):
 *   points:   double[3*n] interleaved (x0,y0,z0,x1,...) like Vec<[f64;3]>
 *   membrane: double[12] per element =
 *             {flaeche, ruhe_inv[4], kette[2], schuss[2], kette_n_m, dehnung_n_m, schub_n_m}
 *   hessen:   double[9]  per element = {f1[3], f2[3], w[3]}
 */
#ifndef GCORACLE_H
#define GCORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- element-level restatements (kontinuum.rs, cpu.rs, biegung.rs, implizit.rs) ---- */
int    or_membran_bauen(const double ruhe[6], const double kanten[3], const double dir2[2],
                        const double material[3], double out_elem[12]);
void   or_membran_dehnung(const double elem[12], const double p[9], double out[3]);
void   or_membran_spannung(const double elem[12], const double p[9], double out[3]);
double or_membran_energie(const double elem[12], const double p[9]);
void   or_membran_krafte(const double elem[12], const double p[9], double out[9]);
void   or_membran_krafte_zug(const double elem[12], const double p[9], double out[9]);
void   or_membran_hessen(const double elem[12], const double p[9], double out_h[9]);
void   or_membran_mal_h_aus(const double elem[12], const double h[9], const double delta[9], double out[9]);
void   or_membran_diag_bloecke_aus(const double elem[12], const double h[9], double out[27]);
void   or_psd2(double a, double b, double c, double out[3]);
int    or_dieder(const double q[12], double* theta, double grad[12]);
double or_winkel_fehler(double theta, double ruhe);
double or_steifigkeit(double weft, double warp, double bias, double theta);
double or_steifigkeit_aus(double c2, double s2, double weft, double warp, double bias);
int    or_cos2_sin2(const double e[2], const double warp[2], double out[2]);
void   or_m3_inv(const double m[9], double out[9]);

/* ---- bodies ---- */
/* Exact signed distance of the ellipsoid-cylinder union used by the drape scenes
 * (ellipsoid by Eberly's bisection on the Lagrange multiplier, capped cylinder
 * closed form, union by min). Used to check the sampled grid. */
double or_koerper_sdf(const double ell_c[3], const double ell_r[3], const double cyl_c[2], double cyl_r,
                      double cyl_z0, double cyl_z1, const double x[3]);
/* Trilinear sample of a grid SDF and its gradient; returns 0 outside the grid. */
int    or_sdf_sample(const float* werte, const int dims[3], const double origin[3], double spacing,
                     const double q[3], double* d, double grad[3]);

/* ---- triangle-mesh body (netzkoerper.rs) ---- */
typedef struct or_netz or_netz;
typedef struct or_naechster {
  double abstand;     /* outside positive */
  double punkt[3];
  double normale[3];
  uint32_t dreieck;
  int32_t region;     /* 0..2 corner k, 3..5 edge k (p0p1, p1p2, p2p0), 6 face */
} or_naechster;
/* Koerpernetz::neu (netzkoerper.rs:297-388); NULL and a message on failure. */
or_netz* or_netz_neu(int32_t n_punkte, const double* punkte, int32_t n_dreiecke, const uint32_t* dreiecke, double suchtiefe,
                     char* fehler, int32_t fehler_len);
void     or_netz_frei(or_netz* k);
int      or_netz_nachfuehren(or_netz* k, const double* punkte, char* fehler, int32_t fehler_len);
int      or_netz_punktgeschwindigkeit(or_netz* k, const double* v, int32_t n); /* n = 0 clears */
/* randkanten, volumen, grid origin[3], cell, counts[3] */
void     or_netz_info(const or_netz* k, int64_t* randkanten, double* volumen, double ursprung[3], double* zelle, int64_t zahl[3]);
/* Grid of triangles (punkte = 0) or of points (punkte = 1): sizes, then the arrays. */
void     or_netz_gitter(const or_netz* k, int32_t punkte, int64_t* n_anfang, int64_t* n_eintraege, uint32_t* anfang, uint32_t* eintraege);
void     or_netz_normalen(const or_netz* k, double* flaechennormale, double* punktnormale, uint32_t* nachbar, double* kugel4);
int      or_netz_naechster(const or_netz* k, const double x[3], double radius, or_naechster* out);
void     or_netz_naechster_alle(const or_netz* k, const double x[3], or_naechster* out);
void     or_netz_abstand_ohne_grenze(const or_netz* k, const double x[3], or_naechster* out);
int      or_netz_durchstoss(const or_netz* k, const double a[3], const double b[3], const double starr[3], double p[3],
                            double n[3], double vk[3]);
double   or_netz_geschwindigkeit_an(const or_netz* k, int32_t t, const double p[3], const double starr[3], double out[3]);
/* Koerpernetz::punkte_gegen_dreiecke on cloth triangles; x and v in place. */
int32_t  or_netz_punkte_gegen_dreiecke(const or_netz* k, double tief, const double versatz[3], const double geschw[3],
                                       int32_t nd, const uint32_t* dreiecke, int32_t n, const double* w, double* x, double* v);
/* Koerper::kontakt of a mesh body at offset `versatz` (koerper.rs:82-90). */
int      or_netz_kontakt(const or_netz* k, const double versatz[3], const double geschw[3], const double x[3], double abstand,
                         double normale[3], double* tiefe, double vk[3]);

/* ---- simulation ---- */
typedef struct or_szene {
  int32_t n;
  const double* x;                 /* 3n */
  const double* v;                 /* 3n */
  const double* masse;             /* n  */
  const double* w;                 /* n  */
  const double* kollisionsabstand; /* n  */
  const double* reibung;           /* n  */
  int32_t n_abstand;
  const uint32_t* abstand_punkte;  /* 2 per constraint */
  const double* abstand_ruhe;
  const double* abstand_nachgiebigkeit;
  const uint8_t* membran_kante;    /* 1 = carried by a membrane (k = 0) */
  int32_t n_naht;
  const uint32_t* naht_punkte;     /* 2 per seam spring */
  const double* naht_ruhe;         /* may be +inf (open seam) */
  int32_t n_biegung;
  const uint32_t* biegung_punkte;  /* 4 per constraint */
  const double* biegung_ruhe;
  const double* biegung_nachgiebigkeit;
  int32_t n_membran;
  const uint32_t* membran_punkte;  /* 3 per element */
  const double* membran_daten;     /* 12 per element */
  int32_t n_dreiecke;              /* cloth triangles (self-collision); may be 0 */
  const uint32_t* dreiecke;
  const uint32_t* teil_von_punkt;  /* piece index per point (self-collision); may be NULL */
  const double* selbst_e_punkt;    /* n: stretch stiffness (N/m) of the point's piece; may be NULL */
  const double* selbst_e_dreieck;  /* n_dreiecke: stretch stiffness of the triangle's piece; may be NULL */
  const double* selbst_a_dreieck;  /* n_dreiecke: rest area (m^2); may be NULL */
  double bild_dt;
  double schwerkraft[3];
  double daempfung;
} or_szene;

typedef struct or_param {
  int32_t schritte_je_bild;
  int32_t cg_max;
  double cg_toleranz;
  double naht_k;
  double kontakt_rand;
  int32_t normalkraft_aus_last;
  int32_t selbstkollision;
  double selbst_zusatzdicke;
  double selbst_steifigkeit;
  int32_t selbst_stoss_iterationen;
  int32_t selbst_schnitte_an_naehten;
  int32_t durchstoss_pruefen;     /* implizit.rs:476, default on */
  int32_t koerperpunkt_durchgaenge; /* implizit.rs:478-479, default 0 (off) */
  double koerperpunkt_tiefe;        /* implizit.rs:480-481, default 0.01 m */
} or_param;

typedef struct or_statistik {
  int64_t cg_iterationen, loesungen, nicht_konvergiert, kontakte, haftend, freigegeben;
  double rest_max;
  int64_t naehe_punkt_dreieck, naehe_kante_kante, selbst_stoesse;
  int64_t durchdringungen, durchdringungsgebiete;
  int64_t durchstoesse;
  int64_t koerperpunkte;
} or_statistik;

typedef struct or_sim or_sim;

/* Contacts of the last XPBD frame (cpu.rs:63-73). */
typedef struct or_xpbd_statistik {
  int64_t kontakte, haftend;
  double tiefe, tangential;
} or_xpbd_statistik;

void    or_param_default(or_param* p);
or_sim* or_sim_neu(const or_szene* s, const or_param* p);
void    or_sim_frei(or_sim* s);
int     or_sim_koerper_zylinder(or_sim* s, const double punkt[3], const double achse[3], double radius, const double geschw[3]);
int     or_sim_koerper_sdf(or_sim* s, const float* werte, const int dims[3], const double origin[3], double spacing,
                           const double versatz[3], const double geschw[3]);
/* A mesh body (a copy of k, shared between the frames like Arc<Koerpernetz>). */
int     or_sim_koerper_netz(or_sim* s, const or_netz* k, const double versatz[3], const double geschw[3]);
void    or_sim_param(or_sim* s, const or_param* p);
void    or_sim_lage(const or_sim* s, double* x, double* v);
void    or_sim_setze_lage(or_sim* s, const double* x, const double* v);
void    or_sim_setze_naht_ruhe(or_sim* s, const double* ruhe);
void    or_sim_setze_w(or_sim* s, const double* w);
void    or_sim_setze_dv_vorher(or_sim* s, const double* dv, int32_t n); /* n = 0 clears the warm start */
int32_t or_sim_dv_vorher(const or_sim* s, double* out);            /* returns its length (0 = none) */
void    or_sim_bild(or_sim* s);
/* One frame of the XPBD path `CpuReferenz::bild` (cpu.rs:226-245) on the
 * sim's scene: `unterschritte` substeps of bild_dt / unterschritte, seams with
 * compliance naht_nachgiebigkeit (szene.rs:545). */
void    or_sim_xpbd_bild(or_sim* s, int32_t unterschritte, double naht_nachgiebigkeit, or_xpbd_statistik* out);
void    or_sim_schritt(or_sim* s, double h);
void    or_sim_statistik(const or_sim* s, or_statistik* out);
void    or_sim_reset_statistik(or_sim* s);

/* System of the next substep at the current state (implizit.rs system_bauen),
 * kept inside the handle for the probes below. */
void    or_sys_bauen(or_sim* s, double h);
int32_t or_sys_n_kanten(const or_sim* s);   /* edges followed by seams */
int32_t or_sys_n_biegungen(const or_sim* s);
void    or_sys_kanten(const or_sim* s, uint32_t* ij, double* k, double* richtung, double* quer, double* ruhe);
void    or_sys_biegungen(const or_sim* s, uint32_t* punkte, double* grad, double* k);
void    or_sys_membran_h(const or_sim* s, double* out_h9);
void    or_sys_vektoren(const or_sim* s, double* rhs, double* f, double* last, double* pinv9);
void    or_sys_diag(const or_sim* s, double* diag9);
void    or_sys_fest(const or_sim* s, int32_t* flag, double* normale, double* vn_soll, double* v_koerper);
void    or_sys_setze_fest(or_sim* s, const int32_t* flag, const double* normale, const double* vn_soll, const double* v_koerper);
void    or_sys_mal_h(const or_sim* s, const double* p, double* out);
void    or_sys_mal_a(const or_sim* s, const double* p, double* out);
/* CpuLoeser::loese on the built system and contact set; dv in = warm start. */
void    or_sys_loese(const or_sim* s, double* dv, int32_t* iterationen, int32_t* nicht_konvergiert, double* rest);
void    or_sys_rest(const or_sim* s, const double* dv, double* out);

/* ---- self-collision (selbstkollision.rs) ---- */
/* Proximity pairs of the current state (finde, :387-398), in the reference's
 * order; no side effects. Writes up to `max` entries, returns the count.
 * art: 0 point-triangle, 1 edge-edge. brute != 0: every entry instead of the
 * hash's buckets (the check of the hash). */
int32_t or_selbst_naehe(or_sim* s, int32_t brute, int32_t max, int32_t* art, uint32_t* punkte4, double* gewichte4,
                        double* normale3, double* abstand, double* ruhe);
/* Springs of the last built system (proximity springs, then push springs). */
int32_t or_sys_federn(const or_sim* s, int32_t max, uint32_t* punkte5, double* gewichte5, double* kraft3, double* block9);
/* Velocities after the last build (the proximity damping changes them). */
void    or_sim_v(const or_sim* s, double* v);
/* Cuts of the current state with their region roots (finde_schnitte, gebiete). */
int32_t or_selbst_schnitte(or_sim* s, int32_t max, uint32_t* kante_dreieck, double* t, double* w3, double* richtung3,
                           double* gewicht, uint32_t* wurzel);
/* Impacts on the path x0 -> x1 of this scene's mesh (finde_stoesse, :1196-1282). */
int32_t or_selbst_stoesse(or_sim* s, const double* x0, const double* x1, int32_t max, int32_t* art, uint32_t* punkte4,
                          double* lage2, double* zeit, double* seite, double* normale3);

/* The same searches on a bare mesh (Stoffnetz::neu from triangles and radii,
 * no seam partners); for the reference's own unit tests. With nt = 0 the
 * mesh has only the nk given edges. */
int32_t or_netz_naehe(int32_t n, const double* x, const double* w, int32_t nt, const uint32_t* dreiecke, const double* r_punkt,
                      int32_t punkt_dreieck, int32_t kante_kante, int32_t max, int32_t* art, uint32_t* punkte4,
                      double* gewichte4, double* normale3, double* abstand, double* ruhe);
/* finde_stoesse on x0 -> x1, then stossantwort on x1 (in place) with masse.
 * Returns the number of impacts found; *korrekturen = the response's count. */
int32_t or_netz_stoesse(int32_t n, const double* x0, double* x1, const double* w, const double* masse, int32_t nt,
                        const uint32_t* dreiecke, int32_t nk, const uint32_t* kanten, int32_t punkt_dreieck,
                        int32_t kante_kante, int32_t iterationen, int32_t max, int32_t* art, uint32_t* punkte4,
                        double* lage2, double* zeit, double* seite, int32_t* korrekturen);

#ifdef __cplusplus
}
#endif
#endif
