#!/usr/bin/env python3
"""Benchmark: implicit backward-Euler steps/s of the 40k-vertex woven drape at
PCG rtol 1e-6 on one MI355X, FP32 bulk with FP64 reductions, cloth
self-collision on.

One step = one ImplizitCg::schritt of dt = 1/120 s (self-collision pairs,
springs, cuts; build; PCG with the contact filter; release pass; friction;
update; impacts) with every input resident in HBM. Prints one JSON line with the throughput, the PCG kernel's roofline and
the FP64 CPU restatement timed on a bounded sample of the same workload.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--graph] [--praezision fp32|fp64] [--ohne-selbstkollision]
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

WURZEL = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, WURZEL)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, 8 TB/s
TRAFFIC_JE_ITERATION = 3585606
TRAFFIC_EINHEIT = "B per PCG iteration (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, KiB x 1024; steps 241-250, 136 x 1024)"


def pcg_bytes(s, praezision: str) -> dict:
    """Algorithmic bytes of one PCG iteration: every record and vector an
    iteration must touch, counted once (t = bytes per float):
    membrane 12 + 21t per triangle (3 indices, f1 f2, w, D^-1, basis,
    stiffness, area), bending 16 + 13t per hinge (4 indices, 4 gradients, k),
    seam or mesh edge with k != 0 8 + 5t, point 41t (p in, A p out, mass:
    7t; the fused vector pass 34t: p, A p, r, dv in, P^-1, filter, dv, r, p
    out). For t = 4 these are 96, 68, 28 and 164 B. The CSR index stream that
    the operator also reads is not counted."""
    t = 4 if praezision == "fp32" else 8
    n = s.n
    nm = s.membran_punkte.shape[0]
    nb = s.biegung_punkte.shape[0]
    nk = int(np.count_nonzero(~s.membran_kante)) + s.naht_punkte.shape[0]
    membran, biegung, kante, punkt = 12 + 21 * t, 16 + 13 * t, 8 + 5 * t, 41 * t
    matvec = nm * membran + nb * biegung + nk * kante + n * 7 * t
    iteration = matvec + n * 34 * t
    return {"iteration": iteration, "matvec": matvec, "start": iteration,
            "membranen": nm, "biegungen": nb, "kanten": nk,
            "je_einheit": {"membrane_B": membran, "bending_B": biegung, "edge_B": kante, "point_B": punkt}}


def cpu_modell() -> str:
    """The host CPU's model name (/proc/cpuinfo), else what platform knows."""
    try:
        with open("/proc/cpuinfo") as f:
            for zeile in f:
                if zeile.lower().startswith("model name"):
                    return zeile.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--praezision", default=None, choices=["fp32", "fp64"],
                    help="fp32 (FP32 bulk, FP64 reductions) by default; adjoint20k defaults to fp64")
    ap.add_argument("--graph", action="store_true", help="replay a captured HIP graph per step")
    ap.add_argument("--cpu-sekunden", type=float, default=25.0, help="time budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--ohne-vergleich", action="store_true",
                    help="stapel64: skip the sequential comparison (for counter runs)")
    ap.add_argument("--stapel-eager", action="store_true",
                    help="stapel64: eager batched steps instead of graph replays (for counter runs)")
    ap.add_argument("--bloecke", type=int, default=0)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--lds-kb", type=int, default=0, help="LDS a PCG tile may fill (KiB); 0 = the library's choice")
    ap.add_argument("--ohne-selbstkollision", action="store_true", help="self-collision off (the plain drape)")
    ap.add_argument("--koerperpunkte", type=int, default=0,
                    help="netz40k: passes of body points against cloth triangles per step (0, the reference's "
                         "default, switches them off)")
    ap.add_argument("--workload", default="drape40k",
                    choices=["drape40k", "stapel64", "adjoint20k", "gross250k", "netz40k", "xpbd40k"],
                    help="drape40k: the 40k-point drape; stapel64: 64 garments of 2k-60k points in one batch; "
                         "adjoint20k: gradient of a 60-step 20k-point drape rollout w.r.t. per-panel stiffnesses "
                         "and friction; gross250k: the 250k-point drape, FP32 bulk against all-FP64; "
                         "netz40k: the 40k drape on the triangle-mesh body (contacts and path entries by mesh "
                         "query); xpbd40k: XPBD frames (CpuReferenz) of the 40k drape")
    ap.add_argument("--vorlauf", type=int, default=120, help="adjoint20k: settling steps before the rollout starts")
    ap.add_argument("--fd-schritt", type=float, default=1e-4, help="adjoint20k: relative step of the central differences")
    ap.add_argument("--fd", type=int, default=0, help="adjoint20k: central differences of the FP64 forward for this "
                                                      "many parameters (largest gradients first, plus friction)")
    args = ap.parse_args()

    if args.praezision is None:
        # the adjoint's default is FP64: at rtol 1e-6 the FP32 and FP64 forwards stop at
        # different points of the tolerance band, their contact decisions part over a
        # rollout, and each adjoint is the derivative of its own branch
        args.praezision = "fp64" if args.workload == "adjoint20k" else "fp32"
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: one garment's step lives on one device, and "
                                   "sharding across GPUs is not built", "n_gpus": args.gpus}))
        return 2

    if args.workload == "stapel64":
        return stapel_lauf(args)
    if args.workload == "adjoint20k":
        return adjoint_lauf(args)
    if args.workload == "gross250k":
        return gross_lauf(args)
    if args.workload == "xpbd40k":
        return xpbd_lauf(args)

    from garmentcraft_amd import szenen
    from garmentcraft_amd.geraet import GeraetSzene, ImplizitParameter

    t0 = time.time()
    netz = args.workload == "netz40k"
    if netz:
        s, bericht = szenen.drape(mit_koerper=False)
        s.koerper = szenen.koerper_netz(kante=0.01)
        bericht["koerper_dreiecke"] = [int(k.dreiecke.shape[0]) for k in s.koerper]
    else:
        s, bericht = szenen.drape(sdf=szenen.koerper_sdf())
    print(f"[bench] scene {bericht} built in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    dt = s.param.bild_dt
    selbst = not args.ohne_selbstkollision
    param = ImplizitParameter(schritte_je_bild=1, cg_max=10000, cg_toleranz=1e-6, naht_k=500.0, selbstkollision=selbst,
                              koerperpunkt_durchgaenge=args.koerperpunkte)
    g = GeraetSzene(s, param, praezision=args.praezision, bloecke=args.bloecke, threads=args.threads,
                    lds_kb=args.lds_kb)
    print(f"[bench] device: {g.name()}", file=sys.stderr, flush=True)

    step = g.schritte_graph if args.graph else g.schritte
    for k in range(args.warmup):
        step(dt, 1)
        if (k + 1) % 50 == 0:
            print(f"[bench] warmup step {k + 1}", file=sys.stderr, flush=True)
    g.statistik(reset=True)
    x_w, v_w = g.lage()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)

    t_a = time.perf_counter()
    step(dt, args.steps)
    t_b = time.perf_counter()
    ms_gesamt, ms_pcg = g.zeiten()
    st = g.statistik()
    wall_ms = (t_b - t_a) * 1e3
    ms_per_step = wall_ms / args.steps
    value = args.steps / (wall_ms / 1e3)
    print(f"[bench] {args.steps} steps in {wall_ms:.1f} ms (device {ms_gesamt:.1f} ms, pcg {ms_pcg:.1f} ms), "
          f"{st['cg_iterationen']} CG iterations, {st['nicht_konvergiert']} unconverged solves, "
          f"longest barrier wait {st['barriere_spins_max']} polls; self-collision: "
          f"{st['naehe_punkt_dreieck']} point-triangle and {st['naehe_kante_kante']} edge-edge pairs, "
          f"{st['durchdringungen']} cuts, {st['selbst_stoesse']} impacts; "
          f"{st['cg_rundungsgrenze']} FP32 solves ended at the rounding floor", file=sys.stderr, flush=True)

    b = pcg_bytes(s, args.praezision)
    iters = st["cg_iterationen"]
    loesungen = st["loesungen"]
    roofline = None
    if not args.graph and ms_pcg == ms_pcg and ms_pcg > 0:
        # both PCG launches of every step (first solve, solve after a
        # release), bracketed by HIP events on the stream they run on
        algo = iters * b["iteration"] + loesungen * b["start"]
        achieved = algo / (ms_pcg * 1e-3) / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_pcg", "bytes_per_iteration": b["iteration"], "bytes_per_unit": b["je_einheit"],
                    "us_per_iteration": round(ms_pcg * 1e3 / max(iters, 1), 3)}
        if args.praezision == "fp32":
            # a fixed reference measurement, not of this run: rocprofv3 --pmc
            # FETCH_SIZE and WRITE_SIZE over the k_pcg dispatches of steps
            # 241-250 of this drape at the default launch shape, scaled by the
            # counters of a plain copy of known size
            # (profiles/k_pcg_drape40k_pmc.md): bytes per PCG iteration
            roofline["traffic"] = TRAFFIC_JE_ITERATION
            roofline["traffic_unit"] = TRAFFIC_EINHEIT

    cpu = None
    if not args.no_cpu:
        import oracle as O
        o = O.OrakelSim(s, O.param(schritte_je_bild=1, cg_max=10000, cg_toleranz=1e-6, naht_k=500.0,
                                   selbstkollision=1 if selbst else 0, koerperpunkt_durchgaenge=args.koerperpunkte))
        o.setze_lage(x_w, v_w)
        # warm start of the oracle: the device's dv of the last warm-up step is
        # not exported; start from h*g like a fresh solver, which costs the CPU
        # at most a few extra iterations on the first sampled step.
        t_c = time.perf_counter()
        n_cpu = 0
        while True:
            o.schritt(dt)
            n_cpu += 1
            if time.perf_counter() - t_c > args.cpu_sekunden or n_cpu >= args.steps:
                break
        t_d = time.perf_counter()
        st_o = o.statistik()
        cpu = {"value": round(n_cpu / (t_d - t_c), 4), "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"{n_cpu} steps of the same drape from the state after {args.warmup} warm-up steps, "
                         f"FP64 C++ restatement of CpuLoeser/ImplizitCg::schritt, {st_o['cg_iterationen']} CG iterations",
               "cpu": cpu_modell(), "build": "g++ -O3 (oracle/Makefile)"}

    out = {
        "metric": "implicit steps/s @40k verts, PCG rtol 1e-6; batched/adjoint vs plain GPU vs CPU",
        "value": round(value, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.praezision == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": "drape: 28 woven panels, 24 seams, " +
                               ("ellipsoid and cylinder as triangle meshes of 1 cm edges (" +
                                " + ".join(str(t) for t in bericht["koerper_dreiecke"]) + " triangles), "
                                if netz else "ellipsoid-cylinder grid SDF 256^3 at 3 mm, ") +
                               "friction 0.3, " + ("self-collision on, " if selbst else "self-collision off, ") +
                               "dt 1/120, rtol 1e-6",
                   "timed_steps": f"{args.warmup + 1}-{args.warmup + args.steps}",
                   "punkte": s.n, "dreiecke": int(s.dreiecke.shape[0]), "biegungen": b["biegungen"],
                   "selbstkollision": selbst, "graph": bool(args.graph), "praezision": args.praezision,
                   "cg_iterationen_je_schritt": round(iters / args.steps, 1), "loesungen": loesungen,
                   "nicht_konvergiert": st["nicht_konvergiert"], "rundungsgrenze": st["cg_rundungsgrenze"],
                   "ms_pcg_je_schritt": round(ms_pcg / args.steps, 3) if ms_pcg == ms_pcg else None,
                   "ms_geraet_je_schritt": round(ms_gesamt / args.steps, 3) if ms_gesamt == ms_gesamt else None,
                   "naehe_je_schritt": round((st["naehe_punkt_dreieck"] + st["naehe_kante_kante"]) / args.steps, 1),
                   "stoesse": st["selbst_stoesse"], "durchdringungen": st["durchdringungen"],
                   **({"koerperpunkt_durchgaenge": args.koerperpunkte, "koerperpunkte": st["koerperpunkte"]}
                      if args.koerperpunkte else {}),
                   "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    return 0


def xpbd_lauf(args) -> int:
    """XPBD frames (CpuReferenz::bild, cpu.rs:226-245) of the 40k drape on the
    grid-SDF body: 20 substeps of a 1/60 s frame, each one round of distance,
    membrane, seam and bending constraints in the reference's Gauss-Seidel
    order, body contact with friction. Against the FP64 restatement on one
    host core, from the same state."""
    from garmentcraft_amd import szenen
    from garmentcraft_amd.geraet import GeraetSzene, ImplizitParameter

    steps = args.steps if args.steps != 300 else 60
    warmup = args.warmup if args.warmup != 10 else 20
    t0 = time.time()
    s, bericht = szenen.drape(sdf=szenen.koerper_sdf())
    s.param.bild_dt = 1.0 / 60.0
    n = s.param.unterschritte
    print(f"[bench] scene {bericht} built in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    g = GeraetSzene(s, ImplizitParameter(), praezision="fp64")
    g.xpbd_topologie(s)
    laeufe = g.xpbd_laeufe()
    print(f"[bench] runs and levels {laeufe}", file=sys.stderr, flush=True)
    for _ in range(warmup):
        g.xpbd_bild(n)
    x_w, v_w = g.lage()
    t_a = time.perf_counter()
    kontakte = 0
    for _ in range(steps):
        kontakte += g.xpbd_bild(n)["kontakte"]
    t_b = time.perf_counter()
    wall_ms = (t_b - t_a) * 1e3
    value = steps / (wall_ms / 1e3)
    cpu = None
    if not args.no_cpu:
        import oracle as O
        o = O.OrakelSim(s, O.param())
        o.setze_lage(x_w, v_w)
        t_c = time.perf_counter()
        n_cpu = 0
        while True:
            o.xpbd_bild(n)
            n_cpu += 1
            if time.perf_counter() - t_c > args.cpu_sekunden or n_cpu >= steps:
                break
        t_d = time.perf_counter()
        cpu = {"value": round(n_cpu / (t_d - t_c), 4), "unit": "frames/s", "cores": 1, "kind": "port",
               "sample": f"{n_cpu} XPBD frames of the same drape from the state after {warmup} warm-up frames, "
                         f"FP64 C++ restatement of CpuReferenz::bild", "cpu": cpu_modell(),
               "build": "g++ -O3 (oracle/Makefile)"}
    out = {"metric": "XPBD frames/s @40k verts (CpuReferenz, 20 substeps of a 1/60 s frame)", "value": round(value, 3),
           "unit": "frames/s", "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": round(wall_ms / steps, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "drape: 28 woven panels on the grid-SDF body, XPBD, 20 substeps per 1/60 s frame",
                      "punkte": s.n, "laeufe": laeufe,
                      "us_je_unterschritt": round(wall_ms * 1e3 / (steps * n), 1),
                      "kontakte_je_bild": round(kontakte / steps, 1), "parallelism": "1 GPU"},
           "roofline": None, "cpu_baseline": cpu}
    print(json.dumps(out))
    return 0


def stapel_lauf(args) -> int:
    """64 garments of 2k-60k points (log-uniform), latex or woven, stiffness
    over one decade, friction 0.1-0.5, all on the drape's body: one batched
    substep per step (one persistent PCG launch per solve pass for all of them,
    replayed from a captured graph, launch shape from the autotuner) against
    the same 64 systems stepped one after another with plain gc_step calls."""
    from garmentcraft_amd import szenen
    from garmentcraft_amd.geraet import GeraetStapel, GeraetSzene, ImplizitParameter

    steps = args.steps if args.steps != 300 else 10
    # 90 warm-up steps: the garments lie on the body, with contacts, sticking
    # and proximity pairs, when the timed steps start
    warmup = args.warmup if args.warmup != 10 else 90
    t0 = time.time()
    sdf = szenen.koerper_sdf()
    satz = szenen.stapel(sdf=sdf, prozesse=min(16, os.cpu_count() or 1))
    punkte = [s.n for s, _ in satz]
    print(f"[bench] {len(satz)} scenes, {sum(punkte)} points ({min(punkte)}-{max(punkte)}) built in "
          f"{time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    dt = 1.0 / 120.0
    selbst = not args.ohne_selbstkollision
    param = ImplizitParameter(schritte_je_bild=1, cg_max=10000, cg_toleranz=1e-6, naht_k=500.0, selbstkollision=selbst)

    def kontexte():
        gs = []
        for s, _ in satz:
            gs.append(GeraetSzene(s, param, praezision=args.praezision, koerper_quelle=gs[0] if gs else None,
                                  lds_kb=args.lds_kb))
        return gs

    A = kontexte()
    st = GeraetStapel(A)
    st.schritte(dt, 1)
    wahl = st.autotune(dt)
    print(f"[bench] autotune: {wahl}", file=sys.stderr, flush=True)
    # graph replays by default; --stapel-eager for counter runs (the profiler's
    # counter collection does not take graph launches)
    schritt = st.schritte if args.stapel_eager else st.schritte_graph
    # the warm-up in chunks of 10 steps, so that it reports progress
    for k in range(1, warmup, 10):
        schritt(dt, min(10, warmup - k))
        print(f"[bench] warm-up steps 1-{min(k + 9, warmup - 1) + 1} done", file=sys.stderr, flush=True)
    nb, threads = st.bloecke()
    for g in A:
        g.statistik(reset=True)
    t_a = time.perf_counter()
    schritt(dt, steps)
    t_b = time.perf_counter()
    wall_ms = (t_b - t_a) * 1e3
    value = steps / (wall_ms / 1e3)
    stat = [g.statistik() for g in A]
    print("[bench] per system (points, blocks, CG iterations): "
          + str([(int(p), int(b), int(q["cg_iterationen"])) for p, b, q in zip(punkte, nb, stat)]), file=sys.stderr, flush=True)
    iters = sum(q["cg_iterationen"] for q in stat)
    print(f"[bench] batch: {steps} graph steps in {wall_ms:.1f} ms, {iters} CG iterations in all, "
          f"max per system {max(q['cg_iterationen'] for q in stat)}", file=sys.stderr, flush=True)

    # roofline: the same number of eager batched steps right after, PCG
    # launches bracketed by HIP events on the batch's stream
    for g in A:
        g.statistik(reset=True)
    st.schritte(dt, steps)
    _, ms_pcg = st.zeiten()
    stat_e = [g.statistik() for g in A]
    algo = 0
    for (s, _), q in zip(satz, stat_e):
        b = pcg_bytes(s, args.praezision)
        algo += q["cg_iterationen"] * b["iteration"] + q["loesungen"] * b["start"]
    achieved = algo / (ms_pcg * 1e-3) / 1e9
    iters_e = sum(q["cg_iterationen"] for q in stat_e)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_pcg_stapel",
                "ms_pcg_je_schritt": round(ms_pcg / steps, 3),
                "eager_schritte": {"cg_iterationen": iters_e, "bytes_algorithmisch": algo}}
    st.close()

    # the same systems, one after another, plain gc_step per system and step
    seq = None
    if not args.ohne_vergleich:
        B = kontexte()
        for g in B:
            g.schritte(dt, warmup)
        t_c = time.perf_counter()
        for _ in range(steps):
            for g in B:
                g.schritte(dt, 1)
        t_d = time.perf_counter()
        seq = steps / (t_d - t_c)
        print(f"[bench] sequential: {steps} steps of {len(B)} systems in {(t_d - t_c) * 1e3:.1f} ms", file=sys.stderr,
              flush=True)
    out = {
        "metric": "implicit steps/s @40k verts, PCG rtol 1e-6; batched/adjoint vs plain GPU vs CPU",
        "value": round(value, 3), "unit": "batched steps/s (64 garments per step)", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": round(wall_ms / steps, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64" if args.praezision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": "stapel64: 64 garments, points log-uniform in [2k, 60k], latex or woven, stiffness "
                               "x10^U(-0.5,0.5), friction U(0.1,0.5), shared 256^3 SDF, "
                               + ("self-collision on, " if selbst else "self-collision off, ") + "dt 1/120, rtol 1e-6",
                   "timed_steps": f"{warmup + 1}-{warmup + steps}",
                   "punkte": sum(punkte), "systeme": len(satz), "graph": not args.stapel_eager, "praezision": args.praezision,
                   "threads": threads, "bloecke": int(nb.sum()), "autotune_ms_loesung": round(wahl["ms_loesung"], 3),
                   "bloecke_je_system_max": int(nb.max()),
                   "sequentiell_steps_s": round(seq, 3) if seq else None,
                   "stapel_gegen_sequentiell": round(value / seq, 2) if seq else None,
                   "cg_iterationen_je_schritt": round(iters / steps, 1),
                   "kontakte_je_schritt": round(sum(q["kontakte"] for q in stat) / steps, 1),
                   "naehe_je_schritt": round(sum(q["naehe_punkt_dreieck"] + q["naehe_kante_kante"] for q in stat) / steps, 1),
                   "parallelism": "1 GPU"},
        "roofline": roofline,
        "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def adjoint_lauf(args) -> int:
    """The adjoint rollout: a 20k-point drape of the drape's 28 panels (cells
    scaled to 20k points), 60 steps of 1/120 s from the drape settled on the
    body, checkpoint every 10 steps, FP64 (--praezision fp32: FP32 bulk with
    FP64 reductions), rtol 1e-6, self-collision off (the adjoint does not
    differentiate it). Target: the final positions
    with per-panel stretch/shear/bend factors and the friction coefficient
    scaled by U(0.7, 1.3). Loss: mean squared point distance at the final step.
    value = adjoint rollout steps/s: steps / (forward with checkpoints + reverse
    pass, which re-runs each segment and steps the adjoint back)."""
    from garmentcraft_amd import szenen
    from garmentcraft_amd.adjungiert import PanelMaterial, Rollout, verlust, verlust_gradient
    from garmentcraft_amd.geraet import GeraetSzene, ImplizitParameter

    schritte = 60 if args.steps == 300 else args.steps
    alle = 10
    h = 1.0 / 120.0
    t0 = time.time()
    sdf = szenen.koerper_sdf()
    zelle = 352.0 * (20000.0 / 40758.0) ** 0.5
    cfg = szenen.DrapeKonfig(zelle_mm=zelle, jitter_mm=30.0 * zelle / 352.0, versatz_mm=(0.0, -0.5 * zelle))
    s, bericht = szenen.drape(cfg, sdf=sdf)
    print(f"[bench] adjoint scene {bericht} built in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    mat = PanelMaterial(s)
    rng = np.random.default_rng(20261205)
    th0 = mat.theta0()
    th_ziel = th0 * rng.uniform(0.7, 1.3, mat.n_param)

    # the rollout starts from the drape settled on the body (the FP64 form,
    # args.vorlauf steps of the flat drop): through the drop's impact the
    # rollout is chaotic (FP32 and FP64 forms part by centimetres) and no
    # gradient describes it
    x_s, v_s, dv_s = s.x, s.v, None
    if args.vorlauf > 0:
        p = ImplizitParameter(schritte_je_bild=1, cg_max=10000, cg_toleranz=1e-8, naht_k=500.0, selbstkollision=False)
        # a fixed launch shape, so the settled state does not depend on the
        # default shape (the FP64 reductions group their block partials by it)
        gv = GeraetSzene(s, p, praezision="fp64", bloecke=64, lds_kb=128)
        gv.schritte(h, args.vorlauf)
        x_s, v_s = gv.lage()
        dv_s = None
        gv.close()
        print(f"[bench] settled {args.vorlauf} steps", file=sys.stderr, flush=True)

    def kontext(praezision, tol):
        p = ImplizitParameter(schritte_je_bild=1, cg_max=10000, cg_toleranz=tol, naht_k=500.0, selbstkollision=False)
        g = GeraetSzene(s, p, praezision=praezision)
        return g, Rollout(g, s, mat, h, schritte, alle=alle, x0=x_s, v0=v_s, dv_vorher=dv_s)

    g32, r32 = kontext(args.praezision, 1e-6)
    ziel = r32.vorwaerts(th_ziel)
    print(f"[bench] target rollout done", file=sys.stderr, flush=True)
    r32.vorwaerts(th0)  # warm-up of both passes
    r32.rueckwaerts(verlust_gradient(r32.g.lage()[0], ziel))
    g32.statistik(reset=True)
    t_a = time.perf_counter()
    x = r32.vorwaerts(th0)
    t_b = time.perf_counter()
    d32 = r32.rueckwaerts(verlust_gradient(x, ziel))
    t_c = time.perf_counter()
    st = g32.statistik()
    ast = r32.statistik
    L0 = verlust(x, ziel)
    # plain steps of the same rollout, no checkpoints
    g32.laden(-1)
    t_p = time.perf_counter()
    g32.schritte(h, schritte)
    t_q = time.perf_counter()
    plain = schritte / (t_q - t_p)
    print(f"[bench] forward {1e3 * (t_b - t_a):.1f} ms, reverse {1e3 * (t_c - t_b):.1f} ms, plain {1e3 * (t_q - t_p):.1f} ms; "
          f"loss {L0:.6e}; adjoint solves {ast}", file=sys.stderr, flush=True)

    # one step's vector-Jacobian product from the same state in both forms, at
    # rtol 1e-10 so that both solves reach the same solution and take the same
    # discrete decisions (what the precision costs, with the tolerance out of
    # the way): a random seed dL/dx' back to dL/dx, dL/dv and the parameter
    # gradients
    vjp = {}
    seed = np.random.default_rng(5).standard_normal((s.n, 3))
    for pr in ("fp32", "fp64"):
        gq, rq = kontext(pr, 1e-10)
        rq._setzen(th0)
        gq.laden(-1)
        gq.adjoint_gradient(s, reset=True)
        gq.adjoint_setzen(seed, None)
        gq.adjoint_schritt(h)
        gx_, gv_ = gq.adjoint_lesen()
        vjp[pr] = (gx_, gv_, mat.zurueck(gq.adjoint_gradient(s, reset=True)), gq.lage()[0])
        gq.close()
    def _rel(a, b):
        return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    vjp_rel = {"dL_dx": _rel(vjp["fp32"][0], vjp["fp64"][0]), "dL_dv": _rel(vjp["fp32"][1], vjp["fp64"][1]),
               "dL_dtheta": _rel(vjp["fp32"][2], vjp["fp64"][2]),
               "x_nach_schritt_m": float(np.max(np.abs(vjp["fp32"][3] - vjp["fp64"][3])))}
    print(f"[bench] one-step VJP fp32 vs fp64: {vjp_rel}", file=sys.stderr, flush=True)

    # the other precision's adjoint on the same rollout (each the derivative of
    # its own forward's branch)
    andere = "fp32" if args.praezision == "fp64" else "fp64"
    g64, r64 = kontext(andere, 1e-6)
    x64 = r64.vorwaerts(th0)
    d64 = r64.rueckwaerts(verlust_gradient(x64, ziel))
    d32_, d64_ = (d32, d64) if args.praezision == "fp32" else (d64, d32)
    rel32 = float(np.linalg.norm(d32_ - d64_) / np.linalg.norm(d64_))
    print(f"[bench] fp32 vs fp64 adjoint gradient: relative {rel32:.3e}; positions {np.max(np.abs(x - x64)):.3e} m",
          file=sys.stderr, flush=True)
    fd = None
    if args.fd > 0:
        # central differences (relative step 1e-4) of the FP64 forward at
        # rtol 1e-10 against the FP64 adjoint at the same tolerance
        g_fd, r_fd = kontext("fp64", 1e-10)
        xf = r_fd.vorwaerts(th0)
        dfd_adj = r_fd.rueckwaerts(verlust_gradient(xf, ziel))
        wahl = list(np.argsort(-np.abs(dfd_adj[:-1]))[: max(0, args.fd - 1)]) + [mat.n_param - 1]
        zeilen = []
        for k in wahl:
            hk = args.fd_schritt * abs(th0[k])
            tp, tm = th0.copy(), th0.copy()
            tp[k] += hk
            tm[k] -= hk
            fdk = (verlust(r_fd.vorwaerts(tp), ziel) - verlust(r_fd.vorwaerts(tm), ziel)) / (2 * hk)
            zeilen.append({"param": int(k), "fd": fdk, "adjoint_fp64": float(dfd_adj[k]), "adjoint_fp32": float(d32_[k]),
                           "rel": abs(dfd_adj[k] - fdk) / max(abs(fdk), 1e-300)})
            print(f"[bench] fd theta[{k}]: {fdk:.6e} adjoint {dfd_adj[k]:.6e} (fp32 at 1e-6: {d32_[k]:.6e})",
                  file=sys.stderr, flush=True)
        v_fd = np.array([z["fd"] for z in zeilen])
        v_ad = np.array([z["adjoint_fp64"] for z in zeilen])
        fd = {"parameter": zeilen, "rel_norm": float(np.linalg.norm(v_ad - v_fd) / np.linalg.norm(v_fd)),
              "rtol": 1e-10, "schritt": f"{args.fd_schritt:g} relative"}
    value = schritte / (t_c - t_a)
    out = {
        "metric": "implicit steps/s @40k verts, PCG rtol 1e-6; batched/adjoint vs plain GPU vs CPU",
        "value": round(value, 3), "unit": "adjoint rollout steps/s (forward with checkpoints + reverse)",
        "n_gpus": 1, "steps": schritte, "warmup": 1, "ms_per_step": round(1e3 * (t_c - t_a) / schritte, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if args.praezision == "fp32" else "f64", "data": "synthetic",
        "config": {"workload": f"adjoint20k: 28-panel woven drape scaled to 20k points, settled {args.vorlauf} steps, "
                               f"then {schritte} steps of 1/120 s, checkpoint every 10, rtol 1e-6, self-collision off; "
                               "gradient of the final mean squared point distance w.r.t. 28x3 per-panel stiffness "
                               "factors and friction",
                   "punkte": s.n, "parameter": mat.n_param, "verlust": L0,
                   "plain_steps_s": round(plain, 3), "adjoint_gegen_plain": round(value / plain, 3),
                   "ms_vorwaerts": round(1e3 * (t_b - t_a), 1), "ms_rueckwaerts": round(1e3 * (t_c - t_b), 1),
                   "cg_iterationen_je_schritt": round(st["cg_iterationen"] / schritte, 1),
                   "adjoint_cg_iterationen_je_schritt": round(ast["cg_iterationen"] / schritte, 1),
                   "adjoint_abweichungen": ast["abweichungen"], "adjoint_eintritte": ast["eintritte"],
                   "fp32_gegen_fp64_gradient": rel32, "ein_schritt_fp32_gegen_fp64": vjp_rel, "fd": fd,
                   "parallelism": "1 GPU"},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


def gross_lauf(args) -> int:
    """The 250k-point drape: the 28 panels meshed at 4 mm (~250k points,
    ~490k triangles), body SDF 384^3 at 2 mm, self-collision on, friction 0.3,
    100 steps of 1/240 s at rtol 1e-7, stepped in lockstep in the FP32-bulk
    form and the all-FP64 form; per step the FP32 form's position error
    against the FP64 form (which equals the FP64 oracle to 1e-9 m on the
    drape tests)."""
    from garmentcraft_amd import szenen
    from garmentcraft_amd.geraet import GeraetSzene, ImplizitParameter

    schritte = 100 if args.steps == 300 else args.steps
    h = 1.0 / 240.0
    t0 = time.time()
    sdf = szenen.koerper_sdf(aufloesung=384, spacing=0.002)
    s, bericht = szenen.drape(szenen.DrapeKonfig(kante_mm=4.0), sdf=sdf)
    print(f"[bench] scene {bericht} built in {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    selbst = not args.ohne_selbstkollision
    p = ImplizitParameter(schritte_je_bild=1, cg_max=20000, cg_toleranz=1e-7, naht_k=500.0, selbstkollision=selbst)
    g = {pr: GeraetSzene(s, p, praezision=pr, bloecke=args.bloecke, lds_kb=args.lds_kb) for pr in ("fp32", "fp64")}
    ms = {"fp32": 0.0, "fp64": 0.0}
    ms_pcg = {"fp32": 0.0, "fp64": 0.0}
    fehler_max, fehler_mittel = [], []
    iters_alt = {pr: 0 for pr in g}
    for k in range(schritte):
        zeile = []
        for pr in ("fp32", "fp64"):
            t_a = time.perf_counter()
            g[pr].schritte(h, 1)
            dt_ms = 1e3 * (time.perf_counter() - t_a)
            ms[pr] += dt_ms
            ms_pcg[pr] += g[pr].zeiten()[1]
            it = g[pr].statistik()["cg_iterationen"]
            zeile.append(f"{pr} {dt_ms:.0f} ms {it - iters_alt[pr]} it")
            iters_alt[pr] = it
        d = np.linalg.norm(g["fp32"].lage()[0] - g["fp64"].lage()[0], axis=1)
        fehler_max.append(float(d.max()))
        fehler_mittel.append(float(d.mean()))
        print(f"[bench] step {k + 1}: {', '.join(zeile)}; position error max {d.max():.3e} m mean {d.mean():.3e} m",
              file=sys.stderr, flush=True)
    st = {pr: g[pr].statistik() for pr in g}
    roof = {}
    for pr in g:
        b = pcg_bytes(s, pr)
        algo = st[pr]["cg_iterationen"] * b["iteration"] + st[pr]["loesungen"] * b["start"]
        ach = algo / (ms_pcg[pr] * 1e-3) / 1e9
        roof[pr] = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": None, "kernel": "k_pcg",
                    "bytes_per_iteration": b["iteration"],
                    "us_per_iteration": round(ms_pcg[pr] * 1e3 / max(st[pr]["cg_iterationen"], 1), 3)}
    value = schritte / (ms["fp32"] * 1e-3)
    out = {
        "metric": "implicit steps/s @40k verts, PCG rtol 1e-6; batched/adjoint vs plain GPU vs CPU",
        "value": round(value, 3), "unit": "steps/s (FP32 bulk, 250k points)", "n_gpus": 1, "steps": schritte,
        "warmup": 0, "ms_per_step": round(ms["fp32"] / schritte, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "gross250k: 28 woven panels at 4 mm edges, SDF 384^3 at 2 mm, friction 0.3, "
                               + ("self-collision on, " if selbst else "self-collision off, ") + "dt 1/240, rtol 1e-7",
                   "punkte": s.n, "dreiecke": int(s.dreiecke.shape[0]),
                   "fp64_steps_s": round(schritte / (ms["fp64"] * 1e-3), 3),
                   "fp32_gegen_fp64_schneller": round(ms["fp64"] / ms["fp32"], 3),
                   "positionsfehler_max_m": max(fehler_max), "positionsfehler_ende_max_m": fehler_max[-1],
                   "positionsfehler_ende_mittel_m": fehler_mittel[-1],
                   "positionsfehler_je_schritt_max_m": [float(f"{e:.3e}") for e in fehler_max],
                   "cg_iterationen_je_schritt": {pr: round(st[pr]["cg_iterationen"] / schritte, 1) for pr in g},
                   "nicht_konvergiert": {pr: st[pr]["nicht_konvergiert"] for pr in g},
                   "rundungsgrenze_fp32": st["fp32"]["cg_rundungsgrenze"], "parallelism": "1 GPU"},
        "roofline": roof["fp32"], "roofline_fp64": roof["fp64"], "cpu_baseline": None,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
