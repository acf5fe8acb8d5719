"""HBM traffic of a kernel from two rocprofv3 --pmc runs (FETCH_SIZE, then
WRITE_SIZE), scaled by what the same counters read for a plain copy of known
size (garmentcraft_amd/kopie_kalibrierung, 512 MiB read and written per
dispatch).

    python -m garmentcraft_amd.zaehler --fetch F.db --write W.db \\
        --kopie-fetch KF.db --kopie-write KW.db --kernel k_pcg --letzte 80 --iterationen N

sums the counters over the last `--letzte` dispatches of the kernel named
`--kernel` (any template arguments, or exactly those given, as in
k_pcg<double>) and divides by `--iterationen` (the PCG iterations those
dispatches ran, from the bench line of the same run). The databases are
rocprofv3's rocpd SQLite output (view counters_collection).
"""
from __future__ import annotations

import argparse
import json
import sqlite3

KOPIE_KIB = 512 * 1024  # bytes a copy dispatch reads, and writes, in KiB


def werte(db: str, zaehler: str, kernel: str) -> list[float]:
    """Per-dispatch values of one counter for the kernel, in dispatch order."""
    c = sqlite3.connect(db)
    try:
        zeilen = c.execute(
            "select dispatch_id, kernel_name, value from counters_collection where counter_name = ? "
            "order by dispatch_id", (zaehler,)).fetchall()
    finally:
        c.close()
    def passt(name: str) -> bool:
        voll = name.split("(")[0].split(" ")[-1]  # with the template arguments, e.g. k_pcg<double>
        return voll == kernel if "<" in kernel else voll.split("<")[0] == kernel
    return [float(v) for _, name, v in zeilen if passt(name)]


def faktor(db: str, zaehler: str) -> float:
    """Bytes per counter KiB: the copy's known KiB over what the counter read
    (median over its dispatches)."""
    v = sorted(werte(db, zaehler, "k_kopie"))
    if not v:
        raise SystemExit(f"{db}: no k_kopie dispatch with {zaehler}")
    return KOPIE_KIB / v[len(v) // 2]


def main(argv=None) -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--fetch", required=True)
    ap.add_argument("--write", required=True)
    ap.add_argument("--kopie-fetch", required=True)
    ap.add_argument("--kopie-write", required=True)
    ap.add_argument("--kernel", default="k_pcg")
    ap.add_argument("--letzte", type=int, required=True)
    ap.add_argument("--iterationen", type=int, required=True)
    ap.add_argument("--algorithmisch", type=float, default=None, help="algorithmic bytes per iteration")
    a = ap.parse_args(argv)
    ff, fw = faktor(a.kopie_fetch, "FETCH_SIZE"), faktor(a.kopie_write, "WRITE_SIZE")
    f = werte(a.fetch, "FETCH_SIZE", a.kernel)[-a.letzte:]
    w = werte(a.write, "WRITE_SIZE", a.kernel)[-a.letzte:]
    if len(f) < a.letzte or len(w) < a.letzte:
        raise SystemExit(f"fewer than {a.letzte} {a.kernel} dispatches")
    gelesen = sum(f) * 1024 * ff / a.iterationen
    geschrieben = sum(w) * 1024 * fw / a.iterationen
    out = {"kernel": a.kernel, "dispatches": a.letzte, "iterations": a.iterationen,
           "fetch_factor": round(ff, 4), "write_factor": round(fw, 4),
           "fetch_kib_per_dispatch_max": round(max(f), 1), "write_kib_per_dispatch_max": round(max(w), 1),
           "read_B_per_iteration": round(gelesen), "written_B_per_iteration": round(geschrieben),
           "traffic_B_per_iteration": round(gelesen + geschrieben)}
    if a.algorithmisch:
        out["traffic_over_algorithmic"] = round((gelesen + geschrieben) / a.algorithmisch, 4)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
