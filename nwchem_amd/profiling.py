"""Profiling helpers (run under rocprofv3, and summaries of its output).

    python3 -m nwchem_amd.profiling copy [--mib 1024] [--repeats 1]

launches the library's calibration copy kernel (stream_copy): 8n bytes read
and 8n bytes written per launch, the yardstick that turns the FETCH_SIZE /
WRITE_SIZE counter readings of the (T) kernel into HBM bytes.

    python3 -m nwchem_amd.profiling stats DB --command CMD --out FILE.txt
    python3 -m nwchem_amd.profiling pmc --kernel-db DB... [--copy-db DB...]
                                        --command CMD --out FILE.txt

summarise a `rocprofv3 --kernel-trace --stats` run (per-kernel calls, total
and average duration) and one or more `rocprofv3 --pmc` runs (counter sums
per dispatch of the fused kernel; with the copy kernel's runs, the HBM bytes
per launch, written beside FILE.txt as FILE.json for bench.py).  DB is the
`run_results.db` that rocprofv3 writes with `-o run`.
"""
from __future__ import annotations

import argparse
import json
import os
import sqlite3
import sys

KERNEL_LIKE = "%fused_triples%"
COPY_LIKE = "%stream_copy%"
N_SIMD = 256 * 4  # MI355X: 256 CUs x 4 SIMDs


def copy_probe(mib: int = 1024, repeats: int = 1) -> int:
    from . import _lib

    n = mib * 1024 * 1024 // 8
    L = _lib.lib()
    L.tce_t_copy_probe.argtypes = [_lib.C.c_long, _lib.C.c_int]
    rc = L.tce_t_copy_probe(n, repeats)
    _lib.check(rc, "tce_t_copy_probe")
    return n * 8


def _rows(db: str, sql: str, args=()):
    con = sqlite3.connect(db)
    try:
        return con.execute(sql, args).fetchall()
    finally:
        con.close()


def kernel_stats(db: str):
    """[(name, calls, total_us, avg_us, percent)] from rocprofv3's top_kernels view."""
    return _rows(db, "select name, total_calls, total_duration, average, percentage from top_kernels")


def pmc_sums(db: str, like: str):
    """{counter: (sum over dispatches, dispatches, mean duration ns)} for kernels matching like."""
    out = {}
    for cn, v, nd, du in _rows(
            db, "select counter_name, sum(counter_value), count(distinct dispatch_id), avg(duration) "
                "from pmc_events where name like ? group by counter_name", (like,)):
        out[cn] = (float(v), int(nd), float(du))
    return out


def summarize_stats(db: str, command: str) -> str:
    L = ["## rocprofv3 --kernel-trace --stats", f"command: {command}", "",
         f"{'kernel':64s} {'calls':>6s} {'total_ms':>12s} {'avg_ms':>12s} {'pct':>7s}"]
    for name, calls, tot, avg, pct in kernel_stats(db):
        L.append(f"{name[:64]:64s} {calls:6d} {tot / 1e3:12.3f} {avg / 1e3:12.3f} {pct:7.3f}")
    return "\n".join(L) + "\n"


def summarize_pmc(kernel_dbs, copy_dbs, command: str):
    """Text summary and a dict of derived figures (per dispatch of the fused kernel)."""
    rows, copy = {}, {}
    dur, nd = None, None
    for db in kernel_dbs:
        for cn, (v, n, du) in pmc_sums(db, KERNEL_LIKE).items():
            rows[cn] = v / n
            dur, nd = du, n
    for db in copy_dbs:
        for cn, (v, n, du) in pmc_sums(db, COPY_LIKE).items():
            copy[cn] = v / n
    L = ["## rocprofv3 --pmc, fused (T) kernel", f"command (one run per counter group): {command}"]
    if copy_dbs:
        L.append("calibration: rocprofv3 --pmc <same group> -- python3 -m nwchem_amd.profiling copy --mib 1024")
    L.append(f"dispatches per run: {nd}; mean kernel duration under PMC: {dur / 1e6:.3f} ms")
    L.append("")
    L.append(f"{'counter (per dispatch)':34s} {'fused kernel':>16s} {'stream_copy 1GiB':>18s}")
    for cn in sorted(set(rows) | set(copy)):
        L.append(f"{cn:34s} {rows.get(cn, float('nan')):16.6g} {copy.get(cn, float('nan')):18.6g}")
    L.append("")
    derived = {"kernel_ms_under_pmc": dur / 1e6}
    nbytes = 1024 ** 3
    if "FETCH_SIZE" in rows and "FETCH_SIZE" in copy and "WRITE_SIZE" in rows and "WRITE_SIZE" in copy:
        ff = nbytes / (copy["FETCH_SIZE"] * 1024)
        wf = nbytes / (copy["WRITE_SIZE"] * 1024)
        fb = rows["FETCH_SIZE"] * 1024 * ff
        wb = rows["WRITE_SIZE"] * 1024 * wf
        L.append(f"calibration: stream_copy reads and writes {nbytes} B; FETCH_SIZE reads "
                 f"{copy['FETCH_SIZE'] * 1024:.0f} B -> read factor {ff:.4f}; WRITE_SIZE reads "
                 f"{copy['WRITE_SIZE'] * 1024:.0f} B -> write factor {wf:.4f}")
        L.append(f"HBM bytes per launch: read {fb:.4e} B, write {wb:.4e} B, total {fb + wb:.4e} B "
                 f"({(fb + wb) / (dur * 1e-9) / 1e12:.3f} TB/s over the launch)")
        derived.update(hbm_bytes_per_launch=fb + wb, read_bytes=fb, write_bytes=wb,
                       fetch_factor=ff, write_factor=wf)
    if "TCC_HIT_sum" in rows and "TCC_MISS_sum" in rows:
        h, m = rows["TCC_HIT_sum"], rows["TCC_MISS_sum"]
        L.append(f"L2 hit rate {h / (h + m):.3f}")
        derived["l2_hit_rate"] = h / (h + m)
    if "SQ_VALU_MFMA_BUSY_CYCLES" in rows and dur:
        # busy cycles summed over SIMDs; the shader clock is taken from
        # SQ_BUSY_CYCLES when present (summed over 32 shader engines)
        f = rows["SQ_VALU_MFMA_BUSY_CYCLES"] / N_SIMD
        if "SQ_BUSY_CYCLES" in rows:
            busy = rows["SQ_BUSY_CYCLES"] / 32
            L.append(f"MFMA busy fraction {f / busy:.3f} (SQ_VALU_MFMA_BUSY_CYCLES / 1024 SIMDs / "
                     f"per-engine SQ_BUSY_CYCLES)")
            derived["mfma_busy"] = f / busy
    if "SQ_WAVE_CYCLES" in rows:
        wc = rows["SQ_WAVE_CYCLES"]
        for cn in ("SQ_WAIT_INST_ANY", "SQ_WAIT_ANY", "SQ_WAIT_INST_LDS"):
            if cn in rows:
                L.append(f"{cn} / SQ_WAVE_CYCLES = {rows[cn] / wc:.3f}")
    if "SQ_LDS_BANK_CONFLICT" in rows and "SQ_BUSY_CYCLES" in rows:
        cu_cycles = rows["SQ_BUSY_CYCLES"] / 32 * 256
        L.append(f"LDS bank-conflict cycles / CU cycles = {rows['SQ_LDS_BANK_CONFLICT'] / cu_cycles:.3f}")
    return "\n".join(L) + "\n", derived


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="what", required=True)
    c = sub.add_parser("copy")
    c.add_argument("--mib", type=int, default=1024)
    c.add_argument("--repeats", type=int, default=1)
    s = sub.add_parser("stats")
    s.add_argument("db")
    s.add_argument("--command", required=True)
    s.add_argument("--out", required=True)
    p = sub.add_parser("pmc")
    p.add_argument("--kernel-db", nargs="+", required=True)
    p.add_argument("--copy-db", nargs="*", default=[])
    p.add_argument("--command", required=True)
    p.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    if a.what == "copy":
        nbytes = copy_probe(a.mib, a.repeats)
        print(f"stream_copy: {a.repeats} launch(es), {nbytes} bytes read and written per launch")
    elif a.what == "stats":
        text = summarize_stats(a.db, a.command)
        with open(a.out, "w") as f:
            f.write(text)
        print(text)
    else:
        text, derived = summarize_pmc(a.kernel_db, a.copy_db, a.command)
        with open(a.out, "w") as f:
            f.write(text)
        with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
            json.dump(derived, f, indent=1)
        print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
