"""Profiling helpers for bench.py (run under rocprofv3, and summaries of its
output).  Measurement tooling, not part of the library.

    python3 profiling.py copy [--mib 1024] [--repeats 1]

launches the calibration copy kernel (stream_copy, in the separate
measurement library nwchem_amd/libtce_t_probe.so): 8n bytes read and 8n
bytes written per launch, the yardstick that turns the FETCH_SIZE /
WRITE_SIZE counter readings of the (T) kernel into HBM bytes.

    python3 profiling.py stats DB --command CMD --out FILE.txt
    python3 profiling.py pmc --kernel-db DB... [--copy-db DB...]
                             [--kernel-like PATTERN] --command CMD --out FILE.txt

summarise a `rocprofv3 --kernel-trace --stats` run (per-kernel calls, total
and average duration) and one or more `rocprofv3 --pmc` runs (counter sums
per dispatch of the kernels whose name matches PATTERN, one column per kernel
name; with the copy kernel's runs, the HBM bytes per launch, written beside
FILE.txt as FILE.json for bench.py).  DB is the `run_results.db` that
rocprofv3 writes with `-o run`.
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


PROBE_LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nwchem_amd", "libtce_t_probe.so")


def copy_probe(mib: int = 1024, repeats: int = 1) -> int:
    import ctypes as C

    n = mib * 1024 * 1024 // 8
    L = C.CDLL(PROBE_LIB)
    L.probe_stream_copy.argtypes = [C.c_long, C.c_int]
    L.probe_stream_copy.restype = C.c_int
    rc = L.probe_stream_copy(n, repeats)
    if rc != 0:
        raise RuntimeError(f"probe_stream_copy failed ({rc})")
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
    """{kernel name: {counter: (sum over dispatches, dispatches, mean duration ns)}}
    for the kernels whose name matches like."""
    out = {}
    for kn, cn, v, nd, du in _rows(
            db, "select name, counter_name, sum(counter_value), count(distinct dispatch_id), "
                "avg(duration) from pmc_events where name like ? group by name, counter_name", (like,)):
        out.setdefault(kn, {})[cn] = (float(v), int(nd), float(du))
    return out


def short_name(kernel: str) -> str:
    """fused_triples_mixed<2, false> from the demangled signature."""
    k = kernel.replace("(anonymous namespace)::", "").split("(")[0]
    return k.split("::")[-1].replace("void ", "")


def summarize_stats(db: str, command: str) -> str:
    L = ["## rocprofv3 --kernel-trace --stats", f"command: {command}", "",
         f"{'kernel':64s} {'calls':>6s} {'total_ms':>12s} {'avg_ms':>12s} {'pct':>7s}"]
    for name, calls, tot, avg, pct in kernel_stats(db):
        L.append(f"{name[:64]:64s} {calls:6d} {tot / 1e3:12.3f} {avg / 1e3:12.3f} {pct:7.3f}")
    return "\n".join(L) + "\n"


def derive(rows: dict, dur_ns: float, copy: dict, factors) -> tuple[list, dict]:
    """Figures derived from one kernel's counters per dispatch."""
    L, d = [], {"kernel_ms_under_pmc": dur_ns / 1e6}
    nbytes = 1024 ** 3
    ff = wf = None
    if "FETCH_SIZE" in copy and "WRITE_SIZE" in copy:
        ff = nbytes / (copy["FETCH_SIZE"] * 1024)
        wf = nbytes / (copy["WRITE_SIZE"] * 1024)
        L.append(f"calibration: stream_copy reads and writes {nbytes} B; FETCH_SIZE reads "
                 f"{copy['FETCH_SIZE'] * 1024:.0f} B -> read factor {ff:.4f}; WRITE_SIZE reads "
                 f"{copy['WRITE_SIZE'] * 1024:.0f} B -> write factor {wf:.4f}")
    elif factors:
        ff, wf = factors
        L.append(f"calibration factors given: read {ff}, write {wf} (stream_copy of 1 GiB, "
                 f"`python3 profiling.py copy` under the same counters)")
    if ff and "FETCH_SIZE" in rows and "WRITE_SIZE" in rows:
        fb = rows["FETCH_SIZE"] * 1024 * ff
        wb = rows["WRITE_SIZE"] * 1024 * wf
        L.append(f"HBM bytes per launch: read {fb:.4e} B, write {wb:.4e} B, total {fb + wb:.4e} B "
                 f"({(fb + wb) / (dur_ns * 1e-9) / 1e12:.3f} TB/s over the launch)")
        d.update(hbm_bytes_per_launch=fb + wb, read_bytes=fb, write_bytes=wb,
                 fetch_factor=ff, write_factor=wf)
    elif ff and "FETCH_SIZE" in rows:
        fb = rows["FETCH_SIZE"] * 1024 * ff
        L.append(f"HBM bytes read per launch: {fb:.4e} B ({fb / (dur_ns * 1e-9) / 1e12:.3f} TB/s)")
        d.update(read_bytes=fb, fetch_factor=ff)
    elif wf and "WRITE_SIZE" in rows:
        wb = rows["WRITE_SIZE"] * 1024 * wf
        L.append(f"HBM bytes written per launch: {wb:.4e} B")
        d.update(write_bytes=wb, write_factor=wf)
    if "TCC_HIT_sum" in rows and "TCC_MISS_sum" in rows:
        h, m = rows["TCC_HIT_sum"], rows["TCC_MISS_sum"]
        L.append(f"L2 hit rate {h / (h + m):.3f}")
        d["l2_hit_rate"] = h / (h + m)
    # MFMA work: each MOPS count is 512 flops (counter_defs.yaml: *_FLOPS)
    for prec in ("F64", "F32", "BF16", "F16"):
        cn = "SQ_INSTS_VALU_MFMA_MOPS_" + prec
        if cn in rows and rows[cn] > 0:
            fl = rows[cn] * 512
            L.append(f"MFMA {prec} flops per launch {fl:.4e} -> {fl / (dur_ns * 1e-9) / 1e12:.2f} "
                     f"TFLOP/s over the launch")
            d["mfma_flops_" + prec.lower()] = fl
    if "SQ_VALU_MFMA_BUSY_CYCLES" in rows:
        # busy cycles summed over SIMDs; the shader clock from GRBM_GUI_ACTIVE
        # (counter_defs.yaml MFMA_UTIL) or SQ_BUSY_CYCLES summed over 32 engines
        f = rows["SQ_VALU_MFMA_BUSY_CYCLES"] / N_SIMD
        if "SQ_BUSY_CYCLES" in rows:
            busy = rows["SQ_BUSY_CYCLES"] / 32
            L.append(f"MFMA busy fraction {f / busy:.3f} (SQ_VALU_MFMA_BUSY_CYCLES / 1024 SIMDs / "
                     f"per-engine SQ_BUSY_CYCLES)")
            d["mfma_busy"] = f / busy
    if "SQ_WAVE_CYCLES" in rows:
        wc = rows["SQ_WAVE_CYCLES"]
        for cn in ("SQ_WAIT_INST_ANY", "SQ_WAIT_ANY", "SQ_WAIT_INST_LDS", "SQ_ACTIVE_INST_LDS",
                   "SQ_ACTIVE_INST_VALU"):
            if cn in rows:
                L.append(f"{cn} / SQ_WAVE_CYCLES = {rows[cn] / wc:.3f}")
                d[cn.lower() + "_frac"] = rows[cn] / wc
    if "SQ_LDS_BANK_CONFLICT" in rows and "SQ_BUSY_CYCLES" in rows:
        cu_cycles = rows["SQ_BUSY_CYCLES"] / 32 * 256
        L.append(f"LDS bank-conflict cycles / CU cycles = {rows['SQ_LDS_BANK_CONFLICT'] / cu_cycles:.3f}")
        d["lds_bank_conflict_frac"] = rows["SQ_LDS_BANK_CONFLICT"] / cu_cycles
    if "SQ_LDS_BANK_CONFLICT" in rows and "SQ_LDS_IDX_ACTIVE" in rows:
        c, a = rows["SQ_LDS_BANK_CONFLICT"], rows["SQ_LDS_IDX_ACTIVE"]
        if a > c:
            L.append(f"LDS bank-conflict cycles / conflict-free LDS cycles = {c / (a - c):.3f}")
    return L, d


def summarize_pmc(kernel_dbs, copy_dbs, command: str, like: str = KERNEL_LIKE, factors=None):
    """Text summary and {kernel: derived figures} (per dispatch of each kernel
    whose name matches like; one column per kernel)."""
    rows, dur, nd = {}, {}, {}
    copy = {}
    for db in kernel_dbs:
        for kn, cs in pmc_sums(db, like).items():
            k = short_name(kn)
            for cn, (v, n, du) in cs.items():
                rows.setdefault(k, {})[cn] = v / n
                dur[k], nd[k] = du, n
    for db in copy_dbs:
        for kn, cs in pmc_sums(db, COPY_LIKE).items():
            for cn, (v, n, du) in cs.items():
                copy[cn] = v / n
    names = sorted(rows)
    L = ["## rocprofv3 --pmc, per dispatch of each kernel",
         f"command (one run per counter group): {command}"]
    if copy_dbs:
        L.append("calibration: rocprofv3 --pmc <same group> -- python3 profiling.py copy --mib 1024")
    for k in names:
        L.append(f"{k}: {nd[k]} dispatch(es) per run; mean duration under PMC {dur[k] / 1e6:.3f} ms")
    L.append("")
    cols = names + (["stream_copy 1GiB"] if copy else [])
    L.append(f"{'counter':34s} " + " ".join(f"{c[-30:]:>30s}" for c in cols))
    for cn in sorted(set().union(*[set(r) for r in rows.values()], set(copy))):
        vals = [rows[k].get(cn, float("nan")) for k in names] + ([copy.get(cn, float("nan"))] if copy else [])
        L.append(f"{cn:34s} " + " ".join(f"{v:30.6g}" for v in vals))
    derived = {}
    for k in names:
        L.append("")
        L.append(f"-- {k}")
        dl, dd = derive(rows[k], dur[k], copy, factors)
        L += dl
        derived[k] = dd
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
    p.add_argument("--kernel-like", default=KERNEL_LIKE)
    p.add_argument("--factors", type=float, nargs=2, default=None,
                   help="FETCH_SIZE and WRITE_SIZE calibration factors when no copy run is given")
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
        text, derived = summarize_pmc(a.kernel_db, a.copy_db, a.command, a.kernel_like, a.factors)
        with open(a.out, "w") as f:
            f.write(text)
        with open(os.path.splitext(a.out)[0] + ".json", "w") as f:
            json.dump(derived, f, indent=1)
        print(text)
    return 0


if __name__ == "__main__":
    sys.exit(main())
