#!/usr/bin/env python3
"""Benchmark of the FP64 (T) job on one MI355X.

A step is one complete (T) job: every tile sextuple of the task loop of
ccsd_t_gpu.F contracted (27 sd_t variants) and reduced to E[T], E(T).  Inputs
(the canonical T1/T2/V2 blocks) are uploaded once before the timed region and
stay resident in HBM.

Default workload: no=40 occupied (tiles of 20), nv=320 virtual (tiles of 40),
480 tasks, FP64, seeded synthetic inputs.

Output: one JSON line.  `value` is algorithmic TFLOP/s of the FP64 job:
F = sum over the kernel tasks run of |t3 tile| * 18 * (no + nv + 1) (the
reference's per-task contraction count, ccsd_t_kernels_omp.F; energy
excluded) over the wall time of the job.  The kernel tasks are the box-sized
sub-tiles of the reference tiles (the library's default), which compute the
same energies without the symmetric duplicates a reference task computes
inside equal tiles, so F is about a third of the reference tiling's count
F_ref at this workload; both are reported, and `reference_tiling` times the
same job with the reference's tiles as kernel tasks (F_ref of work).  The
same line reports the reduced-precision modes (`--modes`): the d1/d2
contractions in FP32 and in 3-term split-BF16, each timed the same way on the
same resident inputs, with energies/s, the speedup over FP64 and |dE| in
micro-Eh against the FP64 result of the run; and the Lambda-CCSD(T)
energies of lambda_ccsd_t.F over the same tiles (`--lambda`), a second
seeded amplitude set taken as the Lambda amplitudes; and CR-CCSD(T)
(`--cr`), with the bare integrals as cr_ccsd_t_N's intermediates so that its
moment sums equal E[T] and E(T) of the run.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "(T) TFLOP/s & energies/s, FP64 vs split-BF16 at ≤1 µEh; serving p50/p99 latency"
PEAK_FP64_TFLOPS = 78.6
# dense MFMA peaks of the reduced-precision modes (MI355X): FP32, and BF16
# for both split modes; products per FP64-equivalent product of each mode
PEAK_MODE_TFLOPS = {"fp32": 157.3, "bf16x3": 2500.0, "bf16x6": 2500.0}
MODE_PRODUCTS = {"fp32": 1, "bf16x3": 3, "bf16x6": 6}


def job_work(tm, precision: int = 0, tiling: int = 2) -> float:
    """Flops a run of this job performs in `tiling` (the kernel tasks the
    library builds for it, tce_t_job_work; 2 = orbital, what the server
    uses); no upload needed."""
    from nwchem_amd import TriplesJob

    with TriplesJob(tm, tiling=tiling) as j:
        return j.work(precision)[0]


def f_ref(tm) -> float:
    """Reference tile-dispatch flops of one job (contractions only)."""
    r = tm.range
    tot = 0
    for (p4, p5, p6, h1, h2, h3) in tm.tasks():
        tot += r[p4 - 1] * r[p5 - 1] * r[p6 - 1] * r[h1 - 1] * r[h2 - 1] * r[h3 - 1]
    return float(tot) * 18.0 * (tm.no + tm.nv + 1)


def traffic_from_profiles(workload: str, kernel: str):
    """HBM bytes per launch of the fused kernel from the committed rocprofv3
    PMC summary (FETCH_SIZE + WRITE_SIZE, scaled by the copy-kernel factor),
    or None when no summary exists for this kernel and workload."""
    # {kernel: figures} per kernel of the run, keyed by the template name
    # (fused_triples_mixed_fp64 is fused_triples_mixed<0, 0, false>:
    # precision, epilogue (0: the energy), narrowed contraction ranges), in
    # the kernel's own summary or in the summary of all precisions
    base, _, prec = kernel.rpartition("_")
    key = f"{base}<{ {'fp64': 0, 'fp32': 1, 'bf16x3': 2, 'bf16x6': 3}.get(prec, -1)}, 0, false>"
    for name in (kernel, base + "_all_precisions"):
        path = os.path.join(ROOT, "profiles", name + "_pmc_" + workload + ".json")
        if not os.path.exists(path):
            continue
        with open(path) as f:
            d = json.load(f)
        if "hbm_bytes_per_launch" in d:
            return d["hbm_bytes_per_launch"]
        v = d.get(key)
        if isinstance(v, dict) and "hbm_bytes_per_launch" in v:
            return v["hbm_bytes_per_launch"]
    return None


def host_cpu() -> dict:
    """The host's CPU model and the CPUs this process may run on."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cpus = len(os.sched_getaffinity(0))
    # physical cores among them: distinct (package, core id) pairs
    cores = set()
    for c in os.sched_getaffinity(0):
        try:
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/core_id") as f:
                core = f.read().strip()
            with open(f"/sys/devices/system/cpu/cpu{c}/topology/physical_package_id") as f:
                pkg = f.read().strip()
            cores.add((pkg, core))
        except OSError:
            cores.add(("?", c))
    return {"model": model, "cpus_in_affinity": cpus, "physical_cores_in_affinity": len(cores)}


def cpu_threads_default() -> int:
    """All physical cores this process may use, capped by OMP_NUM_THREADS when
    the environment sets it (a shared host's share of its CPUs)."""
    n = host_cpu()["physical_cores_in_affinity"]
    env = os.environ.get("OMP_NUM_THREADS")
    if env and env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(tm, threads: int):
    """The reference's own Fortran kernels (oracle/_ref/libsdt_ref_omp.so,
    ccsd_t_kernels_omp.F built -O3 -fopenmp) timed on the host over one full
    task of the job at its exact per-call shapes: all 27 variants, each d2
    variant once per p7 tile (K = that tile), each d1 variant once per h7
    tile, each s1 variant once -- the 9 nvab + 9 noab + 9 calls the
    ccsd_t_gpu.F task loop makes for the job's first task (ccsd_t_doubles_gpu.F,
    ccsd_t_singles_gpu.F).  The rate is that task's F_ref flops over the wall
    time of its calls; OMP_PROC_BIND=close, OMP_PLACES=cores."""
    import ctypes as C

    hc = host_cpu()  # before the OpenMP runtime binds this thread to one core
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    import oracle

    try:
        R = oracle.ref_lib(omp=True)
        kind = "reference"
    except (FileNotFoundError, OSError):
        return None
    p4, p5, p6, h1, h2, h3 = tm.tasks()[0]
    r = tm.range
    dims = [int(r[h3 - 1]), int(r[h2 - 1]), int(r[h1 - 1]), int(r[p6 - 1]), int(r[p5 - 1]), int(r[p4 - 1])]
    h3d, h2d, h1d, p6d, p5d, p4d = dims
    n3 = int(np.prod(dims))
    ks_p = [int(r[k]) for k in range(tm.noab, tm.noab + tm.nvab)]  # one d2 call per p7 tile
    ks_h = [int(r[k]) for k in range(tm.noab)]                     # one d1 call per h7 tile
    rng = np.random.default_rng(0)
    t3 = np.zeros(n3)
    t3 += 1e-3  # first touch outside the timed region
    DP = C.POINTER(C.c_double)
    kp, kh = max(ks_p), max(ks_h)
    ops = {"s1": (rng.standard_normal(p4d * h1d), rng.standard_normal(h3d * h2d * p6d * p5d)),
           "d1": (rng.standard_normal(kh * p4d * p5d * h1d), rng.standard_normal(h3d * h2d * p6d * kh)),
           "d2": (rng.standard_normal(kp * p4d * h1d * h2d), rng.standard_normal(kp * h3d * p6d * p5d))}
    calls = {"s1": [1], "d1": ks_h, "d2": ks_p}
    flops, n_calls = 0.0, 0
    t0 = time.perf_counter()
    for fam in ("d2", "d1", "s1"):
        a, b = ops[fam]
        for v in range(1, 10):
            fn = getattr(R, f"sd_t_{fam}_{v}_")
            for K in calls[fam]:
                args = [C.c_long(x) for x in (dims if fam == "s1" else dims + [K])]
                fn(*[C.byref(x) for x in args], t3.ctypes.data_as(DP), a.ctypes.data_as(DP),
                   b.ctypes.data_as(DP))
                flops += 2.0 * n3 * K
                n_calls += 1
    t_tot = time.perf_counter() - t0
    rate = flops / t_tot
    return {
        "value": rate / 1e12,
        "unit": "TFLOP/s",
        "cores": threads,
        "kind": kind,
        "sample": "one full task (the first) of this job on the reference Fortran kernels: "
                  "sd_t_d2_1..9 x %d p7 tiles (K=%s), sd_t_d1_1..9 x %d h7 tiles (K=%s), "
                  "sd_t_s1_1..9 once, %d calls on a t3 tile of %d elements, %.4e flops in %.1f s "
                  "on %d OpenMP threads; the full job (%d tasks of this shape) at this rate "
                  "would take %.0f s" % (len(ks_p), "/".join(map(str, sorted(set(ks_p)))), len(ks_h),
                                         "/".join(map(str, sorted(set(ks_h)))), n_calls, n3, flops,
                                         t_tot, threads, len(tm.tasks()), f_ref(tm) / rate),
        "seconds": t_tot,
        "cpu_model": hc["model"],
        "host_cpus_in_affinity": hc["cpus_in_affinity"],
        "host_physical_cores_in_affinity": hc["physical_cores_in_affinity"],
        "omp": {"OMP_NUM_THREADS": threads, "OMP_PROC_BIND": os.environ.get("OMP_PROC_BIND"),
                "OMP_PLACES": os.environ.get("OMP_PLACES")},
    }


def ragged(n: int, ts: int) -> list:
    """Tiles of ts with a ragged last tile (n = k*ts + r)."""
    r = [ts] * (n // ts)
    if n % ts:
        r.append(n % ts)
    return r


def small_case(threads: int) -> dict | None:
    """The small case (no=8 in one tile, nv=32 in tiles of 16, 4 tasks) whole: the
    reference CPU path (its compiled Fortran kernels under the ccsd_t_gpu.F
    task loop at usedevice = 0) on one core and on `threads` OpenMP threads,
    and the GPU job on the same inputs."""
    import oracle
    from nwchem_amd import TriplesJob, tiles as tl

    tm = tl.TileModel.synthetic(8, 32, 8, 16, 1234)
    b = tl.synth_blocks(tm, 1234)
    ot = oracle.Tiles(tm.range, tm.noab, tm.evl)
    out = {"workload": "no=8 (1 tile), nv=32 (2 tiles of 16), 4 tasks, seed 1234", "flops": f_ref(tm)}
    try:
        for name, omp in (("cpu_1core", False), ("cpu_all_cores", True)):
            be = oracle.ref_backend(omp=omp)
            t0 = time.perf_counter()
            e = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, backend=be)
            out[name] = {"seconds": time.perf_counter() - t0, "E_bt": e[0], "E_pt": e[1],
                         "build": "-O3 -fopenmp" if omp else "-O2, no OpenMP, no FP contraction"}
    except (FileNotFoundError, OSError):
        return None
    out["cpu_all_cores"]["threads"] = int(os.environ.get("OMP_NUM_THREADS", threads))
    with TriplesJob(tm) as j:
        j.upload(b)
        j.run()  # first launch of the kernel (code load) outside the timing
        t0 = time.perf_counter()
        e = j.run()
        out["gpu"] = {"seconds": time.perf_counter() - t0, "kernel_ms": j.last_kernel_ms,
                      "E_bt": e[0], "E_pt": e[1]}
        # reduced-precision modes at a physical energy scale (|E| ~ 10 Eh):
        # |dE| against the FP64 result, in micro-Hartree
        out["precision_modes_dE_uEh"] = {}
        for name, p in (("fp32", 1), ("bf16x3", 2), ("bf16x6", 3)):
            em = j.run(p)
            out["precision_modes_dE_uEh"][name] = [(em[0] - e[0]) * 1e6, (em[1] - e[1]) * 1e6]
    out["dE_gpu_vs_cpu"] = [e[0] - out["cpu_1core"]["E_bt"], e[1] - out["cpu_1core"]["E_pt"]]
    print(f"[bench] no=8/nv=32: reference CPU {out['cpu_1core']['seconds']:.2f} s on 1 core, "
          f"{out['cpu_all_cores']['seconds']:.2f} s on {out['cpu_all_cores']['threads']} threads; "
          f"GPU {out['gpu']['seconds'] * 1e3:.2f} ms; dE {out['dE_gpu_vs_cpu']}; reduced precision "
          f"dE (uEh) {out['precision_modes_dE_uEh']}", file=sys.stderr, flush=True)
    return out


def bench_batch(njobs: int, seed: int, steps: int, modes=(0, 2, 3), no_range=(6, 30),
                nv_range=(48, 256)) -> dict:
    """Many independent jobs (no ~ U{6..30}, nv ~ U{48..256} by default,
    occupied tiles 6-20, virtual tiles 16-48, ragged last tiles) in one
    persistent launch per run (autotuned, replayed as a hipGraph), against the
    same jobs run one after another.  The jobs use the orbital tiling (E[T],
    E(T) per job)."""
    from nwchem_amd import TILING_ORBITAL, TriplesBatch, TriplesJob, tiles as tl

    rng = np.random.default_rng(seed)
    jobs, F = [], 0.0
    Fm = {p: 0.0 for p in modes}
    t0 = time.perf_counter()
    for k in range(njobs):
        no, nv = int(rng.integers(no_range[0], no_range[1] + 1)), int(rng.integers(nv_range[0], nv_range[1] + 1))
        to, tv = int(rng.integers(6, 21)), int(rng.integers(16, 49))
        tm = tl.TileModel.synthetic(no, nv, 0, 0, seed + k, ragged(no, to), ragged(nv, tv))
        b = tl.synth_blocks(tm, seed + k)
        j = TriplesJob(tm, tiling=TILING_ORBITAL)  # a batch returns each job's E[T], E(T) only
        j.upload(b)
        del b
        jobs.append(j)
        F += f_ref(tm)
        for p in modes:
            Fm[p] += j.work(p)[0]
        print(f"[bench] batch job {k}: no={no} nv={nv} tiles {to}/{tv} ({len(tm.tasks())} tasks) "
              f"ready at {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    t_prep = time.perf_counter() - t0
    out = {"jobs": njobs, "flops_reference_tiling": F, "prep_s": t_prep}
    names = {0: "fp64", 1: "fp32", 2: "bf16x3", 3: "bf16x6"}
    with TriplesBatch(jobs) as batch:
        for p in modes:
            print(f"[bench] batch {names[p]}: planning (one timed run per candidate, graph capture)",
                  file=sys.stderr, flush=True)
            t1 = time.perf_counter()
            batch.run(p)  # plan: tuning runs, graph capture, first replay
            t_plan = time.perf_counter() - t1
            print(f"[bench] batch {names[p]}: planned in {t_plan:.1f} s: {batch.plan(p)}",
                  file=sys.stderr, flush=True)
            ms = []
            t1 = time.perf_counter()
            for st in range(steps):
                batch.run(p)
                ms.append(batch.last_ms)
                print(f"[bench] batch {names[p]}: replay {st} {batch.last_ms:.1f} ms",
                      file=sys.stderr, flush=True)
            wall = (time.perf_counter() - t1) / steps
            seq = 0.0
            for k, j in enumerate(jobs):  # the same jobs one launch each
                t2 = time.perf_counter()
                j.run(p)
                seq += time.perf_counter() - t2
                if k % 16 == 15:
                    print(f"[bench] batch {names[p]}: one by one, {k + 1} jobs in {seq:.1f} s",
                          file=sys.stderr, flush=True)
            out[names[p]] = {"ms_per_run": wall * 1e3, "graph_ms": float(np.mean(ms)),
                             "energies_per_s": njobs / wall, "flops_per_run": Fm[p],
                             "fp64_equiv_tflops": Fm[p] / wall / 1e12,
                             "sequential_ms": seq * 1e3, "speedup_vs_sequential": seq / wall,
                             "plan": batch.plan(p), "plan_s": t_plan}
            print(f"[bench] batch {names[p]}: {wall*1e3:.1f} ms per run of {njobs} jobs "
                  f"({njobs/wall:.1f} energies/s, {Fm[p]/wall/1e12:.1f} TF/s), one by one {seq*1e3:.1f} ms; "
                  f"plan {batch.plan(p)}", file=sys.stderr, flush=True)
    for j in jobs:
        j.close()
    return out


def lambda_inputs(tm, seed: int):
    """Synthetic Lambda inputs of the job's tiles: the y1/y2 blocks of a
    second seeded amplitude set, laid out as tce_y1_offset / tce_y2_offset
    keep them ([h][p], [h][h'][p][p']), and a seeded one-electron operator
    over the tce_mo1e_offset table."""
    from nwchem_amd import tiles as tl

    b = tl.synth_blocks(tm, seed)
    noab, r = tm.noab, tm.range
    nvab = len(r) - noab
    N = noab + nvab

    def table(h):
        n = int(h[0])
        return [(int(h[1 + i]), int(h[1 + n + i])) for i in range(n)]

    t1 = dict(table(b.t1h))
    t2 = dict(table(b.t2h))
    y1h, y2h, f1h = tm.y1_hash(), tm.y2_hash(), tm.f1_hash()
    y1 = np.zeros(b.t1.size)
    for key, off in table(y1h):
        h, p = key // nvab + 1, key % nvab + noab + 1
        rh, rp = int(r[h - 1]), int(r[p - 1])
        to = t1[h - 1 + noab * (p - noab - 1)]
        y1[off:off + rh * rp] = b.t1[to:to + rh * rp].reshape(rp, rh).T.ravel()
    y2 = np.zeros(b.t2.size)
    for key, off in table(y2h):
        p2 = key % nvab + noab + 1
        q = key // nvab
        p1 = q % nvab + noab + 1
        q //= nvab
        h4, h3 = q % noab + 1, q // noab + 1
        shp = tuple(int(r[x - 1]) for x in (p1, p2, h3, h4))
        n = int(np.prod(shp))
        to = t2[h4 - 1 + noab * (h3 - 1 + noab * (p2 - noab - 1 + nvab * (p1 - noab - 1)))]
        y2[off:off + n] = b.t2[to:to + n].reshape(shp).transpose(2, 3, 0, 1).ravel()
    nf = sum(int(r[k // N] * r[k % N]) for k, _ in table(f1h))
    f1 = np.random.default_rng(seed).uniform(-0.05, 0.05, max(nf, 1))
    return b, y1, y1h, y2, y2h, f1, f1h


def bench_lambda(tm, blocks, seed: int, F_run: float, t_step: float):
    """tce_t_lambda_ccsd_t over every task of the job, orbital pairing: two
    passes of the fused kernel over the job's kernel tasks (the left-hand
    job's y3 into a window, the right-hand job's t3d with the energy), so
    the executed work is twice the (T) job's F; reported against two (T)
    energies' time."""
    from nwchem_amd import TriplesJob
    from nwchem_amd.job import LAMBDA_ORBITAL, lambda_ccsd_t

    t0 = time.perf_counter()
    _, y1, y1h, y2, y2h, f1, f1h = lambda_inputs(tm, seed)
    t_in = time.perf_counter() - t0
    modes = {}
    with TriplesJob(tm) as right, TriplesJob(tm) as left:
        right.upload(blocks)
        left.upload_lambda(y1, y1h, y2, y2h, blocks.v2, blocks.v2h)
        t1 = time.perf_counter()
        e1, e2, te = lambda_ccsd_t(right, left, f1, f1h, LAMBDA_ORBITAL)
        wall = time.perf_counter() - t1
        # split-BF16: the first call converts both jobs' operands (untimed)
        for name, prec in (("bf16x6", 3), ("bf16x3", 2)):
            lambda_ccsd_t(right, left, f1, f1h, LAMBDA_ORBITAL, prec)
            t1 = time.perf_counter()
            m1, m2, _ = lambda_ccsd_t(right, left, f1, f1h, LAMBDA_ORBITAL, prec)
            mw = time.perf_counter() - t1
            modes[name] = {"seconds": mw, "speedup_vs_fp64": wall / mw, "energy1": m1,
                           "energy2": m2, "dE_rel": max(abs(m1 - e1) / abs(e1),
                                                        abs(m2 - e2) / abs(e2))}
            print(f"[bench] lambda-ccsd(t) {name}: {mw:.2f} s ({wall / mw:.2f}x FP64), "
                  f"relative dE {modes[name]['dE_rel']:.2e}", file=sys.stderr, flush=True)
    out = {"tasks": int(te.shape[0]), "seconds": wall, "energy1": e1, "energy2": e2,
           "flops": 2.0 * F_run, "tflops": 2.0 * F_run / wall / 1e12,
           "frac_of_fp64_peak": 2.0 * F_run / wall / 1e12 / PEAK_FP64_TFLOPS,
           "vs_two_triples_energies": wall / (2.0 * t_step), "inputs_seconds": t_in,
           "precision_modes": modes}
    print(f"[bench] lambda-ccsd(t): {wall:.2f} s for {out['tasks']} tasks "
          f"({out['tflops']:.1f} TF/s, {wall / (2.0 * t_step):.3f}x two (T) energies), "
          f"energy1={e1:.12e} energy2={e2:.12e}", file=sys.stderr, flush=True)
    return out


def cr_identity_inputs(tm, blocks):
    """The two moment intermediates of cr_ccsd_t_N set to their first terms,
    the bare integrals: i1(h11 p4 h1 h2) = v(h11 p4 h1 h2) and i1(p4 p5 h1
    p12) = v(p4 p5 h1 p12), copied from the job's v2 blocks into the tables
    of OFFSET_cr_ccsd_t_N_1_1 (blocks [p4][h11][h1][h2]) and _2_1 (blocks
    [p4][p5][h1][p12]).  The moment tile is then the (T) doubles tile, so
    num1 = E[T] and num2 = E(T) at any size."""
    noab, r = tm.noab, tm.range
    nvab = len(r) - noab
    N = noab + nvab
    spin, sym = tm.spin, tm.sym

    def table(h):
        n = int(h[0])
        return dict((int(h[1 + i]), int(h[1 + n + i])) for i in range(n))

    def ok(a, b, c, d):
        sp = spin[a - 1] + spin[b - 1] + spin[c - 1] + spin[d - 1]
        return ((not tm.restricted or sp != 8)
                and spin[a - 1] + spin[b - 1] == spin[c - 1] + spin[d - 1]
                and (sym[a - 1] ^ sym[b - 1] ^ sym[c - 1] ^ sym[d - 1]) == 0)

    v2 = table(blocks.v2h)

    def vkey(a, b, c, d):
        return d - 1 + N * (c - 1 + N * (b - 1 + N * (a - 1)))

    def build(entries, src_key, perm):
        keys, offs, size = [], [], 0
        for key, tl4 in entries:
            keys.append(key)
            offs.append(size)
            size += int(np.prod([r[x - 1] for x in tl4]))
        out = np.zeros(max(size, 1))
        for (key, tl4), off in zip(entries, offs):
            shp = [int(r[x - 1]) for x in src_key(tl4)]
            n = int(np.prod(shp))
            so = v2[vkey(*src_key(tl4))]
            out[off:off + n] = blocks.v2[so:so + n].reshape(shp).transpose(perm).ravel()
        h = np.array([len(keys)] + keys + offs, dtype=np.int64)
        return out, h

    e1 = [(h2 - 1 + noab * (h1 - 1 + noab * (h11 - 1 + noab * (p4 - noab - 1))), (p4, h11, h1, h2))
          for p4 in range(noab + 1, N + 1) for h11 in range(1, noab + 1)
          for h1 in range(1, noab + 1) for h2 in range(h1, noab + 1) if ok(h11, p4, h1, h2)]
    e2 = [(p12 - noab - 1 + nvab * (h1 - 1 + noab * (p5 - noab - 1 + nvab * (p4 - noab - 1))),
           (p4, p5, h1, p12))
          for p4 in range(noab + 1, N + 1) for p5 in range(p4, N + 1)
          for h1 in range(1, noab + 1) for p12 in range(noab + 1, N + 1) if ok(p4, p5, h1, p12)]
    i1_1, i1_1h = build(e1, lambda t: (t[1], t[0], t[2], t[3]), (1, 0, 2, 3))
    i1_2, i1_2h = build(e2, lambda t: t, (0, 1, 2, 3))
    return i1_1, i1_1h, i1_2, i1_2h


def bench_cr(tm, blocks, e_bt: float, e_pt: float, t_step: float):
    """tce_t_cr_ccsd_t over every task of the job (FP64, then 6-term
    split-BF16): the left pass over
    the (T) job, the CR right pass over the CR job (moment and denominator
    tiles), with the bare integrals as intermediates so that num1 and num2
    are the job's E[T] and E(T); reported against two (T) energies' time."""
    from nwchem_amd import TriplesJob
    from nwchem_amd.job import cr_ccsd_t

    t0 = time.perf_counter()
    i1_1, i1_1h, i1_2, i1_2h = cr_identity_inputs(tm, blocks)
    t_in = time.perf_counter() - t0
    with TriplesJob(tm) as tj, TriplesJob(tm) as cj:
        tj.upload(blocks)
        cj.upload_cr(blocks.t1, blocks.t1h, blocks.t2, blocks.t2h, i1_1, i1_1h, i1_2, i1_2h)
        t1 = time.perf_counter()
        e1, e2, terms = cr_ccsd_t(tj, cj)
        wall = time.perf_counter() - t1
        # 6-term split-BF16: the first call converts both jobs' operands (untimed)
        cr_ccsd_t(tj, cj, 3)
        t1 = time.perf_counter()
        m1, m2, _ = cr_ccsd_t(tj, cj, 3)
        mw = time.perf_counter() - t1
    rel = max(abs(terms["num1"] - e_bt) / abs(e_bt), abs(terms["num2"] - e_pt) / abs(e_pt))
    bf = {"seconds": mw, "speedup_vs_fp64": wall / mw, "energy1": m1, "energy2": m2,
          "dE_rel": max(abs(m1 - e1) / abs(e1), abs(m2 - e2) / abs(e2))}
    print(f"[bench] cr-ccsd(t) bf16x6: {mw:.2f} s ({wall / mw:.2f}x FP64), relative dE "
          f"{bf['dE_rel']:.2e}", file=sys.stderr, flush=True)
    out = {"seconds": wall, "energy1": e1, "energy2": e2, "terms": terms,
           "vs_two_triples_energies": wall / (2.0 * t_step), "inputs_seconds": t_in,
           "num_vs_triples_rel": rel, "precision_modes": {"bf16x6": bf}}
    print(f"[bench] cr-ccsd(t): {wall:.2f} s ({wall / (2.0 * t_step):.3f}x two (T) energies), "
          f"num1/num2 against E[T]/E(T): rel. diff {rel:.2e}, energy1={e1:.12e} "
          f"energy2={e2:.12e}", file=sys.stderr, flush=True)
    return out


def bench_serve(nreq: int, lam: float, npool: int, seed: int, slice_ms: float,
                rates: dict | None = None) -> dict:
    """Serving: a Poisson stream (rate lam) of requests for a pool of seeded
    systems (no in [6,40], nv in [48,320], log-uniform; tiles of 20/40 by the
    tce_tile rule), each FP64 or 6-term split-BF16 at random; systems are uploaded
    on first use and stay resident.  Latency = finish - submit."""
    from nwchem_amd import TriplesServer, tiles as tl

    rng = np.random.default_rng(seed)
    pool = []
    t0 = time.perf_counter()
    for k in range(npool):
        no = int(round(np.exp(rng.uniform(np.log(6), np.log(40)))))
        nv = int(round(np.exp(rng.uniform(np.log(48), np.log(320)))))
        tm = tl.TileModel.synthetic(no, nv, min(no, 20), min(nv, 40), seed + 100 + k)
        pool.append((tm, tl.synth_blocks(tm, seed + 100 + k), {p: job_work(tm, p) for p in (0, 3)}))
        print(f"[bench] serve system {k}: no={no} nv={nv} ready at {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
    t_prep = time.perf_counter() - t0
    arrivals = np.cumsum(rng.exponential(1.0 / lam, nreq))
    picks = rng.integers(0, npool, nreq)
    precs = rng.choice([0, 3], nreq)  # FP64 or split-BF16 at <= 1 uEh (6 terms)
    lat, F = [], 0.0
    with TriplesServer(slice_ms=slice_ms, max_active=16) as srv:
        sids = [srv.add_system(tm, b) for tm, b, f in pool]
        tickets = []
        t0 = time.perf_counter()
        for i in range(nreq):
            dt = arrivals[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            tickets.append(srv.submit(sids[picks[i]], int(precs[i])))
        print(f"[bench] serve: {nreq} requests submitted in {time.perf_counter() - t0:.1f} s",
              file=sys.stderr, flush=True)
        res = []
        for i, t in enumerate(tickets):
            while True:
                try:
                    res.append(srv.wait(t, 30.0))
                    break
                except TimeoutError:
                    print(f"[bench] serve: waiting for request {i} at {time.perf_counter() - t0:.1f} s",
                          file=sys.stderr, flush=True)
        wall = time.perf_counter() - t0
        st = srv.stats()
    wait = []
    for i, r in enumerate(res):
        lat.append(r["finish"] - r["submit"])
        wait.append(r["admit"] - r["submit"])
        F += pool[picks[i]][2][int(precs[i])]
    lat, wait = np.array(lat), np.array(wait)
    # offered load: the device time the arrivals ask for per second, from
    # each request's work at the rate of its precision measured alone
    # (rates[p] in flops/s, the main run's FP64 and bf16x3 job rates)
    demand = None
    if rates:
        per_req = np.array([pool[picks[i]][2][int(precs[i])] / rates[int(precs[i])] for i in range(nreq)])
        demand = float(per_req.mean() * lam)
    q = max(1, nreq // 4)
    arrival_rate = nreq / float(arrivals[-1]) if arrivals[-1] > 0 else None
    out = {"requests": nreq, "lambda_per_s": lam, "pool_systems": npool,
           "precision_mix": {"fp64": int((precs == 0).sum()), "bf16x6": int((precs == 3).sum())},
           "p50_s": float(np.percentile(lat, 50)), "p99_s": float(np.percentile(lat, 99)),
           "mean_s": float(lat.mean()), "wall_s": wall, "energies_per_s": nreq / wall,
           "arrival_rate_per_s": arrival_rate, "completion_rate_per_s": nreq / wall,
           "offered_load_at_no40_nv320_rates": demand,
           "queue_wait_first_quarter_p50_s": float(np.percentile(wait[:q], 50)),
           "queue_wait_last_quarter_p50_s": float(np.percentile(wait[-q:], 50)),
           "fp64_equiv_tflops": F / wall / 1e12, "slices": st["slices"], "prep_s": t_prep,
           "sizes": [[int(tm.no), int(tm.nv)] for tm, b, f in pool]}
    print(f"[bench] serve: {nreq} requests at {lam}/s over {npool} systems: p50 {out['p50_s']:.3f} s, "
          f"p99 {out['p99_s']:.3f} s, {out['energies_per_s']:.2f} energies/s (arrivals "
          f"{arrival_rate:.2f}/s, offered load {demand}), queue wait p50 first/last quarter "
          f"{out['queue_wait_first_quarter_p50_s']:.3f}/{out['queue_wait_last_quarter_p50_s']:.3f} s, "
          f"{st['slices']} slices", file=sys.stderr, flush=True)
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--no", type=int, default=40)
    ap.add_argument("--nv", type=int, default=320)
    ap.add_argument("--tile-o", type=int, default=20)
    ap.add_argument("--tile-v", type=int, default=40)
    ap.add_argument("--seed", type=int, default=20261202)
    ap.add_argument("--cpu-baseline", type=int, default=1, help="time the reference CPU kernels (0: skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baseline (0: all physical cores of this "
                         "process's affinity, capped by OMP_NUM_THREADS when set)")
    ap.add_argument("--batch-jobs", type=int, default=64,
                    help="batch of independent jobs (64 by default; 0: skip)")
    ap.add_argument("--batch-steps", type=int, default=3,
                    help="timed batch runs per precision (graph replays)")
    ap.add_argument("--batch-small-jobs", type=int, default=128,
                    help="a second batch of small jobs (no in [6,10], nv in [48,64]) that "
                         "underfill the GPU one by one (0: skip)")
    ap.add_argument("--mode-steps", type=int, default=5,
                    help="timed runs per reduced-precision mode (at most --steps)")
    ap.add_argument("--serve", type=int, default=200, help="serving requests (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-low", type=int, default=60,
                    help="requests of a second serving run below the device's capacity (0: skip)")
    ap.add_argument("--serve-low-rate", type=float, default=4.0)
    ap.add_argument("--serve-pool", type=int, default=12)
    ap.add_argument("--slice-ms", type=float, default=50.0)
    ap.add_argument("--reference-tiling", type=int, default=1,
                    help="also time the job with the reference tiles as kernel tasks (0: skip)")
    ap.add_argument("--task-tiles", type=int, default=1,
                    help="time tce_t_job_triples_tiles on this many tasks (0: skip)")
    ap.add_argument("--lambda", dest="lam", type=int, default=1,
                    help="also time tce_t_lambda_ccsd_t over the whole job (1) or skip it (0)")
    ap.add_argument("--cr", type=int, default=1,
                    help="also time tce_t_cr_ccsd_t over the whole job (1) or skip it (0)")
    ap.add_argument("--modes", default="fp32,bf16x3,bf16x6",
                    help="reduced-precision modes to time after FP64 (comma list; '' for none)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the (T) path shards by tasks across GPUs but this "
              "build runs and reports a single MI355X (use --gpus 1).", file=sys.stderr)
        return 2
    if args.steps < 1:
        print("--steps must be >= 1", file=sys.stderr)
        return 2

    from nwchem_amd import TriplesJob, tiles as tl

    t0 = time.perf_counter()
    tm = tl.TileModel.synthetic(args.no, args.nv, args.tile_o, args.tile_v, args.seed)
    blocks = tl.synth_blocks(tm, args.seed)
    t_gen = time.perf_counter() - t0
    F = f_ref(tm)
    workload = f"no{args.no}t{args.tile_o}_nv{args.nv}t{args.tile_v}"
    print(f"[bench] {workload}: {len(tm.tasks())} tasks, F_ref={F:.4e} flops, inputs {t_gen:.1f}s",
          file=sys.stderr, flush=True)

    job = TriplesJob(tm)
    t0 = time.perf_counter()
    job.upload(blocks)
    t_up = time.perf_counter() - t0
    print(f"[bench] upload (pack + H2D) {t_up:.1f}s, device bytes {job.device_bytes/2**30:.2f} GiB",
          file=sys.stderr, flush=True)
    for w in range(args.warmup):
        job.run()
        print(f"[bench] warmup {w} kernel {job.last_kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    kms = []
    t0 = time.perf_counter()
    for s in range(args.steps):
        e_bt, e_pt = job.run()
        kms.append(job.last_kernel_ms)
        print(f"[bench] step {s} kernel {kms[-1]:.1f} ms E[T]={e_bt:.12e} E(T)={e_pt:.12e}",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    kernel = job.kernel
    F_run, items = job.work(0)
    print(f"[bench] kernel tasks: {items} items, F = {F_run:.4e} flops "
          f"({F / F_run:.2f}x less than the reference tiling's F_ref)", file=sys.stderr, flush=True)
    modes = {}
    for name in [m for m in args.modes.split(",") if m]:
        prec = {"fp32": 1, "bf16x3": 2, "bf16x6": 3}[name]
        t1 = time.perf_counter()
        job.run(prec)  # converts the contraction tensors on first use (untimed)
        t_prep = time.perf_counter() - t1
        mk = []
        msteps = max(1, min(args.steps, args.mode_steps))
        t1 = time.perf_counter()
        for s in range(msteps):
            m_bt, m_pt = job.run(prec)
            mk.append(job.last_kernel_ms)
        mwall = time.perf_counter() - t1
        d_bt, d_pt = (m_bt - e_bt) * 1e6, (m_pt - e_pt) * 1e6
        print(f"[bench] {name}: kernel {np.mean(mk):.1f} ms (first run incl. conversion "
              f"{t_prep:.1f} s) E[T]={m_bt:.12e} dE[T]={d_bt:.3f} uEh, E(T)={m_pt:.12e} "
              f"dE(T)={d_pt:.3f} uEh", file=sys.stderr, flush=True)
        modes[name] = {
            "steps": msteps,
            "ms_per_step": mwall / msteps * 1e3,
            "kernel_ms": float(np.mean(mk)),
            "energies_per_s": msteps / mwall,
            "speedup_vs_fp64": (wall / args.steps) / (mwall / msteps),
            "fp64_equiv_tflops": job.work(prec)[0] * msteps / mwall / 1e12,
            "E_bt": m_bt, "E_pt": m_pt,
            "dE_bt_uEh": d_bt, "dE_pt_uEh": d_pt,
            "dE_rel": max(abs(m_bt - e_bt) / abs(e_bt), abs(m_pt - e_pt) / abs(e_pt)) if e_bt and e_pt else None,
            "kernel": f"fused_triples_mixed_{name}",
            "extra_device_bytes": job.precision_bytes(prec),
        }
        # the mode's own MFMA rate: its products (1, 3 or 6 per FP64-equivalent
        # product, contraction padding not counted) over the kernel time
        ach = MODE_PRODUCTS[name] * job.work(prec)[0] / (np.mean(mk) * 1e-3) / 1e12
        modes[name]["roofline"] = {
            "bound": "mfma", "achieved": ach, "peak": PEAK_MODE_TFLOPS[name], "unit": "TFLOP/s",
            "frac": ach / PEAK_MODE_TFLOPS[name], "products_per_fp64_product": MODE_PRODUCTS[name],
            "traffic": traffic_from_profiles(workload, f"fused_triples_mixed_{name}")}
    tiles_out = None
    if args.task_tiles > 0:
        # tce_t_job_triples_tiles: the first tasks' t3s/t3d tiles back on the
        # host, as a CPU consumer of ccsd_t_singles_l / ccsd_t_doubles_l gets them
        tk = tm.tasks()[:args.task_tiles]
        r = tm.range
        elems = [int(np.prod([r[b - 1] for b in t])) for t in tk]
        job.triples_tiles(tk[0][3], tk[0][4], tk[0][5], tk[0][0], tk[0][1], tk[0][2])  # warm
        t1 = time.perf_counter()
        for t in tk:
            job.triples_tiles(t[3], t[4], t[5], t[0], t[1], t[2])
        tw = time.perf_counter() - t1
        ft = sum(n * 18.0 * (args.no + args.nv + 1) for n in elems)
        # the same tiles left in HBM for a consumer running on the GPU
        t1 = time.perf_counter()
        for t in tk:
            job.triples_tiles_device(t[3], t[4], t[5], t[0], t[1], t[2])
        tdv = time.perf_counter() - t1
        tiles_out = {"tasks": len(tk), "seconds_per_task": tw / len(tk),
                     "elements_per_task": float(np.mean(elems)),
                     "task_flops_per_s": ft / tw,
                     "host_bytes_per_s": 16.0 * sum(elems) / tw,
                     "device_seconds_per_task": tdv / len(tk),
                     "device_task_flops_per_s": ft / tdv}
        print(f"[bench] task tiles: {tw / len(tk) * 1e3:.1f} ms per task "
              f"({np.mean(elems):.3g} elements, {ft / tw / 1e12:.1f} TF/s, "
              f"{16.0 * sum(elems) / tw / 1e9:.1f} GB/s of tiles to the host); left in HBM: "
              f"{tdv / len(tk) * 1e3:.1f} ms per task ({ft / tdv / 1e12:.1f} TF/s)",
              file=sys.stderr, flush=True)
    job.close()
    lam = bench_lambda(tm, blocks, args.seed + 1, F_run, wall / args.steps) if args.lam else None
    cr = bench_cr(tm, blocks, e_bt, e_pt, wall / args.steps) if args.cr else None
    reftile = None
    if args.reference_tiling:
        # the same job with the reference's tiles as kernel tasks: the work the
        # reference's own task loop does, timed like the steps above
        os.environ["TCE_T_SUBTILE"] = "0"
        try:
            rj = TriplesJob(tm)
        finally:
            del os.environ["TCE_T_SUBTILE"]
        with rj:
            rj.upload(blocks)
            t1 = time.perf_counter()
            r_bt, r_pt = rj.run()
            rwall = time.perf_counter() - t1
            rk = rj.last_kernel_ms
            rF, ritems = rj.work(0)
        reftile = {"ms_per_step": rwall * 1e3, "kernel_ms": rk, "items": ritems, "flops": rF,
                   "tflops": rF / rwall / 1e12, "kernel_tflops": rF / (rk * 1e-3) / 1e12,
                   "frac_of_peak": rF / (rk * 1e-3) / 1e12 / PEAK_FP64_TFLOPS,
                   "energies_per_s": 1.0 / rwall, "E_bt": r_bt, "E_pt": r_pt,
                   "dE_vs_subtiled_rel": max(abs(r_bt - e_bt) / abs(e_bt), abs(r_pt - e_pt) / abs(e_pt))}
        print(f"[bench] reference tiling: kernel {rk:.1f} ms, {ritems} items, "
              f"{reftile['kernel_tflops']:.2f} TF/s ({100 * reftile['frac_of_peak']:.1f}% of peak), "
              f"E[T]={r_bt:.12e} (rel. diff {reftile['dE_vs_subtiled_rel']:.2e})", file=sys.stderr, flush=True)

    ms_step = wall / args.steps * 1e3
    value = F_run * args.steps / wall / 1e12
    k_avg = float(np.mean(kms))
    achieved = F_run / (k_avg * 1e-3) / 1e12
    roofline = {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_FP64_TFLOPS, "traffic": traffic_from_profiles(workload, kernel),
                "kernel": kernel, "kernel_ms": k_avg, "flops_per_launch": F_run,
                "items_per_launch": items}
    batch = bench_batch(args.batch_jobs, args.seed, max(1, args.batch_steps)) if args.batch_jobs else None
    batch_small = (bench_batch(args.batch_small_jobs, args.seed + 7, max(1, args.batch_steps),
                               no_range=(6, 10), nv_range=(48, 64))
                   if args.batch_small_jobs else None)
    rates = {0: F_run * args.steps / wall}
    if "bf16x6" in modes:
        rates[3] = modes["bf16x6"]["fp64_equiv_tflops"] * 1e12
    serve = bench_serve(args.serve, args.serve_rate, args.serve_pool, args.seed, args.slice_ms,
                        rates if len(rates) == 2 else None) if args.serve else None
    # the same pool and request sequence at an arrival rate the device keeps up with
    serve_low = bench_serve(args.serve_low, args.serve_low_rate, args.serve_pool, args.seed,
                            args.slice_ms, rates if len(rates) == 2 else None) if args.serve_low else None
    cpu = None
    small = None
    if args.cpu_baseline:
        threads = args.cpu_threads or cpu_threads_default()
        print(f"[bench] cpu baseline on {threads} threads ({host_cpu()})", file=sys.stderr, flush=True)
        del blocks
        cpu = cpu_baseline(tm, threads)
        small = small_case(threads)
        if cpu:
            print(f"[bench] cpu baseline {cpu['value']*1e3:.2f} GFLOP/s ({cpu['seconds']:.1f}s)",
                  file=sys.stderr, flush=True)
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Philox inputs, resident in HBM)",
        "config": {"workload": f"(T) FP64 job, no={args.no} (tiles {args.tile_o}), nv={args.nv} "
                               f"(tiles {args.tile_v}), {len(tm.tasks())} tasks",
                   "no": args.no, "nv": args.nv, "tile_o": args.tile_o, "tile_v": args.tile_v,
                   "tasks": len(tm.tasks()), "precision": "fp64"},
        "energies_per_s": args.steps / wall,
        "E_bt": e_bt,
        "E_pt": e_pt,
        "flops_per_job": F_run,
        "flops_reference_tiling": F,
        "roofline": roofline,
        "reference_tiling": reftile,
        "cpu_baseline": cpu,
        "no8_nv32": small,
        "precision_modes": modes or None,
        "task_tiles": tiles_out,
        "lambda_ccsd_t": lam,
        "cr_ccsd_t": cr,
        "batch": batch,
        "batch_small_jobs": batch_small,
        "serving": serve,
        "serving_below_capacity": serve_low,
    }
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
