"""GPU parity: the HIP paths against the CPU oracle on the same seeded inputs.

Tolerance (FP64 path): |dE| <= max(1e-10 Eh, 1e-12 |E|) for E[T], E(T) and
every task's pair.  1e-10 Eh absolute is the stated target; the relative
floor only matters when |E| > 100 Eh (the random amplitudes of the synthetic
inputs give |E| of order 10 Eh at no=8/nv=32 and far more at larger sizes,
where FP64 summation of the reference itself cannot resolve 1e-10 Eh).
"""
import json
import os
import sys

import numpy as np
import pytest

import oracle
from nwchem_amd import TriplesJob, tiles as tl
from nwchem_amd import _lib

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from golden import make_golden as mg  # noqa: E402
import spin_systems  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

pytestmark = pytest.mark.gpu


def tol(e):
    return max(1e-10, 1e-12 * abs(e))


# The reduced-precision modes are checked against the oracle in the same
# operand representation within the accumulation bound of
# tests/precision_model.py (tests/test_gpu_precision.py); here, where a
# case is one of those fixture systems, the same references apply.
FIXTURE_OF_CASE = {(8, 32, 8, 16, 1234): "no8_nv32", (7, 13, 3, 5, 11): "no7_nv13_to3_tv5",
                   (5, 9, 2, 4, 12): "no5_nv9_to2_tv4", (9, 41, 0, 0, 15): "no9_nv41_ragged"}


def precision_fixture(name, mode):
    """(emulated totals, task energies, per-task bounds, total bounds)."""
    with np.load(os.path.join(GOLD, "precision.npz")) as z:
        return tuple(z[f"{name}_{mode}_{k}"] for k in ("total", "tasks", "task_bounds", "total_bounds"))


def system(no, nv, to, tv, seed, ranges_o=None, ranges_v=None):
    tm = tl.TileModel.synthetic(no, nv, to, tv, seed, ranges_o, ranges_v)
    b = tl.synth_blocks(tm, seed)
    ot = oracle.Tiles(tm.range, tm.noab, tm.evl)
    return tm, b, ot


def oracle_energies(ot, b):
    return oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)


def assert_parity(gpu_e, gpu_te, ref):
    e1, e2, te = ref
    assert abs(gpu_e[0] - e1) <= tol(e1), (gpu_e, e1, e2)
    assert abs(gpu_e[1] - e2) <= tol(e2), (gpu_e, e1, e2)
    assert gpu_te.shape == te.shape
    for t in range(te.shape[0]):
        assert abs(gpu_te[t, 0] - te[t, 0]) <= tol(te[t, 0]), t
        assert abs(gpu_te[t, 1] - te[t, 1]) <= tol(te[t, 1]), t


CASES = [
    # (no, nv, tile_o, tile_v, seed, ranges_o, ranges_v)
    (8, 32, 8, 16, 1234, None, None),     # the small case, no=8/nv=32
    (7, 13, 3, 5, 11, None, None),        # ranges not multiples of 4
    (5, 9, 2, 4, 12, None, None),
    (10, 22, 4, 6, 13, None, None),
    (6, 17, 6, 17, 14, None, None),       # one tile per space
    (9, 41, 0, 0, 15, [6, 3], [16, 16, 9]),  # ragged last tiles
    (4, 4, 4, 4, 16, None, None),
]


@pytest.mark.parametrize("kernel", ["wave", "wg", "wg16"])
@pytest.mark.parametrize("case", CASES, ids=[f"no{c[0]}_nv{c[1]}_to{c[2]}_tv{c[3]}" for c in CASES])
def test_job_matches_oracle(case, kernel, monkeypatch):
    """Every FP64 kernel: one wave per 4x4x4 particle cube (fused_wave), the
    LDS-cooperative 8-wave workgroup on 4x4x4 MFMA (fused_wg) and on 16x16x4
    MFMA (fused_mixed's FP64 policy)."""
    monkeypatch.setenv("TCE_T_KERNEL", kernel)
    no, nv, to, tv, seed, ro, rv = case
    tm, b, ot = system(no, nv, to, tv, seed, ro, rv)
    ref = oracle_energies(ot, b)
    print(f"case no={no} nv={nv} ranges={list(tm.range)} tasks={len(ref[2])} kernel={kernel}", flush=True)
    with TriplesJob(tm) as job:
        job.upload(b)
        e = job.run()
        te = job.task_energies()
    assert_parity(e, te, ref)


@pytest.mark.parametrize("kernel", ["wave", "wg16"])
@pytest.mark.parametrize("case", CASES, ids=[f"no{c[0]}_nv{c[1]}_to{c[2]}_tv{c[3]}" for c in CASES])
def test_reference_tiling_matches_oracle(case, kernel, monkeypatch):
    """TCE_T_SUBTILE=0: the reference tiles themselves as kernel tasks (boxes
    over each tile, the symmetric duplicates inside equal tiles computed and
    weighted as the reference does)."""
    monkeypatch.setenv("TCE_T_KERNEL", kernel)
    monkeypatch.setenv("TCE_T_SUBTILE", "0")
    no, nv, to, tv, seed, ro, rv = case
    tm, b, ot = system(no, nv, to, tv, seed, ro, rv)
    ref = oracle_energies(ot, b)
    print(f"case no={no} nv={nv} ranges={list(tm.range)} kernel={kernel} reference tiling", flush=True)
    with TriplesJob(tm) as job:
        job.upload(b)
        e = job.run()
        te = job.task_energies()
        f, n = job.work()
    r = tm.range
    f_ref = sum(float(r[a - 1] * r[b - 1] * r[c - 1] * r[i - 1] * r[j - 1] * r[k - 1])
                for a, b, c, i, j, k in tm.tasks()) * 18 * (no + nv + 1)
    assert f == f_ref and n >= len(ref[2])
    assert_parity(e, te, ref)


@pytest.mark.parametrize("kernel", ["wave", "wg16"])
@pytest.mark.parametrize("case", CASES, ids=[f"no{c[0]}_nv{c[1]}_to{c[2]}_tv{c[3]}" for c in CASES])
def test_orbital_tiling_matches_oracle(case, kernel, monkeypatch):
    """TILING_ORBITAL: box-sized pieces of the whole occupied and virtual
    spaces as kernel tasks, E[T] and E(T) against the oracle's sums over the
    reference tiles; per-task energies and restart blocks are refused."""
    from nwchem_amd import TILING_ORBITAL
    monkeypatch.setenv("TCE_T_KERNEL", kernel)
    no, nv, to, tv, seed, ro, rv = case
    tm, b, ot = system(no, nv, to, tv, seed, ro, rv)
    e1, e2, _ = oracle_energies(ot, b)
    print(f"case no={no} nv={nv} ranges={list(tm.range)} kernel={kernel} orbital tiling", flush=True)
    with TriplesJob(tm, tiling=TILING_ORBITAL) as job:
        job.upload(b)
        e = job.run()
        assert abs(e[0] - e1) <= tol(e1) and abs(e[1] - e2) <= tol(e2), (e, e1, e2)
        with pytest.raises(RuntimeError):
            job.task_energies()
        with pytest.raises(RuntimeError):
            job.run_restart()
        fx = FIXTURE_OF_CASE.get(case[:5])
        if fx:
            eb = np.array(job.run(2))
            tot, _, _, bnd = precision_fixture(fx, "bf16x3")
            assert np.all(np.abs(eb - tot) <= bnd), (eb, tot, bnd)


@pytest.mark.parametrize("prm", [dict(sizes_o=[3, 2], sizes_v=[4, 3], irreps_o=[0, 0], irreps_v=[1, 1], seed=5),
                                 dict(sizes_o=[2, 2], sizes_v=[5, 6], seed=10, flip_symmetric=True,
                                      restricted=True)], ids=["spin_irreps", "restricted"])
def test_orbital_tiling_spin_systems(prm):
    from nwchem_amd import TILING_ORBITAL
    s = spin_systems.SpinSystem(**prm)
    b = s.blocks
    e1, e2, _ = oracle.ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    with TriplesJob(s.tiles, tiling=TILING_ORBITAL) as job:
        job.upload(b)
        e = job.run()
    assert abs(e[0] - e1) <= tol(e1) and abs(e[1] - e2) <= tol(e2), (e, e1, e2)


def test_job_is_deterministic_and_reusable():
    tm, b, ot = system(8, 32, 8, 16, 99)
    with TriplesJob(tm) as job:
        job.upload(b)
        e1 = job.run()
        t1 = job.task_energies()
        e2 = job.run()
        t2 = job.task_energies()
    assert e1 == e2 and np.array_equal(t1, t2)


def test_job_zero_singles_equal_energies():
    tm, b, ot = system(6, 14, 3, 5, 21)
    b.t1[:] = 0
    with TriplesJob(tm) as job:
        job.upload(b)
        e = job.run()
        te = job.task_energies()
    assert e[0] == e[1] and np.array_equal(te[:, 0], te[:, 1])


@pytest.fixture(params=[0, 1], ids=["direct", "ttgt"])
def legacy_backend(request):
    """The per-call d1/d2 back end (tce_t_legacy_backend): the direct kernel
    or TTGT (DGEMM + transpose); restored to direct afterwards."""
    L = _lib.lib()
    assert L.tce_t_legacy_backend(request.param) == request.param
    yield request.param
    L.tce_t_legacy_backend(0)


@pytest.mark.parametrize("case", [CASES[0], CASES[1], CASES[5]],
                         ids=["no8_nv32", "ragged_small", "ragged_tiles"])
def test_legacy_surface_matches_oracle(case, legacy_backend):
    """ccsd_t_gpu.F's usedevice=1 call sequence (dev_mem_*, sd_t_*_cuda_,
    compute_en_, dev_release_, finalizememmodule_) against libtce_t.so, on
    both back ends of the d1/d2 calls."""
    no, nv, to, tv, seed, ro, rv = case
    tm, b, ot = system(no, nv, to, tv, seed, ro, rv)
    ref = oracle_energies(ot, b)
    L = _lib.lib()
    be = oracle.device_backend(L)
    got = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, backend=be)
    assert L.tce_t_legacy_status() == 0
    assert_parity(got[:2], got[2], ref)


def test_cpu_order_kernels_equal_the_reference_kernels(legacy_backend):
    """The CPU surface (sd_t_s1/d1/d2_N_ in ccsd_t_kernels_omp.F's argument
    order, accumulating into a host triplesx) against the tiles the
    reference's own compiled kernels leave on the golden inputs
    (tests/golden/kernels.npz; exact binary fractions, so bit for bit, in
    any summation order), on both back ends."""
    import ctypes as C
    L = _lib.lib()
    DPt = C.POINTER(C.c_double)
    with np.load(os.path.join(GOLD, "kernels.npz")) as z:
        gold = {k: z[k] for k in z.files}
    for si, shape in enumerate(mg.KERNEL_SHAPES):
        dims = [C.c_long(v) for v in shape]
        for fam in ("s1", "d1", "d2"):
            for n in range(1, 10):
                t3, a, b = mg.kernel_inputs(fam, shape, n)
                args = dims[:6] if fam == "s1" else dims
                getattr(L, f"sd_t_{fam}_{n}_")(*[C.byref(d) for d in args], t3.ctypes.data_as(DPt),
                                                a.ctypes.data_as(DPt), b.ctypes.data_as(DPt))
                assert np.array_equal(t3, gold[f"shape{si}_{fam}_{n}"]), (si, fam, n)
    assert L.tce_t_legacy_status() == 0


def test_fortran_caller_matches_oracle():
    """The ISO_C_BINDING module driven from Fortran (nwchem_amd/fortran)."""
    import os
    import re
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "nwchem_amd", "fortran", "tce_t_fdemo")
    if not os.path.exists(exe):
        pytest.skip("Fortran demo not built")
    out = subprocess.run([exe, "8", "32", "8", "16"], check=True, capture_output=True, text=True,
                         timeout=120).stdout
    m = re.search(r"E\[T\]=\s*(\S+)\s+E\(T\)=\s*(\S+)", out)
    assert m, out
    tm, b, ot = system(8, 32, 8, 16, 1234)
    e1, e2, _ = oracle_energies(ot, b)
    assert abs(float(m.group(1)) - e1) <= tol(e1) and abs(float(m.group(2)) - e2) <= tol(e2)
    # the ccsd_t_restart.F call pattern from Fortran: one outer tile per call
    r = re.search(r"restart blocks=\s*(\d+)\s+E\(T\)=\s*(\S+)", out)
    assert r, out
    assert int(r.group(1)) == tm.nvab and abs(float(r.group(2)) - e2) <= tol(e2)
    # tiles from tce_t_job_triples_tiles, the energy summed by the Fortran caller
    t = re.search(r"tiles E\[T\]=\s*(\S+)\s+E\(T\)=\s*(\S+)", out)
    assert t, out
    assert abs(float(t.group(1)) - e1) <= tol(e1) and abs(float(t.group(2)) - e2) <= tol(e2)
    # Lambda-CCSD(T) from Fortran: the left-hand job from y blocks the caller
    # laid out (seed 4321 amplitudes), f1 = 0.05 sin(k); against the oracle
    import bench
    lam = re.search(r"lambda energy1=\s*(\S+)\s+energy2=\s*(\S+)", out)
    assert lam, out
    _, y1, y1h, y2, y2h, f1, f1h = bench.lambda_inputs(tm, 4321)
    f1 = 0.05 * np.sin(np.arange(1, f1.size + 1, dtype=np.float64))
    want = oracle.lambda_ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, y1, y1h, y2, y2h, b.v2, b.v2h, f1,
                                f1h, 0)
    np.testing.assert_allclose([float(lam.group(1)), float(lam.group(2))], want[:2], rtol=1e-11)
    # CR-CCSD(T) from Fortran with the bare integrals as the intermediates (the
    # caller lays out the two i1 tables): num1, num2 are E[T], E(T); the energies
    # follow from the five printed terms as cr_ccsd_t.F:254-258 forms them
    cr = re.search(r"cr terms=\s*(\S+)\s+(\S+)\s+(\S+)\s+(\S+)\s+(\S+)", out)
    ce = re.search(r"cr energy1=\s*(\S+)\s+energy2=\s*(\S+)", out)
    assert cr and ce, out
    num1, num2, den0, den1, den2 = (float(x) for x in cr.groups())
    assert abs(num1 - e1) <= tol(e1) and abs(num2 - e2) <= tol(e2)
    np.testing.assert_allclose([float(ce.group(1)), float(ce.group(2))],
                               [num1 / (1.0 + den1 + den0), num2 / (1.0 + den2 + den0)],
                               rtol=1e-14)


# ---- golden vectors (reference kernels, tests/golden) and spin/irrep/restricted tilings

GOLDEN_SYSTEMS = list(mg.SYSTEMS)


def golden(name):
    with np.load(os.path.join(GOLD, "energies.npz")) as z:
        tot, te = z[name + "_total"], z[name + "_tasks"]
    return float(tot[0]), float(tot[1]), te


@pytest.mark.parametrize("name", GOLDEN_SYSTEMS)
def test_job_matches_golden(name):
    """The default kernel choice against the energies the reference's own
    kernels give on the same inputs (spin-labelled and restricted included)."""
    kind, prm = mg.SYSTEMS[name]
    ot, b, tm = mg.build_system(kind, prm)
    print(f"golden {name}: ranges={list(tm.range)}", flush=True)
    with TriplesJob(tm) as job:
        job.upload(b)
        e = job.run()
        te = job.task_energies()
        print(f"  kernel {job.kernel}", flush=True)
    assert_parity(e, te, golden(name))


SPIN_GPU = [
    dict(sizes_o=[3, 2], sizes_v=[4, 3], irreps_o=[0, 1], irreps_v=[1, 0], seed=5),
    dict(sizes_o=[8], sizes_v=[16, 8], irreps_o=[0], irreps_v=[0, 1], seed=9),
    dict(sizes_o=[4, 4], sizes_v=[8, 8], seed=10, flip_symmetric=True, restricted=True),
]


@pytest.mark.parametrize("kernel", ["wave", "wg", "wg16"])
@pytest.mark.parametrize("prm", SPIN_GPU, ids=["spin_irreps", "spin_8_16", "restricted"])
def test_spin_tilings_match_oracle(prm, kernel, monkeypatch):
    """Tiles labelled by spin and irrep (blocks and tasks that break
    conservation absent), and restricted runs, on every FP64 kernel."""
    monkeypatch.setenv("TCE_T_KERNEL", kernel)
    s = spin_systems.SpinSystem(**prm)
    b = s.blocks
    ref = oracle.ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    print(f"spin case {prm} kernel={kernel} tasks={len(ref[2])}", flush=True)
    with TriplesJob(s.tiles) as job:
        job.upload(b)
        e = job.run()
        te = job.task_energies()
    assert_parity(e, te, ref)


@pytest.mark.parametrize("name", ["c1_ragged", "irreps"])
def test_orbital_form_integrals_through_the_job(name):
    """A closed-shell system given as orbital-form integrals (intorb):
    converted by tce_t_v2_from_intorb, uploaded and run.  The blocks equal
    the spin-orbital ones built directly (CPU tests), so the energies must
    equal the job's on those bit for bit, and the oracle's within the FP64
    tolerance."""
    import test_intorb as ti

    ss, orb, am = ti.build(name)
    no_s, nv_s = ss.no // 2, ss.nv // 2
    ss.W, ss.V, ss.S = ti.spin_orbital_v(orb, no_s, nv_s)
    direct = ss._fill_blocks()
    io_o = oracle.Intorb(**am)
    h, osize = io_o.alpha_offset(2)
    v2orb = io_o.pack(orb, osize)
    T = ss.tiles
    v2 = tl.v2_from_intorb(T, tl.Intorb(**am), v2orb, h, direct.v2h, T.sizes()["v2"])
    conv = tl.Blocks(direct.t1, direct.t1h, direct.t2, direct.t2h, v2, direct.v2h)
    ref = oracle.ccsd_t(ss.oracle_tiles(), conv.t1, conv.t1h, conv.t2, conv.t2h, conv.v2, conv.v2h)
    print(f"intorb case {name}: tasks={len(ref[2])}", flush=True)
    out = []
    for blocks in (conv, direct):
        with TriplesJob(T) as job:
            job.upload(blocks)
            out.append((job.run(), job.task_energies()))
    assert out[0][0] == out[1][0] and np.array_equal(out[0][1], out[1][1])
    assert_parity(out[0][0], out[0][1], ref)


@pytest.mark.parametrize("precision", [0, 2], ids=["fp64", "bf16x3"])
@pytest.mark.parametrize("prm", SPIN_GPU, ids=["spin_irreps", "spin_8_16", "restricted"])
def test_forbidden_blocks_are_skipped_exactly(prm, precision, monkeypatch):
    """Spin- and irrep-labelled tilings: the 16x16 kernel contracts each pass
    only over the orbitals of the spin and irrep its blocks allow (the
    reference skips the other h7b / p7b tiles).  The skipped chunks hold
    zeros, so the energies equal the run over the whole spaces bit for bit."""
    s = spin_systems.SpinSystem(**prm)
    monkeypatch.setenv("TCE_T_KERNEL", "wg16")
    out = {}
    for sparse in ("0", "1"):
        monkeypatch.setenv("TCE_T_SPARSE", sparse)
        with TriplesJob(s.tiles) as job:
            job.upload(s.blocks)
            out[sparse] = (job.run(precision), job.task_energies())
    assert out["0"][0] == out["1"][0]
    assert np.array_equal(out["0"][1], out["1"][1])


def test_legacy_surface_spin_tiling():
    """The per-call surface under the same spin/irrep task loop."""
    s = spin_systems.SpinSystem(**SPIN_GPU[0])
    b = s.blocks
    ref = oracle.ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    L = _lib.lib()
    got = oracle.ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h,
                        backend=oracle.device_backend(L))
    assert L.tce_t_legacy_status() == 0
    assert_parity(got[:2], got[2], ref)


def test_mixed_precision_deterministic_and_fp64_unchanged():
    """Reduced-precision runs repeat bit for bit, and running them leaves
    the FP64 result of the same job untouched."""
    tm, b, ot = system(8, 32, 8, 16, 77)
    with TriplesJob(tm) as job:
        job.upload(b)
        f1 = job.run(0)
        a = [job.run(p) for p in (1, 2, 1, 2)]
        f2 = job.run(0)
    assert a[0] == a[2] and a[1] == a[3]
    assert f1 == f2
    assert a[0] != f1 and a[1] != f1


# ---- batches: several jobs in one persistent launch (hipGraph replays)
BATCH = [(8, 32, 8, 16, 21, None, None), (7, 13, 3, 5, 22, None, None),
         (10, 22, 4, 6, 23, None, None), (9, 41, 0, 0, 24, [6, 3], [16, 16, 9]),
         (6, 17, 6, 17, 25, None, None)]


def _batch_systems():
    out = []
    for no, nv, to, tv, seed, ro, rv in BATCH:
        tm, b, ot = system(no, nv, to, tv, seed, ro, rv)
        out.append((tm, b, ot))
    return out


def test_batch_matches_oracle_and_single_runs():
    """FP64 batch energies match the oracle per job; reduced-precision batch
    energies equal each job's own run bit for bit (same kernel, each job's
    partials reduced in its own fixed order)."""
    from nwchem_amd import TriplesBatch
    systems = _batch_systems()
    jobs = []
    try:
        for tm, b, ot in systems:
            j = TriplesJob(tm)
            j.upload(b)
            jobs.append(j)
        singles = {p: [j.run(p) for j in jobs] for p in (1, 2)}
        with TriplesBatch(jobs) as batch:
            e64 = batch.run(0)
            print("fp64 plan", batch.plan(0), flush=True)
            for k, (tm, b, ot) in enumerate(systems):
                e1, e2, te = oracle_energies(ot, b)
                assert abs(e64[k, 0] - e1) <= tol(e1), (k, e64[k], e1)
                assert abs(e64[k, 1] - e2) <= tol(e2), (k, e64[k], e2)
                assert_parity(tuple(e64[k]), jobs[k].task_energies(), (e1, e2, te))
            for p in (1, 2):
                ep = batch.run(p)
                print(p, "plan", batch.plan(p), flush=True)
                for k in range(len(jobs)):
                    assert tuple(ep[k]) == singles[p][k], (p, k, ep[k], singles[p][k])
    finally:
        for j in jobs:
            j.close()


def test_batch_of_mixed_tilings():
    """Jobs of different tilings (orbital, sub-tile, reference) in one
    batch: FP64 totals match the oracle, reduced-precision totals equal each
    job's own run bit for bit."""
    from nwchem_amd import TILING_ORBITAL, TILING_REFERENCE, TILING_SUBTILE, TriplesBatch
    systems = _batch_systems()
    tilings = [TILING_ORBITAL, TILING_SUBTILE, TILING_REFERENCE, TILING_ORBITAL, TILING_SUBTILE]
    jobs = []
    try:
        for (tm, b, ot), t in zip(systems, tilings):
            j = TriplesJob(tm, tiling=t)
            j.upload(b)
            jobs.append(j)
        singles = [j.run(2) for j in jobs]
        with TriplesBatch(jobs) as batch:
            e64 = batch.run(0)
            for k, (tm, b, ot) in enumerate(systems):
                e1, e2, _ = oracle_energies(ot, b)
                assert abs(e64[k, 0] - e1) <= tol(e1) and abs(e64[k, 1] - e2) <= tol(e2), k
            ep = batch.run(2)
            for k in range(len(jobs)):
                assert tuple(ep[k]) == singles[k], (k, ep[k], singles[k])
    finally:
        for j in jobs:
            j.close()


def test_batch_graph_replays_are_identical():
    """The captured graph replayed several times, with eager single-job runs
    and synchronisations in between, gives the eager result every time."""
    from nwchem_amd import TriplesBatch
    systems = _batch_systems()[:3]
    jobs = []
    try:
        for tm, b, ot in systems:
            j = TriplesJob(tm)
            j.upload(b)
            jobs.append(j)
        with TriplesBatch(jobs) as batch:
            for p in (0, 2):
                first = batch.run(p)  # plans (tuning runs are eager) and replays once
                for rep in range(3):
                    eager = [jobs[k].run(p) for k in range(len(jobs))]
                    again = batch.run(p)
                    assert np.array_equal(again, first), (p, rep)
                    if p == 2:
                        assert [tuple(x) for x in again] == eager
    finally:
        for j in jobs:
            j.close()


# ---- serving: continuous batching of requests
def test_server_results_same_alone_and_in_the_mix():
    """Each request's energies are identical (bit for bit) whether it ran
    alone on the server or overlapped with others of different sizes and
    precisions; FP64 requests also match the oracle."""
    from nwchem_amd import TriplesServer
    systems = _batch_systems()
    reqs = [(k, p) for k in range(len(systems)) for p in (0, 2)] + [(0, 1), (3, 1)]
    alone = {}
    with TriplesServer(slice_ms=2.0, max_active=8, reserve_bytes=0) as srv:
        sids = [srv.add_system(tm, b) for tm, b, ot in systems]
        for k, p in reqs:  # one at a time
            r = srv.wait(srv.submit(sids[k], p))
            alone[(k, p)] = (r["E_bt"], r["E_pt"])
    with TriplesServer(slice_ms=2.0, max_active=8, reserve_bytes=0) as srv:
        sids = [srv.add_system(tm, b) for tm, b, ot in systems]
        tickets = [(k, p, srv.submit(sids[k], p)) for k, p in reqs]  # all at once
        mix = {(k, p): srv.wait(t) for k, p, t in tickets}
        st = srv.stats()
    print("slices", st, flush=True)
    overlapped = 0
    for (k, p), r in mix.items():
        assert (r["E_bt"], r["E_pt"]) == alone[(k, p)], (k, p)
        assert r["submit"] <= r["admit"] <= r["start"] <= r["finish"]
    spans = sorted((r["start"], r["finish"]) for r in mix.values())
    for a, b in zip(spans, spans[1:]):
        overlapped += b[0] < a[1]
    assert overlapped > 0  # requests shared the device
    for k, (tm, b, ot) in enumerate(systems):
        e1, e2, te = oracle_energies(ot, b)
        assert abs(alone[(k, 0)][0] - e1) <= tol(e1)
        assert abs(alone[(k, 0)][1] - e2) <= tol(e2)


def test_server_evicts_idle_systems():
    """With free memory for about two and a half systems and two slots,
    requests for three different systems all complete: the third evicts an
    idle one.  Each result matches the system's own job run."""
    from nwchem_amd import TriplesServer
    from nwchem_amd.job import device_memory
    systems = [system(8, 64, 8, 16, 31 + k) for k in range(3)]
    ref, one = [], 0
    for tm, b, ot in systems:
        with TriplesJob(tm) as j:
            j.upload(b)
            ref.append(j.run())
            one = max(one, j.device_bytes)
    free_b, total_b = device_memory()
    reserve = max(0, free_b - (5 * one) // 2)
    print(f"system bytes {one}, free {free_b}, reserve {reserve}", flush=True)
    with TriplesServer(slice_ms=2.0, max_active=2, reserve_bytes=reserve) as srv:
        sids = [srv.add_system(tm, b) for tm, b, ot in systems]
        rs = [srv.wait(srv.submit(s, 0)) for s in sids[:1]]  # warm: one resident
        ts = [srv.submit(s, 0) for s in sids]
        rs = [srv.wait(t) for t in ts]
        st = srv.stats()
    print("stats", st, flush=True)
    assert st["resident"] <= 3
    for (e1, e2), r in zip(ref, rs):
        assert abs(r["E_bt"] - e1) <= tol(e1)
        assert abs(r["E_pt"] - e2) <= tol(e2)


# ---- restart (ccsd_t_restart.F): blocks of one outer virtual tile
@pytest.mark.parametrize("precision", [0, 2], ids=["fp64", "bf16x3"])
def test_restart_blocks_match_task_sums(precision):
    """Running the job in restartable blocks of outer virtual tiles, two per
    call with the state passed along, fills the table with each tile's
    E(T) (its tasks' oracle energies summed) and ends on the full E(T)."""
    tm, b, ot = system(9, 41, 0, 0, 15, [6, 3], [16, 16, 9])
    if precision == 0:
        e1, e2, te = oracle_energies(ot, b)
        tbound = np.array([[tol(x) for x in row] for row in te])
    else:
        tot, te, tbound, _ = precision_fixture("no9_nv41_ragged", "bf16x3")
        e1, e2 = tot
    tasks = tm.tasks()
    nvab = tm.range.size - tm.noab
    per_tile = np.zeros(nvab)
    tile_bound = np.zeros(nvab)
    for (p4, p5, p6, h1, h2, h3), (t4, t5), (b4, b5) in zip(tasks, te, tbound):
        per_tile[p4 - tm.noab - 1] += t5
        tile_bound[p4 - tm.noab - 1] += b5
    with TriplesJob(tm) as job:
        job.upload(b)
        st, calls = None, 0
        while st is None or "t_energy" not in st:
            st = job.run_restart(st, precision, max_tiles=2)
            calls += 1
        full = job.run(precision)
    assert calls == (nvab + 1) // 2
    assert np.all(np.abs(st["table"] - per_tile) <= tile_bound), (st["table"], per_tile)
    assert abs(st["t_energy"] - e2) <= tile_bound.sum()
    # the blocks are the same items as the whole run, reduced in another order
    assert abs(st["t_energy"] - full[1]) <= 1e-12 * abs(full[1])


# ---- a system without allowed tasks (every block breaks spin conservation)
def _empty_system():
    evl = np.sort(np.random.default_rng(1).uniform(-2.0, 4.0, 12))
    tm = tl.TileModel(range=[2, 2, 4, 4], noab=2, evl=evl, spin=[1, 1, 2, 2])
    return tm, tl.synth_blocks(tm, 3)


def test_empty_system_job_batch_and_server():
    """Holes all alpha, particles all beta: no block is stored and no task
    is allowed.  A job gives zero energies and an empty task list, a batch
    that holds it gives zeros for it, and a server request for it finishes
    without a launch; the other jobs are unaffected."""
    from nwchem_amd import TILING_ORBITAL, TriplesBatch, TriplesServer
    tm, b = _empty_system()
    assert len(tm.tasks()) == 0
    with TriplesJob(tm) as job:
        job.upload(b)
        assert job.run() == (0.0, 0.0)
        assert job.task_energies().shape == (0, 2)
        assert job.run(2) == (0.0, 0.0)
        assert job.work() == (0.0, 0)
    tm2, b2, ot2 = system(8, 32, 8, 16, 1234)
    e1, e2, _ = oracle_energies(ot2, b2)
    jobs = [TriplesJob(tm), TriplesJob(tm2, tiling=TILING_ORBITAL)]
    try:
        jobs[0].upload(b)
        jobs[1].upload(b2)
        with TriplesBatch(jobs) as batch:
            for p in (0, 2):
                e = batch.run(p)
                assert tuple(e[0]) == (0.0, 0.0)
                if p == 0:
                    assert abs(e[1, 0] - e1) <= tol(e1)
                else:
                    tot, _, _, bnd = precision_fixture("no8_nv32", "bf16x3")
                    assert abs(e[1, 0] - tot[0]) <= bnd[0]
    finally:
        for j in jobs:
            j.close()
    with TriplesServer(slice_ms=2.0, max_active=4, reserve_bytes=0) as srv:
        s0, s1 = srv.add_system(tm, b), srv.add_system(tm2, b2)
        ts = [srv.submit(s0, 0), srv.submit(s1, 0), srv.submit(s0, 2)]
        r = [srv.wait(t) for t in ts]
    assert (r[0]["E_bt"], r[0]["E_pt"]) == (0.0, 0.0) and (r[2]["E_bt"], r[2]["E_pt"]) == (0.0, 0.0)
    assert abs(r[1]["E_bt"] - e1) <= tol(e1) and abs(r[1]["E_pt"] - e2) <= tol(e2)


# ---- triples tiles of one task (tce_t_job_triples_tiles)

TILE_SYSTEMS = {
    "spin_irreps": ("spin", SPIN_GPU[0]),
    "restricted": ("spin", SPIN_GPU[2]),
    "ragged": ("synthetic", (9, 41, 0, 0, 15, [6, 3], [16, 16, 9])),
    "no7_nv13": ("synthetic", (7, 13, 3, 5, 11, None, None)),
}
TILE_REL = 1e-12  # FP64; the reduced-precision tiles: tests/test_gpu_precision.py


def _tile_system(name):
    kind, prm = TILE_SYSTEMS[name]
    if kind == "spin":
        s = spin_systems.SpinSystem(**prm)
        return s.tiles, s.blocks, s.oracle_tiles()
    return system(*prm)


def _check_tiles(job, tm, b, ot, precision, tasks):
    """Element-wise bound relative to the largest |t3d| (|t3s|) over the
    checked tasks: some tiles are zero up to cancellation (equal tiles whose
    elements vanish by antisymmetry), so a tile's own maximum is no scale."""
    rel = TILE_REL
    ref = {t: oracle.task_tiles(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, t) for t in tasks}
    ss = max(np.abs(r[0]).max() for r in ref.values())
    sd = max(np.abs(r[1]).max() for r in ref.values())
    assert ss > 0 and sd > 0
    for task in tasks:
        p4, p5, p6, h1, h2, h3 = task
        gs, gd = job.triples_tiles(h1, h2, h3, p4, p5, p6, precision)
        rs, rd = ref[task]
        assert np.abs(gd - rd).max() <= rel * sd, (task, np.abs(gd - rd).max(), sd)
        assert np.abs(gs - rs).max() <= rel * ss, (task, np.abs(gs - rs).max(), ss)


@pytest.mark.parametrize("name", list(TILE_SYSTEMS))
def test_triples_tiles_match_oracle(name):
    """t3s and t3d of every task, element by element, against the oracle's
    singles and doubles dispatch (FP64)."""
    tm, b, ot = _tile_system(name)
    tasks = tm.tasks()
    print(f"tiles {name}: {len(tasks)} tasks", flush=True)
    with TriplesJob(tm) as job:
        job.upload(b)
        _check_tiles(job, tm, b, ot, 0, tasks)


def test_triples_tiles_of_orbital_jobs():
    """A job whose kernel tasks are the orbital tiling exports the reference
    tiles all the same."""
    from nwchem_amd import TILING_ORBITAL
    tm, b, ot = _tile_system("ragged")
    with TriplesJob(tm, tiling=TILING_ORBITAL) as job:
        job.upload(b)
        _check_tiles(job, tm, b, ot, 0, tm.tasks()[::3])


@pytest.mark.parametrize("precision", [0, 2], ids=["fp64", "bf16x3"])
def test_triples_tiles_6dts_slices_equal_the_whole_tiles(precision):
    """The memory-sliced export (ccsd_t_6dts.F's xmem path): every p4 x p5
    slice, ragged last slices included, equals the same part of the whole
    tile bit for bit, in the slice layout triplesx(h3,h2,h1,p6,maxp5,maxp4)
    with the unused part of the leading corner left untouched (zero)."""
    tm, b, ot = _tile_system("ragged")
    with TriplesJob(tm) as job:
        job.upload(b)
        for task in (tm.tasks()[0], tm.tasks()[len(tm.tasks()) // 2], tm.tasks()[-1]):
            p4, p5, p6, h1, h2, h3 = task
            gs, gd = job.triples_tiles(h1, h2, h3, p4, p5, p6, precision)
            r4, r5 = gs.shape[0], gs.shape[1]
            m4, m5 = max(1, (r4 + 2) // 3), max(1, (r5 + 1) // 2)  # uneven slices
            for i0 in range(0, r4, m4):
                for j0 in range(0, r5, m5):
                    i1, j1 = min(r4, i0 + m4), min(r5, j0 + m5)
                    ss, sd = job.triples_tiles_6dts(h1, h2, h3, p4, p5, p6, i0 + 1, i1, j0 + 1, j1,
                                                    m4, m5, precision)
                    assert np.array_equal(sd[:i1 - i0, :j1 - j0], gd[i0:i1, j0:j1]), (task, i0, j0)
                    assert np.array_equal(ss[:i1 - i0, :j1 - j0], gs[i0:i1, j0:j1]), (task, i0, j0)
                    assert not sd[i1 - i0:].any() and not sd[:, j1 - j0:].any()
        with pytest.raises(RuntimeError):  # a slice larger than maxp4
            job.triples_tiles_6dts(h1, h2, h3, p4, p5, p6, 1, 3, 1, 1, 2, 1)


def test_triples_tiles_left_on_the_device_equal_the_host_tiles():
    """tce_t_job_triples_tiles_device: the task's tiles in HBM, read back
    with hipMemcpy, equal what tce_t_job_triples_tiles hands the host, bit
    for bit, in FP64 and split-BF16."""
    import ctypes as C

    hip = C.CDLL("/opt/rocm/lib/libamdhip64.so")
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemcpy.restype = C.c_int
    tm, b, _ = system(7, 13, 3, 5, 11)
    task = tm.tasks()[2]
    print(f"device tiles of task {task}", flush=True)
    with TriplesJob(tm) as job:
        job.upload(b)
        for prec in (0, 2):
            hs, hd = job.triples_tiles(task[3], task[4], task[5], task[0], task[1], task[2], prec)
            ps, pd, n = job.triples_tiles_device(task[3], task[4], task[5], task[0], task[1],
                                                 task[2], prec)
            assert n == hs.size
            for ptr, want in ((ps, hs), (pd, hd)):
                got = np.empty(n)
                assert hip.hipMemcpy(got.ctypes.data, ptr, n * 8, 2) == 0  # device to host
                assert np.array_equal(got, want.ravel()), prec


def test_triples_tiles_accumulate_and_reject_non_tasks():
    """The C entry point adds into the caller's buffers (as the drivers do);
    a sextuple that is not a task, or a job without upload, is an error."""
    import ctypes as C
    tm, b, ot = _tile_system("no7_nv13")
    L = _lib.lib()
    DP = C.POINTER(C.c_double)
    p4, p5, p6, h1, h2, h3 = tm.tasks()[5]
    shape = tuple(int(tm.range[t - 1]) for t in (p4, p5, p6, h1, h2, h3))
    with TriplesJob(tm) as job:
        with pytest.raises(RuntimeError):
            job.triples_tiles(h1, h2, h3, p4, p5, p6)
        job.upload(b)
        s, d = np.zeros(shape), np.zeros(shape)
        for _ in range(2):
            assert L.tce_t_job_triples_tiles(job._h, 0, h1, h2, h3, p4, p5, p6,
                                             s.ctypes.data_as(DP), d.ctypes.data_as(DP)) == 0
        gs, gd = job.triples_tiles(h1, h2, h3, p4, p5, p6)
        assert np.array_equal(s, 2 * gs) and np.array_equal(d, 2 * gd)
        assert L.tce_t_job_triples_tiles(job._h, 0, h1, h2, h3, p4, p5, p6, None,
                                         d.ctypes.data_as(DP)) == 0
        assert np.array_equal(d, 3 * gd)
        assert L.tce_t_job_triples_tiles(job._h, 0, h1, h2, h3, p4, p5, p6,
                                         s.ctypes.data_as(DP), None) == 0
        assert np.array_equal(s, 3 * gs) and np.array_equal(d, 3 * gd)
        n = tm.noab + tm.nvab
        for bad in [(h2, h1, h3, p4, p5, p6) if h1 != h2 else (h1, h2, h3, p5 + 1, p5, p6),
                    (0, h2, h3, p4, p5, p6), (h1, h2, h3, p4, p5, n + 1), (h1, h2, h3, 1, p5, p6)]:
            with pytest.raises(RuntimeError):
                job.triples_tiles(*bad)


# ---- upload: block fetch and sort on the device (pack.hip)

@pytest.mark.parametrize("name", ["ragged", "spin_irreps", "restricted", "no7_nv13"])
def test_device_pack_equals_host_pack(name, monkeypatch):
    """The dense tensors built on the device from the canonical blocks equal
    the host packer's bit for bit: identical energies, task energies and
    task tiles from uploads through either path."""
    tm, b, ot = _tile_system(name)
    tasks = tm.tasks()[::7]
    out = {}
    for mode in ("host", "device"):
        if mode == "host":
            monkeypatch.setenv("TCE_T_PACK", "host")
        else:
            monkeypatch.delenv("TCE_T_PACK", raising=False)
        with TriplesJob(tm) as job:
            job.upload(b)
            e = job.run()
            te = job.task_energies()
            tiles = [job.triples_tiles(t[3], t[4], t[5], t[0], t[1], t[2]) for t in tasks]
        out[mode] = (e, te, tiles)
    h, d = out["host"], out["device"]
    assert h[0] == d[0] and np.array_equal(h[1], d[1])
    for (hs, hd), (ds, dd) in zip(h[2], d[2]):
        assert np.array_equal(hs, ds) and np.array_equal(hd, dd)
    assert_parity(d[0], d[1], oracle_energies(ot, b) if name in ("ragged", "no7_nv13")
                  else oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h))


def _spread(data, hash_, gap):
    """The same blocks in a file with `gap` NaNs before every block (a file
    that holds other data between the blocks the path reads, as the full
    tce_mo2e_offset file holds other integral classes)."""
    L = int(hash_[0])
    offs = hash_[1 + L:1 + 2 * L]
    order = np.argsort(offs)
    ends = np.append(np.sort(offs)[1:], data.size)
    out, new = [], hash_.copy()
    pos = 0
    for k, o_end in zip(order, ends):
        o = offs[k]
        out.append(np.full(gap, np.nan))
        pos += gap
        new[1 + L + k] = pos
        out.append(data[o:o_end])
        pos += o_end - o
    out.append(np.full(gap, np.nan))
    return np.concatenate(out), new


def test_upload_reads_only_the_referenced_blocks():
    """NaNs between the blocks of all three files change nothing: the upload
    copies and reads only the blocks the path needs."""
    tm, b, ot = system(7, 13, 3, 5, 11)
    ref = oracle_energies(ot, b)
    t1, t1h = _spread(b.t1, b.t1h, 3)
    t2, t2h = _spread(b.t2, b.t2h, 5)
    v2, v2h = _spread(b.v2, b.v2h, 7)
    assert np.isnan(v2).sum() > 0
    with TriplesJob(tm) as job:
        job.upload(tl.Blocks(t1, t1h, t2, t2h, v2, v2h))
        e = job.run()
        te = job.task_energies()
    assert np.isfinite(e).all()
    assert_parity(e, te, ref)


def test_device_pack_reports_missing_block():
    """A block the spin/irrep rules require but the table lacks is an error
    of the upload (tce_hash's errquit), before anything reaches the device."""
    tm, b, ot = system(5, 9, 2, 4, 1)
    h = b.t2h.copy()
    L = int(h[0])
    h[1 + L // 2] += 10**9
    with TriplesJob(tm) as job:
        with pytest.raises(RuntimeError, match="missing"):
            job.upload(tl.Blocks(b.t1, b.t1h, b.t2, h, b.v2, b.v2h))


def test_reupload_replaces_every_copy():
    """A second upload into the same job replaces the tensors and every copy
    derived from them (reduced-precision operands): each precision then gives
    what a fresh job on the new blocks gives, bit for bit."""
    tm = tl.TileModel.synthetic(8, 32, 8, 16, 1234)
    b1, b2 = tl.synth_blocks(tm, 1234), tl.synth_blocks(tm, 4321)
    fresh = {}
    with TriplesJob(tm) as job:
        job.upload(b2)
        for p in (0, 1, 2):
            fresh[p] = job.run(p)
    with TriplesJob(tm) as job:
        job.upload(b1)
        first = {p: job.run(p) for p in (0, 1, 2)}
        job.upload(b2)
        for p in (0, 1, 2):
            assert job.run(p) == fresh[p], p
    assert first[0] != fresh[0]


def test_batch_refuses_a_reuploaded_job():
    """A batch's plans hold its jobs' device buffers: after one of its jobs
    is uploaded again the batch reports an error instead of running on freed
    memory."""
    from nwchem_amd import TILING_ORBITAL, TriplesBatch
    tm = tl.TileModel.synthetic(6, 14, 3, 5, 21)
    b = tl.synth_blocks(tm, 21)
    with TriplesJob(tm, tiling=TILING_ORBITAL) as j1, TriplesJob(tm, tiling=TILING_ORBITAL) as j2:
        j1.upload(b)
        j2.upload(b)
        with TriplesBatch([j1, j2]) as batch:
            e = batch.run(0)
            assert tuple(e[0]) == tuple(e[1])
            j2.upload(b)
            with pytest.raises(RuntimeError):
                batch.run(0)


def test_triples_tiles_larger_than_the_host_staging_chunk():
    """One task of 4.7e6 elements (one tile of 8 holes, one of 21
    particles): doubles and singles (9.5e6 doubles) come back through the
    pinned staging in two chunks of 8 Mi doubles, the boundary inside the
    singles half; both halves against the oracle."""
    tm, b, ot = system(8, 21, 8, 21, 31)
    tasks = tm.tasks()
    assert len(tasks) == 1
    p4, p5, p6, h1, h2, h3 = tasks[0]
    rs, rd = oracle.task_tiles(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, tasks[0])
    assert 2 * rd.size > 8 << 20
    with TriplesJob(tm) as job:
        job.upload(b)
        gs, gd = job.triples_tiles(h1, h2, h3, p4, p5, p6)
    assert np.abs(gd - rd).max() <= 1e-12 * np.abs(rd).max()
    assert np.abs(gs - rs).max() <= 1e-12 * np.abs(rs).max()


# ---- the benchmark's own size: no=40 (tiles of 20), nv=320 (tiles of 40), 480 tasks
C2 = None


def no40_nv320_inputs():
    """Inputs of the benchmark job, checked against the digest the golden
    task energies were computed from (tests/golden/no40_nv320_tasks.json)."""
    global C2
    if C2 is None:
        with open(os.path.join(GOLD, "no40_nv320_tasks.json")) as f:
            meta = json.load(f)
        prm = meta["params"]
        tm = tl.TileModel.synthetic(prm["no"], prm["nv"], prm["to"], prm["tv"], prm["seed"])
        b = tl.synth_blocks(tm, prm["seed"])
        print("no=40/nv=320 inputs generated", flush=True)
        assert mg.digest(b.t1, b.t2, b.v2) == meta["inputs_sha256"], "inputs changed"
        C2 = (tm, b, meta)
    return C2


def summation_tol(n, e_bt, e_pt):
    """What the reference's own energy sum of one task can be off by: its
    recursive summation (ccsd_t_gpu.F:170-200 / ccsd_t_dot.F) of n terms,
    probabilistic bound lam sqrt(n) u sum|terms| (Higham & Mary 2019, lam = 6,
    u = 2^-53), with sum|terms| <= 2 max(|E[T]|, |E(T)|) (the E[T] terms share
    one sign, the singles part of E(T) is a small correction)."""
    return 6.0 * np.sqrt(n) * 2.0 ** -53 * 2.0 * max(abs(e_bt), abs(e_pt))


def test_no40_nv320_fp64_tasks_match_reference():
    """Eight tasks of the benchmark job (first, last, equal / two equal /
    distinct tiles) against the energies of the reference's compiled kernels
    under the ccsd_t_gpu.F task loop, for the default sub-tiles and the
    reference tiles."""
    from nwchem_amd import TILING_REFERENCE
    tm, b, meta = no40_nv320_inputs()
    got = {}
    for name, tiling in (("subtiles", None), ("reference tiles", TILING_REFERENCE)):
        print(f"no=40/nv=320: {name}", flush=True)
        with (TriplesJob(tm) if tiling is None else TriplesJob(tm, tiling=tiling)) as job:
            job.upload(b)
            e = job.run()
            got[name] = (e, job.task_energies())
    assert got["subtiles"][1].shape == (meta["tasks_total"], 2)
    for t in meta["tasks"]:
        ref = np.array([float.fromhex(t["E_bt"]), float.fromhex(t["E_pt"])])
        tl_ = summation_tol(t["elements"], *ref)
        for name, (e, te) in got.items():
            d = np.abs(te[t["index"]] - ref)
            print(f"task {t['index']} {t['tiles']} {name}: |dE| = {d} (tol {tl_:.3e}, "
                  f"rel {d / np.abs(ref)})", flush=True)
            assert np.all(d <= tl_), (t, name, te[t["index"]], ref)
    e, te = got["subtiles"]
    er, ter = got["reference tiles"]
    for k in (0, 1):
        assert abs(e[k] - er[k]) <= 1e-12 * abs(er[k]), (e, er)
    assert np.all(np.abs(te - ter) <= 1e-12 * np.abs(ter) + 1e-9)


@pytest.mark.parametrize("mode", ["fp32", "bf16x3"])
def test_no40_nv320_reduced_precision_tasks(mode):
    """The benchmark job in reduced precision: two tasks against the oracle in
    the same representation (tests/golden/precision.json) within the
    accumulation bound of tests/precision_model.py; |dE| against the FP64
    reference printed for all eight fixture tasks."""
    import precision_model as pm
    tm, b, meta = no40_nv320_inputs()
    with open(os.path.join(GOLD, "precision.json")) as f:
        pmeta = json.load(f)["no40_nv320"]
    assert pmeta["inputs_sha256"] == meta["inputs_sha256"]
    with TriplesJob(tm) as job:
        job.upload(b)
        e = job.run(pm.PRECISION_ID[mode])
        te = job.task_energies()
    for t in meta["tasks"]:
        ref = np.array([float.fromhex(t["E_bt"]), float.fromhex(t["E_pt"])])
        print(f"task {t['index']} {t['tiles']} {mode}: E - E_fp64 = {(te[t['index']] - ref) * 1e6} uEh",
              flush=True)
    for t in pmeta["tasks"]:
        em = np.array([float.fromhex(t[mode]["E_bt"]), float.fromhex(t[mode]["E_pt"])])
        bnd = np.array([t[mode]["bound_bt"], t[mode]["bound_pt"]])
        d = np.abs(te[t["index"]] - em)
        print(f"task {t['index']} {mode}: |E - E_emul| = {d} (bound {bnd})", flush=True)
        assert np.all(d <= bnd), (t, te[t["index"]], em, bnd)


def test_no40_nv320_task_through_the_per_call_surface(legacy_backend):
    """The first task of the benchmark job through the reference's per-call
    GPU surface in ccsd_t_gpu.F order (the drop-in of sd_t_total.cu: host
    block fetch and sort per call, operands copied per call, t3 accumulated
    in device tiles, compute_en_), against the reference kernels' energy of
    that task; the time per task is printed (what the binary drop-in costs
    next to the whole-job path)."""
    import time
    tm, b, meta = no40_nv320_inputs()
    ot = oracle.Tiles(tm.range, tm.noab, tm.evl)
    L = _lib.lib()
    be = oracle.device_backend(L)
    t = meta["tasks"][0]
    k = t["index"]
    t0 = time.perf_counter()
    got = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, backend=be, task_begin=k,
                        task_end=k + 1)
    dt = time.perf_counter() - t0
    assert L.tce_t_legacy_status() == 0
    ref = np.array([float.fromhex(t["E_bt"]), float.fromhex(t["E_pt"])])
    d = np.abs(got[2][0] - ref)
    print(f"per-call surface ({['direct', 'ttgt'][legacy_backend]}), no=40/nv=320 task {k} {t['tiles']}: "
          f"{dt:.2f} s, |dE| = {d}", flush=True)
    assert np.all(d <= summation_tol(t["elements"], *ref))


def test_no40_nv320_orbital_tiling_agrees():
    """The orbital tiling decomposes the benchmark job's sum into other items
    (pieces of the whole spaces, other factors and padding): E[T] and E(T)
    agree with the sub-tiled job to 1e-12 relative."""
    from nwchem_amd import TILING_ORBITAL
    tm, b, meta = no40_nv320_inputs()
    with TriplesJob(tm) as job:
        job.upload(b)
        e = job.run()
    with TriplesJob(tm, tiling=TILING_ORBITAL) as job:
        job.upload(b)
        eo = job.run()
    for k in (0, 1):
        assert abs(e[k] - eo[k]) <= 1e-12 * abs(e[k]), (e, eo)
