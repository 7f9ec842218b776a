"""CPU tests of libtce_t.so's host half: integer bookkeeping (bit-exact against
the oracle), the seeded generator, the upload pack (GA gets + sorts + signs ->
dense HBM layouts) and the exported C-ABI surface.  No GPU calls here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle
from nwchem_amd import _lib
from nwchem_amd import tiles as tl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DP = C.POINTER(C.c_double)
LP = C.POINTER(C.c_long)


def test_philox_known_answers(lib):
    # Random123 kat_vectors, philox4x32 with 10 rounds
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, exp in cases:
        c = (C.c_uint * 4)(*ctr)
        k = (C.c_uint * 2)(*key)
        o = (C.c_uint * 4)()
        lib.tce_t_philox4x32(c, k, o)
        assert tuple(o) == exp


@pytest.mark.parametrize("n,ts", [(40, 20), (320, 40), (45, 20), (8, 8), (32, 16), (7, 3), (1, 5), (97, 24)])
def test_tile_ranges_closed_form_and_oracle(n, ts):
    r = tl.tile_ranges(n, ts)
    assert r.sum() == n and len(r) == -(-n // ts) and r.max() - r.min() <= 1 and r.max() <= ts
    ro = np.zeros(len(r), np.int64)
    oracle.lib().ora_tile_ranges(n, ts, ro.ctypes.data_as(LP))
    assert np.array_equal(r, ro)


SYSTEMS = [(8, 32, 8, 16), (40, 320, 20, 40), (7, 13, 3, 5), (12, 30, 5, 7)]


@pytest.mark.parametrize("no,nv,to,tv", SYSTEMS)
def test_hash_tables_bit_exact_vs_oracle(no, nv, to, tv):
    tm = tl.TileModel.synthetic(no, nv, to, tv, 1)
    ot = oracle.Tiles(tm.range, tm.noab, tm.evl)
    for kind, h in (("t1", tm.t1_hash()), ("t2", tm.t2_hash()), ("v2", tm.v2_hash())):
        ho, size = ot.hash(kind, 1)
        assert np.array_equal(h, ho), kind
        L = int(h[0])
        keys, offs = h[1:L + 1], h[L + 1:]
        assert np.all(np.diff(keys) > 0)  # the bisection of tce_hash needs sorted keys
        assert offs[0] == 0 and np.all(np.diff(offs) > 0)
        for i in range(0, L, max(1, L // 50)):
            assert _lib.lib().tce_t_hash(h.ctypes.data_as(LP), int(keys[i])) == offs[i]
            assert oracle.lib().ora_hash(h.ctypes.data_as(LP), int(keys[i])) == offs[i]
        assert _lib.lib().tce_t_hash(h.ctypes.data_as(LP), int(keys[-1]) + 1) == -1


def test_v2_subset_is_subsequence_of_full_mo2e_table():
    tm = tl.TileModel.synthetic(7, 13, 3, 5, 1)
    ot = oracle.Tiles(tm.range, tm.noab, tm.evl)
    full, _ = ot.hash("v2", 0)
    sub = tm.v2_hash()
    Lf, Ls = int(full[0]), int(sub[0])
    fk = list(full[1:Lf + 1])
    pos = [fk.index(k) for k in sub[1:Ls + 1]]
    assert pos == sorted(pos)
    # (T) reads: (h p|h h), (p p|h p), (p p|h h) -- check by brute decoding
    N = tm.noab + tm.nvab
    for k in sub[1:Ls + 1]:
        g2 = k % N + 1
        g1 = (k // N) % N + 1
        g4 = (k // N // N) % N + 1
        g3 = k // N // N // N + 1
        p = [g > tm.noab for g in (g3, g4, g1, g2)]
        assert p in ([False, True, False, False], [True, True, False, True], [True, True, False, False])


def test_generator_exact_antisymmetry_and_moments():
    no, nv = 6, 10
    d = tl.synth_dense(11, no, nv)
    T2, W, V, S = d["T2"], d["W"], d["V"], d["S"]
    assert np.array_equal(T2, -T2.transpose(1, 0, 2, 3))
    assert np.array_equal(T2, -T2.transpose(0, 1, 3, 2))
    assert np.array_equal(W, -W.transpose(0, 1, 3, 2))
    assert np.array_equal(V, -V.transpose(1, 0, 2, 3))
    assert np.array_equal(S, -S.transpose(1, 0, 2, 3)) and np.array_equal(S, -S.transpose(0, 1, 3, 2))
    big = tl.synth_dense(12, 20, 40)
    for name, sig in (("T1", 0.01), ("T2", 0.02), ("W", 0.05), ("V", 0.05), ("S", 0.05)):
        x = big[name]
        x = x[x != 0]
        assert abs(x.std() / sig - 1) < 0.05, name
        assert abs(x.mean()) < 0.05 * sig, name
    evl = np.zeros(30)
    _lib.lib().tce_t_synth_evl(3, 10, 20, evl.ctypes.data_as(DP))
    eo, ev = evl[:10], evl[10:]
    assert np.all(np.diff(eo) >= 0) and np.all(np.diff(ev) >= 0)
    assert eo.min() >= -2.0 and eo.max() <= -0.3 and ev.min() >= 0.2 and ev.max() <= 4.0


@pytest.mark.parametrize("no,nv,to,tv", [(8, 32, 8, 16), (7, 13, 3, 5), (5, 9, 2, 4)])
def test_upload_pack_reproduces_dense_tensors(no, nv, to, tv):
    """The GA-get half of the upload (hash lookup, tce_sort signs, tile
    transposes) must rebuild every dense tensor element the path reads."""
    seed = 314
    tm = tl.TileModel.synthetic(no, nv, to, tv, seed)
    b = tl.synth_blocks(tm, seed)
    d = tl.synth_dense(seed, no, nv)
    Ep, Mp = -(-nv // 8) * 8, -(-no // 8) * 8
    T1 = np.zeros((nv, no))
    T2e = np.zeros((no, no, nv, Ep))
    T2m = np.zeros((no, nv, nv, Mp))
    Wd1 = np.zeros((no, no, nv, Mp))
    Vd2 = np.zeros((no, nv, nv, Ep))
    S = np.zeros((no, no, nv, nv))
    rc = _lib.lib().tce_t_pack_dense(C.byref(tm.c()), b.t1.ctypes.data_as(DP), b.t1h.ctypes.data_as(LP),
                                     b.t2.ctypes.data_as(DP), b.t2h.ctypes.data_as(LP),
                                     b.v2.ctypes.data_as(DP), b.v2h.ctypes.data_as(LP),
                                     *[a.ctypes.data_as(DP) for a in (T1, T2e, T2m, Wd1, Vd2, S)])
    assert rc == 0
    assert np.array_equal(T1, d["T1"])
    assert np.array_equal(T2e[..., :nv], d["T2"].transpose(2, 3, 0, 1))  # [h][h'][p][e]
    assert np.array_equal(T2m[..., :no], d["T2"].transpose(2, 0, 1, 3))  # [h][p][p'][m]
    assert np.array_equal(Wd1[..., :no], d["W"].transpose(2, 3, 1, 0))   # [h][h'][p][m]
    assert np.array_equal(Vd2[..., :nv], d["V"].transpose(2, 0, 1, 3))   # [w][p][p'][e]
    assert np.array_equal(S, d["S"].transpose(2, 3, 0, 1))               # [h][h'][p][p']
    # padding is zero
    assert not T2e[..., nv:].any() and not Vd2[..., nv:].any()
    assert not T2m[..., no:].any() and not Wd1[..., no:].any()


def test_upload_pack_reports_missing_block():
    tm = tl.TileModel.synthetic(5, 9, 2, 4, 1)
    b = tl.synth_blocks(tm, 1)
    h = b.t2h.copy()
    L = int(h[0])
    h[1 + L // 2] += 10**9  # corrupt one key
    rc = _lib.lib().tce_t_pack_dense(C.byref(tm.c()), b.t1.ctypes.data_as(DP), b.t1h.ctypes.data_as(LP),
                                     b.t2.ctypes.data_as(DP), h.ctypes.data_as(LP),
                                     b.v2.ctypes.data_as(DP), b.v2h.ctypes.data_as(LP), *[None] * 6)
    assert rc == -4  # TCE_T_ERR_KEY


def test_library_exports_every_header_symbol():
    hdr = open(os.path.join(ROOT, "include", "tce_t.h")).read()
    hdr = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    hdr = "\n".join(ln for ln in hdr.splitlines() if not ln.lstrip().startswith("#"))
    declared = set(re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", hdr))
    declared -= {"if", "sizeof"}
    exported = set(_lib.exported_symbols())
    missing = sorted(declared - exported)
    assert not missing, missing
    # the GPU surface (sd_t_*_cuda_) and the CPU surface (sd_t_*_N_), 27 each
    assert len([s for s in declared if s.startswith("sd_t_")]) == 54


def test_task_list_matches_oracle():
    for no, nv, to, tv in SYSTEMS:
        tm = tl.TileModel.synthetic(no, nv, to, tv, 1)
        ot = oracle.Tiles(tm.range, tm.noab, tm.evl)
        assert len(tm.tasks()) == ot.count_tasks()
    tm = tl.TileModel.synthetic(40, 320, 20, 40, 1)
    assert len(tm.tasks()) == 480


def _job_work(tm, subtile: bool, precision: int = 0):
    """(reference tasks, flops, items) of a job plan (tce_t_job_create needs
    no device)."""
    L = _lib.lib()
    old = os.environ.get("TCE_T_SUBTILE")
    os.environ["TCE_T_SUBTILE"] = "1" if subtile else "0"
    try:
        h = C.c_void_p()
        _lib.check(L.tce_t_job_create(C.byref(tm.c()), C.byref(h)), "tce_t_job_create")
    finally:
        if old is None:
            del os.environ["TCE_T_SUBTILE"]
        else:
            os.environ["TCE_T_SUBTILE"] = old
    f, n = C.c_double(0), C.c_long(0)
    _lib.check(L.tce_t_job_work(h, precision, C.byref(f), C.byref(n)), "tce_t_job_work")
    nt = L.tce_t_job_num_tasks(h)
    L.tce_t_job_destroy(h)
    return nt, f.value, n.value


def _f_ref(tm):
    r = tm.range
    return float(sum(r[a - 1] * r[b - 1] * r[c - 1] * r[i - 1] * r[j - 1] * r[k - 1]
                     for a, b, c, i, j, k in tm.tasks())) * 18 * (tm.no + tm.nv + 1)


@pytest.mark.parametrize("no,nv,to,tv", [(40, 320, 20, 40), (8, 32, 8, 16), (12, 48, 12, 24)])
def test_subtiled_work_is_the_box_tiling(no, nv, to, tv):
    """Sub-tiling a job whose tiles are whole numbers of boxes does exactly
    the work of the reference task loop run on box-sized tiles (4 occupied,
    8 virtual orbitals): same flops, one item per sub-task; the reference
    task count (per-task energies) is unchanged."""
    tm = tl.TileModel.synthetic(no, nv, to, tv, 1)
    fine = tl.TileModel.synthetic(no, nv, 4, 8, 1)
    nt, f, n = _job_work(tm, True)
    nt0, f0, n0 = _job_work(tm, False)
    assert nt == nt0 == len(tm.tasks())
    assert f0 == _f_ref(tm)
    assert f == _f_ref(fine)
    assert n == len(fine.tasks())
    assert f < f0 and n < n0


def test_subtiled_work_ragged_tiles_counts_every_orbital_sextuple_once():
    """Ragged tiles: the sub-tiled work is the reference count of a tiling
    refined to the sub-tiles, which never exceeds the reference tiling's."""
    for ro, rv in [([6, 3], [16, 16, 9]), ([9], [21, 21, 21, 21, 21, 21]), ([5, 5, 5], [13, 13])]:
        tm = tl.TileModel.synthetic(sum(ro), sum(rv), 0, 0, 3, ro, rv)
        refine = lambda rs, e: [min(e, r - k) for r in rs for k in range(0, r, e)]
        fine = tl.TileModel.synthetic(sum(ro), sum(rv), 0, 0, 3, refine(ro, 4), refine(rv, 8))
        nt, f, n = _job_work(tm, True, precision=2)  # 8-wide particle sub-tiles
        assert f == _f_ref(fine) and n == len(fine.tasks())
        assert f <= _f_ref(tm)


def _work_tiling(tm, tiling: int, precision: int = 0):
    L = _lib.lib()
    h = C.c_void_p()
    _lib.check(L.tce_t_job_create_tiling(C.byref(tm.c()), tiling, C.byref(h)), "create_tiling")
    f, n = C.c_double(0), C.c_long(0)
    _lib.check(L.tce_t_job_work(h, precision, C.byref(f), C.byref(n)), "tce_t_job_work")
    assert L.tce_t_job_tiling(h) == tiling
    L.tce_t_job_destroy(h)
    return f.value, n.value


@pytest.mark.parametrize("ro,rv", [([6, 3], [16, 16, 9]), ([9], [21] * 6), ([13, 13, 2], [42, 21])])
def test_orbital_tiling_is_the_box_tiling_of_each_space(ro, rv):
    """TCE_T_TILING_ORBITAL on a C1 spin-orbital tiling (one spin, one irrep):
    the work of the reference task loop on the whole occupied space cut into
    4s and the whole virtual space into 8s, whatever the reference tiles."""
    no, nv = sum(ro), sum(rv)
    tm = tl.TileModel.synthetic(no, nv, 0, 0, 5, ro, rv)
    cut = lambda n, e: [min(e, n - k) for k in range(0, n, e)]
    fine = tl.TileModel.synthetic(no, nv, 0, 0, 5, cut(no, 4), cut(nv, 8))
    f, n = _work_tiling(tm, 2, precision=2)
    assert f == _f_ref(fine) and n == len(fine.tasks())
    f1, n1 = _work_tiling(tm, 1, precision=2)
    assert n <= n1


def test_orbital_tiling_keeps_spin_and_irrep_runs():
    """Spin-labelled tiles: consecutive tiles of one spin and irrep merge
    into one run before the cut, and pieces never straddle a change of spin
    or irrep, so the task filters still apply piece by piece."""
    ranges = [2, 2, 2, 2, 5, 6, 5, 6]       # holes a a b b, particles a a b b
    spin = [1, 1, 2, 2, 1, 1, 2, 2]
    evl = np.sort(np.random.default_rng(1).uniform(-2, 4, sum(ranges)))
    tm = tl.TileModel(range=ranges, noab=4, evl=evl, spin=spin)
    fine = tl.TileModel(range=[4, 4, 8, 3, 8, 3], noab=2, evl=evl, spin=[1, 2, 1, 1, 2, 2])
    f, n = _work_tiling(tm, 2, precision=2)
    assert f == _f_ref(fine) and n == len(fine.tasks())
    irr = tl.TileModel(range=ranges, noab=4, evl=evl, spin=spin, sym=[0, 1, 0, 1, 0, 0, 0, 0])
    fine_irr = tl.TileModel(range=[2, 2, 2, 2, 8, 3, 8, 3], noab=4, evl=evl, spin=[1, 1, 2, 2, 1, 1, 2, 2],
                            sym=[0, 1, 0, 1, 0, 0, 0, 0])
    f, n = _work_tiling(irr, 2, precision=2)
    assert f == _f_ref(fine_irr) and n == len(fine_irr.tasks())


def test_unknown_tiling_is_an_error():
    L = _lib.lib()
    tm = tl.TileModel.synthetic(8, 32, 8, 16, 1)
    h = C.c_void_p()
    assert L.tce_t_job_create_tiling(C.byref(tm.c()), 7, C.byref(h)) == -1


def test_job_plan_without_allowed_tasks():
    """Holes all alpha, particles all beta: empty block tables, no task, no
    work item, in every tiling."""
    evl = np.sort(np.random.default_rng(1).uniform(-2.0, 4.0, 12))
    tm = tl.TileModel(range=[2, 2, 4, 4], noab=2, evl=evl, spin=[1, 1, 2, 2])
    assert tm.sizes() == {"t1": 0, "t2": 0, "v2": 0} and len(tm.tasks()) == 0
    for tiling in (0, 1, 2):
        for precision in (0, 2):
            assert _work_tiling(tm, tiling, precision) == (0.0, 0)


def _kernel_and_items(tm, tiling, kernel=None):
    L = _lib.lib()
    old = os.environ.pop("TCE_T_KERNEL", None)
    if kernel:
        os.environ["TCE_T_KERNEL"] = kernel
    try:
        h = C.c_void_p()
        _lib.check(L.tce_t_job_create_tiling(C.byref(tm.c()), tiling, C.byref(h)), "create")
    finally:
        os.environ.pop("TCE_T_KERNEL", None)
        if old is not None:
            os.environ["TCE_T_KERNEL"] = old
    f, n = C.c_double(0), C.c_long(0)
    L.tce_t_job_work(h, 0, C.byref(f), C.byref(n))
    k = L.tce_t_job_kernel(h)
    L.tce_t_job_destroy(h)
    return k, n.value


@pytest.mark.parametrize("ro,rv", [([8, 8], [16] * 4), ([9], [21] * 6), ([19], [31] * 4 + [1]),
                                   ([6, 4], [19] * 10 + [1]), ([8], [16, 16])])
@pytest.mark.parametrize("tiling", [0, 1, 2])
def test_particle_box_choice_follows_the_padded_work(ro, rv, tiling):
    """8-wide particle boxes (the 16x16x4 kernel) unless their padded work
    (items x 8^3) exceeds 2.2x that of 4-wide boxes (items x 4^3), counted in
    the job's own kernel tasks."""
    tm = tl.TileModel.synthetic(sum(ro), sum(rv), 0, 0, 2, ro, rv)
    _, n4 = _kernel_and_items(tm, tiling, "wave")
    _, n8 = _kernel_and_items(tm, tiling, "wg16")
    k, n = _kernel_and_items(tm, tiling)
    want8 = 512 * n8 <= 2.2 * 64 * n4
    assert (k == 2) == want8 and n == (n8 if want8 else n4)
