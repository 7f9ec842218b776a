"""CPU oracle of the TCE (T) path -- TEST INFRASTRUCTURE ONLY.

Loaded by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline
leg, never by the product package.  Wraps

  liboracle.so            C restatement of the reference path (sdt_oracle.c)
  _ref/libsdt_ref.so      the reference's own Fortran kernels + ccsd_t_dot,
                          compiled from the reference tree (strict FP, serial)
  _ref/libsdt_ref_omp.so  the same sources, -O3 -fopenmp (CPU baseline)

See sdt_oracle.h for the reference lines each function restates and for what
pins each of them.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LP = C.POINTER(C.c_long)
DP = C.POINTER(C.c_double)

K6 = C.CFUNCTYPE(None, LP, LP, LP, LP, LP, LP, DP, DP, DP)
K7 = C.CFUNCTYPE(None, LP, LP, LP, LP, LP, LP, LP, DP, DP, DP)
V0 = C.CFUNCTYPE(None)
M6 = C.CFUNCTYPE(None, LP, LP, LP, LP, LP, LP)
EN = C.CFUNCTYPE(None, DP, DP, DP, DP, DP, DP, DP, DP, LP, LP, LP, LP, LP, LP, DP, DP)


class OraTiles(C.Structure):
    _fields_ = [("noab", C.c_long), ("nvab", C.c_long), ("spin", LP), ("sym", LP), ("range", LP),
                ("offset", LP), ("alpha", LP), ("evl", DP), ("restricted", C.c_int)]


class OraIntorb(C.Structure):
    _fields_ = [("noa", C.c_long), ("nva", C.c_long), ("b2am", LP), ("spin_alpha", LP),
                ("sym_alpha", LP), ("range_alpha", LP), ("irrep_v", C.c_long)]


class OraBackend(C.Structure):
    _fields_ = [("usedevice", C.c_int), ("s1", K6 * 9), ("d1", K7 * 9), ("d2", K7 * 9),
                ("initmemmodule", V0), ("finalizememmodule", V0), ("dev_mem_s", M6),
                ("dev_mem_d", M6), ("dev_release", V0), ("compute_en", EN)]


_libs: dict = {}


def build(ref: bool = True) -> None:
    """Build liboracle.so (and the reference kernels when the reference tree
    is present)."""
    subprocess.run(["make", "-C", HERE, "oracle"] + (["ref"] if ref else []), check=True)


def _load(name: str) -> C.CDLL:
    if name in _libs:
        return _libs[name]
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build(ref=name.startswith("_ref"))
    if not os.path.exists(path):
        raise FileNotFoundError(path)
    L = C.CDLL(path, mode=C.RTLD_LOCAL)
    _libs[name] = L
    return L


def lib() -> C.CDLL:
    L = _load("liboracle.so")
    if not getattr(L, "_declared", False):
        L.ora_tile_ranges.argtypes = [C.c_long, C.c_long, LP]
        L.ora_tile_ranges.restype = C.c_long
        for f in ("ora_t1_offset", "ora_t2_offset", "ora_y1_offset", "ora_y2_offset",
                  "ora_f1_offset"):
            getattr(L, f).argtypes = [C.POINTER(OraTiles), LP, LP]
            getattr(L, f).restype = C.c_long
        L.ora_v2_offset.argtypes = [C.POINTER(OraTiles), C.c_int, LP, LP]
        L.ora_v2_offset.restype = C.c_long
        L.ora_hash.argtypes = [LP, C.c_long]
        L.ora_hash.restype = C.c_long
        L.ora_sort_2.argtypes = [DP, DP, C.c_long, C.c_long, C.c_long, C.c_long, C.c_double]
        L.ora_sort_4.argtypes = [DP, DP] + [C.c_long] * 8 + [C.c_double]
        L.ora_sd_t_s1.argtypes = [C.c_int] + [C.c_long] * 6 + [DP, DP, DP]
        L.ora_sd_t_d1.argtypes = [C.c_int] + [C.c_long] * 7 + [DP, DP, DP]
        L.ora_sd_t_d2.argtypes = [C.c_int] + [C.c_long] * 7 + [DP, DP, DP]
        L.ora_ccsd_t_dot.argtypes = [DP, DP, C.c_int] + [C.c_long] * 6 + [DP] * 6 + [C.c_long] * 6 + [DP, DP]
        L.ora_count_tasks.argtypes = [C.POINTER(OraTiles)]
        L.ora_count_tasks.restype = C.c_long
        L.ora_ccsd_t_gpu.argtypes = [C.POINTER(OraTiles), DP, LP, DP, LP, DP, LP, C.POINTER(OraBackend),
                                     C.c_long, C.c_long, DP, DP, DP]
        L.ora_ccsd_t_gpu.restype = C.c_long
        L.ora_set_mutation.argtypes = [C.c_int, C.c_int]
        L.ora_set_mutation.restype = None
        L.ora_set_arith.argtypes = [C.c_int]
        L.ora_set_arith.restype = None
        L.ora_split_bf16.argtypes = [DP, C.c_long, C.c_int, C.POINTER(C.c_ushort), C.POINTER(C.c_ushort)]
        L.ora_split_bf16.restype = None
        L.ora_split3_bf16.argtypes = [DP, C.c_long] + [C.POINTER(C.c_ushort)] * 3
        L.ora_split3_bf16.restype = None
        L.ora_brute_force.argtypes = [C.c_long, C.c_long] + [DP] * 7 + [DP, DP]
        L.ora_task_tiles.argtypes = [C.POINTER(OraTiles), DP, LP, DP, LP, DP, LP, LP, DP, DP]
        L.ora_task_tiles.restype = C.c_long
        L.ora_v2_alpha_offset.argtypes = [C.POINTER(OraIntorb), C.c_long, LP, LP]
        L.ora_v2_alpha_offset.restype = C.c_long
        L.ora_hash_v2.argtypes = [C.POINTER(OraIntorb), LP, C.c_long]
        L.ora_hash_v2.restype = C.c_long
        L.ora_v2orb_pack.argtypes = [C.POINTER(OraIntorb), DP, C.c_long, DP]
        L.ora_v2orb_pack.restype = None
        L.ora_v2_from_intorb.argtypes = [C.POINTER(OraTiles), C.POINTER(OraIntorb), DP, LP, LP, DP]
        L.ora_v2_from_intorb.restype = C.c_int
        L.ora_sortacc_6.argtypes = [DP, DP] + [C.c_long] * 12 + [C.c_double]
        L.ora_sortacc_6.restype = None
        L._declared = True
    return L


def ref_lib(omp: bool | str = False) -> C.CDLL:
    """The reference's Fortran kernels (oracle/_ref): omp=False the serial
    strict build, True the -O3 -fopenmp build, "par" the strict-FP OpenMP
    build."""
    if omp == "par":
        return _load("_ref/libsdt_ref_par.so")
    return _load("_ref/libsdt_ref_omp.so" if omp else "_ref/libsdt_ref.so")


def _lp(a):
    return a.ctypes.data_as(LP)


def _dp(a):
    return a.ctypes.data_as(DP)


class Tiles:
    """Owns the arrays behind an OraTiles view."""

    def __init__(self, range_, noab, evl, spin=None, sym=None, alpha=None, restricted=False):
        self.range = np.ascontiguousarray(range_, np.int64)
        n = len(self.range)
        self.noab, self.nvab = int(noab), n - int(noab)
        self.offset = np.concatenate([[0], np.cumsum(self.range)[:-1]]).astype(np.int64)
        self.spin = np.ones(n, np.int64) if spin is None else np.ascontiguousarray(spin, np.int64)
        self.sym = np.zeros(n, np.int64) if sym is None else np.ascontiguousarray(sym, np.int64)
        self.alpha = np.arange(1, n + 1, dtype=np.int64) if alpha is None else np.ascontiguousarray(alpha, np.int64)
        self.evl = np.ascontiguousarray(evl, np.float64)
        self.restricted = bool(restricted)
        self.c = OraTiles(self.noab, self.nvab, _lp(self.spin), _lp(self.sym), _lp(self.range),
                          _lp(self.offset), _lp(self.alpha), _dp(self.evl), int(self.restricted))

    def hash(self, kind: str, subset: int = 1) -> tuple[np.ndarray, int]:
        L = lib()
        size = C.c_long(0)
        if kind == "v2":
            n = L.ora_v2_offset(C.byref(self.c), subset, None, C.byref(size))
            h = np.zeros(2 * n + 1, np.int64)
            L.ora_v2_offset(C.byref(self.c), subset, _lp(h), C.byref(size))
        else:
            f = getattr(L, f"ora_{kind}_offset")
            n = f(C.byref(self.c), None, C.byref(size))
            h = np.zeros(2 * n + 1, np.int64)
            f(C.byref(self.c), _lp(h), C.byref(size))
        return h, int(size.value)

    def count_tasks(self) -> int:
        return int(lib().ora_count_tasks(C.byref(self.c)))


def ref_backend(omp: bool | str = False) -> OraBackend:
    """Kernel back-end made of the reference's compiled Fortran kernels
    (CPU argument order, usedevice = 0); omp: as ref_lib."""
    R = ref_lib(omp)
    be = OraBackend()
    be.usedevice = 0
    for n in range(9):
        be.s1[n] = K6(("sd_t_s1_%d_" % (n + 1), R))
        be.d1[n] = K7(("sd_t_d1_%d_" % (n + 1), R))
        be.d2[n] = K7(("sd_t_d2_%d_" % (n + 1), R))
    be._keep = R
    return be


def device_backend(L: C.CDLL) -> OraBackend:
    """Back-end that drives a library exporting the reference's GPU surface
    (sd_t_*_cuda_, dev_mem_*, compute_en_, ...) the way ccsd_t_gpu.F does
    with usedevice = 1.  The library is the thing under test."""
    be = OraBackend()
    be.usedevice = 1
    for n in range(9):
        be.s1[n] = K6(("sd_t_s1_%d_cuda_" % (n + 1), L))
        be.d1[n] = K7(("sd_t_d1_%d_cuda_" % (n + 1), L))
        be.d2[n] = K7(("sd_t_d2_%d_cuda_" % (n + 1), L))
    be.initmemmodule = V0(("initmemmodule_", L))
    be.finalizememmodule = V0(("finalizememmodule_", L))
    be.dev_mem_s = M6(("dev_mem_s_", L))
    be.dev_mem_d = M6(("dev_mem_d_", L))
    be.dev_release = V0(("dev_release_", L))
    be.compute_en = EN(("compute_en_", L))
    be._keep = L
    return be


def ccsd_t(tiles: Tiles, t1, t1h, t2, t2h, v2, v2h, backend: OraBackend | None = None,
           task_begin: int = 0, task_end: int = -1):
    """ccsd_t_gpu.F task loop.  Returns (E[T], E(T), per-task energies)."""
    L = lib()
    ntot = tiles.count_tasks()
    nsel = (ntot if task_end < 0 else min(task_end, ntot)) - task_begin
    te = np.zeros((max(nsel, 1), 2))
    e1, e2 = C.c_double(0), C.c_double(0)
    rc = L.ora_ccsd_t_gpu(C.byref(tiles.c), _dp(t1), _lp(t1h), _dp(t2), _lp(t2h), _dp(v2), _lp(v2h),
                          C.byref(backend) if backend is not None else None, task_begin, task_end,
                          C.byref(e1), C.byref(e2), _dp(te))
    if rc < 0:
        raise RuntimeError(f"ora_ccsd_t_gpu failed: {rc}")
    return e1.value, e2.value, te[:rc]


def task_tiles(tiles: Tiles, t1, t1h, t2, t2h, v2, v2h, task):
    """(t3s, t3d) of the task with tiles task = (p4b, p5b, p6b, h1b, h2b, h3b),
    1-based, as the reference's singles and doubles drivers leave them:
    arrays of shape (r4, r5, r6, r1, r2, r3) (triplesx(h3,h2,h1,p6,p5,p4))."""
    L = lib()
    tk = np.ascontiguousarray(task, np.int64)
    shape = tuple(int(tiles.range[b - 1]) for b in tk)
    s = np.zeros(shape)
    d = np.zeros(shape)
    rc = L.ora_task_tiles(C.byref(tiles.c), _dp(t1), _lp(t1h), _dp(t2), _lp(t2h), _dp(v2), _lp(v2h),
                          _lp(tk), _dp(s), _dp(d))
    if rc < 0:
        raise ValueError(f"ora_task_tiles({list(tk)}) failed: {rc}")
    return s, d


def brute_force(no, nv, T1, T2, W, V, S, eo, ev):
    L = lib()
    e4, e5 = C.c_double(0), C.c_double(0)
    arrs = [np.ascontiguousarray(a, np.float64) for a in (T1, T2, W, V, S, eo, ev)]
    L.ora_brute_force(no, nv, *[_dp(a) for a in arrs], C.byref(e4), C.byref(e5))
    return e4.value, e5.value


ARITH = {"fp64": 0, "fp32": 1, "bf16x3": 2, "abs": 3, "bf16x2_planted": 4, "bf16x3_trunc_planted": 5,
         "bf16x6": 6}


class arith:
    """Context manager: the built-in d1/d2 kernels in another arithmetic
    (see sdt_oracle.h, ora_set_arith); FP64 again on exit."""

    def __init__(self, mode: str):
        self.mode = ARITH[mode]

    def __enter__(self):
        lib().ora_set_arith(self.mode)
        return self

    def __exit__(self, *a):
        lib().ora_set_arith(0)


def split_bf16(x, truncate_lo: bool = False):
    """The library's operand split of FP64 values: (hi, lo) as uint16 bf16
    bit patterns (hi = bf16(fl32(x)), lo = bf16(fl32(x - hi)), round to
    nearest even; truncate_lo: the planted truncation)."""
    x = np.ascontiguousarray(x, np.float64).ravel()
    hi = np.zeros(x.size, np.uint16)
    lo = np.zeros(x.size, np.uint16)
    lib().ora_split_bf16(_dp(x), x.size, int(truncate_lo), hi.ctypes.data_as(C.POINTER(C.c_ushort)),
                         lo.ctypes.data_as(C.POINTER(C.c_ushort)))
    return hi, lo


def split3_bf16(x):
    """The library's three-part split (hi, mid, lo as uint16 bf16 patterns)."""
    x = np.ascontiguousarray(x, np.float64).ravel()
    parts = [np.zeros(x.size, np.uint16) for _ in range(3)]
    lib().ora_split3_bf16(_dp(x), x.size, *[p.ctypes.data_as(C.POINTER(C.c_ushort)) for p in parts])
    return tuple(parts)


def set_mutation(family: int, n: int) -> None:
    """Plant (family, n >= 1) or remove (family = -1) a sign mistake in one
    built-in kernel variant (test of the checks' sensitivity)."""
    lib().ora_set_mutation(family, n)


class Intorb:
    """Owns the arrays behind an OraIntorb view (orbital-form integrals:
    alpha tiles, k_b2am, irrep of V)."""

    def __init__(self, noa, nva, b2am, spin_alpha, sym_alpha, range_alpha, irrep_v=0):
        self.noa, self.nva, self.irrep_v = int(noa), int(nva), int(irrep_v)
        self.b2am = np.ascontiguousarray(b2am, np.int64)
        self.spin_alpha = np.ascontiguousarray(spin_alpha, np.int64)
        self.sym_alpha = np.ascontiguousarray(sym_alpha, np.int64)
        self.range_alpha = np.ascontiguousarray(range_alpha, np.int64)
        self.c = OraIntorb(self.noa, self.nva, _lp(self.b2am), _lp(self.spin_alpha),
                           _lp(self.sym_alpha), _lp(self.range_alpha), self.irrep_v)

    def alpha_offset(self, idiv2e: int = 2):
        """tce_mo2e_offset_intorb: (k_v2_alpha_offset, orbital-form size)."""
        L = lib()
        size = C.c_long(0)
        n = L.ora_v2_alpha_offset(C.byref(self.c), idiv2e, None, C.byref(size))
        if n < 0:
            raise ValueError("idiv2e leaves no segment")
        h = np.zeros(n, np.int64)
        L.ora_v2_alpha_offset(C.byref(self.c), idiv2e, _lp(h), C.byref(size))
        return h, int(size.value)

    def hash_v2(self, hash_, key: int) -> int:
        return int(lib().ora_hash_v2(C.byref(self.c), _lp(hash_), key))

    def pack(self, orb, size: int):
        """tce_mo2e_trans's orbital-form array from dense (pq|rs)."""
        orb = np.ascontiguousarray(orb, np.float64)
        out = np.zeros(max(size, 1))
        lib().ora_v2orb_pack(C.byref(self.c), _dp(orb), orb.shape[0], _dp(out))
        return out


def v2_from_intorb(tiles: Tiles, io: Intorb, v2orb, alpha_offset, v2h, size: int):
    """get_hash_block_i over every block of v2h (ora_v2_from_intorb)."""
    out = np.zeros(max(size, 1))
    v2orb = np.ascontiguousarray(v2orb, np.float64)
    rc = lib().ora_v2_from_intorb(C.byref(tiles.c), C.byref(io.c), _dp(v2orb), _lp(alpha_offset),
                                  _lp(v2h), _dp(out))
    if rc != 0:
        raise KeyError("orbital-form key not found")
    return out


# ---- Lambda-CCSD(T) (lambda_ccsd_t.F) -----------------------------------

def _table(h):
    n = int(h[0])
    return [(int(h[1 + i]), int(h[1 + n + i])) for i in range(n)]


def lambda_blocks(tiles: Tiles, y1, y1h, y2, y2h):
    """The Lambda amplitudes placed as t1/t2 blocks of the same tiles: y(h p)
    blocks [h][p] (tce_y1_offset.F) -> [p][h] under the t1 key, y(h h' p p')
    blocks [h][h'][p][p'] (tce_y2_offset.F) -> [p][p'][h][h'] under the t2
    key.  Returns (t1, t1h, t2, t2h)."""
    L = lib()
    noab, nvab = tiles.noab, tiles.nvab
    r = tiles.range
    t1h, s1 = tiles.hash("t1")
    t2h, s2 = tiles.hash("t2")
    y1 = np.ascontiguousarray(y1, np.float64)
    y2 = np.ascontiguousarray(y2, np.float64)
    y1h = np.ascontiguousarray(y1h, np.int64)
    y2h = np.ascontiguousarray(y2h, np.int64)
    t1 = np.zeros(max(s1, 1))
    t2 = np.zeros(max(s2, 1))
    for key, off in _table(t1h):
        h, p = key % noab + 1, key // noab + noab + 1
        yo = L.ora_hash(_lp(y1h), p - noab - 1 + nvab * (h - 1))
        if yo < 0:
            raise ValueError("y1 lacks a block of the t1 table")
        rh, rp = int(r[h - 1]), int(r[p - 1])
        t1[off:off + rh * rp] = y1[yo:yo + rh * rp].reshape(rh, rp).T.ravel()
    for key, off in _table(t2h):
        h4 = key % noab + 1
        q = key // noab
        h3 = q % noab + 1
        q //= noab
        p2, p1 = q % nvab + noab + 1, q // nvab + noab + 1
        yo = L.ora_hash(_lp(y2h), p2 - noab - 1 + nvab * (p1 - noab - 1 + nvab * (h4 - 1 + noab * (h3 - 1))))
        if yo < 0:
            raise ValueError("y2 lacks a block of the t2 table")
        r1, r2, r3, r4 = (int(r[b - 1]) for b in (p1, p2, h3, h4))
        n = r1 * r2 * r3 * r4
        t2[off:off + n] = y2[yo:yo + n].reshape(r3, r4, r1, r2).transpose(2, 3, 0, 1).ravel()
    return t1, t1h, t2, t2h


def fock_hole_particle(tiles: Tiles, f1, f1h):
    """F[h][p] over every orbital from the hole-particle blocks of the
    one-electron operator (tce_mo1e_offset.F blocks [g6][g5]; the f(h6 p3)
    reads of lambda_ccsd_t_left.F:397-400, all-beta pairs from their alpha
    image in a restricted run, tce_restricted.F:1-32)."""
    L = lib()
    noab = tiles.noab
    N = noab + tiles.nvab
    r, o = tiles.range, tiles.offset
    no = int(r[:noab].sum())
    nv = int(r[noab:].sum())
    F = np.zeros((no, nv))
    f1 = np.ascontiguousarray(f1, np.float64)
    f1h = np.ascontiguousarray(f1h, np.int64)
    for hb in range(1, noab + 1):
        for pb in range(noab + 1, N + 1):
            if tiles.spin[hb - 1] != tiles.spin[pb - 1] or (tiles.sym[hb - 1] ^ tiles.sym[pb - 1]):
                continue
            h, p = hb, pb
            if tiles.restricted and tiles.spin[hb - 1] + tiles.spin[pb - 1] == 4:
                h, p = int(tiles.alpha[hb - 1]), int(tiles.alpha[pb - 1])
            off = L.ora_hash(_lp(f1h), p - 1 + N * (h - 1))
            if off < 0:
                raise ValueError("f1 lacks a hole-particle block")
            rh, rp = int(r[hb - 1]), int(r[pb - 1])
            oh, op = int(o[hb - 1]), int(o[pb - 1]) - no
            F[oh:oh + rh, op:op + rp] = f1[off:off + rh * rp].reshape(rh, rp)
    return F


def _dense_y2(tiles: Tiles, t2, t2h):
    """Y[a, b, i, j] over every orbital from t2-layout blocks (canonical
    p1 <= p2, h3 <= h4; the other orderings by antisymmetry, all-beta blocks
    from their alpha image in a restricted run)."""
    L = lib()
    noab, nvab = tiles.noab, tiles.nvab
    r, o = tiles.range, tiles.offset
    no = int(r[:noab].sum())
    nv = int(r[noab:].sum())
    Y = np.zeros((nv, nv, no, no))
    sp, sy = tiles.spin, tiles.sym

    def sl(b):
        s = int(o[b - 1]) - (no if b > noab else 0)
        return slice(s, s + int(r[b - 1]))

    for p1 in range(noab + 1, noab + nvab + 1):
        for p2 in range(p1, noab + nvab + 1):
            for h3 in range(1, noab + 1):
                for h4 in range(h3, noab + 1):
                    if sp[p1 - 1] + sp[p2 - 1] != sp[h3 - 1] + sp[h4 - 1]:
                        continue
                    if sy[p1 - 1] ^ sy[p2 - 1] ^ sy[h3 - 1] ^ sy[h4 - 1]:
                        continue
                    q = (p1, p2, h3, h4)
                    if tiles.restricted and sum(int(sp[b - 1]) for b in q) == 8:
                        q = tuple(int(tiles.alpha[b - 1]) for b in q)
                    key = q[3] - 1 + noab * (q[2] - 1 + noab * (q[1] - noab - 1 + nvab * (q[0] - noab - 1)))
                    off = L.ora_hash(_lp(t2h), key)
                    if off < 0:
                        raise ValueError("t2 lacks a block")
                    shp = tuple(int(r[b - 1]) for b in (p1, p2, h3, h4))
                    blk = t2[off:off + int(np.prod(shp))].reshape(shp)
                    Y[sl(p1), sl(p2), sl(h3), sl(h4)] = blk
                    Y[sl(p2), sl(p1), sl(h3), sl(h4)] = -blk.transpose(1, 0, 2, 3)
                    Y[sl(p1), sl(p2), sl(h4), sl(h3)] = -blk.transpose(0, 1, 3, 2)
                    Y[sl(p2), sl(p1), sl(h4), sl(h3)] = blk.transpose(1, 0, 3, 2)
    return Y


def _task_list(tiles: Tiles):
    """Tasks of the (T) loop in order (lambda_ccsd_t.F:57-84, the loop and
    tests of ccsd_t_gpu.F)."""
    noab, nvab = tiles.noab, tiles.nvab
    sp, sy = tiles.spin, tiles.sym
    out = []
    for p4 in range(noab + 1, noab + nvab + 1):
        for p5 in range(p4, noab + nvab + 1):
            for p6 in range(p5, noab + nvab + 1):
                for h1 in range(1, noab + 1):
                    for h2 in range(h1, noab + 1):
                        for h3 in range(h2, noab + 1):
                            q = (p4, p5, p6, h1, h2, h3)
                            if sp[p4 - 1] + sp[p5 - 1] + sp[p6 - 1] != sp[h1 - 1] + sp[h2 - 1] + sp[h3 - 1]:
                                continue
                            if tiles.restricted and sum(int(sp[b - 1]) for b in q) > 8:
                                continue
                            if sy[p4 - 1] ^ sy[p5 - 1] ^ sy[p6 - 1] ^ sy[h1 - 1] ^ sy[h2 - 1] ^ sy[h3 - 1]:
                                continue
                            out.append(q)
    return out


def lambda_ccsd_t(tiles: Tiles, t1, t1h, t2, t2h, y1, y1h, y2, y2h, v2, v2h, f1=None, f1h=None,
                  pairing: int = 0):
    """lambda_ccsd_t.F:56-200: (energy1, energy2, per-task energies).  The
    t3 doubles tile (ccsd_t_doubles) and the left tiles (lambda_ccsd_t_left:
    the y1 and v2 terms as ora_task_tiles over the Lambda amplitudes placed
    as t1/t2, the Fock term P(c/ab) P(k/ij) y(a b i j) f(k c) in numpy),
    then factor * t * y / D summed per task.  pairing 0: elements of the
    same orbitals; 1: element i of the t3 order [p4][p5][p6][h1][h2][h3]
    with element i of the left order [h1][h2][h3][p4][p5][p6]
    (lambda_ccsd_t.F:156-178)."""
    lt1, lt1h, lt2, lt2h = lambda_blocks(tiles, y1, y1h, y2, y2h)
    F = None if f1 is None else fock_hole_particle(tiles, f1, f1h)
    Y = _dense_y2(tiles, lt2, lt2h) if F is not None else None
    noab = tiles.noab
    r, o = tiles.range, tiles.offset
    no = int(r[:noab].sum())
    ev = tiles.evl

    def orb(b):
        s = int(o[b - 1]) - (no if b > noab else 0)
        return np.arange(s, s + int(r[b - 1]))

    tasks = _task_list(tiles)
    te = np.zeros((len(tasks), 2))
    for n, q in enumerate(tasks):
        p4, p5, p6, h1, h2, h3 = q
        _, td = task_tiles(tiles, t1, t1h, t2, t2h, v2, v2h, q)
        ys, yd = task_tiles(tiles, lt1, lt1h, lt2, lt2h, v2, v2h, q)
        A, B, Cc, I, J, K = (orb(b) for b in q)
        if F is not None:
            def yy(p, pp, h, hh):
                return Y[np.ix_(p, pp, h, hh)]

            def ff(h, p):
                return F[np.ix_(h, p)]

            e = np.einsum
            yf = (e("abij,kc->abcijk", yy(A, B, I, J), ff(K, Cc))
                  - e("cbij,ka->abcijk", yy(Cc, B, I, J), ff(K, A))
                  - e("acij,kb->abcijk", yy(A, Cc, I, J), ff(K, B))
                  - e("abkj,ic->abcijk", yy(A, B, K, J), ff(I, Cc))
                  + e("cbkj,ia->abcijk", yy(Cc, B, K, J), ff(I, A))
                  + e("ackj,ib->abcijk", yy(A, Cc, K, J), ff(I, B))
                  - e("abik,jc->abcijk", yy(A, B, I, K), ff(J, Cc))
                  + e("cbik,ja->abcijk", yy(Cc, B, I, K), ff(J, A))
                  + e("acik,jb->abcijk", yy(A, Cc, I, K), ff(J, B)))
            yd = yd + yf
        if pairing == 1:
            shp = yd.shape
            yd = yd.transpose(3, 4, 5, 0, 1, 2).reshape(shp)
            ys = ys.transpose(3, 4, 5, 0, 1, 2).reshape(shp)
        D = (-ev[no + A][:, None, None, None, None, None] - ev[no + B][None, :, None, None, None, None]
             - ev[no + Cc][None, None, :, None, None, None] + ev[I][None, None, None, :, None, None]
             + ev[J][None, None, None, None, :, None] + ev[K][None, None, None, None, None, :])
        f = 2.0 if tiles.restricted else 1.0
        if p4 == p5 == p6:
            f /= 6.0
        elif p4 == p5 or p5 == p6:
            f /= 2.0
        if h1 == h2 == h3:
            f /= 6.0
        elif h1 == h2 or h2 == h3:
            f /= 2.0
        te[n, 0] = f * np.sum(td * yd / D)
        te[n, 1] = f * np.sum(td * (ys + yd) / D)
    return float(te[:, 0].sum()), float(te[:, 1].sum()), te


# ---- CR-CCSD(T) -------------------------------------------------------------

def _dense_t1(tiles: Tiles, t1, t1h):
    """T[a, i] over every orbital from t1-layout blocks ([p][h]; all-beta
    blocks from their alpha image in a restricted run)."""
    L = lib()
    noab, nvab = tiles.noab, tiles.nvab
    r, o = tiles.range, tiles.offset
    no = int(r[:noab].sum())
    nv = int(r[noab:].sum())
    T = np.zeros((nv, no))
    sp, sy = tiles.spin, tiles.sym

    def sl(b):
        s = int(o[b - 1]) - (no if b > noab else 0)
        return slice(s, s + int(r[b - 1]))

    for p in range(noab + 1, noab + nvab + 1):
        for h in range(1, noab + 1):
            if sp[p - 1] != sp[h - 1] or sy[p - 1] ^ sy[h - 1]:
                continue
            q = (p, h)
            if tiles.restricted and sp[p - 1] + sp[h - 1] == 4:
                q = (int(tiles.alpha[p - 1]), int(tiles.alpha[h - 1]))
            off = L.ora_hash(_lp(t1h), q[1] - 1 + noab * (q[0] - noab - 1))
            if off < 0:
                raise ValueError("t1 lacks a block")
            n = int(r[p - 1] * r[h - 1])
            T[sl(p), sl(h)] = t1[off:off + n].reshape(int(r[p - 1]), int(r[h - 1]))
    return T


def cr_operands(tiles: Tiles, t1, t1h, t2, t2h, v2h, i1_1, i1_1h, i1_2, i1_2h):
    """The operands of cr_ccsd_t.F's extra tiles as a v2-layout array over the
    table v2h, so that ccsd_t_singles / ccsd_t_doubles over it give them:
      <h p||h h> block (h11, p4, h1, h2): i1(h11 p4 h1 h2) of cr_ccsd_t_N
        (table OFFSET_cr_ccsd_t_N_1_1, block [p4][h11][h1][h2] after the sorts
        of cr_ccsd_t_N.F:739-752);
      <p p||h p> block (p4, p5, h1, p12): i1(p4 p5 h1 p12) (OFFSET_cr_ccsd_t_N_2_1,
        block [p4][p5][h1][p12], cr_ccsd_t_N.F:3990-3996);
      <p p||h h> block: S = t(p q h g) + (t(p h) t(q g) - t(p g) t(q h)) / 3,
        the denominator of cr_ccsd_t_E.F:7-9 (P(9) t t2 - 2/3 P(9) t i1,
        i1 = -1/4 P(4) t t) with its single t moved to the front.
    Doubles over the array are the moment tile (cr_ccsd_t_N.F:8,26),
    singles the denominator tile."""
    L = lib()
    noab, nvab = tiles.noab, tiles.nvab
    N = noab + nvab
    r, o = tiles.range, tiles.offset
    no = int(r[:noab].sum())
    T1 = _dense_t1(tiles, t1, t1h)
    T2 = _dense_y2(tiles, t2, t2h)
    S = T2 + (np.einsum("ph,qg->pqhg", T1, T1) - np.einsum("pg,qh->pqhg", T1, T1)) / 3.0

    def sl(b):
        s = int(o[b - 1]) - (no if b > noab else 0)
        return slice(s, s + int(r[b - 1]))

    n = int(v2h[0])
    ent = [(int(v2h[1 + k]), int(v2h[1 + n + k])) for k in range(n)]
    size = max([off + int(np.prod([r[b - 1] for b in _vtiles(key, N)])) for key, off in ent] + [1])
    out = np.zeros(size)
    for key, off in ent:
        g3, g4, g1, g2 = _vtiles(key, N)
        shp = tuple(int(r[b - 1]) for b in (g3, g4, g1, g2))
        m = int(np.prod(shp))
        if g3 <= noab < g4 and g1 <= noab and g2 <= noab:
            k1 = g2 - 1 + noab * (g1 - 1 + noab * (g3 - 1 + noab * (g4 - noab - 1)))
            s1 = L.ora_hash(_lp(i1_1h), k1)
            if s1 < 0:
                raise ValueError("i1(h11 p4 h1 h2) lacks a block")
            blk = i1_1[s1:s1 + m].reshape(shp[1], shp[0], shp[2], shp[3]).transpose(1, 0, 2, 3)
        elif g3 > noab and g4 > noab and g1 <= noab < g2:
            k2 = g2 - noab - 1 + nvab * (g1 - 1 + noab * (g4 - noab - 1 + nvab * (g3 - noab - 1)))
            s2 = L.ora_hash(_lp(i1_2h), k2)
            if s2 < 0:
                raise ValueError("i1(p4 p5 h1 p12) lacks a block")
            blk = i1_2[s2:s2 + m].reshape(shp)
        elif g3 > noab and g4 > noab and g1 <= noab and g2 <= noab:
            blk = S[sl(g3), sl(g4), sl(g1), sl(g2)]
        else:
            continue
        out[off:off + m] = blk.ravel()
    return out


def _vtiles(key: int, N: int):
    """(g3, g4, g1, g2) of a tce_mo2e_offset key g2-1 + N (g1-1 + N (g4-1 + N (g3-1)))."""
    g2 = key % N + 1
    key //= N
    g1 = key % N + 1
    key //= N
    return key // N + 1, key % N + 1, g1, g2


def cr_ccsd_t(tiles: Tiles, t1, t1h, t2, t2h, v2, v2h, i1_1, i1_1h, i1_2, i1_2h):
    """cr_ccsd_t.F:70-258 after its intermediates exist: per task of the (T)
    loop the (T) tiles ts, td (ccsd_t_singles_l / ccsd_t_doubles_l), the
    moment m and the denominator tile q (cr_operands), w = factor / D, and
      num1 += w m td, num2 += w m (ts + td), den1 += w q td, den2 += w q (ts + td);
    den0 = sum t(a i)^2 + 1/4 sum t(a b i j)^2 + 1/2 sum t(a i) t(b j) t(a b i j)
    (cr_ccsd_t_D.F:4-7 with c = t); energy_n = num_n / (1 + den_n + den0).
    Returns (energy1, energy2, dict of the five terms)."""
    vc = cr_operands(tiles, t1, t1h, t2, t2h, v2h, i1_1, i1_1h, i1_2, i1_2h)
    noab = tiles.noab
    r, o = tiles.range, tiles.offset
    no = int(r[:noab].sum())
    ev = tiles.evl

    def orb(b):
        s = int(o[b - 1]) - (no if b > noab else 0)
        return np.arange(s, s + int(r[b - 1]))

    num1 = num2 = den1 = den2 = 0.0
    for q in _task_list(tiles):
        p4, p5, p6, h1, h2, h3 = q
        ts, td = task_tiles(tiles, t1, t1h, t2, t2h, v2, v2h, q)
        qd, m = task_tiles(tiles, t1, t1h, t2, t2h, vc, v2h, q)
        A, B, Cc, I, J, K = (orb(b) for b in q)
        D = (-ev[no + A][:, None, None, None, None, None] - ev[no + B][None, :, None, None, None, None]
             - ev[no + Cc][None, None, :, None, None, None] + ev[I][None, None, None, :, None, None]
             + ev[J][None, None, None, None, :, None] + ev[K][None, None, None, None, None, :])
        f = 2.0 if tiles.restricted else 1.0
        if p4 == p5 == p6:
            f /= 6.0
        elif p4 == p5 or p5 == p6:
            f /= 2.0
        if h1 == h2 == h3:
            f /= 6.0
        elif h1 == h2 or h2 == h3:
            f /= 2.0
        num1 += f * np.sum(m * td / D)
        num2 += f * np.sum(m * (ts + td) / D)
        den1 += f * np.sum(qd * td / D)
        den2 += f * np.sum(qd * (ts + td) / D)
    T1 = _dense_t1(tiles, t1, t1h)
    T2 = _dense_y2(tiles, t2, t2h)
    den0 = (float(np.sum(T1 * T1)) + 0.25 * float(np.sum(T2 * T2))
            + 0.5 * float(np.einsum("ai,bj,abij->", T1, T1, T2)))
    terms = dict(num1=float(num1), num2=float(num2), den0=den0, den1=float(den1), den2=float(den2))
    return (terms["num1"] / (1.0 + terms["den1"] + den0),
            terms["num2"] / (1.0 + terms["den2"] + den0), terms)
