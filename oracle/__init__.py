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
        for f in ("ora_t1_offset", "ora_t2_offset"):
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
