"""tce_t_compute_en_noalloc (the energy of compute_en_ without any device
allocation, block partials in the library's fixed scratch array) against
compute_en_ itself, through the C-ABI, on two calls in a row with the second
tile smaller than the first.

Both entry points get the same task tiles: compute_en_ reduces the
library-owned tiles that sd_t_s1_1_cuda_ / sd_t_d2_1_cuda_ filled from host
operands (the per-call surface, as ccsd_t_gpu.F drives it); the no-allocation
call reduces device copies of the same tiles, formed on the host from the
same operands by the kernels' definitions (ccsd_t_kernels_omp.F: s1_1
t3s += t1(p4,h1) v2(h3,h2,p6,p5), d2_1 t3d -= sum_p7 t2(p7,p4,h1,h2)
v2(p7,h3,p6,p5)).  Bar: the FP64 energy tolerance max(1e-10, 1e-12 |E|).
"""
import ctypes as C

import numpy as np
import pytest

from nwchem_amd import _lib

pytestmark = pytest.mark.gpu

DP = C.POINTER(C.c_double)
LP = C.POINTER(C.c_long)


def tol(e):
    return max(1e-10, 1e-12 * abs(e))


def _task(L, dims, factor, seed):
    """Fill the library's task tiles through the per-call surface, run
    compute_en_; return (its energies, host tiles, orbital energies)."""
    h1d, h2d, h3d, p4d, p5d, p6d = dims
    p7d = 7
    rng = np.random.default_rng(seed)
    t1 = rng.standard_normal(p4d * h1d) * 0.1           # Fortran t1(p4,h1)
    v2s = rng.standard_normal(h3d * h2d * p6d * p5d)    # v2(h3,h2,p6,p5)
    t2 = rng.standard_normal(p7d * p4d * h1d * h2d) * 0.1  # t2(p7,p4,h1,h2)
    v2d = rng.standard_normal(p7d * h3d * p6d * p5d)    # v2(p7,h3,p6,p5)
    ev = [np.sort(rng.uniform(-2.0, -0.3, n)) for n in (h1d, h2d, h3d)] + \
         [np.sort(rng.uniform(0.2, 4.0, n)) for n in (p4d, p5d, p6d)]
    d = [C.c_long(x) for x in dims]
    k7 = C.c_long(p7d)
    L.initmemmodule_()
    L.dev_mem_d_(*[C.byref(x) for x in d])
    L.dev_mem_s_(*[C.byref(x) for x in d])
    dummy = np.zeros(1)
    L.sd_t_s1_1_cuda_(C.byref(d[0]), C.byref(d[1]), C.byref(d[2]), C.byref(d[3]), C.byref(d[4]),
                      C.byref(d[5]), dummy.ctypes.data_as(DP), t1.ctypes.data_as(DP),
                      v2s.ctypes.data_as(DP))
    L.sd_t_d2_1_cuda_(C.byref(d[0]), C.byref(d[1]), C.byref(d[2]), C.byref(d[3]), C.byref(d[4]),
                      C.byref(d[5]), C.byref(k7), dummy.ctypes.data_as(DP),
                      t2.ctypes.data_as(DP), v2d.ctypes.data_as(DP))
    energy = np.zeros(2)
    f = C.c_double(factor)
    L.compute_en_(C.byref(f), energy.ctypes.data_as(DP), *[e.ctypes.data_as(DP) for e in ev],
                  *[C.byref(x) for x in d], None, None)
    L.dev_release_()
    assert L.tce_t_legacy_status() == 0, L.tce_t_last_error()
    # the same tiles on the host, C order [p4][p5][p6][h1][h2][h3]
    T1 = t1.reshape(h1d, p4d)                 # [h1][p4]
    VS = v2s.reshape(p5d, p6d, h2d, h3d)      # [p5][p6][h2][h3]
    T2 = t2.reshape(h2d, h1d, p4d, p7d)       # [h2][h1][p4][p7]
    VD = v2d.reshape(p5d, p6d, h3d, p7d)      # [p5][p6][h3][p7]
    t3s = np.einsum("ia,bcjk->abcijk", T1, VS)
    t3d = -np.einsum("jiae,bcke->abcijk", T2, VD)
    return energy, np.ascontiguousarray(t3d), np.ascontiguousarray(t3s), ev


def _hip():
    """The HIP runtime libtce_t.so itself uses (its loaded dependency): the
    test allocates the device inputs and outputs with it.  (torch bundles a
    HIP runtime of its own; two runtimes in one process do not mix.)"""
    H = C.CDLL("libamdhip64.so.7", mode=C.RTLD_GLOBAL)
    H.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    H.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    H.hipFree.argtypes = [C.c_void_p]
    H.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    H.hipStreamSynchronize.argtypes = [C.c_void_p]
    H.hipStreamDestroy.argtypes = [C.c_void_p]
    return H


H2D, D2H = 1, 2


def _noalloc(L, dims, factor, t3d, t3s, ev):
    H = _hip()
    bufs = []

    def upload(a):
        p = C.c_void_p()
        assert H.hipMalloc(C.byref(p), max(8, a.nbytes)) == 0
        assert H.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        bufs.append(p)
        return p

    try:
        g_d, g_s = upload(t3d), upload(t3s)
        g_ev = [upload(e) for e in ev]
        out = upload(np.full(2, np.nan))
        st = C.c_void_p()
        assert H.hipStreamCreate(C.byref(st)) == 0
        rc = L.tce_t_compute_en_noalloc(C.c_double(factor), g_d, g_s, *g_ev,
                                        *[C.c_long(x) for x in dims], out, st)
        assert rc == 0, L.tce_t_last_error()
        assert H.hipStreamSynchronize(st) == 0
        H.hipStreamDestroy(st)
        res = np.zeros(2)
        assert H.hipMemcpy(res.ctypes.data, out, 16, D2H) == 0
        return res
    finally:
        for p in bufs:
            H.hipFree(p)


@pytest.fixture(scope="module")
def L():
    L = _lib.lib()
    L.tce_t_compute_en_noalloc.argtypes = [C.c_double] + [C.c_void_p] * 8 + [C.c_long] * 6 + \
        [C.c_void_p, C.c_void_p]
    L.tce_t_compute_en_noalloc.restype = C.c_int
    return L


def test_noalloc_energy_matches_compute_en_on_two_calls(L):
    # (h1d, h2d, h3d, p4d, p5d, p6d): the second call is smaller than the first
    for dims, factor, seed in (((5, 6, 7, 9, 10, 11), 0.5, 1), ((2, 3, 4, 3, 5, 4), 1.0, 2)):
        print(f"noalloc case dims={dims} factor={factor}", flush=True)
        ref, t3d, t3s, ev = _task(L, dims, factor, seed)
        got = _noalloc(L, dims, factor, t3d, t3s, ev)
        assert abs(got[0] - ref[0]) <= tol(ref[0]), (got, ref)
        assert abs(got[1] - ref[1]) <= tol(ref[1]), (got, ref)
        # the scratch array keeps the last call's partials: a repeat of the
        # same call gives the same bits
        again = _noalloc(L, dims, factor, t3d, t3s, ev)
        assert np.array_equal(again, got)


def test_noalloc_rejects_bad_arguments(L):
    dims = (2, 2, 2, 2, 2, 2)
    x = C.c_void_p(1)
    bad = [C.c_long(v) for v in dims[:5]] + [C.c_long(0)]
    assert L.tce_t_compute_en_noalloc(C.c_double(1.0), *([x] * 8), *bad, x, None) == -1
    assert L.tce_t_compute_en_noalloc(C.c_double(1.0), None, *([x] * 7),
                                      *[C.c_long(v) for v in dims], x, None) == -1
