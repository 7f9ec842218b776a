"""Pins the CPU oracle (oracle/) before anything is checked against it.

  * the 27 kernels and ccsd_t_dot: bit-exact against the reference's own
    Fortran compiled from the reference tree (oracle/_ref);
  * the tiled dispatch of ccsd_t_gpu.F (row tables, de-duplication, filters,
    hash lookups, sorts and their signs, factors): against an independent
    dense textbook (T) (P(9) antisymmetrisers over every orbital sextuple);
  * planted mistakes: negating any one of the 27 variants moves E far outside
    the tolerance, so the cross-check above would catch it;
  * invariants: tile-size independence, t1 = 0 => E(T) == E[T].
"""
import ctypes as C
import os

import numpy as np
import pytest

import oracle
from nwchem_amd import tiles as tl

REF_BUILT = os.path.exists(os.path.join(oracle.HERE, "_ref", "libsdt_ref.so"))
needs_ref = pytest.mark.skipif(not REF_BUILT, reason="reference kernels not built (no reference tree)")

DP = C.POINTER(C.c_double)
LP = C.POINTER(C.c_long)


def _dp(a):
    return a.ctypes.data_as(DP)


def _ref_kernel(fam, n):
    return getattr(oracle.ref_lib(), f"sd_t_{fam}_{n}_")


SHAPES = [  # (h3d, h2d, h1d, p6d, p5d, p4d, K)
    (3, 2, 4, 5, 3, 2, 3),
    (1, 1, 1, 1, 1, 1, 1),
    (4, 4, 4, 6, 6, 6, 7),
    (2, 5, 3, 4, 7, 1, 5),
]


@needs_ref
@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("fam", ["s1", "d1", "d2"])
def test_kernels_bit_exact_vs_reference_fortran(fam, shape):
    rng = np.random.default_rng(hash((fam, shape)) % 2**32)
    h3d, h2d, h1d, p6d, p5d, p4d, K = shape
    n3 = h3d * h2d * h1d * p6d * p5d * p4d
    L = oracle.lib()
    for n in range(1, 10):
        t3a = rng.standard_normal(n3)
        t3_0 = t3a.copy()
        t3b = t3a.copy()
        if fam == "s1":
            a = rng.standard_normal(p4d * h1d)
            b = rng.standard_normal(h3d * h2d * p6d * p5d)
            L.ora_sd_t_s1(n, h3d, h2d, h1d, p6d, p5d, p4d, _dp(t3a), _dp(a), _dp(b))
            dims = [C.c_long(v) for v in (h3d, h2d, h1d, p6d, p5d, p4d)]
        elif fam == "d1":
            a = rng.standard_normal(K * p4d * p5d * h1d)
            b = rng.standard_normal(h3d * h2d * p6d * K)
            L.ora_sd_t_d1(n, h3d, h2d, h1d, p6d, p5d, p4d, K, _dp(t3a), _dp(a), _dp(b))
            dims = [C.c_long(v) for v in (h3d, h2d, h1d, p6d, p5d, p4d, K)]
        else:
            a = rng.standard_normal(K * p4d * h1d * h2d)
            b = rng.standard_normal(K * h3d * p6d * p5d)
            L.ora_sd_t_d2(n, h3d, h2d, h1d, p6d, p5d, p4d, K, _dp(t3a), _dp(a), _dp(b))
            dims = [C.c_long(v) for v in (h3d, h2d, h1d, p6d, p5d, p4d, K)]
        _ref_kernel(fam, n)(*[C.byref(d) for d in dims], _dp(t3b), _dp(a), _dp(b))
        assert np.array_equal(t3a, t3b), f"{fam}_{n} differs from the reference kernel"
        assert not np.array_equal(t3a, t3_0)  # the kernel did update the tile


@needs_ref
def test_ccsd_t_dot_bit_exact_vs_reference_fortran():
    rng = np.random.default_rng(7)
    R = oracle.ref_lib()
    L = oracle.lib()
    for tiles in [(1, 1, 1, 4, 4, 4), (1, 1, 2, 4, 5, 5), (1, 2, 2, 3, 3, 4)]:
        h1b, h2b, h3b, p4b, p5b, p6b = tiles
        r = [3, 2, 4, 5, 3, 2]  # r_h1 r_h2 r_h3 r_p4 r_p5 r_p6
        n = int(np.prod(r))
        s, d = rng.standard_normal(n), rng.standard_normal(n)
        o = [-1.5 + rng.random(r[0]), -1.2 + rng.random(r[1]), -1.1 + rng.random(r[2]),
             1 + rng.random(r[3]), 1.2 + rng.random(r[4]), 1.4 + rng.random(r[5])]
        for restricted in (0, 1):
            e1a, e2a = C.c_double(0.25), C.c_double(-0.5)
            L.ora_ccsd_t_dot(_dp(s), _dp(d), restricted, h1b, h2b, h3b, p4b, p5b, p6b,
                             *[_dp(x) for x in o], *r, C.byref(e1a), C.byref(e2a))
            e1b, e2b = C.c_double(0.25), C.c_double(-0.5)
            lg = C.c_int32(restricted)  # Fortran LOGICAL (default kind, 4 bytes)
            args = [C.c_long(v) for v in (h1b, h2b, h3b, p4b, p5b, p6b)]
            rr = [C.c_long(v) for v in r]
            R.ccsd_t_dot_(_dp(s), _dp(d), C.byref(lg), *[C.byref(a) for a in args],
                          *[_dp(x) for x in o], *[C.byref(v) for v in rr], C.byref(e1b), C.byref(e2b))
            assert e1a.value == e1b.value and e2a.value == e2b.value


def _system(no, nv, to, tv, seed, ranges_o=None, ranges_v=None):
    tm = tl.TileModel.synthetic(no, nv, to, tv, seed, ranges_o, ranges_v)
    b = tl.synth_blocks(tm, seed)
    ot = oracle.Tiles(tm.range, tm.noab, tm.evl)
    return tm, b, ot


def _brute(no, nv, seed, evl, zero_t1=False):
    d = tl.synth_dense(seed, no, nv)
    if zero_t1:
        d["T1"][:] = 0
    return oracle.brute_force(no, nv, d["T1"], d["T2"], d["W"], d["V"], d["S"], evl[:no], evl[no:])


def _close(a, b, rel=1e-12, abs_=1e-10):
    return abs(a - b) <= max(abs_, rel * abs(b))


@pytest.mark.parametrize("no,nv,to,tv", [(4, 8, 2, 3), (6, 10, 3, 4), (5, 9, 5, 9), (7, 11, 3, 5)])
def test_dispatch_matches_textbook_triples(no, nv, to, tv):
    seed = 1000 + no * nv
    tm, b, ot = _system(no, nv, to, tv, seed)
    e = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    bf = _brute(no, nv, seed, tm.evl)
    assert _close(e[0], bf[0]) and _close(e[1], bf[1]), (e[:2], bf)
    # E(T) - E[T] is the singles term: also pinned, on its own scale
    assert abs((e[1] - e[0]) - (bf[1] - bf[0])) <= 1e-12 * abs(bf[0]) + 1e-13


def test_dispatch_ref_backend_bitwise_equals_c_kernels():
    if not REF_BUILT:
        pytest.skip("reference kernels not built")
    tm, b, ot = _system(8, 32, 8, 16, 1234)
    e_c = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    e_r = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, backend=oracle.ref_backend())
    assert e_c[0] == e_r[0] and e_c[1] == e_r[1]
    assert np.array_equal(e_c[2], e_r[2])


@pytest.mark.parametrize("fam,n", [(f, n) for f in range(3) for n in range(1, 10)])
def test_planted_sign_mistake_is_caught(fam, n):
    no, nv, to, tv, seed = 5, 7, 2, 3, 77
    tm, b, ot = _system(no, nv, to, tv, seed)
    bf = _brute(no, nv, seed, tm.evl)
    oracle.set_mutation(fam, n)
    try:
        e = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    finally:
        oracle.set_mutation(-1, 0)
    # s1 mistakes only move E(T); d1/d2 mistakes move both
    quant = e[1] if fam == 0 else e[0]
    ref = bf[1] if fam == 0 else bf[0]
    assert abs(quant - ref) > 1e3 * max(1e-10, 1e-12 * abs(ref)), (fam, n, quant, ref)


def test_tile_size_invariance():
    no, nv, seed = 6, 12, 4242
    es = []
    for to, tv in [(6, 12), (3, 4), (2, 5), (4, 7)]:
        tm, b, ot = _system(no, nv, to, tv, seed)
        es.append(oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)[:2])
    for e in es[1:]:
        assert _close(e[0], es[0][0]) and _close(e[1], es[0][1])


def test_zero_singles_gives_equal_energies():
    tm, b, ot = _system(6, 10, 3, 4, 99)
    t1 = np.zeros_like(b.t1)
    e = oracle.ccsd_t(ot, t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    assert e[0] == e[1]


def test_task_subset_sums_to_total():
    tm, b, ot = _system(6, 12, 3, 4, 5)
    ntot = ot.count_tasks()
    e, e2, te = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    assert te.shape == (ntot, 2)
    half = ntot // 2
    a = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, task_begin=0, task_end=half)
    c = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, task_begin=half, task_end=-1)
    assert np.array_equal(np.vstack([a[2], c[2]]), te)
    assert _close(a[0] + c[0], e)


def test_sorts_match_numpy_transpose():
    L = oracle.lib()
    rng = np.random.default_rng(3)
    a, b, c, d = 3, 4, 2, 5
    x = rng.standard_normal(a * b * c * d)
    X = x.reshape(a, b, c, d)
    import itertools
    for perm in itertools.permutations((1, 2, 3, 4)):
        out = np.zeros_like(x)
        L.ora_sort_4(_dp(x), _dp(out), a, b, c, d, *perm, -1.0)
        exp = -np.transpose(X, [p - 1 for p in perm]).reshape(-1)
        assert np.array_equal(out, exp)
    y = rng.standard_normal(a * b)
    out = np.zeros_like(y)
    L.ora_sort_2(_dp(y), _dp(out), a, b, 2, 1, 1.0)
    assert np.array_equal(out, y.reshape(a, b).T.reshape(-1))


def test_brute_force_closed_form_single_triple():
    """no = nv = 3: one occupied and one virtual triple, so E = t3^2 / D with
    t3 the (abc,ijk) = (012,012) element.  With only t(0 2 0 1) = t (and its
    antisymmetric images) and <1 2||2 2> = v (and images), the d2 term gives
    t3d = -t v; with t1(0 0) = s and <1 2||1 2> = w the singles give
    t3s = s w; d1 integrals are zero."""
    no = nv = 3
    t, v, s, w = 0.3, -0.7, 0.11, 0.9
    T1 = np.zeros((nv, no)); T1[0, 0] = s
    T2 = np.zeros((nv, nv, no, no))
    T2[0, 2, 0, 1], T2[2, 0, 0, 1], T2[0, 2, 1, 0], T2[2, 0, 1, 0] = t, -t, -t, t
    W = np.zeros((no, nv, no, no))
    V = np.zeros((nv, nv, no, nv)); V[1, 2, 2, 2], V[2, 1, 2, 2] = v, -v
    S = np.zeros((nv, nv, no, no))
    S[1, 2, 1, 2], S[2, 1, 1, 2], S[1, 2, 2, 1], S[2, 1, 2, 1] = w, -w, -w, w
    eo = np.array([-1.3, -0.9, -0.6]); ev = np.array([0.4, 1.1, 2.5])
    D = eo.sum() - ev.sum()
    e4, e5 = oracle.brute_force(no, nv, T1, T2, W, V, S, eo, ev)
    t3d, t3s = -t * v, s * w
    assert abs(e4 - t3d * t3d / D) <= 1e-15
    assert abs(e5 - t3d * (t3d + t3s) / D) <= 1e-15
    # the d1 term on its own: <m c||j k> with m=0: Y = -t(a b i m) v(m c j k)
    T2b = np.zeros_like(T2)
    T2b[0, 1, 0, 0] = 0.0  # diagonal hole pair vanishes by antisymmetry
    T2b[0, 1, 0, 1], T2b[1, 0, 0, 1], T2b[0, 1, 1, 0], T2b[1, 0, 1, 0] = t, -t, -t, t
    Wb = np.zeros_like(W)
    Wb[1, 2, 1, 2], Wb[1, 2, 2, 1] = v, -v  # <m=1 c=2||j=1 k=2>
    e4b, _ = oracle.brute_force(no, nv, np.zeros_like(T1), T2b, Wb, np.zeros_like(V), np.zeros_like(S), eo, ev)
    # t3(012,012): P(c/ab)P(i/jk) of -t(a b i m) v(m c j k); only the identity
    # term survives with i=0, m=1: -t(0 1 0 1) v(1 2 1 2) = -t v
    assert abs(e4b - (t * v) ** 2 / D) <= 1e-15
