"""Lambda-CCSD(T) (lambda_ccsd_t.F) on the CPU: the oracle's block tables,
sort and energy composition against brute force and the dense textbook,
and the library's host-side tables and argument checks.

The textbook (SpinSystem.dense_lambda) is written from the term list of
lambda_ccsd_t_left.F:7-10 with the integrals in the orientation written
there; the oracle composes the left tiles as the (T) tiles of a job whose
amplitudes are the Lambda amplitudes (the orientation the (T) path reads,
equal for real orbitals) plus the Fock term.  Agreement pins the
composition, the y-block placement and the Fock block read.
"""
from __future__ import annotations

import ctypes as C
import itertools

import numpy as np
import pytest

import oracle
from nwchem_amd import _lib
from spin_systems import SpinSystem

SPIN_CASES = [
    dict(sizes_o=[3, 2], sizes_v=[4, 3], irreps_o=[0, 1], irreps_v=[1, 0], seed=5),
    dict(sizes_o=[2, 2], sizes_v=[3, 2, 2], irreps_o=[0, 3], irreps_v=[1, 2, 3], seed=6),
    dict(sizes_o=[4], sizes_v=[5, 4], seed=8),
]


def _oracle(s, L, pairing=0, fock=True):
    b = s.blocks
    return oracle.lambda_ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, L["y1"], L["y1h"],
                                L["y2"], L["y2h"], b.v2, b.v2h,
                                L["f1"] if fock else None, L["f1h"] if fock else None, pairing)


def _close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(b), 1e-300)


# ---- block tables ----------------------------------------------------------

def _enumerate(tiles, kind):
    """Blocks of tce_y1_offset.F / tce_y2_offset.F / tce_mo1e_offset.F by
    brute force: (key, size) in storage order."""
    noab, nvab = tiles.noab, len(tiles.range) - tiles.noab
    N = noab + nvab
    sp, sy, r = tiles.spin, tiles.sym, tiles.range
    out = []
    if kind == "y1":
        for h in range(1, noab + 1):
            for p in range(noab + 1, N + 1):
                if sp[h - 1] == sp[p - 1] and not (sy[h - 1] ^ sy[p - 1]) and \
                        not (tiles.restricted and sp[h - 1] + sp[p - 1] == 4):
                    out.append((p - noab - 1 + nvab * (h - 1), r[h - 1] * r[p - 1]))
    elif kind == "y2":
        for h7, h8, p1, p2 in itertools.product(range(1, noab + 1), range(1, noab + 1),
                                                range(noab + 1, N + 1), range(noab + 1, N + 1)):
            if h8 < h7 or p2 < p1:
                continue
            q = (h7, h8, p1, p2)
            if sp[h7 - 1] + sp[h8 - 1] != sp[p1 - 1] + sp[p2 - 1]:
                continue
            if sy[h7 - 1] ^ sy[h8 - 1] ^ sy[p1 - 1] ^ sy[p2 - 1]:
                continue
            if tiles.restricted and sum(sp[b - 1] for b in q) == 8:
                continue
            key = p2 - noab - 1 + nvab * (p1 - noab - 1 + nvab * (h8 - 1 + noab * (h7 - 1)))
            out.append((key, int(np.prod([r[b - 1] for b in q]))))
    else:
        for g6, g5 in itertools.product(range(1, N + 1), range(1, N + 1)):
            if sp[g6 - 1] == sp[g5 - 1] and not (sy[g6 - 1] ^ sy[g5 - 1]) and \
                    not (tiles.restricted and sp[g6 - 1] + sp[g5 - 1] == 4):
                out.append((g5 - 1 + N * (g6 - 1), r[g6 - 1] * r[g5 - 1]))
    return out


@pytest.mark.parametrize("restricted", [False, True])
@pytest.mark.parametrize("prm", SPIN_CASES)
@pytest.mark.parametrize("kind", ["y1", "y2", "f1"])
def test_lambda_tables_match_enumeration(prm, restricted, kind):
    s = SpinSystem(**prm, flip_symmetric=restricted, restricted=restricted)
    blocks = _enumerate(s.tiles, kind)
    want = np.zeros(2 * len(blocks) + 1, np.int64)
    want[0] = len(blocks)
    off = 0
    for n, (key, size) in enumerate(blocks):
        want[1 + n] = key
        want[1 + len(blocks) + n] = off
        off += size
    ora, ora_size = s.oracle_tiles().hash(kind)
    lib = getattr(s.tiles, f"{kind}_hash")()
    np.testing.assert_array_equal(ora, want)
    np.testing.assert_array_equal(lib, want)
    assert ora_size == off
    keys = want[1:1 + len(blocks)]
    assert np.all(np.diff(keys) > 0)  # tce_hash bisects: keys ascend


# ---- sorts and the two element orders ----------------------------------------

def _sortacc_6(x, dims, perm, factor=1.0, into=None):
    out = np.zeros(x.size) if into is None else into
    xs = np.ascontiguousarray(x, np.float64).ravel()
    oracle.lib().ora_sortacc_6(oracle._dp(xs), oracle._dp(out), *[int(d) for d in dims],
                               *[int(p) for p in perm], float(factor))
    return out


def test_sortacc_6_is_a_transpose():
    """ora_sortacc_6 (sort/tce_sortacc.F:126-280) against numpy.transpose for
    every permutation of six axes of distinct extents, with factor and
    accumulation."""
    dims = (2, 3, 1, 4, 2, 3)
    x = np.arange(1, int(np.prod(dims)) + 1, dtype=np.float64).reshape(dims)
    for perm in itertools.permutations(range(1, 7)):
        got = _sortacc_6(x, dims, perm)
        want = np.transpose(x, [p - 1 for p in perm]).ravel()
        np.testing.assert_array_equal(got, want)
    perm = (6, 4, 3, 5, 2, 1)
    base = np.full(x.size, 0.5)
    got = _sortacc_6(x, dims, perm, factor=-2.0, into=base.copy())
    np.testing.assert_array_equal(got, base - 2.0 * np.transpose(x, [5, 3, 2, 4, 1, 0]).ravel())


def test_reference_tile_orders():
    """The orders that the two element pairings name, from the sorts as the
    reference writes them: ccsd_t_doubles.F:209-211 (c_sort [h3][h2][p6][h1]
    [p5][p4], sort 6,5,3,4,2,1) leaves [p4][p5][p6][h1][h2][h3];
    lambda_ccsd_t_left.F:192-194 (c_sort [p3][p2][h6][h5][p1][h4], sort
    6,4,3,5,2,1; the left tile's h4..h6, p1..p3 are the task's h1..h3,
    p4..p6) leaves [h1][h2][h3][p4][p5][p6]."""
    r = dict(p4=2, p5=3, p6=4, h1=1, h2=2, h3=3)

    def tagged(names):
        grids = np.meshgrid(*[np.arange(r[n]) for n in names], indexing="ij")
        return sum(g * 10 ** (5 - "p4 p5 p6 h1 h2 h3".split().index(n))
                   for g, n in zip(grids, names)).astype(np.float64)

    t_sort = ["h3", "h2", "p6", "h1", "p5", "p4"]
    got = _sortacc_6(tagged(t_sort), [r[n] for n in t_sort], (6, 5, 3, 4, 2, 1))
    np.testing.assert_array_equal(got, tagged(["p4", "p5", "p6", "h1", "h2", "h3"]).ravel())
    y_sort = ["p6", "p5", "h3", "h2", "p4", "h1"]
    got = _sortacc_6(tagged(y_sort), [r[n] for n in y_sort], (6, 4, 3, 5, 2, 1))
    np.testing.assert_array_equal(got, tagged(["h1", "h2", "h3", "p4", "p5", "p6"]).ravel())


# ---- energies ----------------------------------------------------------------

@pytest.mark.parametrize("prm", SPIN_CASES)
def test_oracle_matches_textbook(prm):
    """Both energies, with and without the Fock term, equal the dense
    textbook over every spin-orbital sextuple."""
    s = SpinSystem(**prm)
    L = s.lambda_inputs()
    for fock in (True, False):
        e1, e2, te = _oracle(s, L, fock=fock)
        d1, d2 = s.dense_lambda(L, fock=fock)
        assert te.shape[0] > 0
        assert _close(e1, d1) and _close(e2, d2), (fock, (e1, e2), (d1, d2))


@pytest.mark.parametrize("prm", SPIN_CASES)
def test_restricted_matches_unrestricted_and_textbook(prm):
    """Spin-flip symmetric data: the restricted run (all-beta blocks of y,
    f and v from their alpha images, beta-majority tasks skipped, factor 2)
    gives the unrestricted energies and the textbook's."""
    u = SpinSystem(**prm, flip_symmetric=True)
    r = SpinSystem(**prm, flip_symmetric=True, restricted=True)
    Lu, Lr = u.lambda_inputs(), r.lambda_inputs()
    np.testing.assert_array_equal(Lu["Y2"], Lr["Y2"])  # same stream, same data
    eu, er = _oracle(u, Lu), _oracle(r, Lr)
    d = u.dense_lambda(Lu)
    assert er[2].shape[0] < eu[2].shape[0]
    assert Lr["y2"].size < Lu["y2"].size and Lr["f1"].size < Lu["f1"].size
    for k in (0, 1):
        assert _close(er[k], eu[k]) and _close(er[k], d[k]), (er[:2], eu[:2], d)


def test_planted_mistakes_are_caught():
    """A sign flip of the Fock term, a dropped Fock term or a dropped y1
    term each moves the energies far beyond the tolerance."""
    s = SpinSystem(**SPIN_CASES[0])
    L = s.lambda_inputs()
    e1, e2, _ = _oracle(s, L)
    for bad in (s.dense_lambda(L, y3f_sign=-1.0), s.dense_lambda(L, fock=False)):
        assert not _close(e1, bad[0], rel=1e-6) and not _close(e2, bad[1], rel=1e-6)
    assert not _close(e2, e1, rel=1e-6)  # the y1 term contributes to energy2 only


@pytest.mark.parametrize("restricted", [False, True])
@pytest.mark.parametrize("prm", SPIN_CASES)
def test_lambda_with_y_equal_t_is_the_triples_energy(prm, restricted):
    """With the amplitudes as Lambda amplitudes and no Fock term the left
    tiles are the right ones: energy1 = E[T], energy2 = E(T), task by task,
    against the (T) task loop of the oracle."""
    s = SpinSystem(**prm, flip_symmetric=restricted, restricted=restricted)
    b = s.blocks
    y1, y1h, y2, y2h = s.y_equal_t()
    lam = oracle.lambda_ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, y1, y1h, y2, y2h, b.v2,
                               b.v2h)
    ref = oracle.ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    assert _close(lam[0], ref[0]) and _close(lam[1], ref[1]), (lam[:2], ref[:2])
    np.testing.assert_allclose(lam[2], ref[2], rtol=1e-12, atol=1e-300)


def test_pairings_agree_on_one_orbital_tiles_and_differ_otherwise():
    """With one orbital per tile the two element orders coincide, so the
    reference-order pairing equals the orbital pairing; with larger tiles it
    does not."""
    one = SpinSystem(sizes_o=[1, 1], sizes_v=[1, 1, 1], irreps_o=[0, 1], irreps_v=[1, 0, 1],
                     seed=11)
    L = one.lambda_inputs()
    a, b = _oracle(one, L, 0), _oracle(one, L, 1)
    assert a[2].shape[0] > 0
    np.testing.assert_allclose(b[2], a[2], rtol=1e-14, atol=1e-300)
    s = SpinSystem(**SPIN_CASES[2])
    L = s.lambda_inputs()
    a, b = _oracle(s, L, 0), _oracle(s, L, 1)
    assert not _close(b[0], a[0], rel=1e-3)


# ---- library surface without a GPU ---------------------------------------------

def test_lambda_entry_points_reject_bad_arguments():
    """Argument checks that run before any device call."""
    L = _lib.lib()
    e = C.c_double(0.0)
    assert L.tce_t_lambda_ccsd_t(None, None, None, None, 0, 0, C.byref(e), C.byref(e), None,
                                 0) == -1
    assert L.tce_t_job_upload_lambda(None, None, None, None, None, None, None) == -1
