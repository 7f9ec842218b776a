"""CR-CCSD(T) test infrastructure on the CPU: the block tables of the two
moment intermediates against the library's own integral table, and the
dense textbook (written from the term lists of cr_ccsd_t_N.F, cr_ccsd_t_E.F
and cr_ccsd_t_D.F) against what the mathematics fixes:
  - with the intermediates set to the bare integrals (i1 = v, the first
    term of each) the moment tile is the (T) doubles tile, so num1 and num2
    are E[T] and E(T) of the dense (T);
  - the denominator tile is the triples projection of exp(T1 + T2) beyond
    the first order, P(a/bc) P(i/jk) t(a i) t(b c j k) + det[t] over the
    3 x 3 block of t1 (T1 T2 + T1^3 / 6);
  - a planted sign mistake in the denominator's second term fails that.
"""
import itertools
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import spin_systems  # noqa: E402

SYSTEMS = [
    dict(sizes_o=[2, 1], sizes_v=[2, 2], irreps_o=[0, 1], irreps_v=[1, 0], seed=3),
    dict(sizes_o=[2], sizes_v=[3, 2], seed=4, flip_symmetric=True, restricted=True),
]
IDS = ["spin_irreps", "restricted"]


def _table_keys(h):
    L = int(h[0])
    return [int(k) for k in h[1:1 + L]]


@pytest.mark.parametrize("prm", SYSTEMS, ids=IDS)
def test_i1_tables_hold_the_blocks_of_the_integral_table(prm):
    """The i1(h11 p4 h1 h2) and i1(p4 p5 h1 p12) tables list, in key order,
    exactly the tile tuples whose <h p||h h> and <p p||h p> blocks the
    library's tce_mo2e_offset table holds."""
    s = spin_systems.SpinSystem(**prm)
    T = s.tiles
    noab, nvab = T.noab, T.nvab
    N = noab + nvab
    (h1, e1, _), (h2, e2, _) = s.i1_tables()
    k1, k2 = _table_keys(h1), _table_keys(h2)
    assert k1 == sorted(k1) and k2 == sorted(k2)
    v2 = set(_table_keys(T.v2_hash()))

    def vkey(a, b, c, d):
        return d - 1 + N * (c - 1 + N * (b - 1 + N * (a - 1)))

    want1 = {(p4, h11, a, b) for h11 in range(1, noab + 1) for p4 in range(noab + 1, N + 1)
             for a in range(1, noab + 1) for b in range(a, noab + 1) if vkey(h11, p4, a, b) in v2}
    want2 = {(p4, p5, a, p12) for p4 in range(noab + 1, N + 1) for p5 in range(p4, N + 1)
             for a in range(1, noab + 1) for p12 in range(noab + 1, N + 1)
             if vkey(p4, p5, a, p12) in v2}
    assert {t for _, _, t in e1} == want1
    assert {t for _, _, t in e2} == want2


@pytest.mark.parametrize("prm", SYSTEMS, ids=IDS)
def test_moment_with_bare_integrals_is_the_triples_energy(prm):
    s = spin_systems.SpinSystem(**prm)
    cr = s.dense_cr(s.W, s.V)
    ts, td = s.dense_t3()
    eo, ev = s.eo, s.ev
    D = (-ev[:, None, None, None, None, None] - ev[None, :, None, None, None, None]
         - ev[None, None, :, None, None, None] + eo[None, None, None, :, None, None]
         + eo[None, None, None, None, :, None] + eo[None, None, None, None, None, :])
    e_t = float(np.sum(td * td / D)) / 36.0
    e_pt = float(np.sum(td * (ts + td) / D)) / 36.0
    assert cr["num1"] == pytest.approx(e_t, rel=1e-12)
    assert cr["num2"] == pytest.approx(e_pt, rel=1e-12)
    # the brute-force (T) of the oracle agrees as well
    bf = s.brute_force()
    assert e_t == pytest.approx(bf[0], rel=1e-10)


def _q_closed_form(s):
    """P(a/bc) P(i/jk) t(a i) t(b c j k) + det of t1 over (a b c) x (i j k)."""
    t1, t2 = s.T1, s.T2
    nv, no = t1.shape
    q = np.zeros((nv,) * 3 + (no,) * 3)
    for a, b, c in itertools.product(range(nv), repeat=3):
        for i, j, k in itertools.product(range(no), repeat=3):
            m = np.array([[t1[a, i], t1[a, j], t1[a, k]], [t1[b, i], t1[b, j], t1[b, k]],
                          [t1[c, i], t1[c, j], t1[c, k]]])
            x = np.linalg.det(m)
            for (p, pp, ppp, sp) in ((a, b, c, 1), (b, a, c, -1), (c, b, a, -1)):
                for (h, hh, hhh, sh) in ((i, j, k, 1), (j, i, k, -1), (k, j, i, -1)):
                    x += sp * sh * t1[p, h] * t2[pp, ppp, hh, hhh]
            q[a, b, c, i, j, k] = x
    return q


def _q_of(s, den_sign):
    """The denominator tile as dense_cr forms it from the term list of
    cr_ccsd_t_E.F:7-9; den_sign flips its second term."""
    t1, t2 = s.T1, s.T2
    tt = np.einsum("bj,ck->bcjk", t1, t1)
    x = -0.25 * (tt - tt.transpose(1, 0, 2, 3) - tt.transpose(0, 1, 3, 2)
                 + tt.transpose(1, 0, 3, 2))

    def tr(y, pa, ph):
        return np.transpose(y, tuple(pa) + tuple(3 + q for q in ph))

    id3 = ((0, 1, 2), 1.0)
    p_a = [id3, ((1, 0, 2), -1.0), ((2, 1, 0), -1.0)]
    p_c = [id3, ((2, 1, 0), -1.0), ((0, 2, 1), -1.0)]
    h_k = [id3, ((2, 1, 0), -1.0), ((0, 2, 1), -1.0)]
    h_i = [id3, ((1, 0, 2), -1.0), ((2, 1, 0), -1.0)]

    def P(y, pp, hh):
        return sum(sa * sh * tr(y, pa, ph) for pa, sa in pp for ph, sh in hh)

    q = P(np.einsum("abij,ck->abcijk", t2, t1), p_c, h_k)
    return q + den_sign * (-2.0 / 3.0) * P(np.einsum("ai,bcjk->abcijk", t1, x), p_a, h_i)


def test_denominator_tile_is_the_exp_t_projection():
    s = spin_systems.SpinSystem(**SYSTEMS[0])
    want = _q_closed_form(s)
    assert np.allclose(_q_of(s, 1.0), want, rtol=1e-12, atol=1e-16)
    # the sums of dense_cr use the same tile: den1 against a direct sum
    ts, td = s.dense_t3()
    eo, ev = s.eo, s.ev
    D = (-ev[:, None, None, None, None, None] - ev[None, :, None, None, None, None]
         - ev[None, None, :, None, None, None] + eo[None, None, None, :, None, None]
         + eo[None, None, None, None, :, None] + eo[None, None, None, None, None, :])
    cr = s.dense_cr(s.W, s.V)
    assert cr["den1"] == pytest.approx(float(np.sum(want * td / D)) / 36.0, rel=1e-11)
    assert cr["den2"] == pytest.approx(float(np.sum(want * (ts + td) / D)) / 36.0, rel=1e-11)


def test_planted_denominator_sign_mistake_is_caught():
    s = spin_systems.SpinSystem(**SYSTEMS[0])
    want = _q_closed_form(s)
    assert not np.allclose(_q_of(s, -1.0), want, rtol=1e-6, atol=1e-12)
    assert s.dense_cr(s.W, s.V, den_sign=-1.0)["den1"] != pytest.approx(
        s.dense_cr(s.W, s.V)["den1"], rel=1e-6)


def test_overlap_closed_form():
    """den0 = <(1 + T1 + T2)^+ exp(T)> through second order: with t1 = 0 it
    is 1/4 sum t2^2, with t2 = 0 it is sum t1^2."""
    s = spin_systems.SpinSystem(**SYSTEMS[0])
    t1, t2 = s.T1.copy(), s.T2.copy()
    s.T1 = np.zeros_like(t1)
    assert s.dense_cr(s.W, s.V)["den0"] == pytest.approx(0.25 * float(np.sum(t2 * t2)), rel=1e-14)
    s.T1, s.T2 = t1, np.zeros_like(t2)
    assert s.dense_cr(s.W, s.V)["den0"] == pytest.approx(float(np.sum(t1 * t1)), rel=1e-14)


@pytest.mark.parametrize("prm", SYSTEMS, ids=IDS)
def test_oracle_task_loop_equals_the_textbook(prm):
    """oracle.cr_ccsd_t (the task loop of cr_ccsd_t.F over per-task tiles,
    the intermediates in their own tables) against the dense textbook."""
    import oracle

    s = spin_systems.SpinSystem(**prm)
    I = s.cr_inputs()
    b = s.blocks
    e1, e2, terms = oracle.cr_ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h,
                                     I["i1_1"], I["i1_1h"], I["i1_2"], I["i1_2h"])
    want = s.dense_cr(I["I1"], I["I2"])
    for k, v in terms.items():
        assert v == pytest.approx(want[k], rel=1e-12), k
    assert e1 == pytest.approx(want["energy1"], rel=1e-12)
    assert e2 == pytest.approx(want["energy2"], rel=1e-12)


def test_planted_i1_block_order_mistake_is_caught():
    """i1(h11 p4 h1 h2) laid out [h11][p4][h1][h2] (the v2 order) instead of
    [p4][h11][h1][h2]: the moment sums leave the textbook."""
    import oracle

    s = spin_systems.SpinSystem(**SYSTEMS[0])
    I = s.cr_inputs()
    b = s.blocks
    (_, e1, _), _ = s.i1_tables()
    T = s.tiles
    no = s.no

    def sl(t):
        o = int(T.offset[t - 1]) - (no if t > T.noab else 0)
        return slice(o, o + int(T.range[t - 1]))

    bad = I["i1_1"].copy()
    for _, off, (p4, h11, h1, h2) in e1:
        blk = I["I1"][sl(h11), sl(p4), sl(h1), sl(h2)].ravel()
        bad[off:off + blk.size] = blk
    _, _, terms = oracle.cr_ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h,
                                   bad, I["i1_1h"], I["i1_2"], I["i1_2h"])
    want = s.dense_cr(I["I1"], I["I2"])
    assert terms["num1"] != pytest.approx(want["num1"], rel=1e-6)
