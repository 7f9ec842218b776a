"""CR-CCSD(T) on the GPU (tce_t_job_upload_cr, tce_t_cr_ccsd_t) against the
dense textbook of tests/spin_systems.py (written from the term lists of
cr_ccsd_t_N.F, cr_ccsd_t_E.F and cr_ccsd_t_D.F and pinned in
tests/test_cr.py), against the oracle's task loop (oracle.cr_ccsd_t, equal to
the textbook in tests/test_cr.py), and against the library's own (T) energies when the
intermediates are the bare integrals (the moment tile is then the (T)
doubles tile: num1 = E[T], num2 = E(T)).

Tolerance: FP64, |dx| <= 1e-10 |x| + 1e-15 for every sum and energy (the
device sums each task in fixed-order blocks, the textbook over dense
tensors in numpy's order).  Reduced precision, against FP64: the t3, moment
and denominator tiles carry the d1/d2 product error of the mode, so
|dx| <= tol |x| with tol 1e-6 for FP32 and BF16X6 (products to 2^-24) and
3e-5 for BF16X3 (the dropped lo*lo term), as for Lambda-CCSD(T); den0 is
FP64 in every mode.
"""
import os
import sys

import pytest

import oracle
from nwchem_amd import TriplesJob
from nwchem_amd.job import BF16X3, BF16X6, FP32, TILING_ORBITAL, TILING_REFERENCE, cr_ccsd_t

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import spin_systems  # noqa: E402

pytestmark = pytest.mark.gpu

CR_GPU = [
    dict(sizes_o=[3, 2], sizes_v=[4, 3], irreps_o=[0, 1], irreps_v=[1, 0], seed=5),
    dict(sizes_o=[8], sizes_v=[16, 8], irreps_o=[0], irreps_v=[0, 1], seed=9),
    dict(sizes_o=[4, 4], sizes_v=[8, 8], seed=10, flip_symmetric=True, restricted=True),
]
IDS = ["spin_irreps", "spin_8_16", "restricted"]
KEYS = ("num1", "num2", "den0", "den1", "den2")


def _run(s, I, tiling=None):
    b = s.blocks
    print(f"cr: no={s.no} nv={s.nv} tiling={tiling}", flush=True)
    with TriplesJob(s.tiles, tiling=tiling) as tj, TriplesJob(s.tiles, tiling=tiling) as cj:
        tj.upload(b)
        cj.upload_cr(b.t1, b.t1h, b.t2, b.t2h, I["i1_1"], I["i1_1h"], I["i1_2"], I["i1_2h"])
        e1, e2, terms = cr_ccsd_t(tj, cj)
        return e1, e2, terms, tj.run()


def _close(got, want):
    return abs(got - want) <= 1e-10 * abs(want) + 1e-15


@pytest.mark.parametrize("prm", CR_GPU, ids=IDS)
def test_cr_matches_the_dense_textbook(prm):
    s = spin_systems.SpinSystem(**prm)
    I = s.cr_inputs()
    e1, e2, terms, _ = _run(s, I)
    want = s.dense_cr(I["I1"], I["I2"])
    for k in KEYS:
        assert _close(terms[k], want[k]), (k, terms[k], want[k])
    assert _close(e1, want["energy1"]) and _close(e2, want["energy2"])
    b = s.blocks
    o1, o2, ot = oracle.cr_ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h,
                                  I["i1_1"], I["i1_1h"], I["i1_2"], I["i1_2h"])
    for k in KEYS:
        assert _close(terms[k], ot[k]), (k, terms[k], ot[k])
    assert _close(e1, o1) and _close(e2, o2)


@pytest.mark.parametrize("prm", CR_GPU, ids=IDS)
def test_cr_with_bare_integrals_gives_the_triples_energies(prm):
    s = spin_systems.SpinSystem(**prm)
    I = dict(zip(("i1_1", "i1_1h", "i1_2", "i1_2h"), s.i1_blocks(s.W, s.V)))
    _, _, terms, (et, ept) = _run(s, I)
    assert _close(terms["num1"], et) and _close(terms["num2"], ept)


def test_cr_reference_tiling_and_windows_agree(monkeypatch):
    s = spin_systems.SpinSystem(**CR_GPU[1])
    I = s.cr_inputs()
    want = s.dense_cr(I["I1"], I["I2"])
    _, _, terms, _ = _run(s, I, tiling=TILING_REFERENCE)
    for k in KEYS:
        assert _close(terms[k], want[k]), (k, terms[k], want[k])
    monkeypatch.setenv("TCE_T_LAMBDA_WINDOW", "3")
    _, _, terms, _ = _run(s, I)
    for k in KEYS:
        assert _close(terms[k], want[k]), (k, terms[k], want[k])


def test_cr_rejects_wrong_jobs():
    s = spin_systems.SpinSystem(**CR_GPU[0])
    I = s.cr_inputs()
    b = s.blocks
    with TriplesJob(s.tiles) as tj, TriplesJob(s.tiles) as cj:
        tj.upload(b)
        cj.upload_cr(b.t1, b.t1h, b.t2, b.t2h, I["i1_1"], I["i1_1h"], I["i1_2"], I["i1_2h"])
        with pytest.raises(RuntimeError, match="in that order"):
            cr_ccsd_t(cj, tj)
        with pytest.raises(RuntimeError, match="two distinct jobs"):
            cr_ccsd_t(tj, tj)
    with TriplesJob(s.tiles, tiling=TILING_ORBITAL) as tj, \
            TriplesJob(s.tiles, tiling=TILING_ORBITAL) as cj:
        tj.upload(b)
        cj.upload_cr(b.t1, b.t1h, b.t2, b.t2h, I["i1_1"], I["i1_1h"], I["i1_2"], I["i1_2h"])
        with pytest.raises(RuntimeError, match="orbital tiling"):
            cr_ccsd_t(tj, cj)


@pytest.mark.parametrize("mode", [(FP32, 1e-6), (BF16X6, 1e-6), (BF16X3, 3e-5)],
                         ids=["fp32", "bf16x6", "bf16x3"])
def test_cr_precision_modes_against_fp64(mode):
    prec, tol = mode
    s = spin_systems.SpinSystem(**CR_GPU[1])
    I = s.cr_inputs()
    b = s.blocks
    print(f"cr precision {prec}: no={s.no} nv={s.nv}", flush=True)
    with TriplesJob(s.tiles) as tj, TriplesJob(s.tiles) as cj:
        tj.upload(b)
        cj.upload_cr(b.t1, b.t1h, b.t2, b.t2h, I["i1_1"], I["i1_1h"], I["i1_2"], I["i1_2h"])
        f1, f2, ft = cr_ccsd_t(tj, cj)
        m1, m2, mt = cr_ccsd_t(tj, cj, prec)
    assert mt["den0"] == ft["den0"]
    for k in ("num1", "num2", "den1", "den2"):
        assert abs(mt[k] - ft[k]) <= tol * abs(ft[k]), (k, mt[k], ft[k])
    assert abs(m1 - f1) <= tol * abs(f1) and abs(m2 - f2) <= tol * abs(f2)
