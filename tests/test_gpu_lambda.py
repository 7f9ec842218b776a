"""Lambda-CCSD(T) on the GPU (tce_t_job_upload_lambda, tce_t_lambda_ccsd_t)
against the oracle's composition (pinned against the dense textbook in
tests/test_lambda.py) on the same spin-orbital systems.

Tolerance: FP64, |dE| <= 1e-11 |E| + 1e-15 for both energies and every
task's pair (the device sums each task in fixed-order blocks, the oracle in
numpy's order).  Reduced precision, against FP64: the t3 and y3 tiles
carry the d1/d2 product error of the mode, so |dE| <= tol |E| with tol
1e-6 for FP32 and BF16X6 (products to 2^-24) and 3e-5 for BF16X3 (the
dropped lo*lo term, 2^-16 of each product).
"""
import os
import sys

import numpy as np
import pytest

import oracle
from nwchem_amd import TriplesJob
from nwchem_amd.job import (BF16X3, BF16X6, FP32, LAMBDA_ORBITAL, LAMBDA_REFERENCE_ORDER,
                            lambda_ccsd_t)

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import spin_systems  # noqa: E402

pytestmark = pytest.mark.gpu

LAMBDA_GPU = [
    dict(sizes_o=[3, 2], sizes_v=[4, 3], irreps_o=[0, 1], irreps_v=[1, 0], seed=5),
    dict(sizes_o=[8], sizes_v=[16, 8], irreps_o=[0], irreps_v=[0, 1], seed=9),
    dict(sizes_o=[4, 4], sizes_v=[8, 8], seed=10, flip_symmetric=True, restricted=True),
]
IDS = ["spin_irreps", "spin_8_16", "restricted"]


def _run(s, L, pairing, precision=0, fock=True):
    b = s.blocks
    with TriplesJob(s.tiles) as right, TriplesJob(s.tiles) as left:
        right.upload(b)
        left.upload_lambda(L["y1"], L["y1h"], L["y2"], L["y2h"], b.v2, b.v2h)
        return lambda_ccsd_t(right, left, L["f1"] if fock else None, L["f1h"] if fock else None,
                             pairing, precision)


# orbital pairing: the fused passes over the job's kernel tasks (default,
# in one window or in windows of 3 items), or whole tiles per task
PATHS = {"fused": {}, "fused_windows": {"TCE_T_LAMBDA_WINDOW": "3"},
         "tiles": {"TCE_T_LAMBDA": "tiles"}}


@pytest.mark.parametrize("case", [(LAMBDA_ORBITAL, "fused"), (LAMBDA_ORBITAL, "fused_windows"),
                                  (LAMBDA_ORBITAL, "tiles"), (LAMBDA_REFERENCE_ORDER, "tiles")],
                         ids=["orbital_fused", "orbital_fused_windows", "orbital_tiles",
                              "reference_order"])
@pytest.mark.parametrize("prm", LAMBDA_GPU, ids=IDS)
def test_lambda_matches_oracle(prm, case, monkeypatch):
    pairing, path = case
    for k, v in PATHS[path].items():
        monkeypatch.setenv(k, v)
    s = spin_systems.SpinSystem(**prm)
    L = s.lambda_inputs()
    b = s.blocks
    want = oracle.lambda_ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, L["y1"], L["y1h"],
                                L["y2"], L["y2h"], b.v2, b.v2h, L["f1"], L["f1h"], pairing)
    print(f"lambda case {prm} pairing={pairing} path={path} tasks={len(want[2])}", flush=True)
    e1, e2, te = _run(s, L, pairing)
    np.testing.assert_allclose([e1, e2], want[:2], rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(te, want[2], rtol=1e-11, atol=1e-15)


@pytest.mark.parametrize("tiling", ["reference", "orbital", "mixed"])
def test_lambda_other_tilings_match_oracle(tiling):
    """Jobs with the reference tiles as kernel tasks (fused passes over
    them), ORBITAL-tiled jobs and jobs of different tilings (whole tiles
    per task) give the oracle's energies and per-task pairs."""
    from nwchem_amd.job import TILING_ORBITAL, TILING_REFERENCE, TILING_SUBTILE

    tl_r, tl_l = {"reference": (TILING_REFERENCE, TILING_REFERENCE),
                  "orbital": (TILING_ORBITAL, TILING_ORBITAL),
                  "mixed": (TILING_SUBTILE, TILING_REFERENCE)}[tiling]
    s = spin_systems.SpinSystem(**LAMBDA_GPU[0])
    L = s.lambda_inputs()
    b = s.blocks
    want = oracle.lambda_ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, L["y1"], L["y1h"],
                                L["y2"], L["y2h"], b.v2, b.v2h, L["f1"], L["f1h"], LAMBDA_ORBITAL)
    print(f"lambda tiling={tiling}", flush=True)
    with TriplesJob(s.tiles, tiling=tl_r) as right, TriplesJob(s.tiles, tiling=tl_l) as left:
        right.upload(b)
        left.upload_lambda(L["y1"], L["y1h"], L["y2"], L["y2h"], b.v2, b.v2h)
        e1, e2, te = lambda_ccsd_t(right, left, L["f1"], L["f1h"], LAMBDA_ORBITAL)
    np.testing.assert_allclose([e1, e2], want[:2], rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(te, want[2], rtol=1e-11, atol=1e-15)


def test_lambda_without_fock_term_matches_oracle():
    s = spin_systems.SpinSystem(**LAMBDA_GPU[0])
    L = s.lambda_inputs()
    b = s.blocks
    want = oracle.lambda_ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, L["y1"], L["y1h"],
                                L["y2"], L["y2h"], b.v2, b.v2h, None, None, LAMBDA_ORBITAL)
    print("lambda case without f", flush=True)
    e1, e2, te = _run(s, L, LAMBDA_ORBITAL, fock=False)
    np.testing.assert_allclose([e1, e2], want[:2], rtol=1e-11, atol=1e-15)
    np.testing.assert_allclose(te, want[2], rtol=1e-11, atol=1e-15)


@pytest.mark.parametrize("path", ["fused", "tiles"])
@pytest.mark.parametrize("precision,rtol", [(FP32, 1e-6), (BF16X3, 3e-5), (BF16X6, 1e-6)],
                         ids=["fp32", "bf16x3", "bf16x6"])
def test_lambda_reduced_precision_near_fp64(precision, rtol, path, monkeypatch):
    for k, v in PATHS[path].items():
        monkeypatch.setenv(k, v)
    s = spin_systems.SpinSystem(**LAMBDA_GPU[1])
    L = s.lambda_inputs()
    print(f"lambda precision={precision} path={path}", flush=True)
    ref = _run(s, L, LAMBDA_ORBITAL)
    got = _run(s, L, LAMBDA_ORBITAL, precision)
    np.testing.assert_allclose(got[:2], ref[:2], rtol=rtol)


def test_lambda_fused_equals_tiles_on_a_synthetic_job(monkeypatch):
    """A C1 job of ragged tiles (several kernel tasks per reference task,
    sub-tile factors): the fused passes and the whole-tile path give the
    same energies (different summation orders)."""
    import bench
    from nwchem_amd import tiles as tl

    tm = tl.TileModel.synthetic(10, 36, 4, 10, 21)
    b, y1, y1h, y2, y2h, f1, f1h = bench.lambda_inputs(tm, 22)
    print(f"lambda synthetic no=10 nv=36 tasks={len(tm.tasks())}", flush=True)
    out = {}
    for path in ("fused", "tiles"):
        monkeypatch.delenv("TCE_T_LAMBDA", raising=False)
        for k, v in PATHS[path].items():
            monkeypatch.setenv(k, v)
        with TriplesJob(tm) as right, TriplesJob(tm) as left:
            right.upload(b)
            left.upload_lambda(y1, y1h, y2, y2h, b.v2, b.v2h)
            out[path] = lambda_ccsd_t(right, left, f1, f1h)
    np.testing.assert_allclose(out["fused"][:2], out["tiles"][:2], rtol=1e-12)
    np.testing.assert_allclose(out["fused"][2], out["tiles"][2], rtol=1e-10, atol=1e-14)


def test_lambda_rejects_mismatched_jobs():
    """Jobs over different tiles, or one job as both sides, are refused
    before any device work."""
    a = spin_systems.SpinSystem(**LAMBDA_GPU[0])
    c = spin_systems.SpinSystem(**LAMBDA_GPU[2])
    La = a.lambda_inputs()
    with TriplesJob(a.tiles) as ja, TriplesJob(c.tiles) as jc:
        ja.upload(a.blocks)
        jc.upload(c.blocks)
        with pytest.raises(RuntimeError, match="differ"):
            lambda_ccsd_t(ja, jc)
        with pytest.raises(RuntimeError, match="two distinct jobs"):
            lambda_ccsd_t(ja, ja)
        with pytest.raises(RuntimeError, match="y1 lacks"):
            jc.upload_lambda(La["y1"], np.array([0], np.int64), La["y2"], La["y2h"],
                             c.blocks.v2, c.blocks.v2h)


@pytest.mark.parametrize("path", ["fused", "tiles"])
@pytest.mark.parametrize("prm", LAMBDA_GPU, ids=IDS)
def test_lambda_with_y_equal_t_is_the_triples_energy(prm, path, monkeypatch):
    """The amplitudes as Lambda amplitudes, no Fock term: the left tiles are
    the right ones, so energy1 = E[T] and energy2 = E(T) of the same job,
    task by task."""
    for k, v in PATHS[path].items():
        monkeypatch.setenv(k, v)
    s = spin_systems.SpinSystem(**prm)
    b = s.blocks
    y1, y1h, y2, y2h = s.y_equal_t()
    print(f"lambda y=t case {prm} path={path}", flush=True)
    with TriplesJob(s.tiles) as right, TriplesJob(s.tiles) as left:
        right.upload(b)
        e = right.run()
        te_t = right.task_energies()
        left.upload_lambda(y1, y1h, y2, y2h, b.v2, b.v2h)
        e1, e2, te = lambda_ccsd_t(right, left)
    np.testing.assert_allclose([e1, e2], e, rtol=1e-12)
    np.testing.assert_allclose(te, te_t, rtol=1e-11, atol=1e-15)


@pytest.mark.timeout(600)
def test_lambda_at_the_benchmark_size_is_the_triples_energy():
    """no=40 (tiles of 20), nv=320 (tiles of 40), 480 tasks: with y = t and
    no Fock term the fused Lambda passes give the job's own E[T] and E(T)
    for every task, in FP64 and split-BF16 (3 terms).  The job's FP64 task
    energies at this size are pinned against the reference's compiled
    kernels (tests/golden/no40_nv320_tasks.json)."""
    import bench
    from nwchem_amd import tiles as tl

    seed = 20261202
    tm = tl.TileModel.synthetic(40, 320, 20, 40, seed)
    b, y1, y1h, y2, y2h, _, _ = bench.lambda_inputs(tm, seed)  # y from the same blocks
    print("lambda at no=40/nv=320: inputs ready", flush=True)
    with TriplesJob(tm) as right, TriplesJob(tm) as left:
        right.upload(b)
        left.upload_lambda(y1, y1h, y2, y2h, b.v2, b.v2h)
        for prec in (0, BF16X3):
            e = right.run(prec)
            te_t = right.task_energies()
            print(f"lambda at no=40/nv=320: precision {prec} (T) done", flush=True)
            e1, e2, te = lambda_ccsd_t(right, left, precision=prec)
            print(f"lambda at no=40/nv=320: precision {prec} lambda done", flush=True)
            np.testing.assert_allclose([e1, e2], e, rtol=1e-12)
            np.testing.assert_allclose(te, te_t, rtol=1e-11, atol=1e-15)
