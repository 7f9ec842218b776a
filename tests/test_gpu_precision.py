"""GPU parity of the reduced-precision modes (FP32, 3-term and 6-term split-BF16).

The check is split along the two error sources of tests/precision_model.py:

  * the operand conversion is integer work: the job's converted tensors
    (tce_t_job_operands) must equal the oracle's conversion of the dense FP64
    tensors (tce_t_pack_dense) bit for bit, padding included;
  * everything else is compared with the oracle's energies in the SAME
    representation (tests/golden/precision.npz, products exact, FP64
    accumulation): |E_gpu - E_emul| <= the accumulation bound of
    tests/precision_model.py, for E[T], E(T) and every task.

|E_gpu - E_fp64| is printed for every case (the representation error the
modes carry by design).
"""
import ctypes as C
import json
import os
import sys

import numpy as np
import pytest

import oracle
from nwchem_amd import TriplesJob, _lib

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from golden import make_golden as mg  # noqa: E402
import precision_model as pm  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
DP = C.POINTER(C.c_double)
LP = C.POINTER(C.c_long)

pytestmark = pytest.mark.gpu

MODES = ["fp32", "bf16x3", "bf16x6"]


def fixture(name, mode):
    with np.load(os.path.join(GOLD, "precision.npz")) as z:
        return {k: z[f"{name}_{mode}_{k}"] for k in ("total", "tasks", "task_bounds", "total_bounds")}


def fp64_golden(name):
    with np.load(os.path.join(GOLD, "energies.npz")) as z:
        return z[name + "_total"], z[name + "_tasks"]


def checked_system(name):
    """The golden system's inputs, verified against the fixture digests."""
    kind, prm = mg.SYSTEMS[name]
    ot, b, tm = mg.build_system(kind, prm)
    with open(os.path.join(GOLD, "precision.json")) as f:
        meta = json.load(f)[name]
    assert mg.digest(b.t1, b.t2, b.v2) == meta["inputs_sha256"], "inputs changed"
    return ot, b, tm


def dense(tm, b):
    """tce_t_pack_dense: T2e, Vd2, Wd1, T2m (FP64, contraction index last)."""
    no, nv = tm.no, tm.nv
    Ep, Mp = -(-nv // 8) * 8, -(-no // 8) * 8
    T2e = np.zeros((no, no, nv, Ep))
    T2m = np.zeros((no, nv, nv, Mp))
    Wd1 = np.zeros((no, no, nv, Mp))
    Vd2 = np.zeros((no, nv, nv, Ep))
    rc = _lib.lib().tce_t_pack_dense(C.byref(tm.c()), b.t1.ctypes.data_as(DP), b.t1h.ctypes.data_as(LP),
                                     b.t2.ctypes.data_as(DP), b.t2h.ctypes.data_as(LP),
                                     b.v2.ctypes.data_as(DP), b.v2h.ctypes.data_as(LP), None,
                                     T2e.ctypes.data_as(DP), T2m.ctypes.data_as(DP),
                                     Wd1.ctypes.data_as(DP), Vd2.ctypes.data_as(DP), None)
    assert rc == 0
    # tensor ids of tce_t_job_operands: 0 T2e, 1 Vd2 (K = nv), 2 Wd1, 3 T2m (K = no)
    return [(T2e, nv), (Vd2, nv), (Wd1, no), (T2m, no)]


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", ["no9_nv41_ragged", "spin_sym_unrestricted"])
def test_operands_are_the_oracle_conversion(name, mode):
    ot, b, tm = checked_system(name)
    prec = pm.PRECISION_ID[mode]
    with TriplesJob(tm) as job:
        job.upload(b)
        for tid, (X, K) in enumerate(dense(tm, b)):
            got = job.operands(prec, tid)
            rows = X.shape[0] * X.shape[1] * X.shape[2]
            x = X.reshape(rows, X.shape[3])[:, :K]
            if mode == "fp32":
                assert got.shape[0] == rows and got.shape[1] >= K and got.shape[1] % 16 == 0
                assert np.array_equal(got[:, :K].view(np.uint32), x.astype(np.float32).view(np.uint32))
                assert not got[:, K:].any()
            else:
                ref = oracle.split_bf16(x) if mode == "bf16x3" else oracle.split3_bf16(x)
                assert len(got) == len(ref)
                for part, r in zip(got, ref):
                    assert part.shape[0] == rows and part.shape[1] >= K and part.shape[1] % 32 == 0
                    assert np.array_equal(part[:, :K].ravel(), r), (name, tid)
                    assert not part[:, K:].any()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("name", mg.PRECISION_SYSTEMS)
def test_reduced_precision_matches_emulation(name, mode):
    ot, b, tm = checked_system(name)
    fx = fixture(name, mode)
    prec = pm.PRECISION_ID[mode]
    with TriplesJob(tm) as job:
        job.upload(b)
        e = np.array(job.run(prec))
        te = job.task_energies()
    e64, te64 = fp64_golden(name)
    d = np.abs(e - fx["total"])
    print(f"{name} {mode}: |E - E_emul| = {d} (bound {fx['total_bounds']}); "
          f"E - E_fp64 = {(e - e64) * 1e6} uEh", flush=True)
    assert np.all(d <= fx["total_bounds"]), (e, fx["total"], fx["total_bounds"])
    assert te.shape == fx["tasks"].shape
    assert np.all(np.abs(te - fx["tasks"]) <= fx["task_bounds"])


@pytest.mark.parametrize("mode", MODES)
def test_reduced_precision_tiles_element_bound(mode):
    """The task tiles of a reduced-precision run element by element: t3d
    against the oracle's tile in the same representation within the
    per-element accumulation bound lam sqrt(N_r) u A_i, t3s (FP64 in every
    mode) within 1e-12 of its largest element."""
    from nwchem_amd import TILING_ORBITAL
    ot, b, tm = checked_system("no9_nv41_ragged")
    prec = pm.PRECISION_ID[mode]
    tasks = tm.tasks()[::4]
    with TriplesJob(tm, tiling=TILING_ORBITAL) as job:
        job.upload(b)
        for t in tasks:
            p4, p5, p6, h1, h2, h3 = t
            gs, gd = job.triples_tiles(h1, h2, h3, p4, p5, p6, prec)
            rs, _ = oracle.task_tiles(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, t)
            with oracle.arith(mode):
                _, rd = oracle.task_tiles(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, t)
            with oracle.arith("abs"):
                _, ra = oracle.task_tiles(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, t)
            bound = pm.element_bound(tm, ra, mode)
            assert np.all(np.abs(gd - rd) <= bound), (t, np.max(np.abs(gd - rd) - bound))
            assert np.abs(gs - rs).max() <= 1e-12 * max(np.abs(rs).max(), 1e-300)
