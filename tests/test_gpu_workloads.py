"""GPU tests at the sizes of the batch and serving workloads.

Batch (64 independent jobs with no in [6, 30], nv in [48, 256], occupied
tiles 6-20, virtual tiles 16-48, ragged last tiles): jobs drawn from that
distribution in one batch; the small end against the oracle (FP64 within
the FP64 tolerance, split-BF16 against the oracle in the same representation
within the accumulation bound of tests/precision_model.py), every job's batch
energies equal to its own run bit for bit.

Serving (requests of no in [6, 40], nv in [48, 320], FP64 or split-BF16): a
request of the largest size (no=40, nv=320) together with small FP64 and
split-BF16 requests and a medium one: each result the same alone and in the
mix, bit for bit; the small FP64 ones against the oracle.
"""
import os
import sys

import numpy as np
import pytest

import oracle
from nwchem_amd import TILING_ORBITAL, TriplesBatch, TriplesJob, TriplesServer, tiles as tl

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import precision_model as pm  # noqa: E402

pytestmark = pytest.mark.gpu


def tol(e):
    return max(1e-10, 1e-12 * abs(e))


def ragged(n, ts):
    r = [ts] * (n // ts)
    if n % ts:
        r.append(n % ts)
    return r


def make(no, nv, to, tv, seed):
    tm = tl.TileModel.synthetic(no, nv, 0, 0, seed, ragged(no, to), ragged(nv, tv))
    return tm, tl.synth_blocks(tm, seed), None


# (no, nv, occupied tile, virtual tile): the small end (oracle-checked) and
# larger jobs of the batch distribution, every one with a ragged last tile
SMALL = [(6, 48, 6, 17), (7, 50, 6, 16), (9, 53, 8, 18)]
LARGE = [(13, 97, 9, 29), (21, 160, 16, 37), (30, 256, 20, 48)]


def test_batch_at_the_batch_workload_sizes():
    systems = [make(*s, seed=900 + k) for k, s in enumerate(SMALL + LARGE)]
    jobs = []
    try:
        for tm, b, _ in systems:
            j = TriplesJob(tm, tiling=TILING_ORBITAL)
            j.upload(b)
            jobs.append(j)
        print("batch jobs:", [(tm.no, tm.nv, list(tm.range)) for tm, b, _ in systems], flush=True)
        single = {p: [j.run(p) for j in jobs] for p in (0, 2)}
        with TriplesBatch(jobs) as batch:
            for p in (0, 2):
                e = batch.run(p)
                plan = batch.plan(p)
                print(f"precision {p}: plan {plan}, {batch.last_ms:.1f} ms", flush=True)
                # the autotuner's schedule: a work-queue slice of 32, 64 or 128
                # items, with its tuning runs when the 16x16 kernel was chosen
                # (none when the plan came from the process-wide cache)
                assert plan["slice"] in (32, 64, 128)
                if plan["kernel"].startswith("fused_triples_mixed") and plan["tuned_ms"]:
                    assert len(plan["slice_ms"]) == 3 and min(plan["slice_ms"]) > 0
                for k in range(len(jobs)):
                    assert tuple(e[k]) == single[p][k], (p, k, e[k], single[p][k])
        for k, (tm, b, _) in enumerate(systems[:len(SMALL)]):
            ot = oracle.Tiles(tm.range, tm.noab, tm.evl)
            e1, e2, _ = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
            g = single[0][k]
            assert abs(g[0] - e1) <= tol(e1) and abs(g[1] - e2) <= tol(e2), (k, g, e1, e2)
            em = pm.emulated(tm, b, ot, "bf16x3")
            _, bnd = pm.accumulation_bounds(tm, b, ot, "bf16x3")
            g = np.array(single[2][k])
            print(f"job {k}: bf16x3 |E - E_emul| {np.abs(g - em[:2])} bound {bnd}", flush=True)
            assert np.all(np.abs(g - np.array(em[:2])) <= bnd), (k, g, em[:2], bnd)
    finally:
        for j in jobs:
            j.close()


def test_server_largest_request_alone_and_in_the_mix():
    big = make(40, 320, 20, 40, 950)
    medium = make(24, 160, 12, 40, 951)
    small = [make(*s, seed=960 + k) for k, s in enumerate(SMALL)]
    systems = [big, medium] + small
    reqs = [(0, 2), (1, 0)] + [(2 + k, p) for k in range(len(small)) for p in (0, 2)]
    alone = {}
    with TriplesServer(slice_ms=20.0, max_active=8, reserve_bytes=0) as srv:
        sids = [srv.add_system(tm, b) for tm, b, _ in systems]
        for k, p in reqs:
            r = srv.wait(srv.submit(sids[k], p))
            alone[(k, p)] = (r["E_bt"], r["E_pt"])
            print(f"alone {k} {p}: {r['finish'] - r['start']:.3f} s", flush=True)
    with TriplesServer(slice_ms=20.0, max_active=8, reserve_bytes=0) as srv:
        sids = [srv.add_system(tm, b) for tm, b, _ in systems]
        tickets = [(k, p, srv.submit(sids[k], p)) for k, p in reqs]
        mix = {(k, p): srv.wait(t) for k, p, t in tickets}
        st = srv.stats()
    print("mix", st, {kp: round(r["finish"] - r["submit"], 3) for kp, r in mix.items()}, flush=True)
    for kp, r in mix.items():
        assert (r["E_bt"], r["E_pt"]) == alone[kp], kp
    # the small requests finished while the largest one was still running
    big_finish = mix[(0, 2)]["finish"]
    assert any(mix[(2 + k, 0)]["finish"] < big_finish for k in range(len(small)))
    for k, (tm, b, _) in enumerate(small):
        ot = oracle.Tiles(tm.range, tm.noab, tm.evl)
        e1, e2, _ = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
        g = alone[(2 + k, 0)]
        assert abs(g[0] - e1) <= tol(e1) and abs(g[1] - e2) <= tol(e2)
