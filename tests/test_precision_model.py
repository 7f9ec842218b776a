"""CPU checks of the reduced-precision error model (tests/precision_model.py)
and of the oracle's restatement of the GPU's operand arithmetic.

  * the operand split is round-to-nearest-even bf16, checked against the
    definition (nearest of the two neighbouring bf16 values, ties to the
    even significand);
  * the emulated modes reduce to FP64 exactly where the inputs are already
    representable (every product exact either way, same sums);
  * the committed references (tests/golden/precision.npz) are what the
    oracle computes now;
  * the accumulation bound is small enough to fail planted mistakes: a
    2-term split (lo*hi' dropped) and a truncated lo.
"""
import os
import sys

import numpy as np
import pytest

import oracle

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from golden import make_golden as mg  # noqa: E402
import precision_model as pm  # noqa: E402

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def bf16_to_f64(h):
    return (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def rne_bf16_by_definition(f):
    """bf16 of float32 values f: the nearer of the truncation and the next
    bf16 value up in magnitude, ties to the even significand."""
    u = f.view(np.uint32)
    lo = (u >> 16).astype(np.uint16)          # truncation toward zero
    hi = (lo.astype(np.uint32) + 1).astype(np.uint16)
    dlo = np.abs(f.astype(np.float64) - bf16_to_f64(lo))
    dhi = np.abs(f.astype(np.float64) - bf16_to_f64(hi))
    pick_hi = (dhi < dlo) | ((dhi == dlo) & (lo & 1 == 1))
    return np.where(pick_hi, hi, lo)


def test_split_is_round_to_nearest_even_bf16():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(200000) * 10.0 ** rng.integers(-6, 6, 200000),
                        # exact ties of the hi rounding (half a bf16 ulp above 1 + 2m/128)
                        1.0 + (2 * np.arange(64) + 1) * 2.0 ** -8,
                        # exact ties of the lo rounding: x - hi with 9 significant bits
                        1.0 + (257 + 2 * np.arange(128)) * 2.0 ** -21,
                        np.array([0.0, -0.0, 1.0, -3.5, 2.0 ** -120])])
    hi, lo = oracle.split_bf16(x)
    assert np.array_equal(hi, rne_bf16_by_definition(x.astype(np.float32)))
    r = (x - bf16_to_f64(hi)).astype(np.float32)
    assert np.array_equal(lo, rne_bf16_by_definition(r))
    # the pair represents x to 2^-17 relative (two 8-bit roundings plus fl32)
    rep = bf16_to_f64(hi) + bf16_to_f64(lo)
    nz = x != 0
    assert np.all(np.abs(rep[nz] - x[nz]) <= 2.0 ** -17 * np.abs(x[nz]))


def test_three_part_split_is_exact_to_2_to_the_24():
    """hi, mid, lo: each the round-to-nearest-even bf16 of the remainder
    left by the parts before it (in FP32); together x to 2^-24 relative."""
    x = np.random.default_rng(5).standard_normal(200000) * 10.0 ** np.random.default_rng(6).integers(-5, 5, 200000)
    hi, mid, lo = oracle.split3_bf16(x)
    assert np.array_equal(hi, rne_bf16_by_definition(x.astype(np.float32)))
    r1 = x - bf16_to_f64(hi)
    assert np.array_equal(mid, rne_bf16_by_definition(r1.astype(np.float32)))
    r2 = r1 - bf16_to_f64(mid)
    assert np.array_equal(lo, rne_bf16_by_definition(r2.astype(np.float32)))
    rep = bf16_to_f64(hi) + bf16_to_f64(mid) + bf16_to_f64(lo)
    assert np.all(np.abs(rep - x) <= 2.0 ** -24 * np.abs(x))


def test_planted_truncation_changes_the_split():
    x = np.random.default_rng(4).standard_normal(100000)
    _, lo = oracle.split_bf16(x)
    _, lo_t = oracle.split_bf16(x, truncate_lo=True)
    assert np.mean(lo != lo_t) > 0.3  # a bit-exact comparison of the split catches it


def _system(name):
    kind, prm = mg.SYSTEMS[name]
    return mg.build_system(kind, prm)


@pytest.mark.parametrize("mode,dtype", [("fp32", np.float32), ("bf16x3", "bf16"), ("bf16x6", "bf16")])
def test_emulation_is_fp64_on_representable_inputs(mode, dtype):
    ot, b, tm = _system("no5_nv9_to2_tv4")
    if dtype == "bf16":
        rnd = lambda a: bf16_to_f64(oracle.split_bf16(a)[0])  # noqa: E731
    else:
        rnd = lambda a: a.astype(np.float32).astype(np.float64)  # noqa: E731
    t2, v2 = rnd(b.t2), rnd(b.v2)
    e64 = oracle.ccsd_t(ot, b.t1, b.t1h, t2, b.t2h, v2, b.v2h)
    with oracle.arith(mode):
        em = oracle.ccsd_t(ot, b.t1, b.t1h, t2, b.t2h, v2, b.v2h)
    assert em[0] == e64[0] and em[1] == e64[1] and np.array_equal(em[2], e64[2])
    # and on the unrounded inputs the modes differ from FP64
    with oracle.arith(mode):
        eu = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    assert eu[0] != oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)[0]


def test_abs_mode_bounds_every_element():
    ot, b, tm = _system("no7_nv13_to3_tv5")
    for t in tm.tasks()[::7]:
        _, d = oracle.task_tiles(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, t)
        with oracle.arith("abs"):
            _, a = oracle.task_tiles(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, t)
        assert np.all(a >= np.abs(d) * (1 - 1e-12)) and a.max() > 0


@pytest.mark.parametrize("name", ["no8_nv32", "no5_nv9_to2_tv4", "spin_flip_restricted"])
def test_precision_references_are_current(name):
    ot, b, tm = _system(name)
    with np.load(os.path.join(GOLD, "precision.npz")) as z:
        for mode in ("fp32", "bf16x3", "bf16x6"):
            e1, e2, te = pm.emulated(tm, b, ot, mode)
            assert np.array_equal(z[f"{name}_{mode}_total"], [e1, e2])
            assert np.array_equal(z[f"{name}_{mode}_tasks"], te)


@pytest.mark.parametrize("name", ["no8_nv32", "no9_nv41_ragged"])
def test_bound_fails_a_two_term_split(name):
    ot, b, tm = _system(name)
    with np.load(os.path.join(GOLD, "precision.npz")) as z:
        tot, bnd = z[f"{name}_bf16x3_total"], z[f"{name}_bf16x3_total_bounds"]
    e1, e2, _ = pm.emulated(tm, b, ot, "bf16x2_planted")
    assert np.all(np.abs(np.array([e1, e2]) - tot) > 10 * bnd)


def test_bound_fails_a_truncated_lo():
    """At no=8/nv=32 a truncated lo moves some task's energy beyond its bound
    (it also fails the bit-exact split check of the GPU tests)."""
    name = "no8_nv32"
    ot, b, tm = _system(name)
    with np.load(os.path.join(GOLD, "precision.npz")) as z:
        te, tb = z[f"{name}_bf16x3_tasks"], z[f"{name}_bf16x3_task_bounds"]
    _, _, te_bad = pm.emulated(tm, b, ot, "bf16x3_trunc_planted")
    assert np.any(np.abs(te_bad - te) > tb)
