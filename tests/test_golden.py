"""The oracle against the committed golden vectors (tests/golden/), which
were written by the reference's own compiled kernels (make_golden.py), and
the spin/symmetry/restricted rules of the task loop against closed forms.

These run without the reference tree: the vectors travel with the tests.
"""
import os
import sys

import numpy as np
import pytest

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
from golden import make_golden as mg  # noqa: E402
import spin_systems  # noqa: E402

import ctypes as C  # noqa: E402

GOLD = os.path.join(HERE, "golden")
DP = C.POINTER(C.c_double)


@pytest.fixture(scope="module")
def kernels():
    with np.load(os.path.join(GOLD, "kernels.npz")) as z:
        return {k: z[k] for k in z.files}


@pytest.fixture(scope="module")
def energies():
    import json

    with np.load(os.path.join(GOLD, "energies.npz")) as z:
        en = {k: z[k] for k in z.files}
    with open(os.path.join(GOLD, "energies.json")) as f:
        return en, json.load(f)


def test_golden_shapes_are_the_generator_shapes(kernels):
    assert kernels["shapes"].tolist() == [list(s) for s in mg.KERNEL_SHAPES]


@pytest.mark.parametrize("si", range(len(mg.KERNEL_SHAPES)))
@pytest.mark.parametrize("fam", ["s1", "d1", "d2"])
def test_oracle_kernels_bit_exact_vs_golden(kernels, fam, si):
    """ora_sd_t_* (sdt_oracle.c) against the tiles the reference's
    ccsd_t_kernels_omp.F left behind, bit for bit, for all 9 variants."""
    shape = mg.KERNEL_SHAPES[si]
    h3d, h2d, h1d, p6d, p5d, p4d, K = shape
    L = oracle.lib()
    for n in range(1, 10):
        t3, a, b = mg.kernel_inputs(fam, shape, n)
        start = t3.copy()
        ptrs = [x.ctypes.data_as(DP) for x in (t3, a, b)]
        if fam == "s1":
            L.ora_sd_t_s1(n, h3d, h2d, h1d, p6d, p5d, p4d, *ptrs)
        else:
            getattr(L, f"ora_sd_t_{fam}")(n, h3d, h2d, h1d, p6d, p5d, p4d, K, *ptrs)
        gold = kernels[f"shape{si}_{fam}_{n}"]
        assert not np.array_equal(gold, start)
        assert np.array_equal(t3, gold), f"{fam}_{n} shape {shape}"


@pytest.mark.parametrize("name", list(mg.SYSTEMS))
def test_oracle_energies_vs_golden(energies, name):
    """The oracle's task loop with its C kernels reproduces the golden
    energies (reference kernels under the same loop) bit for bit; the inputs
    are checked first, so a generator change cannot pass as a result."""
    en, meta = energies
    kind, prm = mg.SYSTEMS[name]
    ot, b, tm = mg.build_system(kind, prm)
    m = meta[name]
    assert mg.digest(b.t1h, b.t2h, b.v2h, tm.range, tm.evl) == m["tables_sha256"]
    assert mg.digest(b.t1, b.t2, b.v2) == m["inputs_sha256"]
    e1, e2, te = oracle.ccsd_t(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)
    assert te.shape[0] == m["tasks"]
    assert e1 == float.fromhex(m["E_bt"]) and e2 == float.fromhex(m["E_pt"])
    assert np.array_equal(te, en[name + "_tasks"])


# ---- spin, irreps, restricted (ccsd_t_gpu.F:94-112,153-167, tce_restricted.F)

SPIN_CASES = [
    dict(sizes_o=[3, 2], sizes_v=[4, 3], irreps_o=[0, 1], irreps_v=[1, 0], seed=5),
    dict(sizes_o=[2, 2], sizes_v=[3, 2, 2], irreps_o=[0, 3], irreps_v=[1, 2, 3], seed=6),
    dict(sizes_o=[4], sizes_v=[5, 4], seed=8),
]


def _energies(s):
    b = s.blocks
    return oracle.ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h)


def _close(a, b, rel=1e-12):
    return abs(a - b) <= rel * max(abs(a), abs(b)) + 1e-15


@pytest.mark.parametrize("prm", SPIN_CASES)
def test_spin_and_irrep_tiling_matches_textbook(prm):
    """Spin- and irrep-labelled tiles: blocks and tasks that break
    conservation are skipped; the sum equals the dense textbook (T) over
    every spin-orbital sextuple of tensors that vanish there."""
    s = spin_systems.SpinSystem(**prm)
    e1, e2, te = _energies(s)
    bf = s.brute_force()
    assert te.shape[0] > 0
    assert _close(e1, bf[0]) and _close(e2, bf[1]), ((e1, e2), bf)


@pytest.mark.parametrize("prm", SPIN_CASES)
def test_restricted_equals_unrestricted_on_closed_shell_data(prm):
    """Spin-flip symmetric data: the restricted run (all-beta blocks read
    from their alpha images, tasks with a beta majority skipped, factor 2)
    gives the unrestricted result over every task, and the textbook one."""
    u = spin_systems.SpinSystem(**prm, flip_symmetric=True)
    r = spin_systems.SpinSystem(**prm, flip_symmetric=True, restricted=True)
    eu, er = _energies(u), _energies(r)
    assert er[2].shape[0] < eu[2].shape[0]
    assert r.blocks.v2.size < u.blocks.v2.size  # all-beta blocks are not stored
    bf = u.brute_force()
    for k in (0, 1):
        assert _close(er[k], eu[k]) and _close(er[k], bf[k]), (er[:2], eu[:2], bf)


def test_restricted_factor_is_needed():
    """Planted mistake: dropping the restricted factor 2 (counting the
    alpha-majority tasks once) misses half the energy, so the check above
    would fail."""
    prm = SPIN_CASES[0]
    u = spin_systems.SpinSystem(**prm, flip_symmetric=True)
    r = spin_systems.SpinSystem(**prm, flip_symmetric=True, restricted=True)
    eu = _energies(u)
    te = _energies(r)[2]
    half = te.sum(axis=0) / 2.0  # what factor 1 would give
    assert not _close(half[0], eu[0], rel=1e-3)
    assert _close(2 * half[0], eu[0])


def test_beta_block_missing_is_an_error():
    """A permitted block absent from its table is an error of the task loop
    (the reference's tce_hash would errquit), never a silent zero."""
    s = spin_systems.SpinSystem(**SPIN_CASES[0])
    b = s.blocks
    t2h = b.t2h.copy()
    L = int(t2h[0])
    t2h[1 + L - 1] = 10**12  # the last key no longer matches its block
    with pytest.raises(RuntimeError):
        oracle.ccsd_t(s.oracle_tiles(), b.t1, b.t1h, b.t2, t2h, b.v2, b.v2h)


# ---- triples tiles of one task (the drivers' output to other consumers)

TILE_CASES = SPIN_CASES + [dict(sizes_o=[3, 2], sizes_v=[4, 3], seed=7, flip_symmetric=True,
                                restricted=True)]


@pytest.mark.parametrize("prm", TILE_CASES, ids=["irreps", "irreps_3v", "c1_ragged", "restricted"])
def test_oracle_task_tiles_equal_textbook_t3(prm):
    """ora_task_tiles (the singles and doubles dispatch of one task) against
    the dense textbook t3s / t3d restricted to the task's tiles, for every
    task: spin/irrep labels, ragged tiles, restricted runs (beta blocks read
    from their alpha images)."""
    s = spin_systems.SpinSystem(**prm)
    b = s.blocks
    ds, dd = s.dense_t3()
    ot = s.oracle_tiles()
    tasks = s.tiles.tasks()
    assert len(tasks) > 0
    scale = np.abs(dd).max()
    for task in tasks:
        ts, td = oracle.task_tiles(ot, b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, task)
        sl = s.tile_slices(task)
        assert np.abs(td - dd[sl]).max() <= 1e-14 * scale, task
        assert np.abs(ts - ds[sl]).max() <= 1e-14 * scale, task


def test_task_tiles_textbook_check_has_teeth():
    """Planted mistake: the textbook t3d with the d1 term's sign flipped
    differs from the oracle's tiles, so the check above can fail."""
    s = spin_systems.SpinSystem(**SPIN_CASES[0])
    b = s.blocks
    _, bad = s.dense_t3(d1_sign=-1.0)
    worst = 0.0
    for task in s.tiles.tasks():
        _, td = oracle.task_tiles(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h, task)
        worst = max(worst, np.abs(td - bad[s.tile_slices(task)]).max())
    assert worst > 1e-6


def test_oracle_task_tiles_rejects_non_tasks():
    s = spin_systems.SpinSystem(**SPIN_CASES[0])
    b = s.blocks
    T = s.tiles
    n = T.noab + T.nvab
    with pytest.raises(ValueError):  # particle tiles out of order
        oracle.task_tiles(s.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, b.v2, b.v2h,
                          (n, T.noab + 1, n, 1, 1, 1))
