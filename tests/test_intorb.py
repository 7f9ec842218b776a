"""Orbital-form integral input (intorb): the library's tce_t_v2_alpha_offset
and tce_t_v2_from_intorb against the oracle's literal restatements of
tce_mo2e_offset_intorb.F, tce_hash_v2 (tce_hash.F) and get_block_ind_i
(get_block_ind.F:818-1470), and both against mathematics:

  * the block table: block count and total size by brute-force enumeration,
    every stored key found by the reference's segment walk at the running
    sum of block sizes;
  * the spin-orbital blocks: on integrals with the 8-fold symmetry of real
    orbitals, stored as the four-index transformation stores them
    (tce_mo2e_trans.F), every block equals the dense
    <pq||rs> = (pr|qs) d(sp,sr) d(sq,ss) - (ps|qr) d(sp,ss) d(sq,sr),
    bit for bit (each value is one copy or one difference);
  * on stored blocks without any symmetry the library's orientation rule
    equals the reference's sixteen sorts element for element.

CPU only: the conversion is host code ahead of tce_t_job_upload.  The GPU
test of the energies is in test_gpu_parity.py.
"""
import os
import sys

import numpy as np
import pytest

import oracle
from nwchem_amd import tiles as tl
from nwchem_amd import _lib

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
from spin_systems import SpinSystem  # noqa: E402


def index_pair(i, j):
    return i * (i - 1) // 2 + j


def alpha_model(sizes_o, sizes_v, irreps_o, irreps_v):
    """Alpha tile list (occupied then virtual) and k_b2am of the spin-orbital
    tiles of SpinSystem (alpha holes, beta holes, alpha particles, beta
    particles), as tce_tile.F:1156-1212 maps them without an active space."""
    noa, nva = len(sizes_o), len(sizes_v)
    b2am = (list(range(1, noa + 1)) * 2 + list(range(noa + 1, noa + nva + 1)) * 2)
    return dict(noa=noa, nva=nva, b2am=b2am, spin_alpha=[1] * (noa + nva),
                sym_alpha=list(irreps_o) + list(irreps_v), range_alpha=list(sizes_o) + list(sizes_v))


def orbital_integrals(sizes_o, sizes_v, irreps_o, irreps_v, seed, symmetric=True):
    """Dense (pq|rs) over the spatial orbitals (occupied then virtual): the
    8-fold symmetry of real orbitals, zero where the irreps' XOR is not 0."""
    lab = []
    for r, y in zip(list(sizes_o) + list(sizes_v), list(irreps_o) + list(irreps_v)):
        lab += [y] * r
    lab = np.array(lab)
    n = len(lab)
    x = np.random.default_rng(seed).standard_normal((n, n, n, n)) * 0.05
    if symmetric:
        x = x + x.transpose(1, 0, 2, 3)
        x = x + x.transpose(0, 1, 3, 2)
        x = x + x.transpose(2, 3, 0, 1)
    g = np.ix_(lab, lab, lab, lab)
    return np.where((g[0] ^ g[1] ^ g[2] ^ g[3]) == 0, x, 0.0)


def spin_orbital_v(orb, no_s, nv_s, sign2=-1.0):
    """Dense <PQ||RS> over the spin-orbitals of SpinSystem's order (occupied
    alpha, occupied beta, virtual alpha, virtual beta)."""
    no, nv = 2 * no_s, 2 * nv_s
    sp = np.array([0] * no_s + [1] * no_s + [0] * nv_s + [1] * nv_s)
    spat = np.concatenate([np.arange(no_s), np.arange(no_s), no_s + np.arange(nv_s),
                           no_s + np.arange(nv_s)])
    O = orb[np.ix_(spat, spat, spat, spat)]               # O[P,Q,R,S] = (p q|r s)
    d = (sp[:, None] == sp[None, :]).astype(float)
    first = np.einsum("prqs->pqrs", O) * d[:, None, :, None] * d[None, :, None, :]
    second = np.einsum("psqr->pqrs", O) * d[:, None, None, :] * d[None, :, :, None]
    V = first + sign2 * second
    o, v = slice(0, no), slice(no, no + nv)
    return V[o, v, o, o], V[v, v, o, v], V[v, v, o, o]


CASES = {
    # (occupied alpha tiles, virtual alpha tiles, irreps, restricted)
    "c1_ragged": ([3, 2], [4, 5, 3], None, None, True),
    "irreps": ([2, 3, 1], [3, 2, 4, 2], [0, 1, 1], [0, 1, 2, 3], True),
    "unrestricted_equal": ([3], [4, 4], None, None, False),
    "one_tile_each": ([4], [6], None, None, True),
}


def build(name, symmetric=True, seed=7):
    so, sv, yo, yv, restricted = CASES[name]
    yo = [0] * len(so) if yo is None else yo
    yv = [0] * len(sv) if yv is None else yv
    ss = SpinSystem(so, sv, yo, yv, seed=seed, flip_symmetric=True, restricted=restricted)
    orb = orbital_integrals(so, sv, yo, yv, seed + 1, symmetric)
    return ss, orb, alpha_model(so, sv, yo, yv)


@pytest.mark.parametrize("idiv2e", [1, 2, 3, 7])
@pytest.mark.parametrize("name", list(CASES))
def test_alpha_offset_table(name, idiv2e):
    ss, orb, am = build(name)
    io_p = tl.Intorb(**am)
    io_o = oracle.Intorb(**am)
    # brute force: blocks g3 <= g4, g1 <= g2 conserving spin and irrep with
    # INDEX_PAIR(g4,g3) >= INDEX_PAIR(g2,g1), their count and total size
    N = am["noa"] + am["nva"]
    r, y = am["range_alpha"], am["sym_alpha"]
    blocks = [(a, b, c, d) for a in range(1, N + 1) for b in range(a, N + 1)
              for c in range(1, N + 1) for d in range(c, N + 1)
              if (y[a - 1] ^ y[b - 1] ^ y[c - 1] ^ y[d - 1]) == 0
              and index_pair(b, a) >= index_pair(d, c)]
    if len(blocks) < idiv2e:
        # the reference leaves entries of such a table unset: the library
        # refuses it
        with pytest.raises(RuntimeError):
            io_p.alpha_offset(idiv2e)
        return
    ho, so = io_o.alpha_offset(idiv2e)
    hp, sp = io_p.alpha_offset(idiv2e)
    assert np.array_equal(hp, ho) and sp == so
    assert so == sum(r[a - 1] * r[b - 1] * r[c - 1] * r[d - 1] for a, b, c, d in blocks)
    L1 = int(ho[0])
    assert L1 == (idiv2e if len(blocks) % idiv2e == 0 else idiv2e + 1)
    # segment entries: the first block and the last one are entries
    key = lambda a, b, c, d: d - 1 + N * (c - 1 + N * (b - 1 + N * (a - 1)))  # noqa: E731
    keys = [int(k) for k in ho[1:L1 + 2]]
    assert keys[0] == key(*blocks[0]) and keys[-1] == key(*blocks[-1])
    assert keys == sorted(keys)


@pytest.mark.parametrize("idiv2e", [1, 2, 5])
@pytest.mark.parametrize("name", list(CASES))
def test_hash_v2_walk_finds_every_block(name, idiv2e):
    ss, orb, am = build(name)
    io = oracle.Intorb(**am)
    N = am["noa"] + am["nva"]
    r, y = am["range_alpha"], am["sym_alpha"]
    nblocks = sum(1 for a in range(1, N + 1) for b in range(a, N + 1) for c in range(1, N + 1)
                  for d in range(c, N + 1) if (y[a - 1] ^ y[b - 1] ^ y[c - 1] ^ y[d - 1]) == 0
                  and index_pair(b, a) >= index_pair(d, c))
    if nblocks < idiv2e:
        pytest.skip("fewer blocks than idiv2e: the reference leaves table entries unset")
    h, size = io.alpha_offset(idiv2e)
    off = 0
    for a in range(1, N + 1):
        for b in range(a, N + 1):
            for c in range(1, N + 1):
                for d in range(c, N + 1):
                    if (y[a - 1] ^ y[b - 1] ^ y[c - 1] ^ y[d - 1]) != 0:
                        continue
                    if index_pair(b, a) < index_pair(d, c):
                        continue
                    k = d - 1 + N * (c - 1 + N * (b - 1 + N * (a - 1)))
                    assert io.hash_v2(h, k) == off, (a, b, c, d)
                    off += r[a - 1] * r[b - 1] * r[c - 1] * r[d - 1]
    assert off == size
    assert io.hash_v2(h, -1) == -1


def _convert(ss, orb, am, idiv2e=2):
    io_o = oracle.Intorb(**am)
    h, size = io_o.alpha_offset(idiv2e)
    v2orb = io_o.pack(orb, size)
    T = ss.tiles
    v2h = T.v2_hash()
    vsize = T.sizes()["v2"]
    got = tl.v2_from_intorb(T, tl.Intorb(**am), v2orb, h, v2h, vsize)
    lit = oracle.v2_from_intorb(ss.oracle_tiles(), io_o, v2orb, h, v2h, vsize)
    return got, lit, v2h


@pytest.mark.parametrize("name", list(CASES))
def test_blocks_equal_the_dense_antisymmetrised_integrals(name):
    ss, orb, am = build(name)
    no_s, nv_s = ss.no // 2, ss.nv // 2
    ss.W, ss.V, ss.S = spin_orbital_v(orb, no_s, nv_s)
    direct = ss._fill_blocks().v2
    got, lit, _ = _convert(ss, orb, am)
    assert np.array_equal(lit, direct)
    assert np.array_equal(got, direct)
    # a planted sign of the exchange term changes the blocks
    ss.W, ss.V, ss.S = spin_orbital_v(orb, no_s, nv_s, sign2=+1.0)
    assert not np.array_equal(got, ss._fill_blocks().v2)


@pytest.mark.parametrize("name", list(CASES))
def test_orientation_rule_equals_the_reference_sorts(name):
    """Stored blocks without any symmetry: which stored element lands where
    is then visible in every value."""
    ss, _, am = build(name)
    io = oracle.Intorb(**am)
    h, size = io.alpha_offset(2)
    v2orb = np.random.default_rng(3).standard_normal(size)
    T = ss.tiles
    v2h = T.v2_hash()
    vsize = T.sizes()["v2"]
    got = tl.v2_from_intorb(T, tl.Intorb(**am), v2orb, h, v2h, vsize)
    lit = oracle.v2_from_intorb(ss.oracle_tiles(), io, v2orb, h, v2h, vsize)
    assert np.array_equal(got, lit)
    # and without the caller's table (the library's own enumeration)
    assert np.array_equal(tl.v2_from_intorb(T, tl.Intorb(**am), v2orb, None, v2h, vsize), lit)


def _all_blocks(T):
    """A block table over every spin- and irrep-conserving quadruple
    (g3, g4, g1, g2) of spin-orbital tiles, in any order: the blocks other
    TCE code fetches with get_hash_block_i as well, which reach all sixteen
    sorts of get_block_ind_i (the three classes of the (T) path reach nine)."""
    N = T.noab + T.nvab
    sp, sy, rg = T.spin, T.sym, T.range
    keys, offs, off = [], [], 0
    for g3 in range(1, N + 1):
        for g4 in range(1, N + 1):
            for g1 in range(1, N + 1):
                for g2 in range(1, N + 1):
                    if sp[g3 - 1] + sp[g4 - 1] != sp[g1 - 1] + sp[g2 - 1]:
                        continue
                    if (sy[g3 - 1] ^ sy[g4 - 1] ^ sy[g1 - 1] ^ sy[g2 - 1]) != 0:
                        continue
                    keys.append(g2 - 1 + N * (g1 - 1 + N * (g4 - 1 + N * (g3 - 1))))
                    offs.append(off)
                    off += int(rg[g3 - 1] * rg[g4 - 1] * rg[g1 - 1] * rg[g2 - 1])
    return np.array([len(keys)] + keys + offs, np.int64), off


@pytest.mark.parametrize("name", ["c1_ragged", "irreps"])
def test_every_block_order_equals_the_dense_integrals(name):
    ss, orb, am = build(name)
    T = ss.tiles
    h_all, size = _all_blocks(T)
    io_o = oracle.Intorb(**am)
    h, osize = io_o.alpha_offset(2)
    v2orb = io_o.pack(orb, osize)
    got = tl.v2_from_intorb(T, tl.Intorb(**am), v2orb, h, h_all, size)
    lit = oracle.v2_from_intorb(ss.oracle_tiles(), io_o, v2orb, h, h_all, size)
    assert np.array_equal(got, lit)
    # the dense tensor over all spin-orbitals, sliced per block
    no_s, nv_s = ss.no // 2, ss.nv // 2
    sp = np.array([0] * no_s + [1] * no_s + [0] * nv_s + [1] * nv_s)
    spat = np.concatenate([np.arange(no_s), np.arange(no_s), no_s + np.arange(nv_s),
                           no_s + np.arange(nv_s)])
    O = orb[np.ix_(spat, spat, spat, spat)]
    d = (sp[:, None] == sp[None, :]).astype(float)
    V = (np.einsum("prqs->pqrs", O) * d[:, None, :, None] * d[None, :, None, :]
         - np.einsum("psqr->pqrs", O) * d[:, None, None, :] * d[None, :, :, None])
    N = T.noab + T.nvab
    L = int(h_all[0])
    ofs = np.concatenate([[0], np.cumsum(T.range)[:-1]])
    sl = lambda g: slice(int(ofs[g - 1]), int(ofs[g - 1] + T.range[g - 1]))  # noqa: E731
    for x in range(L):
        k, off = int(h_all[1 + x]), int(h_all[1 + L + x])
        g2 = k % N + 1
        k //= N
        g1 = k % N + 1
        k //= N
        g4 = k % N + 1
        g3 = k // N + 1
        blk = V[sl(g3), sl(g4), sl(g1), sl(g2)].ravel()
        assert np.array_equal(got[off:off + blk.size], blk), (g3, g4, g1, g2)


def test_energy_from_orbital_form_equals_the_textbook():
    """The (T) energy of the oracle's task loop over intorb-converted blocks
    against the dense brute force of the same spin-orbital system."""
    ss, orb, am = build("irreps")
    no_s, nv_s = ss.no // 2, ss.nv // 2
    ss.W, ss.V, ss.S = spin_orbital_v(orb, no_s, nv_s)
    got, _, v2h = _convert(ss, orb, am)
    b = ss.blocks
    e1, e2, _ = oracle.ccsd_t(ss.oracle_tiles(), b.t1, b.t1h, b.t2, b.t2h, got, v2h)
    r1, r2 = ss.brute_force()
    assert abs(e1 - r1) <= 1e-12 * max(1.0, abs(r1)) and abs(e2 - r2) <= 1e-12 * max(1.0, abs(r2))


def test_bad_inputs_are_rejected():
    ss, orb, am = build("c1_ragged")
    io = tl.Intorb(**am)
    h, size = io.alpha_offset(2)
    v2orb = np.zeros(size)
    T = ss.tiles
    v2h = T.v2_hash()
    bad = h.copy()
    bad[int(h[0]) + 2] += 1  # an entry's offset
    with pytest.raises(RuntimeError):
        tl.v2_from_intorb(T, io, v2orb, bad, v2h, T.sizes()["v2"])
    am2 = dict(am, range_alpha=[r + 1 for r in am["range_alpha"]])
    with pytest.raises(RuntimeError):
        tl.v2_from_intorb(T, tl.Intorb(**am2), v2orb, None, v2h, T.sizes()["v2"])
    assert _lib.lib().tce_t_v2_alpha_offset(None, 2, None, None) == -1
    badh = v2h.copy()
    badh[1] = (T.noab + T.nvab) ** 4  # a key past the last tile quadruple
    with pytest.raises(RuntimeError):
        tl.v2_from_intorb(T, io, v2orb, h, badh, T.sizes()["v2"])


def test_fortran_caller_converts_the_same_blocks():
    """tce_t_v2_alpha_offset and tce_t_v2_from_intorb through the Fortran
    module (nwchem_amd/fortran, `tce_t_fdemo intorb`): a closed-shell tiling
    with alpha holes (3, 2) and particles (4, 5, 3), orbital-form values
    mod(7919 k, 1000) / 1024 (exact, so the checksums are exact), against the
    Python binding on the same inputs."""
    import re
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                       "nwchem_amd", "fortran", "tce_t_fdemo")
    if not os.path.exists(exe):
        pytest.skip("Fortran demo not built")
    out = subprocess.run([exe, "intorb"], check=True, capture_output=True, text=True,
                         timeout=60).stdout
    m = re.search(r"intorb blocks=(\d+) size=(\d+) sum=\s*(\S+) check=\s*(\S+)", out)
    assert m, out
    r = [3, 2, 3, 2, 4, 5, 3, 4, 5, 3]
    T = tl.TileModel(range=np.array(r), noab=4, evl=-2.0 + 0.125 * np.arange(1, 35),
                     spin=np.array([1, 1, 2, 2, 1, 1, 1, 2, 2, 2]), sym=np.zeros(10, int),
                     alpha=np.array([1, 2, 1, 2, 5, 6, 7, 5, 6, 7]), restricted=True)
    io = tl.Intorb(noa=2, nva=3, b2am=[1, 2, 1, 2, 3, 4, 5, 3, 4, 5], spin_alpha=[1] * 5,
                   sym_alpha=[0] * 5, range_alpha=[3, 2, 4, 5, 3])
    h, osize = io.alpha_offset(2)
    k = np.arange(1, osize + 1, dtype=np.int64)
    orb = ((k * 7919) % 1000).astype(np.float64) / 1024.0
    v2h = T.v2_hash()
    vsize = T.sizes()["v2"]
    v = tl.v2_from_intorb(T, io, orb, h, v2h, vsize)
    assert int(m.group(1)) == int(v2h[0]) and int(m.group(2)) == vsize
    w = (np.arange(1, vsize + 1) % 17).astype(np.float64)
    assert float(m.group(3)) == float(np.sum(v)) and float(m.group(4)) == float(np.sum(v * w))
