"""Spin- and symmetry-labelled test systems, built from dense spin-orbital
tensors so that a textbook (T) over the dense tensors pins the tiled result.

NWChem tiles spin-orbitals by spin and irrep (tce_tile.F): alpha holes,
beta holes, alpha particles, beta particles, each space split into tiles
that carry one spin (k_spin: 1 alpha, 2 beta) and one irrep (k_sym, D2h
subgroup labels combined by XOR).  Blocks that break spin or symmetry
conservation are not stored (tce_t1_offset.F, tce_t2_offset.F,
tce_mo2e_offset.F), and tasks that break it are skipped (ccsd_t_gpu.F:94-112).
In a restricted run (closed-shell reference) all-beta blocks are read from
their alpha images (tce_restricted.F), tasks with more beta than alpha
spin-orbitals are skipped and the rest count twice (ccsd_t_gpu.F:153-157).

Dense tensors here are random, antisymmetric, and zero wherever spin or
irrep is not conserved, so the spin-adapted tiled sum must equal the dense
sum over every orbital sextuple (the skipped blocks and tasks contribute
exactly zero).  With spin-flip symmetric tensors (alpha <-> beta images
equal) the restricted run must equal the unrestricted one.
"""
from __future__ import annotations

import numpy as np

from nwchem_amd import tiles as tl


def _spin_orbitals(sizes_o, sizes_v, irreps_o, irreps_v):
    """Tile list (alpha then beta in each space): ranges, spins, irreps,
    alpha images (1-based) and per-orbital spin/irrep/spatial labels."""
    ro = list(sizes_o) * 2
    rv = list(sizes_v) * 2
    spin = [1] * len(sizes_o) + [2] * len(sizes_o) + [1] * len(sizes_v) + [2] * len(sizes_v)
    sym = list(irreps_o) * 2 + list(irreps_v) * 2
    no_t, nv_t = len(sizes_o), len(sizes_v)
    alpha = (list(range(1, no_t + 1)) * 2
             + list(range(2 * no_t + 1, 2 * no_t + nv_t + 1)) * 2)
    return np.array(ro + rv, np.int64), 2 * no_t, np.array(spin), np.array(sym), np.array(alpha)


def _orbital_labels(ranges, spin, sym):
    s, y = [], []
    for r, sp, sy in zip(ranges, spin, sym):
        s += [sp] * int(r)
        y += [sy] * int(r)
    return np.array(s), np.array(y)


def _antisym(x, pairs):
    for a, b in pairs:
        x = x - np.swapaxes(x, a, b)
    return x


class SpinSystem:
    """One spin-orbital (T) system: tiles, dense tensors, TCE blocks."""

    def __init__(self, sizes_o, sizes_v, irreps_o=None, irreps_v=None, seed=0,
                 flip_symmetric=False, restricted=False):
        irreps_o = [0] * len(sizes_o) if irreps_o is None else list(irreps_o)
        irreps_v = [0] * len(sizes_v) if irreps_v is None else list(irreps_v)
        ranges, noab, spin, sym, alpha = _spin_orbitals(sizes_o, sizes_v, irreps_o, irreps_v)
        self.no = int(ranges[:noab].sum())
        self.nv = int(ranges[noab:].sum())
        no, nv = self.no, self.nv
        so, yo = _orbital_labels(ranges[:noab], spin[:noab], sym[:noab])
        sv, yv = _orbital_labels(ranges[noab:], spin[noab:], sym[noab:])
        rng = np.random.default_rng(seed)
        no_s, nv_s = no // 2, nv // 2

        # orbital energies: beta copies of the alpha ones
        eo_a = np.sort(rng.uniform(-2.0, -0.3, no_s))
        ev_a = np.sort(rng.uniform(0.2, 4.0, nv_s))
        self.eo = np.concatenate([eo_a, eo_a])
        self.ev = np.concatenate([ev_a, ev_a])

        def flip(x, spaces):
            """Image of x under alpha <-> beta on every index."""
            for ax, sp in enumerate(spaces):
                n = no_s if sp == "o" else nv_s
                x = np.roll(x, n, axis=ax)
            return x

        def mask(spaces, signs):
            """spin and irrep conservation over the indices (sign +1: ket
            side, -1: bra side)."""
            lab_s = [so if sp == "o" else sv for sp in spaces]
            lab_y = [yo if sp == "o" else yv for sp in spaces]
            grids_s = np.ix_(*lab_s)
            grids_y = np.ix_(*lab_y)
            tot = sum(sg * g for sg, g in zip(signs, grids_s))
            xr = grids_y[0]
            for g in grids_y[1:]:
                xr = xr ^ g
            return (tot == 0) & (xr == 0)

        def make(shape_spaces, signs, pairs, sigma):
            shape = tuple(no if sp == "o" else nv for sp in shape_spaces)
            x = sigma * rng.standard_normal(shape)
            x = _antisym(x, pairs)
            if flip_symmetric:
                x = 0.5 * (x + flip(x, shape_spaces))
            return np.where(mask(shape_spaces, signs), x, 0.0)

        # T1[p,h], T2[p,p',h,h'], W[m,p,h,h'] = <m p||h h'>,
        # V[p,p',h,e] = <p p'||h e>, S[p,p',h,h'] = <p p'||h h'>
        self.T1 = make("vo", (1, -1), [], 0.01)
        self.T2 = make("vvoo", (1, 1, -1, -1), [(0, 1), (2, 3)], 0.02)
        self.W = make("ovoo", (1, 1, -1, -1), [(2, 3)], 0.05)
        self.V = make("vvov", (1, 1, -1, -1), [(0, 1)], 0.05)
        self.S = make("vvoo", (1, 1, -1, -1), [(0, 1), (2, 3)], 0.05)

        self._make = make  # later tensors (Lambda inputs) continue the same stream

        evl = np.concatenate([self.eo, self.ev])
        self.tiles = tl.TileModel(range=ranges, noab=noab, evl=evl, spin=spin, sym=sym,
                                  alpha=alpha, restricted=restricted)
        self.blocks = self._fill_blocks()

    def _fill_blocks(self) -> tl.Blocks:
        """Canonical TCE blocks (C order over the block's index list) from
        the dense tensors, at the offsets of the hash tables the library
        builds (same key formulas as tce_t1/t2/mo2e_offset)."""
        T = self.tiles
        noab, nvab = T.noab, T.nvab
        N = noab + nvab
        rng_, ofs = T.range, T.offset
        no = self.no

        def sl(b):  # orbital slice of tile b (1-based) within its space
            o = ofs[b - 1] - (no if b > noab else 0)
            return slice(int(o), int(o + rng_[b - 1]))

        def table(h):
            L = int(h[0])
            return [(int(h[1 + i]), int(h[1 + L + i])) for i in range(L)]

        t1h, t2h, v2h = T.t1_hash(), T.t2_hash(), T.v2_hash()
        s = T.sizes()
        t1 = np.zeros(max(s["t1"], 1))
        t2 = np.zeros(max(s["t2"], 1))
        v2 = np.zeros(max(s["v2"], 1))
        for key, off in table(t1h):
            h6, p5 = key % noab + 1, key // noab + noab + 1
            blk = self.T1[sl(p5), sl(h6)].ravel()
            t1[off:off + blk.size] = blk
        for key, off in table(t2h):
            h4 = key % noab + 1
            r = key // noab
            h3 = r % noab + 1
            r //= noab
            p2 = r % nvab + noab + 1
            p1 = r // nvab + noab + 1
            blk = self.T2[sl(p1), sl(p2), sl(h3), sl(h4)].ravel()
            t2[off:off + blk.size] = blk
        for key, off in table(v2h):
            g2 = key % N + 1
            r = key // N
            g1 = r % N + 1
            r //= N
            g4 = r % N + 1
            g3 = r // N + 1
            if g3 <= noab:
                D = self.W
            elif g2 > noab:
                D = self.V
            else:
                D = self.S
            blk = D[sl(g3), sl(g4), sl(g1), sl(g2)].ravel()
            v2[off:off + blk.size] = blk
        return tl.Blocks(t1, t1h, t2, t2h, v2, v2h)

    def oracle_tiles(self):
        import oracle

        T = self.tiles
        return oracle.Tiles(T.range, T.noab, T.evl, T.spin, T.sym, T.alpha, T.restricted)

    def dense_t3(self, d1_sign=1.0):
        """Textbook (t3s, t3d)[a, b, c, i, j, k] over all spin-orbitals:
            t3d = P(a/bc) P(k/ij) [-sum_e t(a e i j) <b c||k e>]
                + P(c/ab) P(i/jk) [-sum_m t(a b i m) <m c||j k>]
            t3s = P(a/bc) P(i/jk) t(a i) <b c||j k>
        with P(x/yz) f = f - f(x<->y) - f(x<->z).  d1_sign flips the second
        term (a planted mistake for the tests)."""
        def tr(x, pa, ph):
            return np.transpose(x, tuple(pa) + tuple(3 + q for q in ph))

        id3 = ((0, 1, 2), 1.0)
        p_a = [id3, ((1, 0, 2), -1.0), ((2, 1, 0), -1.0)]   # P(a/bc) on (a,b,c)
        p_c = [id3, ((2, 1, 0), -1.0), ((0, 2, 1), -1.0)]   # P(c/ab)
        h_k = [id3, ((2, 1, 0), -1.0), ((0, 2, 1), -1.0)]   # P(k/ij) on (i,j,k)
        h_i = [id3, ((1, 0, 2), -1.0), ((2, 1, 0), -1.0)]   # P(i/jk)
        X = -np.einsum("aeij,bcke->abcijk", self.T2, self.V)
        Y = -np.einsum("abim,mcjk->abcijk", self.T2, self.W)
        Z = np.einsum("ai,bcjk->abcijk", self.T1, self.S)
        d = sum(sa * sh * tr(X, pa, ph) for pa, sa in p_a for ph, sh in h_k)
        d = d + d1_sign * sum(sa * sh * tr(Y, pa, ph) for pa, sa in p_c for ph, sh in h_i)
        s = sum(sa * sh * tr(Z, pa, ph) for pa, sa in p_a for ph, sh in h_i)
        return s, d

    def tile_slices(self, task):
        """Orbital slices (a, b, c, i, j, k) of task = (p4b, p5b, p6b, h1b,
        h2b, h3b) (1-based tiles) in the dense tensors."""
        T = self.tiles

        def sl(b):
            o = int(T.offset[b - 1]) - (self.no if b > T.noab else 0)
            return slice(o, o + int(T.range[b - 1]))

        return tuple(sl(b) for b in task)

    def brute_force(self):
        import oracle

        return oracle.brute_force(self.no, self.nv, self.T1, self.T2, self.W, self.V, self.S,
                                  self.eo, self.ev)

    # ---- Lambda-CCSD(T) inputs ------------------------------------------

    def lambda_inputs(self):
        """Dense Lambda amplitudes Y1[i, a] = y(i a), Y2[i, j, a, b] =
        y(i j a b) and one-electron operator blocks Fhh, Fhp[i, a] = f(i a),
        Fph, Fpp (not symmetric, so a transposed read shows), plus their TCE
        blocks at the offsets of the tce_y1/y2/mo1e tables.  Returns a dict
        with the dense tensors and (y1, y1h, y2, y2h, f1, f1h)."""
        T = self.tiles
        mk = self._make
        d = dict(Y1=mk("ov", (1, -1), [], 0.02),
                 Y2=mk("oovv", (1, 1, -1, -1), [(0, 1), (2, 3)], 0.03),
                 Fhh=mk("oo", (1, -1), [], 0.1), Fhp=mk("ov", (1, -1), [], 0.05),
                 Fph=mk("vo", (1, -1), [], 0.05), Fpp=mk("vv", (1, -1), [], 0.1))
        noab, nvab = T.noab, T.nvab
        N = noab + nvab
        rng_, ofs = T.range, T.offset
        no = self.no

        def sl(b):
            o = ofs[b - 1] - (no if b > noab else 0)
            return slice(int(o), int(o + rng_[b - 1]))

        def table(h):
            L = int(h[0])
            return [(int(h[1 + i]), int(h[1 + L + i])) for i in range(L)]

        y1, y1h, y2, y2h = self.y_blocks(d["Y1"], d["Y2"])
        f1h = T.f1_hash()
        f_blocks = []
        for key, off in table(f1h):
            g5, g6 = key % N + 1, key // N + 1
            dense = d[("Fh" if g6 <= noab else "Fp") + ("h" if g5 <= noab else "p")]
            f_blocks.append((off, dense[sl(g6), sl(g5)].ravel()))
        f1 = np.zeros(max(sum(b.size for _, b in f_blocks), 1))
        for off, b in f_blocks:
            f1[off:off + b.size] = b
        d.update(y1=y1, y1h=y1h, y2=y2, y2h=y2h, f1=f1, f1h=f1h)
        return d

    def y_blocks(self, Y1, Y2):
        """Dense Y1[i, a], Y2[i, j, a, b] laid out as the tce_y1_offset /
        tce_y2_offset tables keep them: (y1, y1h, y2, y2h)."""
        T = self.tiles
        noab, nvab = T.noab, T.nvab
        rng_, ofs = T.range, T.offset
        no = self.no

        def sl(b):
            o = ofs[b - 1] - (no if b > noab else 0)
            return slice(int(o), int(o + rng_[b - 1]))

        def table(h):
            L = int(h[0])
            return [(int(h[1 + i]), int(h[1 + L + i])) for i in range(L)]

        y1h, y2h = T.y1_hash(), T.y2_hash()
        y1_blocks = []
        for key, off in table(y1h):
            h, p = key // nvab + 1, key % nvab + noab + 1
            y1_blocks.append((off, Y1[sl(h), sl(p)].ravel()))
        y1 = np.zeros(max(sum(b.size for _, b in y1_blocks), 1))
        for off, b in y1_blocks:
            y1[off:off + b.size] = b
        y2_blocks = []
        for key, off in table(y2h):
            p2 = key % nvab + noab + 1
            q = key // nvab
            p1 = q % nvab + noab + 1
            q //= nvab
            h8, h7 = q % noab + 1, q // noab + 1
            y2_blocks.append((off, Y2[sl(h7), sl(h8), sl(p1), sl(p2)].ravel()))
        y2 = np.zeros(max(sum(b.size for _, b in y2_blocks), 1))
        for off, b in y2_blocks:
            y2[off:off + b.size] = b
        return y1, y1h, y2, y2h

    def y_equal_t(self):
        """The amplitudes themselves as Lambda amplitudes: y(i a) = t(a i),
        y(i j a b) = t(a b i j), in the y tables' layout."""
        return self.y_blocks(self.T1.T, self.T2.transpose(2, 3, 0, 1))

    def dense_lambda(self, L, fock=True, y3f_sign=1.0):
        """Textbook Lambda-CCSD(T) energies over every spin-orbital sextuple,
        written from the term list of lambda_ccsd_t_left.F:7-10 with the
        integrals in the orientation written there (real orbitals:
        <jk||bc> = S[b,c,j,k], <jk||mc> = W[m,c,j,k], <ke||bc> = V[b,c,k,e]):
            ys  = P(i/jk) P(a/bc) y(i a) <jk||bc>
            yf  = P(k/ij) P(c/ab) y(i j a b) f(k c)
            yd1 = -P(i/jk) P(c/ab) sum_m y(i m a b) <jk||mc>
            yd2 = -P(k/ij) P(a/bc) sum_e y(i j a e) <ke||bc>
        paired with t3d of the same orbitals (lambda_ccsd_t.F:156-178):
            E1 = 1/36 sum t3d (yd1 + yd2 + yf) / D
            E2 = 1/36 sum t3d (ys + yd1 + yd2 + yf) / D."""
        def tr(x, pa, ph):
            return np.transpose(x, tuple(pa) + tuple(3 + q for q in ph))

        id3 = ((0, 1, 2), 1.0)
        p_a = [id3, ((1, 0, 2), -1.0), ((2, 1, 0), -1.0)]
        p_c = [id3, ((2, 1, 0), -1.0), ((0, 2, 1), -1.0)]
        h_k = [id3, ((2, 1, 0), -1.0), ((0, 2, 1), -1.0)]
        h_i = [id3, ((1, 0, 2), -1.0), ((2, 1, 0), -1.0)]

        def P(x, pp, hh):
            return sum(sa * sh * tr(x, pa, ph) for pa, sa in pp for ph, sh in hh)

        v_oovv = self.S.transpose(2, 3, 0, 1)
        v_ooov = self.W.transpose(2, 3, 0, 1)
        v_ovvv = self.V.transpose(2, 3, 0, 1)
        Y1, Y2 = L["Y1"], L["Y2"]
        ys = P(np.einsum("ia,jkbc->abcijk", Y1, v_oovv), p_a, h_i)
        yd = P(-np.einsum("imab,jkmc->abcijk", Y2, v_ooov), p_c, h_i)
        yd = yd + P(-np.einsum("ijae,kebc->abcijk", Y2, v_ovvv), p_a, h_k)
        if fock:
            yd = yd + y3f_sign * P(np.einsum("ijab,kc->abcijk", Y2, L["Fhp"]), p_c, h_k)
        _, t3d = self.dense_t3()
        eo, ev = self.eo, self.ev
        D = (-ev[:, None, None, None, None, None] - ev[None, :, None, None, None, None]
             - ev[None, None, :, None, None, None] + eo[None, None, None, :, None, None]
             + eo[None, None, None, None, :, None] + eo[None, None, None, None, None, :])
        e1 = float(np.sum(t3d * yd / D)) / 36.0
        e2 = float(np.sum(t3d * (ys + yd) / D)) / 36.0
        return e1, e2

    # ---- CR-CCSD(T) inputs ----------------------------------------------

    def i1_tables(self):
        """The block tables of the two moment intermediates, as
        OFFSET_cr_ccsd_t_N_1_1 (i1(h11 p4 h1 h2): loops p4, h11, h1, h2 >= h1;
        key h2-1 + noab (h1-1 + noab (h11-1 + noab (p4-noab-1))), block
        [p4][h11][h1][h2]) and OFFSET_cr_ccsd_t_N_2_1 (i1(p4 p5 h1 p12): loops
        p4, p5 >= p4, h1, p12; key p12-noab-1 + nvab (h1-1 + noab (p5-noab-1 +
        nvab (p4-noab-1))), block [p4][p5][h1][p12]) build them
        (cr_ccsd_t_N.F:773-830, 4011-4070): spin and irrep conservation, no
        all-beta block when restricted.  Returns ((hash, [(key, offset, tiles)])
        for each)."""
        T = self.tiles
        noab, nvab = T.noab, T.nvab
        N = noab + nvab
        spin, sym, rng_ = T.spin, T.sym, T.range

        def ok(a, b, c, d):
            sp = spin[a - 1] + spin[b - 1] + spin[c - 1] + spin[d - 1]
            return ((not T.restricted or sp != 8)
                    and spin[a - 1] + spin[b - 1] == spin[c - 1] + spin[d - 1]
                    and (sym[a - 1] ^ sym[b - 1] ^ sym[c - 1] ^ sym[d - 1]) == 0)

        def build(entries):
            L = len(entries)
            h = np.zeros(1 + 2 * L, dtype=np.int64)
            h[0] = L
            out, size = [], 0
            for n, (key, tiles) in enumerate(entries):
                h[1 + n], h[1 + L + n] = key, size
                out.append((key, size, tiles))
                size += int(np.prod([rng_[b - 1] for b in tiles]))
            return h, out, size

        e1 = [(h2 - 1 + noab * (h1 - 1 + noab * (h11 - 1 + noab * (p4 - noab - 1))),
               (p4, h11, h1, h2))
              for p4 in range(noab + 1, N + 1) for h11 in range(1, noab + 1)
              for h1 in range(1, noab + 1) for h2 in range(h1, noab + 1) if ok(h11, p4, h1, h2)]
        e2 = [(p12 - noab - 1 + nvab * (h1 - 1 + noab * (p5 - noab - 1 + nvab * (p4 - noab - 1))),
               (p4, p5, h1, p12))
              for p4 in range(noab + 1, N + 1) for p5 in range(p4, N + 1)
              for h1 in range(1, noab + 1) for p12 in range(noab + 1, N + 1) if ok(p4, p5, h1, p12)]
        return build(e1), build(e2)

    def i1_blocks(self, I1, I2):
        """Dense I1[m, c, j, k] = i1(h11=m p4=c h1=j h2=k) and I2[b, c, k, e]
        = i1(p4=b p5=c h1=k p12=e) as the two tables keep them: (i1_1,
        i1_1h, i1_2, i1_2h)."""
        T = self.tiles
        noab = T.noab
        rng_, ofs = T.range, T.offset
        no = self.no

        def sl(b):
            o = ofs[b - 1] - (no if b > noab else 0)
            return slice(int(o), int(o + rng_[b - 1]))

        (h1, e1, n1), (h2, e2, n2) = self.i1_tables()
        a1, a2 = np.zeros(max(n1, 1)), np.zeros(max(n2, 1))
        for _, off, (p4, h11, hh1, hh2) in e1:
            blk = I1[sl(h11), sl(p4), sl(hh1), sl(hh2)].transpose(1, 0, 2, 3).ravel()
            a1[off:off + blk.size] = blk
        for _, off, (p4, p5, hh1, p12) in e2:
            blk = I2[sl(p4), sl(p5), sl(hh1), sl(p12)].ravel()
            a2[off:off + blk.size] = blk
        return a1, h1, a2, h2

    def cr_inputs(self):
        """Random moment intermediates with the symmetry cr_ccsd_t_N gives
        them (antisymmetric in the pair the tables keep sorted) and their
        blocks: dict(I1, I2, i1_1, i1_1h, i1_2, i1_2h)."""
        mk = self._make
        I1 = mk("ovoo", (1, 1, -1, -1), [(2, 3)], 0.05)
        I2 = mk("vvov", (1, 1, -1, -1), [(0, 1)], 0.05)
        a1, h1, a2, h2 = self.i1_blocks(I1, I2)
        return dict(I1=I1, I2=I2, i1_1=a1, i1_1h=h1, i1_2=a2, i1_2h=h2)

    def dense_cr(self, I1, I2, den_sign=1.0):
        """Textbook CR-CCSD(T) over every spin-orbital sextuple, from the term
        lists of cr_ccsd_t_N.F:8,26, cr_ccsd_t_E.F:7-9 and cr_ccsd_t_D.F:4-7
        (c = t, cr_ccsd_t.F:58-59), with P(x/yz) as in dense_t3 and P(4) the
        four signed exchanges of (a/b)(i/j):
            m   = -P(c/ab) P(i/jk) sum_m t(a b i m) I1(m c j k)
                  -P(a/bc) P(k/ij) sum_e t(a e i j) I2(b c k e)
            q   = P(c/ab) P(k/ij) t(a b i j) t(c k) - 2/3 P(a/bc) P(i/jk) t(a i) x(b c j k),
                  x(b c j k) = -1/4 P(4) t(b j) t(c k)
            den0 = sum t(a i)^2 + 1/4 sum t(a b i j)^2 + 1/2 sum t(a i) t(b j) t(a b i j)
        paired with the (T) tiles (cr_ccsd_t.F:150-205):
            num1 = 1/36 sum m td / D, num2 = 1/36 sum m (ts + td) / D,
            den1, den2 the same with q; energy_n = num_n / (1 + den_n + den0).
        den_sign flips the second term of q (a planted mistake)."""
        def tr(x, pa, ph):
            return np.transpose(x, tuple(pa) + tuple(3 + q for q in ph))

        id3 = ((0, 1, 2), 1.0)
        p_a = [id3, ((1, 0, 2), -1.0), ((2, 1, 0), -1.0)]
        p_c = [id3, ((2, 1, 0), -1.0), ((0, 2, 1), -1.0)]
        h_k = [id3, ((2, 1, 0), -1.0), ((0, 2, 1), -1.0)]
        h_i = [id3, ((1, 0, 2), -1.0), ((2, 1, 0), -1.0)]

        def P(x, pp, hh):
            return sum(sa * sh * tr(x, pa, ph) for pa, sa in pp for ph, sh in hh)

        t1, t2 = self.T1, self.T2
        m = P(-np.einsum("abim,mcjk->abcijk", t2, I1), p_c, h_i)
        m = m + P(-np.einsum("aeij,bcke->abcijk", t2, I2), p_a, h_k)
        tt = np.einsum("bj,ck->bcjk", t1, t1)
        x = -0.25 * (tt - tt.transpose(1, 0, 2, 3) - tt.transpose(0, 1, 3, 2)
                     + tt.transpose(1, 0, 3, 2))
        q = P(np.einsum("abij,ck->abcijk", t2, t1), p_c, h_k)
        q = q + den_sign * (-2.0 / 3.0) * P(np.einsum("ai,bcjk->abcijk", t1, x), p_a, h_i)
        den0 = (float(np.sum(t1 * t1)) + 0.25 * float(np.sum(t2 * t2))
                + 0.5 * float(np.einsum("ai,bj,abij->", t1, t1, t2)))
        ts, td = self.dense_t3()
        eo, ev = self.eo, self.ev
        D = (-ev[:, None, None, None, None, None] - ev[None, :, None, None, None, None]
             - ev[None, None, :, None, None, None] + eo[None, None, None, :, None, None]
             + eo[None, None, None, None, :, None] + eo[None, None, None, None, None, :])
        out = dict(num1=float(np.sum(m * td / D)) / 36.0,
                   num2=float(np.sum(m * (ts + td) / D)) / 36.0,
                   den1=float(np.sum(q * td / D)) / 36.0,
                   den2=float(np.sum(q * (ts + td) / D)) / 36.0, den0=den0)
        out["energy1"] = out["num1"] / (1.0 + out["den1"] + den0)
        out["energy2"] = out["num2"] / (1.0 + out["den2"] + den0)
        return out
