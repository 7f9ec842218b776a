"""Tile model and TCE block bookkeeping of the (T) path (host side).

Mirrors the tce.fh / MA arrays the reference's ccsd_t_gpu.F reads
(k_spin, k_sym, k_range, k_offset, k_alpha, k_evl_sorted; noab, nvab,
restricted) and the canonical block tables built by tce_t1_offset.F,
tce_t2_offset.F and tce_mo2e_offset.F.  All integer work is done by
libtce_t.so so that the device path and these helpers share one
implementation; arrays are numpy int64 (the reference's INTEGER*8).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import DP, LP, TilesC, check


def _lp(a: np.ndarray):
    return a.ctypes.data_as(LP)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(DP)


def tile_ranges(n: int, tilesize: int) -> np.ndarray:
    """tce_tile.F:330-355: ceil(n/tilesize) tiles differing by at most one."""
    L = _lib.lib()
    nb = L.tce_t_tile_ranges(n, tilesize, None)
    out = np.zeros(max(nb, 1), dtype=np.int64)
    L.tce_t_tile_ranges(n, tilesize, _lp(out))
    return out[:nb]


@dataclass
class TileModel:
    """Tiles of one (T) job; tiles are numbered from 1 (holes first)."""

    range: np.ndarray
    noab: int
    evl: np.ndarray
    spin: np.ndarray | None = None
    sym: np.ndarray | None = None
    alpha: np.ndarray | None = None
    restricted: bool = False
    _c: TilesC | None = field(default=None, init=False, repr=False)

    def __post_init__(self) -> None:
        self.range = np.ascontiguousarray(self.range, dtype=np.int64)
        n = len(self.range)
        self.nvab = n - self.noab
        if self.noab <= 0 or self.nvab <= 0:
            raise ValueError("need at least one hole and one particle tile")
        self.offset = np.concatenate([[0], np.cumsum(self.range)[:-1]]).astype(np.int64)
        self.spin = np.ones(n, np.int64) if self.spin is None else np.ascontiguousarray(self.spin, np.int64)
        self.sym = np.zeros(n, np.int64) if self.sym is None else np.ascontiguousarray(self.sym, np.int64)
        self.alpha = (np.arange(1, n + 1, dtype=np.int64) if self.alpha is None
                      else np.ascontiguousarray(self.alpha, np.int64))
        self.evl = np.ascontiguousarray(self.evl, dtype=np.float64)
        if len(self.evl) != int(self.range.sum()):
            raise ValueError("evl must hold one energy per orbital")

    # ---- sizes
    @property
    def no(self) -> int:
        return int(self.range[: self.noab].sum())

    @property
    def nv(self) -> int:
        return int(self.range[self.noab:].sum())

    @classmethod
    def synthetic(cls, no: int, nv: int, tile_o: int, tile_v: int, seed: int,
                  ranges_o=None, ranges_v=None) -> "TileModel":
        """C1 spin-orbital tiling: tce_tile ranges (or explicit ones) and
        seeded sorted orbital energies."""
        ro = tile_ranges(no, tile_o) if ranges_o is None else np.asarray(ranges_o, np.int64)
        rv = tile_ranges(nv, tile_v) if ranges_v is None else np.asarray(ranges_v, np.int64)
        if ro.sum() != no or rv.sum() != nv:
            raise ValueError("tile ranges do not add up to no / nv")
        evl = np.zeros(no + nv)
        check(_lib.lib().tce_t_synth_evl(seed, no, nv, _dp(evl)), "tce_t_synth_evl")
        return cls(range=np.concatenate([ro, rv]), noab=len(ro), evl=evl)

    def c(self) -> TilesC:
        """tce_t_tiles view (arrays stay owned by this object)."""
        if self._c is None:
            self._c = TilesC(self.noab, self.nvab, _lp(self.spin), _lp(self.sym), _lp(self.range),
                             _lp(self.offset), _lp(self.alpha), _dp(self.evl), int(self.restricted))
        return self._c

    # ---- canonical block tables
    def _hash(self, fname: str) -> np.ndarray:
        L = _lib.lib()
        f = getattr(L, fname)
        size = C.c_long(0)
        n = f(C.byref(self.c()), None, C.byref(size))
        if n < 0:
            raise RuntimeError(f"{fname}: invalid tiles")
        h = np.zeros(2 * n + 1, np.int64)
        f(C.byref(self.c()), _lp(h), C.byref(size))
        return h

    def t1_hash(self) -> np.ndarray:
        return self._hash("tce_t_t1_offset")

    def t2_hash(self) -> np.ndarray:
        return self._hash("tce_t_t2_offset")

    def v2_hash(self) -> np.ndarray:
        return self._hash("tce_t_v2_offset")

    def y1_hash(self) -> np.ndarray:
        """tce_y1_offset.F: Lambda singles y(h p), blocks [h][p]."""
        return self._hash("tce_t_y1_offset")

    def y2_hash(self) -> np.ndarray:
        """tce_y2_offset.F: Lambda doubles y(h h' p p'), blocks [h][h'][p][p']."""
        return self._hash("tce_t_y2_offset")

    def f1_hash(self) -> np.ndarray:
        """tce_mo1e_offset.F: one-electron operator f(g g'), blocks [g][g']."""
        return self._hash("tce_t_f1_offset")

    @staticmethod
    def block_size(hash_: np.ndarray, tiles: "TileModel", kind: str) -> int:
        f = {"t1": "tce_t_t1_offset", "t2": "tce_t_t2_offset", "v2": "tce_t_v2_offset"}[kind]
        size = C.c_long(0)
        getattr(_lib.lib(), f)(C.byref(tiles.c()), None, C.byref(size))
        return int(size.value)

    def sizes(self) -> dict:
        L = _lib.lib()
        out = {}
        for k, f in (("t1", "tce_t_t1_offset"), ("t2", "tce_t_t2_offset"), ("v2", "tce_t_v2_offset")):
            s = C.c_long(0)
            getattr(L, f)(C.byref(self.c()), None, C.byref(s))
            out[k] = int(s.value)
        return out

    def tasks(self) -> list[tuple[int, int, int, int, int, int]]:
        """Tile sextuples (p4b,p5b,p6b,h1b,h2b,h3b) in the order of the task
        loop of ccsd_t_gpu.F:87-112."""
        out = []
        no, nv = self.noab, self.nvab
        sp, sy = self.spin, self.sym
        for p4 in range(no + 1, no + nv + 1):
            for p5 in range(p4, no + nv + 1):
                for p6 in range(p5, no + nv + 1):
                    for h1 in range(1, no + 1):
                        for h2 in range(h1, no + 1):
                            for h3 in range(h2, no + 1):
                                s_p = sp[p4 - 1] + sp[p5 - 1] + sp[p6 - 1]
                                s_h = sp[h1 - 1] + sp[h2 - 1] + sp[h3 - 1]
                                if s_p != s_h:
                                    continue
                                if self.restricted and s_p + s_h > 8:
                                    continue
                                if (sy[p4 - 1] ^ sy[p5 - 1] ^ sy[p6 - 1] ^ sy[h1 - 1] ^ sy[h2 - 1]
                                        ^ sy[h3 - 1]) != 0:
                                    continue
                                out.append((p4, p5, p6, h1, h2, h3))
        return out


@dataclass
class Blocks:
    """Canonical TCE blocks (host copies of d_t1, d_t2, d_v2) and their hash
    tables (k_t1_offset, k_t2_offset, k_v2_offset)."""

    t1: np.ndarray
    t1h: np.ndarray
    t2: np.ndarray
    t2h: np.ndarray
    v2: np.ndarray
    v2h: np.ndarray


def synth_blocks(tiles: TileModel, seed: int) -> Blocks:
    """Seeded inputs as canonical TCE blocks (tce_t_synth_blocks)."""
    t1h, t2h, v2h = tiles.t1_hash(), tiles.t2_hash(), tiles.v2_hash()
    s = tiles.sizes()
    t1 = np.zeros(max(s["t1"], 1))
    t2 = np.zeros(max(s["t2"], 1))
    v2 = np.zeros(max(s["v2"], 1))
    check(_lib.lib().tce_t_synth_blocks(C.byref(tiles.c()), seed, _dp(t1), _lp(t1h), _dp(t2),
                                        _lp(t2h), _dp(v2), _lp(v2h)), "tce_t_synth_blocks")
    return Blocks(t1, t1h, t2, t2h, v2, v2h)


def synth_dense(seed: int, no: int, nv: int) -> dict:
    """Dense forms of the same seeded tensors (T1[p,h], T2[p,p',h,h'],
    W[m,p,h,h'] = <m p||h h'>, V[p,p',h,e] = <p p'||h e>, S[p,p',h,h'])."""
    T1 = np.zeros((nv, no))
    T2 = np.zeros((nv, nv, no, no))
    W = np.zeros((no, nv, no, no))
    V = np.zeros((nv, nv, no, nv))
    S = np.zeros((nv, nv, no, no))
    check(_lib.lib().tce_t_synth_dense(seed, no, nv, _dp(T1), _dp(T2), _dp(W), _dp(V), _dp(S)),
          "tce_t_synth_dense")
    return {"T1": T1, "T2": T2, "W": W, "V": V, "S": S}


@dataclass
class Intorb:
    """Orbital-form integral description (tce_t_intorb): the alpha tile list
    (k_spin_alpha, k_sym_alpha, k_range_alpha, noa occupied then nva
    virtual), the alpha tile of every spin-orbital tile (k_b2am) and the
    irrep of V."""

    noa: int
    nva: int
    b2am: np.ndarray
    spin_alpha: np.ndarray
    sym_alpha: np.ndarray
    range_alpha: np.ndarray
    irrep_v: int = 0

    def __post_init__(self) -> None:
        for f in ("b2am", "spin_alpha", "sym_alpha", "range_alpha"):
            setattr(self, f, np.ascontiguousarray(getattr(self, f), np.int64))

    def c(self):
        return _lib.IntorbC(self.noa, self.nva, _lp(self.b2am), _lp(self.spin_alpha),
                            _lp(self.sym_alpha), _lp(self.range_alpha), self.irrep_v)

    def alpha_offset(self, idiv2e: int = 2) -> tuple[np.ndarray, int]:
        """k_v2_alpha_offset and the orbital-form array size
        (tce_t_v2_alpha_offset)."""
        L = _lib.lib()
        size = C.c_long(0)
        n = L.tce_t_v2_alpha_offset(C.byref(self.c()), idiv2e, None, C.byref(size))
        if n < 0:
            raise RuntimeError(L.tce_t_last_error().decode())
        h = np.zeros(n, np.int64)
        L.tce_t_v2_alpha_offset(C.byref(self.c()), idiv2e, _lp(h), C.byref(size))
        return h, int(size.value)


def v2_from_intorb(tiles: TileModel, io: Intorb, v2orb: np.ndarray, alpha_offset, v2h,
                   size: int) -> np.ndarray:
    """The spin-orbital V2 blocks of v2h from orbital-form integrals
    (tce_t_v2_from_intorb, get_hash_block_i)."""
    v2 = np.zeros(max(size, 1))
    v2orb = np.ascontiguousarray(v2orb, np.float64)
    check(_lib.lib().tce_t_v2_from_intorb(C.byref(tiles.c()), C.byref(io.c()), _dp(v2orb),
                                          None if alpha_offset is None else _lp(alpha_offset),
                                          _lp(v2h), _dp(v2)), "tce_t_v2_from_intorb")
    return v2
