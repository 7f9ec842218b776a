// intorb.cpp -- orbital-form two-electron integrals (the intorb input of a
// closed-shell 2eorb run) turned into the spin-orbital V2 blocks the upload
// reads.  Host-only C++ compiled into libtce_t.so.
//
//   tce_t_v2_alpha_offset  tce_mo2e_offset_intorb.F:1-110: the orbital-form
//                          block table k_v2_alpha_offset (segmented)
//   tce_t_v2_from_intorb   get_hash_block_i (get_block_ind.F:818-1470,
//                          ioalg 2) for every block of a tce_mo2e_offset
//                          table, the lookup of tce_hash_v2 (tce_hash.F:1-140)
//
// Orbital-form storage (as tce_mo2e_trans.F:686-745 writes it): one block per
// alpha-tile quadruple g3 <= g4, g1 <= g2 that conserves spin and irrep and
// has INDEX_PAIR(g4,g3) >= INDEX_PAIR(g2,g1), in loop order g3, g4, g1, g2,
// each holding the integrals (g3 g4 | g1 g2) at [g1][g2][g3][g4] (C order,
// the g4 orbital fastest).  A spin-orbital block v(g3 g4 g1 g2) is
//   (g3 g1 | g4 g2)  when spin(g3) = spin(g1) and spin(g4) = spin(g2)
// minus
//   (g3 g2 | g4 g1)  when spin(g3) = spin(g2) and spin(g4) = spin(g1),
// each read from the stored block of its two orbital pairs: a pair is
// ordered by alpha tile (lower tile first, ties keep the pair's first
// orbital second), and the pairs are swapped when the first pair's
// INDEX_PAIR is the smaller.  This is the single rule behind the sixteen
// tce_sortacc_4 calls of get_block_ind_i.
#include "internal.h"

#include <algorithm>
#include <cstring>
#include <thread>
#include <unordered_map>
#include <vector>

namespace tce_t {
namespace {

struct Alpha {
  long n = 0;  // alpha tiles (occupied then virtual)
  std::vector<long> spin, sym, range, b2am;
  long irrep_v = 0;
  long sp(long a) const { return spin[a - 1]; }
  long sy(long a) const { return sym[a - 1]; }
  long rg(long a) const { return range[a - 1]; }
};

inline long index_pair(long i, long j) { return i * (i - 1) / 2 + j; }

bool load_alpha(const tce_t_intorb *io, long nso, Alpha &A) {
  if (!io || io->noa < 0 || io->nva < 0 || io->noa + io->nva <= 0 || !io->spin_alpha ||
      !io->sym_alpha || !io->range_alpha)
    return false;
  A.n = io->noa + io->nva;
  A.spin.assign(io->spin_alpha, io->spin_alpha + A.n);
  A.sym.assign(io->sym_alpha, io->sym_alpha + A.n);
  A.range.assign(io->range_alpha, io->range_alpha + A.n);
  for (long r : A.range)
    if (r <= 0) return false;
  A.irrep_v = io->irrep_v;
  if (nso > 0) {
    if (!io->b2am) return false;
    A.b2am.assign(io->b2am, io->b2am + nso);
    for (long a : A.b2am)
      if (a < 1 || a > A.n) return false;
  }
  return true;
}

// Orbital-form blocks in storage order: f(key, g3, g4, g1, g2, offset).
template <class F> long each_orb_block(const Alpha &A, F &&f) {
  const long N = A.n;
  long size = 0;
  for (long g3 = 1; g3 <= N; ++g3)
    for (long g4 = g3; g4 <= N; ++g4)
      for (long g1 = 1; g1 <= N; ++g1)
        for (long g2 = g1; g2 <= N; ++g2) {
          if (A.sp(g3) + A.sp(g4) != A.sp(g1) + A.sp(g2)) continue;
          if ((A.sy(g3) ^ A.sy(g4) ^ A.sy(g1) ^ A.sy(g2)) != A.irrep_v) continue;
          if (index_pair(g4, g3) < index_pair(g2, g1)) continue;
          const long key = g2 - 1 + N * (g1 - 1 + N * (g4 - 1 + N * (g3 - 1)));
          f(key, g3, g4, g1, g2, size);
          size += A.rg(g3) * A.rg(g4) * A.rg(g1) * A.rg(g2);
        }
  return size;
}

// One orbital pair of a spin-orbital block: tiles (alpha) and which of the
// block's four element indices (0 a = g3, 1 b = g4, 2 c = g1, 3 d = g2)
// runs over each orbital.
struct Pair {
  long t1, t2;  // alpha tiles of the pair's first and second orbital
  int e1, e2;   // element index of each
};

// Source of one half of a block: stored block offset and the stride of each
// element index in it.
struct Source {
  long offset = -1;
  long stride[4] = {0, 0, 0, 0};
};

bool half_source(const Alpha &A, const std::unordered_map<long, long> &offs, Pair x, Pair y,
                 Source &s) {
  // order each pair: lower alpha tile first (the stored g3' <= g4' and
  // g1' <= g2'); a pair whose first tile is >= its second keeps the second
  // orbital first (get_block_ind.F: l31s/l42s false when ig3b >= ig1b)
  auto order = [](Pair p, long &lo_t, int &lo_e, long &hi_t, int &hi_e) {
    if (p.t1 >= p.t2) {
      lo_t = p.t2, lo_e = p.e2, hi_t = p.t1, hi_e = p.e1;
    } else {
      lo_t = p.t1, lo_e = p.e1, hi_t = p.t2, hi_e = p.e2;
    }
  };
  long xl, xh, yl, yh;
  int xle, xhe, yle, yhe;
  order(x, xl, xle, xh, xhe);
  order(y, yl, yle, yh, yhe);
  // the pair with the larger INDEX_PAIR comes first (lp31p42 false when >=)
  long k3, k4, k1, k2;
  int e3, e4, e1, e2;
  if (index_pair(xh, xl) >= index_pair(yh, yl)) {
    k3 = xl, e3 = xle, k4 = xh, e4 = xhe, k1 = yl, e1 = yle, k2 = yh, e2 = yhe;
  } else {
    k3 = yl, e3 = yle, k4 = yh, e4 = yhe, k1 = xl, e1 = xle, k2 = xh, e2 = xhe;
  }
  const long N = A.n;
  const long key = k2 - 1 + N * (k1 - 1 + N * (k4 - 1 + N * (k3 - 1)));
  auto it = offs.find(key);
  if (it == offs.end()) return false;
  s.offset = it->second;
  // stored [g1'][g2'][g3'][g4']
  const long r2 = A.rg(k2), r3 = A.rg(k3), r4 = A.rg(k4);
  s.stride[e4] = 1;
  s.stride[e3] = r4;
  s.stride[e2] = r3 * r4;
  s.stride[e1] = r2 * r3 * r4;
  return true;
}

}  // namespace
}  // namespace tce_t

using namespace tce_t;

extern "C" long tce_t_v2_alpha_offset(const tce_t_intorb *io, long idiv2e, long *hash,
                                      long *size) {
  Alpha A;
  if (!load_alpha(io, 0, A) || idiv2e <= 0) {
    set_error("tce_t_v2_alpha_offset: invalid alpha tiles or idiv2e");
    return -1;
  }
  std::vector<long> keys, offs, g3s, g4s, g1s, g2s;
  const long total = each_orb_block(A, [&](long key, long g3, long g4, long g1, long g2, long off) {
    keys.push_back(key);
    offs.push_back(off);
    g3s.push_back(g3);
    g4s.push_back(g4);
    g1s.push_back(g1);
    g2s.push_back(g2);
  });
  if (size) *size = total;
  const long length = (long)keys.size();
  if (length < idiv2e) {
    // the reference leaves entries of such a table unset
    set_error("tce_t_v2_alpha_offset: fewer orbital-form blocks than idiv2e");
    return -1;
  }
  // segments of ipiece blocks (tce_mo2e_offset_intorb.F:45-110): an entry at
  // the first block, at every ipiece-th block, and at the last block when
  // the count does not divide evenly
  const long ipiece = length / idiv2e;
  const long length1 = ipiece * idiv2e == length ? idiv2e : idiv2e + 1;
  const long words = 6 * (length1 + 1) + 1;
  if (!hash) return words;
  std::fill(hash, hash + words, 0L);
  hash[0] = length1;
  long ipos = 1, count = 0;
  auto put = [&](long b) {
    hash[ipos] = keys[b];
    hash[(length1 + 1) + ipos] = offs[b];
    hash[2 * (length1 + 1) + ipos] = g3s[b];
    hash[3 * (length1 + 1) + ipos] = g4s[b];
    hash[4 * (length1 + 1) + ipos] = g1s[b];
    hash[5 * (length1 + 1) + ipos] = g2s[b];
    ++ipos;
  };
  for (long b = 0; b < length; ++b) {
    ++count;
    if (b == 0) put(b);
    if (count == ipiece) {
      put(b);
      count = 0;
    }
  }
  if (count != 0) put(length - 1);
  return words;
}

extern "C" int tce_t_v2_from_intorb(const tce_t_tiles *tiles, const tce_t_intorb *io,
                                    const double *v2orb, const long *v2_alpha_offset,
                                    const long *v2_hash, double *v2) {
  Tiles T;
  if (!T.load(tiles)) {
    set_error("tce_t_v2_from_intorb: invalid tile description");
    return TCE_T_ERR_ARG;
  }
  const long nso = T.noab + T.nvab;
  Alpha A;
  if (!load_alpha(io, nso, A) || !v2orb || !v2_hash || !v2) {
    set_error("tce_t_v2_from_intorb: invalid arguments");
    return TCE_T_ERR_ARG;
  }
  for (long g = 1; g <= nso; ++g)
    if (A.rg(A.b2am[g - 1]) != T.rng(g) || A.sy(A.b2am[g - 1]) != T.sy(g)) {
      set_error("tce_t_v2_from_intorb: a tile's range or irrep differs from its alpha tile's");
      return TCE_T_ERR_ARG;
    }
  std::unordered_map<long, long> offs;
  each_orb_block(A, [&](long key, long, long, long, long, long off) { offs[key] = off; });
  if (v2_alpha_offset) {
    // the caller's segmented table (tce_hash_v2 walks it from these entries)
    // must describe the same storage
    const long length1 = v2_alpha_offset[0];
    for (long ipos = 1; ipos <= length1 + 1; ++ipos) {
      const long key = v2_alpha_offset[ipos], off = v2_alpha_offset[(length1 + 1) + ipos];
      auto it = offs.find(key);
      if (it == offs.end() || it->second != off) {
        set_error("tce_t_v2_from_intorb: k_v2_alpha_offset does not match the alpha tiles");
        return TCE_T_ERR_ARG;
      }
    }
  }
  const long L = v2_hash[0];
  const long nkeys = nso * nso * nso * nso;
  if (L < 0) {
    set_error("tce_t_v2_from_intorb: invalid block table");
    return TCE_T_ERR_ARG;
  }
  for (long i = 0; i < L; ++i)
    if (v2_hash[1 + i] < 0 || v2_hash[1 + i] >= nkeys || v2_hash[1 + L + i] < 0) {
      set_error("tce_t_v2_from_intorb: a block key or offset is outside the tiles");
      return TCE_T_ERR_ARG;
    }
  int err = 0;
  auto work = [&](long i0, long i1) {
    for (long i = i0; i < i1; ++i) {
      const long key = v2_hash[1 + i], off = v2_hash[1 + L + i];
      long r = key;
      const long g2 = r % nso + 1;
      r /= nso;
      const long g1 = r % nso + 1;
      r /= nso;
      const long g4 = r % nso + 1;
      const long g3 = r / nso + 1;
      const long s3 = T.spn(g3), s4 = T.spn(g4), s1 = T.spn(g1), s2 = T.spn(g2);
      const bool first = s3 == s1 && s4 == s2, second = s3 == s2 && s4 == s1;
      const long a3 = A.b2am[g3 - 1], a4 = A.b2am[g4 - 1], a1 = A.b2am[g1 - 1],
                 a2 = A.b2am[g2 - 1];
      Source P, Q;
      // element indices: 0 = g3, 1 = g4, 2 = g1, 3 = g2
      if (first && !half_source(A, offs, Pair{a3, a1, 0, 2}, Pair{a4, a2, 1, 3}, P)) {
        __atomic_store_n(&err, 1, __ATOMIC_RELAXED);
        continue;
      }
      if (second && !half_source(A, offs, Pair{a3, a2, 0, 3}, Pair{a4, a1, 1, 2}, Q)) {
        __atomic_store_n(&err, 1, __ATOMIC_RELAXED);
        continue;
      }
      const long r3 = T.rng(g3), r4 = T.rng(g4), r1 = T.rng(g1), r2 = T.rng(g2);
      double *out = v2 + off;
      long n = 0;
      for (long a = 0; a < r3; ++a)
        for (long b = 0; b < r4; ++b)
          for (long c = 0; c < r1; ++c)
            for (long d = 0; d < r2; ++d, ++n) {
              // the reference zeroes the block, adds (g3 g1|g4 g2) with
              // factor 1, then (g3 g2|g4 g1) with factor -1
              double v = 0.0;
              if (first)
                v = v + v2orb[P.offset + a * P.stride[0] + b * P.stride[1] + c * P.stride[2] +
                              d * P.stride[3]];
              if (second)
                v = v + v2orb[Q.offset + a * Q.stride[0] + b * Q.stride[1] + c * Q.stride[2] +
                              d * Q.stride[3]] *
                            -1.0;
              out[n] = v;
            }
    }
  };
  const long nth = std::max(1L, std::min<long>(L, (long)std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (long t = 0; t < nth; ++t) th.emplace_back(work, L * t / nth, L * (t + 1) / nth);
  for (auto &x : th) x.join();
  if (err) {
    set_error("tce_t_v2_from_intorb: a block's orbital-form source is not stored");
    return TCE_T_ERR_KEY;
  }
  return TCE_T_OK;
}
