// host.cpp -- integer bookkeeping of the TCE (T) path and the seeded
// synthetic generator.  Host-only C++ compiled into libtce_t.so.
//
//   tce_t_tile_ranges   tce_tile.F:330-355
//   tce_t_t1_offset     tce_t1_offset.F:20-56
//   tce_t_t2_offset     tce_t2_offset.F:1-70
//   tce_t_v2_offset     tce_mo2e_offset.F:1-70, keeping the three integral
//                       classes the (T) path reads (same keys, same order)
//   tce_t_hash          tce_hash.F:271-316
#include "internal.h"

#include <cmath>
#include <cstring>
#include <thread>
#include <vector>
#include <algorithm>

namespace tce_t {

long tile_ranges(long n, long tilesize, long *ranges) {
  if (n <= 0 || tilesize <= 0) return 0;
  long nb = (n + tilesize - 1) / tilesize;
  long done = 0;
  for (long k = 1; k <= nb; ++k) {
    long r = k * n / nb - done;
    if (ranges) ranges[k - 1] = r;
    done += r;
  }
  return nb;
}

bool Tiles::load(const tce_t_tiles *t) {
  if (!t || t->noab <= 0 || t->nvab <= 0 || !t->range || !t->offset || !t->spin ||
      !t->sym || !t->evl)
    return false;
  noab = t->noab;
  nvab = t->nvab;
  restricted = t->restricted != 0;
  long n = noab + nvab;
  range.assign(t->range, t->range + n);
  offset.assign(t->offset, t->offset + n);
  spin.assign(t->spin, t->spin + n);
  sym.assign(t->sym, t->sym + n);
  alpha.resize(n);
  for (long i = 0; i < n; ++i) alpha[i] = t->alpha ? t->alpha[i] : i + 1;
  norb = 0;
  for (long i = 0; i < n; ++i) {
    if (range[i] <= 0) return false;
    if (offset[i] != norb) return false;  // tiles must tile the orbitals in order
    norb += range[i];
  }
  no = offset[noab];
  nv = norb - no;
  evl.assign(t->evl, t->evl + norb);
  return true;
}

// Canonical block visitor for t(p1 p2 h3 h4); the enumeration order defines
// the offsets (tce_t2_offset.F:39-67).
template <class F> static void each_t2_block(const Tiles &T, F &&f) {
  long noab = T.noab, nvab = T.nvab;
  for (long p1 = noab + 1; p1 <= noab + nvab; ++p1)
    for (long p2 = p1; p2 <= noab + nvab; ++p2)
      for (long h3 = 1; h3 <= noab; ++h3)
        for (long h4 = h3; h4 <= noab; ++h4) {
          if (T.spn(p1) + T.spn(p2) != T.spn(h3) + T.spn(h4)) continue;
          if ((T.sy(p1) ^ T.sy(p2) ^ T.sy(h3) ^ T.sy(h4)) != 0) continue;
          if (T.restricted && T.spn(p1) + T.spn(p2) + T.spn(h3) + T.spn(h4) == 8) continue;
          long key = h4 - 1 + noab * (h3 - 1 + noab * (p2 - noab - 1 + nvab * (p1 - noab - 1)));
          f(key, p1, p2, h3, h4);
        }
}

template <class F> static void each_t1_block(const Tiles &T, F &&f) {
  long noab = T.noab, nvab = T.nvab;
  for (long p5 = noab + 1; p5 <= noab + nvab; ++p5)
    for (long h6 = 1; h6 <= noab; ++h6) {
      if (T.spn(p5) != T.spn(h6)) continue;
      if ((T.sy(p5) ^ T.sy(h6)) != 0) continue;
      if (T.restricted && T.spn(p5) + T.spn(h6) == 4) continue;
      f(h6 - 1 + noab * (p5 - noab - 1), p5, h6);
    }
}

// v(g3 g4 g1 g2) blocks read by the path: (h p|h h) d1, (p p|h p) d2,
// (p p|h h) s1.
template <class F> static void each_v2_block(const Tiles &T, F &&f) {
  long noab = T.noab, N = T.noab + T.nvab;
  auto part = [&](long b) { return b > noab; };
  for (long g3 = 1; g3 <= N; ++g3)
    for (long g4 = g3; g4 <= N; ++g4)
      for (long g1 = 1; g1 <= N; ++g1)
        for (long g2 = g1; g2 <= N; ++g2) {
          bool keep = (!part(g3) && part(g4) && !part(g1) && !part(g2)) ||
                      (part(g3) && part(g4) && !part(g1) && part(g2)) ||
                      (part(g3) && part(g4) && !part(g1) && !part(g2));
          if (!keep) continue;
          if (T.spn(g3) + T.spn(g4) != T.spn(g1) + T.spn(g2)) continue;
          if ((T.sy(g3) ^ T.sy(g4) ^ T.sy(g1) ^ T.sy(g2)) != 0) continue;
          if (T.restricted && T.spn(g3) + T.spn(g4) + T.spn(g1) + T.spn(g2) == 8) continue;
          long key = g2 - 1 + N * (g1 - 1 + N * (g4 - 1 + N * (g3 - 1)));
          f(key, g3, g4, g1, g2);
        }
}

// De-excitation amplitudes y(h p) (tce_y1_offset.F:39-58): hole tile outer,
// key p - noab - 1 + nvab (h - 1), block [h][p].
template <class F> static void each_y1_block(const Tiles &T, F &&f) {
  long noab = T.noab, nvab = T.nvab;
  for (long h4 = 1; h4 <= noab; ++h4)
    for (long p1 = noab + 1; p1 <= noab + nvab; ++p1) {
      if (T.spn(h4) != T.spn(p1)) continue;
      if ((T.sy(h4) ^ T.sy(p1)) != 0) continue;
      if (T.restricted && T.spn(h4) + T.spn(p1) == 4) continue;
      f(p1 - noab - 1 + nvab * (h4 - 1), h4, p1, 0, 0);
    }
}

// y(h7 h8 p1 p2), h7 <= h8, p1 <= p2 (tce_y2_offset.F:43-62), block
// [h7][h8][p1][p2].
template <class F> static void each_y2_block(const Tiles &T, F &&f) {
  long noab = T.noab, nvab = T.nvab;
  for (long h7 = 1; h7 <= noab; ++h7)
    for (long h8 = h7; h8 <= noab; ++h8)
      for (long p1 = noab + 1; p1 <= noab + nvab; ++p1)
        for (long p2 = p1; p2 <= noab + nvab; ++p2) {
          if (T.spn(h7) + T.spn(h8) != T.spn(p1) + T.spn(p2)) continue;
          if ((T.sy(h7) ^ T.sy(h8) ^ T.sy(p1) ^ T.sy(p2)) != 0) continue;
          if (T.restricted && T.spn(h7) + T.spn(h8) + T.spn(p1) + T.spn(p2) == 8) continue;
          f(p2 - noab - 1 + nvab * (p1 - noab - 1 + nvab * (h8 - 1 + noab * (h7 - 1))), h7, h8,
            p1, p2);
        }
}

// One-electron operator f(g6 g5) over every tile pair (tce_mo1e_offset.F:
// 35-50), key g5 - 1 + (noab + nvab)(g6 - 1), block [g6][g5].
template <class F> static void each_f1_block(const Tiles &T, F &&f) {
  long N = T.noab + T.nvab;
  for (long g6 = 1; g6 <= N; ++g6)
    for (long g5 = 1; g5 <= N; ++g5) {
      if (T.spn(g6) != T.spn(g5)) continue;
      if ((T.sy(g6) ^ T.sy(g5)) != 0) continue;
      if (T.restricted && T.spn(g6) + T.spn(g5) == 4) continue;
      f(g5 - 1 + N * (g6 - 1), g6, g5, 0, 0);
    }
}

template <class Each> static long build_hash(Each each, long *hash, long *size,
                                             const Tiles &T, int nidx) {
  std::vector<long> keys, sizes;
  each([&](long key, long a, long b, long c, long d) {
    keys.push_back(key);
    long s = T.rng(a) * T.rng(b);
    if (nidx == 4) s *= T.rng(c) * T.rng(d);
    sizes.push_back(s);
  });
  long L = (long)keys.size(), off = 0;
  if (hash) {
    hash[0] = L;
    for (long i = 0; i < L; ++i) {
      hash[1 + i] = keys[i];
      hash[1 + L + i] = off;
      off += sizes[i];
    }
  } else {
    for (long s : sizes) off += s;
  }
  if (size) *size = off;
  return L;
}

long t1_offset(const Tiles &T, long *hash, long *size) {
  return build_hash([&](auto &&g) { each_t1_block(T, [&](long k, long p, long h) { g(k, p, h, 0, 0); }); },
                    hash, size, T, 2);
}
long t2_offset(const Tiles &T, long *hash, long *size) {
  return build_hash([&](auto &&g) { each_t2_block(T, g); }, hash, size, T, 4);
}
long v2_offset(const Tiles &T, long *hash, long *size) {
  return build_hash([&](auto &&g) { each_v2_block(T, g); }, hash, size, T, 4);
}

long y1_offset(const Tiles &T, long *hash, long *size) {
  return build_hash([&](auto &&g) { each_y1_block(T, g); }, hash, size, T, 2);
}
long y2_offset(const Tiles &T, long *hash, long *size) {
  return build_hash([&](auto &&g) { each_y2_block(T, g); }, hash, size, T, 4);
}
long f1_offset(const Tiles &T, long *hash, long *size) {
  return build_hash([&](auto &&g) { each_f1_block(T, g); }, hash, size, T, 2);
}

long hash_lookup(const long *hash, long key) {
  // Bisection while the window holds more than five entries, then a scan:
  // the probe sequence of tce_hash.F:271-316.
  long L = hash[0];
  if (L <= 0) return -1;
  long lo = 1, hi = L, mid;
  for (;;) {
    if (hi - lo <= 4) {
      mid = -1;
      for (long i = lo; i <= hi; ++i)
        if (hash[i] == key) mid = i;
      if (mid < 0) return -1;
      break;
    }
    mid = (lo + hi) / 2;
    if (hash[mid] == key) break;
    if (hash[mid] > key)
      hi = mid;
    else
      lo = mid;
  }
  return hash[L + mid];
}

// ------------------------------------------------------------ generator --

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).
void philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) {
      k0 += W0;
      k1 += W1;
    }
    uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    c0 = hi1 ^ c1 ^ k0;
    c1 = lo1;
    c2 = hi0 ^ c3 ^ k1;
    c3 = lo0;
  }
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

static inline void draw(uint64_t seed, uint32_t cls, uint64_t idx, double &u1, double &u2) {
  uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), cls, 0u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t o[4];
  philox4x32(ctr, key, o);
  const double inv53 = 1.0 / 9007199254740992.0;
  u1 = ((double)(o[0] >> 5) * 67108864.0 + (double)(o[1] >> 6)) * inv53;  // [0,1)
  u2 = ((double)(o[2] >> 5) * 67108864.0 + (double)(o[3] >> 6)) * inv53;
}

double gen_uniform(uint64_t seed, uint32_t cls, uint64_t idx) {
  double u1, u2;
  draw(seed, cls, idx, u1, u2);
  return u1;
}

double gen_normal(uint64_t seed, uint32_t cls, uint64_t idx) {
  double u1, u2;
  draw(seed, cls, idx, u1, u2);
  const double two_pi = 6.283185307179586476925286766559;
  return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(two_pi * u2);
}

// Element functions on global indices: holes h in [0,no), particles p in
// [0,nv).  Antisymmetrisation by exact differences keeps every element an
// exact negative of its transposes.
struct Synth {
  uint64_t seed;
  long no, nv;
  static constexpr double kSigT1 = 0.01, kSigT2 = 0.02, kSigV = 0.05;
  static constexpr double kRsqrt2 = 0.70710678118654752440;
  double t1(long p, long h) const { return kSigT1 * gen_normal(seed, kClsT1, (uint64_t)p * no + h); }
  double x2(long p, long q, long h, long g) const {
    return kSigT2 * gen_normal(seed, kClsT2, (((uint64_t)p * nv + q) * no + h) * no + g);
  }
  double t2(long p, long q, long h, long g) const {
    double u1 = x2(p, q, h, g) - x2(q, p, h, g);
    double u2 = x2(p, q, g, h) - x2(q, p, g, h);
    return (u1 - u2) * 0.5;
  }
  double y1(long m, long p, long h, long g) const {  // <m p || h g>
    return kSigV * gen_normal(seed, kClsVd1, (((uint64_t)m * nv + p) * no + h) * no + g);
  }
  double w(long m, long p, long h, long g) const { return (y1(m, p, h, g) - y1(m, p, g, h)) * kRsqrt2; }
  double z2(long p, long q, long h, long e) const {  // <p q || h e>
    return kSigV * gen_normal(seed, kClsVd2, (((uint64_t)p * nv + q) * no + h) * nv + e);
  }
  double v(long p, long q, long h, long e) const { return (z2(p, q, h, e) - z2(q, p, h, e)) * kRsqrt2; }
  double zs(long p, long q, long h, long g) const {  // <p q || h g>
    return kSigV * gen_normal(seed, kClsVs1, (((uint64_t)p * nv + q) * no + h) * no + g);
  }
  double s(long p, long q, long h, long g) const {
    double u1 = zs(p, q, h, g) - zs(q, p, h, g);
    double u2 = zs(p, q, g, h) - zs(q, p, g, h);
    return (u1 - u2) * 0.5;
  }
};

void synth_evl(uint64_t seed, long no, long nv, double *evl) {
  std::vector<double> o(no), v(nv);
  for (long i = 0; i < no; ++i) o[i] = -2.0 + 1.7 * gen_uniform(seed, kClsEvlO, i);
  for (long a = 0; a < nv; ++a) v[a] = 0.2 + 3.8 * gen_uniform(seed, kClsEvlV, a);
  std::sort(o.begin(), o.end());
  std::sort(v.begin(), v.end());
  std::copy(o.begin(), o.end(), evl);
  std::copy(v.begin(), v.end(), evl + no);
}

template <class F> static void parallel_for(long n, F &&f) {
  unsigned nt = std::max(1u, std::min(64u, std::thread::hardware_concurrency()));
  if (n < 64 || nt == 1) {
    for (long i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  std::atomic<long> next(0);
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&] {
      for (long i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto &x : th) x.join();
}

int synth_blocks(const Tiles &T, uint64_t seed, double *t1, const long *t1h, double *t2,
                 const long *t2h, double *v2, const long *v2h) {
  Synth G{seed, T.no, T.nv};
  long no = T.no;
  // Collect blocks first, then fill them in parallel.
  struct Blk {
    int cls;
    long key, a, b, c, d;
  };
  std::vector<Blk> blocks;
  if (t1)
    each_t1_block(T, [&](long k, long p, long h) { blocks.push_back({0, k, p, h, 0, 0}); });
  if (t2) each_t2_block(T, [&](long k, long a, long b, long c, long d) { blocks.push_back({1, k, a, b, c, d}); });
  if (v2) each_v2_block(T, [&](long k, long a, long b, long c, long d) { blocks.push_back({2, k, a, b, c, d}); });
  std::atomic<int> bad(0);
  parallel_for((long)blocks.size(), [&](long bi) {
    const Blk &B = blocks[bi];
    if (B.cls == 0) {
      long off = hash_lookup(t1h, B.key);
      if (off < 0) { bad = 1; return; }
      long rp = T.rng(B.a), rh = T.rng(B.b), op = T.ofs(B.a) - no, oh = T.ofs(B.b);
      double *dst = t1 + off;
      for (long p = 0; p < rp; ++p)
        for (long h = 0; h < rh; ++h) dst[p * rh + h] = G.t1(op + p, oh + h);
    } else if (B.cls == 1) {
      long off = hash_lookup(t2h, B.key);
      if (off < 0) { bad = 1; return; }
      long r1 = T.rng(B.a), r2 = T.rng(B.b), r3 = T.rng(B.c), r4 = T.rng(B.d);
      long o1 = T.ofs(B.a) - no, o2 = T.ofs(B.b) - no, o3 = T.ofs(B.c), o4 = T.ofs(B.d);
      double *dst = t2 + off;
      for (long i = 0; i < r1; ++i)
        for (long j = 0; j < r2; ++j)
          for (long k = 0; k < r3; ++k)
            for (long l = 0; l < r4; ++l)
              *dst++ = G.t2(o1 + i, o2 + j, o3 + k, o4 + l);
    } else {
      long off = hash_lookup(v2h, B.key);
      if (off < 0) { bad = 1; return; }
      long r1 = T.rng(B.a), r2 = T.rng(B.b), r3 = T.rng(B.c), r4 = T.rng(B.d);
      double *dst = v2 + off;
      bool p3 = B.a > T.noab, p1 = B.c > T.noab, p2 = B.d > T.noab;
      long o1 = T.ofs(B.a) - (p3 ? no : 0), o2 = T.ofs(B.b) - no, o3 = T.ofs(B.c) - (p1 ? no : 0),
           o4 = T.ofs(B.d) - (p2 ? no : 0);
      for (long i = 0; i < r1; ++i)
        for (long j = 0; j < r2; ++j)
          for (long k = 0; k < r3; ++k)
            for (long l = 0; l < r4; ++l) {
              double val;
              if (!p3)  // (h p | h h): <m p || h g>
                val = G.w(o1 + i, o2 + j, o3 + k, o4 + l);
              else if (p2)  // (p p | h p): <p q || h e>
                val = G.v(o1 + i, o2 + j, o3 + k, o4 + l);
              else  // (p p | h h)
                val = G.s(o1 + i, o2 + j, o3 + k, o4 + l);
              *dst++ = val;
            }
    }
  });
  return bad ? TCE_T_ERR_KEY : TCE_T_OK;
}

int synth_dense(uint64_t seed, long no, long nv, double *T1, double *T2, double *W, double *V,
                double *S) {
  Synth G{seed, no, nv};
  if (T1)
    for (long p = 0; p < nv; ++p)
      for (long h = 0; h < no; ++h) T1[p * no + h] = G.t1(p, h);
  if (T2)
    parallel_for(nv, [&](long p) {
      for (long q = 0; q < nv; ++q)
        for (long h = 0; h < no; ++h)
          for (long g = 0; g < no; ++g) T2[((p * nv + q) * no + h) * no + g] = G.t2(p, q, h, g);
    });
  if (W)
    parallel_for(no, [&](long m) {
      for (long p = 0; p < nv; ++p)
        for (long h = 0; h < no; ++h)
          for (long g = 0; g < no; ++g) W[((m * nv + p) * no + h) * no + g] = G.w(m, p, h, g);
    });
  if (V)
    parallel_for(nv, [&](long p) {
      for (long q = 0; q < nv; ++q)
        for (long h = 0; h < no; ++h)
          for (long e = 0; e < nv; ++e) V[((p * nv + q) * no + h) * nv + e] = G.v(p, q, h, e);
    });
  if (S)
    parallel_for(nv, [&](long p) {
      for (long q = 0; q < nv; ++q)
        for (long h = 0; h < no; ++h)
          for (long g = 0; g < no; ++g) S[((p * nv + q) * no + h) * no + g] = G.s(p, q, h, g);
    });
  return TCE_T_OK;
}

}  // namespace tce_t

// ------------------------------------------------------------- C-ABI ---------
using namespace tce_t;

extern "C" long tce_t_tile_ranges(long n, long tilesize, long *ranges) {
  return tile_ranges(n, tilesize, ranges);
}

#define WITH_TILES(t, fail)              \
  Tiles T;                               \
  if (!T.load(t)) {                      \
    set_error("invalid tile description"); \
    return fail;                         \
  }

extern "C" long tce_t_t1_offset(const tce_t_tiles *t, long *hash, long *size) {
  WITH_TILES(t, -1);
  return t1_offset(T, hash, size);
}
extern "C" long tce_t_t2_offset(const tce_t_tiles *t, long *hash, long *size) {
  WITH_TILES(t, -1);
  return t2_offset(T, hash, size);
}
extern "C" long tce_t_v2_offset(const tce_t_tiles *t, long *hash, long *size) {
  WITH_TILES(t, -1);
  return v2_offset(T, hash, size);
}
extern "C" long tce_t_y1_offset(const tce_t_tiles *t, long *hash, long *size) {
  WITH_TILES(t, -1);
  return y1_offset(T, hash, size);
}
extern "C" long tce_t_y2_offset(const tce_t_tiles *t, long *hash, long *size) {
  WITH_TILES(t, -1);
  return y2_offset(T, hash, size);
}
extern "C" long tce_t_f1_offset(const tce_t_tiles *t, long *hash, long *size) {
  WITH_TILES(t, -1);
  return f1_offset(T, hash, size);
}
extern "C" long tce_t_hash(const long *hash, long key) { return hash ? hash_lookup(hash, key) : -1; }

extern "C" int tce_t_synth_evl(unsigned long long seed, long no, long nv, double *evl) {
  if (no <= 0 || nv <= 0 || !evl) return TCE_T_ERR_ARG;
  synth_evl(seed, no, nv, evl);
  return TCE_T_OK;
}
extern "C" int tce_t_synth_blocks(const tce_t_tiles *t, unsigned long long seed, double *t1,
                                  const long *t1h, double *t2, const long *t2h, double *v2,
                                  const long *v2h) {
  WITH_TILES(t, TCE_T_ERR_ARG);
  if ((t1 && !t1h) || (t2 && !t2h) || (v2 && !v2h)) return TCE_T_ERR_ARG;
  return synth_blocks(T, seed, t1, t1h, t2, t2h, v2, v2h);
}
extern "C" int tce_t_synth_dense(unsigned long long seed, long no, long nv, double *T1,
                                 double *T2, double *W, double *V, double *S) {
  if (no <= 0 || nv <= 0) return TCE_T_ERR_ARG;
  return synth_dense(seed, no, nv, T1, T2, W, V, S);
}
extern "C" void tce_t_philox4x32(const unsigned int ctr[4], const unsigned int key[2],
                                 unsigned int out[4]) {
  philox4x32(ctr, key, out);
}
