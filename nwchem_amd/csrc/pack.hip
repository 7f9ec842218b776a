// pack.hip -- the upload's block fetch and sort on the device.
//
// The reference fetches every canonical TCE block a task needs through the
// offset hash tables (get_hash_block.F:1-44, tce_hash.F:271-316), maps
// all-beta blocks to their alpha images in restricted runs
// (tce_restricted.F:1-71) and sorts each block into the kernel's index order
// with a sign per swapped index pair (tce_hashnsort.F:47-64,129-144,
// sort/tce_sort.F:15-125, the sorts of ccsd_t_doubles_gpu.F:282-304,942-976).
// The job does that once per upload, for every block, into the dense HBM
// layouts the fused kernels read (T1, T2e, T2m, Wd1, Vd2, S: job.hip).
//
// Here the host only resolves each tile tuple of the five tensors to the
// offset of the canonical block to read (hash lookup after the tile order
// and restricted mapping; -1 where the spin/irrep rules leave the block out;
// a block the rules require but the table lacks is an error, as tce_hash's
// errquit): tables of one entry per tile tuple.  The byte ranges of the
// three files that those blocks cover are copied to the device once (runs of
// adjacent blocks, so a file holding other integral classes costs nothing
// extra), and one gather kernel per tensor writes its dense layout: one
// thread per dense element, tile and in-tile position of each orbital index
// from small maps, index order and sign from the order of the tiles.  The
// values are those of the host packer (pack_dense in job.hip) bit for bit:
// the same element of the same block times +-1.
//
// Bound: HBM.  Dense bytes written (14.5 GiB at no=40/nv=320) plus the
// canonical bytes read (7.1 GB, each read once per dense copy it feeds:
// T2 twice, the rest once); a few ms, against seconds for the host scatter.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#include "job.h"

namespace tce_t {

namespace {

// Orbital maps: tile (1-based, holes first) and position inside the tile of
// orbital o (holes 0..no-1, particles no..norb-1); tile ranges by tile.
struct PackGeo {
  const int *tile, *loc, *rng;
  int no, nv, noab, nvab;
};

constexpr int PACK_THREADS = 256;
constexpr int PACK_BLOCKS = 4096;

__device__ __forceinline__ long gid() { return (long)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ long gstride() { return (long)gridDim.x * blockDim.x; }

// T1[p][h] = t(p h), block (pb, hb) in C order [p][h].
__global__ void __launch_bounds__(PACK_THREADS)
    pack_t1(PackGeo G, const long *tab, const double *src, double *out) {
  const long n = (long)G.nv * G.no;
  for (long x = gid(); x < n; x += gstride()) {
    const int p = (int)(x / G.no), h = (int)(x % G.no);
    const int pb = G.tile[G.no + p], hb = G.tile[h];
    const long off = tab[(long)(pb - G.noab - 1) * G.noab + (hb - 1)];
    out[x] = off < 0 ? 0.0 : src[off + (long)G.loc[G.no + p] * G.rng[hb] + G.loc[h]];
  }
}

// t(p e h g) for the tiles of p, e, h, g: canonical block of the sorted
// tile pairs, sign -1 per swapped pair.
__device__ __forceinline__ double t2_value(const PackGeo &G, const long *tab, const double *src,
                                           int p, int e, int h, int g) {
  const int P1 = G.tile[G.no + p], P2 = G.tile[G.no + e], H1 = G.tile[h], H2 = G.tile[g];
  const long off = tab[(((long)(P1 - G.noab - 1) * G.nvab + (P2 - G.noab - 1)) * G.noab + (H1 - 1)) *
                           G.noab +
                       (H2 - 1)];
  if (off < 0) return 0.0;
  const int i = G.loc[G.no + p], j = G.loc[G.no + e], k = G.loc[h], l = G.loc[g];
  const bool sp = P1 > P2, sh = H1 > H2;
  const long cb = sp ? G.rng[P1] : G.rng[P2], cc = sh ? G.rng[H2] : G.rng[H1],
             cd = sh ? G.rng[H1] : G.rng[H2];
  const long ia = sp ? j : i, ib = sp ? i : j, ic = sh ? l : k, id = sh ? k : l;
  const double v = src[off + ((ia * cb + ib) * cc + ic) * cd + id];
  return (sp != sh) ? -v : v;
}

// T2e[h][g][p][e] (e padded to Ep) = t(p e h g)
__global__ void __launch_bounds__(PACK_THREADS)
    pack_t2e(PackGeo G, long Ep, const long *tab, const double *src, double *out) {
  const long n = (long)G.no * G.no * G.nv * Ep;
  for (long x = gid(); x < n; x += gstride()) {
    const int e = (int)(x % Ep);
    long r = x / Ep;
    const int p = (int)(r % G.nv);
    r /= G.nv;
    const int g = (int)(r % G.no), h = (int)(r / G.no);
    out[x] = e < G.nv ? t2_value(G, tab, src, p, e, h, g) : 0.0;
  }
}

// T2m[h][p][e][m] (m padded to Mp) = t(p e h m)
__global__ void __launch_bounds__(PACK_THREADS)
    pack_t2m(PackGeo G, long Mp, const long *tab, const double *src, double *out) {
  const long n = (long)G.no * G.nv * G.nv * Mp;
  for (long x = gid(); x < n; x += gstride()) {
    const int m = (int)(x % Mp);
    long r = x / Mp;
    const int e = (int)(r % G.nv);
    r /= G.nv;
    const int p = (int)(r % G.nv), h = (int)(r / G.nv);
    out[x] = m < G.no ? t2_value(G, tab, src, p, e, h, m) : 0.0;
  }
}

// Wd1[h][g][p][m] (m padded to Mp) = <m p||h g> from v(H7 P6 H2 H3) with the
// hole pair (H2, H3) sorted, sign -1 when swapped.  PM: the block holds
// [P6][H7][H2][H3] instead, as cr_ccsd_t_N.F:749-752 leaves the moment
// intermediate i1(h11 p4 h1 h2).
template <bool PM>
__global__ void __launch_bounds__(PACK_THREADS)
    pack_wd1(PackGeo G, long Mp, const long *tab, const double *src, double *out) {
  const long n = (long)G.no * G.no * G.nv * Mp;
  for (long x = gid(); x < n; x += gstride()) {
    const int m = (int)(x % Mp);
    if (m >= G.no) {
      out[x] = 0.0;
      continue;
    }
    long r = x / Mp;
    const int p = (int)(r % G.nv);
    r /= G.nv;
    const int g = (int)(r % G.no), h = (int)(r / G.no);
    const int H7 = G.tile[m], P6 = G.tile[G.no + p], H2 = G.tile[h], H3 = G.tile[g];
    const long off = tab[(((long)(H7 - 1) * G.nvab + (P6 - G.noab - 1)) * G.noab + (H2 - 1)) *
                             G.noab +
                         (H3 - 1)];
    if (off < 0) {
      out[x] = 0.0;
      continue;
    }
    const bool sw = H2 > H3;
    const long r6 = G.rng[P6], r7 = G.rng[H7], r2 = G.rng[H2], r3 = G.rng[H3];
    const long lm = G.loc[m], lp = G.loc[G.no + p], lh = G.loc[h], lg = G.loc[g];
    const long ih = sw ? lg : lh, ig = sw ? lh : lg, c2 = sw ? r2 : r3;
    const long pair = PM ? lp * r7 + lm : lm * r6 + lp;
    const double v = src[off + (pair * (sw ? r3 : r2) + ih) * c2 + ig];
    out[x] = sw ? -v : v;
  }
}

// Vd2[w][p][q][e] (e padded to Ep) = <p q||w e> from v(P5 P6 H3 P7) with the
// particle pair (P5, P6) sorted, sign -1 when swapped.
__global__ void __launch_bounds__(PACK_THREADS)
    pack_vd2(PackGeo G, long Ep, const long *tab, const double *src, double *out) {
  const long n = (long)G.no * G.nv * G.nv * Ep;
  for (long x = gid(); x < n; x += gstride()) {
    const int e = (int)(x % Ep);
    if (e >= G.nv) {
      out[x] = 0.0;
      continue;
    }
    long r = x / Ep;
    const int q = (int)(r % G.nv);
    r /= G.nv;
    const int p = (int)(r % G.nv), w = (int)(r / G.nv);
    const int P5 = G.tile[G.no + p], P6 = G.tile[G.no + q], H3 = G.tile[w], P7 = G.tile[G.no + e];
    const long off = tab[(((long)(P5 - G.noab - 1) * G.nvab + (P6 - G.noab - 1)) * G.noab + (H3 - 1)) *
                             G.nvab +
                         (P7 - G.noab - 1)];
    if (off < 0) {
      out[x] = 0.0;
      continue;
    }
    const bool sw = P5 > P6;
    const long cb = sw ? G.rng[P5] : G.rng[P6], r3 = G.rng[H3], r7 = G.rng[P7];
    const long i = G.loc[G.no + p], j = G.loc[G.no + q];
    const long ia = sw ? j : i, ib = sw ? i : j;
    const double v = src[off + ((ia * cb + ib) * r3 + G.loc[w]) * r7 + G.loc[G.no + e]];
    out[x] = sw ? -v : v;
  }
}

// S[h][g][p][q] = <p q||h g> from v(P5 P6 H2 H3), both pairs sorted.
__global__ void __launch_bounds__(PACK_THREADS)
    pack_s(PackGeo G, const long *tab, const double *src, double *out) {
  const long n = (long)G.no * G.no * G.nv * G.nv;
  for (long x = gid(); x < n; x += gstride()) {
    const int q = (int)(x % G.nv);
    long r = x / G.nv;
    const int p = (int)(r % G.nv);
    r /= G.nv;
    const int g = (int)(r % G.no), h = (int)(r / G.no);
    const int P5 = G.tile[G.no + p], P6 = G.tile[G.no + q], H2 = G.tile[h], H3 = G.tile[g];
    const long off = tab[(((long)(P5 - G.noab - 1) * G.nvab + (P6 - G.noab - 1)) * G.noab + (H2 - 1)) *
                             G.noab +
                         (H3 - 1)];
    if (off < 0) {
      out[x] = 0.0;
      continue;
    }
    const bool sp = P5 > P6, sh = H2 > H3;
    const long r5 = G.rng[P5], r6 = G.rng[P6], r2 = G.rng[H2], r3 = G.rng[H3];
    const long i = G.loc[G.no + p], j = G.loc[G.no + q], lh = G.loc[h], lg = G.loc[g];
    const long cb = sp ? r5 : r6, cd = sh ? r2 : r3;
    const long ia = sp ? j : i, ib = sp ? i : j, ic = sh ? lg : lh, id = sh ? lh : lg;
    const double v = src[off + ((ia * cb + ib) * (sh ? r3 : r2) + ic) * cd + id];
    out[x] = (sp != sh) ? -v : v;
  }
}

// S of a CR-CCSD(T) job: S[h][g][p][q] = t(p q h g) + (t(p h) t(q g) -
// t(p g) t(q h)) / 3 from the t2 blocks and the packed T1[p][h], so that the
// job's singles P(9) t(a i) S(b c j k) are the denominator tile of
// cr_ccsd_t_E.F:7-9, P(9) t(p4 p5 h1 h2) t(p6 h3) - 2/3 P(9) t(p4 h1)
// i1(p5 p6 h2 h3) with i1 = -1/4 P(4) t(p4 h1) t(p5 h2) (the P(9) sums
// agree after moving the single t to the front: even permutations).
__global__ void __launch_bounds__(PACK_THREADS)
    pack_s_cr(PackGeo G, const long *tab, const double *src, const double *T1, double *out) {
  const long n = (long)G.no * G.no * G.nv * G.nv;
  for (long x = gid(); x < n; x += gstride()) {
    const int q = (int)(x % G.nv);
    long r = x / G.nv;
    const int p = (int)(r % G.nv);
    r /= G.nv;
    const int g = (int)(r % G.no), h = (int)(r / G.no);
    const double ph = T1[(long)p * G.no + h], qg = T1[(long)q * G.no + g],
                 pg = T1[(long)p * G.no + g], qh = T1[(long)q * G.no + h];
    out[x] = t2_value(G, tab, src, p, q, h, g) + (ph * qg - pg * qh) / 3.0;
  }
}

// Block tables of one upload: per tensor, the file offset of the canonical
// block behind each tile tuple (-1: none), and the extent (doubles) that the
// gather reads from it.
struct Table {
  std::vector<long> off, len;
  void add(long o, long n) {
    off.push_back(o);
    len.push_back(o < 0 ? 0 : n);
  }
};

// Adjacent or overlapping [off, off + len) ranges of one file merged into
// runs; compact offset of every referenced block in the device copy.
struct Runs {
  std::vector<long> start, end, base;  // file range [start, end) at device offset base
  long total = 0;
  void build(const std::vector<const Table *> &tabs) {
    std::vector<std::pair<long, long>> iv;
    for (const Table *t : tabs)
      for (size_t i = 0; i < t->off.size(); ++i)
        if (t->off[i] >= 0 && t->len[i] > 0) iv.push_back({t->off[i], t->off[i] + t->len[i]});
    std::sort(iv.begin(), iv.end());
    for (const auto &r : iv) {
      if (!end.empty() && r.first <= end.back()) {
        end.back() = std::max(end.back(), r.second);
      } else {
        start.push_back(r.first);
        end.push_back(r.second);
      }
    }
    for (size_t k = 0; k < start.size(); ++k) {
      base.push_back(total);
      total += end[k] - start[k];
    }
  }
  long remap(long off) const {
    const size_t k = std::upper_bound(start.begin(), start.end(), off) - start.begin() - 1;
    return base[k] + (off - start[k]);
  }
};

}  // namespace

// Device upload of the dense tensors (tce_t_job_upload): tables on the host,
// referenced file ranges to the device, gathers on the job's stream.  The
// dense buffers J.dT1 .. J.dS must be allocated; returns a status.
int pack_on_device(Job &J, const double *t1, const long *t1h, const double *t2, const long *t2h,
                   const double *v2, const long *v2h, int kind, const double *x2,
                   const long *x2h) {
  const bool cr = kind == PACK_CR;
  const Tiles &T = J.T;
  const long no = T.no, nv = T.nv, noab = T.noab, nvab = T.nvab, N = noab + nvab;
  bool missing = false;
  auto look = [&](const long *hash, long key, bool need) {
    const long o = hash_lookup(hash, key);
    if (o < 0 && need) missing = true;
    return o;
  };
  // T1: (pb, hb); all-beta pairs read from the alpha image when restricted
  Table t1tab;
  for (long pb = noab + 1; pb <= N; ++pb)
    for (long hb = 1; hb <= noab; ++hb) {
      long p1 = pb, h1 = hb;
      if (T.restricted && T.spn(pb) + T.spn(hb) == 4) p1 = T.alp(pb), h1 = T.alp(hb);
      const bool need = T.spn(pb) == T.spn(hb) && (T.sy(pb) ^ T.sy(hb)) == 0;
      t1tab.add(look(t1h, h1 - 1 + noab * (p1 - noab - 1), need), T.rng(pb) * T.rng(hb));
    }
  // T2: (P1, P2, H1, H2), both pairs sorted for the key
  Table t2tab;
  for (long P1 = noab + 1; P1 <= N; ++P1)
    for (long P2 = noab + 1; P2 <= N; ++P2)
      for (long H1 = 1; H1 <= noab; ++H1)
        for (long H2 = 1; H2 <= noab; ++H2) {
          long a = std::min(P1, P2), b = std::max(P1, P2), c = std::min(H1, H2), d = std::max(H1, H2);
          const bool need = allowed4(T, a, b, c, d);
          restrict4(T, a, b, c, d);
          const long key = d - 1 + noab * (c - 1 + noab * (b - noab - 1 + nvab * (a - noab - 1)));
          t2tab.add(look(t2h, key, need), T.rng(P1) * T.rng(P2) * T.rng(H1) * T.rng(H2));
        }
  auto vkey = [&](long g3, long g4, long g1, long g2) {
    return g2 - 1 + N * (g1 - 1 + N * (g4 - 1 + N * (g3 - 1)));
  };
  // Wd1: v(H7 P6 H2 H3), (H2, H3) sorted
  Table wtab;
  for (long H7 = 1; H7 <= noab; ++H7)
    for (long P6 = noab + 1; P6 <= N; ++P6)
      for (long H2 = 1; H2 <= noab; ++H2)
        for (long H3 = 1; H3 <= noab; ++H3) {
          long g3 = H7, g4 = P6, g1 = std::min(H2, H3), g2 = std::max(H2, H3);
          const bool need = allowed4(T, g3, g4, g1, g2);
          restrict4(T, g3, g4, g1, g2);
          // CR: i1(h11 p4 h1 h2) keyed as OFFSET_cr_ccsd_t_N_1_1 (cr_ccsd_t_N.F:800-830)
          const long key = cr ? g2 - 1 + noab * (g1 - 1 + noab * (g3 - 1 + noab * (g4 - noab - 1)))
                              : vkey(g3, g4, g1, g2);
          wtab.add(look(v2h, key, need), T.rng(H7) * T.rng(P6) * T.rng(H2) * T.rng(H3));
        }
  // Vd2: v(P5 P6 H3 P7), (P5, P6) sorted
  Table vtab;
  for (long P5 = noab + 1; P5 <= N; ++P5)
    for (long P6 = noab + 1; P6 <= N; ++P6)
      for (long H3 = 1; H3 <= noab; ++H3)
        for (long P7 = noab + 1; P7 <= N; ++P7) {
          long g3 = std::min(P5, P6), g4 = std::max(P5, P6), g1 = H3, g2 = P7;
          const bool need = allowed4(T, g3, g4, g1, g2);
          restrict4(T, g3, g4, g1, g2);
          // CR: i1(p4 p5 h1 p12) keyed as OFFSET_cr_ccsd_t_N_2_1, same block order
          // [P5][P6][H3][P7] (cr_ccsd_t_N.F:3990-3996)
          if (cr)
            vtab.add(look(x2h, g2 - noab - 1 + nvab * (g1 - 1 + noab * (g4 - noab - 1 + nvab * (g3 - noab - 1))),
                          need),
                     T.rng(P5) * T.rng(P6) * T.rng(H3) * T.rng(P7));
          else
            vtab.add(look(v2h, vkey(g3, g4, g1, g2), need), T.rng(P5) * T.rng(P6) * T.rng(H3) * T.rng(P7));
        }
  // S: v(P5 P6 H2 H3), both pairs sorted (CR: from t2 and t1, pack_s_cr)
  Table stab;
  for (long P5 = noab + 1; P5 <= (cr ? noab : N); ++P5)
    for (long P6 = noab + 1; P6 <= N; ++P6)
      for (long H2 = 1; H2 <= noab; ++H2)
        for (long H3 = 1; H3 <= noab; ++H3) {
          long g3 = std::min(P5, P6), g4 = std::max(P5, P6), g1 = std::min(H2, H3), g2 = std::max(H2, H3);
          const bool need = allowed4(T, g3, g4, g1, g2);
          restrict4(T, g3, g4, g1, g2);
          stab.add(look(v2h, vkey(g3, g4, g1, g2), need), T.rng(P5) * T.rng(P6) * T.rng(H2) * T.rng(H3));
        }
  if (missing) {
    set_error("a block needed by the (T) path is missing from a hash table");
    return TCE_T_ERR_KEY;
  }

  // the referenced ranges of each file, compacted on the device
  // (CR: Vd2's blocks from the second intermediate, file x)
  Runs r1, r2, rv, rx;
  r1.build({&t1tab});
  r2.build({&t2tab});
  if (cr) {
    rv.build({&wtab});
    rx.build({&vtab});
  } else {
    rv.build({&wtab, &vtab, &stab});
  }
  for (long &o : t1tab.off)
    if (o >= 0) o = r1.remap(o);
  for (long &o : t2tab.off)
    if (o >= 0) o = r2.remap(o);
  for (Table *t : {&wtab, &vtab, &stab})
    for (long &o : t->off)
      if (o >= 0) o = (cr && t == &vtab) ? rx.remap(o) : rv.remap(o);

  // one device allocation: maps, tables, then the three compacted files
  std::vector<int> tile(T.norb), loc(T.norb), rng(N + 1, 0);
  for (long b = 1; b <= N; ++b) {
    rng[b] = (int)T.rng(b);
    for (long k = 0; k < T.rng(b); ++k) {
      tile[T.ofs(b) + k] = (int)b;
      loc[T.ofs(b) + k] = (int)k;
    }
  }
  auto al = [](size_t n) { return (n + 255) / 256 * 256; };
  const size_t b_maps = al((2 * T.norb + N + 1) * sizeof(int));
  const Table *tabs[5] = {&t1tab, &t2tab, &wtab, &vtab, &stab};
  size_t b_tabs = 0, tab_at[5];
  for (int k = 0; k < 5; ++k) {
    tab_at[k] = b_maps + b_tabs;
    b_tabs += al(std::max<size_t>(1, tabs[k]->off.size()) * sizeof(long));
  }
  const size_t f1_at = b_maps + b_tabs, f2_at = f1_at + al(r1.total * 8),
               fv_at = f2_at + al(r2.total * 8), fx_at = fv_at + al(rv.total * 8),
               bytes = fx_at + al(rx.total * 8);
  std::vector<char> head(f1_at, 0);
  std::memcpy(head.data(), tile.data(), T.norb * sizeof(int));
  std::memcpy(head.data() + T.norb * sizeof(int), loc.data(), T.norb * sizeof(int));
  std::memcpy(head.data() + 2 * T.norb * sizeof(int), rng.data(), (N + 1) * sizeof(int));
  for (int k = 0; k < 5; ++k)
    std::memcpy(head.data() + tab_at[k], tabs[k]->off.data(), tabs[k]->off.size() * sizeof(long));
  char *d = nullptr;
  HIPCHK(hipMalloc(&d, std::max<size_t>(bytes, 256)), "upload staging");
  int rc = TCE_T_OK;
  // The referenced ranges of a file form one stream of R.total doubles (the
  // runs in order, at device offsets base[k]); it goes to the device through
  // two pinned buffers: host threads fill one while the DMA of the other runs.
  constexpr long CH = 8L << 20;  // doubles per staging buffer (64 MiB)
  double *stage = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (!hip_ok(hipHostMalloc(&stage, 2 * CH * sizeof(double)), "upload staging") ||
      !hip_ok(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming), "event") ||
      !hip_ok(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming), "event"))
    rc = TCE_T_ERR_HIP;
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  long piece = 0;  // staging buffers used so far (alternating)
  // copy stream positions [p0, p0 + n) of file runs R into dst
  auto gather = [](const Runs &R, const double *file, long p0, long n, double *dst) {
    size_t k = std::upper_bound(R.base.begin(), R.base.end(), p0) - R.base.begin() - 1;
    while (n > 0) {
      const long in = p0 - R.base[k], len = std::min(n, R.end[k] - R.start[k] - in);
      std::memcpy(dst, file + R.start[k] + in, len * sizeof(double));
      dst += len;
      p0 += len;
      n -= len;
      ++k;
    }
  };
  auto stream_file = [&](const Runs &R, const double *file, size_t at) {
    for (long p0 = 0; p0 < R.total && rc == TCE_T_OK; p0 += CH, ++piece) {
      const long n = std::min(CH, R.total - p0);
      double *buf = stage + (piece & 1) * CH;
      // the DMA that last read this buffer must be done
      if (piece >= 2 && !hip_ok(hipEventSynchronize(ev[piece & 1]), "upload blocks")) {
        rc = TCE_T_ERR_HIP;
        break;
      }
      std::vector<std::thread> th;
      const long part = (n + nt - 1) / nt;
      for (unsigned t = 1; t < nt && (long)t * part < n; ++t)
        th.emplace_back(gather, std::cref(R), file, p0 + (long)t * part,
                        std::min(n, (long)(t + 1) * part) - (long)t * part, buf + (long)t * part);
      gather(R, file, p0, std::min(n, part), buf);
      for (auto &x : th) x.join();
      if (!hip_ok(hipMemcpyAsync(d + at + p0 * 8, buf, n * sizeof(double), hipMemcpyHostToDevice,
                                 J.stream),
                  "upload blocks") ||
          !hip_ok(hipEventRecord(ev[piece & 1], J.stream), "event"))
        rc = TCE_T_ERR_HIP;
    }
  };
  if (rc == TCE_T_OK &&
      !hip_ok(hipMemcpyAsync(d, head.data(), head.size(), hipMemcpyHostToDevice, J.stream),
              "upload tables"))
    rc = TCE_T_ERR_HIP;
  stream_file(r1, t1, f1_at);
  stream_file(r2, t2, f2_at);
  stream_file(rv, v2, fv_at);
  if (cr) stream_file(rx, x2, fx_at);
  if (rc == TCE_T_OK) {
    PackGeo G;
    G.tile = (const int *)d;
    G.loc = (const int *)(d + T.norb * sizeof(int));
    G.rng = (const int *)(d + 2 * T.norb * sizeof(int));
    G.no = (int)no;
    G.nv = (int)nv;
    G.noab = (int)noab;
    G.nvab = (int)nvab;
    const long *tb[5];
    for (int k = 0; k < 5; ++k) tb[k] = (const long *)(d + tab_at[k]);
    const double *f1 = (const double *)(d + f1_at), *f2 = (const double *)(d + f2_at),
                 *fv = (const double *)(d + fv_at), *fx = (const double *)(d + fx_at);
    const dim3 grid(PACK_BLOCKS), blk(PACK_THREADS);
    hipLaunchKernelGGL(pack_t1, grid, blk, 0, J.stream, G, tb[0], f1, J.dT1);
    hipLaunchKernelGGL(pack_t2e, grid, blk, 0, J.stream, G, J.Ep, tb[1], f2, J.dT2e);
    hipLaunchKernelGGL(pack_t2m, grid, blk, 0, J.stream, G, J.Mp, tb[1], f2, J.dT2m);
    if (cr) {
      hipLaunchKernelGGL(pack_wd1<true>, grid, blk, 0, J.stream, G, J.Mp, tb[2], fv, J.dWd1);
      hipLaunchKernelGGL(pack_vd2, grid, blk, 0, J.stream, G, J.Ep, tb[3], fx, J.dVd2);
      hipLaunchKernelGGL(pack_s_cr, grid, blk, 0, J.stream, G, tb[1], f2, J.dT1, J.dS);
    } else {
      hipLaunchKernelGGL(pack_wd1<false>, grid, blk, 0, J.stream, G, J.Mp, tb[2], fv, J.dWd1);
      hipLaunchKernelGGL(pack_vd2, grid, blk, 0, J.stream, G, J.Ep, tb[3], fv, J.dVd2);
      hipLaunchKernelGGL(pack_s, grid, blk, 0, J.stream, G, tb[4], fv, J.dS);
    }
    if (!hip_ok(hipGetLastError(), "pack launch")) rc = TCE_T_ERR_HIP;
  }
  // the copies read the staging buffers (and `head`) until they complete
  if (!hip_ok(hipStreamSynchronize(J.stream), "pack")) rc = TCE_T_ERR_HIP;
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  if (stage) (void)hipHostFree(stage);
  (void)hipFree(d);
  return rc;
}

}  // namespace tce_t
