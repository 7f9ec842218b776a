// internal.h -- shared declarations of libtce_t.so (not part of the C-ABI).
#pragma once

#include "../../include/tce_t.h"

#include <atomic>
#include <cstdint>
#include <string>
#include <vector>

namespace tce_t {

// Philox counter classes of the synthetic generator (one per tensor).
enum : uint32_t {
  kClsEvlO = 1,
  kClsEvlV = 2,
  kClsT1 = 3,
  kClsT2 = 4,
  kClsVd1 = 5,
  kClsVd2 = 6,
  kClsVs1 = 7,
};

// Host copy of the tile model (tce.fh / MA arrays), tiles numbered from 1.
struct Tiles {
  long noab = 0, nvab = 0;
  long no = 0, nv = 0, norb = 0;  // occupied / virtual / all orbitals
  bool restricted = false;
  std::vector<long> range, offset, spin, sym, alpha;
  std::vector<double> evl;
  bool load(const tce_t_tiles *t);
  long rng(long b) const { return range[b - 1]; }
  long ofs(long b) const { return offset[b - 1]; }
  long spn(long b) const { return spin[b - 1]; }
  long sy(long b) const { return sym[b - 1]; }
  long alp(long b) const { return alpha[b - 1]; }
};

long tile_ranges(long n, long tilesize, long *ranges);
long t1_offset(const Tiles &T, long *hash, long *size);
long t2_offset(const Tiles &T, long *hash, long *size);
long v2_offset(const Tiles &T, long *hash, long *size);
long y1_offset(const Tiles &T, long *hash, long *size);
long y2_offset(const Tiles &T, long *hash, long *size);
long f1_offset(const Tiles &T, long *hash, long *size);
long hash_lookup(const long *hash, long key);

void philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double gen_uniform(uint64_t seed, uint32_t cls, uint64_t idx);
double gen_normal(uint64_t seed, uint32_t cls, uint64_t idx);
void synth_evl(uint64_t seed, long no, long nv, double *evl);
int synth_blocks(const Tiles &T, uint64_t seed, double *t1, const long *t1h, double *t2,
                 const long *t2h, double *v2, const long *v2h);
int synth_dense(uint64_t seed, long no, long nv, double *T1, double *T2, double *W, double *V,
                double *S);

void set_error(const std::string &msg);

// Task factor of ccsd_t_gpu.F:153-167 / ccsd_t_dot.F:51-65.
inline double task_factor(bool restricted, long p4b, long p5b, long p6b, long h1b, long h2b,
                          long h3b) {
  double f = restricted ? 2.0 : 1.0;
  if (p4b == p5b && p5b == p6b)
    f = f / 6.0;
  else if (p4b == p5b || p5b == p6b)
    f = f / 2.0;
  if (h1b == h2b && h2b == h3b)
    f = f / 6.0;
  else if (h1b == h2b || h2b == h3b)
    f = f / 2.0;
  return f;
}

// Task filter of ccsd_t_gpu.F:94-112.
// Precision modes of the d1/d2 contractions (TCE_T_FP64 .. TCE_T_BF16X6)
// and the number of operand parts of a reduced-precision copy.
constexpr int NPREC = 4;
inline bool valid_precision(int p) { return p >= TCE_T_FP64 && p <= TCE_T_BF16X6; }
inline int mixed_parts(int p) { return p == TCE_T_FP32 ? 1 : (p == TCE_T_BF16X3 ? 2 : 3); }

inline bool task_allowed(const Tiles &T, long p4b, long p5b, long p6b, long h1b, long h2b,
                         long h3b) {
  long sp = T.spn(p4b) + T.spn(p5b) + T.spn(p6b), sh = T.spn(h1b) + T.spn(h2b) + T.spn(h3b);
  if (sp != sh) return false;
  if (T.restricted && sp + sh > 8) return false;
  if ((T.sy(p4b) ^ T.sy(p5b) ^ T.sy(p6b) ^ T.sy(h1b) ^ T.sy(h2b) ^ T.sy(h3b)) != 0) return false;
  return true;
}

// Spin and irrep conservation of the block with tiles (a, b, c, d)
// (tce_t2_offset.F, tce_mo2e_offset.F): stored blocks satisfy it.
inline bool allowed4(const Tiles &T, long a, long b, long c, long d) {
  return T.spn(a) + T.spn(b) == T.spn(c) + T.spn(d) && (T.sy(a) ^ T.sy(b) ^ T.sy(c) ^ T.sy(d)) == 0;
}

// tce_restricted_4 (tce_restricted.F:36-71): in a restricted run an all-beta
// block is read from its alpha image.
inline void restrict4(const Tiles &T, long &a, long &b, long &c, long &d) {
  if (T.restricted && T.spn(a) + T.spn(b) + T.spn(c) + T.spn(d) == 8) {
    a = T.alp(a);
    b = T.alp(b);
    c = T.alp(c);
    d = T.alp(d);
  }
}

}  // namespace tce_t
