// fused_mixed.hip -- the fused (T) kernel on 16x16 MFMA tiles: the d1/d2
// contractions in FP64 or in reduced precision; the denominators, their
// division, the singles term and every energy sum in FP64.
//
//   TCE_T_FP64    v_mfma_f64_16x16x4_f64, FP64 operands and accumulation
//                 (the default FP64 kernel of 8-wide tilings); its D layout
//                 puts row i in lane group i%4, register i/4, so its X rows
//                 are ordered 4r + L' (the roles of the two fields swapped)
//   TCE_T_FP32    v_mfma_f32_16x16x4_f32, FP32 operands and accumulation
//   TCE_T_BF16X3  v_mfma_f32_16x16x32_bf16, each operand x split into
//                 hi = bf16(x), lo = bf16(x - hi); per product hi*hi + hi*lo
//                 + lo*hi (3 terms), FP32 accumulation
//   TCE_T_BF16X6  the same MFMA, x = hi + mid + lo (three bf16, 24 bits);
//                 per product the 6 terms down to 2^-24 relative (hi*hi,
//                 hi*mid, mid*hi, hi*lo, mid*mid, lo*hi), FP32 accumulation;
//                 two LDS chunk buffers (three parts of 16 KB per chunk)
//
// Work item and workgroup as fused_wg.hip: one task x an 8x8x8 particle box x
// a 4x4x4 hole box, 512 threads (8 waves, 2 per SIMD); wave w owns the 4x4x4
// particle cube (w>>2&1, w>>1&1, w&1) by the hole box, 4096 t3 elements in
// 64 FP32 accumulators per lane.
//
// MFMA lane layout (16x16, probed on the device): A[i][k] at lane i+16(k/E),
// B[k][j] at lane j+16(k/E) (E = k values per lane: 1 for f32, 8 for bf16),
// D[i][j] at lane j+16(i/4), register i%4.  Every d1/d2 contraction is
// X(L, r, s; e) * Y(p1, p2, h; e): lone particle L with the hole pair (r, s)
// on one operand, the particle pair (p1, p2) with the lone hole h on the
// other.  With L fixed, row i = 4L' + r and col j = 4p1' + p2', so
//   lane = 4 p1' + p2' + 16 L', accumulator quad (s, h), component r.
// The lane keeps element (i, j, k) of its hole box at acc[16i + 4j + k]; a
// pass only chooses which 4 of them form each MFMA quad (r = i for lone hole
// j or k, r = j for lone hole i), so the three contractions of one L share
// the lane layout and changing L (a -> b -> c) is one lane permutation.
// Pass order: L=a {h=k, h=j, h=i} x {d2, d1}, L=b, L=c (18 passes).
//
// Operands stream through LDS in chunks of KC contraction values (8 f64, 16
// f32 or 32 bf16 = four 16-byte quads per row) with three buffers: direct
// global->LDS loads (global_load_lds_dwordx4) fetch chunk j+2 (j+3 in the
// FP32 loop) while the waves run chunk j; the workgroup waits for its own
// loads of chunk j+1 and synchronises once per chunk (pass_step).  A pass
// with a minus sign negates the accumulators before and after it.
//
// LDS layout of one buffer part (hi or lo; quads, kq = quad 0..3 of a row):
//   X  [4 s][8 L slots][4 r][4 kq ^ ((slot & 3) ^ (r & 2))]         512 quads
//   Y  [4 h][8 p1 slots][8 p2 slots][4 kq ^ ((p1 & 3) ^ (p2 & 2))] 1024 quads
// Fragment reads (16 lanes of one kq: 4 row/col slots x 4 inner indices)
// touch 16 distinct 4-bank groups; each load instruction writes 64
// consecutive quads and 4 consecutive lanes read one 64-byte row segment.
#include <type_traits>

#include "fused.h"

namespace tce_t {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));

template <int PREC>
struct Prec;
template <>
struct Prec<TCE_T_FP64> {
  static constexpr int KC = 8;     // contraction values per chunk
  static constexpr int NPART = 1;  // operand parts
  static constexpr int ES = 8;     // bytes per element
  // LDS chunk buffers (24 KB each), chunks loaded NBUF - 1 ahead.  At no=40/nv=320
  // (no=40, nv=320), kernel ms: 3 buffers 8283, 4 buffers 8176, 5 buffers 8251.
  static constexpr int NBUF = 4;
  static constexpr bool SPLIT = false;  // chunk loop: see pass_step
  typedef double acc;
};
template <>
struct Prec<TCE_T_FP32> {
  static constexpr int KC = 16;
  static constexpr int NPART = 1;
  static constexpr int ES = 4;
  static constexpr int NBUF = 3;  // kernel ms at no=40/nv=320: 3 buffers 4310, 4 buffers 4349
  static constexpr bool SPLIT = true;  // measured 0.7% faster than the direct loop
  typedef float acc;
};
template <>
struct Prec<TCE_T_BF16X3> {
  static constexpr int KC = 32;
  static constexpr int NPART = 2;  // hi, lo
  static constexpr int ES = 2;
  static constexpr int NBUF = 3;
  static constexpr bool SPLIT = false;  // measured 20% slower split (64 fragment VGPRs)
  typedef float acc;
};
template <>
struct Prec<TCE_T_BF16X6> {
  static constexpr int KC = 32;
  static constexpr int NPART = 3;  // hi, mid, lo
  static constexpr int ES = 2;
  static constexpr int NBUF = 2;   // 3 x 48 KB do not fit the 160 KB of LDS
  static constexpr bool SPLIT = false;
  typedef float acc;
};
typedef double d4 __attribute__((ext_vector_type(4)));
typedef double dv2 __attribute__((ext_vector_type(2)));

constexpr int X_QUADS = 512, Y_QUADS = 1024, PART_QUADS = X_QUADS + Y_QUADS;

// Epilogue staging (doubles), overlaying the chunk buffers after the last
// chunk:
//   S   [hole pair t: (j,k) (i,k) (i,j)][16 pairs u*4+v]
//       [particle pair s: (b,c) (a,c) (a,b)][8 x 8 slots]           9216
//   T1  [particle box a b c][hole box i j k][8 slots][4 holes]        288
constexpr int EPI_S = 3 * 16 * 3 * 64;
constexpr int EPI_T1 = 9 * 32;
constexpr int EPI_QUADS = (EPI_S + EPI_T1) / 2;

template <int PREC>
constexpr int buf_quads() {
  return Prec<PREC>::NPART * PART_QUADS;
}
template <int PREC>
constexpr int lds_quads() {
  constexpr int nb = Prec<PREC>::NBUF;
  return nb * buf_quads<PREC>() > EPI_QUADS ? nb * buf_quads<PREC>() : EPI_QUADS;
}

// The 18 passes (fused.h, MIXED_PASS): kind (0 = d2, 1 = d1), lone particle
// L (a b c), lone hole h (i j k), sign.  d2: X = t(L e r s) (T2e), Y =
// <p1 p2||h e> (Vd2); d1: X = <m L||r s> (Wd1), Y = t(p1 p2 h m) (T2m);
// signs and pairs as the variant table in fused.h (ccsd_t_kernels_omp.F
// layouts).
typedef PassDef PassInfo;
#define PASS MIXED_PASS

// Hole box of an item: global hole o + min(b0 + u, rm), u = 0..3.
struct HoleBox {
  int o, b0, rm;
};
__device__ __forceinline__ int hole(const HoleBox &h, int u) { return h.o + min(h.b0 + u, h.rm); }
// Particle range of the box: global index g0 + min(slot, sm), slot = 0..7.
struct PRange {
  long g0;
  int sm;
};
struct ItemGeo {
  HoleBox I, J, K;
  PRange a, b, c;
};
template <int T>
__device__ __forceinline__ int hole_t(const ItemGeo &G, int u) {
  if constexpr (T == 0) return hole(G.I, u);
  else if constexpr (T == 1) return hole(G.J, u);
  else return hole(G.K, u);
}
template <int T>
__device__ __forceinline__ long particle_t(const ItemGeo &G, int slot) {
  if constexpr (T == 0) return G.a.g0 + min(slot, G.a.sm);
  else if constexpr (T == 1) return G.b.g0 + min(slot, G.b.sm);
  else return G.c.g0 + min(slot, G.c.sm);
}
// Per-lane selections (ternaries on scalars).
__device__ __forceinline__ int hole_of(const ItemGeo &G, int t, int u) {
  return t == 0 ? hole(G.I, u) : (t == 1 ? hole(G.J, u) : hole(G.K, u));
}
__device__ __forceinline__ long particle_of(const ItemGeo &G, int t, int slot) {
  return t == 0 ? G.a.g0 + min(slot, G.a.sm)
                : (t == 1 ? G.b.g0 + min(slot, G.b.sm) : G.c.g0 + min(slot, G.c.sm));
}
// This wave's cube coordinate along particle x (0 = a, 1 = b, 2 = c).
__device__ __forceinline__ int cube(int wave, int x) { return (wave >> (2 - x)) & 1; }

// Quad swizzles of the operand rows: LDS position k of a row holds the row's
// global quad k ^ swz, for both the loads (which write each row linearly)
// and the fragment reads.  Measured with gfx950's SQ_LDS_BANK_CONFLICT over
// every swizzle that is linear in the row bits: a ds_read_b128 whose 16
// lanes of one kq read rows (slot or p1 from lane bits 2-3, r or p2 from
// bits 0-1) is conflict-free with s(row) = -(slot & 3) mod 4 (FP32 and bf16
// X reads, every Y read), while (slot & 3) ^ (r & 2) took two passes (4
// conflict cycles per read).  The FP64 X reads put slot on lane bits 0-1 and
// r on bits 2-3, for which (slot & 3) ^ (r & 2) is conflict-free.
template <int PREC>
__device__ __forceinline__ int x_swz(int slot, int r) {
  return PREC == TCE_T_FP64 ? (slot & 3) ^ (r & 2) : (4 - (slot & 3)) & 3;
}
__device__ __forceinline__ int y_swz(int p1, int) { return (4 - (p1 & 3)) & 3; }

// Direct global->LDS load of 16 bytes per lane: lane l writes LDS bytes
// [m0 + 16 l, m0 + 16 l + 16).  Issued by inline assembly so that the
// compiler does not tie later LDS reads of other buffers to it; the kernel
// waits for these loads itself (vmcnt) before the barrier that publishes a
// buffer.
__device__ __forceinline__ void lds_load16(const void *g, unsigned lds_byte) {
  asm volatile("s_mov_b32 m0, %0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(lds_byte), "v"(g)
               : "memory");
}
// s_waitcnt vmcnt(N) with the other counters left alone (gfx9 encoding).
template <int N>
__device__ __forceinline__ void wait_vm() {
  __builtin_amdgcn_s_waitcnt((N & 15) | ((N >> 4) << 14) | (7 << 4) | (15 << 8));
}
__device__ __forceinline__ void block_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// One thread's load sources of a pass, per operand part: the X quad of LDS
// block `wave`, the Y quads of blocks 2w and 2w+1 (byte pointers at k = 0).
template <int PREC>
struct Src {
  const char *x[Prec<PREC>::NPART], *y0[Prec<PREC>::NPART], *y1[Prec<PREC>::NPART];
};

template <int PREC, int P>
__device__ __forceinline__ Src<PREC> pass_src(const MixedArgsG &M, const ItemGeo &G, int wave,
                                              int lane) {
  using Pr = Prec<PREC>;
  constexpr PassInfo pi = PASS[P];
  constexpr int L = pi.L, h = pi.h;
  constexpr int r = pair_r(h), s = pair_s(h), p1 = L == 0 ? 1 : 0, p2 = L == 2 ? 1 : 2;
  constexpr int QE = 16 / Pr::ES;  // elements per quad
  const int no = M.f.no, nv = M.f.nv;
  const long Kp = pi.kind ? M.Mk : M.Ek;
  Src<PREC> S;
  // X block `wave`: s-index wave>>1, slots 4*(wave&1) + lane>>4
  const int xs = wave >> 1, slot = 4 * (wave & 1) + (lane >> 4), xr = (lane >> 2) & 3;
  const long xe = (((long)hole_t<r>(G, xr) * no + hole_t<s>(G, xs)) * nv + particle_t<L>(G, slot)) *
                      Kp + QE * ((lane & 3) ^ x_swz<PREC>(slot, xr));
  // Y blocks 2w+t: h-index bq>>2, p1 slots 2*(bq&3) + lane>>5
  long ye[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int bq = 2 * wave + t;
    const int yh = bq >> 2, p1s = 2 * (bq & 3) + (lane >> 5), p2s = (lane >> 2) & 7;
    const long row = particle_t<p1>(G, p1s) * nv + particle_t<p2>(G, p2s);
    ye[t] = ((long)hole_t<h>(G, yh) * nv * nv + row) * Kp + QE * ((lane & 3) ^ y_swz(p1s, p2s));
  }
#pragma unroll
  for (int part = 0; part < Pr::NPART; ++part) {
    const char *X = (const char *)(pi.kind ? M.x1[part] : M.x2[part]);
    const char *Y = (const char *)(pi.kind ? M.y1[part] : M.y2[part]);
    S.x[part] = X + xe * Pr::ES;
    S.y0[part] = Y + ye[0] * Pr::ES;
    S.y1[part] = Y + ye[1] * Pr::ES;
  }
  return S;
}

// Issue the loads of one chunk (byte offset kb within the rows) into buffer b.
template <int PREC>
__device__ __forceinline__ void load_chunk(const Src<PREC> &S, long kb, unsigned lds_base, int b,
                                           int wave) {
#pragma unroll
  for (int part = 0; part < Prec<PREC>::NPART; ++part) {
    const unsigned base = lds_base + (unsigned)(b * buf_quads<PREC>() + part * PART_QUADS) * 16u;
    lds_load16(S.x[part] + kb, __builtin_amdgcn_readfirstlane(base + wave * 1024u));
    lds_load16(S.y0[part] + kb,
               __builtin_amdgcn_readfirstlane(base + X_QUADS * 16u + 2 * wave * 1024u));
    lds_load16(S.y1[part] + kb,
               __builtin_amdgcn_readfirstlane(base + X_QUADS * 16u + (2 * wave + 1) * 1024u));
  }
}

// Accumulator index of quad (s, h), component r for row hole RH (0 = i,
// 1 = j) and lone hole H: element (i, j, k) at acc[16 i + 4 j + k].
template <int H>
__device__ __forceinline__ constexpr int acc_at(int s, int h, int r) {
  return H == 0 ? 16 * h + 4 * r + s      // r = j, s = k, h = i
                : (H == 2 ? 16 * r + 4 * s + h   // r = i, s = j, h = k
                          : 16 * r + 4 * h + s);  // r = i, s = k, h = j
}

// Operand fragments of one chunk in registers: X quads s = 0..3 and Y quads
// h = 0..3 of this lane (bf16: hi and lo parts).
template <int PREC>
struct Frag;
template <>
struct Frag<TCE_T_FP64> {
  dv2 x[4], y[4];
};
template <>
struct Frag<TCE_T_FP32> {
  f4 x[4], y[4];
};
template <>
struct Frag<TCE_T_BF16X3> {
  bf8 x[4], y[4], xl[4], yl[4];
};
template <>
struct Frag<TCE_T_BF16X6> {
  bf8 x[4], y[4], xm[4], ym[4], xl[4], yl[4];
};

// X quads S0..S1-1 (offset xa for s = 0, s adds 128 quads) and the Y quads
// (offset yb for h = 0, h adds 256) of the chunk in buffer cb.
template <int PREC, int S0, int S1>
__device__ __forceinline__ void read_x(Frag<PREC> &F, const f4 *cb, int xa) {
#pragma unroll
  for (int s = S0; s < S1; ++s) {
    if constexpr (PREC == TCE_T_FP64) {
      F.x[s] = reinterpret_cast<const dv2 *>(cb)[xa + 128 * s];
    } else if constexpr (PREC == TCE_T_FP32) {
      F.x[s] = cb[xa + 128 * s];
    } else if constexpr (PREC == TCE_T_BF16X3) {
      F.x[s] = reinterpret_cast<const bf8 *>(cb)[xa + 128 * s];
      F.xl[s] = reinterpret_cast<const bf8 *>(cb)[PART_QUADS + xa + 128 * s];
    } else {
      F.x[s] = reinterpret_cast<const bf8 *>(cb)[xa + 128 * s];
      F.xm[s] = reinterpret_cast<const bf8 *>(cb)[PART_QUADS + xa + 128 * s];
      F.xl[s] = reinterpret_cast<const bf8 *>(cb)[2 * PART_QUADS + xa + 128 * s];
    }
  }
}
template <int PREC>
__device__ __forceinline__ void read_y(Frag<PREC> &F, const f4 *cb, int yb) {
#pragma unroll
  for (int h = 0; h < 4; ++h) {
    if constexpr (PREC == TCE_T_FP64) {
      F.y[h] = reinterpret_cast<const dv2 *>(cb)[X_QUADS + yb + 256 * h];
    } else if constexpr (PREC == TCE_T_FP32) {
      F.y[h] = cb[X_QUADS + yb + 256 * h];
    } else if constexpr (PREC == TCE_T_BF16X3) {
      F.y[h] = reinterpret_cast<const bf8 *>(cb)[X_QUADS + yb + 256 * h];
      F.yl[h] = reinterpret_cast<const bf8 *>(cb)[PART_QUADS + X_QUADS + yb + 256 * h];
    } else {
      F.y[h] = reinterpret_cast<const bf8 *>(cb)[X_QUADS + yb + 256 * h];
      F.ym[h] = reinterpret_cast<const bf8 *>(cb)[PART_QUADS + X_QUADS + yb + 256 * h];
      F.yl[h] = reinterpret_cast<const bf8 *>(cb)[2 * PART_QUADS + X_QUADS + yb + 256 * h];
    }
  }
}

// FP64: the MFMAs of X quads S0..S1-1 with every Y quad (lone hole H) on
// the scalar accumulators, each quad gathered for its MFMA (the compiler
// keeps the FP64 chunk loops free of moves and spills this way; the quad
// form below spills there).
template <int H, int S0, int S1>
__device__ __forceinline__ void mma_x64(double (&a)[64], const Frag<TCE_T_FP64> &F) {
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int s = S0; s < S1; ++s)
#pragma unroll
      for (int h = 0; h < 4; ++h) {
        d4 c = {a[acc_at<H>(s, h, 0)], a[acc_at<H>(s, h, 1)], a[acc_at<H>(s, h, 2)],
                a[acc_at<H>(s, h, 3)]};
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(F.x[s][t], F.y[h][t], c, 0, 0, 0);
#pragma unroll
        for (int r = 0; r < 4; ++r) a[acc_at<H>(s, h, r)] = c[r];
      }
}

// FP32 and split-BF16: accumulator quads of a pass, q[4 s + h] holding
// components r = 0..3 of MFMA quad (s, h), i.e. acc[acc_at<H>(s, h, r)].
// Inside a pass's chunk loop the 16 quads are the MFMA C/D operands
// themselves (4 consecutive registers each); the scattered acc[] order of
// the other passes is restored only at the pass boundary, so the chunk loop
// needs no register moves (before: 65-136 moves per bf16 chunk).
template <int PREC>
struct AccQ;
template <>
struct AccQ<TCE_T_FP64> {
  typedef double type[4][4];  // unused: FP64 keeps the scalar accumulators
};
template <>
struct AccQ<TCE_T_FP32> {
  typedef f4 type;
};
template <>
struct AccQ<TCE_T_BF16X3> {
  typedef f4 type;
};
template <>
struct AccQ<TCE_T_BF16X6> {
  typedef f4 type;
};

template <int PREC, int H>
__device__ __forceinline__ void to_quads(typename AccQ<PREC>::type (&q)[16],
                                         const typename Prec<PREC>::acc (&a)[64]) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int h = 0; h < 4; ++h)
#pragma unroll
      for (int r = 0; r < 4; ++r) q[4 * s + h][r] = a[acc_at<H>(s, h, r)];
}
template <int PREC, int H>
__device__ __forceinline__ void from_quads(typename Prec<PREC>::acc (&a)[64],
                                           const typename AccQ<PREC>::type (&q)[16]) {
#pragma unroll
  for (int s = 0; s < 4; ++s)
#pragma unroll
    for (int h = 0; h < 4; ++h)
#pragma unroll
      for (int r = 0; r < 4; ++r) a[acc_at<H>(s, h, r)] = q[4 * s + h][r];
}

// The MFMAs of X quads S0..S1-1 with every Y quad.  Split-BF16 issues the
// three terms of all 16 quads term by term (lo*hi', hi*lo', hi*hi' for each
// quad, in that order as before), so dependent MFMAs on one quad are 16
// instructions apart.
template <int PREC, int S0, int S1>
__device__ __forceinline__ void mma_q(typename AccQ<PREC>::type (&q)[16], const Frag<PREC> &F) {
  if constexpr (PREC == TCE_T_BF16X3) {
#pragma unroll
    for (int term = 0; term < 3; ++term)
#pragma unroll
      for (int s = S0; s < S1; ++s)
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          const bf8 xa = term == 0 ? F.xl[s] : F.x[s];
          const bf8 yb = term == 1 ? F.yl[h] : F.y[h];
          q[4 * s + h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa, yb, q[4 * s + h], 0, 0, 0);
        }
  } else if constexpr (PREC == TCE_T_BF16X6) {
    // smallest terms first: lo*hi', hi*lo', mid*mid', mid*hi', hi*mid', hi*hi'
#pragma unroll
    for (int term = 0; term < 6; ++term)
#pragma unroll
      for (int s = S0; s < S1; ++s)
#pragma unroll
        for (int h = 0; h < 4; ++h) {
          const bf8 xa = term == 0 ? F.xl[s] : ((term == 2 || term == 3) ? F.xm[s] : F.x[s]);
          const bf8 yb = term == 1 ? F.yl[h] : ((term == 2 || term == 4) ? F.ym[h] : F.y[h]);
          q[4 * s + h] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(xa, yb, q[4 * s + h], 0, 0, 0);
        }
  } else {
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int s = S0; s < S1; ++s)
#pragma unroll
        for (int h = 0; h < 4; ++h)
          q[4 * s + h] = __builtin_amdgcn_mfma_f32_16x16x4f32(F.x[s][t], F.y[h][t], q[4 * s + h], 0, 0, 0);
  }
}

// This lane's fragment offsets in pass P: X quad (s = 0), Y quad (h = 0).
// X row (lane & 15) = 4L' + r (L' = hi, r = lo), quad kq.  The f64 MFMA
// writes row i to lane group i % 4, register i / 4 (the f32 and bf16 ones
// to group i / 4, register i % 4), so for FP64 the rows are ordered 4r + L'
// instead: either way the result lands at lane 16L' + col, component r.
// Y col (lane & 15) = 4p1' + p2' (p1' = hi, p2' = lo).
template <int PREC, int P>
__device__ __forceinline__ int2 frag_offsets(int wave, int lane) {
  constexpr PassInfo pi = PASS[P];
  constexpr int Lp = pi.L, p1 = Lp == 0 ? 1 : 0, p2 = Lp == 2 ? 1 : 2;
  const int kq = lane >> 4, hi = (lane & 15) >> 2, lo = lane & 3;
  const int xL = PREC == TCE_T_FP64 ? lo : hi, xr = PREC == TCE_T_FP64 ? hi : lo;
  const int slot = 4 * cube(wave, Lp) + xL;
  const int p1s = 4 * cube(wave, p1) + hi, p2s = 4 * cube(wave, p2) + lo;
  return make_int2((slot * 4 + xr) * 4 + (kq ^ x_swz<PREC>(slot, xr)),
                   (p1s * 8 + p2s) * 4 + (kq ^ y_swz(p1s, p2s)));
}

__device__ __forceinline__ void permute_lanes(float (&a)[64], int src) {
#pragma unroll
  for (int r = 0; r < 64; ++r)
    a[r] = __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(a[r])));
}
__device__ __forceinline__ void permute_lanes(double (&a)[64], int src) {
#pragma unroll
  for (int r = 0; r < 64; ++r) a[r] = perm_lane(a[r], src);
}

template <class T>
__device__ __forceinline__ void negate(T (&a)[64]) {
#pragma unroll
  for (int r = 0; r < 64; ++r) a[r] = -a[r];
}

// Chunks of pass p of the item's task: the chunks covering the task's
// contraction range of the pass (every operand outside it is zero), first
// chunk c0, at least one (a pass without allowed orbitals then contracts
// a chunk of zeros).
struct PassChunks {
  int c0, n;
};
// SP = false (jobs without spin or irrep labels: every range is the whole
// space) takes the padded extents and never reads the ranges.
template <int PREC, bool SP>
__device__ __forceinline__ PassChunks chunks_of(const MixedArgsG &M, gptr<const TaskDev> T, int p) {
  constexpr int KC = Prec<PREC>::KC;
  if (p >= 18) return PassChunks{0, 0};
  if constexpr (!SP) {
    return PassChunks{0, (PASS[p].kind ? M.Mk : M.Ek) / KC};
  } else {
    const int c0 = T->x0[p] / KC, c1 = (T->x1[p] + KC - 1) / KC;
    return PassChunks{c0, c1 > c0 ? c1 - c0 : 1};
  }
}

// Load global chunk `c2` of the pass sequence that starts at pass P (chunk
// c2 of pass P, or of one of the next D passes when P has fewer chunks;
// every pass has at least one) into buffer nb.
template <int PREC, bool SP, int P, int D>
__device__ __forceinline__ void load_ahead(const Src<PREC> &s0, const MixedArgsG &M,
                                           gptr<const TaskDev> T, const ItemGeo &G, int wave,
                                           int lane, int c2, unsigned lds_base, int nb) {
  using Pr = Prec<PREC>;
  const PassChunks pc = chunks_of<PREC, SP>(M, T, P);
  if (c2 < pc.n) {
    load_chunk<PREC>(s0, (long)(pc.c0 + c2) * Pr::KC * Pr::ES, lds_base, nb, wave);
  } else if constexpr (D > 0 && P + 1 < 18) {
    load_ahead<PREC, SP, P + 1, D - 1>(pass_src<PREC, P + 1>(M, G, wave, lane), M, T, G, wave,
                                       lane, c2 - pc.n, lds_base, nb);
  }
}

// s_waitcnt lgkmcnt(0) with the other counters left alone (gfx9 encoding).
__device__ __forceinline__ void wait_lgkm0() { __builtin_amdgcn_s_waitcnt(0xC07F); }

// Wait until at most min(p, MAXP) chunks of NL loads each are outstanding
// (p <= 0: none); the chunks complete in issue order.
template <int NL, int MAXP>
__device__ __forceinline__ void wait_pending(int p) {
  if constexpr (MAXP >= 3) {
    if (p >= 3) {
      wait_vm<3 * NL>();
      return;
    }
  }
  if constexpr (MAXP >= 2) {
    if (p >= 2) {
      wait_vm<2 * NL>();
      return;
    }
  }
  if (p >= 1)
    wait_vm<NL>();
  else
    wait_vm<0>();
}

// Pass P: layout change, then its chunks, with three LDS chunk buffers.
//
// SPLIT (FP32): on entry F holds the fragments of the pass's
// first chunk j.  Per chunk: the MFMAs of X quads 0-1; this thread's loads
// of chunk j+1 and its LDS reads of chunk j complete, then the barrier that
// publishes chunk j+1 and frees buffer j % NBUF, which receives the loads of
// chunk j+NBUF; the fragments of X quads 0-1 of chunk j+1 are read while the
// MFMAs of X quads 2-3 of chunk j run, then the rest of chunk j+1's, so the
// matrix core never waits for an LDS read after the barrier.
//
// Otherwise (FP64, whose 64 double accumulators leave no room to keep a
// chunk's fragments live across the barrier, and split-BF16, whose hi and lo
// fragments take 64 registers): per chunk the fragments of chunk j are read
// and its MFMAs run, with the loads of chunk j+NBUF-1 (into the buffer freed
// by the last barrier) issued halfway through them; then this thread's
// loads of chunk j+1 complete and the barrier publishes it.
template <int PREC, bool SP, int P>
__device__ __forceinline__ void pass_step(typename Prec<PREC>::acc (&acc)[64], const f4 *lds,
                                          unsigned lds_base, const MixedArgsG &M,
                                          gptr<const TaskDev> T, const ItemGeo &G, int wave,
                                          int lane, int &j, int &buf, int total,
                                          Frag<PREC> &F) {
  using Pr = Prec<PREC>;
  constexpr PassInfo pi = PASS[P];
  constexpr int H = pi.h;
  constexpr int NLOAD = 3 * Pr::NPART;  // load instructions per chunk
  const int kq = lane >> 4, hi = (lane & 15) >> 2, lo = lane & 3;
  if constexpr (P == 6) {  // L: a -> b; lane (4b'+c') + 16a' -> (4a'+c') + 16b'
    permute_lanes(acc, 4 * kq + lo + 16 * hi);
  } else if constexpr (P == 12) {  // L: b -> c; lane (4a'+c') + 16b' -> (4a'+b') + 16c'
    permute_lanes(acc, 4 * hi + kq + 16 * lo);
  }
  if constexpr (pi.minus) negate(acc);
  typename AccQ<PREC>::type q[16];
  if constexpr (PREC != TCE_T_FP64) to_quads<PREC, H>(q, acc);
  // the MFMAs of X quads S0..S1-1 of fragments f on this pass's accumulators
  auto mma = [&](auto s0_, auto s1_, const Frag<PREC> &f) {
    constexpr int S0 = decltype(s0_)::value, S1 = decltype(s1_)::value;
    if constexpr (PREC == TCE_T_FP64)
      mma_x64<H, S0, S1>(acc, f);
    else
      mma_q<PREC, S0, S1>(q, f);
  };
  using I0 = std::integral_constant<int, 0>;
  using I2 = std::integral_constant<int, 2>;
  using I4 = std::integral_constant<int, 4>;
  const int nch0 = chunks_of<PREC, SP>(M, T, P).n;
  if constexpr (Pr::SPLIT) {
    // load sources of this pass; those of the next passes are formed only in
    // the branches that need them (the pass's last chunks), which keeps the
    // chunk loop's live registers to the accumulators, fragments and s0
    const Src<PREC> s0 = pass_src<PREC, P>(M, G, wave, lane);
    constexpr int NB = Pr::NBUF;
    for (int c = 0; c < nch0; ++c, ++j) {
      const bool next = j + 1 < total;
      mma(I0(), I2(), F);
      // chunk j+1 landed; chunks j+2 .. j+NB-1 may be pending
      wait_pending<NLOAD, NB - 2>(total - j - 2);
      wait_lgkm0();
      block_barrier();
      // every wave has read chunk j: its buffer takes chunk j+NB
      if (j + NB < total) load_ahead<PREC, SP, P, NB>(s0, M, T, G, wave, lane, c + NB, lds_base, buf);
      const int nbuf = buf + 1 < NB ? buf + 1 : 0;
      const f4 *cbn = lds + nbuf * buf_quads<PREC>();
      // chunk j+1: this pass or the first chunk of pass P+1
      const int2 o = c + 1 < nch0 ? frag_offsets<PREC, P>(wave, lane)
                                  : frag_offsets<PREC, (P + 1 < 18 ? P + 1 : 17)>(wave, lane);
      if (next) read_x<PREC, 0, 2>(F, cbn, o.x);
      mma(I2(), I4(), F);
      if (next) {
        read_x<PREC, 2, 4>(F, cbn, o.x);
        read_y<PREC>(F, cbn, o.y);
      }
      buf = nbuf;
    }
  } else {
    // Split-BF16: the lane's LDS fragment offsets (and, with two buffers, the
    // load sources) are formed again in every chunk from a lane index the
    // compiler cannot see through.  Kept live across the loop they were
    // spilled, and a reload's s_waitcnt also drains the chunk loads in flight
    // (vmcnt counts them in issue order).  Kernel ms at no=40/nv=320, kept live /
    // formed per chunk (two boxes each): bf16x6 2865-2915 / 2736-2823, bf16x3
    // 1598-1604 / 1589-1600, FP64 8200 / 8293 (FP64 keeps them live; a few
    // of its passes reload them).
    constexpr bool per_chunk = PREC != TCE_T_FP64;
    constexpr bool src_per_chunk = per_chunk && Pr::NBUF < 3;
    const Src<PREC> s0 = pass_src<PREC, P>(M, G, wave, lane);
    const int2 o0 = frag_offsets<PREC, P>(wave, lane);
    for (int c = 0; c < nch0; ++c, ++j) {
      int ln = lane;
      if constexpr (per_chunk) asm volatile("" : "+v"(ln));
      const int2 o = per_chunk ? frag_offsets<PREC, P>(wave, ln) : o0;
      const bool more = j + 2 < total;
      Frag<PREC> f;
      const f4 *cb = lds + buf * buf_quads<PREC>();
      read_x<PREC, 0, 4>(f, cb, o.x);
      read_y<PREC>(f, cb, o.y);
      if constexpr (Pr::NBUF >= 3) {
        // NBUF buffers, AH = NBUF - 1 chunks ahead: the loads of chunk j+AH
        // (buffer (j + AH) % NBUF, which held chunk j-1 and is free since the
        // last barrier) go between the two halves of the MFMAs, so their
        // address arithmetic issues while the matrix core works (three
        // buffers: 0.7% faster than before the MFMAs, 3% faster than after
        // them; four buffers, FP64 at no=40/nv=320: 8200 ms here, 8414 before)
        constexpr int AH = Pr::NBUF - 1;
        mma(I0(), I2(), f);
        if (j + AH < total) {
          const int nb = buf + AH < Pr::NBUF ? buf + AH : buf + AH - Pr::NBUF;
          if constexpr (src_per_chunk)
            load_ahead<PREC, SP, P, AH>(pass_src<PREC, P>(M, G, wave, ln), M, T, G, wave, ln,
                                        c + AH, lds_base, nb);
          else
            load_ahead<PREC, SP, P, AH>(s0, M, T, G, wave, lane, c + AH, lds_base, nb);
        }
        mma(I2(), I4(), f);
        // chunk j+1 landed; chunks j+2 .. j+AH (those issued) may be pending
        wait_pending<NLOAD, AH - 1>(total - j - 2);
        block_barrier();
        buf = buf + 1 < Pr::NBUF ? buf + 1 : 0;
      } else {
        // two buffers: chunk j+2 goes into this chunk's buffer once every
        // wave has read it (after the barrier), with the whole of chunk
        // j+1 to land
        mma(I0(), I2(), f);
        mma(I2(), I4(), f);
        wait_vm<0>();  // chunk j+1 landed (no later chunk is in flight)
        block_barrier();
        if (more)
          load_ahead<PREC, SP, P, 2>(pass_src<PREC, P>(M, G, wave, ln), M, T, G, wave, ln, c + 2,
                                     lds_base, buf);
        buf ^= 1;
      }
    }
  }
  if constexpr (PREC != TCE_T_FP64) from_quads<PREC, H>(acc, q);
  if constexpr (pi.minus) negate(acc);
}

template <int PREC, bool SP, int P>
__device__ __forceinline__ void all_passes(typename Prec<PREC>::acc (&acc)[64], const f4 *lds,
                                           unsigned lds_base, const MixedArgsG &M,
                                           gptr<const TaskDev> T, const ItemGeo &G, int wave,
                                           int lane, int &j, int &buf, int total,
                                           Frag<PREC> &F) {
  pass_step<PREC, SP, P>(acc, lds, lds_base, M, T, G, wave, lane, j, buf, total, F);
  if constexpr (P + 1 < 18)
    all_passes<PREC, SP, P + 1>(acc, lds, lds_base, M, T, G, wave, lane, j, buf, total, F);
}

// Stage the item's singles operands into LDS (FP64, as fused_wg.hip).
__device__ __forceinline__ void stage_epilogue(double *e, const FusedArgsG &A, int tid,
                                               const ItemGeo &G) {
  const int no = A.no, nv = A.nv;
#pragma unroll
  for (int rb = 0; rb < 18; rb += 6) {
    double v[6];
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      const int idx = tid + 512 * (rb + r);
      const int q = idx >> 6, p1 = (idx >> 3) & 7, p2 = idx & 7;
      const int sp = q % 3, tp = q / 3, t = tp >> 4, u = (tp >> 2) & 3, w = tp & 3;
      const int h1 = hole_of(G, t == 0 ? 1 : 0, u), h2 = hole_of(G, t == 2 ? 1 : 2, w);
      const long g1 = particle_of(G, sp == 0 ? 1 : 0, p1), g2 = particle_of(G, sp == 2 ? 1 : 2, p2);
      v[r] = A.S[(((long)h1 * no + h2) * nv + g1) * nv + g2];
    }
#pragma unroll
    for (int r = 0; r < 6; ++r) e[tid + 512 * (rb + r)] = v[r];
  }
  if (tid < EPI_T1) {
    const int u = tid & 3, slot = (tid >> 2) & 7, pt = (tid >> 5) / 3, ht = (tid >> 5) % 3;
    e[EPI_S + tid] = A.T1[particle_of(G, pt, slot) * no + hole_of(G, ht, u)];
  }
}

// FP64 energy of one wave's 4096 elements: d[16 i + 4 j + k] (converted
// from the FP32 accumulators); den = d * factor / (e_h1 + e_h2 + e_h3 -
// e_p4 - e_p5 - e_p6), E[T] += den * d, E(T) += den * (d + s) with the 9
// singles products (ccsd_t_dot.F:107 and the s1 table, as fused_wg.hip).
template <class AccT>
__device__ __forceinline__ void energy(const AccT (&acc)[64], const double *e,
                                       const FusedArgsG &A, const TaskDev &T, int sa, int sb,
                                       int sc, int la, int lb, int lc, int i0, int j0, int k0,
                                       const ItemGeo &G, double &e4, double &e5) {
  e4 = 0.0;
  e5 = 0.0;
  if (!(la < T.r4 && lb < T.r5 && lc < T.r6)) return;
  const int no = A.no;
  const double denom_0 =
      -(A.evl[no + T.op4 + la] + A.evl[no + T.op5 + lb] + A.evl[no + T.op6 + lc]);
  double eh[3][4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    eh[0][u] = A.evl[hole(G.I, u)];
    eh[1][u] = A.evl[hole(G.J, u)];
    eh[2][u] = A.evl[hole(G.K, u)];
  }
  double den[64];
#pragma unroll
  for (int ii = 0; ii < 4; ++ii)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int r = 16 * ii + 4 * jj + kk;
        const bool ok = i0 + ii < T.r1 && j0 + jj < T.r2 && k0 + kk < T.r3;
        const double d = (double)acc[r];
        const double dn = d * T.factor / (eh[0][ii] + eh[1][jj] + eh[2][kk] + denom_0);
        e4 += ok ? dn * d : 0.0;
        den[r] = ok ? dn : 0.0;
      }
  const int obc = 0 * 64 + sb * 8 + sc, oac = 1 * 64 + sa * 8 + sc, oab = 2 * 64 + sa * 8 + sb;
  const double *t1 = e + EPI_S;
  double x = 0.0;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    double ta[4], tb[4], tc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ta[u] = t1[((0 * 3 + t) * 8 + sa) * 4 + u];
      tb[u] = t1[((1 * 3 + t) * 8 + sb) * 4 + u];
      tc[u] = t1[((2 * 3 + t) * 8 + sc) * 4 + u];
    }
    const double sg = t == 1 ? -1.0 : 1.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const double *sp = e + (t * 16 + u * 4 + w) * 192;
        const double sbc = sg * sp[obc], sac = sg * sp[oac], sab = sg * sp[oab];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const int r = t == 0 ? m * 16 + u * 4 + w : (t == 1 ? u * 16 + m * 4 + w : u * 16 + w * 4 + m);
          x += den[r] * (ta[m] * sbc - tb[m] * sac + tc[m] * sab);
        }
      }
  }
  e5 = e4 + x;
}

// Tile export: the task's doubles d and singles s (the products of the
// energy's singles term, without the denominator), element (a, b, c, i, j, k)
// of the task at ((((a r5 + b) r6 + c) r1 + i) r2 + j) r3 + k, i.e. the
// reference's triplesx(h3,h2,h1,p6,p5,p4); doubles at A.partials, singles
// n = |tile| doubles after them.  One task per launch (A.tasks[0]).
// The 9 singles products of the lane's 64 elements (the energy's singles
// term without the denominator), from the staged epilogue operands.
__device__ __forceinline__ void singles_products(const double *e, int sa, int sb, int sc,
                                                 double (&sv)[64]) {
#pragma unroll
  for (int r = 0; r < 64; ++r) sv[r] = 0.0;
  const int obc = 0 * 64 + sb * 8 + sc, oac = 1 * 64 + sa * 8 + sc, oab = 2 * 64 + sa * 8 + sb;
  const double *t1 = e + EPI_S;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    double ta[4], tb[4], tc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      ta[u] = t1[((0 * 3 + t) * 8 + sa) * 4 + u];
      tb[u] = t1[((1 * 3 + t) * 8 + sb) * 4 + u];
      tc[u] = t1[((2 * 3 + t) * 8 + sc) * 4 + u];
    }
    const double sg = t == 1 ? -1.0 : 1.0;
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int w = 0; w < 4; ++w) {
        const double *sp = e + (t * 16 + u * 4 + w) * 192;
        const double sbc = sg * sp[obc], sac = sg * sp[oac], sab = sg * sp[oab];
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const int r = t == 0 ? m * 16 + u * 4 + w : (t == 1 ? u * 16 + m * 4 + w : u * 16 + w * 4 + m);
          sv[r] += ta[m] * sbc - tb[m] * sac + tc[m] * sab;
        }
      }
  }
}

template <class AccT>
__device__ __forceinline__ void store_tiles(const AccT (&acc)[64], const double *e,
                                            const FusedArgsG &A, const TaskDev &T, int sa, int sb,
                                            int sc, int la, int lb, int lc, int i0, int j0,
                                            int k0) {
  if (!(la < T.r4 && lb < T.r5 && lc < T.r6)) return;
  const long n = (long)T.r1 * T.r2 * T.r3 * T.r4 * T.r5 * T.r6;
  const gptr<double> td = A.partials, ts = A.partials + n;
  const long base = (((long)la * T.r5 + lb) * T.r6 + lc) * T.r1;
#pragma unroll
  for (int ii = 0; ii < 4; ++ii)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
        if (i0 + ii < T.r1 && j0 + jj < T.r2 && k0 + kk < T.r3)
          td[((base + i0 + ii) * T.r2 + j0 + jj) * T.r3 + k0 + kk] = (double)acc[16 * ii + 4 * jj + kk];
  double sv[64];
  singles_products(e, sa, sb, sc, sv);
#pragma unroll
  for (int ii = 0; ii < 4; ++ii)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk)
        if (i0 + ii < T.r1 && j0 + jj < T.r2 && k0 + kk < T.r3)
          ts[((base + i0 + ii) * T.r2 + j0 + jj) * T.r3 + k0 + kk] = sv[16 * ii + 4 * jj + kk];
}

// Epilogues of the item loop: each item's energy partials, one task's t3
// tiles (store_tiles), or the two Lambda-CCSD(T) passes (lambda.hip).
enum { EPI_ENERGY = 0, EPI_TILES = 1, EPI_LAMBDA_LEFT = 2, EPI_LAMBDA_RIGHT = 3, EPI_CR_RIGHT = 4 };

// Left-hand pass: the wave's y3 doubles (the accumulators plus the Fock
// term P(c/ab) P(k/ij) y(a b i j) f(k c), y from the job's dense T2e,
// T2e[i][j][a][b] = y(a b i j)) and y3 singles, register r of the lane at
// out[r * 64 + lane] and out[(64 + r) * 64 + lane].  Elements outside the
// task are written too (the right-hand pass skips them).
template <class AccT>
__device__ __forceinline__ void lambda_left(const AccT (&acc)[64], const double *e,
                                            const MixedArgsG &M, const TaskDev &T, int sa, int sb,
                                            int sc, int la, int lb, int lc, const ItemGeo &G,
                                            gptr<double> out, int lane) {
  const FusedArgsG &A = M.f;
  double yd[64];
#pragma unroll
  for (int r = 0; r < 64; ++r) yd[r] = (double)acc[r];
  if (M.lam_f && la < T.r4 && lb < T.r5 && lc < T.r6) {
    const int no = A.no, nv = A.nv;
    const long Ep = A.Ep;
    const int a = T.op4 + la, b = T.op5 + lb, c = T.op6 + lc;
    int hI[4], hJ[4], hK[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      hI[u] = hole(G.I, u);
      hJ[u] = hole(G.J, u);
      hK[u] = hole(G.K, u);
    }
    const gptr<const double> Y = A.T2e, F = M.lam_f;
    auto y = [&](int p, int q, int h, int g) { return Y[(((long)h * no + g) * nv + p) * Ep + q]; };
    // g(h1, h2, h3) = y(a b h1 h2) f(h3 c) - y(c b h1 h2) f(h3 a) - y(a c h1 h2) f(h3 b);
    // the term is g(i, j, k) - g(k, j, i) - g(i, k, j).  T = 0: g(i, j, k);
    // T = 1: -g(k, j, i); T = 2: -g(i, k, j).  Outer loops over (h1, h2),
    // inner over h3, whose three f values are loaded once per term.
    // hole u of set 0 (i), 1 (j) or 2 (k); every index is a constant after unrolling
    auto H = [&](int set, int u) { return set == 0 ? hI[u] : (set == 1 ? hJ[u] : hK[u]); };
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const int s1 = t == 1 ? 2 : 0, s2 = t == 2 ? 2 : 1, s3 = t == 0 ? 2 : (t == 1 ? 0 : 1);
      const double sg = t == 0 ? 1.0 : -1.0;
      double fc[4], fa[4], fb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        fc[u] = F[(long)H(s3, u) * nv + c];
        fa[u] = F[(long)H(s3, u) * nv + a];
        fb[u] = F[(long)H(s3, u) * nv + b];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          const int h1 = H(s1, u), h2 = H(s2, w);
          const double yab = sg * y(a, b, h1, h2), ycb = sg * y(c, b, h1, h2),
                       yac = sg * y(a, c, h1, h2);
#pragma unroll
          for (int m = 0; m < 4; ++m) {
            // (h1, h2, h3) = (u, w, m) in the order of (i, j, k) for the term
            const int ii = t == 1 ? m : u, jj = t == 2 ? m : w, kk = t == 0 ? m : (t == 1 ? u : w);
            yd[16 * ii + 4 * jj + kk] += yab * fc[m] - ycb * fa[m] - yac * fb[m];
          }
        }
    }
  }
  // doubles out before the singles are formed: one 64-value array live at a time
#pragma unroll
  for (int r = 0; r < 64; ++r) out[r * 64 + lane] = yd[r];
  double sv[64];
  singles_products(e, sa, sb, sc, sv);
#pragma unroll
  for (int r = 0; r < 64; ++r) out[(64 + r) * 64 + lane] = sv[r];
}

// Right-hand pass: den = t3d * factor / D as in energy(), e1 += den * yd,
// e2 += den * (yd + ys) with the left-hand values of the same elements.
template <class AccT>
__device__ __forceinline__ void lambda_right(const AccT (&acc)[64], const FusedArgsG &A,
                                             const TaskDev &T, int la, int lb, int lc, int i0,
                                             int j0, int k0, const ItemGeo &G,
                                             gptr<const double> in, int lane, double &e1,
                                             double &e2) {
  e1 = 0.0;
  e2 = 0.0;
  if (!(la < T.r4 && lb < T.r5 && lc < T.r6)) return;
  const int no = A.no;
  const double denom_0 =
      -(A.evl[no + T.op4 + la] + A.evl[no + T.op5 + lb] + A.evl[no + T.op6 + lc]);
  double eh[3][4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    eh[0][u] = A.evl[hole(G.I, u)];
    eh[1][u] = A.evl[hole(G.J, u)];
    eh[2][u] = A.evl[hole(G.K, u)];
  }
#pragma unroll
  for (int ii = 0; ii < 4; ++ii)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int r = 16 * ii + 4 * jj + kk;
        const bool ok = i0 + ii < T.r1 && j0 + jj < T.r2 && k0 + kk < T.r3;
        const double dn = (double)acc[r] * T.factor / (eh[0][ii] + eh[1][jj] + eh[2][kk] + denom_0);
        const double yd = in[r * 64 + lane], ys = in[(64 + r) * 64 + lane];
        e1 += ok ? dn * yd : 0.0;
        e2 += ok ? dn * (yd + ys) : 0.0;
      }
}

// CR-CCSD(T) right-hand pass (cr_ccsd_t.F:150-205): with w = factor / D,
// the right-hand job's doubles r (the moment) and singles q (the
// denominator tile) against the left-hand t3 doubles td and singles ts of
// the same elements: s0 += w r td, s1 += w r (td + ts), s2 += w q td,
// s3 += w q (td + ts).
template <class AccT>
__device__ __forceinline__ void cr_right(const AccT (&acc)[64], const double *e,
                                         const FusedArgsG &A, const TaskDev &T, int sa, int sb,
                                         int sc, int la, int lb, int lc, int i0, int j0, int k0,
                                         const ItemGeo &G, gptr<const double> in, int lane,
                                         double (&s)[4]) {
  s[0] = s[1] = s[2] = s[3] = 0.0;
  if (!(la < T.r4 && lb < T.r5 && lc < T.r6)) return;
  double sv[64];
  singles_products(e, sa, sb, sc, sv);
  const int no = A.no;
  const double denom_0 =
      -(A.evl[no + T.op4 + la] + A.evl[no + T.op5 + lb] + A.evl[no + T.op6 + lc]);
  double eh[3][4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    eh[0][u] = A.evl[hole(G.I, u)];
    eh[1][u] = A.evl[hole(G.J, u)];
    eh[2][u] = A.evl[hole(G.K, u)];
  }
#pragma unroll
  for (int ii = 0; ii < 4; ++ii)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int r = 16 * ii + 4 * jj + kk;
        const bool ok = i0 + ii < T.r1 && j0 + jj < T.r2 && k0 + kk < T.r3;
        const double w = T.factor / (eh[0][ii] + eh[1][jj] + eh[2][kk] + denom_0);
        const double rd = (double)acc[r] * w, rs = sv[r] * w;
        const double td = in[r * 64 + lane], ts = in[(64 + r) * 64 + lane];
        s[0] += ok ? rd * td : 0.0;
        s[1] += ok ? rd * (td + ts) : 0.0;
        s[2] += ok ? rs * td : 0.0;
        s[3] += ok ? rs * (td + ts) : 0.0;
      }
}

// The item loop of the kernels below, with the epilogue EPI.
template <int PREC, int EPI, bool SP>
__global__ void __launch_bounds__(512) fused_triples_mixed(const MixedSegment *segs, int nseg,
                                                           long nglobal,
                                                           unsigned long long *sched) {
  __shared__ f4 lds[lds_quads<PREC>()];
  __shared__ long next_item;  // dynamic schedule: the workgroup's next item
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const unsigned lds_base = (unsigned)(unsigned long)(__attribute__((address_space(3))) f4 *)lds;
  const int kq = lane >> 4, hi = (lane & 15) >> 2, lo = lane & 3;
  using Pr = Prec<PREC>;
  // Persistent loop over the launch's items, uniform across the workgroup
  // (every barrier is reached by all 8 waves): the dynamic schedule of
  // sched_item when `sched` is given, else a static round-robin.
  const bool dyn = sched != nullptr;
  const int sh = dyn ? (int)sched[9] : 0;  // log2 of the slice width
  unsigned q = blockIdx.x & 7u;  // thread 0: the queue it draws from
  unsigned long long ticket = 0;
  long g = 0;
  if (dyn) {
    if (tid == 0) next_item = sched_item(sched, q, atomicAdd(&sched[q], 1ull), nglobal, sh);
    __syncthreads();
    g = next_item;
  }
  int seg = 0;  // segment of the workgroup's current item
  for (long r = 0;; ++r) {
    if (dyn) {
      if (g < 0) break;
      // the next item's ticket is drawn now and read at the end of this item
      if (tid == 0) ticket = atomicAdd(&sched[q], 1ull);
    } else {
      if (r * (long)gridDim.x >= nglobal) break;
      g = round_item(r, blockIdx.x, gridDim.x, nseg > 1);
      if (g >= nglobal) continue;  // last round: uniform per workgroup
    }
    seg = next_segment(segs, nseg, seg, g);
    const MixedSegment &sg = segs[seg];
    const long item = sg.item0 + (g - sg.g0);
    const MixedArgsG M = load_args(sg.args);
    const FusedArgsG &A = M.f;
    const gptr<const TaskDev> Tp = A.tasks + find_task(A.tasks, A.ntasks, item);
    const TaskDev T = *(const TaskDev *)Tp;
    int total = 0;  // chunks of the item (the passes' contraction ranges)
#pragma unroll
    for (int p = 0; p < 18; ++p) total += chunks_of<PREC, SP>(M, Tp, p).n;
    // boxes of a multi-box task (reference and orbital tilings): hole boxes
    // outermost, particle c fastest.  Measured against hole k fastest on the
    // orbital-tiling jobs of the 64-job batch (up to 32 particle boxes per
    // index against at most 8 hole boxes): 6-7% faster in split-BF16, 2% in
    // FP64.  (Sub-tile tables have one box per task and order their tasks
    // themselves.)
    long box = item - T.item0;
    const int bc = box % T.nb6; box /= T.nb6;
    const int bb = box % T.nb5; box /= T.nb5;
    const int ba = box % T.nb4; box /= T.nb4;
    const int bk = box % T.nb3; box /= T.nb3;
    const int bj = box % T.nb2; box /= T.nb2;
    const int bi = (int)box;
    const int a0 = 8 * ba, b0 = 8 * bb, c0 = 8 * bc;
    ItemGeo G;
    G.I = HoleBox{T.oh1, 4 * bi, T.r1 - 1};
    G.J = HoleBox{T.oh2, 4 * bj, T.r2 - 1};
    G.K = HoleBox{T.oh3, 4 * bk, T.r3 - 1};
    G.a = PRange{T.op4 + a0, min(7, T.r4 - 1 - a0)};
    G.b = PRange{T.op5 + b0, min(7, T.r5 - 1 - b0)};
    G.c = PRange{T.op6 + c0, min(7, T.r6 - 1 - c0)};

    typename Pr::acc acc[64];  // element (i, j, k) of the lane at acc[16 i + 4 j + k]
#pragma unroll
    for (int r = 0; r < 64; ++r) acc[r] = 0;

    Frag<PREC> F;
    {  // chunks 0 .. 2 (SPLIT) or 0 .. NBUF-2 of the pass sequence into
       // buffers 0 ..; SPLIT: then the fragments of chunk 0
      constexpr int PRO = Pr::SPLIT ? Pr::NBUF : (Pr::NBUF > 3 ? Pr::NBUF - 1 : 2);
      const Src<PREC> s0 = pass_src<PREC, 0>(M, G, wave, lane);
#pragma unroll
      for (int k = 0; k < PRO; ++k)
        if (k < total) load_ahead<PREC, SP, 0, PRO>(s0, M, Tp, G, wave, lane, k, lds_base, k);
      // chunk 0 has landed once at most the loads of the later ones are
      // pending (fewer chunks in the item: wait for all)
      if (total >= PRO)
        wait_vm<3 * Pr::NPART * (PRO - 1)>();
      else
        wait_vm<0>();
      block_barrier();
      if constexpr (Pr::SPLIT) {
        const int2 o = frag_offsets<PREC, 0>(wave, lane);
        read_x<PREC, 0, 4>(F, lds, o.x);
        read_y<PREC>(F, lds, o.y);
      }
    }
    int j = 0, buf = 0;
    all_passes<PREC, SP, 0>(acc, lds, lds_base, M, Tp, G, wave, lane, j, buf, total, F);

    // Epilogue: the chunk buffers are free (the last chunk ended with a
    // barrier and no load is outstanding).  The right-hand Lambda pass needs
    // no singles operands.
    double *epi = reinterpret_cast<double *>(lds);
    if constexpr (EPI != EPI_LAMBDA_RIGHT) stage_epilogue(epi, A, tid, G);
    __syncthreads();
    // final layout: L = c, lane = 4a' + b' + 16c'
    const int sa = 4 * cube(wave, 0) + hi, sb = 4 * cube(wave, 1) + lo, sc = 4 * cube(wave, 2) + kq;
    if constexpr (EPI == EPI_TILES) {
      store_tiles(acc, epi, A, T, sa, sb, sc, a0 + sa, b0 + sb, c0 + sc, 4 * bi, 4 * bj, 4 * bk);
    } else if constexpr (EPI == EPI_LAMBDA_LEFT) {
      lambda_left(acc, epi, M, T, sa, sb, sc, a0 + sa, b0 + sb, c0 + sc, G,
                  A.partials + ((item - M.lam_item0) * 8 + wave) * 8192, lane);
    } else if constexpr (EPI == EPI_CR_RIGHT) {
      double cs[4];
      cr_right(acc, epi, A, T, sa, sb, sc, a0 + sa, b0 + sb, c0 + sc, 4 * bi, 4 * bj, 4 * bk, G,
               M.lam_y + ((item - M.lam_item0) * 8 + wave) * 8192, lane, cs);
      wave_sum(cs[0], cs[1]);
      wave_sum(cs[2], cs[3]);
      if (lane == 0) {
        A.partials[2 * (item * 8 + wave)] = cs[0];
        A.partials[2 * (item * 8 + wave) + 1] = cs[1];
        M.cr_den[2 * (item * 8 + wave)] = cs[2];
        M.cr_den[2 * (item * 8 + wave) + 1] = cs[3];
      }
    } else {
      double e4, e5;
      if constexpr (EPI == EPI_LAMBDA_RIGHT)
        lambda_right(acc, A, T, a0 + sa, b0 + sb, c0 + sc, 4 * bi, 4 * bj, 4 * bk, G,
                     M.lam_y + ((item - M.lam_item0) * 8 + wave) * 8192, lane, e4, e5);
      else
        energy(acc, epi, A, T, sa, sb, sc, a0 + sa, b0 + sb, c0 + sc, 4 * bi, 4 * bj, 4 * bk, G,
               e4, e5);
      wave_sum(e4, e5);
      if (lane == 0) {
        A.partials[2 * (item * 8 + wave)] = e4;
        A.partials[2 * (item * 8 + wave) + 1] = e5;
      }
    }
    if (dyn && tid == 0) next_item = sched_item(sched, q, ticket, nglobal, sh);
    __syncthreads();  // the next item's first loads overwrite the staged data
    if (dyn) g = next_item;
  }
  if (dyn && tid == 0) sched_leave(sched);
}

}  // namespace

// The contraction extents of every job in the table must be multiples of
// the precision's chunk (16 FP32 / 32 bf16 values); the job layer pads them.
// sparse: some job of the table has contraction ranges narrower than its
// spaces (spin / irrep labels), which the SP = true kernels follow.
template <int PREC, int EPI, bool SP>
static void launch_one(const MixedSegment *segs, int nseg, long total, int grid, hipStream_t s,
                       unsigned long long *sched) {
  hipLaunchKernelGGL(HIP_KERNEL_NAME(fused_triples_mixed<PREC, EPI, SP>), dim3((unsigned)grid),
                     dim3(512), 0, s, segs, nseg, total, sched);
}
template <int EPI>
static int launch_prec(const MixedSegment *segs, int nseg, long total, int precision, int grid,
                       bool sparse, hipStream_t s, unsigned long long *q) {
  if (precision == TCE_T_FP64)
    sparse ? launch_one<TCE_T_FP64, EPI, true>(segs, nseg, total, grid, s, q)
           : launch_one<TCE_T_FP64, EPI, false>(segs, nseg, total, grid, s, q);
  else if (precision == TCE_T_FP32)
    sparse ? launch_one<TCE_T_FP32, EPI, true>(segs, nseg, total, grid, s, q)
           : launch_one<TCE_T_FP32, EPI, false>(segs, nseg, total, grid, s, q);
  else if (precision == TCE_T_BF16X3)
    sparse ? launch_one<TCE_T_BF16X3, EPI, true>(segs, nseg, total, grid, s, q)
           : launch_one<TCE_T_BF16X3, EPI, false>(segs, nseg, total, grid, s, q);
  else if (precision == TCE_T_BF16X6)
    sparse ? launch_one<TCE_T_BF16X6, EPI, true>(segs, nseg, total, grid, s, q)
           : launch_one<TCE_T_BF16X6, EPI, false>(segs, nseg, total, grid, s, q);
  else
    return -1;
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int launch_fused_mixed(const MixedSegment *segs, int nseg, long total, int precision, int grid,
                       hipStream_t s, bool sparse, unsigned long long *sched) {
  return launch_prec<EPI_ENERGY>(segs, nseg, total, precision, grid, sparse, s, sched);
}

int launch_fused_tiles(const MixedSegment *segs, long total, int precision, int grid,
                       hipStream_t s, bool sparse, unsigned long long *sched) {
  return launch_prec<EPI_TILES>(segs, 1, total, precision, grid, sparse, s, sched);
}

int launch_fused_lambda(const MixedSegment *segs, long total, int precision, int grid,
                        hipStream_t s, bool sparse, unsigned long long *sched, bool left) {
  return left ? launch_prec<EPI_LAMBDA_LEFT>(segs, 1, total, precision, grid, sparse, s, sched)
              : launch_prec<EPI_LAMBDA_RIGHT>(segs, 1, total, precision, grid, sparse, s, sched);
}

int launch_fused_cr_right(const MixedSegment *segs, long total, int precision, int grid,
                          hipStream_t s, bool sparse, unsigned long long *sched) {
  return launch_prec<EPI_CR_RIGHT>(segs, 1, total, precision, grid, sparse, s, sched);
}


}  // namespace tce_t
