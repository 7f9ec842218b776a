// fused_wg.hip -- cooperative fused (T) kernel: a 512-thread workgroup (8
// waves, 2 per SIMD) owns an 8x8x8 particle box by a 4x4x4 hole box of one
// task's t3 tile; wave w holds the 4x4x4 particle cube (w>>2&1, w>>1&1, w&1)
// in 64 FP64 accumulators (layout and variant table in fused.h).
//
// For each of the 18 d1/d2 contractions the workgroup streams the two
// operand slices through LDS in chunks of 8 contraction values, double
// buffered: every thread fetches 3 x 16 B of the next chunk from HBM/L2 while
// the waves run the 128 MFMAs of the current chunk from LDS, then stores them
// and the workgroup synchronises once per chunk.  Each operand element is
// read from global memory once per workgroup instead of once per wave.
//
// LDS layout of one buffer (16-byte quads; a quad holds the two contraction
// values a lane feeds to two consecutive MFMAs):
//   P  [16 hole pairs][4 k-pairs kq][8 particle slots ^ 2kq]           512 quads
//   Q  [4 singles][4 k-pairs kq][8 y slots][8 x slots ^ swz(kq, y)]   1024 quads
// with swz(kq, y) = 4*(y & 1) ^ (2kq + kq/2).  Reading a fragment, 8
// consecutive lanes (x = 0..3, y = 0..1) touch 8 distinct 4-bank groups
// (P: 4 quads broadcast over y; Q: bit 2 follows y, bits 0-1 follow x), and
// so do 8 consecutive lanes storing a staged chunk (4 k-pairs x 2 slots).
#include "fused.h"

namespace tce_t {
namespace {

constexpr int KC = 8;                 // contraction values per chunk
constexpr int P_QUADS = 16 * 4 * 8;   // 512
constexpr int Q_BLK = 64;             // quads per (single, k-pair)
constexpr int Q_QUADS = 4 * 4 * Q_BLK;
constexpr int BUF_QUADS = P_QUADS + Q_QUADS;  // 1536 quads = 24 KiB

// Bank swizzles of the P particle slot and the Q x slot (k-pair kq, y slot).
__device__ __forceinline__ constexpr int p_swz(int slot, int kq) { return slot ^ (2 * kq); }
__device__ __forceinline__ constexpr int q_swz(int sx, int sy, int kq) {
  return sx ^ (4 * (sy & 1)) ^ (2 * kq + (kq >> 1));
}

// Epilogue staging (doubles), overlaying both chunk buffers:
//   S   [hole pair t: (j,k) (i,k) (i,j)][16 pairs u*4+v]
//       [particle pair s: (b,c) (a,c) (a,b)][8 x 8 slots]           9216
//   T1  [particle box a b c][hole box i j k][8 slots][4 holes]        288
constexpr int EPI_S = 3 * 16 * 3 * 64;
constexpr int EPI_T1 = 9 * 32;
constexpr int EPI_DOUBLES = EPI_S + EPI_T1;
constexpr int LDS_QUADS = (2 * BUF_QUADS > EPI_DOUBLES / 2) ? 2 * BUF_QUADS : EPI_DOUBLES / 2;

// Hole box of an item: global hole o + min(b0 + u, rm), u = 0..3.
struct HoleBox {
  int o, b0, rm;
};
__device__ __forceinline__ int hole(const HoleBox &h, int u) { return h.o + min(h.b0 + u, h.rm); }

// Particle range of a workgroup box: global index g0 + min(slot, sm),
// slot = 0..7 (clamped into the tile at a ragged edge).
struct PRange {
  long g0;
  int sm;
};

// Uniform description of one contraction's operand slices:
// P(pair, particle, k) = P[(h1*no+h2)*pS + particle*Kp + k] and
// Q(single, first, second, k) = Q[h*qS + (first*nv + second)*Kp + k].
// pair: 0 = holes (i,j) with single k, 1 = (j,k) with i, 2 = (i,k) with j.
// qx_first: the lane-x particle of Q is its first index (phase 1) or its
// second (phase 2).
struct PassDesc {
  gptr<const double> P;
  long pS;
  PRange pr;
  gptr<const double> Q;
  long qS;
  PRange fr, gr;
  int pair;
  bool qx_first;
  int Kp;
};

// One thread's share of staging a chunk of a pass: three 16-byte global
// reads (P row quad, two Q row quads) and their LDS quad slots.
struct Stage {
  gptr<const double> pg, qg0, qg1;
  int pl, ql0, ql1;
};

__device__ __forceinline__ Stage make_stage(const PassDesc &d, int tid, const HoleBox &hI,
                                            const HoleBox &hJ, const HoleBox &hK, int no, int nv) {
  Stage st;
  const int prow = tid >> 2, pq = tid & 3;
  const int ppair = prow >> 3, ps = prow & 7;
  const int pu = ppair >> 2, pv = ppair & 3;
  const int ph1 = d.pair == 1 ? hole(hJ, pu) : hole(hI, pu);
  const int ph2 = d.pair == 0 ? hole(hJ, pv) : hole(hK, pv);
  st.pg = d.P + ((long)ph1 * no + ph2) * d.pS + (d.pr.g0 + min(ps, d.pr.sm)) * d.Kp + 2 * pq;
  st.pl = ppair * 32 + pq * 8 + p_swz(ps, pq);
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int idx = tid + 512 * r;
    const int row = idx >> 2, q = idx & 3;
    const int w = row >> 6, pp = row & 63, fs = pp >> 3, gs = pp & 7;
    const int hs = d.pair == 0 ? hole(hK, w) : (d.pair == 1 ? hole(hI, w) : hole(hJ, w));
    const long f = d.fr.g0 + min(fs, d.fr.sm), g = d.gr.g0 + min(gs, d.gr.sm);
    const gptr<const double> src = d.Q + (long)hs * d.qS + (f * nv + g) * d.Kp + 2 * q;
    const int sx = d.qx_first ? fs : gs, sy = d.qx_first ? gs : fs;
    const int dst = P_QUADS + (w * 4 + q) * Q_BLK + sy * 8 + q_swz(sx, sy, q);
    if (r == 0) {
      st.qg0 = src;
      st.ql0 = dst;
    } else {
      st.qg1 = src;
      st.ql1 = dst;
    }
  }
  return st;
}

typedef double dv2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ double2 load_quad(gptr<const double> p) {
  const dv2 v = *(gptr<const dv2>)p;
  return double2(v.x, v.y);
}

__device__ __forceinline__ void stage_load(const Stage &st, int k, double2 &a, double2 &b,
                                           double2 &c) {
  a = load_quad(st.pg + k);
  b = load_quad(st.qg0 + k);
  c = load_quad(st.qg1 + k);
}

__device__ __forceinline__ void stage_store(double2 *buf, int pl, int ql0, int ql1, double2 a,
                                            double2 b, double2 c) {
  buf[pl] = a;
  buf[ql0] = b;
  buf[ql1] = c;
}

// One contraction.  On entry chunk 0 of `cur` sits in LDS buffer `buf`
// (visible to all waves).  Every chunk: prefetch the next chunk -- of this
// pass, or chunk 0 of the next pass during the last chunk -- into registers,
// run the 128 MFMAs of the current chunk from LDS, store the prefetch into
// the other buffer, synchronise once.  pslot: this lane's P particle slot and
// qoff its Q quad offset within a (single, k-pair) block, both swizzled for
// the lane's k-pair z.
template <int PAIR, bool P_IS_R>
__device__ __forceinline__ void pass(acc_t (&acc)[64], double2 *lds, int &buf, const Stage &cur,
                                     int Kp, const Stage &next, bool has_next, int pslot, int qoff,
                                     int z, double sign) {
  const int nch = Kp / KC;
  for (int c = 0; c < nch; ++c) {
    const double2 *cb = lds + buf * BUF_QUADS;
    double2 *nb = lds + (buf ^ 1) * BUF_QUADS;
    const bool more = c + 1 < nch;
    double2 r0, r1, r2;
    if (more)
      stage_load(cur, (c + 1) * KC, r0, r1, r2);
    else if (has_next)
      stage_load(next, 0, r0, r1, r2);
    double2 qf[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const double2 t = cb[P_QUADS + (w * 4 + z) * Q_BLK + qoff];
      qf[w].x = sign * t.x;
      qf[w].y = sign * t.y;
    }
    // P fragments in two halves of the pair index keep 32 VGPRs live
    // instead of 64 (LDS latency is short; the other wave on the SIMD
    // covers it).
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      double2 pf[2][4];
#pragma unroll
      for (int uu = 0; uu < 2; ++uu)
#pragma unroll
        for (int v = 0; v < 4; ++v) pf[uu][v] = cb[((2 * h + uu) * 4 + v) * 32 + z * 8 + pslot];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int uu = 0; uu < 2; ++uu)
#pragma unroll
          for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int w = 0; w < 4; ++w) {
              const int r = acc_index<PAIR>(2 * h + uu, v, w);
              const double pv = t ? pf[uu][v].y : pf[uu][v].x;
              const double qv = t ? qf[w].y : qf[w].x;
              acc[r] = P_IS_R ? mfma4(pv, qv, acc[r]) : mfma4(qv, pv, acc[r]);
            }
    }
    if (more)
      stage_store(nb, cur.pl, cur.ql0, cur.ql1, r0, r1, r2);
    else if (has_next)
      stage_store(nb, next.pl, next.ql0, next.ql1, r0, r1, r2);
    __syncthreads();
    buf ^= 1;
  }
}

// Stage the item's singles operands into LDS: every S(h h' | p p') the 4x4x4
// hole box by 8x8x8 particle box needs and the T1(p, h) slices.  18 + 1
// independent loads per thread, issued together.  Slots past a ragged tile
// edge read clamped (in-range) indices; their elements are masked later.
__device__ __forceinline__ void stage_epilogue(double *e, const FusedArgsG &A, int tid,
                                               const HoleBox (&H)[3], const PRange (&R)[3], int no,
                                               int nv) {
  double v[18];
#pragma unroll
  for (int r = 0; r < 18; ++r) {
    const int idx = tid + 512 * r;
    const int q = idx >> 6, p1 = (idx >> 3) & 7, p2 = idx & 7;
    const int sp = q % 3, tp = q / 3, t = tp >> 4, u = (tp >> 2) & 3, w = tp & 3;
    const int h1 = hole(H[t == 0 ? 1 : 0], u), h2 = hole(H[t == 2 ? 1 : 2], w);
    const PRange &R1 = R[sp == 0 ? 1 : 0], &R2 = R[sp == 2 ? 1 : 2];
    const long g1 = R1.g0 + min(p1, R1.sm), g2 = R2.g0 + min(p2, R2.sm);
    v[r] = A.S[(((long)h1 * no + h2) * nv + g1) * nv + g2];
  }
  double t1 = 0.0;
  if (tid < EPI_T1) {
    const int u = tid & 3, slot = (tid >> 2) & 7, pt = (tid >> 5) / 3, ht = (tid >> 5) % 3;
    t1 = A.T1[(R[pt].g0 + min(slot, R[pt].sm)) * no + hole(H[ht], u)];
  }
#pragma unroll
  for (int r = 0; r < 18; ++r) e[tid + 512 * r] = v[r];
  if (tid < EPI_T1) e[EPI_S + tid] = t1;
}

// Energy of one wave's 4096 elements (phase-2 layout) from the staged LDS
// operands.  sa/sb/sc: the lane's particle slots in the box; la/lb/lc its
// particle indices within the task's tiles; i0/j0/k0 the box's first holes
// within the tiles.  Per element, den = d * factor / (e_h1 + e_h2 + e_h3 -
// e_p4 - e_p5 - e_p6) and E[T] += den * d, E(T) += den * (d + s) with
//   s = t(a i)S(jk|bc) - t(a j)S(ik|bc) + t(a k)S(ij|bc)
//     - t(b i)S(jk|ac) + t(b j)S(ik|ac) - t(b k)S(ij|ac)
//     + t(c i)S(jk|ab) - t(c j)S(ik|ab) + t(c k)S(ij|ab)
// (ccsd_t_dot.F:107 and the s1 table).  The d registers are overwritten
// with den, and E(T) - E[T] = sum den * s is accumulated hole pair by hole
// pair so that each staged S value is read once per lane.
__device__ __forceinline__ void wg_energy(acc_t (&acc)[64], const double *e, const FusedArgsG &A,
                                          const TaskDev &T, int sa, int sb, int sc, int la, int lb,
                                          int lc, int i0, int j0, int k0, const HoleBox (&H)[3],
                                          double &e4, double &e5) {
  e4 = 0.0;
  e5 = 0.0;
  if (!(la < T.r4 && lb < T.r5 && lc < T.r6)) return;
  const int no = A.no;
  const double denom_0 = -(A.evl[no + T.op4 + la] + A.evl[no + T.op5 + lb] + A.evl[no + T.op6 + lc]);
  double eh[3][4];
#pragma unroll
  for (int t = 0; t < 3; ++t)
#pragma unroll
    for (int u = 0; u < 4; ++u) eh[t][u] = A.evl[hole(H[t], u)];
#pragma unroll
  for (int ii = 0; ii < 4; ++ii)
#pragma unroll
    for (int jj = 0; jj < 4; ++jj)
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int r = ii * 16 + jj * 4 + kk;
        const bool ok = i0 + ii < T.r1 && j0 + jj < T.r2 && k0 + kk < T.r3;
        const double d = acc[r];
        const double den = d * T.factor / (eh[0][ii] + eh[1][jj] + eh[2][kk] + denom_0);
        e4 += ok ? den * d : 0.0;
        acc[r] = ok ? den : 0.0;
      }
  // particle-pair offsets of this lane: (b,c), (a,c), (a,b)
  const int obc = 0 * 64 + sb * 8 + sc, oac = 1 * 64 + sa * 8 + sc, oab = 2 * 64 + sa * 8 + sb;
  const double *t1 = e + EPI_S;
  double x = 0.0;
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    // hole t is the one outside the pair: t = 0 -> i (pair j,k), 1 -> j, 2 -> k
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
          // element (ii, jj, kk): the lone hole m, the pair (u, w) in order
          const int r = t == 0 ? m * 16 + u * 4 + w : (t == 1 ? u * 16 + m * 4 + w : u * 16 + w * 4 + m);
          x += acc[r] * (ta[m] * sbc - tb[m] * sac + tc[m] * sab);
        }
      }
  }
  e5 = e4 + x;
}

__global__ void __launch_bounds__(512) fused_triples_wg(const Segment *segs, int nseg, long total) {
  __shared__ double2 lds[LDS_QUADS];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int x = lane & 3, y = (lane >> 2) & 3, z = lane >> 4;
  const int ca = (wave >> 2) & 1, cb = (wave >> 1) & 1, cc = wave & 1;
  // Persistent loop, static round-robin over the launch's items (uniform
  // across the workgroup, so every barrier below is reached by all 8 waves).
  int seg = 0;  // segment of the workgroup's current item
  for (long r = 0; r * (long)gridDim.x < total; ++r) {
    const long g = round_item(r, blockIdx.x, gridDim.x, nseg > 1);
    if (g >= total) continue;  // last round: uniform per workgroup
    seg = next_segment(segs, nseg, seg, g);
    const Segment &sg = segs[seg];
    const long item = sg.item0 + (g - sg.g0);
    const FusedArgsG A = load_args(sg.args);
    const int no = A.no, nv = A.nv, Ep = A.Ep, Mp = A.Mp;
    const long sT2e = (long)nv * Ep;      // stride of (h*no+h') in T2e
    const long sW = (long)nv * Mp;        // stride of (h*no+h') in Wd1
    const long sV = (long)nv * nv * Ep;   // stride of w in Vd2
    const long sT2m = (long)nv * nv * Mp; // stride of h in T2m
    const TaskDev T = *(const TaskDev *)(A.tasks + find_task(A.tasks, A.ntasks, item));
    long box = item - T.item0;
    const int bc = box % T.nb6; box /= T.nb6;
    const int bb = box % T.nb5; box /= T.nb5;
    const int ba = box % T.nb4; box /= T.nb4;
    const int bk = box % T.nb3; box /= T.nb3;
    const int bj = box % T.nb2; box /= T.nb2;
    const int bi = (int)box;
    const int a0 = 8 * ba, b0 = 8 * bb, c0 = 8 * bc;
    const HoleBox HI{T.oh1, 4 * bi, T.r1 - 1}, HJ{T.oh2, 4 * bj, T.r2 - 1}, HK{T.oh3, 4 * bk, T.r3 - 1};
    const PRange RA{T.op4 + a0, min(7, T.r4 - 1 - a0)}, RB{T.op5 + b0, min(7, T.r5 - 1 - b0)},
        RC{T.op6 + c0, min(7, T.r6 - 1 - c0)};

    acc_t acc[64];
#pragma unroll
    for (int r = 0; r < 64; ++r) acc[r] = 0.0;

    // The 18 contractions in order; descriptor k+1 is staged during the
    // last chunk of pass k.
    const PassDesc D[6] = {
        {A.T2e, sT2e, RA, A.Vd2, sV, RB, RC, 0, true, Ep},   // d2 N=1..3: t(a e..), <b c||..>
        {A.T2e, sT2e, RB, A.Vd2, sV, RA, RC, 0, true, Ep},   // d2 N=4..6: t(b e..), <a c||..>
        {A.Wd1, sW, RA, A.T2m, sT2m, RB, RC, 0, true, Mp},   // d1 N=4..6: <m a||..>, t(b c..)
        {A.Wd1, sW, RB, A.T2m, sT2m, RA, RC, 0, true, Mp},   // d1 N=7..9: <m b||..>, t(a c..)
        {A.T2e, sT2e, RC, A.Vd2, sV, RA, RB, 0, false, Ep},  // d2 N=7..9: t(c e..), <a b||..>
        {A.Wd1, sW, RC, A.T2m, sT2m, RA, RB, 0, false, Mp},  // d1 N=1..3: <m c||..>, t(a b..)
    };
    auto desc = [&](int g, int pair) {
      PassDesc d = D[g];
      d.pair = pair;
      return d;
    };
    int buf = 0;
    Stage s0 = make_stage(desc(0, 0), tid, HI, HJ, HK, no, nv), s1;
    {
      double2 r0, r1, r2;
      stage_load(s0, 0, r0, r1, r2);
      stage_store(lds, s0.pl, s0.ql0, s0.ql1, r0, r1, r2);
      __syncthreads();
    }
    // per-group lane slots: P particle slot and Q quad offset
    const int psA = p_swz(4 * ca + x, z), psB = p_swz(4 * cb + x, z), psC = p_swz(4 * cc + x, z);
    // Q(f, g): phase 1 stores the x slot from f (lane x = first index),
    // phase 2 from g (lane x = second index, b)
    const int qoBC = (4 * cc + y) * 8 + q_swz(4 * cb + x, 4 * cc + y, z);
    const int qoAC = (4 * cc + y) * 8 + q_swz(4 * ca + x, 4 * cc + y, z);
    const int qoAB = (4 * ca + y) * 8 + q_swz(4 * cb + x, 4 * ca + y, z);
#define TCE_T_PASS(PAIR, PIR, G, NG, NPAIR, PS, QO, SIGN, KP, LAST)                    \
  {                                                                                  \
    if (!(LAST)) s1 = make_stage(desc(NG, NPAIR), tid, HI, HJ, HK, no, nv);          \
    pass<PAIR, PIR>(acc, lds, buf, s0, KP, s1, !(LAST), PS, QO, z, SIGN);            \
    s0 = s1;                                                                         \
  }
    // ---- phase 1 (rows a = z, cols b = x, blocks c = y)
    TCE_T_PASS(0, true, 0, 0, 1, psA, qoBC, -1.0, Ep, false)   // d2 N=1
    TCE_T_PASS(1, true, 0, 0, 2, psA, qoBC, -1.0, Ep, false)   // d2 N=2
    TCE_T_PASS(2, true, 0, 1, 0, psA, qoBC, +1.0, Ep, false)   // d2 N=3
    TCE_T_PASS(0, false, 1, 1, 1, psB, qoAC, +1.0, Ep, false)  // d2 N=4
    TCE_T_PASS(1, false, 1, 1, 2, psB, qoAC, +1.0, Ep, false)  // d2 N=5
    TCE_T_PASS(2, false, 1, 2, 1, psB, qoAC, -1.0, Ep, false)  // d2 N=6
    TCE_T_PASS(1, true, 2, 2, 2, psA, qoBC, -1.0, Mp, false)   // d1 N=4
    TCE_T_PASS(2, true, 2, 2, 0, psA, qoBC, +1.0, Mp, false)   // d1 N=5
    TCE_T_PASS(0, true, 2, 3, 1, psA, qoBC, -1.0, Mp, false)   // d1 N=6
    TCE_T_PASS(1, false, 3, 3, 2, psB, qoAC, +1.0, Mp, false)  // d1 N=7
    TCE_T_PASS(2, false, 3, 3, 0, psB, qoAC, -1.0, Mp, false)  // d1 N=8
    TCE_T_PASS(0, false, 3, 4, 0, psB, qoAC, +1.0, Mp, false)  // d1 N=9
    switch_layout(acc, x, y, z);
    // ---- phase 2 (rows c = z, cols b = x, blocks a = y)
    TCE_T_PASS(0, true, 4, 4, 1, psC, qoAB, -1.0, Ep, false)   // d2 N=7
    TCE_T_PASS(1, true, 4, 4, 2, psC, qoAB, -1.0, Ep, false)   // d2 N=8
    TCE_T_PASS(2, true, 4, 5, 1, psC, qoAB, +1.0, Ep, false)   // d2 N=9
    TCE_T_PASS(1, true, 5, 5, 2, psC, qoAB, -1.0, Mp, false)   // d1 N=1
    TCE_T_PASS(2, true, 5, 5, 0, psC, qoAB, +1.0, Mp, false)   // d1 N=2
    TCE_T_PASS(0, true, 5, 5, 0, psC, qoAB, -1.0, Mp, true)    // d1 N=3
#undef TCE_T_PASS
    // Epilogue: both chunk buffers are free (the last chunk ended with a
    // barrier); stage the singles operands over them.
    const HoleBox H3[3] = {HI, HJ, HK};
    const PRange R3[3] = {RA, RB, RC};
    double *epi = reinterpret_cast<double *>(lds);
    stage_epilogue(epi, A, tid, H3, R3, no, nv);
    __syncthreads();
    double e4, e5;
    wg_energy(acc, epi, A, T, 4 * ca + y, 4 * cb + x, 4 * cc + z, a0 + 4 * ca + y, b0 + 4 * cb + x,
              c0 + 4 * cc + z, 4 * bi, 4 * bj, 4 * bk, H3, e4, e5);
    wave_sum(e4, e5);
    if (lane == 0) {
      A.partials[2 * (item * 8 + wave)] = e4;
      A.partials[2 * (item * 8 + wave) + 1] = e5;
    }
    __syncthreads();  // the next item's prologue overwrites the staged data
  }
}

}  // namespace

int launch_fused_wg(const Segment *segs, int nseg, long total, int grid, hipStream_t s) {
  hipLaunchKernelGGL(fused_triples_wg, dim3((unsigned)grid), dim3(512), 0, s, segs, nseg, total);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tce_t
