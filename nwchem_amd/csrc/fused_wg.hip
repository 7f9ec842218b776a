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
//   P  [16 hole pairs][4 k-pairs][8 particle slots]               512 quads
//   Q  [4 singles][4 k-pairs][8 lane-y slots][12-quad pitch]     1536 quads
// A fragment read of 8 consecutive lanes touches 8 distinct 4-bank groups
// (P: 4 consecutive quads, broadcast over y; Q: x + 12y mod 8 distinct).
#include "fused.h"

namespace tce_t {
namespace {

constexpr int KC = 8;                 // contraction values per chunk
constexpr int P_QUADS = 16 * 4 * 8;   // 512
constexpr int Q_PITCH = 12;
constexpr int Q_BLK = 8 * Q_PITCH;    // quads per (single, k-pair)
constexpr int Q_QUADS = 4 * 4 * Q_BLK;
constexpr int BUF_QUADS = P_QUADS + Q_QUADS;  // 2048 quads = 32 KiB

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

// One contraction with LDS staging.  P(pair, particle, k) =
// P[(h1*no+h2)*pS + particle*Kp + k]; Q(single, first, second, k) =
// Q[h*qS + (first*nv + second)*Kp + k].  QX: the lane-x particle of Q is its
// first index (phase 1) or its second (phase 2).  pslot: this lane's P
// particle slot; qoff: its Q quad offset (y-slot * pitch + x-slot).
template <int PAIR, bool P_IS_R, bool QX_FIRST>
__device__ __forceinline__ void pass(acc_t (&acc)[64], double2 *lds, int tid,
                                     const double *__restrict__ P, long pS, PRange pr,
                                     const double *__restrict__ Q, long qS, PRange fr, PRange gr,
                                     const HoleBox &hI, const HoleBox &hJ, const HoleBox &hK,
                                     int no, int nv, int Kp, int pslot, int qoff, int z,
                                     double sign) {
  // this thread's share of the staging (same for every chunk of the pass)
  const int prow = tid >> 2, pq = tid & 3;
  const int ppair = prow >> 3, ps = prow & 7;
  const int pu = ppair >> 2, pv = ppair & 3;
  const int ph1 = PAIR == 1 ? hole(hJ, pu) : hole(hI, pu);
  const int ph2 = PAIR == 0 ? hole(hJ, pv) : hole(hK, pv);
  const double *pg = P + ((long)ph1 * no + ph2) * pS + (pr.g0 + min(ps, pr.sm)) * Kp + 2 * pq;
  const int pl = ppair * 32 + pq * 8 + ps;
  const double *qg[2];
  int ql[2];
#pragma unroll
  for (int r = 0; r < 2; ++r) {
    const int idx = tid + 512 * r;
    const int row = idx >> 2, q = idx & 3;
    const int w = row >> 6, pp = row & 63, fs = pp >> 3, gs = pp & 7;
    const int hs = PAIR == 0 ? hole(hK, w) : (PAIR == 1 ? hole(hI, w) : hole(hJ, w));
    const long f = fr.g0 + min(fs, fr.sm), g = gr.g0 + min(gs, gr.sm);
    qg[r] = Q + (long)hs * qS + (f * nv + g) * Kp + 2 * q;
    const int sx = QX_FIRST ? fs : gs, sy = QX_FIRST ? gs : fs;
    ql[r] = P_QUADS + (w * 4 + q) * Q_BLK + sy * Q_PITCH + sx;
  }
  const int nch = Kp / KC;
  double2 sp = *(const double2 *)pg, sq0 = *(const double2 *)qg[0], sq1 = *(const double2 *)qg[1];
  lds[pl] = sp;
  lds[ql[0]] = sq0;
  lds[ql[1]] = sq1;
  __syncthreads();
  for (int c = 0; c < nch; ++c) {
    const double2 *cur = lds + (c & 1) * BUF_QUADS;
    double2 *nxt = lds + ((c + 1) & 1) * BUF_QUADS;
    const bool more = c + 1 < nch;
    if (more) {
      const int k = (c + 1) * KC;
      sp = *(const double2 *)(pg + k);
      sq0 = *(const double2 *)(qg[0] + k);
      sq1 = *(const double2 *)(qg[1] + k);
    }
    double2 qf[4];
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      const double2 t = cur[P_QUADS + (w * 4 + z) * Q_BLK + qoff];
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
        for (int v = 0; v < 4; ++v) pf[uu][v] = cur[((2 * h + uu) * 4 + v) * 32 + z * 8 + pslot];
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
    if (more) {
      nxt[pl] = sp;
      nxt[ql[0]] = sq0;
      nxt[ql[1]] = sq1;
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(512) fused_triples_wg(FusedArgs A) {
  __shared__ double2 lds[2 * BUF_QUADS];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int x = lane & 3, y = (lane >> 2) & 3, z = lane >> 4;
  const int ca = (wave >> 2) & 1, cb = (wave >> 1) & 1, cc = wave & 1;
  const int no = A.no, nv = A.nv, Ep = A.Ep, Mp = A.Mp;
  const long sT2e = (long)nv * Ep;      // stride of (h*no+h') in T2e
  const long sW = (long)nv * Mp;        // stride of (h*no+h') in Wd1
  const long sV = (long)nv * nv * Ep;   // stride of w in Vd2
  const long sT2m = (long)nv * nv * Mp; // stride of h in T2m
  // Persistent loop, static round-robin over work items (uniform across the
  // workgroup, so every barrier below is reached by all 8 waves).
  for (long item = blockIdx.x; item < A.nitems; item += gridDim.x) {
    const TaskDev &T = A.tasks[find_task(A.tasks, A.ntasks, item)];
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

    // ---- phase 1 (rows a = z, cols b = x, blocks c = y)
    {  // d2 N=1..3: P = t(a e h h'), Q = <b c||h e>
      const int ps = 4 * ca + x, qo = (4 * cc + y) * Q_PITCH + 4 * cb + x;
      pass<0, true, true>(acc, lds, tid, A.T2e, sT2e, RA, A.Vd2, sV, RB, RC, HI, HJ, HK, no, nv, Ep, ps, qo, z, -1.0);
      pass<1, true, true>(acc, lds, tid, A.T2e, sT2e, RA, A.Vd2, sV, RB, RC, HI, HJ, HK, no, nv, Ep, ps, qo, z, -1.0);
      pass<2, true, true>(acc, lds, tid, A.T2e, sT2e, RA, A.Vd2, sV, RB, RC, HI, HJ, HK, no, nv, Ep, ps, qo, z, +1.0);
    }
    {  // d2 N=4..6: P = t(b e h h'), Q = <a c||h e>
      const int ps = 4 * cb + x, qo = (4 * cc + y) * Q_PITCH + 4 * ca + x;
      pass<0, false, true>(acc, lds, tid, A.T2e, sT2e, RB, A.Vd2, sV, RA, RC, HI, HJ, HK, no, nv, Ep, ps, qo, z, +1.0);
      pass<1, false, true>(acc, lds, tid, A.T2e, sT2e, RB, A.Vd2, sV, RA, RC, HI, HJ, HK, no, nv, Ep, ps, qo, z, +1.0);
      pass<2, false, true>(acc, lds, tid, A.T2e, sT2e, RB, A.Vd2, sV, RA, RC, HI, HJ, HK, no, nv, Ep, ps, qo, z, -1.0);
    }
    {  // d1 N=4..6: P = <m a||h h'>, Q = t(b c h m)
      const int ps = 4 * ca + x, qo = (4 * cc + y) * Q_PITCH + 4 * cb + x;
      pass<1, true, true>(acc, lds, tid, A.Wd1, sW, RA, A.T2m, sT2m, RB, RC, HI, HJ, HK, no, nv, Mp, ps, qo, z, -1.0);
      pass<2, true, true>(acc, lds, tid, A.Wd1, sW, RA, A.T2m, sT2m, RB, RC, HI, HJ, HK, no, nv, Mp, ps, qo, z, +1.0);
      pass<0, true, true>(acc, lds, tid, A.Wd1, sW, RA, A.T2m, sT2m, RB, RC, HI, HJ, HK, no, nv, Mp, ps, qo, z, -1.0);
    }
    {  // d1 N=7..9: P = <m b||h h'>, Q = t(a c h m)
      const int ps = 4 * cb + x, qo = (4 * cc + y) * Q_PITCH + 4 * ca + x;
      pass<1, false, true>(acc, lds, tid, A.Wd1, sW, RB, A.T2m, sT2m, RA, RC, HI, HJ, HK, no, nv, Mp, ps, qo, z, +1.0);
      pass<2, false, true>(acc, lds, tid, A.Wd1, sW, RB, A.T2m, sT2m, RA, RC, HI, HJ, HK, no, nv, Mp, ps, qo, z, -1.0);
      pass<0, false, true>(acc, lds, tid, A.Wd1, sW, RB, A.T2m, sT2m, RA, RC, HI, HJ, HK, no, nv, Mp, ps, qo, z, +1.0);
    }
    switch_layout(acc, x, y, z);
    // ---- phase 2 (rows c = z, cols b = x, blocks a = y)
    {  // d2 N=7..9: P = t(c e h h'), Q = <a b||h e> (second index b on lane x)
      const int ps = 4 * cc + x, qo = (4 * ca + y) * Q_PITCH + 4 * cb + x;
      pass<0, true, false>(acc, lds, tid, A.T2e, sT2e, RC, A.Vd2, sV, RA, RB, HI, HJ, HK, no, nv, Ep, ps, qo, z, -1.0);
      pass<1, true, false>(acc, lds, tid, A.T2e, sT2e, RC, A.Vd2, sV, RA, RB, HI, HJ, HK, no, nv, Ep, ps, qo, z, -1.0);
      pass<2, true, false>(acc, lds, tid, A.T2e, sT2e, RC, A.Vd2, sV, RA, RB, HI, HJ, HK, no, nv, Ep, ps, qo, z, +1.0);
    }
    {  // d1 N=1..3: P = <m c||h h'>, Q = t(a b h m)
      const int ps = 4 * cc + x, qo = (4 * ca + y) * Q_PITCH + 4 * cb + x;
      pass<1, true, false>(acc, lds, tid, A.Wd1, sW, RC, A.T2m, sT2m, RA, RB, HI, HJ, HK, no, nv, Mp, ps, qo, z, -1.0);
      pass<2, true, false>(acc, lds, tid, A.Wd1, sW, RC, A.T2m, sT2m, RA, RB, HI, HJ, HK, no, nv, Mp, ps, qo, z, +1.0);
      pass<0, true, false>(acc, lds, tid, A.Wd1, sW, RC, A.T2m, sT2m, RA, RB, HI, HJ, HK, no, nv, Mp, ps, qo, z, -1.0);
    }
    int hI[4], hJ[4], hK[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      hI[u] = hole(HI, u);
      hJ[u] = hole(HJ, u);
      hK[u] = hole(HK, u);
    }
    double e4, e5;
    item_energy(acc, A, T, a0 + 4 * ca + y, b0 + 4 * cb + x, c0 + 4 * cc + z, 4 * bi, 4 * bj,
                4 * bk, hI, hJ, hK, e4, e5);
    wave_sum(e4, e5);
    if (lane == 0) {
      A.partials[2 * (item * 8 + wave)] = e4;
      A.partials[2 * (item * 8 + wave) + 1] = e5;
    }
  }
}

}  // namespace

int launch_fused_wg(const FusedArgs &A, int grid, hipStream_t s) {
  hipLaunchKernelGGL(fused_triples_wg, dim3((unsigned)grid), dim3(512), 0, s, A);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tce_t
