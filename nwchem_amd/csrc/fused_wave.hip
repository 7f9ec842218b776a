// fused_wave.hip -- one wavefront per work item (particle box 4x4x4), operand
// fragments read straight from global memory (L1/L2).  Used for small or
// ragged tilings where the 8-wide particle boxes of fused_wg.hip would waste
// work; see fused.h for the register layout and the variant table.
#include "fused.h"

namespace tce_t {
namespace {

// One contraction on the accumulators.  PAIR selects the two hole indices
// the "pair" operand P carries (0: i,j  1: j,k  2: i,k); the third hole
// indexes the "single" operand Q.  P_IS_R: P supplies the MFMA rows (A).
// Pb/Qb are wave-uniform pointers (operand origin of the item's particle
// cube); pLane/qLane are the lane's small offsets (doubles) from there, so
// every load is SGPR base + 32-bit VGPR offset.  The contraction index runs
// contiguously, two values per lane per load (lane z holds k0+2z, k0+2z+1).
template <int PAIR, bool P_IS_R>
__device__ __forceinline__ void contract(acc_t (&acc)[64], const double *__restrict__ Pb,
                                         unsigned pLane, long pS, const double *__restrict__ Qb,
                                         unsigned qLane, long qS, const int (&hI)[4],
                                         const int (&hJ)[4], const int (&hK)[4], int no, int Kp,
                                         int z, double sign) {
  const char *pp[4][4], *qp[4];
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      int h1 = PAIR == 1 ? hJ[u] : hI[u];
      int h2 = PAIR == 0 ? hJ[v] : hK[v];
      pp[u][v] = (const char *)(Pb + ((long)h1 * no + h2) * pS);
    }
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    int hs = PAIR == 0 ? hK[w] : (PAIR == 1 ? hI[w] : hJ[w]);
    qp[w] = (const char *)(Qb + (long)hs * qS);
  }
  const unsigned po = (pLane + 2u * z) * 8u, qo = (qLane + 2u * z) * 8u;
  for (int k0 = 0; k0 < Kp; k0 += 8) {
    const unsigned kb = (unsigned)k0 * 8u;
    double2 pf[4][4], qf[4];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) pf[u][v] = *(const double2 *)(pp[u][v] + (po + kb));
#pragma unroll
    for (int w = 0; w < 4; ++w) {
      double2 t = *(const double2 *)(qp[w] + (qo + kb));
      qf[w].x = sign * t.x;
      qf[w].y = sign * t.y;
    }
    mma_step<PAIR, P_IS_R>(acc, pf, qf);
  }
}

__global__ void __launch_bounds__(64, 2) fused_triples(const Segment *segs, int nseg, long total) {
  const int lane = threadIdx.x;
  const int x = lane & 3, y = (lane >> 2) & 3, z = lane >> 4;
  // Persistent loop, static round-robin over the launch's items: the item
  // index is wave-uniform by construction and every wave leaves after at
  // most ceil(total / gridDim.x) items.
  int seg = 0;  // segment of the workgroup's current item
  for (long r = 0; r * (long)gridDim.x < total; ++r) {
    const long g = round_item(r, blockIdx.x, gridDim.x, nseg > 1);
    if (g >= total) continue;  // last round: uniform per workgroup
    seg = next_segment(segs, nseg, seg, g);
    const Segment &sg = segs[seg];
    const long item = sg.item0 + (g - sg.g0);
    const FusedArgsG A = load_args(sg.args);
    const int no = A.no, nv = A.nv, Ep = A.Ep, Mp = A.Mp;
    const long nvE = (long)nv * Ep, nvM = (long)nv * Mp;
    const TaskDev T = *(const TaskDev *)(A.tasks + find_task(A.tasks, A.ntasks, item));
    long box = item - T.item0;
    int bc = box % T.nb6; box /= T.nb6;
    int bb = box % T.nb5; box /= T.nb5;
    int ba = box % T.nb4; box /= T.nb4;
    int bk = box % T.nb3; box /= T.nb3;
    int bj = box % T.nb2; box /= T.nb2;
    int bi = (int)box;
    const int a0 = 4 * ba, b0 = 4 * bb, c0 = 4 * bc;
    int hI[4], hJ[4], hK[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      hI[u] = T.oh1 + min(4 * bi + u, T.r1 - 1);
      hJ[u] = T.oh2 + min(4 * bj + u, T.r2 - 1);
      hK[u] = T.oh3 + min(4 * bk + u, T.r3 - 1);
    }
    // Cube origin (global particle indices, wave-uniform) and the lane's
    // offsets from it, clamped into the tile (lanes past a ragged tile edge
    // compute duplicates that the energy mask drops).
    const long A0 = T.op4 + a0, B0 = T.op5 + b0, C0 = T.op6 + c0;
    const unsigned dxa = min(x, T.r4 - 1 - a0), dya = min(y, T.r4 - 1 - a0);
    const unsigned dxb = min(x, T.r5 - 1 - b0);
    const unsigned dxc = min(x, T.r6 - 1 - c0), dyc = min(y, T.r6 - 1 - c0);

    acc_t acc[64];
#pragma unroll
    for (int r = 0; r < 64; ++r) acc[r] = 0.0;

    const long sT2e = nvE;             // stride of (h*no+h') in T2e
    const long sW = nvM;               // stride of (h*no+h') in Wd1
    const long sV = (long)nv * nvE;    // stride of w in Vd2
    const long sT2m = (long)nv * nvM;  // stride of h in T2m
    const unsigned uEp = Ep, uMp = Mp, unv = nv;
    // ---- phase 1: lone particle a or b
    {  // d2 N=1..3: rows a from t(a e h h'), cols b / blocks c from <b c||h e>
      const double *P = (const double *)(A.T2e + A0 * Ep), *Q = (const double *)(A.Vd2 + (B0 * nv + C0) * Ep);
      const unsigned pl = dxa * uEp, ql = (dxb * unv + dyc) * uEp;
      contract<0, true>(acc, P, pl, sT2e, Q, ql, sV, hI, hJ, hK, no, Ep, z, -1.0);
      contract<1, true>(acc, P, pl, sT2e, Q, ql, sV, hI, hJ, hK, no, Ep, z, -1.0);
      contract<2, true>(acc, P, pl, sT2e, Q, ql, sV, hI, hJ, hK, no, Ep, z, +1.0);
    }
    {  // d2 N=4..6: rows a / blocks c from <a c||h e>, cols b from t(b e h h')
      const double *P = (const double *)(A.T2e + B0 * Ep), *Q = (const double *)(A.Vd2 + (A0 * nv + C0) * Ep);
      const unsigned pl = dxb * uEp, ql = (dxa * unv + dyc) * uEp;
      contract<0, false>(acc, P, pl, sT2e, Q, ql, sV, hI, hJ, hK, no, Ep, z, +1.0);
      contract<1, false>(acc, P, pl, sT2e, Q, ql, sV, hI, hJ, hK, no, Ep, z, +1.0);
      contract<2, false>(acc, P, pl, sT2e, Q, ql, sV, hI, hJ, hK, no, Ep, z, -1.0);
    }
    {  // d1 N=4..6: rows a from <m a||h h'>, cols b / blocks c from t(b c h m)
      const double *P = (const double *)(A.Wd1 + A0 * Mp), *Q = (const double *)(A.T2m + (B0 * nv + C0) * Mp);
      const unsigned pl = dxa * uMp, ql = (dxb * unv + dyc) * uMp;
      contract<1, true>(acc, P, pl, sW, Q, ql, sT2m, hI, hJ, hK, no, Mp, z, -1.0);
      contract<2, true>(acc, P, pl, sW, Q, ql, sT2m, hI, hJ, hK, no, Mp, z, +1.0);
      contract<0, true>(acc, P, pl, sW, Q, ql, sT2m, hI, hJ, hK, no, Mp, z, -1.0);
    }
    {  // d1 N=7..9: rows a / blocks c from t(a c h m), cols b from <m b||h h'>
      const double *P = (const double *)(A.Wd1 + B0 * Mp), *Q = (const double *)(A.T2m + (A0 * nv + C0) * Mp);
      const unsigned pl = dxb * uMp, ql = (dxa * unv + dyc) * uMp;
      contract<1, false>(acc, P, pl, sW, Q, ql, sT2m, hI, hJ, hK, no, Mp, z, +1.0);
      contract<2, false>(acc, P, pl, sW, Q, ql, sT2m, hI, hJ, hK, no, Mp, z, -1.0);
      contract<0, false>(acc, P, pl, sW, Q, ql, sT2m, hI, hJ, hK, no, Mp, z, +1.0);
    }
    switch_layout(acc, x, y, z);
    // ---- phase 2: lone particle c
    {  // d2 N=7..9: rows c from t(c e h h'), cols b / blocks a from <a b||h e>
      const double *P = (const double *)(A.T2e + C0 * Ep), *Q = (const double *)(A.Vd2 + (A0 * nv + B0) * Ep);
      const unsigned pl = dxc * uEp, ql = (dya * unv + dxb) * uEp;
      contract<0, true>(acc, P, pl, sT2e, Q, ql, sV, hI, hJ, hK, no, Ep, z, -1.0);
      contract<1, true>(acc, P, pl, sT2e, Q, ql, sV, hI, hJ, hK, no, Ep, z, -1.0);
      contract<2, true>(acc, P, pl, sT2e, Q, ql, sV, hI, hJ, hK, no, Ep, z, +1.0);
    }
    {  // d1 N=1..3: rows c from <m c||h h'>, cols b / blocks a from t(a b h m)
      const double *P = (const double *)(A.Wd1 + C0 * Mp), *Q = (const double *)(A.T2m + (A0 * nv + B0) * Mp);
      const unsigned pl = dxc * uMp, ql = (dya * unv + dxb) * uMp;
      contract<1, true>(acc, P, pl, sW, Q, ql, sT2m, hI, hJ, hK, no, Mp, z, -1.0);
      contract<2, true>(acc, P, pl, sW, Q, ql, sT2m, hI, hJ, hK, no, Mp, z, +1.0);
      contract<0, true>(acc, P, pl, sW, Q, ql, sT2m, hI, hJ, hK, no, Mp, z, -1.0);
    }
    double e4, e5;
    item_energy(acc, A, T, a0 + y, b0 + x, c0 + z, 4 * bi, 4 * bj, 4 * bk, hI, hJ, hK, e4, e5);
    wave_sum(e4, e5);
    if (lane == 0) {
      A.partials[2 * item] = e4;
      A.partials[2 * item + 1] = e5;
    }
  }
}

}  // namespace

int launch_fused_wave(const Segment *segs, int nseg, long total, int grid, hipStream_t s) {
  hipLaunchKernelGGL(fused_triples, dim3((unsigned)grid), dim3(64), 0, s, segs, nseg, total);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // namespace tce_t
