// fused.h -- device-side pieces shared by the fused (T) kernels.
//
// A work item is (task, particle box, 4x4x4 hole box) of one task's t3 tile.
// Every wavefront holds a 4x4x4 particle cube by 4x4x4 hole box (4096 t3
// elements) in 64 FP64 accumulators:
//   lane = x + 4y + 16z, register r = 16*ii + 4*jj + kk (hole offsets)
//   phase 1: (a,b,c) = (z, x, y)    rows a, cols b, blocks c
//   phase 2: (a,b,c) = (y, x, z)    rows c, cols b, blocks a
// v_mfma_f64_4x4x4_4b reads A[i][k] of block q from lane i+4q+16k, B[k][j]
// from lane j+4q+16k and writes D[i][j] to lane j+4q+16i, so a contraction
// runs on the accumulators iff its row index and column index sit on
// different operands; the block index may sit on either (the other operand
// is replicated).  Phase 1 carries the 12 contractions whose lone particle
// is a or b (d2 N=1..6, d1 N=4..9), phase 2 the 6 whose lone particle is c
// (d2 N=7..9, d1 N=1..3); one lane permutation switches the layouts.
//
// The 27 reference variants in task indices (a,b,c = p4,p5,p6 tiles,
// i,j,k = h1,h2,h3 tiles; t = t2 with the tce_sort signs folded in):
//   d2 N=1..3: -t(a e i j)<b c||k e>, -t(a e j k)<b c||i e>, +t(a e i k)<b c||j e>
//   d2 N=4..6: +t(b e i j)<a c||k e>, +t(b e j k)<a c||i e>, -t(b e i k)<a c||j e>
//   d2 N=7..9: -t(c e i j)<a b||k e>, -t(c e j k)<a b||i e>, +t(c e i k)<a b||j e>
//   d1 N=1..3: -t(a b i m)<m c||j k>, +t(a b j m)<m c||i k>, -t(a b k m)<m c||i j>
//   d1 N=4..6: -t(b c i m)<m a||j k>, +t(b c j m)<m a||i k>, -t(b c k m)<m a||i j>
//   d1 N=7..9: +t(a c i m)<m b||j k>, -t(a c j m)<m b||i k>, +t(a c k m)<m b||i j>
//   s1 N=1..9: +t(a i)<b c||j k>, -t(a j)<b c||i k>, +t(a k)<b c||i j>,
//              -t(b i)<a c||j k>, +t(b j)<a c||i k>, -t(b k)<a c||i j>,
//              +t(c i)<a b||j k>, -t(c j)<a b||i k>, +t(c k)<a b||i j>
// (ccsd_t_kernels_omp.F layouts and signs under the row tables of
// ccsd_t_singles_gpu.F / ccsd_t_doubles_gpu.F).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/tce_t.h"

namespace tce_t {

// The 18 d1/d2 contractions of fused_mixed.hip in its pass order: kind
// (0 = d2, 1 = d1), lone particle L (0 a, 1 b, 2 c), lone hole h (0 i, 1 j,
// 2 k), sign (1 = minus).  The contraction index x (e of d2, m of d1) pairs
// with L and the hole pair (r, s) = the two holes other than h on one
// operand (t(L e r s), <m L||r s>), so spin and irrep conservation of that
// block fix x's spin to spin(r) + spin(s) - spin(L) and its irrep to
// irrep(L) ^ irrep(r) ^ irrep(s).
struct PassDef {
  int kind, L, h, minus;
};
constexpr PassDef MIXED_PASS[18] = {
    {0, 0, 2, 1},  // d2 N=1: -t(a e i j)<b c||k e>
    {1, 0, 2, 1},  // d1 N=6: -t(b c k m)<m a||i j>
    {0, 0, 1, 0},  // d2 N=3: +t(a e i k)<b c||j e>
    {1, 0, 1, 0},  // d1 N=5: +t(b c j m)<m a||i k>
    {0, 0, 0, 1},  // d2 N=2: -t(a e j k)<b c||i e>
    {1, 0, 0, 1},  // d1 N=4: -t(b c i m)<m a||j k>
    {0, 1, 0, 0},  // d2 N=5: +t(b e j k)<a c||i e>
    {1, 1, 0, 0},  // d1 N=7: +t(a c i m)<m b||j k>
    {0, 1, 2, 0},  // d2 N=4: +t(b e i j)<a c||k e>
    {1, 1, 2, 0},  // d1 N=9: +t(a c k m)<m b||i j>
    {0, 1, 1, 1},  // d2 N=6: -t(b e i k)<a c||j e>
    {1, 1, 1, 1},  // d1 N=8: -t(a c j m)<m b||i k>
    {0, 2, 2, 1},  // d2 N=7: -t(c e i j)<a b||k e>
    {1, 2, 2, 1},  // d1 N=3: -t(a b k m)<m c||i j>
    {0, 2, 1, 0},  // d2 N=9: +t(c e i k)<a b||j e>
    {1, 2, 1, 0},  // d1 N=2: +t(a b j m)<m c||i k>
    {0, 2, 0, 1},  // d2 N=8: -t(c e j k)<a b||i e>
    {1, 2, 0, 1},  // d1 N=1: -t(a b i m)<m c||j k>
};
// hole pair (r, s) of lone hole h
constexpr int pair_r(int h) { return h == 0 ? 1 : 0; }
constexpr int pair_s(int h) { return h == 2 ? 1 : 2; }

struct TaskDev {
  int op4, op5, op6;  // particle-space offsets of the task's particle tiles
  int oh1, oh2, oh3;  // hole offsets
  int r4, r5, r6, r1, r2, r3;
  int nb4, nb5, nb6, nb1, nb2, nb3;  // boxes per index (particle box edge pbox, hole 4)
  long item0;                        // first work item of the task
  double factor;
  // Contraction range of pass p of MIXED_PASS: orbitals [x0[p], x1[p]) of
  // the virtual (d2) or occupied (d1) space hold every orbital of the spin
  // and irrep the pass's blocks allow (the reference skips the other
  // h7b / p7b tiles, ccsd_t_doubles_gpu.F:260-263, 924-927); outside it
  // every operand of the pass is zero.  x0 = x1 = 0: no orbital allowed.
  unsigned short x0[18], x1[18];
};

struct FusedArgs {
  const double *T1, *T2e, *T2m, *Vd2, *Wd1, *S, *evl;
  const TaskDev *tasks;
  int ntasks;
  long nitems;
  int no, nv, Ep, Mp;  // orbitals, padded contraction extents
  double *partials;    // {E[T], E(T)} per wave of each item
};

// Reduced-precision operands of the d1/d2 contractions (fused_mixed.hip):
// the dense tensors of FusedArgs converted to FP32 (part 0) or split into
// bf16 hi (part 0) and lo (part 1), or hi, mid, lo (parts 0-2), contraction
// index padded with zeros to Ek (d2) / Mk (d1) elements.
struct MixedArgs {
  FusedArgs f;
  const void *x2[3], *y2[3];  // d2: T2e, Vd2 (parts: FP32 | hi, lo | hi, mid, lo)
  const void *x1[3], *y1[3];  // d1: Wd1, T2m
  int Ek, Mk;
  // Lambda-CCSD(T) passes (lambda.hip): the window of left-hand tiles of
  // items [lam_item0, lam_item0 + window) read by the right-hand pass, and
  // the Fock hole-particle block F[h][p] (or null) the left-hand pass adds.
  const double *lam_y = nullptr, *lam_f = nullptr;
  long lam_item0 = 0;
  // CR-CCSD(T) right-hand pass: the denominator partials (2 per wave, as
  // f.partials holds the moment partials).
  double *cr_den = nullptr;
};

// A launch covers a table of segments: work items [item0, item1) of the job
// whose arguments sit at *args (device memory); g0 is the segment's first
// index in the launch's global item range (segments sorted by g0).  One job
// is one segment; a batch of jobs, or a slice of several jobs' remaining
// items, is one launch over several.
template <class Args>
struct SegmentT {
  const Args *args;
  long item0, item1, g0;
};
typedef SegmentT<FusedArgs> Segment;
typedef SegmentT<MixedArgs> MixedSegment;

// Pointers read from memory are generic to the compiler, which then issues
// flat (per-lane) loads and treats what they return as divergent.  Every
// pointer of the argument tables is global memory: the kernels read the
// tables into these mirrors (same layout, pointers typed global), which keeps
// the loads scalar/global and the values uniform.
template <class T>
using gptr = __attribute__((address_space(1))) T *;

struct FusedArgsG {
  gptr<const double> T1, T2e, T2m, Vd2, Wd1, S, evl;
  gptr<const TaskDev> tasks;
  int ntasks;
  long nitems;
  int no, nv, Ep, Mp;
  gptr<double> partials;
};
struct MixedArgsG {
  FusedArgsG f;
  gptr<const void> x2[3], y2[3];
  gptr<const void> x1[3], y1[3];
  int Ek, Mk;
  gptr<const double> lam_y, lam_f;
  long lam_item0;
  gptr<double> cr_den;
};
static_assert(sizeof(FusedArgsG) == sizeof(FusedArgs), "argument mirror");
static_assert(sizeof(MixedArgsG) == sizeof(MixedArgs), "argument mirror");

__device__ __forceinline__ FusedArgsG load_args(const FusedArgs *p) {
  const gptr<const FusedArgs> q = (gptr<const FusedArgs>)p;
  FusedArgsG A;
  A.T1 = (gptr<const double>)q->T1;
  A.T2e = (gptr<const double>)q->T2e;
  A.T2m = (gptr<const double>)q->T2m;
  A.Vd2 = (gptr<const double>)q->Vd2;
  A.Wd1 = (gptr<const double>)q->Wd1;
  A.S = (gptr<const double>)q->S;
  A.evl = (gptr<const double>)q->evl;
  A.tasks = (gptr<const TaskDev>)q->tasks;
  A.ntasks = q->ntasks;
  A.nitems = q->nitems;
  A.no = q->no;
  A.nv = q->nv;
  A.Ep = q->Ep;
  A.Mp = q->Mp;
  A.partials = (gptr<double>)q->partials;
  return A;
}

__device__ __forceinline__ MixedArgsG load_args(const MixedArgs *p) {
  const gptr<const MixedArgs> q = (gptr<const MixedArgs>)p;
  MixedArgsG M;
  M.f = load_args(&p->f);
  for (int i = 0; i < 3; ++i) {
    M.x2[i] = (gptr<const void>)q->x2[i];
    M.y2[i] = (gptr<const void>)q->y2[i];
    M.x1[i] = (gptr<const void>)q->x1[i];
    M.y1[i] = (gptr<const void>)q->y1[i];
  }
  M.Ek = q->Ek;
  M.Mk = q->Mk;
  M.lam_y = (gptr<const double>)q->lam_y;
  M.lam_f = (gptr<const double>)q->lam_f;
  M.lam_item0 = q->lam_item0;
  M.cr_den = (gptr<double>)q->cr_den;
  return M;
}

// First item of workgroup b in a persistent launch of G workgroups.  The
// hardware deals workgroups to the 8 XCDs round-robin (b -> XCD b % 8), and
// each XCD has its own L2; mapping b to (b % 8) * (G / 8) + b / 8 hands each
// XCD a contiguous run of items, so the items that share operand slices
// (consecutive boxes of one task) meet in the same L2.  The loop then steps
// by G as before; any G works (a G that is not a multiple of 8 keeps b).
__device__ __forceinline__ long first_item(unsigned b, unsigned G) {
  return (G % 8u) ? (long)b : (long)(b % 8u) * (G / 8u) + b / 8u;
}

// Item of workgroup b in round r of a persistent launch (round r covers
// items r*G .. r*G + G-1): the XCD-aware position, mirrored in odd rounds
// when `mirror` (launches over several jobs).  In a batch the items of one
// job follow those of the previous (costlier, in cost order) job, so in each
// round where the job changes the low positions draw the costlier items;
// mirroring every other round keeps any workgroup from drawing them every
// time.  One job's items cost the same, and there the unmirrored order keeps
// each XCD on the same slice of consecutive rounds (half the L2 misses).
// Uniform across the workgroup.
__device__ __forceinline__ long round_item(long r, unsigned b, unsigned G, bool mirror) {
  const long f = first_item(b, G);
  return r * (long)G + ((mirror && (r & 1)) ? (long)G - 1 - f : f);
}

// Dynamic schedule of a persistent launch (fused_triples_mixed): the global
// items are cut into windows of 8 slices, and queue q (one per XCD) holds
// slice q of every window, in window order.  A workgroup draws tickets
// from the queue of its own XCD (b % 8) with a device-scope atomic add, so
// the 32 CUs of an XCD share one slice of consecutive items at a time (their
// operand rows meet in that XCD's L2) while each CU takes the next item when
// it is free: no CU drifts ahead of the others over a long launch, and
// items of uneven cost balance themselves.  A workgroup whose queue is
// exhausted takes tickets from the next queues in turn, and leaves when all
// eight are.  Counters: sched[0..7] the queues' next tickets, sched[8] the
// workgroups that have left; they are zero before a launch, and the last
// workgroup to leave zeroes them for the next launch on the same stream.
// sched[9] holds log2 of the slice width (set once with the counters).
constexpr int SCHED_WORDS = 9;     // counters the last workgroup zeroes
constexpr int SCHED_ALLOC = 10;    // counters + the slice width

__device__ __forceinline__ long queue_item(unsigned long long k, unsigned q, int sh) {
  return (long)((k >> sh) << (sh + 3)) + ((long)q << sh) + (long)(k & ((1ull << sh) - 1));
}

// Item of ticket k of queue q, or of a ticket of the next queues when q is
// exhausted (q then moves on); -1 when every queue is.  One thread.
__device__ __forceinline__ long sched_item(unsigned long long *sched, unsigned &q,
                                           unsigned long long k, long n, int sh) {
  for (int tried = 1;; ++tried) {
    const long g = queue_item(k, q, sh);
    if (g < n) return g;
    if (tried == 8) return -1;
    q = (q + 1) & 7u;
    k = atomicAdd(&sched[q], 1ull);
  }
}

// A workgroup leaves the dynamic schedule (one thread, after its last
// ticket): the last one to leave zeroes the counters.
__device__ __forceinline__ void sched_leave(unsigned long long *sched) {
  __threadfence();
  if (atomicAdd(&sched[8], 1ull) == (unsigned long long)gridDim.x - 1) {
    for (int i = 0; i < SCHED_WORDS; ++i) atomicExch(&sched[i], 0ull);
    __threadfence();
  }
}

// Segment of global index g for a workgroup whose previous item lay in
// segment s (uniform).  Under the static schedule, and within one queue of
// the dynamic one, a workgroup's global indices increase from item to item,
// so its segment only moves forward: one read of the next segment's start
// per item replaces a binary search of the table.  An item taken from
// another queue can lie before the previous one: the scan restarts.  Empty
// segments are stepped over.
template <class Args>
__device__ __forceinline__ int next_segment(const SegmentT<Args> *segs, int nseg, int s, long g) {
  if (g < segs[s].g0) s = 0;
  while (s + 1 < nseg && segs[s + 1].g0 <= g) ++s;
  return s;
}

typedef double acc_t;

__device__ __forceinline__ double mfma4(double a, double b, double c) {
  return __builtin_amdgcn_mfma_f64_4x4x4f64(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ double perm_lane(double v, int src_lane) {
  long long b = __double_as_longlong(v);
  int lo = __builtin_amdgcn_ds_bpermute(src_lane << 2, (int)(b & 0xffffffffLL));
  int hi = __builtin_amdgcn_ds_bpermute(src_lane << 2, (int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Accumulator index of (pair (u,v), single w) for the three pair choices
// (0: pair (i,j) single k, 1: pair (j,k) single i, 2: pair (i,k) single j).
template <int PAIR>
__device__ __forceinline__ constexpr int acc_index(int u, int v, int w) {
  return PAIR == 0 ? (u * 4 + v) * 4 + w : (PAIR == 1 ? (w * 4 + u) * 4 + v : (u * 4 + w) * 4 + v);
}

// 64 MFMAs of one contraction step: P fragments (pairs) x Q fragments
// (singles), two K values per lane held in .x / .y.
template <int PAIR, bool P_IS_R>
__device__ __forceinline__ void mma_step(acc_t (&acc)[64], const double2 (&pf)[4][4],
                                         const double2 (&qf)[4]) {
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int w = 0; w < 4; ++w) {
          const int r = acc_index<PAIR>(u, v, w);
          const double pv = t ? pf[u][v].y : pf[u][v].x;
          const double qv = t ? qf[w].y : qf[w].x;
          acc[r] = P_IS_R ? mfma4(pv, qv, acc[r]) : mfma4(qv, pv, acc[r]);
        }
}

// Move the accumulators from the phase-1 to the phase-2 layout: element
// (a,b,c) goes from lane b+4c+16a to lane b+4a+16c.
__device__ __forceinline__ void switch_layout(acc_t (&acc)[64], int x, int y, int z) {
  const int src = x + 4 * z + 16 * y;
#pragma unroll
  for (int r = 0; r < 64; ++r) acc[r] = perm_lane(acc[r], src);
}

// Singles, denominators and the energy of one wave's 4096 elements (phase-2
// layout).  la/lb/lc: this lane's particle indices within the task's tiles;
// i0/j0/k0: first hole of the box within the tiles; hI/hJ/hK: global hole
// indices (clamped).  Elements past a ragged tile edge are skipped.
// denom = d * factor / (e_h1 + e_h2 + e_h3 + denom_0) as ccsd_t_dot.F:107.
__device__ __forceinline__ void item_energy(const acc_t (&acc)[64], const FusedArgsG &A,
                                            const TaskDev &T, int la, int lb, int lc, int i0,
                                            int j0, int k0, const int (&hI)[4], const int (&hJ)[4],
                                            const int (&hK)[4], double &e4, double &e5) {
  e4 = 0.0;
  e5 = 0.0;
  if (!(la < T.r4 && lb < T.r5 && lc < T.r6)) return;
  const int no = A.no, nv = A.nv;
  const long pa = T.op4 + la, pb = T.op5 + lb, pc = T.op6 + lc;
  const double ea = A.evl[no + pa], eb = A.evl[no + pb], ec = A.evl[no + pc];
  const double denom_0 = -(ea + eb + ec);
  const gptr<const double> T1 = A.T1, S = A.S;
  const long sSh = (long)nv * nv;  // stride of (h*no+h') in S
  const long pbc = pb * nv + pc, pac = pa * nv + pc, pab = pa * nv + pb;
#pragma unroll
  for (int ii = 0; ii < 4; ++ii) {
    if (i0 + ii >= T.r1) continue;
    const int hi = hI[ii];
#pragma unroll
    for (int jj = 0; jj < 4; ++jj) {
      if (j0 + jj >= T.r2) continue;
      const int hj = hJ[jj];
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        if (k0 + kk >= T.r3) continue;
        const int hk = hK[kk];
        const double d = acc[ii * 16 + jj * 4 + kk];
        const long jk = ((long)hj * no + hk) * sSh, ik = ((long)hi * no + hk) * sSh,
                   ij = ((long)hi * no + hj) * sSh;
        double s = T1[pa * no + hi] * S[jk + pbc];
        s -= T1[pa * no + hj] * S[ik + pbc];
        s += T1[pa * no + hk] * S[ij + pbc];
        s -= T1[pb * no + hi] * S[jk + pac];
        s += T1[pb * no + hj] * S[ik + pac];
        s -= T1[pb * no + hk] * S[ij + pac];
        s += T1[pc * no + hi] * S[jk + pab];
        s -= T1[pc * no + hj] * S[ik + pab];
        s += T1[pc * no + hk] * S[ij + pab];
        const double den = d * T.factor / (A.evl[hi] + A.evl[hj] + A.evl[hk] + denom_0);
        e4 += den * d;
        e5 += den * (d + s);
      }
    }
  }
}

__device__ __forceinline__ void wave_sum(double &e4, double &e5) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    e4 += __shfl_xor(e4, off);
    e5 += __shfl_xor(e5, off);
  }
}

// Locate the task of a work item (tasks sorted by item0, item0[t] >= t) --
// uniform.  Sub-tiled tables have one item per task (item0[t] == t): one
// probe finds it.
__device__ __forceinline__ int find_task(gptr<const TaskDev> tasks, int ntasks, long item) {
  if (item < ntasks && tasks[item].item0 == item) return (int)item;
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (tasks[mid].item0 <= item)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Host launchers (one file per kernel family): `total` items over `nseg`
// segments (device table), `grid` persistent workgroups.
int launch_fused_wave(const Segment *segs, int nseg, long total, int grid, hipStream_t s);
int launch_fused_wg(const Segment *segs, int nseg, long total, int grid, hipStream_t s);
// sparse: some job of the table narrows its passes' contraction ranges
// (TaskDev::x0/x1, spin- or irrep-labelled tiles).
// sched: the stream's counters of the dynamic schedule (sched_alloc), or
// null for the static round-robin.
int launch_fused_mixed(const MixedSegment *segs, int nseg, long total, int precision, int grid,
                       hipStream_t s, bool sparse, unsigned long long *sched);
// One task's t3 tiles (fused_mixed.hip): `total` items of the single task at
// args->f.tasks[0], tiles written to args->f.partials (doubles, then singles).
int launch_fused_tiles(const MixedSegment *segs, long total, int precision, int grid,
                       hipStream_t s, bool sparse, unsigned long long *sched);
// The Lambda-CCSD(T) passes over one segment of a job's items (8-wide
// boxes).  left = true: the left-hand job's doubles (plus the Fock term of
// args->lam_f) and singles of each wave's 4096 elements, in register order,
// into args->f.partials at window slot (item - lam_item0) * 8 + wave (8192
// doubles per slot).  left = false: the right-hand job's energy partials
// factor * t3d * (y3 doubles) / D and factor * t3d * (y3 doubles + singles)
// / D with the left-hand values read from args->lam_y at the same slots.
int launch_fused_lambda(const MixedSegment *segs, long total, int precision, int grid,
                        hipStream_t s, bool sparse, unsigned long long *sched, bool left);
// The CR-CCSD(T) right-hand pass (in `precision`): the right-hand job's t3 doubles
// (the moment M3 of cr_ccsd_t_N) and singles (the denominator tile of
// cr_ccsd_t_E) against the left-hand job's t3 doubles and singles in the
// window at args->lam_y: per wave factor * M3 * td / D and factor * M3 *
// (td + ts) / D into args->f.partials, the same with the denominator tile
// into args->cr_den.
int launch_fused_cr_right(const MixedSegment *segs, long total, int precision, int grid,
                          hipStream_t s, bool sparse, unsigned long long *sched);
// Zeroed counters of the dynamic schedule for one launching stream (every
// launch leaves them zero; launches that may overlap need their own), or
// null on failure (error set).  Free with hipFree.
unsigned long long *sched_alloc();
// The counters to pass for a stream's counters `own`: null (the static
// round-robin) when TCE_T_SCHED=static.
unsigned long long *sched_use(unsigned long long *own);

}  // namespace tce_t
