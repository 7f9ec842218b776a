// job.hip -- whole-job (T): the tile-sextuple loop of ccsd_t_gpu.F:87-241 as
// one device-resident job on MI355X.
//
// Data path
//   upload   T1/T2/V2 canonical TCE blocks are read through the reference's
//            offset hash tables (get_hash_block semantics, tce_hashnsort.F
//            signs, tce_restricted mapping) and re-laid out once in HBM as
//            dense tensors whose contraction index is contiguous:
//              T2e[h][h'][p][e]  = t(p e h h')       (d2 operand, e = p7)
//              T2m[h][p][p'][m]  = t(p p' h m)       (d1 operand, m = h7)
//              Vd2[w][p][p'][e]  = <p p'||w e>       (d2 integral)
//              Wd1[h][h'][p][m]  = <m p||h h'>       (d1 integral)
//              S  [h][h'][p][p'] = <p p'||h h'>      (s1 integral)
//              T1 [p][h]         = t(p h)
//            (contraction extents padded with zeros to a multiple of 8).
//   run      one persistent fused kernel over the job's work items (task x
//            particle box x 4x4x4 hole box): fused_triples_wg (8-wide
//            particle boxes, LDS-staged operands), fused_triples (4-wide
//            boxes) for small or ragged tilings, or fused_triples_mixed in a
//            reduced precision.  Each item's t3 block lives in registers for
//            all 18 d1/d2 contractions, then gets its 9 s1 products and the
//            denominator-weighted energy in FP64; one {E[T], E(T)} partial
//            per wave.  A second kernel sums each task's partials in a fixed
//            order (bit-reproducible task energies); the host sums tasks in
//            task-loop order.
//   reduced  the first run in FP32 / split-BF16 converts the four
//            contraction tensors on the device (zero-padded to the kernel's
//            chunk) and builds the 8-wide task table when the job's own
//            kernel uses 4-wide boxes.
//
// Kernel launches take a table of (job arguments, item range) segments in
// device memory (one segment here; many in batch.hip and server.cpp).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <thread>
#include <vector>

#include "job.h"

namespace tce_t {
namespace {

// Per-task fixed-order reduction of the partials (ppi partials per item)
// over the items [bounds[t], bounds[t+1]) of reference task t.
__global__ void __launch_bounds__(256) task_reduce(const long *__restrict__ bounds, int ntasks,
                                                   int ppi, const double *__restrict__ part,
                                                   double *__restrict__ out) {
  const int t = blockIdx.x;
  if (t >= ntasks) return;
  const long i0 = bounds[t] * ppi, i1 = bounds[t + 1] * ppi;
  double s4 = 0.0, s5 = 0.0;
  for (long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    s4 += part[2 * i];
    s5 += part[2 * i + 1];
  }
  __shared__ double r4[256], r5[256];
  r4[threadIdx.x] = s4;
  r5[threadIdx.x] = s5;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r4[threadIdx.x] += r4[threadIdx.x + w];
      r5[threadIdx.x] += r5[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * t] = r4[0];
    out[2 * t + 1] = r5[0];
  }
}

// Rows of K valid FP64 values (row stride kin) -> rows of kout reduced-
// precision values, zero-padded: FP32 (out0), or bf16 hi = bf16(x) (out0)
// and lo = bf16(x - hi) (out1), or hi, mid = bf16(x - hi), lo = bf16(x -
// hi - mid) (out0..2), all rounded to nearest even.
__device__ __forceinline__ unsigned short bf16_rne(float f) {
  unsigned u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

__global__ void convert_rows(const double *__restrict__ in, long rows, int K, int kin, int kout,
                             int precision, void *out0, void *out1, void *out2) {
  const long n = rows * kout;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const long r = i / kout;
    const int k = (int)(i - r * kout);
    const double x = k < K ? in[r * kin + k] : 0.0;
    if (precision == TCE_T_FP32) {
      ((float *)out0)[i] = (float)x;
    } else if (precision == TCE_T_BF16X3) {
      const unsigned short h = bf16_rne((float)x);
      const double hv = (double)__uint_as_float((unsigned)h << 16);
      ((unsigned short *)out0)[i] = h;
      ((unsigned short *)out1)[i] = bf16_rne((float)(x - hv));
    } else {  // BF16X6: hi, mid, lo; each remainder exact in FP64
      const unsigned short h = bf16_rne((float)x);
      const double r1 = x - (double)__uint_as_float((unsigned)h << 16);
      const unsigned short m = bf16_rne((float)r1);
      const double r2 = r1 - (double)__uint_as_float((unsigned)m << 16);
      ((unsigned short *)out0)[i] = h;
      ((unsigned short *)out1)[i] = m;
      ((unsigned short *)out2)[i] = bf16_rne((float)r2);
    }
  }
}

template <class F> void parallel_for(long n, F &&f) {
  unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  std::atomic<long> next(0);
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([&] {
      for (long i; (i = next.fetch_add(1)) < n;) f(i);
    });
  for (auto &x : th) x.join();
}

long round8(long n) { return (n + 7) / 8 * 8; }

}  // namespace

struct DenseHost {
  std::vector<double> T1, T2e, T2m, Wd1, Vd2, S;
};


}  // namespace tce_t

using namespace tce_t;

namespace tce_t {
bool hip_ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  set_error(buf);
  return false;
}

int launch_task_reduce(const long *bounds, int ntasks, int ppi, const double *part, double *out,
                       hipStream_t s) {
  if (ntasks <= 0) return 0;
  hipLaunchKernelGGL(task_reduce, dim3((unsigned)ntasks), dim3(256), 0, s, bounds, ntasks, ppi,
                     part, out);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

unsigned long long *sched_alloc() {
  // slice width 2^shift items per XCD and window (TCE_T_SCHED_SLICE, a power
  // of two from 1 to 1024; default 32)
  static const unsigned long long shift = [] {
    unsigned long long s = 5;
    if (const char *e = getenv("TCE_T_SCHED_SLICE")) {
      const long v = atol(e);
      for (unsigned long long b = 0; b <= 10; ++b)
        if (v == (1L << b)) s = b;
    }
    return s;
  }();
  unsigned long long *p = nullptr;
  if (!hip_ok(hipMalloc(&p, SCHED_ALLOC * sizeof(unsigned long long)), "schedule counters"))
    return nullptr;
  unsigned long long init[SCHED_ALLOC] = {};
  init[9] = shift;
  // synchronous copy from pageable memory: complete on return, before any
  // launch on the streams that use the counters
  if (!hip_ok(hipMemcpy(p, init, sizeof init, hipMemcpyHostToDevice), "schedule counters")) {
    (void)hipFree(p);
    return nullptr;
  }
  return p;
}

unsigned long long *sched_use(unsigned long long *own) {
  static const bool stat = [] {
    const char *e = getenv("TCE_T_SCHED");
    return e && !strcmp(e, "static");
  }();
  return stat ? nullptr : own;
}
}  // namespace tce_t

extern "C" int tce_t_device_memory(long *free_bytes, long *total_bytes) {
  size_t fr = 0, tot = 0;
  HIPCHK(hipMemGetInfo(&fr, &tot), "hipMemGetInfo");
  if (free_bytes) *free_bytes = (long)fr;
  if (total_bytes) *total_bytes = (long)tot;
  return TCE_T_OK;
}

extern "C" int tce_t_init(int device) {
  HIPCHK(hipSetDevice(device), "hipSetDevice");
  return TCE_T_OK;
}

// Contraction ranges of the 18 passes of fused_mixed (TaskDev::x0/x1) for a
// task with tiles (p4b, p5b, p6b, h1b, h2b, h3b): the hull of the orbitals
// whose spin and irrep the pass's blocks allow, in the occupied (d1) or
// virtual (d2) space.  The reference applies the same spin and irrep tests
// to each h7b / p7b tile (ccsd_t_doubles_gpu.F:260-263, 924-927) and skips
// the tiles that fail them; here the forbidden orbitals outside the hull are
// skipped (those inside it hold zeros when classes interleave).
struct ClassRanges {
  // [lo, hi) per (space 0 = occupied / 1 = virtual, spin 0..3, irrep 0..7);
  // hi = 0: no orbital of that class
  long lo[2][4][8], hi[2][4][8];
  explicit ClassRanges(const Tiles &T) {
    for (int sp = 0; sp < 2; ++sp)
      for (int a = 0; a < 4; ++a)
        for (int g = 0; g < 8; ++g) lo[sp][a][g] = 0, hi[sp][a][g] = 0;
    for (long b = 1; b <= T.noab + T.nvab; ++b) {
      const int sp = b > T.noab ? 1 : 0;
      const long spin = T.spn(b), sym = T.sy(b);
      if (spin < 0 || spin > 3 || sym < 0 || sym > 7) continue;
      const long o = T.ofs(b) - (sp ? T.no : 0), e = o + T.rng(b);
      long &l = lo[sp][spin][sym], &h = hi[sp][spin][sym];
      if (h == 0) l = o, h = e;
      else l = std::min(l, o), h = std::max(h, e);
    }
  }
  void set(const Tiles &T, const long tiles[6], TaskDev &d) const {
    for (int p = 0; p < 18; ++p) {
      const PassDef &q = MIXED_PASS[p];
      const long L = tiles[q.L], r = tiles[3 + pair_r(q.h)], s = tiles[3 + pair_s(q.h)];
      const long spin = T.spn(r) + T.spn(s) - T.spn(L), sym = T.sy(L) ^ T.sy(r) ^ T.sy(s);
      const int sp = q.kind == 1 ? 0 : 1;
      const bool ok = spin >= 0 && spin <= 3 && sym >= 0 && sym <= 7;
      d.x0[p] = ok ? (unsigned short)lo[sp][spin][sym] : 0;
      d.x1[p] = ok ? (unsigned short)hi[sp][spin][sym] : 0;
    }
  }
};

// Kernel task table in the order of the reference task loop
// (ccsd_t_gpu.F:87-92): work items of pb-wide particle boxes and 4-wide hole
// boxes.  With sub = true each reference task is replaced by the kernel
// tasks of its box-sized sub-tiles (one item each): sub-tiles qa <= qb <= qc
// of equal reference tiles only, factors by sub-tile equality (1/2, 1/6) as
// ccsd_t_gpu.F:153-167 applies them to tiles, so the reference task's energy
// is the sum of its sub-tasks' and no symmetric duplicate is computed.
// bound[t] = first item of reference task t (bound[ntask] = items).
static long build_tasks(const Tiles &T, int pb, bool sub, std::vector<TaskDev> &tasks,
                        std::vector<long> &bound, std::vector<long> *task_tiles) {
  const long noab = T.noab, nvab = T.nvab;
  long item = 0;
  tasks.clear();
  bound.clear();
  if (task_tiles) task_tiles->clear();
  const ClassRanges cr(T);
  struct Piece {
    int o, r;
  };
  // sub-tiles of tile b with edge e (the last one ragged)
  auto pieces = [&](long b, int e, bool hole) {
    std::vector<Piece> v;
    const int o = (int)(T.ofs(b) - (hole ? 0 : T.no)), r = (int)T.rng(b);
    for (int k = 0; k < r; k += e) v.push_back({o + k, std::min(e, r - k)});
    return v;
  };
  auto sym_factor = [](bool e12, bool e23) {
    return (e12 && e23) ? 1.0 / 6.0 : ((e12 || e23) ? 0.5 : 1.0);
  };
  // Sub-tile order within a reference task: particle boxes outermost and
  // the hole boxes fastest, so that the items an XCD runs together (32
  // consecutive ones) share the particle-pair rows of V2 and T2, the larger
  // operand.  Measured at no=40/nv=320 (tiles of 20/40: 5 boxes per tile in
  // either space) against particle c fastest, one box: L2 misses 20-35%
  // lower, split-BF16 5-6% faster, FP32/FP64 0.2-0.4%.
  // TCE_T_ITEM_ORDER=particles restores c fastest.
  const char *ord = getenv("TCE_T_ITEM_ORDER");
  const bool holes_inner = !(ord && !strcmp(ord, "particles"));
  for (long p4b = noab + 1; p4b <= noab + nvab; ++p4b)
    for (long p5b = p4b; p5b <= noab + nvab; ++p5b)
      for (long p6b = p5b; p6b <= noab + nvab; ++p6b)
        for (long h1b = 1; h1b <= noab; ++h1b)
          for (long h2b = h1b; h2b <= noab; ++h2b)
            for (long h3b = h2b; h3b <= noab; ++h3b) {
              if (!task_allowed(T, p4b, p5b, p6b, h1b, h2b, h3b)) continue;
              bound.push_back(item);
              const long tiles6[6] = {p4b, p5b, p6b, h1b, h2b, h3b};
              TaskDev ranges;
              cr.set(T, tiles6, ranges);
              if (task_tiles) {
                long tt[6] = {p4b, p5b, p6b, h1b, h2b, h3b};
                task_tiles->insert(task_tiles->end(), tt, tt + 6);
              }
              if (!sub) {
                TaskDev d;
                d.op4 = (int)(T.ofs(p4b) - T.no);
                d.op5 = (int)(T.ofs(p5b) - T.no);
                d.op6 = (int)(T.ofs(p6b) - T.no);
                d.oh1 = (int)T.ofs(h1b);
                d.oh2 = (int)T.ofs(h2b);
                d.oh3 = (int)T.ofs(h3b);
                d.r4 = (int)T.rng(p4b);
                d.r5 = (int)T.rng(p5b);
                d.r6 = (int)T.rng(p6b);
                d.r1 = (int)T.rng(h1b);
                d.r2 = (int)T.rng(h2b);
                d.r3 = (int)T.rng(h3b);
                d.nb4 = (d.r4 + pb - 1) / pb;
                d.nb5 = (d.r5 + pb - 1) / pb;
                d.nb6 = (d.r6 + pb - 1) / pb;
                d.nb1 = (d.r1 + 3) / 4;
                d.nb2 = (d.r2 + 3) / 4;
                d.nb3 = (d.r3 + 3) / 4;
                d.item0 = item;
                d.factor = task_factor(T.restricted, p4b, p5b, p6b, h1b, h2b, h3b);
                std::copy(ranges.x0, ranges.x0 + 18, d.x0);
                std::copy(ranges.x1, ranges.x1 + 18, d.x1);
                item += (long)d.nb4 * d.nb5 * d.nb6 * d.nb1 * d.nb2 * d.nb3;
                tasks.push_back(d);
                continue;
              }
              const auto A = pieces(p4b, pb, false), B = pieces(p5b, pb, false),
                         C = pieces(p6b, pb, false), I = pieces(h1b, 4, true),
                         J = pieces(h2b, 4, true), K = pieces(h3b, 4, true);
              const bool pab = p4b == p5b, pbc = p5b == p6b, hij = h1b == h2b, hjk = h2b == h3b;
              const double rf = T.restricted ? 2.0 : 1.0;
              auto emit = [&](size_t a, size_t b, size_t c, size_t i, size_t j, size_t k) {
                TaskDev d;
                d.op4 = A[a].o, d.op5 = B[b].o, d.op6 = C[c].o;
                d.oh1 = I[i].o, d.oh2 = J[j].o, d.oh3 = K[k].o;
                d.r4 = A[a].r, d.r5 = B[b].r, d.r6 = C[c].r;
                d.r1 = I[i].r, d.r2 = J[j].r, d.r3 = K[k].r;
                d.nb4 = d.nb5 = d.nb6 = d.nb1 = d.nb2 = d.nb3 = 1;
                d.item0 = item++;
                d.factor = rf * sym_factor(hij && i == j, hjk && j == k) *
                           sym_factor(pab && a == b, pbc && b == c);
                std::copy(ranges.x0, ranges.x0 + 18, d.x0);
                std::copy(ranges.x1, ranges.x1 + 18, d.x1);
                tasks.push_back(d);
              };
              if (holes_inner) {
                // particles outermost, hole k fastest
                for (size_t a = 0; a < A.size(); ++a)
                  for (size_t b = pab ? a : 0; b < B.size(); ++b)
                    for (size_t c = pbc ? b : 0; c < C.size(); ++c)
                      for (size_t i = 0; i < I.size(); ++i)
                        for (size_t j = hij ? i : 0; j < J.size(); ++j)
                          for (size_t k = hjk ? j : 0; k < K.size(); ++k) emit(a, b, c, i, j, k);
              } else {
                // holes outermost, particle c fastest
                for (size_t i = 0; i < I.size(); ++i)
                  for (size_t j = hij ? i : 0; j < J.size(); ++j)
                    for (size_t k = hjk ? j : 0; k < K.size(); ++k)
                      for (size_t a = 0; a < A.size(); ++a)
                        for (size_t b = pab ? a : 0; b < B.size(); ++b)
                          for (size_t c = pbc ? b : 0; c < C.size(); ++c) emit(a, b, c, i, j, k);
              }
            }
  bound.push_back(item);
  return item;
}

// Orbital tiling: each run of consecutive tiles with one spin and irrep
// (a contiguous orbital range) cut into pieces of e4 occupied or ep virtual
// orbitals, the last one ragged; spin, irrep, orbital energies and the
// restricted flag carry over, so the task filters and factors of the task
// loop apply to the pieces unchanged.
static Tiles orbital_tiles(const Tiles &T, int ep) {
  Tiles R;
  R.restricted = T.restricted;
  R.no = T.no;
  R.nv = T.nv;
  R.norb = T.norb;
  R.evl = T.evl;
  auto space = [&](long first, long last, int e) {
    for (long b = first; b <= last;) {
      long c = b;
      while (c + 1 <= last && T.spn(c + 1) == T.spn(b) && T.sy(c + 1) == T.sy(b)) ++c;
      const long o = T.ofs(b), end = T.ofs(c) + T.rng(c);
      for (long k = o; k < end; k += e) {
        R.range.push_back(std::min<long>(e, end - k));
        R.offset.push_back(k);
        R.spin.push_back(T.spn(b));
        R.sym.push_back(T.sy(b));
      }
      b = c + 1;
    }
  };
  space(1, T.noab, 4);
  R.noab = (long)R.range.size();
  space(T.noab + 1, T.noab + T.nvab, ep);
  R.nvab = (long)R.range.size() - R.noab;
  for (long i = 1; i <= R.noab + R.nvab; ++i) R.alpha.push_back(i);
  return R;
}

// Kernel task table of a job in its tiling mode (pb: particle box).
static long job_tasks(const Job &J, int pb, std::vector<TaskDev> &tasks, std::vector<long> &bound) {
  if (J.tiling == TCE_T_TILING_ORBITAL) {
    const long n = build_tasks(orbital_tiles(J.T, pb), pb, false, tasks, bound, nullptr);
    bound.assign({0L, n});
    return n;
  }
  return build_tasks(J.T, pb, J.tiling == TCE_T_TILING_SUBTILE, tasks, bound, nullptr);
}

// Work items of a job's kernel tasks with pb-wide particle boxes (counted,
// not stored).
static long count_items(const Job &J, int pb) {
  const Tiles R = J.tiling == TCE_T_TILING_ORBITAL ? orbital_tiles(J.T, pb) : J.T;
  const bool sub = J.tiling == TCE_T_TILING_SUBTILE;
  const long noab = R.noab, nvab = R.nvab;
  auto nb = [](long r, long e) { return (r + e - 1) / e; };
  // pieces x (ways to choose ordered sub-tiles) per reference task: for
  // sub-tiles within equal tiles only the canonical (non-decreasing) choices
  auto choose = [](long n, int k) {  // multisets of size k from n pieces
    return k == 1 ? n : (k == 2 ? n * (n + 1) / 2 : n * (n + 1) * (n + 2) / 6);
  };
  long items = 0;
  for (long p4b = noab + 1; p4b <= noab + nvab; ++p4b)
    for (long p5b = p4b; p5b <= noab + nvab; ++p5b)
      for (long p6b = p5b; p6b <= noab + nvab; ++p6b)
        for (long h1b = 1; h1b <= noab; ++h1b)
          for (long h2b = h1b; h2b <= noab; ++h2b)
            for (long h3b = h2b; h3b <= noab; ++h3b) {
              if (!task_allowed(R, p4b, p5b, p6b, h1b, h2b, h3b)) continue;
              const long a = nb(R.rng(p4b), pb), b = nb(R.rng(p5b), pb), c = nb(R.rng(p6b), pb);
              const long i = nb(R.rng(h1b), 4), j = nb(R.rng(h2b), 4), k = nb(R.rng(h3b), 4);
              if (!sub) {
                items += a * b * c * i * j * k;
                continue;
              }
              long pp, hh;
              if (p4b == p5b && p5b == p6b) pp = choose(a, 3);
              else if (p4b == p5b) pp = choose(a, 2) * c;
              else if (p5b == p6b) pp = a * choose(b, 2);
              else pp = a * b * c;
              if (h1b == h2b && h2b == h3b) hh = choose(i, 3);
              else if (h1b == h2b) hh = choose(i, 2) * k;
              else if (h2b == h3b) hh = i * choose(j, 2);
              else hh = i * j * k;
              items += pp * hh;
            }
  return items;
}

extern "C" int tce_t_job_create(const tce_t_tiles *tiles, tce_t_job **out) {
  int tiling = TCE_T_TILING_SUBTILE;
  if (const char *e = getenv("TCE_T_SUBTILE")) tiling = atoi(e);
  return tce_t_job_create_tiling(tiles, tiling, out);
}

extern "C" int tce_t_job_tiling(const tce_t_job *job) {
  return job ? reinterpret_cast<const Job *>(job)->tiling : TCE_T_ERR_ARG;
}

extern "C" int tce_t_job_create_tiling(const tce_t_tiles *tiles, int tiling, tce_t_job **out) {
  if (!out) return TCE_T_ERR_ARG;
  if (tiling < TCE_T_TILING_REFERENCE || tiling > TCE_T_TILING_ORBITAL) {
    set_error("unknown tiling mode");
    return TCE_T_ERR_ARG;
  }
  auto J = std::make_unique<Job>();
  J->tiling = tiling;
  if (!J->T.load(tiles)) {
    set_error("invalid tile description");
    return TCE_T_ERR_ARG;
  }
  const Tiles &T = J->T;
  // Kernel choice: 8-wide particle boxes (cooperative workgroups, operands
  // shared through LDS, the 16x16x4 FP64 kernel) unless their padded work
  // exceeds that of 4-wide boxes (one wave per box, fused_triples) by more
  // than the measured rate ratio of the two kernels (63.4 vs 27.9 TF/s at
  // no=40/nv=320: 2.2x), both counted in the job's own kernel tasks.
  // TCE_T_KERNEL = wave | wg | wg16 forces fused_triples, fused_triples_wg
  // or the FP64 policy of fused_triples_mixed.
  {
    const double cells4 = 64.0 * count_items(*J, 4), cells8 = 512.0 * count_items(*J, 8);
    J->pbox = (cells8 <= 2.2 * cells4) ? 8 : 4;
    // 8-wide boxes run on the 16x16x4 FP64 kernel (measured 3% faster than
    // the 4x4x4 one at no=40/nv=320)
    J->k16 = J->pbox == 8;
    if (const char *k = getenv("TCE_T_KERNEL")) {
      if (!strcmp(k, "wave")) J->pbox = 4, J->k16 = false;
      if (!strcmp(k, "wg")) J->pbox = 8, J->k16 = false;
      if (!strcmp(k, "wg16")) J->pbox = 8, J->k16 = true;
    }
    J->ppi = J->pbox == 8 ? 8 : 1;
  }
  {  // the reference task list (per-task energies, restart blocks)
    std::vector<TaskDev> ref;
    std::vector<long> rb;
    build_tasks(T, J->pbox, false, ref, rb, &J->task_tiles);
  }
  J->nitems = job_tasks(*J, J->pbox, J->tasks, J->bound);
  if (J->pbox == 8)
    J->tasks8 = J->tasks, J->bound8 = J->bound, J->nitems8 = J->nitems;
  else
    J->nitems8 = job_tasks(*J, 8, J->tasks8, J->bound8);
  J->Ep = round8(T.nv);
  J->Mp = round8(T.no);
  for (const TaskDev &d : J->tasks8)
    for (int p = 0; p < 18 && !J->sparse; ++p) {
      const long full = MIXED_PASS[p].kind ? T.no : T.nv;
      J->sparse = d.x0[p] > 0 || d.x1[p] < full;
    }
  if (const char *e = getenv("TCE_T_SPARSE"))  // 0: always the whole spaces
    if (!strcmp(e, "0")) J->sparse = false;
  *out = reinterpret_cast<tce_t_job *>(J.release());
  return TCE_T_OK;
}

extern "C" long tce_t_job_num_tasks(const tce_t_job *job) {
  return job ? reinterpret_cast<const Job *>(job)->ntask() : -1;
}

extern "C" int tce_t_job_work(const tce_t_job *job, int precision, double *flops, long *items) {
  if (!job || !valid_precision(precision)) return TCE_T_ERR_ARG;
  const Job &J = *reinterpret_cast<const Job *>(job);
  const bool wide = precision != TCE_T_FP64 || J.pbox == 8;
  const std::vector<TaskDev> &tab = wide ? J.tasks8 : J.tasks;
  double f = 0.0;
  for (const TaskDev &d : tab)
    f += (double)d.r4 * d.r5 * d.r6 * d.r1 * d.r2 * d.r3;
  if (flops) *flops = f * 18.0 * (double)(J.T.no + J.T.nv + 1);
  if (items) *items = wide ? J.nitems8 : J.nitems;
  return TCE_T_OK;
}

static long dense_bytes(const Job &J) {
  long no = J.T.no, nv = J.T.nv;
  long n = nv * no                       // T1
           + no * no * nv * J.Ep         // T2e
           + no * nv * nv * J.Mp         // T2m
           + no * nv * nv * J.Ep         // Vd2
           + no * no * nv * J.Mp         // Wd1
           + no * no * nv * nv           // S
           + J.T.norb;                   // evl
  return n * (long)sizeof(double);
}

extern "C" long tce_t_job_device_bytes(const tce_t_job *job) {
  if (!job) return -1;
  const Job &J = *reinterpret_cast<const Job *>(job);
  return dense_bytes(J) + J.nitems * J.ppi * 16 + (long)J.tasks.size() * sizeof(TaskDev) +
         (J.ngroups() + 1) * 8 + J.ngroups() * 16 + 64 + (long)J.dtile_bytes;
}

namespace {

struct Packer {
  const Tiles &T;
  int err = 0;
  // Fetch the canonical block with tiles (g3,g4,g1,g2) (already ordered),
  // returns pointer or nullptr (forbidden block -> zeros).  Sets err when a
  // needed key is absent.
  const double *fetch(const double *file, const long *hash, long key, bool needed) {
    long off = hash_lookup(hash, key);
    if (off < 0) {
      if (needed) err = TCE_T_ERR_KEY;
      return nullptr;
    }
    return file + off;
  }
};

}  // namespace

// Host side of the upload: every block the path reads is fetched through the
// hash tables and scattered into the dense device layouts (see the header of
// this file).  Blocks forbidden by spin/symmetry stay zero; a permitted block
// missing from a table is an error (the reference would errquit in tce_hash).
static int pack_dense(const Tiles &T, long Ep, long Mp, const double *t1, const long *t1h,
                      const double *t2, const long *t2h, const double *v2, const long *v2h,
                      DenseHost &D) {
  const long no = T.no, nv = T.nv, noab = T.noab, nvab = T.nvab, N = noab + nvab;
  Packer pk{T};
  std::atomic<int> err(0);

  // ---- T1[p][h]: t(p h) blocks, key h1b-1 + noab*(p4b-noab-1) (ccsd_t_singles_gpu.F:233)
  std::vector<double> &hT1 = D.T1;
  hT1.assign(nv * no, 0.0);
  for (long pb = noab + 1; pb <= noab + nvab; ++pb)
    for (long hb = 1; hb <= noab; ++hb) {
      long p1 = pb, h1 = hb;
      if (T.restricted && T.spn(pb) + T.spn(hb) == 4) {
        p1 = T.alp(pb);
        h1 = T.alp(hb);
      }
      bool need = T.spn(pb) == T.spn(hb) && (T.sy(pb) ^ T.sy(hb)) == 0;
      const double *blk = pk.fetch(t1, t1h, h1 - 1 + noab * (p1 - noab - 1), need);
      if (!blk) continue;
      long rp = T.rng(pb), rh = T.rng(hb), op = T.ofs(pb) - no, oh = T.ofs(hb);
      for (long p = 0; p < rp; ++p)
        for (long h = 0; h < rh; ++h) hT1[(op + p) * no + oh + h] = blk[p * rh + h];
    }

  // ---- T2 in both layouts.  t(p p' h h') for any tile order: canonical block
  // of the sorted tiles, sign -1 per swapped pair (tce_hashnsort.F:47-64,
  // 129-144 and the sorts of ccsd_t_doubles_gpu.F:282-304, 942-976).
  std::vector<double> &hT2e = D.T2e, &hT2m = D.T2m;
  hT2e.assign((size_t)no * no * nv * Ep, 0.0);
  hT2m.assign((size_t)no * nv * nv * Mp, 0.0);
  {
    std::vector<std::pair<long, long>> ppairs;
    for (long P1 = noab + 1; P1 <= N; ++P1)
      for (long P2 = noab + 1; P2 <= N; ++P2) ppairs.push_back({P1, P2});
    parallel_for((long)ppairs.size(), [&](long q) {
      long P1 = ppairs[q].first, P2 = ppairs[q].second;
      for (long H1 = 1; H1 <= noab; ++H1)
        for (long H2 = 1; H2 <= noab; ++H2) {
          bool sp = P1 > P2, sh = H1 > H2;
          long a = sp ? P2 : P1, b = sp ? P1 : P2, c = sh ? H2 : H1, d = sh ? H1 : H2;
          bool need = allowed4(T, a, b, c, d);
          restrict4(T, a, b, c, d);
          long key = d - 1 + noab * (c - 1 + noab * (b - noab - 1 + nvab * (a - noab - 1)));
          const double *blk = pk.fetch(t2, t2h, key, need);
          if (!blk) {
            if (need) err = TCE_T_ERR_KEY;
            continue;
          }
          long r1 = T.rng(P1), r2 = T.rng(P2), r3 = T.rng(H1), r4 = T.rng(H2);
          long o1 = T.ofs(P1) - no, o2 = T.ofs(P2) - no, o3 = T.ofs(H1), o4 = T.ofs(H2);
          // canonical block extents
          long ca = sp ? r2 : r1, cb = sp ? r1 : r2, cc = sh ? r4 : r3, cd = sh ? r3 : r4;
          (void)ca;
          double sgn = (sp ? -1.0 : 1.0) * (sh ? -1.0 : 1.0);
          for (long i = 0; i < r1; ++i)
            for (long j = 0; j < r2; ++j)
              for (long k = 0; k < r3; ++k)
                for (long l = 0; l < r4; ++l) {
                  long ia = sp ? j : i, ib = sp ? i : j, ic = sh ? l : k, id = sh ? k : l;
                  double val = sgn * blk[((ia * cb + ib) * cc + ic) * cd + id];
                  long p = o1 + i, e = o2 + j, h = o3 + k, g = o4 + l;
                  hT2e[((h * no + g) * nv + p) * Ep + e] = val;  // t(p e h g)
                  hT2m[((h * nv + p) * nv + e) * Mp + g] = val;  // t(p e h m=g)
                }
        }
    });
  }

  // ---- <h p||h h> (d1): v(h7 p6 h2 h3), key h3-1+N*(h2-1+N*(p6-1+N*(h7-1)))
  std::vector<double> &hW = D.Wd1;
  hW.assign((size_t)no * no * nv * Mp, 0.0);
  for (long H7 = 1; H7 <= noab; ++H7)
    for (long P6 = noab + 1; P6 <= N; ++P6)
      for (long H2 = 1; H2 <= noab; ++H2)
        for (long H3 = 1; H3 <= noab; ++H3) {
          bool sw = H2 > H3;
          long g3 = H7, g4 = P6, g1 = sw ? H3 : H2, g2 = sw ? H2 : H3;
          bool need = allowed4(T, g3, g4, g1, g2);
          restrict4(T, g3, g4, g1, g2);
          const double *blk = pk.fetch(v2, v2h, g2 - 1 + N * (g1 - 1 + N * (g4 - 1 + N * (g3 - 1))), need);
          if (!blk) {
            if (need) err = TCE_T_ERR_KEY;
            continue;
          }
          long r7 = T.rng(H7), r6 = T.rng(P6), r2 = T.rng(H2), r3 = T.rng(H3);
          long c1 = sw ? r3 : r2, c2 = sw ? r2 : r3;
          (void)c1;
          double sgn = sw ? -1.0 : 1.0;
          for (long m = 0; m < r7; ++m)
            for (long p = 0; p < r6; ++p)
              for (long h = 0; h < r2; ++h)
                for (long g = 0; g < r3; ++g) {
                  long ih = sw ? g : h, ig = sw ? h : g;
                  double val = sgn * blk[((m * r6 + p) * (sw ? r3 : r2) + ih) * c2 + ig];
                  long M = T.ofs(H7) + m, Pp = T.ofs(P6) - no + p, Hh = T.ofs(H2) + h, G = T.ofs(H3) + g;
                  hW[((Hh * no + G) * nv + Pp) * Mp + M] = val;
                }
        }

  // ---- <p p||h p> (d2): v(p5 p6 h3 p7), key p7-1+N*(h3-1+N*(p6-1+N*(p5-1)))
  std::vector<double> &hV = D.Vd2;
  hV.assign((size_t)no * nv * nv * Ep, 0.0);
  {
    std::vector<std::pair<long, long>> ppairs;
    for (long P5 = noab + 1; P5 <= N; ++P5)
      for (long P6 = noab + 1; P6 <= N; ++P6) ppairs.push_back({P5, P6});
    parallel_for((long)ppairs.size(), [&](long q) {
      long P5 = ppairs[q].first, P6 = ppairs[q].second;
      bool sw = P5 > P6;
      for (long H3 = 1; H3 <= noab; ++H3)
        for (long P7 = noab + 1; P7 <= N; ++P7) {
          long g3 = sw ? P6 : P5, g4 = sw ? P5 : P6, g1 = H3, g2 = P7;
          bool need = allowed4(T, g3, g4, g1, g2);
          restrict4(T, g3, g4, g1, g2);
          const double *blk = pk.fetch(v2, v2h, g2 - 1 + N * (g1 - 1 + N * (g4 - 1 + N * (g3 - 1))), need);
          if (!blk) {
            if (need) err = TCE_T_ERR_KEY;
            continue;
          }
          long r5 = T.rng(P5), r6 = T.rng(P6), r3 = T.rng(H3), r7 = T.rng(P7);
          long cb = sw ? r5 : r6;
          double sgn = sw ? -1.0 : 1.0;
          for (long i = 0; i < r5; ++i)
            for (long j = 0; j < r6; ++j)
              for (long h = 0; h < r3; ++h)
                for (long e = 0; e < r7; ++e) {
                  long ia = sw ? j : i, ib = sw ? i : j;
                  double val = sgn * blk[((ia * cb + ib) * r3 + h) * r7 + e];
                  long p = T.ofs(P5) - no + i, pp = T.ofs(P6) - no + j, w = T.ofs(H3) + h,
                       E = T.ofs(P7) - no + e;
                  hV[((w * nv + p) * nv + pp) * Ep + E] = val;
                }
        }
    });
  }

  // ---- <p p||h h> (s1): v(p5 p6 h2 h3), key h3-1+N*(h2-1+N*(p6-1+N*(p5-1)))
  std::vector<double> &hS = D.S;
  hS.assign((size_t)no * no * nv * nv, 0.0);
  for (long P5 = noab + 1; P5 <= N; ++P5)
    for (long P6 = noab + 1; P6 <= N; ++P6)
      for (long H2 = 1; H2 <= noab; ++H2)
        for (long H3 = 1; H3 <= noab; ++H3) {
          bool sp = P5 > P6, sh = H2 > H3;
          long g3 = sp ? P6 : P5, g4 = sp ? P5 : P6, g1 = sh ? H3 : H2, g2 = sh ? H2 : H3;
          bool need = allowed4(T, g3, g4, g1, g2);
          restrict4(T, g3, g4, g1, g2);
          const double *blk = pk.fetch(v2, v2h, g2 - 1 + N * (g1 - 1 + N * (g4 - 1 + N * (g3 - 1))), need);
          if (!blk) {
            if (need) err = TCE_T_ERR_KEY;
            continue;
          }
          long r5 = T.rng(P5), r6 = T.rng(P6), r2 = T.rng(H2), r3 = T.rng(H3);
          long cb = sp ? r5 : r6, cd = sh ? r2 : r3;
          double sgn = (sp ? -1.0 : 1.0) * (sh ? -1.0 : 1.0);
          for (long i = 0; i < r5; ++i)
            for (long j = 0; j < r6; ++j)
              for (long h = 0; h < r2; ++h)
                for (long g = 0; g < r3; ++g) {
                  long ia = sp ? j : i, ib = sp ? i : j, ic = sh ? g : h, id = sh ? h : g;
                  double val = sgn * blk[((ia * cb + ib) * (sh ? r3 : r2) + ic) * cd + id];
                  long p = T.ofs(P5) - no + i, pp = T.ofs(P6) - no + j, H = T.ofs(H2) + h,
                       G = T.ofs(H3) + g;
                  hS[((H * no + G) * nv + p) * nv + pp] = val;
                }
        }
  if (pk.err || err) {
    set_error("a block needed by the (T) path is missing from a hash table");
    return TCE_T_ERR_KEY;
  }
  return TCE_T_OK;
}

// The upload of a (T) job (kind PACK_TRIPLES) or of a CR-CCSD(T) job
// (PACK_CR: v2 and x2 hold the two moment intermediates).
static int job_upload(tce_t_job *job, const double *t1, const long *t1h, const double *t2,
                      const long *t2h, const double *v2, const long *v2h, int kind,
                      const double *x2, const long *x2h) {
  Job &J = *reinterpret_cast<Job *>(job);
  const Tiles &T = J.T;
  // a new upload replaces the tensors and everything derived from them
  // (reduced-precision copies, 8-wide tables, tile buffer)
  if (J.uploaded) job_release(J);
  auto up = [&](double *&d, const std::vector<double> &h, const char *what) -> bool {
    if (d) (void)hipFree(d);
    d = nullptr;
    if (!hip_ok(hipMalloc(&d, h.size() * sizeof(double)), what)) return false;
    return hip_ok(hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice), what);
  };
  if (!J.stream) HIPCHK(hipStreamCreateWithFlags(&J.stream, hipStreamNonBlocking), "stream");
  if (!J.dsched && !(J.dsched = sched_alloc())) return TCE_T_ERR_HIP;
  if (!J.ev0) HIPCHK(hipEventCreate(&J.ev0), "event");
  if (!J.ev1) HIPCHK(hipEventCreate(&J.ev1), "event");
  // Dense tensors: fetched and sorted on the device (pack.hip) from the
  // referenced ranges of the three files; TCE_T_PACK=host builds them with
  // the host packer instead (same values bit for bit) and copies them over.
  const char *pk = getenv("TCE_T_PACK");
  J.cr = kind == PACK_CR;
  if (pk && !strcmp(pk, "host") && kind == PACK_TRIPLES) {
    DenseHost D;
    const int rc = pack_dense(T, J.Ep, J.Mp, t1, t1h, t2, t2h, v2, v2h, D);
    if (rc != TCE_T_OK) return rc;
    if (!up(J.dT1, D.T1, "T1") || !up(J.dT2e, D.T2e, "T2e") || !up(J.dT2m, D.T2m, "T2m") ||
        !up(J.dVd2, D.Vd2, "Vd2") || !up(J.dWd1, D.Wd1, "Wd1") || !up(J.dS, D.S, "S"))
      return TCE_T_ERR_HIP;
  } else {
    const long no = T.no, nv = T.nv;
    auto alloc = [&](double *&d, long n, const char *what) -> bool {
      if (d) (void)hipFree(d);
      d = nullptr;
      return hip_ok(hipMalloc(&d, std::max(1L, n) * sizeof(double)), what);
    };
    if (!alloc(J.dT1, nv * no, "T1") || !alloc(J.dT2e, no * no * nv * J.Ep, "T2e") ||
        !alloc(J.dT2m, no * nv * nv * J.Mp, "T2m") || !alloc(J.dVd2, no * nv * nv * J.Ep, "Vd2") ||
        !alloc(J.dWd1, no * no * nv * J.Mp, "Wd1") || !alloc(J.dS, no * no * nv * nv, "S"))
      return TCE_T_ERR_HIP;
    const int rc = pack_on_device(J, t1, t1h, t2, t2h, v2, v2h, kind, x2, x2h);
    if (rc != TCE_T_OK) return rc;
  }
  if (!up(J.devl, T.evl, "evl")) return TCE_T_ERR_HIP;
  if (J.dtasks) (void)hipFree(J.dtasks);
  HIPCHK(hipMalloc(&J.dtasks, std::max<size_t>(1, J.tasks.size()) * sizeof(TaskDev)), "tasks");
  HIPCHK(hipMemcpy(J.dtasks, J.tasks.data(), J.tasks.size() * sizeof(TaskDev), hipMemcpyHostToDevice),
         "tasks");
  if (J.dbound) (void)hipFree(J.dbound);
  HIPCHK(hipMalloc(&J.dbound, J.bound.size() * sizeof(long)), "bounds");
  HIPCHK(hipMemcpy(J.dbound, J.bound.data(), J.bound.size() * sizeof(long), hipMemcpyHostToDevice),
         "bounds");
  if (J.dpart) (void)hipFree(J.dpart);
  HIPCHK(hipMalloc(&J.dpart, std::max(1L, J.nitems) * J.ppi * 2 * sizeof(double)), "partials");
  if (J.dtaskE) (void)hipFree(J.dtaskE);
  HIPCHK(hipMalloc(&J.dtaskE, std::max(1L, J.ngroups()) * 2 * sizeof(double)), "taskE");
  int cus = 0;
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, J.device), "attr");
  // persistent grid: 8 waves per CU (2 per SIMD)
  J.grid = std::max(1, cus) * (J.pbox == 8 ? 1 : 8);
  J.uploaded = true;
  ++J.gen;
  if (J.pbox == 8) {
    J.dtasks8 = J.dtasks;
    J.dpart8 = J.dpart;
    J.dbound8 = J.dbound;
  }
  return TCE_T_OK;
}

extern "C" int tce_t_job_upload(tce_t_job *job, const double *t1, const long *t1h, const double *t2,
                                const long *t2h, const double *v2, const long *v2h) {
  if (!job || !t1 || !t1h || !t2 || !t2h || !v2 || !v2h) return TCE_T_ERR_ARG;
  return job_upload(job, t1, t1h, t2, t2h, v2, v2h, PACK_TRIPLES, nullptr, nullptr);
}

// A CR-CCSD(T) job (cr_ccsd_t.F): t1 and t2 as a (T) job's, the doubles
// operands from the two intermediates that cr_ccsd_t_N builds before its
// task loop (toggle 1: d_i1_1 with k_i1_offset_1, i1(h11 p4 h1 h2); d_i1_2
// with k_i1_offset_2, i1(p4 p5 h1 p12)).  Its t3 doubles are the moment
// tile of cr_ccsd_t_N (toggle 2), its singles the denominator tile of
// cr_ccsd_t_E; tce_t_cr_ccsd_t pairs them with a (T) job's.
extern "C" int tce_t_job_upload_cr(tce_t_job *job, const double *t1, const long *k_t1_offset,
                                   const double *t2, const long *k_t2_offset, const double *i1_1,
                                   const long *k_i1_offset_1, const double *i1_2,
                                   const long *k_i1_offset_2) {
  if (!job || !t1 || !k_t1_offset || !t2 || !k_t2_offset || !i1_1 || !k_i1_offset_1 || !i1_2 ||
      !k_i1_offset_2) {
    set_error("tce_t_job_upload_cr: null argument");
    return TCE_T_ERR_ARG;
  }
  return job_upload(job, t1, k_t1_offset, t2, k_t2_offset, i1_1, k_i1_offset_1, PACK_CR, i1_2,
                    k_i1_offset_2);
}

// Room for one job's launch table: arguments and one segment, FP64 or
// reduced precision (the larger of the two layouts).
constexpr size_t TAB_BYTES = 1024;
static_assert(sizeof(MixedArgs) + sizeof(MixedSegment) <= TAB_BYTES, "launch table");
static_assert(sizeof(FusedArgs) + sizeof(Segment) <= TAB_BYTES, "launch table");

static int ensure_table(Job &J) {
  if (!J.tab_host) HIPCHK(hipHostMalloc(&J.tab_host, TAB_BYTES), "launch table");
  if (!J.tab_dev) HIPCHK(hipMalloc(&J.tab_dev, TAB_BYTES), "launch table");
  return TCE_T_OK;
}

// Copy `args` and one segment covering items [item0, item1) into the job's
// launch table (stream-ordered); returns the device segment pointer.
template <class Args>
static const SegmentT<Args> *stage_table(Job &J, const Args &args, long item1, long item0 = 0) {
  char *h = (char *)J.tab_host;
  char *d = (char *)J.tab_dev;
  std::memcpy(h, &args, sizeof(Args));
  SegmentT<Args> sg;
  sg.args = (const Args *)d;
  sg.item0 = item0;
  sg.item1 = item1;
  sg.g0 = 0;
  const size_t off = (sizeof(Args) + 15) / 16 * 16;
  std::memcpy(h + off, &sg, sizeof sg);
  if (hipMemcpyAsync(d, h, off + sizeof sg, hipMemcpyHostToDevice, J.stream) != hipSuccess)
    return nullptr;
  return (const SegmentT<Args> *)(d + off);
}

const MixedSegment *tce_t::stage_mixed(Job &J, const MixedArgs &M, long item1, long item0) {
  if (ensure_table(J) != TCE_T_OK) return nullptr;
  return stage_table(J, M, item1, item0);
}

// Contraction extent of a reduced-precision mode: K padded to the kernel's
// chunk (16 FP32 or 32 bf16 values).
static int mixed_extent(long K, int precision) {
  const int kc = precision == TCE_T_FP32 ? 16 : 32;
  return (int)((K + kc - 1) / kc * kc);
}

static long mixed_bytes(const Job &J, int precision) {
  const long no = J.T.no, nv = J.T.nv;
  const long Ek = mixed_extent(nv, precision), Mk = mixed_extent(no, precision);
  const long elems = no * no * nv * Ek + no * nv * nv * Ek + no * no * nv * Mk + no * nv * nv * Mk;
  return elems * (precision == TCE_T_BF16X6 ? 6 : 4);  // FP32, bf16 hi + lo, or hi + mid + lo
}

// Convert the four contraction tensors once per precision (device to
// device, on the job's stream) and set up the pbox-8 task table.
int tce_t::ensure_tasks8(Job &J) {
  if (J.dtasks8) return TCE_T_OK;
  if (J.pbox == 8) {
    J.dtasks8 = J.dtasks;
    J.dpart8 = J.dpart;
    J.dbound8 = J.dbound;
    return TCE_T_OK;
  }
  HIPCHK(hipMalloc(&J.dtasks8, std::max<size_t>(1, J.tasks8.size()) * sizeof(TaskDev)), "tasks8");
  HIPCHK(hipMemcpy(J.dtasks8, J.tasks8.data(), J.tasks8.size() * sizeof(TaskDev),
                   hipMemcpyHostToDevice),
         "tasks8");
  HIPCHK(hipMalloc(&J.dbound8, J.bound8.size() * sizeof(long)), "bounds8");
  HIPCHK(hipMemcpy(J.dbound8, J.bound8.data(), J.bound8.size() * sizeof(long),
                   hipMemcpyHostToDevice),
         "bounds8");
  HIPCHK(hipMalloc(&J.dpart8, std::max(1L, J.nitems8) * 8 * 2 * sizeof(double)), "partials8");
  return TCE_T_OK;
}

// Convert the four contraction tensors once per precision (device to
// device, on the job's stream) and set up the 8-wide task table.
int tce_t::prepare_mixed(Job &J, int precision) {
  Job::Mixed &m = J.mx[precision];
  if (m.ready) return TCE_T_OK;
  const long no = J.T.no, nv = J.T.nv;
  m.Ek = mixed_extent(nv, precision);
  m.Mk = mixed_extent(no, precision);
  const int nparts = mixed_parts(precision);
  const size_t es = precision == TCE_T_FP32 ? 4 : 2;
  struct Conv {
    const double *src;
    long rows;
    int K, kin, kout;
    void **dst;
  } ts[4] = {{J.dT2e, no * no * nv, (int)nv, (int)J.Ep, m.Ek, m.x2},
             {J.dVd2, no * nv * nv, (int)nv, (int)J.Ep, m.Ek, m.y2},
             {J.dWd1, no * no * nv, (int)no, (int)J.Mp, m.Mk, m.x1},
             {J.dT2m, no * nv * nv, (int)no, (int)J.Mp, m.Mk, m.y1}};
  for (auto &t : ts) {
    for (int p = 0; p < nparts; ++p)
      HIPCHK(hipMalloc(&t.dst[p], std::max<size_t>(16, (size_t)t.rows * t.kout * es)), "mixed tensors");
    hipLaunchKernelGGL(convert_rows, dim3(4096), dim3(256), 0, J.stream, t.src, t.rows, t.K, t.kin,
                       t.kout, precision, t.dst[0], nparts > 1 ? t.dst[1] : nullptr,
                       nparts > 2 ? t.dst[2] : nullptr);
    HIPCHK(hipGetLastError(), "convert launch");
  }
  {
    const int rc = ensure_tasks8(J);
    if (rc != TCE_T_OK) return rc;
  }
  HIPCHK(hipStreamSynchronize(J.stream), "convert");
  m.ready = true;
  return TCE_T_OK;
}

FusedArgs tce_t::job_args(const Job &J, bool wide) {
  FusedArgs A;
  A.T1 = J.dT1;
  A.T2e = J.dT2e;
  A.T2m = J.dT2m;
  A.Vd2 = J.dVd2;
  A.Wd1 = J.dWd1;
  A.S = J.dS;
  A.evl = J.devl;
  A.tasks = wide ? J.dtasks8 : J.dtasks;
  A.ntasks = (int)(wide ? J.tasks8.size() : J.tasks.size());
  A.nitems = wide ? J.nitems8 : J.nitems;
  A.no = (int)J.T.no;
  A.nv = (int)J.T.nv;
  A.Ep = (int)J.Ep;
  A.Mp = (int)J.Mp;
  A.partials = wide ? J.dpart8 : J.dpart;
  return A;
}

MixedArgs tce_t::job_margs(const Job &J, int precision) {
  MixedArgs M;
  M.f = job_args(J, true);
  if (precision == TCE_T_FP64) {  // the 16x16 kernel on the FP64 tensors
    M.x2[0] = J.dT2e, M.y2[0] = J.dVd2, M.x1[0] = J.dWd1, M.y1[0] = J.dT2m;
    for (int p = 1; p < 3; ++p) M.x2[p] = M.y2[p] = M.x1[p] = M.y1[p] = nullptr;
    M.Ek = (int)J.Ep;
    M.Mk = (int)J.Mp;
    return M;
  }
  const Job::Mixed &m = J.mx[precision];
  for (int p = 0; p < 3; ++p) {
    M.x2[p] = m.x2[p];
    M.y2[p] = m.y2[p];
    M.x1[p] = m.x1[p];
    M.y1[p] = m.y1[p];
  }
  M.Ek = m.Ek;
  M.Mk = m.Mk;
  return M;
}

double tce_t::item_cost(const Job &J, int precision) {
  if (precision == TCE_T_FP64) return (double)(J.Ep + J.Mp);
  const Job::Mixed &m = J.mx[precision];
  return (double)(m.Ek + m.Mk) * (precision == TCE_T_FP32 ? 0.5 : (precision == TCE_T_BF16X3 ? 0.25 : 0.4));
}

extern "C" long tce_t_job_precision_bytes(const tce_t_job *job, int precision) {
  if (!job) return -1;
  const Job &J = *reinterpret_cast<const Job *>(job);
  if (precision == TCE_T_FP64) return 0;
  if (!valid_precision(precision)) return -1;
  long extra = mixed_bytes(J, precision);
  if (J.pbox != 8) extra += (long)J.tasks8.size() * sizeof(TaskDev) + J.nitems8 * 8 * 16;
  return extra;
}

extern "C" int tce_t_job_operands(tce_t_job *job, int precision, int tensor, void *out0,
                                  void *out1, void *out2, long *rows, long *kout) {
  if (!job || tensor < 0 || tensor > 3) return TCE_T_ERR_ARG;
  if (precision == TCE_T_FP64 || !valid_precision(precision)) return TCE_T_ERR_ARG;
  Job &J = *reinterpret_cast<Job *>(job);
  const long no = J.T.no, nv = J.T.nv;
  const long nrows[4] = {no * no * nv, no * nv * nv, no * no * nv, no * nv * nv};
  const long k = tensor < 2 ? mixed_extent(nv, precision) : mixed_extent(no, precision);
  if (rows) *rows = nrows[tensor];
  if (kout) *kout = k;
  if (!out0 && !out1 && !out2) return TCE_T_OK;
  if (!J.uploaded) {
    set_error("tce_t_job_operands: job not uploaded");
    return TCE_T_ERR_STATE;
  }
  if ((mixed_parts(precision) < 2 && out1) || (mixed_parts(precision) < 3 && out2))
    return TCE_T_ERR_ARG;
  const int rc = prepare_mixed(J, precision);
  if (rc != TCE_T_OK) return rc;
  Job::Mixed &m = J.mx[precision];
  void *const *src = tensor == 0 ? m.x2 : tensor == 1 ? m.y2 : tensor == 2 ? m.x1 : m.y1;
  const size_t bytes = (size_t)nrows[tensor] * k * (precision == TCE_T_FP32 ? 4 : 2);
  void *out[3] = {out0, out1, out2};
  for (int p = 0; p < 3; ++p)
    if (out[p] && bytes)
      HIPCHK(hipMemcpyAsync(out[p], src[p], bytes, hipMemcpyDeviceToHost, J.stream), "operands D2H");
  HIPCHK(hipStreamSynchronize(J.stream), "operands D2H");
  return TCE_T_OK;
}

extern "C" int tce_t_job_run(tce_t_job *job, int precision, double *e_bt, double *e_pt) {
  if (!job) return TCE_T_ERR_ARG;
  Job &J = *reinterpret_cast<Job *>(job);
  if (!J.uploaded) {
    set_error("tce_t_job_run before tce_t_job_upload");
    return TCE_T_ERR_STATE;
  }
  if (!valid_precision(precision)) {
    set_error("unsupported precision mode");
    return TCE_T_ERR_ARG;
  }
  // the 16x16 kernel: reduced precision, or FP64 when the job asks for it
  const bool mixed = precision != TCE_T_FP64 || J.k16;
  if (precision != TCE_T_FP64) {
    const int rc = prepare_mixed(J, precision);
    if (rc != TCE_T_OK) return rc;
  } else if (J.k16) {
    const int rc = ensure_tasks8(J);
    if (rc != TCE_T_OK) return rc;
  }
  const long ntasks = J.ngroups();
  J.taskE.assign(2 * ntasks, 0.0);
  if (ntasks == 0 || J.nitems == 0) {
    if (e_bt) *e_bt = 0.0;
    if (e_pt) *e_pt = 0.0;
    return TCE_T_OK;
  }
  int cus = 0;
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, J.device), "attr");
  {
    const int rc = ensure_table(J);
    if (rc != TCE_T_OK) return rc;
  }
  int lrc;
  const bool wide = mixed || J.pbox == 8;
  const FusedArgs A = job_args(J, wide);
  if (mixed) {
    const MixedSegment *sg = stage_table(J, job_margs(J, precision), J.nitems8);
    if (!sg) return TCE_T_ERR_HIP;
    HIPCHK(hipEventRecord(J.ev0, J.stream), "event");
    lrc = launch_fused_mixed(sg, 1, J.nitems8, precision,
                             (int)std::min<long>(std::max(1, cus), J.nitems8), J.stream, J.sparse,
                             sched_use(J.dsched));
  } else {
    const long grid = std::min<long>(J.grid, J.nitems);
    const Segment *sg = stage_table(J, A, J.nitems);
    if (!sg) return TCE_T_ERR_HIP;
    HIPCHK(hipEventRecord(J.ev0, J.stream), "event");
    lrc = J.pbox == 8 ? launch_fused_wg(sg, 1, J.nitems, (int)grid, J.stream)
                      : launch_fused_wave(sg, 1, J.nitems, (int)grid, J.stream);
  }
  if (lrc != 0) {
    set_error("fused kernel launch failed");
    return TCE_T_ERR_HIP;
  }
  HIPCHK(hipEventRecord(J.ev1, J.stream), "event");
  if (launch_task_reduce(job_bounds(J, wide), (int)ntasks, wide ? 8 : J.ppi, A.partials,
                         J.dtaskE, J.stream) != 0) {
    set_error("task_reduce launch failed");
    return TCE_T_ERR_HIP;
  }
  HIPCHK(hipMemcpyAsync(J.taskE.data(), J.dtaskE, 2 * ntasks * sizeof(double), hipMemcpyDeviceToHost,
                        J.stream),
         "D2H");
  HIPCHK(hipStreamSynchronize(J.stream), "fused_triples");
  HIPCHK(hipEventElapsedTime(&J.last_ms, J.ev0, J.ev1), "elapsed");
  double s4 = 0.0, s5 = 0.0;
  for (long t = 0; t < ntasks; ++t) {
    s4 += J.taskE[2 * t];
    s5 += J.taskE[2 * t + 1];
  }
  if (e_bt) *e_bt = s4;
  if (e_pt) *e_pt = s5;
  return TCE_T_OK;
}

// Restartable (T), ccsd_t_restart.F:58-262: the tasks in blocks of one outer
// virtual tile p4b = noab + b, b = *begin .. nvab.  Each block is one
// contiguous range of work items (p4b is the outermost task index); its
// E(T) (sum of its tasks' E(T) in task order) goes into table[b - 1] and
// *begin becomes b + 1, as the reference writes tce:restart_triples_table
// and tce:ccsd_t_restart_begin to the rtdb.  At most max_tiles blocks per
// call (<= 0: all).  When every block is done, *t_energy = sum of table.
extern "C" int tce_t_job_run_restart(tce_t_job *job, int precision, long *begin, double *table,
                                     long max_tiles, double *t_energy) {
  if (!job || !begin || !table) return TCE_T_ERR_ARG;
  Job &J = *reinterpret_cast<Job *>(job);
  if (!J.uploaded) {
    set_error("tce_t_job_run_restart before tce_t_job_upload");
    return TCE_T_ERR_STATE;
  }
  if (!valid_precision(precision)) {
    set_error("unsupported precision mode");
    return TCE_T_ERR_ARG;
  }
  if (J.tiling == TCE_T_TILING_ORBITAL) {
    set_error("restart blocks need the reference tiles: the job uses TCE_T_TILING_ORBITAL");
    return TCE_T_ERR_STATE;
  }
  const long noab = J.T.noab, nvab = J.T.nvab;
  if (*begin < 1) *begin = 1;
  const bool mixed = precision != TCE_T_FP64 || J.k16;
  if (precision != TCE_T_FP64) {
    const int rc = prepare_mixed(J, precision);
    if (rc != TCE_T_OK) return rc;
  } else if (J.k16) {
    const int rc = ensure_tasks8(J);
    if (rc != TCE_T_OK) return rc;
  }
  {
    const int rc = ensure_table(J);
    if (rc != TCE_T_OK) return rc;
  }
  int cus = 0;
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, J.device), "attr");
  const bool wide = mixed || J.pbox == 8;
  const FusedArgs A = job_args(J, wide);
  const std::vector<long> &bnd = wide ? J.bound8 : J.bound;
  const long ntasks = J.ntask();
  J.taskE.resize(2 * ntasks);
  long done = 0;
  while (*begin <= nvab && (max_tiles <= 0 || done < max_tiles)) {
    const long p4b = noab + *begin;
    long t0 = 0;
    while (t0 < ntasks && J.task_tiles[6 * t0] < p4b) ++t0;
    long t1 = t0;
    while (t1 < ntasks && J.task_tiles[6 * t1] == p4b) ++t1;
    double e = 0.0;
    if (t1 > t0) {
      const long i0 = bnd[t0], i1 = bnd[t1];
      int lrc;
      if (mixed) {
        const MixedSegment *sg = stage_table(J, job_margs(J, precision), i1, i0);
        if (!sg) return TCE_T_ERR_HIP;
        lrc = launch_fused_mixed(sg, 1, i1 - i0, precision,
                                 (int)std::min<long>(std::max(1, cus), i1 - i0), J.stream,
                                 J.sparse, sched_use(J.dsched));
      } else {
        const Segment *sg = stage_table(J, A, i1, i0);
        if (!sg) return TCE_T_ERR_HIP;
        const int grid = (int)std::min<long>(J.grid, i1 - i0);
        lrc = J.pbox == 8 ? launch_fused_wg(sg, 1, i1 - i0, grid, J.stream)
                          : launch_fused_wave(sg, 1, i1 - i0, grid, J.stream);
      }
      if (lrc != 0) {
        set_error("fused kernel launch failed");
        return TCE_T_ERR_HIP;
      }
      if (launch_task_reduce(job_bounds(J, wide) + t0, (int)(t1 - t0), wide ? 8 : J.ppi,
                             A.partials, J.dtaskE + 2 * t0, J.stream) != 0) {
        set_error("task_reduce launch failed");
        return TCE_T_ERR_HIP;
      }
      HIPCHK(hipMemcpyAsync(J.taskE.data() + 2 * t0, J.dtaskE + 2 * t0, 2 * (t1 - t0) * sizeof(double),
                            hipMemcpyDeviceToHost, J.stream),
             "D2H");
      HIPCHK(hipStreamSynchronize(J.stream), "restart block");
      for (long t = t0; t < t1; ++t) e += J.taskE[2 * t + 1];
    }
    table[*begin - 1] = e;
    ++*begin;
    ++done;
  }
  if (*begin > nvab && t_energy) {
    double s5 = 0.0;
    for (long b = 0; b < nvab; ++b) s5 += table[b];
    *t_energy = s5;
  }
  return TCE_T_OK;
}

// Triples tiles of one task for the other consumers of the reference's
// drivers: cr_ccsd_t.F:139-144, lambda_ccsd_t.F (ccsd_t_doubles),
// emb/emb2.F:293 and mrcc/ccsd_t/tce_mrcc_ccsdpt_subg.F:202 call
// ccsd_t_singles_l (ccsd_t_singles_l.F:1-28) and ccsd_t_doubles_l
// (ccsd_t_doubles_l.F:1-100, toggle 2) to add t3s / t3d of the task with
// tiles (t_h1b..t_p6b) into their triplesx(h3,h2,h1,p6,p5,p4) buffers, then
// apply their own functionals.  Here the task's items run through the fused
// kernel on the resident tensors (store_tiles epilogue) and the tiles come
// back to the host, where they are added to *singles / *doubles (either may
// be NULL).  Synchronous on the job's stream.
// The task's t3s / t3d tiles, or the slice p4 in [i0, i0 + ni), p5 in
// [j0, j0 + nj) of them (0-based within the tiles), added into the caller's
// arrays; ld5 = the p5 extent of the caller's layout (>= nj): element
// (p4', p5', p6, h1, h2, h3) of the slice at
// ((((p4' ld5 + p5') r6 + p6) r1 + h1) r2 + h2) r3 + h3.
int tce_t::tile_export_device(Job &J, int precision, long h1b, long h2b, long h3b, long p4b,
                              long p5b, long p6b, long i0, long ni, long j0, long nj, long ld5,
                              TaskDev &d, double **tiles) {
  if (!J.uploaded) {
    set_error("tce_t_job_triples_tiles before tce_t_job_upload");
    return TCE_T_ERR_STATE;
  }
  if (!valid_precision(precision)) {
    set_error("unsupported precision mode");
    return TCE_T_ERR_ARG;
  }
  const Tiles &T = J.T;
  const long noab = T.noab, nvab = T.nvab;
  if (!(1 <= h1b && h1b <= h2b && h2b <= h3b && h3b <= noab && noab < p4b && p4b <= p5b &&
        p5b <= p6b && p6b <= noab + nvab) ||
      !task_allowed(T, p4b, p5b, p6b, h1b, h2b, h3b)) {
    set_error("tile sextuple is not a task of the job");
    return TCE_T_ERR_ARG;
  }
  if (precision != TCE_T_FP64) {
    const int rc = prepare_mixed(J, precision);
    if (rc != TCE_T_OK) return rc;
  }
  if (ni <= 0 || nj <= 0 || i0 < 0 || j0 < 0 || i0 + ni > T.rng(p4b) || j0 + nj > T.rng(p5b) ||
      ld5 < nj) {
    set_error("tile slice outside the task's tiles");
    return TCE_T_ERR_ARG;
  }
  d = TaskDev();
  d.op4 = (int)(T.ofs(p4b) - T.no + i0);
  d.op5 = (int)(T.ofs(p5b) - T.no + j0);
  d.op6 = (int)(T.ofs(p6b) - T.no);
  d.oh1 = (int)T.ofs(h1b);
  d.oh2 = (int)T.ofs(h2b);
  d.oh3 = (int)T.ofs(h3b);
  d.r4 = (int)ni;
  d.r5 = (int)nj;
  d.r6 = (int)T.rng(p6b);
  d.r1 = (int)T.rng(h1b);
  d.r2 = (int)T.rng(h2b);
  d.r3 = (int)T.rng(h3b);
  d.nb4 = (d.r4 + 7) / 8;
  d.nb5 = (d.r5 + 7) / 8;
  d.nb6 = (d.r6 + 7) / 8;
  d.nb1 = (d.r1 + 3) / 4;
  d.nb2 = (d.r2 + 3) / 4;
  d.nb3 = (d.r3 + 3) / 4;
  d.item0 = 0;
  d.factor = 1.0;
  {
    const long tiles6[6] = {p4b, p5b, p6b, h1b, h2b, h3b};
    ClassRanges(T).set(T, tiles6, d);
  }
  const long n = (long)d.r1 * d.r2 * d.r3 * d.r4 * d.r5 * d.r6;
  const long items = (long)d.nb4 * d.nb5 * d.nb6 * d.nb1 * d.nb2 * d.nb3;
  constexpr size_t HEAD = 256;  // the task entry, then the tiles
  static_assert(sizeof(TaskDev) <= HEAD, "tile export header");
  const size_t need = HEAD + 2 * (size_t)n * sizeof(double);
  if (J.dtile_bytes < need) {
    if (J.dtile) HIPCHK(hipFree(J.dtile), "tile buffer");
    J.dtile = nullptr;
    J.dtile_bytes = 0;
    HIPCHK(hipMalloc(&J.dtile, need), "tile buffer");
    J.dtile_bytes = need;
  }
  {
    const int rc = ensure_table(J);
    if (rc != TCE_T_OK) return rc;
  }
  // the task entry travels in the pinned launch table, after the arguments
  // and the segment
  constexpr size_t TASK_AT = 512;
  static_assert((sizeof(MixedArgs) + 15) / 16 * 16 + sizeof(MixedSegment) <= TASK_AT &&
                    TASK_AT + sizeof(TaskDev) <= TAB_BYTES,
                "tile export table");
  std::memcpy((char *)J.tab_host + TASK_AT, &d, sizeof d);
  HIPCHK(hipMemcpyAsync(J.dtile, (char *)J.tab_host + TASK_AT, sizeof d, hipMemcpyHostToDevice,
                        J.stream),
         "task entry");
  MixedArgs M = job_margs(J, precision);
  M.f.tasks = (const TaskDev *)J.dtile;
  M.f.ntasks = 1;
  M.f.nitems = items;
  M.f.partials = (double *)((char *)J.dtile + HEAD);
  const MixedSegment *sg = stage_table(J, M, items);
  if (!sg) {
    set_error("launch table copy failed");
    return TCE_T_ERR_HIP;
  }
  int cus = 0;
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, J.device), "attr");
  if (launch_fused_tiles(sg, items, precision, (int)std::min<long>(std::max(1, cus), items),
                         J.stream, J.sparse, sched_use(J.dsched)) != 0) {
    set_error("tile kernel launch failed");
    return TCE_T_ERR_HIP;
  }
  *tiles = M.f.partials;
  return TCE_T_OK;
}

static int triples_tiles(Job &J, int precision, long h1b, long h2b, long h3b, long p4b, long p5b,
                         long p6b, long i0, long ni, long j0, long nj, long ld5, double *singles,
                         double *doubles) {
  TaskDev d;
  double *dev = nullptr;
  {
    const int rc = tile_export_device(J, precision, h1b, h2b, h3b, p4b, p5b, p6b, i0, ni, j0, nj,
                                      ld5, d, &dev);
    if (rc != TCE_T_OK) return rc;
  }
  const long n = (long)d.r1 * d.r2 * d.r3 * d.r4 * d.r5 * d.r6;
  // Back to the host in chunks through two pinned buffers: the copy of
  // chunk k+1 runs while host threads add chunk k into the caller's arrays.
  // Device layout: doubles [0, n), singles [n, 2n); a NULL array's half is
  // not copied.
  const long first = doubles ? 0 : n, last = singles ? 2 * n : n;
  if (first >= last) {
    HIPCHK(hipStreamSynchronize(J.stream), "fused_triples_tiles");
    return TCE_T_OK;
  }
  constexpr long CH = 8L << 20;  // doubles per chunk (64 MiB)
  double *stage = nullptr;
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = TCE_T_OK;
  if (!hip_ok(hipHostMalloc(&stage, 2 * CH * sizeof(double)), "tile staging") ||
      !hip_ok(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming), "event") ||
      !hip_ok(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming), "event"))
    rc = TCE_T_ERR_HIP;
  const long nch = (last - first + CH - 1) / CH;
  auto issue = [&](long k) {
    const long off = first + k * CH, len = std::min(CH, last - off);
    return hip_ok(hipMemcpyAsync(stage + (k & 1) * CH, dev + off, len * sizeof(double),
                                 hipMemcpyDeviceToHost, J.stream),
                  "D2H") &&
           hip_ok(hipEventRecord(ev[k & 1], J.stream), "event");
  };
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (rc == TCE_T_OK && !issue(0)) rc = TCE_T_ERR_HIP;
  for (long k = 0; k < nch && rc == TCE_T_OK; ++k) {
    if (!hip_ok(hipEventSynchronize(ev[k & 1]), "D2H")) {
      rc = TCE_T_ERR_HIP;
      break;
    }
    if (k + 1 < nch && !issue(k + 1)) rc = TCE_T_ERR_HIP;  // into the buffer of chunk k-1
    const long off = first + k * CH, len = std::min(CH, last - off);
    const double *src = stage + (k & 1) * CH;
    // element g of the device tile (p5 extent nj) at g + p4'(ld5 - nj) inner
    // of the caller's layout (p5 extent ld5)
    const long inner = (long)d.r6 * d.r1 * d.r2 * d.r3, per4 = nj * inner, gap = (ld5 - nj) * inner;
    auto add = [&](long a, long b) {  // chunk positions [a, b)
      for (long i = a; i < b; ++i) {
        const long g = off + i;
        if (g < n)
          doubles[g + (g / per4) * gap] += src[i];
        else
          singles[(g - n) + ((g - n) / per4) * gap] += src[i];
      }
    };
    std::vector<std::thread> th;
    const long part = (len + nt - 1) / nt;
    for (unsigned t = 1; t < nt && (long)t * part < len; ++t)
      th.emplace_back(add, (long)t * part, std::min(len, (long)(t + 1) * part));
    add(0, std::min(len, part));
    for (auto &x : th) x.join();
  }
  if (rc != TCE_T_OK) (void)hipStreamSynchronize(J.stream);  // no copy left in flight
  for (hipEvent_t e : ev)
    if (e) (void)hipEventDestroy(e);
  if (stage) (void)hipHostFree(stage);
  return rc;
}

extern "C" int tce_t_job_triples_tiles(tce_t_job *job, int precision, long h1b, long h2b,
                                       long h3b, long p4b, long p5b, long p6b, double *singles,
                                       double *doubles) {
  if (!job) return TCE_T_ERR_ARG;
  Job &J = *reinterpret_cast<Job *>(job);
  if (p4b < 1 || p5b < 1 || p4b > J.T.noab + J.T.nvab || p5b > J.T.noab + J.T.nvab) {
    set_error("tile sextuple is not a task of the job");
    return TCE_T_ERR_ARG;
  }
  const long r4 = J.T.rng(p4b), r5 = J.T.rng(p5b);
  return triples_tiles(J, precision, h1b, h2b, h3b, p4b, p5b, p6b, 0, r4, 0, r5, r5, singles,
                       doubles);
}

extern "C" int tce_t_job_triples_tiles_device(tce_t_job *job, int precision, long h1b, long h2b,
                                              long h3b, long p4b, long p5b, long p6b,
                                              const double **singles, const double **doubles,
                                              long *elements) {
  if (!job || !singles || !doubles) return TCE_T_ERR_ARG;
  Job &J = *reinterpret_cast<Job *>(job);
  if (p4b < 1 || p5b < 1 || p4b > J.T.noab + J.T.nvab || p5b > J.T.noab + J.T.nvab) {
    set_error("tile sextuple is not a task of the job");
    return TCE_T_ERR_ARG;
  }
  const long r5 = J.T.rng(p5b);
  TaskDev d;
  double *dev = nullptr;
  const int rc = tile_export_device(J, precision, h1b, h2b, h3b, p4b, p5b, p6b, 0, J.T.rng(p4b), 0,
                                    r5, r5, d, &dev);
  if (rc != TCE_T_OK) return rc;
  HIPCHK(hipStreamSynchronize(J.stream), "fused_triples_tiles");
  const long n = (long)d.r1 * d.r2 * d.r3 * d.r4 * d.r5 * d.r6;
  *doubles = dev;
  *singles = dev + n;
  if (elements) *elements = n;
  return TCE_T_OK;
}

extern "C" int tce_t_job_triples_tiles_6dts(tce_t_job *job, int precision, long h1b, long h2b,
                                            long h3b, long p4b, long p5b, long p6b, long istart,
                                            long istop, long jstart, long jstop, long maxp4,
                                            long maxp5, double *singles, double *doubles) {
  if (!job) return TCE_T_ERR_ARG;
  Job &J = *reinterpret_cast<Job *>(job);
  if (istop - istart + 1 > maxp4 || jstop - jstart + 1 > maxp5) {
    set_error("6dts slice larger than maxp4 x maxp5");
    return TCE_T_ERR_ARG;
  }
  if (p4b < 1 || p5b < 1 || p4b > J.T.noab + J.T.nvab || p5b > J.T.noab + J.T.nvab) {
    set_error("tile sextuple is not a task of the job");
    return TCE_T_ERR_ARG;
  }
  return triples_tiles(J, precision, h1b, h2b, h3b, p4b, p5b, p6b, istart - 1,
                       istop - istart + 1, jstart - 1, jstop - jstart + 1, maxp5, singles, doubles);
}

extern "C" int tce_t_job_task_energies(const tce_t_job *job, double *out, long ntasks) {
  if (!job || !out) return TCE_T_ERR_ARG;
  const Job &J = *reinterpret_cast<const Job *>(job);
  if (J.tiling == TCE_T_TILING_ORBITAL) {
    set_error("per-task energies need the reference tiles: the job uses TCE_T_TILING_ORBITAL");
    return TCE_T_ERR_STATE;
  }
  if (J.ntask() != ntasks || (long)J.taskE.size() != 2 * ntasks) return TCE_T_ERR_STATE;
  std::memcpy(out, J.taskE.data(), 2 * ntasks * sizeof(double));
  return TCE_T_OK;
}

extern "C" double tce_t_job_last_kernel_ms(const tce_t_job *job) {
  return job ? (double)reinterpret_cast<const Job *>(job)->last_ms : -1.0;
}

extern "C" int tce_t_job_kernel(const tce_t_job *job) {
  if (!job) return TCE_T_ERR_ARG;
  const Job &J = *reinterpret_cast<const Job *>(job);
  return J.k16 ? TCE_T_KERNEL_WG16 : (J.pbox == 8 ? TCE_T_KERNEL_WG : TCE_T_KERNEL_WAVE);
}

void tce_t::job_release(Job &J) {
  if (J.stream) (void)hipStreamSynchronize(J.stream);
  if (J.dtasks8 && J.dtasks8 != J.dtasks) (void)hipFree(J.dtasks8);
  if (J.dpart8 && J.dpart8 != J.dpart) (void)hipFree(J.dpart8);
  if (J.dbound8 && J.dbound8 != J.dbound) (void)hipFree(J.dbound8);
  J.dtasks8 = nullptr;
  J.dpart8 = nullptr;
  J.dbound8 = nullptr;
  if (J.dbound) (void)hipFree(J.dbound);
  J.dbound = nullptr;
  double **bufs[] = {&J.dT1, &J.dT2e, &J.dT2m, &J.dVd2, &J.dWd1, &J.dS, &J.devl, &J.dpart, &J.dtaskE};
  for (double **b : bufs) {
    if (*b) (void)hipFree(*b);
    *b = nullptr;
  }
  if (J.dtasks) (void)hipFree(J.dtasks);
  J.dtasks = nullptr;
  if (J.dtile) (void)hipFree(J.dtile);
  J.dtile = nullptr;
  J.dtile_bytes = 0;
  for (auto &m : J.mx) {
    for (int p = 0; p < 3; ++p)
      for (void **b : {&m.x2[p], &m.y2[p], &m.x1[p], &m.y1[p]}) {
        if (*b) (void)hipFree(*b);
        *b = nullptr;
      }
    m.ready = false;
  }
  J.uploaded = false;
}

extern "C" int tce_t_job_destroy(tce_t_job *job) {
  if (!job) return TCE_T_OK;
  Job *J = reinterpret_cast<Job *>(job);
  job_release(*J);
  if (J->tab_dev) (void)hipFree(J->tab_dev);
  if (J->tab_host) (void)hipHostFree(J->tab_host);
  if (J->ev0) (void)hipEventDestroy(J->ev0);
  if (J->ev1) (void)hipEventDestroy(J->ev1);
  if (J->dsched) (void)hipFree(J->dsched);
  if (J->stream) (void)hipStreamDestroy(J->stream);
  delete J;
  return TCE_T_OK;
}

extern "C" int tce_t_pack_dense(const tce_t_tiles *tiles, const double *t1, const long *t1h,
                                const double *t2, const long *t2h, const double *v2,
                                const long *v2h, double *T1, double *T2e, double *T2m, double *Wd1,
                                double *Vd2, double *S) {
  Tiles T;
  if (!T.load(tiles) || !t1 || !t1h || !t2 || !t2h || !v2 || !v2h) return TCE_T_ERR_ARG;
  DenseHost D;
  int rc = pack_dense(T, round8(T.nv), round8(T.no), t1, t1h, t2, t2h, v2, v2h, D);
  if (rc != TCE_T_OK) return rc;
  auto put = [](double *dst, const std::vector<double> &v) {
    if (dst) std::memcpy(dst, v.data(), v.size() * sizeof(double));
  };
  put(T1, D.T1);
  put(T2e, D.T2e);
  put(T2m, D.T2m);
  put(Wd1, D.Wd1);
  put(Vd2, D.Vd2);
  put(S, D.S);
  return TCE_T_OK;
}
