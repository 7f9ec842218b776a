// lambda.hip -- the Lambda-CCSD(T) energies of lambda_ccsd_t.F from two
// uploaded jobs on one device.
//
// lambda_ccsd_t.F:56-200 runs the (T) task loop and forms, per task,
//   tdoubles  = ccsd_t_doubles (the t3 doubles tile from t2 and v2),
//   ysingles  = lambda_ccsd_t_left toggle 1:  P(9) y(h4 p1) v(h5 h6 p2 p3),
//   ydoubles  = lambda_ccsd_t_left toggle 2:  P(9) y(h4 h5 p1 p2) f(h6 p3)
//               - P(9) Sum(h7) y(h4 h7 p1 p2) v(h5 h6 h7 p3)
//               - P(9) Sum(p7) y(h4 h5 p1 p7) v(h6 p7 p2 p3)
// (lambda_ccsd_t_left.F:7-10), and sums
//   energy1 += factor * tdoubles * ydoubles / D,
//   energy2 += factor * tdoubles * (ysingles + ydoubles) / D.
// For real orbitals <pq||rs> = <rs||pq>, so the y1 term and the two v2
// terms of y2 are the s1, d1 and d2 terms of the t3 tiles (ccsd_t_singles.F,
// ccsd_t_doubles.F) with t1, t2 replaced by y1, y2: a job uploaded with the
// Lambda amplitudes in place of t1, t2 (tce_t_job_upload_lambda) computes
// them with the fused kernel from the (T) orientation of the integrals.
// The Fock term is formed here element by element.  Per task: both jobs'
// tiles on the device (tile_export_device), then lambda_energy over the
// tile's elements with fixed-order block partials, summed on the host in
// task order.
//
// Element pairing.  ccsd_t_doubles leaves tdoubles in the order
// [p4][p5][p6][h1][h2][h3] (ccsd_t_doubles.F:209-211, sort 6,5,3,4,2,1),
// lambda_ccsd_t_left leaves its tiles in [h1][h2][h3][p4][p5][p6]
// (lambda_ccsd_t_left.F:192-194, sort 6,4,3,5,2,1 of [p3][p2][h6][h5][p1][h4];
// the same order in every term), and the energy loop (lambda_ccsd_t.F:
// 156-178) multiplies element i of one by element i of the other.
// TCE_T_LAMBDA_ORBITAL pairs every t3 element with the y3 element of the
// same six orbitals (the Lambda-CCSD(T) energy); TCE_T_LAMBDA_REFERENCE_ORDER
// pairs them by position in the two orders, as that loop does.  The two
// agree when every tile holds one orbital.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "job.h"

namespace tce_t {
namespace {

constexpr int LB = 256;        // threads per block
constexpr int LBLOCKS = 1024;  // blocks per task (grid-stride over the tile)

struct LamTile {
  int op4, op5, op6, oh1, oh2, oh3;  // orbital offsets within the spaces
  int r4, r5, r6, r1, r2, r3;        // tile ranges
};

// Element e of the t3 tile, ((((a r5 + b) r6 + c) r1 + i) r2 + j) r3 + k
// (local offsets; triplesx(h3,h2,h1,p6,p5,p4) of the reference), is paired
// with the y3 element of local offsets (a', b', c', i', j', k'): the same
// ones, or (ref_order) those at position e of ((((i' r2 + j') r3 + k') r4 +
// a') r5 + b') r6 + c'.  Ld, Ls: the left job's tiles in the t3 order.  y2
// from the left job's dense T2e (T2e[i][j][a][b] = y(a b i j), b padded to
// Ep); F the hole-particle Fock block F[h][p] or null.
__global__ void __launch_bounds__(LB) lambda_energy(const double *R, const double *Ld,
                                                    const double *Ls, long n, LamTile t,
                                                    int ref_order, const double *evl, int no,
                                                    int nv, const double *Y2, long Ep,
                                                    const double *F, double factor,
                                                    double *part) {
  double e1 = 0.0, e2 = 0.0;
  for (long e = (long)blockIdx.x * LB + threadIdx.x; e < n; e += (long)gridDim.x * LB) {
    long x = e;
    const int lk = (int)(x % t.r3); x /= t.r3;
    const int lj = (int)(x % t.r2); x /= t.r2;
    const int li = (int)(x % t.r1); x /= t.r1;
    const int lc = (int)(x % t.r6); x /= t.r6;
    const int lb = (int)(x % t.r5); x /= t.r5;
    const int la = (int)x;
    const double den = -evl[no + t.op4 + la] - evl[no + t.op5 + lb] - evl[no + t.op6 + lc] +
                       evl[t.oh1 + li] + evl[t.oh2 + lj] + evl[t.oh3 + lk];
    int ya = la, yb = lb, yc = lc, yi = li, yj = lj, yk = lk;
    long ye = e;
    if (ref_order) {
      x = e;
      yc = (int)(x % t.r6); x /= t.r6;
      yb = (int)(x % t.r5); x /= t.r5;
      ya = (int)(x % t.r4); x /= t.r4;
      yk = (int)(x % t.r3); x /= t.r3;
      yj = (int)(x % t.r2); x /= t.r2;
      yi = (int)x;
      ye = (((((long)ya * t.r5 + yb) * t.r6 + yc) * t.r1 + yi) * t.r2 + yj) * t.r3 + yk;
    }
    double yf = 0.0;
    if (F) {
      // P(c/ab) P(k/ij) y(a b i j) f(k c) at the y3 element's orbitals
      const int a = t.op4 + ya, b = t.op5 + yb, c = t.op6 + yc;
      const int i = t.oh1 + yi, j = t.oh2 + yj, k = t.oh3 + yk;
      auto y = [&](int p, int q, int h, int g) {
        return Y2[(((long)h * no + g) * nv + p) * Ep + q];
      };
      auto f = [&](int h, int p) { return F[(long)h * nv + p]; };
      yf = y(a, b, i, j) * f(k, c) - y(c, b, i, j) * f(k, a) - y(a, c, i, j) * f(k, b) -
           y(a, b, k, j) * f(i, c) + y(c, b, k, j) * f(i, a) + y(a, c, k, j) * f(i, b) -
           y(a, b, i, k) * f(j, c) + y(c, b, i, k) * f(j, a) + y(a, c, i, k) * f(j, b);
    }
    const double r = R[e], yd = Ld[ye] + yf;
    e1 += factor * r * yd / den;
    e2 += factor * r * (Ls[ye] + yd) / den;
  }
  __shared__ double s1[LB], s2[LB];
  s1[threadIdx.x] = e1;
  s2[threadIdx.x] = e2;
  __syncthreads();
  for (int w = LB / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s1[threadIdx.x] += s1[threadIdx.x + w];
      s2[threadIdx.x] += s2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s1[0];
    part[2 * blockIdx.x + 1] = s2[0];
  }
}

bool same_tiles(const Tiles &A, const Tiles &B) {
  return A.noab == B.noab && A.nvab == B.nvab && A.restricted == B.restricted &&
         A.range == B.range && A.spin == B.spin && A.sym == B.sym && A.evl == B.evl;
}

// tce_restricted_2 (tce_restricted.F:1-32): in a restricted run an all-beta
// pair is read from its alpha image.
void restrict2(const Tiles &T, long &a, long &b) {
  if (T.restricted && T.spn(a) + T.spn(b) == 4) {
    a = T.alp(a);
    b = T.alp(b);
  }
}

// The hole-particle block of the hashed one-electron operator (the
// f(h6 p3) blocks lambda_ccsd_t_left_2 reads, lambda_ccsd_t_left.F:397-400:
// key p3 - 1 + (noab + nvab)(h6 - 1), block [h6][p3]) as F[h][p] over
// every orbital; blocks that break spin or irrep conservation are zero.
int fock_hole_particle(const Tiles &T, const double *f1, const long *f1h,
                       std::vector<double> &F) {
  const long noab = T.noab, N = T.noab + T.nvab;
  F.assign((size_t)T.no * T.nv, 0.0);
  for (long hb = 1; hb <= noab; ++hb)
    for (long pb = noab + 1; pb <= N; ++pb) {
      if (T.spn(hb) != T.spn(pb) || (T.sy(hb) ^ T.sy(pb)) != 0) continue;
      long h = hb, p = pb;
      restrict2(T, h, p);
      const long off = hash_lookup(f1h, p - 1 + N * (h - 1));
      if (off < 0) {
        set_error("tce_t_lambda_ccsd_t: the Fock operator lacks a hole-particle block");
        return TCE_T_ERR_ARG;
      }
      const long rh = T.rng(hb), rp = T.rng(pb);
      const long oh = T.ofs(hb), op = T.ofs(pb) - T.no;
      for (long i = 0; i < rh; ++i)
        for (long a = 0; a < rp; ++a) F[(size_t)(oh + i) * T.nv + op + a] = f1[off + i * rp + a];
    }
  return TCE_T_OK;
}

// Orbital pairing on the job's own kernel tasks (sub-tiles of the reference
// tiles, each box once with its symmetry factor: the product t3 y3 / D is
// symmetric under the permutations within equal tiles, as t3 t3 / D is).
// The items run in windows of W: the left-hand pass leaves each item's y3
// doubles (Fock term added) and singles in a window buffer in register
// order, the right-hand pass forms the energy partials of the same items
// from its t3 doubles and the window; the job's task reduction then gives
// the per-task sums.  No tile leaves the registers in the tile layout.
// cr_den (CR-CCSD(T)): the right-hand pass is the CR one; te gets the
// moment sums, *te_den the denominator sums, per task.
int lambda_fused(Job &JR, Job &JL, const double *dF, int precision, std::vector<double> &te,
                 double *cr_den = nullptr, std::vector<double> *te_den = nullptr) {
  for (Job *J : {&JR, &JL}) {
    int rc = ensure_tasks8(*J);
    if (rc == TCE_T_OK && precision != TCE_T_FP64) rc = prepare_mixed(*J, precision);
    if (rc != TCE_T_OK) return rc;
  }
  if (JR.nitems8 != JL.nitems8 || JR.tasks8.size() != JL.tasks8.size()) {
    set_error("tce_t_lambda_ccsd_t: the jobs' kernel tasks differ");
    return TCE_T_ERR_ARG;
  }
  const long N = JR.nitems8, nt = JR.ntask();
  constexpr long SLOT = 8L * 8192 * sizeof(double);  // window bytes per item (8 waves)
  size_t free_b = 0, total_b = 0;
  HIPCHK(hipMemGetInfo(&free_b, &total_b), "memory info");
  const long budget = std::min<long>(16L << 30, (long)(free_b / 4));
  long W = std::max(1L, std::min(N, budget / SLOT));
  if (const char *w = getenv("TCE_T_LAMBDA_WINDOW")) {  // items per window (tests, tuning)
    const long x = atol(w);
    if (x > 0) W = std::min(W, x);
  }
  int cus = 0;
  HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, JR.device), "attr");
  double *win = nullptr;
  HIPCHK(hipMalloc(&win, (size_t)W * SLOT), "lambda window");
  int rc = TCE_T_OK;
  for (long i0 = 0; i0 < N && rc == TCE_T_OK; i0 += W) {
    const long i1 = std::min(N, i0 + W);
    const int grid = (int)std::min<long>(std::max(1, cus), i1 - i0);
    MixedArgs ML = job_margs(JL, precision);
    ML.f.partials = win;
    ML.lam_f = dF;
    ML.lam_item0 = i0;
    const MixedSegment *sl = stage_mixed(JL, ML, i1, i0);
    if (!sl || launch_fused_lambda(sl, i1 - i0, precision, grid, JL.stream, JL.sparse,
                                   sched_use(JL.dsched), true) != 0) {
      set_error("lambda left pass: launch failed");
      rc = TCE_T_ERR_HIP;
      break;
    }
    if (!hip_ok(hipStreamSynchronize(JL.stream), "lambda left pass")) {
      rc = TCE_T_ERR_HIP;
      break;
    }
    MixedArgs MR = job_margs(JR, precision);
    MR.lam_y = win;
    MR.lam_item0 = i0;
    MR.cr_den = cr_den;
    const MixedSegment *sr = stage_mixed(JR, MR, i1, i0);
    if (!sr || (cr_den ? launch_fused_cr_right(sr, i1 - i0, precision, grid, JR.stream, JR.sparse,
                                               sched_use(JR.dsched))
                       : launch_fused_lambda(sr, i1 - i0, precision, grid, JR.stream, JR.sparse,
                                             sched_use(JR.dsched), false)) != 0) {
      set_error("lambda right pass: launch failed");
      rc = TCE_T_ERR_HIP;
      break;
    }
    if (!hip_ok(hipStreamSynchronize(JR.stream), "lambda right pass")) rc = TCE_T_ERR_HIP;
  }
  // per-task sums of one partials buffer, in the order of the task list
  auto reduce = [&](const double *part, std::vector<double> &out) {
    out.assign(2 * nt, 0.0);
    if (launch_task_reduce(job_bounds(JR, true), (int)nt, 8, part, JR.dtaskE, JR.stream) != 0) {
      set_error("task_reduce launch failed");
      rc = TCE_T_ERR_HIP;
    } else if (!hip_ok(hipMemcpyAsync(out.data(), JR.dtaskE, 2 * nt * sizeof(double),
                                      hipMemcpyDeviceToHost, JR.stream),
                       "D2H") ||
               !hip_ok(hipStreamSynchronize(JR.stream), "lambda task sums")) {
      rc = TCE_T_ERR_HIP;
    }
  };
  if (rc == TCE_T_OK) reduce(JR.dpart8, te);
  if (rc == TCE_T_OK && cr_den) reduce(cr_den, *te_den);
  if (rc != TCE_T_OK) {  // nothing left in flight on either stream
    (void)hipStreamSynchronize(JL.stream);
    (void)hipStreamSynchronize(JR.stream);
  }
  (void)hipFree(win);
  return rc;
}

// The overlap den0 of cr_ccsd_t_D.F:4-7 with c = t (cr_ccsd_t.F:58-59):
//   sum_(a i) t(a i)^2 + 1/4 sum_(a b i j) t(a b i j)^2
//   + 1/2 sum_(a b i j) t(a i) t(b j) t(a b i j),
// all spin-orbital sums, from the job's dense T1[a][i] and T2e[i][j][a][b]
// (b padded to Ep).  One partial per block, summed on the host in order.
constexpr int OV_BLOCKS = 1024, OV_THREADS = 256;
__global__ void __launch_bounds__(OV_THREADS)
    cr_overlap(const double *T1, const double *T2e, int no, int nv, long Ep, double *part) {
  __shared__ double red[OV_THREADS];
  const long stride = (long)gridDim.x * OV_THREADS, x0 = (long)blockIdx.x * OV_THREADS + threadIdx.x;
  double acc = 0.0;
  for (long x = x0; x < (long)nv * no; x += stride) acc += T1[x] * T1[x];
  const long n2 = (long)no * no * nv * nv;
  for (long x = x0; x < n2; x += stride) {
    const int b = (int)(x % nv);
    long r = x / nv;
    const int a = (int)(r % nv);
    r /= nv;
    const int j = (int)(r % no), i = (int)(r / no);
    const double t = T2e[(((long)i * no + j) * nv + a) * Ep + b];
    acc += 0.25 * t * t + 0.5 * T1[(long)a * no + i] * T1[(long)b * no + j] * t;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = OV_THREADS / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

}  // namespace
}  // namespace tce_t

using namespace tce_t;

// The Lambda amplitudes in the excitation layout of t1 and t2: every block
// of the t1 table (tce_t1_offset.F, [p][h]) and t2 table (tce_t2_offset.F,
// [p1][p2][h3][h4]) filled from the y block of the same tiles
// (tce_y1_offset.F, [h][p]; tce_y2_offset.F, [h3][h4][p1][p2]), then one
// tce_t_job_upload.
extern "C" int tce_t_job_upload_lambda(tce_t_job *job, const double *y1, const long *k_y1_offset,
                                       const double *y2, const long *k_y2_offset,
                                       const double *v2, const long *k_v2_offset) {
  if (!job || !y1 || !k_y1_offset || !y2 || !k_y2_offset || !v2 || !k_v2_offset) {
    set_error("tce_t_job_upload_lambda: null argument");
    return TCE_T_ERR_ARG;
  }
  const Tiles &T = reinterpret_cast<Job *>(job)->T;
  const long noab = T.noab, nvab = T.nvab;
  long s1 = 0, s2 = 0;
  const long L1 = t1_offset(T, nullptr, &s1), L2 = t2_offset(T, nullptr, &s2);
  std::vector<long> h1(1 + 2 * L1), h2(1 + 2 * L2);
  t1_offset(T, h1.data(), nullptr);
  t2_offset(T, h2.data(), nullptr);
  std::vector<double> t1(std::max(1L, s1)), t2(std::max(1L, s2));
  for (long n = 0; n < L1; ++n) {
    const long key = h1[1 + n], off = h1[1 + L1 + n];
    const long h = key % noab + 1, p = key / noab + noab + 1;
    const long yo = hash_lookup(k_y1_offset, p - noab - 1 + nvab * (h - 1));
    if (yo < 0) {
      set_error("tce_t_job_upload_lambda: y1 lacks a block of the t1 table");
      return TCE_T_ERR_ARG;
    }
    const long rp = T.rng(p), rh = T.rng(h);
    for (long a = 0; a < rp; ++a)
      for (long i = 0; i < rh; ++i) t1[off + a * rh + i] = y1[yo + i * rp + a];
  }
  for (long n = 0; n < L2; ++n) {
    long key = h2[1 + n];
    const long off = h2[1 + L2 + n];
    const long h4 = key % noab + 1;
    key /= noab;
    const long h3 = key % noab + 1;
    key /= noab;
    const long p2 = key % nvab + noab + 1, p1 = key / nvab + noab + 1;
    const long yo = hash_lookup(
        k_y2_offset, p2 - noab - 1 + nvab * (p1 - noab - 1 + nvab * (h4 - 1 + noab * (h3 - 1))));
    if (yo < 0) {
      set_error("tce_t_job_upload_lambda: y2 lacks a block of the t2 table");
      return TCE_T_ERR_ARG;
    }
    const long r1 = T.rng(p1), r2 = T.rng(p2), r3 = T.rng(h3), r4 = T.rng(h4);
    for (long a = 0; a < r1; ++a)
      for (long b = 0; b < r2; ++b)
        for (long i = 0; i < r3; ++i)
          for (long j = 0; j < r4; ++j)
            t2[off + ((a * r2 + b) * r3 + i) * r4 + j] =
                y2[yo + ((i * r4 + j) * r1 + a) * r2 + b];
  }
  return tce_t_job_upload(job, t1.data(), h1.data(), t2.data(), h2.data(), v2, k_v2_offset);
}

extern "C" int tce_t_lambda_ccsd_t(tce_t_job *right, tce_t_job *left, const double *f1,
                                   const long *k_f1_offset, int pairing, int precision,
                                   double *energy1, double *energy2, double *task_energies,
                                   long ntasks) {
  if (!right || !left || right == left) {
    set_error("tce_t_lambda_ccsd_t: two distinct jobs are needed");
    return TCE_T_ERR_ARG;
  }
  Job &JR = *reinterpret_cast<Job *>(right);
  Job &JL = *reinterpret_cast<Job *>(left);
  if (!JR.uploaded || !JL.uploaded) {
    set_error("tce_t_lambda_ccsd_t before tce_t_job_upload");
    return TCE_T_ERR_STATE;
  }
  if (!valid_precision(precision)) {
    set_error("unsupported precision mode");
    return TCE_T_ERR_ARG;
  }
  if (pairing != TCE_T_LAMBDA_ORBITAL && pairing != TCE_T_LAMBDA_REFERENCE_ORDER) {
    set_error("tce_t_lambda_ccsd_t: unknown element pairing");
    return TCE_T_ERR_ARG;
  }
  if (!same_tiles(JR.T, JL.T) || JR.device != JL.device) {
    set_error("tce_t_lambda_ccsd_t: the jobs' tiles (or devices) differ");
    return TCE_T_ERR_ARG;
  }
  if (!f1 != !k_f1_offset) {
    set_error("tce_t_lambda_ccsd_t: f1 and k_f1_offset go together");
    return TCE_T_ERR_ARG;
  }
  const Tiles &T = JR.T;
  const long nt = JR.ntask();
  if (task_energies && ntasks < nt) {
    set_error("tce_t_lambda_ccsd_t: task_energies holds fewer than the job's tasks");
    return TCE_T_ERR_ARG;
  }
  HIPCHK(hipSetDevice(JR.device), "device");
  std::vector<double> Fh;
  if (f1) {
    const int rc = fock_hole_particle(T, f1, k_f1_offset, Fh);
    if (rc != TCE_T_OK) return rc;
  }
  double *dF = nullptr, *dpart = nullptr;
  std::vector<double> hpart(2 * LBLOCKS);
  auto release = [&]() {
    if (dF) (void)hipFree(dF);
    if (dpart) (void)hipFree(dpart);
  };
  if (f1 && (!hip_ok(hipMalloc(&dF, Fh.size() * sizeof(double)), "Fock block") ||
             !hip_ok(hipMemcpy(dF, Fh.data(), Fh.size() * sizeof(double), hipMemcpyHostToDevice),
                     "Fock block"))) {
    release();
    return TCE_T_ERR_HIP;
  }
  const char *mode = getenv("TCE_T_LAMBDA");
  if (pairing == TCE_T_LAMBDA_ORBITAL && JR.tiling == JL.tiling &&
      JR.tiling != TCE_T_TILING_ORBITAL && !(mode && !strcmp(mode, "tiles"))) {
    std::vector<double> te;
    const int rc = lambda_fused(JR, JL, dF, precision, te);
    release();
    if (rc != TCE_T_OK) return rc;
    double s1 = 0.0, s2 = 0.0;
    for (long t = 0; t < nt; ++t) {
      s1 += te[2 * t];
      s2 += te[2 * t + 1];
    }
    if (task_energies) std::copy(te.begin(), te.end(), task_energies);
    if (energy1) *energy1 = s1;
    if (energy2) *energy2 = s2;
    return TCE_T_OK;
  }
  // Whole tiles per task: the reference-order pairing (and ORBITAL-tiled jobs,
  // whose kernel tasks do not follow the reference tasks)
  if (!hip_ok(hipMalloc(&dpart, 2 * LBLOCKS * sizeof(double)), "lambda partials")) {
    release();
    return TCE_T_ERR_HIP;
  }
  double s1 = 0.0, s2 = 0.0;
  for (long t = 0; t < nt; ++t) {
    const long *tt = &JR.task_tiles[6 * t];  // p4b p5b p6b h1b h2b h3b
    const long p4b = tt[0], p5b = tt[1], p6b = tt[2], h1b = tt[3], h2b = tt[4], h3b = tt[5];
    TaskDev dr, dl;
    double *R = nullptr, *L = nullptr;
    int rc = tile_export_device(JR, precision, h1b, h2b, h3b, p4b, p5b, p6b, 0, T.rng(p4b), 0,
                                T.rng(p5b), T.rng(p5b), dr, &R);
    if (rc == TCE_T_OK)
      rc = tile_export_device(JL, precision, h1b, h2b, h3b, p4b, p5b, p6b, 0, T.rng(p4b), 0,
                              T.rng(p5b), T.rng(p5b), dl, &L);
    if (rc != TCE_T_OK) {
      (void)hipStreamSynchronize(JR.stream);
      (void)hipStreamSynchronize(JL.stream);
      release();
      return rc;
    }
    // the energy kernel runs on the right job's stream, after its tiles;
    // the left job's tiles are waited for here
    if (!hip_ok(hipStreamSynchronize(JL.stream), "left tiles")) {
      (void)hipStreamSynchronize(JR.stream);
      release();
      return TCE_T_ERR_HIP;
    }
    const long n = (long)dr.r1 * dr.r2 * dr.r3 * dr.r4 * dr.r5 * dr.r6;
    const LamTile lt{dr.op4, dr.op5, dr.op6, dr.oh1, dr.oh2, dr.oh3,
                     dr.r4,  dr.r5,  dr.r6,  dr.r1,  dr.r2,  dr.r3};
    const int grid = (int)std::max(1L, std::min<long>(LBLOCKS, (n + LB - 1) / LB));
    const double factor = task_factor(T.restricted, p4b, p5b, p6b, h1b, h2b, h3b);
    hipLaunchKernelGGL(lambda_energy, dim3(grid), dim3(LB), 0, JR.stream, R, L, L + n, n, lt,
                       pairing == TCE_T_LAMBDA_REFERENCE_ORDER ? 1 : 0, JR.devl, (int)T.no,
                       (int)T.nv, JL.dT2e, JL.Ep, dF, factor, dpart);
    if (!hip_ok(hipGetLastError(), "lambda_energy launch") ||
        !hip_ok(hipMemcpyAsync(hpart.data(), dpart, 2 * grid * sizeof(double),
                               hipMemcpyDeviceToHost, JR.stream),
                "lambda partials") ||
        !hip_ok(hipStreamSynchronize(JR.stream), "lambda_energy")) {
      (void)hipStreamSynchronize(JR.stream);
      release();
      return TCE_T_ERR_HIP;
    }
    double t1 = 0.0, t2 = 0.0;
    for (int g = 0; g < grid; ++g) {
      t1 += hpart[2 * g];
      t2 += hpart[2 * g + 1];
    }
    if (task_energies) {
      task_energies[2 * t] = t1;
      task_energies[2 * t + 1] = t2;
    }
    s1 += t1;
    s2 += t2;
  }
  release();
  if (energy1) *energy1 = s1;
  if (energy2) *energy2 = s2;
  return TCE_T_OK;
}

// cr_ccsd_t (cr_ccsd_t.F:1-269) from a (T) job and a CR job of the same
// tiles: the task loop's four sums over the t3 tiles of the (T) job (left
// pass) and the moment and denominator tiles of the CR job (right pass),
// their d1/d2 contractions in `precision`, den0 (FP64) from the CR job's t1
// and t2, then energy1 = num1 / (1 + den1 + den0), energy2 = num2 / (1 +
// den2 + den0) (cr_ccsd_t.F:254-258).
extern "C" int tce_t_cr_ccsd_t(tce_t_job *triples, tce_t_job *cr, int precision,
                               double *energy1, double *energy2, double *terms) {
  if (!triples || !cr || triples == cr) {
    set_error("tce_t_cr_ccsd_t: two distinct jobs are needed");
    return TCE_T_ERR_ARG;
  }
  Job &JL = *reinterpret_cast<Job *>(triples);
  Job &JR = *reinterpret_cast<Job *>(cr);
  if (!JR.uploaded || !JL.uploaded) {
    set_error("tce_t_cr_ccsd_t before the uploads");
    return TCE_T_ERR_STATE;
  }
  if (JL.cr || !JR.cr) {
    set_error("tce_t_cr_ccsd_t: a (T) job (tce_t_job_upload) and a CR job (tce_t_job_upload_cr) are needed, in that order");
    return TCE_T_ERR_ARG;
  }
  if (!same_tiles(JR.T, JL.T) || JR.device != JL.device || JR.tiling != JL.tiling) {
    set_error("tce_t_cr_ccsd_t: the jobs' tiles, tilings (or devices) differ");
    return TCE_T_ERR_ARG;
  }
  if (JR.tiling == TCE_T_TILING_ORBITAL) {
    set_error("tce_t_cr_ccsd_t: not available for the orbital tiling");
    return TCE_T_ERR_ARG;
  }
  if (!valid_precision(precision)) {
    set_error("unsupported precision mode");
    return TCE_T_ERR_ARG;
  }
  HIPCHK(hipSetDevice(JR.device), "device");
  const long nt = JR.ntask();
  int rc = ensure_tasks8(JR);
  if (rc != TCE_T_OK) return rc;
  double *den = nullptr, *ov = nullptr;
  HIPCHK(hipMalloc(&den, std::max(1L, JR.nitems8) * 16 * sizeof(double)), "CR partials");
  if (!hip_ok(hipMalloc(&ov, OV_BLOCKS * sizeof(double)), "CR overlap")) {
    (void)hipFree(den);
    return TCE_T_ERR_HIP;
  }
  std::vector<double> te, td, hov(OV_BLOCKS);
  rc = lambda_fused(JR, JL, nullptr, precision, te, den, &td);
  if (rc == TCE_T_OK) {
    hipLaunchKernelGGL(cr_overlap, dim3(OV_BLOCKS), dim3(OV_THREADS), 0, JR.stream, JR.dT1,
                       JR.dT2e, (int)JR.T.no, (int)JR.T.nv, JR.Ep, ov);
    if (!hip_ok(hipGetLastError(), "cr_overlap launch") ||
        !hip_ok(hipMemcpyAsync(hov.data(), ov, OV_BLOCKS * sizeof(double), hipMemcpyDeviceToHost,
                               JR.stream),
                "D2H") ||
        !hip_ok(hipStreamSynchronize(JR.stream), "cr_overlap"))
      rc = TCE_T_ERR_HIP;
  }
  (void)hipFree(den);
  (void)hipFree(ov);
  if (rc != TCE_T_OK) return rc;
  double num1 = 0.0, num2 = 0.0, den1 = 0.0, den2 = 0.0, den0 = 0.0;
  for (long t = 0; t < nt; ++t) {
    num1 += te[2 * t];
    num2 += te[2 * t + 1];
    den1 += td[2 * t];
    den2 += td[2 * t + 1];
  }
  for (double x : hov) den0 += x;
  if (terms) {
    terms[0] = num1;
    terms[1] = num2;
    terms[2] = den0;
    terms[3] = den1;
    terms[4] = den2;
  }
  if (energy1) *energy1 = num1 / (1.0 + den1 + den0);
  if (energy2) *energy2 = num2 / (1.0 + den2 + den0);
  return TCE_T_OK;
}
