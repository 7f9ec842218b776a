// legacy.hip -- the reference's per-call GPU surface, MI355X-native.
//
// Drop-in for the functions of the same names in src/tce/ccsd_t/sd_t_total.cu
// (27 contraction entry points, compute_en_, dev_mem_s_/dev_mem_d_,
// dev_release_), memory.cu (initmemmodule_/finalizememmodule_ and the
// size-keyed device pool behind them) and hybrid.c (check_device_,
// device_init_).  The drivers ccsd_t_gpu.F / ccsd_t_singles_gpu.F /
// ccsd_t_doubles_gpu.F call these unchanged: operand blocks arrive as host
// arrays per call, results accumulate into library-owned device tiles, and
// compute_en_ reduces a task's tiles to {E[T], E(T)}.
//
// Differences in mechanism (not in results): one library stream instead of a
// stream created and destroyed per call; operands staged through a reusable
// pinned buffer; one generic kernel per family (d1, d2, s1) parameterised by
// the output permutation of the variant instead of 25 hand-expanded kernels;
// a two-pass deterministic energy reduction instead of one-thread blocks and a
// host-side sum.  Errors are recorded (tce_t_last_error) and reported on
// stderr rather than calling exit(1).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <set>

#include "internal.h"

namespace tce_t {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

namespace {

// ------------------------------------------------------------ device pool --
// memory.cu keeps freed blocks in per-size free lists until
// finalizememmodule; the same policy here (exact-size reuse).
struct Pool {
  std::map<size_t, std::set<void *>> free_list;
  std::map<void *, size_t> live;
  bool init = false;
  void *get(size_t bytes) {
    auto it = free_list.find(bytes);
    if (it != free_list.end() && !it->second.empty()) {
      void *p = *it->second.begin();
      it->second.erase(it->second.begin());
      if (it->second.empty()) free_list.erase(it);
      live[p] = bytes;
      return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      // morecore (memory.cu:40-52): flush the free lists and retry once
      (void)hipGetLastError();
      clear();
      if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
    }
    live[p] = bytes;
    return p;
  }
  void put(void *p) {
    auto it = live.find(p);
    if (it == live.end()) return;
    free_list[it->second].insert(p);
    live.erase(it);
  }
  void clear() {
    for (auto &kv : free_list)
      for (void *p : kv.second) (void)hipFree(p);
    free_list.clear();
  }
};

struct Legacy {
  std::mutex mu;
  Pool pool;
  hipStream_t stream = nullptr;
  int backend = -1;                  // TCE_T_LEGACY_DIRECT / _TTGT (-1: from the environment)
  rocblas_handle blas = nullptr;     // TTGT: DGEMM on the library stream
  double *t3_d = nullptr, *t3_s = nullptr;  // task tiles (doubles, singles)
  long dims_d[6] = {0}, dims_s[6] = {0};    // h1,h2,h3,p4,p5,p6 of each tile
  double *pinned = nullptr;
  size_t pinned_bytes = 0;
  bool failed = false;
};
Legacy &L() {
  static Legacy s;
  return s;
}

bool ok(hipError_t e, const char *what) {
  if (e == hipSuccess) return true;
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", what, hipGetErrorString(e));
  set_error(buf);
  fprintf(stderr, "tce_t legacy: %s\n", buf);
  L().failed = true;
  return false;
}

bool ensure_stream() {
  if (L().stream) return true;
  return ok(hipStreamCreateWithFlags(&L().stream, hipStreamNonBlocking), "hipStreamCreate");
}

int backend_of(Legacy &S) {
  if (S.backend < 0) {
    const char *e = getenv("TCE_T_LEGACY");
    S.backend = (e && !strcmp(e, "ttgt")) ? TCE_T_LEGACY_TTGT : TCE_T_LEGACY_DIRECT;
  }
  return S.backend;
}

bool ensure_blas(Legacy &S) {
  if (S.blas) return true;
  if (rocblas_create_handle(&S.blas) != rocblas_status_success ||
      rocblas_set_stream(S.blas, S.stream) != rocblas_status_success) {
    set_error("rocBLAS handle for the TTGT back end");
    fprintf(stderr, "tce_t legacy: %s\n", g_err.c_str());
    S.failed = true;
    return false;
  }
  return true;
}

// Row-index codes of the reference kernels' loop variables.
enum { P4 = 0, P5 = 1, P6 = 2, H1 = 3, H2 = 4, H3 = 5 };

// Output layouts (Fortran order, fastest first) and signs of the 27
// variants: the triplesx declarations and update signs of
// ccsd_t_kernels_omp.F:5-1198 (the GPU kernels of sd_t_total.cu write the
// same layouts).
const int kS1L[9][6] = {{H3, H2, H1, P6, P5, P4}, {H3, H1, H2, P6, P5, P4}, {H1, H3, H2, P6, P5, P4},
                        {H3, H2, H1, P6, P4, P5}, {H3, H1, H2, P6, P4, P5}, {H1, H3, H2, P6, P4, P5},
                        {H3, H2, H1, P4, P6, P5}, {H3, H1, H2, P4, P6, P5}, {H1, H3, H2, P4, P6, P5}};
const int kS1S[9] = {+1, -1, +1, -1, +1, -1, +1, -1, +1};
const int kD1L[9][6] = {{H3, H2, H1, P6, P5, P4}, {H3, H1, H2, P6, P5, P4}, {H1, H3, H2, P6, P5, P4},
                        {H3, H2, H1, P5, P4, P6}, {H3, H1, H2, P5, P4, P6}, {H1, H3, H2, P5, P4, P6},
                        {H3, H2, H1, P5, P6, P4}, {H3, H1, H2, P5, P6, P4}, {H1, H3, H2, P5, P6, P4}};
const int kD1S[9] = {-1, +1, -1, -1, +1, -1, +1, -1, +1};
const int kD2L[9][6] = {{H3, H2, H1, P6, P5, P4}, {H2, H1, H3, P6, P5, P4}, {H2, H3, H1, P6, P5, P4},
                        {H3, H2, H1, P6, P4, P5}, {H2, H1, H3, P6, P4, P5}, {H2, H3, H1, P6, P4, P5},
                        {H3, H2, H1, P4, P6, P5}, {H2, H1, H3, P4, P6, P5}, {H2, H3, H1, P4, P6, P5}};
const int kD2S[9] = {-1, -1, +1, +1, +1, -1, -1, -1, +1};

// One launch description: the output tile is walked in its own (layout)
// order so stores are coalesced; every thread owns one t3 element and runs
// the whole contraction for it.
struct CallArgs {
  long dimL[6];   // extents in output order
  int codeL[6];   // row code at each output position
  long total;     // elements of the output tile
  long K;         // contraction length (1 for s1)
  // operand strides by row code (elements), contraction stride = 1 or 0
  long sa[6], sb[6];
  long ka, kb;    // stride of the contraction index in a / b
  double sign;
};

__global__ void __launch_bounds__(256) contract_kernel(CallArgs c, double *__restrict__ t3,
                                                       const double *__restrict__ a,
                                                       const double *__restrict__ b) {
  long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= c.total) return;
  long r = o, oa = 0, ob = 0;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    long v = r % c.dimL[q];
    r /= c.dimL[q];
    oa += v * c.sa[c.codeL[q]];
    ob += v * c.sb[c.codeL[q]];
  }
  double acc = 0.0;
  for (long k = 0; k < c.K; ++k) acc = fma(a[oa + k * c.ka], b[ob + k * c.kb], acc);
  t3[o] += c.sign * acc;
}

// TTGT back end (the alternative of sd_t_total_ttlg.cu: GEMM plus
// transposes): the d1/d2 contraction as one DGEMM into C[m + M n] (rocBLAS,
// FP64 matrix cores), m and n the output indices that sit on the t2 and the
// v2 operand; this kernel adds C into the task tile in the variant's layout
// (the transpose).  Strides of C by row code in c.sa; c.sb, c.ka, c.kb and
// c.K unused.
__global__ void __launch_bounds__(256) permute_add_kernel(CallArgs c, double *__restrict__ t3,
                                                          const double *__restrict__ C) {
  long o = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= c.total) return;
  long r = o, oc = 0;
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    long v = r % c.dimL[q];
    r /= c.dimL[q];
    oc += v * c.sa[c.codeL[q]];
  }
  t3[o] += c.sign * C[oc];
}

// family: 0 = s1, 1 = d1, 2 = d2.  dims by row code; K contraction length.
// host_t3 == nullptr: accumulate into the library's device tile of the task
// (the GPU surface); otherwise into the caller's host triplesx (the CPU
// surface), staged through the device for the call.
void launch_call(int family, int n, const long dims[6], long K, const double *ha, long na,
                 const double *hb, long nb, bool singles, double *host_t3 = nullptr) {
  Legacy &S = L();
  std::lock_guard<std::mutex> lk(S.mu);
  if (!ensure_stream()) return;
  double *t3 = host_t3 ? nullptr : (singles ? S.t3_s : S.t3_d);
  if (!host_t3 && !t3) {
    set_error("sd_t kernel called before dev_mem_s_/dev_mem_d_");
    fprintf(stderr, "tce_t legacy: %s\n", g_err.c_str());
    S.failed = true;
    return;
  }
  const int *Lo = family == 0 ? kS1L[n - 1] : family == 1 ? kD1L[n - 1] : kD2L[n - 1];
  int sign = family == 0 ? kS1S[n - 1] : family == 1 ? kD1S[n - 1] : kD2S[n - 1];
  CallArgs c;
  c.total = 1;
  for (int q = 0; q < 6; ++q) {
    c.codeL[q] = Lo[q];
    c.dimL[q] = dims[Lo[q]];
    c.total *= c.dimL[q];
  }
  // The task tile holds exactly these elements (same extents, permuted).
  const long *td = singles ? S.dims_s : S.dims_d;
  long tsize = td[0] * td[1] * td[2] * td[3] * td[4] * td[5];
  if (!host_t3 && tsize != c.total) {
    set_error("operand dimensions do not match the task tile");
    fprintf(stderr, "tce_t legacy: %s\n", g_err.c_str());
    S.failed = true;
    return;
  }
  c.K = K;
  c.sign = (double)sign;
  for (int q = 0; q < 6; ++q) c.sa[q] = c.sb[q] = 0;
  long p4d = dims[P4], p5d = dims[P5], p6d = dims[P6], h1d = dims[H1], h2d = dims[H2], h3d = dims[H3];
  if (family == 0) {  // t1sub(p4,h1), v2sub(h3,h2,p6,p5)
    c.sa[P4] = 1;
    c.sa[H1] = p4d;
    c.sb[H3] = 1;
    c.sb[H2] = h3d;
    c.sb[P6] = h3d * h2d;
    c.sb[P5] = h3d * h2d * p6d;
    c.ka = c.kb = 0;
  } else if (family == 1) {  // t2sub(h7,p4,p5,h1), v2sub(h3,h2,p6,h7)
    c.ka = 1;
    c.sa[P4] = K;
    c.sa[P5] = K * p4d;
    c.sa[H1] = K * p4d * p5d;
    c.sb[H3] = 1;
    c.sb[H2] = h3d;
    c.sb[P6] = h3d * h2d;
    c.kb = h3d * h2d * p6d;
  } else {  // t2sub(p7,p4,h1,h2), v2sub(p7,h3,p6,p5)
    c.ka = 1;
    c.sa[P4] = K;
    c.sa[H1] = K * p4d;
    c.sa[H2] = K * p4d * h1d;
    c.kb = 1;
    c.sb[H3] = K;
    c.sb[P6] = K * h3d;
    c.sb[P5] = K * h3d * p6d;
  }
  (void)p5d;
  // Stage both operands through one pinned buffer, copy, launch.  The
  // synchronisation before returning keeps the caller's MA stack semantics
  // (operands may be popped right after the call).
  size_t bytes = (size_t)(na + nb) * sizeof(double);
  if (bytes > S.pinned_bytes) {
    if (S.pinned) (void)hipHostFree(S.pinned);
    S.pinned = nullptr;
    S.pinned_bytes = 0;
    if (!ok(hipHostMalloc((void **)&S.pinned, bytes), "hipHostMalloc")) return;
    S.pinned_bytes = bytes;
  }
  memcpy(S.pinned, ha, na * sizeof(double));
  memcpy(S.pinned + na, hb, nb * sizeof(double));
  double *dops = (double *)S.pool.get(bytes);
  if (!dops) {
    set_error("device pool allocation failed");
    S.failed = true;
    return;
  }
  if (!ok(hipMemcpyAsync(dops, S.pinned, bytes, hipMemcpyHostToDevice, S.stream), "H2D")) return;
  const size_t tbytes = (size_t)c.total * sizeof(double);
  if (host_t3) {
    t3 = (double *)S.pool.get(tbytes);
    if (!t3) {
      set_error("device pool allocation failed");
      S.failed = true;
      S.pool.put(dops);
      return;
    }
    // pageable copies: the caller's array is read and written in place
    if (!ok(hipMemcpyAsync(t3, host_t3, tbytes, hipMemcpyHostToDevice, S.stream), "H2D t3")) return;
  }
  int threads = 256;
  long blocks = (c.total + threads - 1) / threads;
  double *dC = nullptr;
  if (family != 0 && backend_of(S) == TCE_T_LEGACY_TTGT && ensure_blas(S)) {
    // m: the indices on t2sub, n: those on v2sub (C[m + M n])
    long M, N;
    CallArgs p = c;
    for (int q = 0; q < 6; ++q) p.sa[q] = 0;
    if (family == 2) {  // t2sub(p7,p4,h1,h2), v2sub(p7,h3,p6,p5)
      M = p4d * h1d * h2d, N = h3d * p6d * p5d;
      p.sa[P4] = 1, p.sa[H1] = p4d, p.sa[H2] = p4d * h1d;
      p.sa[H3] = M, p.sa[P6] = M * h3d, p.sa[P5] = M * h3d * p6d;
    } else {  // t2sub(h7,p4,p5,h1), v2sub(h3,h2,p6,h7)
      M = p4d * p5d * h1d, N = h3d * h2d * p6d;
      p.sa[P4] = 1, p.sa[P5] = p4d, p.sa[H1] = p4d * p5d;
      p.sa[H3] = M, p.sa[H2] = M * h3d, p.sa[P6] = M * h3d * h2d;
    }
    dC = (double *)S.pool.get((size_t)M * N * sizeof(double));
    if (!dC) {
      set_error("device pool allocation failed");
      S.failed = true;
      S.pool.put(dops);
      if (host_t3) S.pool.put(t3);
      return;
    }
    const double one = 1.0, zero = 0.0;
    const rocblas_status st =
        family == 2
            ? rocblas_dgemm(S.blas, rocblas_operation_transpose, rocblas_operation_none, (int)M,
                            (int)N, (int)K, &one, dops, (int)K, dops + na, (int)K, &zero, dC, (int)M)
            : rocblas_dgemm(S.blas, rocblas_operation_transpose, rocblas_operation_transpose,
                            (int)M, (int)N, (int)K, &one, dops, (int)K, dops + na, (int)N, &zero,
                            dC, (int)M);
    if (st != rocblas_status_success) {
      set_error("rocblas_dgemm failed");
      fprintf(stderr, "tce_t legacy: %s\n", g_err.c_str());
      S.failed = true;
      return;
    }
    hipLaunchKernelGGL(permute_add_kernel, dim3((unsigned)blocks), dim3(threads), 0, S.stream, p,
                       t3, dC);
    if (!ok(hipGetLastError(), "permute_add_kernel launch")) return;
  } else {
    hipLaunchKernelGGL(contract_kernel, dim3((unsigned)blocks), dim3(threads), 0, S.stream, c, t3,
                       dops, dops + na);
    if (!ok(hipGetLastError(), "contract_kernel launch")) return;
  }
  if (host_t3 && !ok(hipMemcpyAsync(host_t3, t3, tbytes, hipMemcpyDeviceToHost, S.stream), "D2H t3"))
    return;
  ok(hipStreamSynchronize(S.stream), "contract_kernel");
  S.pool.put(dops);
  if (dC) S.pool.put(dC);
  if (host_t3) S.pool.put(t3);
}

// Energy of one task tile, layout (h3 fastest, h2, h1, p6, p5, p4 slowest)
// as in ccsd_t_dot.F:89-110; first pass: one partial pair per block.  The
// orbital energies of the six tiles are e1..e6 (h1, h2, h3, p4, p5, p6).
__device__ __forceinline__ void energy_block(long h1d, long h2d, long h3d, long p4d, long p5d,
                                             long p6d, const double *__restrict__ e1,
                                             const double *__restrict__ e2,
                                             const double *__restrict__ e3,
                                             const double *__restrict__ e4,
                                             const double *__restrict__ e5,
                                             const double *__restrict__ e6, double factor,
                                             const double *__restrict__ t3d,
                                             const double *__restrict__ t3s,
                                             double *__restrict__ part) {
  long total = h1d * h2d * h3d * p4d * p5d * p6d;
  double s1 = 0.0, s2 = 0.0;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < total;
       i += (long)gridDim.x * blockDim.x) {
    long r = i;
    long h3 = r % h3d; r /= h3d;
    long h2 = r % h2d; r /= h2d;
    long h1 = r % h1d; r /= h1d;
    long p6 = r % p6d; r /= p6d;
    long p5 = r % p5d; r /= p5d;
    long p4 = r;
    double doub = t3d[i], sing = t3s[i];
    double denom_0 = -(e4[p4] + e5[p5] + e6[p6]);
    double denom = doub * factor / (e1[h1] + e2[h2] + e3[h3] + denom_0);
    s1 += denom * doub;
    s2 += denom * (doub + sing);
  }
  __shared__ double r1[256], r2[256];
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      r1[threadIdx.x] += r1[threadIdx.x + w];
      r2[threadIdx.x] += r2[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = r1[0];
    part[2 * blockIdx.x + 1] = r2[0];
  }
}

// compute_en_: the six energy vectors packed in one device buffer.
__global__ void __launch_bounds__(256) energy_partial(
    long h1d, long h2d, long h3d, long p4d, long p5d, long p6d, const double *__restrict__ ev,
    double factor, const double *__restrict__ t3d, const double *__restrict__ t3s,
    double *__restrict__ part) {
  const double *e1 = ev, *e2 = e1 + h1d, *e3 = e2 + h2d, *e4 = e3 + h3d, *e5 = e4 + p4d, *e6 = e5 + p5d;
  energy_block(h1d, h2d, h3d, p4d, p5d, p6d, e1, e2, e3, e4, e5, e6, factor, t3d, t3s, part);
}

// tce_t_compute_en_noalloc: six separate device vectors.
__global__ void __launch_bounds__(256) energy_partial_split(
    long h1d, long h2d, long h3d, long p4d, long p5d, long p6d, const double *__restrict__ e1,
    const double *__restrict__ e2, const double *__restrict__ e3, const double *__restrict__ e4,
    const double *__restrict__ e5, const double *__restrict__ e6, double factor,
    const double *__restrict__ t3d, const double *__restrict__ t3s, double *__restrict__ part) {
  energy_block(h1d, h2d, h3d, p4d, p5d, p6d, e1, e2, e3, e4, e5, e6, factor, t3d, t3s, part);
}

// Grid of the energy reduction: compute_en_ and tce_t_compute_en_noalloc
// both run energy_partial on this many blocks, so their sums are the same.
constexpr int kEnBlocks = 1024;
constexpr int kEnThreads = 256;

// The block partials of tce_t_compute_en_noalloc: 2 doubles per block,
// 16 KiB, one copy per device for the life of the process.  Every call
// writes all of them before reading them, so nothing needs clearing between
// calls.  Host code never passes this name to a kernel (in host code it is
// a stand-in, not a device address); it uses the address
// hipGetSymbolAddress gives on each device.
constexpr size_t kScratchDoubles = 2 * kEnBlocks;
static __device__ __align__(256) double scratch[kScratchDoubles];

// Second pass: fixed-order sum of the block partials.
__global__ void energy_final(const double *__restrict__ part, int nblk, double *__restrict__ out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double a = 0.0, b = 0.0;
    for (int i = 0; i < nblk; ++i) {
      a += part[2 * i];
      b += part[2 * i + 1];
    }
    out[0] = a;
    out[1] = b;
  }
}

void alloc_tile(bool singles, const long d[6]) {
  Legacy &S = L();
  std::lock_guard<std::mutex> lk(S.mu);
  if (!ensure_stream()) return;
  size_t n = (size_t)d[0] * d[1] * d[2] * d[3] * d[4] * d[5];
  double *&t = singles ? S.t3_s : S.t3_d;
  long *dims = singles ? S.dims_s : S.dims_d;
  if (t) S.pool.put(t);
  t = (double *)S.pool.get(n * sizeof(double));
  if (!t) {
    set_error("device tile allocation failed");
    fprintf(stderr, "tce_t legacy: %s\n", g_err.c_str());
    S.failed = true;
    return;
  }
  for (int q = 0; q < 6; ++q) dims[q] = d[q];
  ok(hipMemsetAsync(t, 0, n * sizeof(double), S.stream), "hipMemset t3");
}

int local_rank() {
  const char *vars[] = {"OMPI_COMM_WORLD_LOCAL_RANK", "MPI_LOCALRANKID", "SLURM_LOCALID",
                        "LOCAL_RANK"};
  for (const char *v : vars) {
    const char *s = getenv(v);
    if (s && *s) return atoi(s);
  }
  return 0;
}

}  // namespace
}  // namespace tce_t

using namespace tce_t;

extern "C" const char *tce_t_last_error(void) { return g_err.c_str(); }

// ------------------------------------------------------ entry points -------
#define S1_ENTRY(N)                                                                       \
  extern "C" void sd_t_s1_##N##_cuda_(long *h1d, long *h2d, long *h3d, long *p4d,        \
                                      long *p5d, long *p6d, double *t3, double *t1,      \
                                      double *v2) {                                      \
    (void)t3;                                                                             \
    long dims[6] = {*p4d, *p5d, *p6d, *h1d, *h2d, *h3d};                                  \
    launch_call(0, N, dims, 1, t1, (*p4d) * (*h1d), v2, (*h3d) * (*h2d) * (*p6d) * (*p5d), \
                true);                                                                    \
  }
#define D1_ENTRY(N)                                                                       \
  extern "C" void sd_t_d1_##N##_cuda_(long *h1d, long *h2d, long *h3d, long *h7d,        \
                                      long *p4d, long *p5d, long *p6d, double *t3,       \
                                      double *t2, double *v2) {                          \
    (void)t3;                                                                             \
    long dims[6] = {*p4d, *p5d, *p6d, *h1d, *h2d, *h3d};                                  \
    launch_call(1, N, dims, *h7d, t2, (*h7d) * (*p4d) * (*p5d) * (*h1d), v2,              \
                (*h3d) * (*h2d) * (*p6d) * (*h7d), false);                                \
  }
#define D2_ENTRY(N)                                                                       \
  extern "C" void sd_t_d2_##N##_cuda_(long *h1d, long *h2d, long *h3d, long *p4d,        \
                                      long *p5d, long *p6d, long *p7d, double *t3,       \
                                      double *t2, double *v2) {                          \
    (void)t3;                                                                             \
    long dims[6] = {*p4d, *p5d, *p6d, *h1d, *h2d, *h3d};                                  \
    launch_call(2, N, dims, *p7d, t2, (*p7d) * (*p4d) * (*h1d) * (*h2d), v2,              \
                (*p7d) * (*h3d) * (*p6d) * (*p5d), false);                                \
  }
#define ALL9(M) M(1) M(2) M(3) M(4) M(5) M(6) M(7) M(8) M(9)
ALL9(S1_ENTRY)
ALL9(D1_ENTRY)
ALL9(D2_ENTRY)

// The CPU surface (ccsd_t_kernels_omp.F argument order: h3d, h2d, h1d, p6d,
// p5d, p4d[, h7d / p7d]): the same contractions accumulated into the
// caller's host triplesx (not zeroed), for the CPU drivers
// (ccsd_t_singles_l.F, offl_ccsd_t_doubles_l.F) and their other callers.
#define S1_CPU(N)                                                                          \
  extern "C" void sd_t_s1_##N##_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d,    \
                                 long *p4d, double *triplesx, double *t1sub, double *v2sub) { \
    long dims[6] = {*p4d, *p5d, *p6d, *h1d, *h2d, *h3d};                                   \
    launch_call(0, N, dims, 1, t1sub, (*p4d) * (*h1d), v2sub,                              \
                (*h3d) * (*h2d) * (*p6d) * (*p5d), true, triplesx);                        \
  }
#define D1_CPU(N)                                                                          \
  extern "C" void sd_t_d1_##N##_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d,    \
                                 long *p4d, long *h7d, double *triplesx, double *t2sub,   \
                                 double *v2sub) {                                          \
    long dims[6] = {*p4d, *p5d, *p6d, *h1d, *h2d, *h3d};                                   \
    launch_call(1, N, dims, *h7d, t2sub, (*h7d) * (*p4d) * (*p5d) * (*h1d), v2sub,         \
                (*h3d) * (*h2d) * (*p6d) * (*h7d), false, triplesx);                       \
  }
#define D2_CPU(N)                                                                          \
  extern "C" void sd_t_d2_##N##_(long *h3d, long *h2d, long *h1d, long *p6d, long *p5d,    \
                                 long *p4d, long *p7d, double *triplesx, double *t2sub,   \
                                 double *v2sub) {                                          \
    long dims[6] = {*p4d, *p5d, *p6d, *h1d, *h2d, *h3d};                                   \
    launch_call(2, N, dims, *p7d, t2sub, (*p7d) * (*p4d) * (*h1d) * (*h2d), v2sub,         \
                (*p7d) * (*h3d) * (*p6d) * (*p5d), false, triplesx);                       \
  }
ALL9(S1_CPU)
ALL9(D1_CPU)
ALL9(D2_CPU)

extern "C" void dev_mem_d_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d) {
  long d[6] = {*h1d, *h2d, *h3d, *p4d, *p5d, *p6d};
  alloc_tile(false, d);
}
extern "C" void dev_mem_s_(long *h1d, long *h2d, long *h3d, long *p4d, long *p5d, long *p6d) {
  long d[6] = {*h1d, *h2d, *h3d, *p4d, *p5d, *p6d};
  alloc_tile(true, d);
}
extern "C" void dev_release_(void) {
  Legacy &S = L();
  std::lock_guard<std::mutex> lk(S.mu);
  if (S.stream) (void)hipStreamSynchronize(S.stream);
  if (S.t3_d) S.pool.put(S.t3_d);
  if (S.t3_s) S.pool.put(S.t3_s);
  S.t3_d = S.t3_s = nullptr;
}
extern "C" void initmemmodule_(void) {
  Legacy &S = L();
  std::lock_guard<std::mutex> lk(S.mu);
  S.pool.init = true;
}
extern "C" void finalizememmodule_(void) {
  Legacy &S = L();
  std::lock_guard<std::mutex> lk(S.mu);
  if (S.stream) (void)hipStreamSynchronize(S.stream);
  S.pool.clear();
  S.pool.init = false;
}

extern "C" void compute_en_(double *factor, double *energy, double *eval1, double *eval2,
                            double *eval3, double *eval4, double *eval5, double *eval6, long *h1d,
                            long *h2d, long *h3d, long *p4d, long *p5d, long *p6d, double *host1,
                            double *host2) {
  (void)host1;
  (void)host2;
  Legacy &S = L();
  std::lock_guard<std::mutex> lk(S.mu);
  energy[0] = energy[1] = 0.0;
  if (!ensure_stream()) return;
  if (!S.t3_d || !S.t3_s) {
    set_error("compute_en_ called without device tiles");
    fprintf(stderr, "tce_t legacy: %s\n", g_err.c_str());
    S.failed = true;
    return;
  }
  long d[6] = {*h1d, *h2d, *h3d, *p4d, *p5d, *p6d};
  long ne = d[0] + d[1] + d[2] + d[3] + d[4] + d[5];
  const int nblk = kEnBlocks, threads = kEnThreads;
  size_t bytes = (ne + 2 * nblk + 2) * sizeof(double);
  double *dbuf = (double *)S.pool.get(bytes);
  if (!dbuf) {
    set_error("device pool allocation failed");
    S.failed = true;
    return;
  }
  // Orbital energies of the six tiles, packed h1,h2,h3,p4,p5,p6.
  if (S.pinned_bytes < (size_t)(ne + 2) * sizeof(double)) {
    if (S.pinned) (void)hipHostFree(S.pinned);
    S.pinned = nullptr;
    S.pinned_bytes = 0;
    if (!ok(hipHostMalloc((void **)&S.pinned, (ne + 2) * sizeof(double)), "hipHostMalloc")) return;
    S.pinned_bytes = (ne + 2) * sizeof(double);
  }
  double *src[6] = {eval1, eval2, eval3, eval4, eval5, eval6};
  long at = 0;
  for (int q = 0; q < 6; ++q) {
    memcpy(S.pinned + at, src[q], d[q] * sizeof(double));
    at += d[q];
  }
  if (!ok(hipMemcpyAsync(dbuf, S.pinned, ne * sizeof(double), hipMemcpyHostToDevice, S.stream), "H2D"))
    return;
  double *part = dbuf + ne, *out = part + 2 * nblk;
  hipLaunchKernelGGL(energy_partial, dim3(nblk), dim3(threads), 0, S.stream, d[0], d[1], d[2], d[3],
                     d[4], d[5], dbuf, *factor, S.t3_d, S.t3_s, part);
  hipLaunchKernelGGL(energy_final, dim3(1), dim3(64), 0, S.stream, part, nblk, out);
  if (!ok(hipGetLastError(), "energy launch")) return;
  if (!ok(hipMemcpyAsync(S.pinned, out, 2 * sizeof(double), hipMemcpyDeviceToHost, S.stream), "D2H"))
    return;
  if (!ok(hipStreamSynchronize(S.stream), "energy")) return;
  energy[0] = S.pinned[0];
  energy[1] = S.pinned[1];
  S.pool.put(dbuf);
}

// ------------------------------------------------ no-allocation energy -----
namespace tce_t {
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {nullptr};  // device address of `scratch` per device

// Address of this device's `scratch` (looked up once per device, with the
// device current); nullptr on failure.
double *scratch_on(int dev) {
  if (dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch[dev] = (double *)p;
  }
  return g_scratch[dev];
}
}  // namespace
}  // namespace tce_t

extern "C" int tce_t_compute_en_noalloc(double factor, const double *t3d, const double *t3s,
                                        const double *eval1, const double *eval2,
                                        const double *eval3, const double *eval4,
                                        const double *eval5, const double *eval6, long h1d,
                                        long h2d, long h3d, long p4d, long p5d, long p6d,
                                        double *energy, void *stream) {
  if (!t3d || !t3s || !eval1 || !eval2 || !eval3 || !eval4 || !eval5 || !eval6 || !energy) {
    set_error("tce_t_compute_en_noalloc: NULL pointer");
    return TCE_T_ERR_ARG;
  }
  const long d[6] = {h1d, h2d, h3d, p4d, p5d, p6d};
  long total = 1;
  for (long x : d) {
    if (x <= 0 || total > (1L << 62) / x) {
      set_error("tce_t_compute_en_noalloc: tile extents must be positive and their product "
                "below 2^62");
      return TCE_T_ERR_ARG;
    }
    total *= x;
  }
  // the partials of kEnBlocks blocks fill the fixed scratch array exactly
  static_assert(2 * kEnBlocks <= (int)kScratchDoubles, "scratch holds every block partial");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) {
    (void)hipGetLastError();
    set_error("tce_t_compute_en_noalloc: hipGetDevice failed");
    return TCE_T_ERR_HIP;
  }
  double *part = scratch_on(dev);
  if (!part) {
    set_error("tce_t_compute_en_noalloc: hipGetSymbolAddress(scratch) failed");
    return TCE_T_ERR_HIP;
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(energy_partial_split, dim3(kEnBlocks), dim3(kEnThreads), 0, st, h1d, h2d, h3d,
                     p4d, p5d, p6d, eval1, eval2, eval3, eval4, eval5, eval6, factor, t3d, t3s,
                     part);
  hipLaunchKernelGGL(energy_final, dim3(1), dim3(64), 0, st, part, kEnBlocks, energy);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string("tce_t_compute_en_noalloc launch: ") + hipGetErrorString(e));
    return TCE_T_ERR_HIP;
  }
  return TCE_T_OK;
}

extern "C" int check_device_(long *icuda) { return local_rank() < *icuda ? 1 : 0; }

extern "C" int device_init_(long *icuda, long *cuda_device_number) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  if (n < *icuda) {
    printf("Warning: Please check whether you have %ld cuda devices per node\n", *icuda);
    fflush(stdout);
    *cuda_device_number = 30;
  } else {
    (void)hipSetDevice(local_rank());
  }
  return 1;
}

extern "C" int tce_t_legacy_backend(int backend) {
  Legacy &S = L();
  std::lock_guard<std::mutex> lk(S.mu);
  if (backend == TCE_T_LEGACY_QUERY) return backend_of(S);
  if (backend != TCE_T_LEGACY_DIRECT && backend != TCE_T_LEGACY_TTGT) return TCE_T_ERR_ARG;
  S.backend = backend;
  return backend;
}

// Sticky failure flag of the per-call surface (the reference exits instead).
extern "C" int tce_t_legacy_status(void) {
  Legacy &S = L();
  int f = S.failed ? TCE_T_ERR_HIP : TCE_T_OK;
  S.failed = false;
  return f;
}
