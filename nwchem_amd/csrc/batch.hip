// batch.hip -- many independent (T) jobs in one persistent launch.
//
// A batch holds uploaded jobs (owned by the caller).  Running it in a
// precision launches one fused kernel over a segment table -- one segment
// per job, all jobs' work items in one global range -- followed by each
// job's fixed-order task reduction and the copy of its task energies to
// pinned host memory.  The first run in a precision plans the launch:
//
//   autotuner  candidates are the kernel family for FP64 (fused_triples_wg
//              with 8-wide particle boxes on 4x4x4 MFMA, fused_triples_mixed
//              on 16x16x4, fused_triples with 4-wide boxes) and
//              the segment order (as given; costliest items first).  For
//              fused_triples_mixed the slice width of its work queues (the
//              consecutive items an XCD takes at a time: 32, 64 or 128) is
//              then tuned on the best of those.  Each candidate runs once
//              on the same sample (the first eighth of every job's items,
//              or all of them for small batches), timed with HIP events; the
//              fastest is kept and remembered per batch shape (the sorted
//              list of job dimensions and tilings) for later batches in the
//              process.
//   graph      the chosen launch sequence is captured once into a hipGraph
//              (kernel, reductions, device-to-host copies; no memset nodes)
//              and replayed by every later run.
//
// Each job's energies are bit-identical to its own tce_t_job_run in the
// same precision and kernel family: work items do not interact and every
// job's partials are reduced in its own fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "job.h"

namespace tce_t {
namespace {

enum Family { FAM_WAVE = 0, FAM_WG = 1, FAM_MIXED = 2, FAM_WG16 = 3 };
constexpr int NCAND = 6;  // at most 3 families x 2 orders

struct Choice {
  int family = FAM_WG;
  bool cost_order = true;
  int shift = 5;  // log2 of the work-queue slice width (fused_triples_mixed)
};
constexpr int NSLICE = 3;  // slice widths tried: 2^5, 2^6, 2^7

std::mutex g_tune_mu;
std::map<std::string, Choice> g_tuned;  // batch shape -> choice

struct Plan {
  bool ready = false;
  Choice choice;
  void *tab = nullptr;       // device: job arguments + segments
  long total = 0;            // work items of the launch
  int nseg = 0, grid = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double tuned_ms[NCAND] = {-1, -1, -1, -1, -1, -1};  // candidate times (family x order)
  double slice_ms[NSLICE] = {-1, -1, -1};             // slice widths 32, 64, 128
};

struct Batch {
  std::vector<Job *> jobs;
  hipStream_t stream = nullptr;
  unsigned long long *sched = nullptr;  // dynamic-schedule counters of the stream
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double *hE = nullptr;            // pinned: task energies of all jobs
  std::vector<long> eoff;          // per job: first task slot in hE
  std::vector<long> gens;          // per job: upload generation the plans were made for
  long ntasks = 0;
  Plan plan[NPREC];
  float last_ms = 0.f;
  int cus = 1;
};

std::string shape_key(const Batch &B, int precision) {
  std::vector<std::string> parts;
  for (const Job *J : B.jobs) {
    std::string s = std::to_string(J->T.no) + "/" + std::to_string(J->T.nv) + ":";
    for (long t = 1; t <= J->T.noab + J->T.nvab; ++t) s += std::to_string(J->T.rng(t)) + ",";
    parts.push_back(s);
  }
  std::sort(parts.begin(), parts.end());
  std::string key = std::to_string(precision) + "|";
  for (auto &p : parts) key += p + ";";
  return key;
}

// Segment order of a choice: job indices, costliest items first or as given.
std::vector<int> job_order(const Batch &B, int precision, bool cost_order) {
  std::vector<int> ord(B.jobs.size());
  std::iota(ord.begin(), ord.end(), 0);
  if (cost_order)
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
      return item_cost(*B.jobs[a], precision) > item_cost(*B.jobs[b], precision);
    });
  return ord;
}

// Build the device table (job arguments, then segments) of a choice; frac < 1
// keeps only the first ceil(frac * items) work items of every job (the
// tuning sample: the same items under every candidate).
int build_table(Batch &B, int precision, const Choice &c, Plan &P, double frac = 1.0) {
  // the 16x16 kernel (reduced precision, or FP64 with family WG16) takes MixedArgs
  const bool mixed = precision != TCE_T_FP64 || c.family == FAM_WG16;
  const bool wide = mixed || c.family == FAM_WG;
  const std::vector<int> ord = job_order(B, precision, c.cost_order);
  const size_t asz = ((mixed ? sizeof(MixedArgs) : sizeof(FusedArgs)) + 15) / 16 * 16;
  const size_t ssz = mixed ? sizeof(MixedSegment) : sizeof(Segment);
  const size_t n = B.jobs.size();
  std::vector<char> h(n * asz + n * ssz, 0);
  if (P.tab) (void)hipFree(P.tab);
  P.tab = nullptr;
  HIPCHK(hipMalloc(&P.tab, h.size()), "batch table");
  char *d = (char *)P.tab;
  long g = 0;
  int nseg = 0;
  for (size_t k = 0; k < n; ++k) {
    const Job &J = *B.jobs[ord[k]];
    const long all = wide ? J.nitems8 : J.nitems;
    if (all == 0) continue;
    const long items = frac >= 1.0 ? all : std::max(1L, (long)std::ceil(frac * (double)all));
    char *ha = h.data() + (size_t)nseg * asz;
    if (mixed) {
      MixedArgs M = job_margs(J, precision);
      std::memcpy(ha, &M, sizeof M);
      MixedSegment s{(const MixedArgs *)(d + (size_t)nseg * asz), 0, items, g};
      std::memcpy(h.data() + n * asz + (size_t)nseg * ssz, &s, sizeof s);
    } else {
      FusedArgs A = job_args(J, wide);
      std::memcpy(ha, &A, sizeof A);
      Segment s{(const FusedArgs *)(d + (size_t)nseg * asz), 0, items, g};
      std::memcpy(h.data() + n * asz + (size_t)nseg * ssz, &s, sizeof s);
    }
    g += items;
    ++nseg;
  }
  HIPCHK(hipMemcpy(P.tab, h.data(), h.size(), hipMemcpyHostToDevice), "batch table");
  P.total = g;
  P.nseg = nseg;
  const int per_cu = (!mixed && c.family == FAM_WAVE) ? 8 : 1;
  P.grid = (int)std::max<long>(1, std::min<long>((long)B.cus * per_cu, g));
  P.choice = c;
  return TCE_T_OK;
}

// Enqueue the launch sequence of a plan on the batch stream.
int enqueue(Batch &B, int precision, const Plan &P) {
  const bool mixed = precision != TCE_T_FP64 || P.choice.family == FAM_WG16;
  const bool wide = mixed || P.choice.family == FAM_WG;
  const size_t n = B.jobs.size();
  const size_t asz = ((mixed ? sizeof(MixedArgs) : sizeof(FusedArgs)) + 15) / 16 * 16;
  const char *segs = (const char *)P.tab + n * asz;
  if (P.total > 0) {
    int rc;
    if (mixed) {
      bool sparse = false;
      for (const Job *j : B.jobs) sparse = sparse || j->sparse;
      rc = launch_fused_mixed((const MixedSegment *)segs, P.nseg, P.total, precision, P.grid,
                              B.stream, sparse, sched_use(B.sched));
    }
    else if (P.choice.family == FAM_WG)
      rc = launch_fused_wg((const Segment *)segs, P.nseg, P.total, P.grid, B.stream);
    else
      rc = launch_fused_wave((const Segment *)segs, P.nseg, P.total, P.grid, B.stream);
    if (rc != 0) {
      set_error("batch kernel launch failed");
      return TCE_T_ERR_HIP;
    }
  }
  for (size_t j = 0; j < n; ++j) {
    const Job &J = *B.jobs[j];
    const long nt = J.ngroups();
    if (nt == 0) continue;
    const FusedArgs A = job_args(J, wide);
    if (launch_task_reduce(job_bounds(J, wide), (int)nt, wide ? 8 : J.ppi, A.partials, J.dtaskE,
                           B.stream) != 0) {
      set_error("batch task_reduce launch failed");
      return TCE_T_ERR_HIP;
    }
    HIPCHK(hipMemcpyAsync(B.hE + 2 * B.eoff[j], J.dtaskE, 2 * nt * sizeof(double),
                          hipMemcpyDeviceToHost, B.stream),
           "batch D2H");
  }
  return TCE_T_OK;
}

// Slice width of the batch stream's work queues (read by every later
// launch of fused_triples_mixed on it, graph replays included).
int set_slice(Batch &B, int shift) {
  const unsigned long long v = (unsigned long long)shift;
  HIPCHK(hipMemcpy(B.sched + 9, &v, sizeof v, hipMemcpyHostToDevice), "slice width");
  return TCE_T_OK;
}

// Time one eager run of a plan (kernel and reductions).
int time_plan(Batch &B, int precision, const Plan &P, float &ms) {
  HIPCHK(hipEventRecord(B.ev0, B.stream), "event");
  const int rc = enqueue(B, precision, P);
  if (rc != TCE_T_OK) return rc;
  HIPCHK(hipEventRecord(B.ev1, B.stream), "event");
  HIPCHK(hipStreamSynchronize(B.stream), "batch run");
  HIPCHK(hipEventElapsedTime(&ms, B.ev0, B.ev1), "elapsed");
  return TCE_T_OK;
}

int plan_batch(Batch &B, int precision) {
  Plan &P = B.plan[precision];
  const bool mixed = precision != TCE_T_FP64;
  for (Job *J : B.jobs) {
    int rc = mixed ? prepare_mixed(*J, precision) : ensure_tasks8(*J);
    if (rc != TCE_T_OK) return rc;
  }
  const std::string key = shape_key(B, precision);
  Choice best;
  bool known = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tuned.find(key);
    if (it != g_tuned.end()) best = it->second, known = true;
  }
  if (!known) {
    // 4-wide boxes need every job's 4-wide task table (jobs whose own kernel
    // is fused_triples)
    bool all4 = true;
    for (const Job *J : B.jobs) all4 = all4 && J->pbox == 4;
    std::vector<int> fams = mixed ? std::vector<int>{FAM_MIXED}
                                  : (all4 ? std::vector<int>{FAM_WG, FAM_WG16, FAM_WAVE}
                                          : std::vector<int>{FAM_WG, FAM_WG16});
    std::vector<Choice> cands;
    for (int fam : fams)
      for (bool co : {true, false}) cands.push_back(Choice{fam, co});
    // Candidates run on the first eighth of every job's items when that
    // still gives the GPU 8 items per CU (the same sample under every
    // candidate); a full run per candidate cost 6-7x the batch itself.
    // TCE_T_BATCH_TUNE_FRACTION sets the fraction (0 < f <= 1).
    long items8 = 0;
    for (const Job *J : B.jobs) items8 += J->nitems8;
    double frac = items8 >= 64L * B.cus ? 0.125 : 1.0;
    if (const char *f = getenv("TCE_T_BATCH_TUNE_FRACTION")) {
      const double x = atof(f);
      if (x > 0.0 && x <= 1.0) frac = x;
    }
    float best_ms = 0.f;
    int rc = set_slice(B, 5);
    if (rc != TCE_T_OK) return rc;
    for (size_t i = 0; i < cands.size(); ++i) {
      rc = build_table(B, precision, cands[i], P, frac);
      if (rc != TCE_T_OK) return rc;
      float ms = 0.f;
      rc = time_plan(B, precision, P, ms);
      if (rc != TCE_T_OK) return rc;
      P.tuned_ms[i] = ms;
      if (i == 0 || ms < best_ms) best_ms = ms, best = cands[i];
    }
    // the work-queue slice width of the 16x16 kernel, on the best plan
    if (mixed || best.family == FAM_WG16) {
      P.slice_ms[0] = best_ms;
      rc = build_table(B, precision, best, P, frac);
      if (rc != TCE_T_OK) return rc;
      const Choice base = best;
      for (int k = 1; k < NSLICE; ++k) {
        Choice c = base;
        c.shift = 5 + k;
        if ((rc = set_slice(B, c.shift)) != TCE_T_OK) return rc;
        float ms = 0.f;
        if ((rc = time_plan(B, precision, P, ms)) != TCE_T_OK) return rc;
        P.slice_ms[k] = ms;
        if (ms < best_ms) best_ms = ms, best = c;
      }
    }
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tuned[key] = best;
  }
  int rc = set_slice(B, best.shift);
  if (rc != TCE_T_OK) return rc;
  rc = build_table(B, precision, best, P);
  if (rc != TCE_T_OK) return rc;
  HIPCHK(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal), "capture");
  rc = enqueue(B, precision, P);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(B.stream, &g);
  if (rc != TCE_T_OK) return rc;
  HIPCHK(e, "end capture");
  P.graph = g;
  HIPCHK(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0), "instantiate");
  P.ready = true;
  return TCE_T_OK;
}

}  // namespace
}  // namespace tce_t

using namespace tce_t;

extern "C" int tce_t_batch_create(tce_t_job *const *jobs, int njobs, tce_t_batch **out) {
  if (!out || njobs < 1 || !jobs) return TCE_T_ERR_ARG;
  auto B = new Batch;
  int dev = 0;
  for (int j = 0; j < njobs; ++j) {
    Job *J = reinterpret_cast<Job *>(jobs[j]);
    if (!J || !J->uploaded) {
      delete B;
      set_error("tce_t_batch_create: every job must be uploaded");
      return TCE_T_ERR_STATE;
    }
    B->jobs.push_back(J);
    B->gens.push_back(J->gen);
    B->eoff.push_back(B->ntasks);
    B->ntasks += J->ngroups();
    dev = J->device;
  }
  auto fail = [&](hipError_t e, const char *what) {
    hip_ok(e, what);
    if (B->stream) (void)hipStreamDestroy(B->stream);
    delete B;
    return TCE_T_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking)) != hipSuccess)
    return fail(e, "batch stream");
  if (!(B->sched = sched_alloc())) {
    (void)hipStreamDestroy(B->stream);
    delete B;
    return TCE_T_ERR_HIP;
  }
  if ((e = hipEventCreate(&B->ev0)) != hipSuccess) return fail(e, "event");
  if ((e = hipEventCreate(&B->ev1)) != hipSuccess) return fail(e, "event");
  if ((e = hipHostMalloc(&B->hE, std::max<long>(1, 2 * B->ntasks) * sizeof(double))) != hipSuccess)
    return fail(e, "batch energies");
  if ((e = hipDeviceGetAttribute(&B->cus, hipDeviceAttributeMultiprocessorCount, dev)) != hipSuccess)
    return fail(e, "attr");
  *out = reinterpret_cast<tce_t_batch *>(B);
  return TCE_T_OK;
}

extern "C" int tce_t_batch_run(tce_t_batch *batch, int precision, double *energies) {
  if (!batch) return TCE_T_ERR_ARG;
  if (!valid_precision(precision)) {
    set_error("unsupported precision mode");
    return TCE_T_ERR_ARG;
  }
  Batch &B = *reinterpret_cast<Batch *>(batch);
  // the plans hold the jobs' device pointers: a job uploaded again (or
  // destroyed) since the batch was made would leave them dangling
  for (size_t j = 0; j < B.jobs.size(); ++j)
    if (!B.jobs[j]->uploaded || B.jobs[j]->gen != B.gens[j]) {
      set_error("tce_t_batch_run: a job was uploaded again or released after tce_t_batch_create");
      return TCE_T_ERR_STATE;
    }
  Plan &P = B.plan[precision];
  if (!P.ready) {
    const int rc = plan_batch(B, precision);
    if (rc != TCE_T_OK) return rc;
  }
  HIPCHK(hipEventRecord(B.ev0, B.stream), "event");
  HIPCHK(hipGraphLaunch(P.exec, B.stream), "graph launch");
  HIPCHK(hipEventRecord(B.ev1, B.stream), "event");
  HIPCHK(hipStreamSynchronize(B.stream), "batch run");
  HIPCHK(hipEventElapsedTime(&B.last_ms, B.ev0, B.ev1), "elapsed");
  for (size_t j = 0; j < B.jobs.size(); ++j) {
    Job &J = *B.jobs[j];
    const long nt = J.ngroups();
    J.taskE.assign(B.hE + 2 * B.eoff[j], B.hE + 2 * (B.eoff[j] + nt));
    double s4 = 0.0, s5 = 0.0;
    for (long t = 0; t < nt; ++t) {
      s4 += J.taskE[2 * t];
      s5 += J.taskE[2 * t + 1];
    }
    if (energies) {
      energies[2 * j] = s4;
      energies[2 * j + 1] = s5;
    }
  }
  return TCE_T_OK;
}

extern "C" double tce_t_batch_last_ms(const tce_t_batch *batch) {
  return batch ? (double)reinterpret_cast<const Batch *>(batch)->last_ms : -1.0;
}

extern "C" int tce_t_batch_plan(const tce_t_batch *batch, int precision, int *family,
                                int *cost_order, double *tuned_ms) {
  if (!batch || !valid_precision(precision)) return TCE_T_ERR_ARG;
  const Plan &P = reinterpret_cast<const Batch *>(batch)->plan[precision];
  if (!P.ready) return TCE_T_ERR_STATE;
  if (family) *family = P.choice.family;
  if (cost_order) *cost_order = P.choice.cost_order ? 1 : 0;
  if (tuned_ms)
    for (int i = 0; i < NCAND; ++i) tuned_ms[i] = P.tuned_ms[i];
  return TCE_T_OK;
}

extern "C" int tce_t_batch_plan_schedule(const tce_t_batch *batch, int precision, int *slice,
                                         double *slice_ms) {
  if (!batch || !valid_precision(precision)) return TCE_T_ERR_ARG;
  const Plan &P = reinterpret_cast<const Batch *>(batch)->plan[precision];
  if (!P.ready) return TCE_T_ERR_STATE;
  if (slice) *slice = 1 << P.choice.shift;
  if (slice_ms)
    for (int i = 0; i < NSLICE; ++i) slice_ms[i] = P.slice_ms[i];
  return TCE_T_OK;
}

extern "C" int tce_t_batch_destroy(tce_t_batch *batch) {
  if (!batch) return TCE_T_OK;
  Batch *B = reinterpret_cast<Batch *>(batch);
  if (B->stream) (void)hipStreamSynchronize(B->stream);
  for (Plan &P : B->plan) {
    if (P.exec) (void)hipGraphExecDestroy(P.exec);
    if (P.graph) (void)hipGraphDestroy(P.graph);
    if (P.tab) (void)hipFree(P.tab);
  }
  if (B->hE) (void)hipHostFree(B->hE);
  if (B->sched) (void)hipFree(B->sched);
  if (B->ev0) (void)hipEventDestroy(B->ev0);
  if (B->ev1) (void)hipEventDestroy(B->ev1);
  if (B->stream) (void)hipStreamDestroy(B->stream);
  delete B;
  return TCE_T_OK;
}
