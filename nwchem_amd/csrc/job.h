// job.h -- the device-resident (T) job shared by the whole-job (job.hip)
// and batch (batch.hip) layers.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "fused.h"
#include "internal.h"

namespace tce_t {

// Kernel task tables (tiling modes TCE_T_TILING_*).  A reference task (tile
// sextuple of the reference task loop) runs as one kernel task with boxes
// over its tiles (REFERENCE), or as the kernel tasks of the box-sized
// sub-tiles of its tiles (SUBTILE, the default: canonical sub-tile order,
// factors by sub-tile equality), which skips the symmetric duplicates and
// the padding inside equal tiles; either way reference task t owns the
// contiguous items [bound[t], bound[t+1]), so its energy is a fixed-order
// sum of its own items.  ORBITAL re-tiles each run of consecutive tiles of
// one spin and irrep into box-sized pieces regardless of the reference tile
// edges (one ragged piece per run instead of one per tile); only the job's
// totals are defined then (bound = {0, items}, one reduction group).
struct Job {
  Tiles T;
  std::vector<TaskDev> tasks;
  std::vector<long> bound;       // item boundaries of the reduction groups in `tasks`
  std::vector<long> task_tiles;  // 6 per reference task: p4b,p5b,p6b,h1b,h2b,h3b
  long nitems = 0;
  int tiling = TCE_T_TILING_SUBTILE;
  long ntask() const { return (long)task_tiles.size() / 6; }
  // reduction groups: the reference tasks, or the whole job (ORBITAL)
  long ngroups() const { return (long)bound.size() - 1; }
  int device = 0;
  hipStream_t stream = nullptr;
  unsigned long long *dsched = nullptr;  // dynamic-schedule counters of the job's stream
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  double *dT1 = nullptr, *dT2e = nullptr, *dT2m = nullptr, *dVd2 = nullptr, *dWd1 = nullptr,
         *dS = nullptr, *devl = nullptr, *dpart = nullptr, *dtaskE = nullptr;
  TaskDev *dtasks = nullptr;
  bool uploaded = false;
  bool cr = false;  // the last upload was a CR-CCSD(T) job (tce_t_job_upload_cr)
  long gen = 0;  // completed uploads: every upload replaces all device buffers
  float last_ms = 0.f;
  std::vector<double> taskE;
  long Ep = 0, Mp = 0;
  int grid = 0;
  int pbox = 4;  // particle box edge of a work item: 4 (fused_wave) or 8 (fused_wg)
  bool k16 = false;  // FP64 on the 16x16 MFMA kernel (fused_triples_mixed, FP64 policy)
  bool sparse = false;  // some task's contraction ranges are narrower than the spaces
  int ppi = 1;   // energy partials per work item (waves per item)
  // reduced-precision path (fused_mixed.hip): 8-wide particle boxes always
  std::vector<TaskDev> tasks8;
  std::vector<long> bound8;
  long nitems8 = 0;
  TaskDev *dtasks8 = nullptr;  // == dtasks when pbox == 8
  double *dpart8 = nullptr;    // == dpart when pbox == 8
  long *dbound = nullptr, *dbound8 = nullptr;  // dbound8 == dbound when pbox == 8
  // launch tables (job arguments + segments): pinned host staging, device copy
  void *tab_host = nullptr, *tab_dev = nullptr;
  void *dtile = nullptr;  // tile export: task entry, then doubles and singles tiles
  size_t dtile_bytes = 0;
  struct Mixed {
    void *x2[3] = {nullptr, nullptr, nullptr}, *y2[3] = {nullptr, nullptr, nullptr};  // T2e, Vd2
    void *x1[3] = {nullptr, nullptr, nullptr}, *y1[3] = {nullptr, nullptr, nullptr};  // Wd1, T2m
    int Ek = 0, Mk = 0;
    bool ready = false;
  } mx[4];  // indexed by precision (1 = FP32, 2 = BF16X3, 3 = BF16X6)
};

// Records a HIP failure (tce_t_last_error) and returns false.
bool hip_ok(hipError_t e, const char *what);
// One task's (or one p4 x p5 slice's) t3 tiles of an uploaded job computed
// on the device, without the copy to the host: doubles at *tiles, then
// singles (n doubles later, n = the tile's elements, from d's ranges), in
// the layout of tce_t_job_triples_tiles; queued on J.stream.  d receives
// the task entry (orbital offsets and ranges of the slice).
int tile_export_device(Job &J, int precision, long h1b, long h2b, long h3b, long p4b, long p5b,
                       long p6b, long i0, long ni, long j0, long nj, long ld5, TaskDev &d,
                       double **tiles);
#define HIPCHK(x, what)                             \
  do {                                              \
    if (!::tce_t::hip_ok((x), what)) return TCE_T_ERR_HIP; \
  } while (0)

// Fixed-order sums of a job's partials (ppi per item) over the items
// [bounds[t], bounds[t+1]) of each of ntasks reference tasks into out
// (2 doubles per task), on stream s.
int launch_task_reduce(const long *bounds, int ntasks, int ppi, const double *part, double *out,
                       hipStream_t s);
// Device task table and partials of 8-wide particle boxes (the cooperative
// and reduced-precision kernels); a no-op when the job already uses them.
int ensure_tasks8(Job &J);
// Reduced-precision copies of the contraction tensors (converted once).
int prepare_mixed(Job &J, int precision);
// The arguments M and one segment of items [item0, item1) in the job's
// launch table, copied on J.stream (the host staging is reused by the next
// call: synchronise J.stream first); the device segment, or null.
const MixedSegment *stage_mixed(Job &J, const MixedArgs &M, long item1, long item0);
// Kernel arguments of a job: FP64 with the job's own task table (wide =
// false) or the 8-wide one (wide = true); reduced precision.
FusedArgs job_args(const Job &J, bool wide);
MixedArgs job_margs(const Job &J, int precision);
// Dense tensors of an upload built on the device from the canonical blocks
// (pack.hip); the dense buffers must be allocated.  Returns a status.
// PACK_CR: a CR-CCSD(T) job, Wd1 from the moment intermediate i1(h11 p4 h1
// h2) in v2/v2h, Vd2 from i1(p4 p5 h1 p12) in x2/x2h (cr_ccsd_t_N.F), S from
// t2 and t1 (the denominator tile of cr_ccsd_t_E.F).
enum { PACK_TRIPLES = 0, PACK_CR = 1 };
int pack_on_device(Job &J, const double *t1, const long *t1h, const double *t2, const long *t2h,
                   const double *v2, const long *v2h, int kind = PACK_TRIPLES,
                   const double *x2 = nullptr, const long *x2h = nullptr);
// Free every device buffer of the job (it can be uploaded again).
void job_release(Job &J);
// Per-item cost weight (contraction chunks of an item) for scheduling.
double item_cost(const Job &J, int precision);
// Device item boundaries of the reference tasks for the table a run uses.
inline const long *job_bounds(const Job &J, bool wide) { return wide ? J.dbound8 : J.dbound; }

}  // namespace tce_t
