// server.cpp -- a (T) serving loop with continuous batching on one GPU.
//
// Systems (tile description + canonical T1/T2/V2 blocks, host memory owned
// by the caller) are registered once; the first request for a system
// uploads it, and it stays device-resident for later requests until memory
// is needed (idle systems are evicted least-recently-used first).  A
// request is (system, precision).  One worker thread runs the device:
//
//   admission  queued requests, first come first served, become active when
//              a slot is free (max_active) and the device has room for what
//              they add: the upload of their system when it is not resident,
//              the reduced-precision copies when missing, and their own
//              partial buffers, with `reserve` bytes kept free (hipMemGetInfo).
//   slices     every iteration launches, per precision, one persistent
//              fused kernel over a segment table holding the next work items
//              of every active request of that precision: equal shares of a
//              budget sized so the launch takes about slice_ms (the budget is
//              fitted from the measured time per item-cost unit).  Requests
//              that arrive while others run join at the next slice, so their
//              work overlaps on the device.
//   results    a request whose items are done gets its fixed-order task
//              reduction; its energies, like a lone run's, depend only on its
//              own work items, never on what shared the slices.
//
// Every request runs on fused_triples_mixed (8-wide particle boxes; FP64 by
// its FP64 policy).  A system without allowed tasks has no items: its
// requests finish at their first slice without a launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "job.h"

namespace tce_t {
namespace {

using Clock = std::chrono::steady_clock;

struct System {
  tce_t_job *job = nullptr;
  const double *t1, *t2, *v2;
  const long *t1h, *t2h, *v2h;
  int active = 0;      // requests holding it
  long last_use = 0;   // admission counter of the last use (LRU)
};

struct Request {
  long id = 0;
  int sys = 0, precision = 0;
  System *sy = nullptr;  // set at submit under the lock; Systems never move
  double t_submit = 0, t_admit = -1, t_start = -1, t_finish = -1;
  long cursor = 0, nitems = 0;
  double *dpart = nullptr, *dtaskE = nullptr;
  double e4 = 0, e5 = 0;          // published with status (under the lock)
  double e4_out = 0, e5_out = 0;  // the worker's reduction (worker only)
  int status = 1;  // 1 pending, 0 done, < 0 error
};

constexpr int MAX_ACTIVE = 64;

struct Server {
  int device = 0;
  double slice_ms = 50;
  long reserve = 0;
  int max_active = 16;
  Clock::time_point t0 = Clock::now();
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  // heap-allocated so a System never moves when another is added: the
  // worker uses a request's System through Request::sy without the lock,
  // while tce_t_server_add_system may grow this vector (under the lock)
  std::vector<std::unique_ptr<System>> systems;
  std::deque<long> queue;
  std::vector<long> active;
  std::map<long, Request> reqs;
  long next_id = 0, uses = 0;
  bool stop = false;
  std::thread worker;
  hipStream_t stream = nullptr;
  unsigned long long *sched = nullptr;  // dynamic-schedule counters of the stream
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  void *tab_host = nullptr, *tab_dev = nullptr;  // launch tables (pinned / device)
  size_t tab_bytes = 0;
  int cus = 1;
  double ms_per_cost[NPREC] = {0, 0, 0, 0};  // fitted launch time per item-cost unit
  long slices = 0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

Job &job_of(const Request &r) { return *reinterpret_cast<Job *>(r.sy->job); }

long request_bytes(const Job &J) { return std::max(1L, J.nitems8) * 8 * 16 + J.ngroups() * 16; }

// Bytes request r would add to the device (caller holds the lock).
long admission_bytes(const Request &r) {
  Job &J = job_of(r);
  long need = request_bytes(J);
  if (!J.uploaded) need += tce_t_job_device_bytes(reinterpret_cast<tce_t_job *>(&J));
  if (r.precision != TCE_T_FP64 && !(J.uploaded && J.mx[r.precision].ready))
    need += tce_t_job_precision_bytes(reinterpret_cast<tce_t_job *>(&J), r.precision);
  return need;
}

void fail_request(Server &S, Request &r, int rc) {
  r.status = rc;
  r.t_finish = S.now();
}

// Make request r active: upload / convert / allocate (called without the lock).
int activate(Request &r) {
  System &sy = *r.sy;
  Job &J = job_of(r);
  if (!J.uploaded) {
    int rc = tce_t_job_upload(sy.job, sy.t1, sy.t1h, sy.t2, sy.t2h, sy.v2, sy.v2h);
    if (rc != TCE_T_OK) return rc;
  }
  int rc = r.precision == TCE_T_FP64 ? ensure_tasks8(J) : prepare_mixed(J, r.precision);
  if (rc != TCE_T_OK) return rc;
  r.nitems = J.nitems8;
  HIPCHK(hipMalloc(&r.dpart, std::max(1L, J.nitems8) * 8 * 2 * sizeof(double)), "request partials");
  HIPCHK(hipMalloc(&r.dtaskE, std::max(1L, J.ngroups()) * 2 * sizeof(double)),
         "request energies");
  return TCE_T_OK;
}

// Evict idle resident systems (LRU) until `need` bytes fit; caller holds the
// lock, no kernel is in flight (the worker is between slices).
bool make_room(Server &S, long need, const System *keep) {
  for (;;) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return false;
    if ((long)fr - need >= S.reserve) return true;
    System *victim = nullptr;
    for (const auto &p : S.systems) {
      System &sy = *p;
      if (&sy == keep || sy.active > 0 || !reinterpret_cast<Job *>(sy.job)->uploaded) continue;
      if (!victim || sy.last_use < victim->last_use) victim = &sy;
    }
    if (!victim) return false;
    job_release(*reinterpret_cast<Job *>(victim->job));
  }
}

// One slice of the requests of one precision: returns the launch time (ms).
int run_slice(Server &S, int precision, const std::vector<Request *> &rs, float &ms) {
  // every precision on the 16x16 kernel (FP64: its FP64 policy, as a job
  // with 8-wide boxes runs it)
  const size_t asz = (sizeof(MixedArgs) + 15) / 16 * 16;
  const size_t ssz = sizeof(MixedSegment);
  const size_t n = rs.size();
  char *h = (char *)S.tab_host;
  char *d = (char *)S.tab_dev;
  // item budget: equal cost shares of a slice_ms launch (first slice: 2
  // items per CU for each request)
  double total_cost = 0;
  long g = 0;
  bool sparse = false;
  size_t ns = 0;  // segments: requests with items left (a system without
                  // allowed tasks has none and finishes without a launch)
  for (size_t k = 0; k < n; ++k) {
    Request &r = *rs[k];
    const Job &J = job_of(r);
    const long left = r.nitems - r.cursor;
    if (left <= 0) continue;
    const double c = item_cost(J, precision);
    long take;
    if (S.ms_per_cost[precision] > 0)
      take = (long)(S.slice_ms / S.ms_per_cost[precision] / (double)n / c);
    else
      take = 2L * S.cus;
    take = std::max(1L, std::min(take, left));
    MixedArgs M = job_margs(J, precision);
    M.f.partials = r.dpart;
    sparse = sparse || J.sparse;
    std::memcpy(h + ns * asz, &M, sizeof M);
    MixedSegment sg{(const MixedArgs *)(d + ns * asz), r.cursor, r.cursor + take, g};
    std::memcpy(h + n * asz + ns * ssz, &sg, sizeof sg);
    ++ns;
    r.cursor += take;
    g += take;
    total_cost += c * (double)take;
  }
  ms = 0.f;
  if (ns == 0) return TCE_T_OK;
  // the segments follow the argument slots of all n requests (n * asz)
  HIPCHK(hipMemcpyAsync(d, h, n * asz + ns * ssz, hipMemcpyHostToDevice, S.stream), "slice table");
  HIPCHK(hipEventRecord(S.ev0, S.stream), "event");
  const int grid = (int)std::min<long>(S.cus, g);
  const int rc =
      launch_fused_mixed((const MixedSegment *)(d + n * asz), (int)ns, g, precision, grid, S.stream,
                         sparse, sched_use(S.sched));
  if (rc != 0) {
    set_error("slice launch failed");
    return TCE_T_ERR_HIP;
  }
  HIPCHK(hipEventRecord(S.ev1, S.stream), "event");
  HIPCHK(hipStreamSynchronize(S.stream), "slice");
  HIPCHK(hipEventElapsedTime(&ms, S.ev0, S.ev1), "elapsed");
  const double per = ms / std::max(1e-9, total_cost);
  S.ms_per_cost[precision] = S.ms_per_cost[precision] > 0 ? 0.5 * S.ms_per_cost[precision] + 0.5 * per : per;
  ++S.slices;
  return TCE_T_OK;
}

// Reduce a finished request's partials and read its energies.
int finish(Server &S, Request &r) {
  Job &J = job_of(r);
  const long nt = J.ngroups();
  std::vector<double> te(2 * std::max(1L, nt));
  if (nt > 0) {
    if (launch_task_reduce(J.dbound8, (int)nt, 8, r.dpart, r.dtaskE, S.stream) != 0) {
      set_error("request task_reduce failed");
      return TCE_T_ERR_HIP;
    }
    HIPCHK(hipMemcpyAsync(te.data(), r.dtaskE, 2 * nt * sizeof(double), hipMemcpyDeviceToHost,
                          S.stream),
           "request D2H");
    HIPCHK(hipStreamSynchronize(S.stream), "request reduce");
  }
  double s4 = 0, s5 = 0;
  for (long t = 0; t < nt; ++t) {
    s4 += te[2 * t];
    s5 += te[2 * t + 1];
  }
  r.e4_out = s4;
  r.e5_out = s5;
  return TCE_T_OK;
}

void release_request(Request &r) {
  if (r.dpart) (void)hipFree(r.dpart);
  if (r.dtaskE) (void)hipFree(r.dtaskE);
  r.dpart = r.dtaskE = nullptr;
}

void worker_loop(Server &S) {
  (void)hipSetDevice(S.device);
  std::unique_lock<std::mutex> lk(S.mu);
  for (;;) {
    S.cv_work.wait(lk, [&] { return S.stop || !S.queue.empty() || !S.active.empty(); });
    if (S.stop && S.queue.empty() && S.active.empty()) break;
    // ---- admission (first come, first served)
    while (!S.queue.empty() && (int)S.active.size() < S.max_active) {
      Request &r = S.reqs[S.queue.front()];
      if (!make_room(S, admission_bytes(r), r.sy)) break;
      S.queue.pop_front();
      r.sy->active++;
      r.sy->last_use = ++S.uses;
      r.t_admit = S.now();
      lk.unlock();
      const int rc = activate(r);
      lk.lock();
      if (rc != TCE_T_OK) {
        release_request(r);
        r.sy->active--;
        fail_request(S, r, rc);
        S.cv_done.notify_all();
        continue;
      }
      S.active.push_back(r.id);
    }
    if (S.active.empty()) {
      if (!S.queue.empty()) {  // nothing fits even with everything idle evicted
        Request &r = S.reqs[S.queue.front()];
        S.queue.pop_front();
        set_error("request does not fit in device memory");
        fail_request(S, r, TCE_T_ERR_HIP);
        S.cv_done.notify_all();
      }
      continue;
    }
    // ---- one slice per precision present
    std::vector<Request *> group[NPREC];
    for (long id : S.active) group[S.reqs[id].precision].push_back(&S.reqs[id]);
    lk.unlock();
    int rc = TCE_T_OK;
    for (int p = 0; p < NPREC && rc == TCE_T_OK; ++p) {
      if (group[p].empty()) continue;
      const double ts = S.now();
      for (Request *r : group[p])
        if (r->t_start < 0) r->t_start = ts;
      float ms = 0;
      rc = run_slice(S, p, group[p], ms);
    }
    // ---- completions: reduce without the lock, publish under it (a waiter
    // reads status, energies and times under the lock)
    std::vector<Request *> done;
    std::vector<int> frc;
    for (int p = 0; p < NPREC; ++p)
      for (Request *r : group[p])
        if (rc != TCE_T_OK || r->cursor >= r->nitems) done.push_back(r);
    for (Request *r : done) {
      frc.push_back(rc != TCE_T_OK ? rc : finish(S, *r));
      release_request(*r);
    }
    lk.lock();
    for (size_t k = 0; k < done.size(); ++k) {
      Request *r = done[k];
      r->e4 = r->e4_out;
      r->e5 = r->e5_out;
      r->t_finish = S.now();
      r->status = frc[k] != TCE_T_OK ? frc[k] : 0;
      r->sy->active--;
      S.active.erase(std::find(S.active.begin(), S.active.end(), r->id));
    }
    if (!done.empty()) S.cv_done.notify_all();
  }
}

}  // namespace
}  // namespace tce_t

using namespace tce_t;

extern "C" int tce_t_server_create(int device, double slice_ms, int max_active, long reserve_bytes,
                                   tce_t_server **out) {
  if (!out || slice_ms <= 0 || max_active < 1 || max_active > MAX_ACTIVE) return TCE_T_ERR_ARG;
  HIPCHK(hipSetDevice(device), "hipSetDevice");
  auto S = new Server;
  S->device = device;
  S->slice_ms = slice_ms;
  S->max_active = max_active;
  S->reserve = std::max(0L, reserve_bytes);
  auto fail = [&](hipError_t e, const char *what) {
    hip_ok(e, what);
    delete S;
    return TCE_T_ERR_HIP;
  };
  hipError_t e;
  if ((e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking)) != hipSuccess) return fail(e, "stream");
  if ((e = hipEventCreate(&S->ev0)) != hipSuccess) return fail(e, "event");
  if ((e = hipEventCreate(&S->ev1)) != hipSuccess) return fail(e, "event");
  if (!(S->sched = sched_alloc())) {
    delete S;
    return TCE_T_ERR_HIP;
  }
  S->tab_bytes = (size_t)MAX_ACTIVE * (sizeof(MixedArgs) + 16 + sizeof(MixedSegment)) + 256;
  if ((e = hipHostMalloc(&S->tab_host, S->tab_bytes)) != hipSuccess) return fail(e, "slice table");
  if ((e = hipMalloc(&S->tab_dev, S->tab_bytes)) != hipSuccess) return fail(e, "slice table");
  if ((e = hipDeviceGetAttribute(&S->cus, hipDeviceAttributeMultiprocessorCount, device)) != hipSuccess)
    return fail(e, "attr");
  S->worker = std::thread(worker_loop, std::ref(*S));
  *out = reinterpret_cast<tce_t_server *>(S);
  return TCE_T_OK;
}

extern "C" int tce_t_server_add_system(tce_t_server *server, const tce_t_tiles *tiles,
                                       const double *t1, const long *t1h, const double *t2,
                                       const long *t2h, const double *v2, const long *v2h) {
  if (!server || !tiles || !t1 || !t1h || !t2 || !t2h || !v2 || !v2h) return TCE_T_ERR_ARG;
  Server &S = *reinterpret_cast<Server *>(server);
  auto sy = std::make_unique<System>();
  // requests return E[T] and E(T) only: the orbital tiling (least padding)
  // unless TCE_T_SUBTILE names another
  const int rc = getenv("TCE_T_SUBTILE")
                     ? tce_t_job_create(tiles, &sy->job)
                     : tce_t_job_create_tiling(tiles, TCE_T_TILING_ORBITAL, &sy->job);
  if (rc != TCE_T_OK) return rc;
  sy->t1 = t1, sy->t1h = t1h, sy->t2 = t2, sy->t2h = t2h, sy->v2 = v2, sy->v2h = v2h;
  std::lock_guard<std::mutex> lk(S.mu);
  S.systems.push_back(std::move(sy));
  return (int)S.systems.size() - 1;
}

extern "C" long tce_t_server_submit(tce_t_server *server, int system, int precision) {
  if (!server) return TCE_T_ERR_ARG;
  if (!valid_precision(precision)) return TCE_T_ERR_ARG;
  Server &S = *reinterpret_cast<Server *>(server);
  std::lock_guard<std::mutex> lk(S.mu);
  if (system < 0 || system >= (int)S.systems.size() || S.stop) return TCE_T_ERR_ARG;
  Request r;
  r.id = S.next_id++;
  r.sys = system;
  r.sy = S.systems[system].get();
  r.precision = precision;
  r.t_submit = S.now();
  S.reqs[r.id] = r;
  S.queue.push_back(r.id);
  S.cv_work.notify_one();
  return r.id;
}

extern "C" int tce_t_server_wait(tce_t_server *server, long ticket, double timeout_s, double *e_bt,
                                 double *e_pt, double *times) {
  if (!server) return TCE_T_ERR_ARG;
  Server &S = *reinterpret_cast<Server *>(server);
  std::unique_lock<std::mutex> lk(S.mu);
  auto it = S.reqs.find(ticket);
  if (it == S.reqs.end()) return TCE_T_ERR_ARG;
  auto pending = [&] { return S.reqs[ticket].status == 1; };
  if (timeout_s < 0)
    S.cv_done.wait(lk, [&] { return !pending(); });
  else if (!S.cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return !pending(); }))
    return TCE_T_ERR_STATE;
  const Request &r = S.reqs[ticket];
  if (e_bt) *e_bt = r.e4;
  if (e_pt) *e_pt = r.e5;
  if (times) {
    times[0] = r.t_submit;
    times[1] = r.t_admit;
    times[2] = r.t_start;
    times[3] = r.t_finish;
  }
  return r.status;
}

extern "C" int tce_t_server_stats(tce_t_server *server, long *slices, int *resident) {
  if (!server) return TCE_T_ERR_ARG;
  Server &S = *reinterpret_cast<Server *>(server);
  std::lock_guard<std::mutex> lk(S.mu);
  if (slices) *slices = S.slices;
  if (resident) {
    int n = 0;
    for (const auto &sy : S.systems) n += reinterpret_cast<Job *>(sy->job)->uploaded ? 1 : 0;
    *resident = n;
  }
  return TCE_T_OK;
}

extern "C" int tce_t_server_destroy(tce_t_server *server) {
  if (!server) return TCE_T_OK;
  Server *S = reinterpret_cast<Server *>(server);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
  }
  S->cv_work.notify_all();
  if (S->worker.joinable()) S->worker.join();
  for (auto &sy : S->systems) tce_t_job_destroy(sy->job);
  if (S->tab_dev) (void)hipFree(S->tab_dev);
  if (S->tab_host) (void)hipHostFree(S->tab_host);
  if (S->sched) (void)hipFree(S->sched);
  if (S->ev0) (void)hipEventDestroy(S->ev0);
  if (S->ev1) (void)hipEventDestroy(S->ev1);
  if (S->stream) (void)hipStreamDestroy(S->stream);
  delete S;
  return TCE_T_OK;
}
