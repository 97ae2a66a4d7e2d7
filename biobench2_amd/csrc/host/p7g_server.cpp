// Serving loop: hmmsearch queries against one resident target database,
// admitted as they arrive and run concurrently.
//
// The reference's resident-database search (hmmpgmd: HMMER/src/hmmdmstr.c:701-790,
// hmmdwrkr.c:184-644) pops one query at a time and gives it the whole machine.
// Here a fixed set of slots, each a P7_GPU_PIPELINE with its own HIP stream and
// device work area, takes queries from a FIFO as slots free up, so the stage
// kernels of several queries overlap on the device. The number of slots is
// bounded by a device-memory budget (admission by memory). Every query's
// result is the same as a lone p7_gpu_Pipeline call: the slots share only
// read-only data (the database and the device copies of the profiles).
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>
#include <string.h>

#include "p7_gpu.h"
#include "p7g_internal.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  int64_t                ticket = 0;
  const P7_GPU_OPROFILE *om = nullptr;
  P7_GPU_HITS            hits;
  bool                   has_hits = false;
  Clock::time_point      t_submit, t_done;
  bool                   done = false;
  int                    status = p7G_OK;
  char                   errmsg[256] = { 0 };
  P7_GPU_STATS           st;
};

}  // namespace

struct p7g_server_s {
  const P7_GPU_SEQDB *db = nullptr;
  std::vector<P7_GPU_PIPELINE *> slots;
  std::vector<std::thread> workers;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<std::shared_ptr<Request>> queue;
  std::map<int64_t, std::shared_ptr<Request>> live;
  int64_t next_ticket = 1;
  bool stop = false;
};

static void
worker_main(P7_GPU_SERVER *srv, P7_GPU_PIPELINE *pli)
{
  for (;;) {
    std::shared_ptr<Request> rq;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      srv->cv_work.wait(lk, [&] { return srv->stop || ! srv->queue.empty(); });
      if (srv->queue.empty()) return;                 // stopping, nothing left
      rq = srv->queue.front();
      srv->queue.pop_front();
    }
    p7g_pipeline_ResetStats(pli);
    const int status = p7_gpu_Pipeline(pli, rq->om, srv->db, rq->has_hits ? &rq->hits : nullptr);
    P7_GPU_STATS st;
    p7_gpu_pipeline_GetStats(pli, &st);
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      rq->status = status;
      if (status != p7G_OK) { strncpy(rq->errmsg, p7_gpu_LastError(), sizeof(rq->errmsg) - 1); }
      rq->st = st;
      rq->t_done = Clock::now();
      rq->done = true;
    }
    srv->cv_done.notify_all();
  }
}

extern "C" P7_GPU_SERVER *
p7_gpu_server_Create(const P7_GPU_SEQDB *db, int nslots, size_t hbm_budget,
                     double F1, double F2, double F3, int do_biasfilter)
{
  if (! db || nslots < 1) { p7g_set_error("server needs a database and at least one slot"); return nullptr; }
  if (p7_gpu_Init(p7g_current_device()) != p7G_OK) return nullptr;
  // admission by device memory: a slot's work area is fixed by the database size
  const size_t per_slot = p7g_pipeline_work_bytes(p7_gpu_seqdb_NSeq(db), p7_gpu_seqdb_NRes(db));
  size_t avail = 0;
  if (p7g_device_free_bytes(&avail) != p7G_OK) return nullptr;
  size_t budget = (size_t) (0.9 * (double) avail);
  if (hbm_budget > 0 && hbm_budget < budget) budget = hbm_budget;
  const int fit = (int) (budget / (per_slot ? per_slot : 1));
  if (fit < 1) {
    p7g_set_error("device memory budget (%zu bytes) holds no query slot (%zu bytes each)", budget, per_slot);
    return nullptr;
  }
  if (nslots > fit) nslots = fit;

  P7_GPU_SERVER *srv = new (std::nothrow) p7g_server_s();
  if (! srv) return nullptr;
  srv->db = db;
  for (int i = 0; i < nslots; i++) {
    P7_GPU_PIPELINE *pli = p7_gpu_pipeline_Create(F1, F2, F3, do_biasfilter, 0);
    if (! pli || p7g_pipeline_Reserve(pli, db) != p7G_OK) {
      if (pli) p7_gpu_pipeline_Destroy(pli);
      p7_gpu_server_Destroy(srv);
      return nullptr;
    }
    srv->slots.push_back(pli);
  }
  for (P7_GPU_PIPELINE *pli : srv->slots) srv->workers.emplace_back(worker_main, srv, pli);
  return srv;
}

extern "C" int
p7_gpu_server_Slots(const P7_GPU_SERVER *srv)
{
  return srv ? (int) srv->slots.size() : 0;
}

extern "C" int
p7_gpu_server_Submit(P7_GPU_SERVER *srv, const P7_GPU_OPROFILE *om, P7_GPU_HITS *hits, int64_t *ret_ticket)
{
  if (! srv || ! om || ! ret_ticket) { p7g_set_error("server_Submit: null argument"); return p7G_EINVAL; }
  auto rq = std::make_shared<Request>();
  memset(&rq->hits, 0, sizeof(rq->hits));
  if (hits) { rq->hits = *hits; rq->has_hits = true; }
  rq->om = om;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    if (srv->stop) { p7g_set_error("server is shutting down"); return p7G_EINVAL; }
    rq->ticket = srv->next_ticket++;
    rq->t_submit = Clock::now();
    srv->live[rq->ticket] = rq;
    srv->queue.push_back(rq);
    *ret_ticket = rq->ticket;
  }
  srv->cv_work.notify_one();
  return p7G_OK;
}

extern "C" int
p7_gpu_server_Wait(P7_GPU_SERVER *srv, int64_t ticket, P7_GPU_STATS *ret_stats, double *ret_latency_ms)
{
  if (! srv) { p7g_set_error("server_Wait: null server"); return p7G_EINVAL; }
  std::shared_ptr<Request> rq;
  {
    std::unique_lock<std::mutex> lk(srv->mu);
    auto it = srv->live.find(ticket);
    if (it == srv->live.end()) { p7g_set_error("unknown ticket %lld", (long long) ticket); return p7G_ENOTFOUND; }
    rq = it->second;
    srv->cv_done.wait(lk, [&] { return rq->done; });
    srv->live.erase(ticket);
  }
  if (ret_stats) *ret_stats = rq->st;
  if (ret_latency_ms) *ret_latency_ms = std::chrono::duration<double, std::milli>(rq->t_done - rq->t_submit).count();
  if (rq->status != p7G_OK) p7g_set_error("%s", rq->errmsg);
  return rq->status;
}

extern "C" void
p7_gpu_server_Destroy(P7_GPU_SERVER *srv)
{
  if (! srv) return;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stop = true;                 // workers finish what is queued, then leave
  }
  srv->cv_work.notify_all();
  for (std::thread &t : srv->workers) t.join();
  for (P7_GPU_PIPELINE *pli : srv->slots) p7_gpu_pipeline_Destroy(pli);
  delete srv;
}
