// Serving loop: hmmsearch queries against one resident target database,
// admitted as they arrive and run concurrently.
//
// The reference's resident-database search (hmmpgmd: HMMER/src/hmmdmstr.c:701-790,
// hmmdwrkr.c:184-644) pops one query at a time and gives it the whole machine.
// Here a fixed set of slots, each a P7_GPU_PIPELINE with its own HIP streams
// and device work areas, takes queries from a FIFO as slots free up. With
// coalescing on (p7_gpu_server_SetCoalesce), a slot that finds several
// queries waiting takes up to `coalesce` of them at once and runs them as one
// p7_gpu_PipelineBatch call: one persistent ungapped-pass launch per kernel
// shape for all of them, then their later stages on concurrent streams;
// otherwise, and with one query waiting, it runs p7_gpu_Pipeline. Memory
// admits the queries: the slots' work areas are set aside when the server is
// made (the budget bounds their number), and each query in flight holds its
// profile's device tables, which must fit in what the budget has left (a
// query that does not fit waits for others to finish; one alone is always
// admitted). Every query's result is the same as a lone p7_gpu_Pipeline call:
// queries share only read-only data (the database and the profiles' device
// copies).
#include <algorithm>
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

#define SERVER_COALESCE_MAX 4      // queries one slot runs as one batch

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
  size_t                 bytes = 0;          // its profile's device tables, while in flight
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
  int coalesce = 1;
  size_t slot_bytes = 0;                     // device memory set aside per slot
  size_t query_budget = 0;                   // device memory for the profiles of queries in flight
  size_t in_flight_bytes = 0;
  int in_flight = 0;
};

// The queries a free slot takes, under the lock: the queue's front ones, up
// to `coalesce`, while their profiles' tables fit in the memory left. The
// sizes come from the profiles' device copies, made on first use.
static std::vector<std::shared_ptr<Request>>
take_queries(P7_GPU_SERVER *srv, std::unique_lock<std::mutex> &lk)
{
  std::vector<std::shared_ptr<Request>> take;
  for (;;) {
    srv->cv_work.wait(lk, [&] { return srv->stop || ! srv->queue.empty(); });
    if (srv->queue.empty()) return take;                // stopping, nothing left
    while (! srv->queue.empty() && (int) take.size() < srv->coalesce) {
      std::shared_ptr<Request> rq = srv->queue.front();
      if (rq->bytes == 0) {
        size_t b = 0;
        lk.unlock();
        const int st = p7_gpu_oprofile_DeviceBytes(rq->om, srv->db, &b);
        lk.lock();
        if (st != p7G_OK) b = 1;                         // the run itself reports the error
        rq->bytes = b;
        if (srv->queue.empty() || srv->queue.front() != rq) continue;   // another slot took it meanwhile
      }
      const bool alone = srv->in_flight == 0 && take.empty();
      if (! alone && srv->in_flight_bytes + rq->bytes > srv->query_budget) break;
      srv->queue.pop_front();
      srv->in_flight_bytes += rq->bytes;
      srv->in_flight++;
      take.push_back(rq);
    }
    if (! take.empty()) return take;
    if (srv->queue.empty()) continue;                    // others took them: wait for more
    srv->cv_work.wait(lk);                               // the front waits for memory to free up
  }
}

static void
worker_main(P7_GPU_SERVER *srv, P7_GPU_PIPELINE *pli)
{
  for (;;) {
    std::vector<std::shared_ptr<Request>> take;
    {
      std::unique_lock<std::mutex> lk(srv->mu);
      take = take_queries(srv, lk);
      if (take.empty()) return;
    }
    const int n = (int) take.size();
    p7g_pipeline_ResetStats(pli);
    std::vector<P7_GPU_STATS> st((size_t) n);
    int status;
    if (n == 1) {
      status = p7_gpu_Pipeline(pli, take[0]->om, srv->db, take[0]->has_hits ? &take[0]->hits : nullptr);
      p7_gpu_pipeline_GetStats(pli, &st[0]);
    } else {
      std::vector<const P7_GPU_OPROFILE *> oms((size_t) n);
      std::vector<P7_GPU_HITS> hits((size_t) n);
      for (int j = 0; j < n; j++) {
        oms[(size_t) j] = take[(size_t) j]->om;
        if (take[(size_t) j]->has_hits) hits[(size_t) j] = take[(size_t) j]->hits;
        else memset(&hits[(size_t) j], 0, sizeof(P7_GPU_HITS));
      }
      status = p7_gpu_PipelineBatch(pli, oms.data(), n, srv->db, hits.data(), st.data(), P7_GPU_BATCH_EAGER);
    }
    {
      std::lock_guard<std::mutex> lk(srv->mu);
      const Clock::time_point now = Clock::now();
      for (int j = 0; j < n; j++) {
        Request &rq = *take[(size_t) j];
        rq.status = status;
        if (status != p7G_OK) { strncpy(rq.errmsg, p7_gpu_LastError(), sizeof(rq.errmsg) - 1); }
        rq.st = st[(size_t) j];
        rq.t_done = now;
        rq.done = true;
        srv->in_flight_bytes -= rq.bytes;
        srv->in_flight--;
      }
    }
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();                           // memory freed: a waiting front may fit now
  }
}

extern "C" P7_GPU_SERVER *
p7_gpu_server_Create(const P7_GPU_SEQDB *db, int nslots, size_t hbm_budget,
                     double F1, double F2, double F3, int do_biasfilter)
{
  if (! db || nslots < 1) { p7g_set_error("server needs a database and at least one slot"); return nullptr; }
  if (p7_gpu_Init(p7g_current_device()) != p7G_OK) return nullptr;
  // admission by device memory: a slot holds its pipeline's work area and, for
  // a batch of up to SERVER_COALESCE_MAX queries, one chain work area, the
  // ungapped pass's results and the caller-order outputs (stage and six scores
  // per sequence) per query, all made here; the rest of the budget holds the
  // profiles of the queries in flight
  const int64_t nseq = p7_gpu_seqdb_NSeq(db);
  const size_t work = p7g_pipeline_work_bytes(nseq, p7_gpu_seqdb_NRes(db));
  const size_t per_slot = work * (1 + SERVER_COALESCE_MAX) + (size_t) SERVER_COALESCE_MAX * (size_t) nseq * 27;
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
  srv->slot_bytes = per_slot;
  srv->query_budget = budget - (size_t) nslots * per_slot;
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
p7_gpu_server_Memory(const P7_GPU_SERVER *srv, size_t *ret_slot_bytes, size_t *ret_query_budget)
{
  if (! srv) { p7g_set_error("server_Memory: null server"); return p7G_EINVAL; }
  if (ret_slot_bytes) *ret_slot_bytes = srv->slot_bytes;
  if (ret_query_budget) *ret_query_budget = srv->query_budget;
  return p7G_OK;
}

extern "C" int
p7_gpu_server_SetCoalesce(P7_GPU_SERVER *srv, int n)
{
  if (! srv || n < 1 || n > SERVER_COALESCE_MAX) {
    p7g_set_error("server_SetCoalesce: n must be 1..%d", SERVER_COALESCE_MAX);
    return p7G_EINVAL;
  }
  // the slots' batch work areas (set aside in the budget) are made now, so
  // that no query waits for them; each slot is taken when it is idle
  if (n > 1)
    for (P7_GPU_PIPELINE *pli : srv->slots)
      if (p7g_pipeline_ReserveBatch(pli, srv->db, SERVER_COALESCE_MAX) != p7G_OK) return p7G_EMEM;
  std::lock_guard<std::mutex> lk(srv->mu);
  srv->coalesce = n;
  return p7G_OK;
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
