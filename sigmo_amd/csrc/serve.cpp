// Serving loop: continuous batching of arriving requests over a resident molecule database.
//
// The reference has no serving path (its driver, library/src/sigmo.cpp:102-159, runs one query
// file per process); this loop reuses the batched form for every admitted batch.  The
// scheduler is independent of the device: `sigmo_serve_schedule` takes an executor callback,
// and `sigmo_serve` plugs in one that runs sigmo_batch_run and sigmo_batch_request_stats.
//
// Scheduling, per worker thread:
//   1. Under the lock, settle the head of the queue: a request whose own cost exceeds the
//      memory budget is rejected, one that waited past the deadline expires.
//   2. If the head has not arrived, sleep until it does (or until a batch completes).
//   3. Otherwise admit arrived requests in order while the batch stays within its request and
//      query-graph limits and the bytes in flight stay within the budget.  If not even the head
//      fits beside the batches in flight, wait for one of them to complete.
//   4. Run the batch outside the lock, then record its requests' results and release its bytes.
// Arrival order is kept: no request is admitted before an earlier one that is still waiting.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "host_graphs.hpp"

namespace sigmo_internal {
int graph_reupload(sigmo_ctx* ctx, const sigmo_host_graphs* host, sigmo_graph** g);
sigmo_ctx* batch_ctx(const sigmo_batch* b);
}  // namespace sigmo_internal

namespace {

using Clock = std::chrono::steady_clock;

constexpr uint32_t kMaxRequestsPerBatch = 256;  // the per-request statistics kernel's limit

struct Scheduler {
  // inputs
  const uint32_t* rgo;
  const double* arrival;
  uint32_t n;
  sigmo_serve_options opt;
  sigmo_serve_executor exec;
  void* user;
  sigmo_serve_result* res;
  // state under `mu`
  std::mutex mu;
  std::condition_variable cv;
  uint32_t next = 0;
  uint64_t inflight_bytes = 0, max_inflight = 0;
  uint32_t batches = 0, active = 0;
  double busy_start = 0, busy = 0, last_done = 0;
  uint64_t matched = 0, embeddings = 0, gmcr = 0;
  bool failed = false;
  int err_code = SIGMO_OK;
  std::string err_msg;
  Clock::time_point t0;

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
  uint32_t graphs(uint32_t r) const { return rgo[r + 1] - rgo[r]; }
  uint64_t cost(uint64_t graph_count) const { return graph_count * opt.bytes_per_graph + opt.bytes_per_batch; }

  void settle(uint32_t r, int status, double t) {
    res[r].status = status;
    res[r].done_s = t;
    res[r].latency_ms = (t - arrival[r]) * 1e3;
  }

  // Step 1 for the head of the queue; returns false if the head was settled (rejected/expired).
  bool head_runnable(double t) {
    const uint32_t r = next;
    if (opt.memory_budget && cost(graphs(r)) > opt.memory_budget) {
      settle(r, SIGMO_SERVE_REJECTED, std::max(t, arrival[r]));
      ++next;
      return false;
    }
    if (opt.deadline_ms > 0 && arrival[r] <= t && (t - arrival[r]) * 1e3 > opt.deadline_ms) {
      settle(r, SIGMO_SERVE_EXPIRED, t);
      ++next;
      return false;
    }
    return true;
  }

  void worker(uint32_t w) {
    std::vector<uint32_t> reqs;
    std::vector<sigmo_request_stats> stats;
    for (;;) {
      uint64_t batch_bytes = 0;
      uint32_t batch_id = 0;
      {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
          if (failed || next >= n) return;
          const double t = now();
          if (!head_runnable(t)) {
            cv.notify_all();
            continue;
          }
          if (arrival[next] > t) {
            cv.wait_until(lk, t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival[next])));
            continue;
          }
          reqs.clear();
          uint64_t g = 0;
          const uint32_t max_req = (opt.max_batch_requests == 0 || opt.max_batch_requests > kMaxRequestsPerBatch) ? kMaxRequestsPerBatch
                                                                                                              : opt.max_batch_requests;
          while (next < n && arrival[next] <= t && reqs.size() < max_req) {
            if (!reqs.empty() && !head_runnable(t)) continue;  // settled; look at the next one
            const uint64_t g2 = g + graphs(next);
            if (!reqs.empty() && opt.max_batch_graphs && g2 > opt.max_batch_graphs) break;
            if (opt.memory_budget && inflight_bytes + cost(g2) > opt.memory_budget) break;
            reqs.push_back(next++);
            g = g2;
          }
          if (reqs.empty()) {  // the head waits for bytes in flight to be released
            cv.wait(lk);
            continue;
          }
          batch_bytes = cost(g);
          inflight_bytes += batch_bytes;
          max_inflight = std::max(max_inflight, inflight_bytes);
          batch_id = batches++;
          if (active++ == 0) busy_start = t;
          for (uint32_t r : reqs) {
            res[r].admit_s = t;
            res[r].batch = batch_id;
            res[r].batch_requests = static_cast<uint32_t>(reqs.size());
            res[r].worker = w;
          }
          break;
        }
      }
      stats.assign(reqs.size(), sigmo_request_stats{0, 0, 0});
      uint64_t gp = 0;
      const int st = exec(user, w, reqs.data(), static_cast<uint32_t>(reqs.size()), stats.data(), &gp);
      std::string msg = st == SIGMO_OK ? std::string() : std::string(sigmo_last_error());
      {
        std::lock_guard<std::mutex> lk(mu);
        const double t = now();
        inflight_bytes -= batch_bytes;
        if (--active == 0) busy += t - busy_start;
        if (st != SIGMO_OK) {
          if (!failed) {
            failed = true;
            err_code = st;
            err_msg = "serving batch " + std::to_string(batch_id) + ": " + msg;
          }
        } else {
          for (size_t i = 0; i < reqs.size(); ++i) {
            const uint32_t r = reqs[i];
            settle(r, SIGMO_SERVE_DONE, t);
            res[r].matched_pairs = stats[i].matched_pairs;
            res[r].embeddings = stats[i].embeddings;
            res[r].hash = stats[i].hash;
            matched += stats[i].matched_pairs;
            embeddings += stats[i].embeddings;
          }
          gmcr += gp;
          last_done = std::max(last_done, t);
        }
        cv.notify_all();
      }
    }
  }
};

// Executor of sigmo_serve: one reusable query-graph upload per worker.
struct DeviceServe {
  sigmo_batch* const* workers;
  const sigmo_host_graphs* queries;
  const uint32_t* rgo;
  int find_first;
  std::vector<sigmo_graph*> graphs;
  std::vector<sigmo_host_graphs> host;
  std::vector<std::vector<uint32_t>> local;

  static int run(void* user, uint32_t w, const uint32_t* requests, uint32_t count, sigmo_request_stats* out, uint64_t* gmcr_pairs) {
    auto* self = static_cast<DeviceServe*>(user);
    sigmo_host_graphs& hb = self->host[w];
    std::vector<uint32_t>& lo = self->local[w];
    hb = sigmo_host_graphs();
    lo.assign(1, 0);
    for (uint32_t i = 0; i < count; ++i) {
      const uint32_t r = requests[i];
      sigmo_host::append_batch(hb, *self->queries, self->rgo[r], self->rgo[r + 1] - self->rgo[r]);
      lo.push_back(hb.num_graphs());
    }
    sigmo_batch* b = self->workers[w];
    int st = sigmo_internal::graph_reupload(sigmo_internal::batch_ctx(b), &hb, &self->graphs[w]);
    if (st != SIGMO_OK) return st;
    sigmo_batch_result br;
    st = sigmo_batch_run(b, self->graphs[w], self->find_first, &br);
    if (st != SIGMO_OK) return st;
    *gmcr_pairs = br.gmcr_pairs;
    return sigmo_batch_request_stats(b, lo.data(), count, out);
  }
};

int check_inputs(const uint32_t* rgo, const double* arrival, uint32_t n, const sigmo_serve_options* opt, sigmo_serve_result* results) {
  if (!rgo || !arrival || !opt || !results) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  for (uint32_t r = 0; r < n; ++r) {
    if (rgo[r + 1] <= rgo[r]) return sigmo_err::set(SIGMO_ERR_INVALID, "request " + std::to_string(r) + " carries no query graph");
    if (!(arrival[r] >= 0) || (r && arrival[r] < arrival[r - 1]))
      return sigmo_err::set(SIGMO_ERR_INVALID, "arrival times must be non-negative and non-decreasing");
  }
  return SIGMO_OK;
}

}  // namespace

extern "C" {

int sigmo_serve_schedule(uint32_t num_workers, const uint32_t* request_graph_offsets, const double* arrival_s, uint32_t num_requests,
                         const sigmo_serve_options* options, sigmo_serve_executor executor, void* user, sigmo_serve_result* results,
                         sigmo_serve_summary* summary) {
  if (num_workers == 0 || num_workers > 16) return sigmo_err::set(SIGMO_ERR_INVALID, "1..16 workers");
  if (!executor) return sigmo_err::set(SIGMO_ERR_INVALID, "null executor");
  const int chk = check_inputs(request_graph_offsets, arrival_s, num_requests, options, results);
  if (chk != SIGMO_OK) return chk;
  std::memset(results, 0, sizeof(sigmo_serve_result) * num_requests);
  Scheduler sch;
  sch.rgo = request_graph_offsets;
  sch.arrival = arrival_s;
  sch.n = num_requests;
  sch.opt = *options;
  sch.exec = executor;
  sch.user = user;
  sch.res = results;
  sch.t0 = Clock::now();
  std::vector<std::thread> threads;
  for (uint32_t w = 1; w < num_workers; ++w) threads.emplace_back([&sch, w] { sch.worker(w); });
  sch.worker(0);
  for (auto& t : threads) t.join();
  if (sch.failed) return sigmo_err::set(sch.err_code, sch.err_msg);
  if (summary) {
    std::memset(summary, 0, sizeof(*summary));
    summary->batches = sch.batches;
    for (uint32_t r = 0; r < num_requests; ++r) {
      summary->done += results[r].status == SIGMO_SERVE_DONE;
      summary->rejected += results[r].status == SIGMO_SERVE_REJECTED;
      summary->expired += results[r].status == SIGMO_SERVE_EXPIRED;
      sch.last_done = std::max(sch.last_done, results[r].done_s);
    }
    summary->wall_s = sch.last_done;
    summary->busy_s = sch.busy;
    summary->max_inflight_bytes = sch.max_inflight;
    summary->matched_pairs = sch.matched;
    summary->embeddings = sch.embeddings;
    summary->gmcr_pairs = sch.gmcr;
  }
  return SIGMO_OK;
}

int sigmo_serve(sigmo_batch* const* workers, uint32_t num_workers, const sigmo_host_graphs* queries, const uint32_t* request_graph_offsets,
                const double* arrival_s, uint32_t num_requests, const sigmo_serve_options* options, sigmo_serve_result* results,
                sigmo_serve_summary* summary) {
  if (!workers || !queries || !options) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (num_workers == 0 || num_workers > 16) return sigmo_err::set(SIGMO_ERR_INVALID, "1..16 workers");
  for (uint32_t w = 0; w < num_workers; ++w)
    if (!workers[w]) return sigmo_err::set(SIGMO_ERR_INVALID, "null worker batch");
  const int chk = check_inputs(request_graph_offsets, arrival_s, num_requests, options, results);
  if (chk != SIGMO_OK) return chk;
  if (num_requests && request_graph_offsets[num_requests] > queries->num_graphs())
    return sigmo_err::set(SIGMO_ERR_INVALID, "request graph offsets exceed the query graphs");
  sigmo_serve_options opt = *options;
  if (opt.bytes_per_graph == 0 && num_requests) {
    // price a query graph at the batch's mean size, from the workspace estimate of a
    // 64-graph batch (the per-pair offsets dominate and grow linearly)
    const uint32_t G = request_graph_offsets[num_requests] - request_graph_offsets[0];
    const uint64_t nodes = queries->graph_offsets[request_graph_offsets[num_requests]] - queries->graph_offsets[request_graph_offsets[0]];
    uint64_t b64 = 0;
    const int st = sigmo_batch_workspace_bytes(workers[0], 64, (nodes * 64 + G - 1) / G, &b64);
    if (st != SIGMO_OK) return st;
    opt.bytes_per_graph = (b64 + 63) / 64;
  }
  DeviceServe ds;
  ds.workers = workers;
  ds.queries = queries;
  ds.rgo = request_graph_offsets;
  ds.find_first = opt.find_first;
  ds.graphs.assign(num_workers, nullptr);
  ds.host.resize(num_workers);
  ds.local.resize(num_workers);
  const int st = sigmo_serve_schedule(num_workers, request_graph_offsets, arrival_s, num_requests, &opt, &DeviceServe::run, &ds, results,
                                      summary);
  for (sigmo_graph* g : ds.graphs) sigmo_graph_free(g);
  return st;
}

}  // extern "C"
