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
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "host_graphs.hpp"

namespace sigmo_internal {
int graph_reupload(sigmo_ctx* ctx, const sigmo_host_graphs* host, sigmo_graph** g);
sigmo_ctx* batch_ctx(const sigmo_batch* b);
}  // namespace sigmo_internal

// Named (not anonymous) so that the C ABI's sigmo_server can hold a Scheduler.
namespace sigmo_serve_detail {

using Clock = std::chrono::steady_clock;

constexpr uint32_t kMaxRequestsPerBatch = 256;  // the per-request statistics kernel's limit

struct Scheduler {
  sigmo_serve_options opt;
  sigmo_serve_executor exec;
  void* user;
  // requests, append-only; everything below is guarded by `mu`
  std::mutex mu;
  std::condition_variable cv;
  std::vector<uint32_t> graph_count;      // query graphs per request
  std::vector<double> arrival;            // seconds from t0
  std::vector<uint64_t> bytes;            // the request's share of a batch's device bytes
  std::vector<sigmo_serve_result> res;
  bool open_ended = false;  // online: workers wait for more requests until `closing`
  bool closing = false;
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
  uint32_t size() const { return static_cast<uint32_t>(graph_count.size()); }

  void push(uint32_t graphs, double at, uint64_t request_bytes) {
    graph_count.push_back(graphs);
    arrival.push_back(at);
    bytes.push_back(request_bytes);
    sigmo_serve_result r{};
    r.status = SIGMO_SERVE_PENDING;
    res.push_back(r);
  }

  void settle(uint32_t r, int status, double t) {
    res[r].status = status;
    res[r].done_s = t;
    res[r].latency_ms = (t - arrival[r]) * 1e3;
  }

  // Step 1 for the head of the queue; returns false if the head was settled (rejected/expired).
  bool head_runnable(double t) {
    const uint32_t r = next;
    if (opt.memory_budget && bytes[r] + opt.bytes_per_batch > opt.memory_budget) {
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
          if (failed) return;
          if (next >= size()) {
            if (!open_ended || closing) return;
            cv.wait(lk);  // an online server waits for the next submission
            continue;
          }
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
          uint64_t g = 0, b = opt.bytes_per_batch;
          const uint32_t max_req = (opt.max_batch_requests == 0 || opt.max_batch_requests > kMaxRequestsPerBatch) ? kMaxRequestsPerBatch
                                                                                                              : opt.max_batch_requests;
          while (next < size() && arrival[next] <= t && reqs.size() < max_req) {
            if (!reqs.empty() && !head_runnable(t)) continue;  // settled; look at the next one
            const uint64_t g2 = g + graph_count[next], b2 = b + bytes[next];
            if (!reqs.empty() && opt.max_batch_graphs && g2 > opt.max_batch_graphs) break;
            if (opt.memory_budget && inflight_bytes + b2 > opt.memory_budget) break;
            reqs.push_back(next++);
            g = g2;
            b = b2;
          }
          if (reqs.empty()) {  // the head waits for bytes in flight to be released
            cv.wait(lk);
            continue;
          }
          batch_bytes = b;
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

  void summarize(sigmo_serve_summary* summary) {
    std::memset(summary, 0, sizeof(*summary));
    summary->batches = batches;
    for (const auto& r : res) {
      summary->done += r.status == SIGMO_SERVE_DONE;
      summary->rejected += r.status == SIGMO_SERVE_REJECTED;
      summary->expired += r.status == SIGMO_SERVE_EXPIRED;
      last_done = std::max(last_done, r.done_s);
    }
    summary->wall_s = last_done;
    summary->busy_s = busy;
    summary->max_inflight_bytes = max_inflight;
    summary->matched_pairs = matched;
    summary->embeddings = embeddings;
    summary->gmcr_pairs = gmcr;
  }
};

// Uploads one batch's query graphs (the requests' graphs, in order) into the worker's reusable
// device graph, runs it and fills the per-request statistics.
int run_device_batch(sigmo_batch* b, sigmo_graph** g, const sigmo_host_graphs& hb, const std::vector<uint32_t>& local, int find_first,
                     uint32_t count, sigmo_request_stats* out, uint64_t* gmcr_pairs) {
  int st = sigmo_internal::graph_reupload(sigmo_internal::batch_ctx(b), &hb, g);
  if (st != SIGMO_OK) return st;
  sigmo_batch_result br;
  st = sigmo_batch_run(b, *g, find_first, &br);
  if (st != SIGMO_OK) return st;
  *gmcr_pairs = br.gmcr_pairs;
  return sigmo_batch_request_stats(b, local.data(), count, out);
}

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
    return run_device_batch(self->workers[w], &self->graphs[w], hb, lo, self->find_first, count, out, gmcr_pairs);
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

}  // namespace sigmo_serve_detail

using namespace sigmo_serve_detail;

extern "C" {

int sigmo_serve_schedule(uint32_t num_workers, const uint32_t* request_graph_offsets, const double* arrival_s, uint32_t num_requests,
                         const sigmo_serve_options* options, sigmo_serve_executor executor, void* user, sigmo_serve_result* results,
                         sigmo_serve_summary* summary) {
  if (num_workers == 0 || num_workers > 16) return sigmo_err::set(SIGMO_ERR_INVALID, "1..16 workers");
  if (!executor) return sigmo_err::set(SIGMO_ERR_INVALID, "null executor");
  const int chk = check_inputs(request_graph_offsets, arrival_s, num_requests, options, results);
  if (chk != SIGMO_OK) return chk;
  Scheduler sch;
  sch.opt = *options;
  sch.exec = executor;
  sch.user = user;
  for (uint32_t r = 0; r < num_requests; ++r) {
    const uint32_t g = request_graph_offsets[r + 1] - request_graph_offsets[r];
    sch.push(g, arrival_s[r], uint64_t(g) * options->bytes_per_graph);
  }
  sch.t0 = Clock::now();
  std::vector<std::thread> threads;
  for (uint32_t w = 1; w < num_workers; ++w) threads.emplace_back([&sch, w] { sch.worker(w); });
  sch.worker(0);
  for (auto& t : threads) t.join();
  if (num_requests) std::memcpy(results, sch.res.data(), sizeof(sigmo_serve_result) * num_requests);
  if (sch.failed) return sigmo_err::set(sch.err_code, sch.err_msg);
  if (summary) sch.summarize(summary);
  return SIGMO_OK;
}

}  // extern "C"

namespace {
// Workers run concurrently, each on its batch's context (its stream, join statistics and pinned
// readback): two workers on one context, or one batch listed twice, would share those.
int check_workers(sigmo_batch* const* workers, uint32_t n) {
  for (uint32_t w = 0; w < n; ++w) {
    if (!workers[w]) return sigmo_err::set(SIGMO_ERR_INVALID, "null worker batch");
    for (uint32_t v = 0; v < w; ++v)
      if (workers[v] == workers[w] || sigmo_internal::batch_ctx(workers[v]) == sigmo_internal::batch_ctx(workers[w]))
        return sigmo_err::set(SIGMO_ERR_INVALID, "workers " + std::to_string(v) + " and " + std::to_string(w) +
                                                     " share a context: each worker needs its own");
  }
  return SIGMO_OK;
}
}  // namespace

extern "C" {

int sigmo_serve(sigmo_batch* const* workers, uint32_t num_workers, const sigmo_host_graphs* queries, const uint32_t* request_graph_offsets,
                const double* arrival_s, uint32_t num_requests, const sigmo_serve_options* options, sigmo_serve_result* results,
                sigmo_serve_summary* summary) {
  if (!workers || !queries || !options) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (num_workers == 0 || num_workers > 16) return sigmo_err::set(SIGMO_ERR_INVALID, "1..16 workers");
  if (const int wk = check_workers(workers, num_workers); wk != SIGMO_OK) return wk;
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

// ---- online serving ------------------------------------------------------------------------

struct sigmo_server {
  Scheduler sch;
  std::vector<sigmo_batch*> workers;
  std::vector<std::unique_ptr<sigmo_host_graphs>> graphs;  // each request's query graphs (guarded by sch.mu)
  std::vector<sigmo_graph*> dev;                           // per worker: its reusable device upload
  std::vector<sigmo_host_graphs> host;                     // per worker: the batch being assembled
  std::vector<std::vector<uint32_t>> local;
  std::vector<std::thread> threads;
  bool device = true;       // false: a caller-supplied executor (sigmo_server_open_schedule)
  bool price_each = false;  // bytes_per_graph was 0: price every request from its own size

  static int run(void* user, uint32_t w, const uint32_t* requests, uint32_t count, sigmo_request_stats* out, uint64_t* gmcr_pairs) {
    auto* self = static_cast<sigmo_server*>(user);
    std::vector<const sigmo_host_graphs*> src(count);
    {
      std::lock_guard<std::mutex> lk(self->sch.mu);  // submissions may grow `graphs` meanwhile
      for (uint32_t i = 0; i < count; ++i) src[i] = self->graphs[requests[i]].get();
    }
    sigmo_host_graphs& hb = self->host[w];
    std::vector<uint32_t>& lo = self->local[w];
    hb = sigmo_host_graphs();
    lo.assign(1, 0);
    for (const sigmo_host_graphs* g : src) {
      sigmo_host::append_batch(hb, *g, 0, g->num_graphs());
      lo.push_back(hb.num_graphs());
    }
    {
      // the batch holds its own copy now: release the requests' graphs (a long-running server
      // keeps only each request's result)
      std::lock_guard<std::mutex> lk(self->sch.mu);
      for (uint32_t i = 0; i < count; ++i) self->graphs[requests[i]].reset();
    }
    // sch.opt is fixed once the server is open
    return run_device_batch(self->workers[w], &self->dev[w], hb, lo, self->sch.opt.find_first, count, out, gmcr_pairs);
  }
};

namespace {
sigmo_server* server_start(sigmo_server* srv, uint32_t num_workers, const sigmo_serve_options* options) {
  srv->sch.opt = *options;
  srv->sch.open_ended = true;
  srv->sch.t0 = Clock::now();
  for (uint32_t w = 0; w < num_workers; ++w) srv->threads.emplace_back([srv, w] { srv->sch.worker(w); });
  return srv;
}

int server_push(sigmo_server* srv, uint32_t graphs, uint64_t request_bytes, std::unique_ptr<sigmo_host_graphs> copy, uint64_t* request_id) {
  std::lock_guard<std::mutex> lk(srv->sch.mu);
  if (srv->sch.closing) return sigmo_err::set(SIGMO_ERR_INVALID, "server is closing");
  if (srv->sch.failed) return sigmo_err::set(srv->sch.err_code, srv->sch.err_msg);
  if (srv->sch.size() == UINT32_MAX) return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "too many requests for one server");
  *request_id = srv->sch.size();
  srv->graphs.push_back(std::move(copy));
  srv->sch.push(graphs, srv->sch.now(), request_bytes);
  srv->sch.cv.notify_all();
  return SIGMO_OK;
}
}  // namespace

extern "C" {

int sigmo_server_open(sigmo_batch* const* workers, uint32_t num_workers, const sigmo_serve_options* options, sigmo_server** out) {
  if (!workers || !options || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (num_workers == 0 || num_workers > 16) return sigmo_err::set(SIGMO_ERR_INVALID, "1..16 workers");
  if (const int wk = check_workers(workers, num_workers); wk != SIGMO_OK) return wk;
  auto* srv = new sigmo_server();
  srv->workers.assign(workers, workers + num_workers);
  srv->dev.assign(num_workers, nullptr);
  srv->host.resize(num_workers);
  srv->local.resize(num_workers);
  srv->price_each = options->bytes_per_graph == 0;
  srv->sch.exec = &sigmo_server::run;
  srv->sch.user = srv;
  *out = server_start(srv, num_workers, options);
  return SIGMO_OK;
}

int sigmo_server_open_schedule(uint32_t num_workers, const sigmo_serve_options* options, sigmo_serve_executor executor, void* user,
                               sigmo_server** out) {
  if (!options || !executor || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (num_workers == 0 || num_workers > 16) return sigmo_err::set(SIGMO_ERR_INVALID, "1..16 workers");
  auto* srv = new sigmo_server();
  srv->device = false;
  srv->sch.exec = executor;
  srv->sch.user = user;
  *out = server_start(srv, num_workers, options);
  return SIGMO_OK;
}

int sigmo_server_submit_count(sigmo_server* srv, uint32_t num_query_graphs, uint64_t* request_id) {
  if (!srv || !request_id) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (srv->device) return sigmo_err::set(SIGMO_ERR_INVALID, "a device server takes query graphs (sigmo_server_submit)");
  if (num_query_graphs == 0) return sigmo_err::set(SIGMO_ERR_INVALID, "a request must carry at least one query graph");
  return server_push(srv, num_query_graphs, uint64_t(num_query_graphs) * srv->sch.opt.bytes_per_graph, nullptr, request_id);
}

int sigmo_server_submit(sigmo_server* srv, const sigmo_host_graphs* query_graphs, uint64_t* request_id) {
  if (!srv || !query_graphs || !request_id) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (!srv->device) return sigmo_err::set(SIGMO_ERR_INVALID, "a scheduling server takes counts (sigmo_server_submit_count)");
  const uint32_t g = query_graphs->num_graphs();
  if (g == 0) return sigmo_err::set(SIGMO_ERR_INVALID, "a request must carry at least one query graph");
  const std::string bad = sigmo_host::validate(*query_graphs);
  if (!bad.empty()) return sigmo_err::set(SIGMO_ERR_INVALID, bad);
  uint64_t request_bytes = uint64_t(g) * srv->sch.opt.bytes_per_graph;
  if (srv->price_each) {  // the workspace of a batch holding just this request
    const int st = sigmo_batch_workspace_bytes(srv->workers[0], g, query_graphs->total_nodes(), &request_bytes);
    if (st != SIGMO_OK) return st;
  }
  return server_push(srv, g, request_bytes, std::make_unique<sigmo_host_graphs>(*query_graphs), request_id);
}

int sigmo_server_poll(sigmo_server* srv, uint64_t request_id, double timeout_ms, sigmo_serve_result* result) {
  if (!srv || !result) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  std::unique_lock<std::mutex> lk(srv->sch.mu);
  if (request_id >= srv->sch.size()) return sigmo_err::set(SIGMO_ERR_INVALID, "unknown request id");
  const uint32_t r = static_cast<uint32_t>(request_id);
  auto settled = [&] { return srv->sch.res[r].status != SIGMO_SERVE_PENDING || srv->sch.failed; };
  if (timeout_ms < 0)
    srv->sch.cv.wait(lk, settled);
  else if (timeout_ms > 0)
    srv->sch.cv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), settled);
  *result = srv->sch.res[r];
  if (srv->sch.failed && result->status == SIGMO_SERVE_PENDING) return sigmo_err::set(srv->sch.err_code, srv->sch.err_msg);
  return SIGMO_OK;
}

int sigmo_server_close(sigmo_server* srv, sigmo_serve_summary* summary) {
  if (!srv) return SIGMO_OK;
  {
    std::lock_guard<std::mutex> lk(srv->sch.mu);
    srv->sch.closing = true;
    srv->sch.cv.notify_all();
  }
  for (auto& t : srv->threads) t.join();
  int st = SIGMO_OK;
  if (srv->sch.failed) st = sigmo_err::set(srv->sch.err_code, srv->sch.err_msg);
  else if (summary) srv->sch.summarize(summary);
  for (sigmo_graph* g : srv->dev) sigmo_graph_free(g);
  delete srv;
  return st;
}

}  // extern "C"
