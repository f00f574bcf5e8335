// Streaming: a molecule set larger than one device pass, run in chunks of molecules.
//
// The reference's multi-node run gives each MPI rank a byte range of the data file and runs
// the whole driver on it (library/src/sigmo_mpi.cpp:19-80, 110-159); here one GPU takes the
// role of a rank whose shard does not fit the candidate memory in one pass.  The shard stays in
// host memory; chunk i+1 is uploaded (its ids rebased on the device) by a loader thread on its
// own stream (its own context) while chunk i runs through the plain per-phase pipeline or the
// batched form on the caller's stream.  Two device graph buffers alternate; a buffer is reloaded only after the
// chunk that used it has returned its results.  With W > 1 workers (batched form) chunk i runs on
// worker i mod W, each with its own context (stream), database and batch, so one chunk's kernels
// fill the device while another chunk's host synchronisations and tail-bound join passes run;
// W + 1 graph buffers rotate, and the chunks' results are handed to the callback in chunk order.  Every chunk's GMCR and counts are those of a
// single run over that chunk (molecule ids chunk-local), so the concatenation over chunks equals
// a single run over the whole shard.
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
int graph_upload_range(sigmo_ctx* ctx, const sigmo_host_graphs* host, uint32_t first, uint32_t count, uint32_t reserve_G, uint64_t reserve_N,
                       uint64_t reserve_E, sigmo_graph** g);
int pipeline_rebind(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, uint64_t reserve_N, sigmo_pipeline** p);
int db_rebind(sigmo_db* db, const sigmo_graph* data);
int ctx_device(const sigmo_ctx* ctx);
int ctx_merge_timings(sigmo_ctx* dst, sigmo_ctx* src);
}  // namespace sigmo_internal

namespace {

using Clock = std::chrono::steady_clock;
double ms_since(Clock::time_point t) { return std::chrono::duration<double, std::milli>(Clock::now() - t).count(); }

struct Loader {
  sigmo_ctx* ctx = nullptr;  // the loader's own stream
  const sigmo_host_graphs* data = nullptr;
  uint32_t chunk = 0, chunks = 0, G = 0;
  uint64_t max_N = 0, max_E = 0;  // the largest chunk: buffers are sized for it once
  static constexpr int kMaxSlots = 5;
  int slots = 2;
  sigmo_graph* slot[kMaxSlots] = {};
  int loaded[kMaxSlots] = {-1, -1, -1, -1, -1};  // chunk held by each slot, -1 if free
  std::mutex mu;
  std::condition_variable cv;
  bool stop = false;
  int err = SIGMO_OK;
  std::string msg;
  double load_ms = 0;

  uint32_t first(uint32_t i) const { return i * chunk; }
  uint32_t count(uint32_t i) const { return std::min<uint32_t>(chunk, G - first(i)); }

  // A pass starts with both slots free.
  void reset() {
    for (int& l : loaded) l = -1;
    stop = false;
    err = SIGMO_OK;
    msg.clear();
    load_ms = 0;
  }

  // Upload chunk i of the host shard into `slot[i % slots]` (ids rebased on the device).
  int load(uint32_t i) {
    const auto t = Clock::now();
    const int st = sigmo_internal::graph_upload_range(ctx, data, first(i), count(i), chunk, max_N, max_E, &slot[i % slots]);
    load_ms += ms_since(t);
    return st;
  }

  void run() {
    for (uint32_t i = 0; i < chunks; ++i) {
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return stop || loaded[i % slots] < 0; });
        if (stop) return;
      }
      const int st = load(i);
      std::lock_guard<std::mutex> lk(mu);
      if (st != SIGMO_OK) {
        err = st;
        msg = sigmo_last_error();
        stop = true;
      } else {
        loaded[i % slots] = static_cast<int>(i);
      }
      cv.notify_all();
      if (stop) return;
    }
  }
};

}  // namespace

// A streaming session: the chunk plan, the loader's context and upload slots, and the form's
// device objects (plain pipeline, or database and batch), kept from one pass to the next.
struct sigmo_stream {
  static constexpr int kMaxWorkers = 4;
  sigmo_ctx* ctx = nullptr;
  const sigmo_graph* query = nullptr;
  sigmo_stream_options opt{};
  Loader L;
  sigmo_pipeline* pipe = nullptr;
  sigmo_db* db = nullptr;
  sigmo_batch* batch = nullptr;
  std::vector<uint32_t> offsets, indices;
  std::vector<int32_t> counts;
  // workers 1.. of a multi-worker batched pass (worker 0 uses ctx, db and batch above)
  int workers = 1;
  sigmo_ctx* wctx[kMaxWorkers] = {};
  sigmo_db* wdb[kMaxWorkers] = {};
  sigmo_batch* wbatch[kMaxWorkers] = {};
  ~sigmo_stream() {
    sigmo_batch_free(batch);
    sigmo_db_free(db);
    for (int w = 1; w < kMaxWorkers; ++w) {
      sigmo_batch_free(wbatch[w]);
      sigmo_db_free(wdb[w]);
      if (wctx[w]) sigmo_ctx_destroy(wctx[w]);
    }
    sigmo_pipeline_free(pipe);
    for (sigmo_graph* g : L.slot) sigmo_graph_free(g);
    sigmo_ctx_destroy(L.ctx);
  }
};

extern "C" {

int sigmo_stream_open(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_host_graphs* data, const sigmo_stream_options* opt,
                      sigmo_stream** out) {
  if (!ctx || !query || !data || !opt || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (opt->form != SIGMO_FORM_PLAIN && opt->form != SIGMO_FORM_BATCHED) return sigmo_err::set(SIGMO_ERR_INVALID, "unknown form");
  const uint32_t G = data->num_graphs();
  if (G == 0) return sigmo_err::set(SIGMO_ERR_INVALID, "empty molecule shard");
  auto* st = new sigmo_stream();
  st->ctx = ctx;
  st->query = query;
  st->opt = *opt;
  Loader& L = st->L;
  L.data = data;
  L.G = G;
  L.chunk = opt->chunk_molecules ? std::min(opt->chunk_molecules, G) : G;
  L.chunks = (G + L.chunk - 1) / L.chunk;
  for (uint32_t i = 0; i < L.chunks; ++i) {
    const uint32_t lo = data->graph_offsets[L.first(i)], hi = data->graph_offsets[L.first(i) + L.count(i)];
    L.max_N = std::max<uint64_t>(L.max_N, hi - lo);
    L.max_E = std::max<uint64_t>(L.max_E, data->row_offsets[hi] - data->row_offsets[lo]);
  }
  const int rc = sigmo_ctx_create(sigmo_internal::ctx_device(ctx), &L.ctx);
  if (rc != SIGMO_OK) {
    delete st;
    return rc;
  }
  if (opt->form == SIGMO_FORM_BATCHED && opt->overlap && opt->workers > 1 && L.chunks > 1) {
    st->workers = static_cast<int>(std::min<uint32_t>({opt->workers, static_cast<uint32_t>(sigmo_stream::kMaxWorkers), L.chunks}));
    L.slots = st->workers + 1;
    for (int w = 1; w < st->workers; ++w) {
      const int r2 = sigmo_ctx_create(sigmo_internal::ctx_device(ctx), &st->wctx[w]);
      if (r2 != SIGMO_OK) {
        delete st;
        return r2;
      }
    }
  }
  *out = st;
  return SIGMO_OK;
}

namespace {
// A multi-worker batched pass: worker w runs chunks w, w+W, ... on its own context.
int stream_pass_workers(sigmo_stream* S, sigmo_stream_chunk_fn on_chunk, void* user, sigmo_stream_result* res) {
  Loader& L = S->L;
  const sigmo_stream_options* opt = &S->opt;
  const int W = S->workers;
  std::mutex mu;  // results, errors and callback order
  std::condition_variable cv;
  uint32_t next_cb = 0;
  bool failed = false;
  int err = SIGMO_OK;
  std::string msg;
  double busy_ms[sigmo_stream::kMaxWorkers] = {}, wait_ms[sigmo_stream::kMaxWorkers] = {};
  auto stop_loader = [&] {
    std::lock_guard<std::mutex> lk(L.mu);
    L.stop = true;
    L.cv.notify_all();
  };
  auto fail = [&](int st) {
    {
      std::lock_guard<std::mutex> lk(mu);
      if (!failed) {
        failed = true;
        err = st;
        msg = sigmo_last_error();
      }
      cv.notify_all();
    }
    stop_loader();  // a worker waiting for a chunk, or the loader for a buffer, must not wait on this one
  };
  auto work = [&](int w) {
    sigmo_ctx* ctx = w == 0 ? S->ctx : S->wctx[w];
    sigmo_db*& db = w == 0 ? S->db : S->wdb[w];
    sigmo_batch*& batch = w == 0 ? S->batch : S->wbatch[w];
    std::vector<uint32_t> offsets, indices;
    std::vector<int32_t> counts;
    for (uint32_t i = static_cast<uint32_t>(w); i < L.chunks; i += static_cast<uint32_t>(W)) {
      const int s = static_cast<int>(i % static_cast<uint32_t>(L.slots));
      {
        const auto tw = Clock::now();
        std::unique_lock<std::mutex> lk(L.mu);
        L.cv.wait(lk, [&] { return L.stop || L.loaded[s] == static_cast<int>(i); });
        wait_ms[w] += ms_since(tw);
        if (L.loaded[s] != static_cast<int>(i)) {
          lk.unlock();
          fail(sigmo_err::set(L.err != SIGMO_OK ? L.err : SIGMO_ERR_INVALID, "chunk load: " + L.msg));
          return;
        }
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        if (failed) return;
      }
      const sigmo_graph* g = L.slot[s];
      const auto tc = Clock::now();
      int st = SIGMO_OK;
      if (!db) {
        st = sigmo_db_prepare(ctx, g, opt->refinement_steps, &db);
        if (st == SIGMO_OK) st = sigmo_batch_create(ctx, db, &batch);
      } else {
        st = sigmo_internal::db_rebind(db, g);
      }
      sigmo_batch_result br;
      if (st == SIGMO_OK) st = sigmo_batch_run(batch, S->query, opt->find_first, &br);
      if (st == SIGMO_OK && on_chunk) {
        const uint32_t n = L.count(i);
        offsets.resize(uint64_t(n) + 1);
        indices.resize(br.gmcr_pairs);
        counts.resize(br.gmcr_pairs);
        st = sigmo_batch_pairs(batch, offsets.data(), offsets.size(), indices.data(), counts.data(), br.gmcr_pairs);
      }
      if (st != SIGMO_OK) {
        fail(st);
        return;
      }
      {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return failed || next_cb == i; });  // chunks report in order
        if (failed) return;
        res->num_matches += br.num_matches;
        res->matched_pairs += br.matched_pairs;
        res->gmcr_pairs += br.gmcr_pairs;
        res->dfs_steps += br.dfs_steps;
        if (on_chunk && on_chunk(user, i, L.first(i), L.count(i), offsets.data(), indices.data(), counts.data(), br.gmcr_pairs) != 0) {
          failed = true;
          err = sigmo_err::set(SIGMO_ERR_INVALID, "chunk callback failed on chunk " + std::to_string(i));
          msg = sigmo_last_error();
          cv.notify_all();
          lk.unlock();
          stop_loader();
          return;
        }
        res->chunks = i + 1;
        next_cb = i + 1;
        cv.notify_all();
      }
      busy_ms[w] += ms_since(tc);
      std::lock_guard<std::mutex> lk(L.mu);
      L.loaded[s] = -1;  // the chunk's results are on the host: the loader may refill this buffer
      L.cv.notify_all();
    }
  };
  std::thread loader([&L] { L.run(); });
  std::vector<std::thread> threads;
  for (int w = 0; w < W; ++w) threads.emplace_back(work, w);
  for (auto& t : threads) t.join();
  stop_loader();
  loader.join();
  for (int w = 0; w < W; ++w) {
    res->compute_ms = std::max(res->compute_ms, busy_ms[w]);
    res->load_wait_ms = std::max(res->load_wait_ms, wait_ms[w]);
  }
  if (failed) return sigmo_err::set(err, msg);
  for (int w = 1; w < W; ++w) {
    const int st = sigmo_internal::ctx_merge_timings(S->ctx, S->wctx[w]);
    if (st != SIGMO_OK) return st;
  }
  return SIGMO_OK;
}
}  // namespace

int sigmo_stream_pass(sigmo_stream* S, sigmo_stream_chunk_fn on_chunk, void* user, sigmo_stream_result* res) {
  if (!S || !res) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  std::memset(res, 0, sizeof(*res));
  sigmo_ctx* ctx = S->ctx;
  const sigmo_graph* query = S->query;
  const sigmo_stream_options* opt = &S->opt;
  Loader& L = S->L;
  L.reset();
  int st = SIGMO_OK;
  const auto t0 = Clock::now();
  if (S->workers > 1) {
    st = stream_pass_workers(S, on_chunk, user, res);
    res->wall_ms = ms_since(t0);
    res->load_ms = L.load_ms;
    return st;
  }
  std::thread loader;
  if (opt->overlap) loader = std::thread([&L] { L.run(); });

  for (uint32_t i = 0; i < L.chunks && st == SIGMO_OK; ++i) {
    const int s = static_cast<int>(i % 2);
    if (opt->overlap) {
      const auto tw = Clock::now();
      std::unique_lock<std::mutex> lk(L.mu);
      L.cv.wait(lk, [&] { return L.stop || L.loaded[s] == static_cast<int>(i); });
      res->load_wait_ms += ms_since(tw);
      if (L.loaded[s] != static_cast<int>(i)) {
        st = sigmo_err::set(L.err != SIGMO_OK ? L.err : SIGMO_ERR_INVALID, "chunk load: " + L.msg);
        break;
      }
    } else {
      const auto tw = Clock::now();
      st = L.load(i);
      res->load_wait_ms += ms_since(tw);
      if (st != SIGMO_OK) break;
    }
    const sigmo_graph* g = L.slot[s];
    const auto tc = Clock::now();
    uint64_t P = 0;
    if (opt->form == SIGMO_FORM_PLAIN) {
      st = sigmo_internal::pipeline_rebind(ctx, query, g, L.max_N, &S->pipe);
      sigmo_pipeline_result pr;
      if (st == SIGMO_OK) st = sigmo_pipeline_execute(S->pipe, opt->refinement_steps, opt->find_first, &pr);
      if (st == SIGMO_OK) {
        res->num_matches += pr.num_matches;
        res->matched_pairs += pr.matched_pairs;
        res->gmcr_pairs += pr.gmcr_pairs;
        res->dfs_steps += pr.dfs_steps;
        P = pr.gmcr_pairs;
      }
    } else {
      if (!S->db) {
        st = sigmo_db_prepare(ctx, g, opt->refinement_steps, &S->db);
        if (st == SIGMO_OK) st = sigmo_batch_create(ctx, S->db, &S->batch);
      } else {
        st = sigmo_internal::db_rebind(S->db, g);
      }
      sigmo_batch_result br;
      if (st == SIGMO_OK) st = sigmo_batch_run(S->batch, query, opt->find_first, &br);
      if (st == SIGMO_OK) {
        res->num_matches += br.num_matches;
        res->matched_pairs += br.matched_pairs;
        res->gmcr_pairs += br.gmcr_pairs;
        res->dfs_steps += br.dfs_steps;
        P = br.gmcr_pairs;
      }
    }
    if (st == SIGMO_OK && on_chunk) {
      const uint32_t n = L.count(i);
      S->offsets.resize(uint64_t(n) + 1);
      S->indices.resize(P);
      S->counts.resize(P);
      st = opt->form == SIGMO_FORM_PLAIN
               ? sigmo_pipeline_pairs(S->pipe, S->offsets.data(), S->offsets.size(), S->indices.data(), S->counts.data(), P)
               : sigmo_batch_pairs(S->batch, S->offsets.data(), S->offsets.size(), S->indices.data(), S->counts.data(), P);
      if (st == SIGMO_OK && on_chunk(user, i, L.first(i), n, S->offsets.data(), S->indices.data(), S->counts.data(), P) != 0)
        st = sigmo_err::set(SIGMO_ERR_INVALID, "chunk callback failed on chunk " + std::to_string(i));
    }
    res->compute_ms += ms_since(tc);
    res->chunks = i + 1;
    std::lock_guard<std::mutex> lk(L.mu);
    L.loaded[s] = -1;  // results are on the host: the loader may refill this buffer
    L.cv.notify_all();
  }
  std::string err_msg = st == SIGMO_OK ? std::string() : std::string(sigmo_last_error());
  if (opt->overlap) {
    {
      std::lock_guard<std::mutex> lk(L.mu);
      L.stop = true;
      L.cv.notify_all();
    }
    loader.join();
  }
  res->wall_ms = ms_since(t0);
  res->load_ms = L.load_ms;
  if (st != SIGMO_OK) return sigmo_err::set(st, err_msg);
  return SIGMO_OK;
}

int sigmo_stream_close(sigmo_stream* S) {
  delete S;
  return SIGMO_OK;
}

int sigmo_stream_run(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_host_graphs* data, const sigmo_stream_options* opt,
                     sigmo_stream_chunk_fn on_chunk, void* user, sigmo_stream_result* res) {
  if (!res) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  std::memset(res, 0, sizeof(*res));
  sigmo_stream* S = nullptr;
  int st = sigmo_stream_open(ctx, query, data, opt, &S);
  if (st != SIGMO_OK) return st;
  st = sigmo_stream_pass(S, on_chunk, user, res);
  const std::string err_msg = st == SIGMO_OK ? std::string() : std::string(sigmo_last_error());
  sigmo_stream_close(S);
  if (st != SIGMO_OK) return sigmo_err::set(st, err_msg);
  return SIGMO_OK;
}

}  // extern "C"
