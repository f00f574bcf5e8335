// Thread-sanitizer harness for the serving scheduler (sigmo_serve_schedule, and the online
// server: sigmo_server_open_schedule / submit_count / poll / close) with a stand-in executor.
// Built with -fsanitize=thread together with serve.cpp and host_graphs.cpp (sigmo_amd/build.py,
// target `tsan`); the device entry points serve.cpp also references come from the product
// library, and this harness never reaches them.  Exits 0 when every request is served with the
// executor's statistics; the sanitizer reports any data race on stderr and fails the run.
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <thread>
#include <vector>

#include "../../include/sigmo_hip.h"

namespace {

struct Recorder {
  std::atomic<int> batches{0};
  std::atomic<int> active{0};
  std::atomic<int> max_active{0};
};

int execute(void* user, uint32_t worker, const uint32_t* requests, uint32_t count, sigmo_request_stats* out, uint64_t* gmcr_pairs) {
  auto* rec = static_cast<Recorder*>(user);
  const int now = ++rec->active;
  int seen = rec->max_active.load();
  while (now > seen && !rec->max_active.compare_exchange_weak(seen, now)) {
  }
  rec->batches++;
  std::this_thread::sleep_for(std::chrono::microseconds(200 + 50 * worker));
  for (uint32_t i = 0; i < count; ++i) {
    out[i].matched_pairs = requests[i];
    out[i].embeddings = 2ull * requests[i];
    out[i].hash = 0x9E3779B97F4A7C15ull * (requests[i] + 1);
  }
  *gmcr_pairs = count;
  --rec->active;
  return SIGMO_OK;
}

int fail(const char* what) {
  std::fprintf(stderr, "serve_schedule_race: %s (%s)\n", what, sigmo_last_error());
  return 1;
}

}  // namespace

int main() {
  // scheduled form: 400 requests of 1..7 graphs arriving 0.1 ms apart, 4 workers, a budget
  const uint32_t n = 400;
  std::vector<uint32_t> rgo(n + 1, 0);
  std::vector<double> arr(n);
  for (uint32_t r = 0; r < n; ++r) {
    rgo[r + 1] = rgo[r] + 1 + r % 7;
    arr[r] = 1e-4 * r;
  }
  sigmo_serve_options opt;
  std::memset(&opt, 0, sizeof(opt));
  opt.max_batch_requests = 16;
  opt.memory_budget = 64;
  opt.bytes_per_graph = 1;
  opt.bytes_per_batch = 1;
  opt.find_first = 1;
  std::vector<sigmo_serve_result> res(n);
  sigmo_serve_summary sum;
  Recorder rec;
  if (sigmo_serve_schedule(4, rgo.data(), arr.data(), n, &opt, execute, &rec, res.data(), &sum) != SIGMO_OK) return fail("schedule");
  for (uint32_t r = 0; r < n; ++r)
    if (res[r].status != SIGMO_SERVE_DONE || res[r].matched_pairs != r) return fail("scheduled request result");
  if (sum.done != n) return fail("scheduled summary");

  // online form: 4 client threads submit 100 requests each while 3 workers run, then poll
  sigmo_server* srv = nullptr;
  Recorder rec2;
  if (sigmo_server_open_schedule(3, &opt, execute, &rec2, &srv) != SIGMO_OK) return fail("open");
  std::vector<uint64_t> ids(400);
  std::vector<std::thread> clients;
  std::atomic<int> errors{0};
  for (int c = 0; c < 4; ++c)
    clients.emplace_back([&, c] {
      for (int i = 0; i < 100; ++i) {
        if (sigmo_server_submit_count(srv, 1 + (i % 5), &ids[c * 100 + i]) != SIGMO_OK) errors++;
        if (i % 10 == 0) std::this_thread::sleep_for(std::chrono::microseconds(100));
      }
    });
  for (auto& t : clients) t.join();
  for (uint32_t i = 0; i < ids.size(); ++i) {
    sigmo_serve_result r;
    if (sigmo_server_poll(srv, ids[i], -1.0, &r) != SIGMO_OK || r.status != SIGMO_SERVE_DONE || r.matched_pairs != ids[i]) errors++;
  }
  sigmo_serve_summary s2;
  if (sigmo_server_close(srv, &s2) != SIGMO_OK) return fail("close");
  if (errors.load() || s2.done != ids.size()) return fail("online results");
  std::printf("serve_schedule_race: ok (%d + %d batches, up to %d and %d in flight)\n", rec.batches.load(), rec2.batches.load(),
              rec.max_active.load(), rec2.max_active.load());
  return 0;
}
