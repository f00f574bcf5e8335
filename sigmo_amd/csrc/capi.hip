// C ABI over the gfx950 kernels: context, device graphs, signatures, candidates, GMCR, join
// and the plain per-phase pipeline.  Mirrors the reference's library/include classes
// (paths relative to the SIGMo repository); every device buffer is owned by a handle and
// every entry point reports failures as status codes plus sigmo_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sigmo_hip.h"
#include "host_graphs.hpp"
#include "kernels.hpp"

namespace {

#define SIGMO_HIP_TRY(expr)                                                                          \
  do {                                                                                               \
    hipError_t _e = (expr);                                                                          \
    if (_e != hipSuccess)                                                                            \
      return sigmo_err::set(_e == hipErrorOutOfMemory ? SIGMO_ERR_NOMEM : SIGMO_ERR_HIP,             \
                            std::string(#expr) + ": " + hipGetErrorString(_e));                      \
  } while (0)

#define SIGMO_TRY(expr)       \
  do {                        \
    int _s = (expr);          \
    if (_s != SIGMO_OK) return _s; \
  } while (0)

template <typename T>
int dev_alloc(T** p, uint64_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return sigmo_err::set(SIGMO_ERR_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed: " + hipGetErrorString(e));
  }
  return SIGMO_OK;
}

template <typename T>
void dev_free(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

}  // namespace

// ---- objects -----------------------------------------------------------------------------------

struct sigmo_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 256;
  std::string name;
  uint64_t global_mem = 0;
  struct Timed {
    hipEvent_t a, b;
    int phase, kernel;
    double bytes;
    uint64_t seq;
    bool shared_a;  // `a` is the previous phase's end event (returned to the pool with it)
  };
  uint64_t launch_seq = 0;          // sequence number of the next timed launch
  uint64_t launch_base = 0;         // sequence number of launch_ms[0]
  std::vector<double> launch_ms;    // per-launch kernel ms, indexed by seq - launch_base
  std::vector<Timed> pending;
  double accum_ms[SIGMO_PHASE_COUNT] = {0, 0, 0, 0, 0};
  uint64_t k_launches[SIGMO_KERNEL_COUNT] = {};
  double k_ms[SIGMO_KERNEL_COUNT] = {};
  double k_bytes[SIGMO_KERNEL_COUNT] = {};
  std::vector<hipEvent_t> pool;
  bool capturing = false;  // recording a graph: no timing events, no host synchronisation
  // Back-to-back phases (the plain pipeline's, which enqueue no other work and no host wait
  // between them) share one event at each boundary: the next phase starts at the previous
  // phase's end event.
  bool chain = false;
  hipEvent_t chain_end = nullptr;
  unsigned long long* stats = nullptr;       // join counters [8]: matches, matched pairs, cursor, overflow, steps
  unsigned long long* stats_host = nullptr;  // pinned mirror

  int event(hipEvent_t* e) {
    if (!pool.empty()) {
      *e = pool.back();
      pool.pop_back();
      return SIGMO_OK;
    }
    // timing only: the events are read after a stream synchronisation, so they need no
    // system-scope fence (whose cache writeback and invalidation would slow the next phase)
    SIGMO_HIP_TRY(hipEventCreateWithFlags(e, hipEventDisableSystemFence));
    return SIGMO_OK;
  }
  int collect() {
    SIGMO_HIP_TRY(hipStreamSynchronize(stream));
    for (auto& t : pending) {
      float ms = 0;
      SIGMO_HIP_TRY(hipEventElapsedTime(&ms, t.a, t.b));
      accum_ms[t.phase] += ms;
      if (t.seq >= launch_base) {
        const uint64_t i = t.seq - launch_base;
        if (launch_ms.size() <= i) launch_ms.resize(i + 1, 0.0);
        launch_ms[i] = ms;
      }
      if (t.kernel >= 0) {
        k_launches[t.kernel] += 1;
        k_ms[t.kernel] += ms;
        k_bytes[t.kernel] += t.bytes;
      }
      if (!t.shared_a) pool.push_back(t.a);
      pool.push_back(t.b);
    }
    pending.clear();
    chain_end = nullptr;
    constexpr size_t kKeep = 1u << 20;  // bounded history for long-running contexts
    if (launch_ms.size() > 2 * kKeep) {
      const size_t drop = launch_ms.size() - kKeep;
      launch_ms.erase(launch_ms.begin(), launch_ms.begin() + drop);
      launch_base += drop;
    }
    return SIGMO_OK;
  }
};

// Records hipEvents around the launches of one phase (utils.hpp:127-161 BatchedEvent).
// `kernel` (SIGMO_KERNEL_*) and `bytes` (algorithmic HBM bytes of the launch) feed
// sigmo_kernel_stats; kernel < 0 records phase time only.
struct PhaseScope {
  sigmo_ctx* c;
  int phase, kernel;
  double bytes;
  hipEvent_t a = nullptr;
  bool shared_a = false;
  int status = SIGMO_OK;
  PhaseScope(sigmo_ctx* ctx, int ph, int k = -1, double by = 0) : c(ctx), phase(ph), kernel(k), bytes(by) {
    if (c->capturing) return;  // a graph records no timing events
    if (c->chain && c->chain_end) {
      a = c->chain_end;
      shared_a = true;
      c->chain_end = nullptr;
      return;
    }
    status = c->event(&a);
    if (status == SIGMO_OK && hipEventRecord(a, c->stream) != hipSuccess) status = SIGMO_ERR_HIP;
  }
  int close() {
    if (c->capturing) return SIGMO_OK;
    if (status != SIGMO_OK) return status;
    hipEvent_t b;
    SIGMO_TRY(c->event(&b));
    SIGMO_HIP_TRY(hipEventRecord(b, c->stream));
    c->pending.push_back({a, b, phase, kernel, bytes, c->launch_seq++, shared_a});
    if (c->chain) c->chain_end = b;
    a = nullptr;
    return SIGMO_OK;
  }
  ~PhaseScope() {
    if (a && !shared_a) c->pool.push_back(a);
  }
};

struct sigmo_graph {
  sigmo_ctx* ctx = nullptr;
  uint32_t G = 0;
  uint64_t N = 0, E = 0;
  uint32_t max_nodes = 0, min_nodes = 0;
  std::vector<uint32_t> host_go;
  uint32_t *go = nullptr, *ro = nullptr, *ci = nullptr, *node_graph = nullptr;
  uint8_t *nl = nullptr, *el = nullptr;
  uint64_t *adj = nullptr, *fold = nullptr, *lplanes = nullptr, *lmask = nullptr;
  sigmo_k::LabelSlots slots{};
  uint64_t cap_G = 0, cap_N = 0, cap_E = 0;  // device buffer capacities (a re-upload reuses them)

  sigmo_k::GraphDev dev() const {
    sigmo_k::GraphDev g;
    g.go = go;
    g.ro = ro;
    g.ci = ci;
    g.nl = nl;
    g.node_graph = node_graph;
    g.adj = adj;
    g.fold = fold;
    g.lplanes = lplanes;
    g.lmask = lmask;
    g.G = G;
    g.N = N;
    return g;
  }
  uint64_t csr_bytes() const { return N * 1 + (uint64_t(G) + 1) * 4 + (N + 1) * 4 + E * 4 + E * 1; }
  uint64_t device_bytes() const {
    return csr_bytes() + N * 4 + N * 24 + ((N + 63) / 64) * 64 + uint64_t(G) * 128;
  }
  ~sigmo_graph() {
    dev_free(go);
    dev_free(ro);
    dev_free(ci);
    dev_free(node_graph);
    dev_free(nl);
    dev_free(el);
    dev_free(adj);
    dev_free(fold);
    dev_free(lplanes);
    dev_free(lmask);
  }
};

// A batch of adjacency-matrix graphs (DeviceBatchedAMGraph, graph.hpp:114-148).
struct sigmo_am_graph {
  sigmo_ctx* ctx = nullptr;
  uint32_t G = 0;
  uint64_t N = 0, words = 0;
  uint32_t *go = nullptr, *nn = nullptr;
  uint64_t* adj = nullptr;
  uint8_t* nl = nullptr;

  sigmo_k::AMGraphDev dev() const {
    sigmo_k::AMGraphDev g;
    g.go = go;
    g.nn = nn;
    g.adj = adj;
    g.nl = nl;
    g.G = G;
    g.N = N;
    return g;
  }
  ~sigmo_am_graph() {
    dev_free(go);
    dev_free(nn);
    dev_free(adj);
    dev_free(nl);
  }
};

// Per-scope refinement history.  The reachables always equal ball(M) minus the start
// position for the largest view size M refined so far (balls only grow), so when the last
// BFS layer is kept (`frontier` valid) a call with view_size M+1 is one BFS step.
struct RefineState {
  uint32_t max_view = 0;
  bool frontier_valid = false;
  bool tainted = false;  // refined with more than one graph: always the full BFS
  const sigmo_graph* graph = nullptr;
  void reset() { *this = RefineState(); }
};

struct sigmo_signatures {
  sigmo_ctx* ctx = nullptr;
  uint64_t D = 0, Q = 0;
  uint64_t capD = 0;  // data-side capacity in nodes (a rebind to a smaller data graph reuses it)
  uint64_t *dsig = nullptr, *qsig = nullptr, *dreach = nullptr, *qreach = nullptr, *dplanes = nullptr;
  uint64_t *dfront = nullptr, *qfront = nullptr;
  uint64_t* tmp = nullptr;  // ViewBased refinement: the signatures before the pass
  uint64_t cap_tmp = 0;
  RefineState dstate, qstate;
  ~sigmo_signatures() {
    dev_free(dsig);
    dev_free(qsig);
    dev_free(dreach);
    dev_free(qreach);
    dev_free(dplanes);
    dev_free(dfront);
    dev_free(qfront);
    dev_free(tmp);
  }
};

struct sigmo_candidates {
  sigmo_ctx* ctx = nullptr;
  uint64_t Q = 0, D = 0, W64 = 0;
  uint64_t cap_words = 0;  // allocated words (rows x W64 of the largest target so far)
  uint64_t* bits = nullptr;
  ~sigmo_candidates() { dev_free(bits); }
};

struct sigmo_gmcr {
  sigmo_ctx* ctx = nullptr;
  uint32_t Gq = 0, Gd = 0;
  uint64_t P = 0;
  bool built = false;
  uint32_t *offsets = nullptr, *counts = nullptr, *indices = nullptr, *pair_graph = nullptr;
  uint64_t *pair_bits = nullptr, *scan_tmp = nullptr, *total = nullptr;
  uint64_t cap_indices = 0, cap_pair_graph = 0, cap_counts_g = 0, cap_offsets = 0, cap_pair_bits = 0, cap_scan = 0;
  int32_t* pair_counts = nullptr;
  uint64_t cap_counts = 0;
  uint32_t* wide_pairs = nullptr;  // pairs of molecules with > 64 nodes, for the two-word lane join
  uint64_t cap_wide = 0;
  uint32_t* deep_pairs = nullptr;  // pairs of queries deeper than the lane join's first pass holds
  uint64_t cap_deep = 0;
  uint64_t last_wide = 0;  // their count in the last join
  // The plain pipeline hands the pair count to the join on the device: P then holds an upper
  // bound (Gq x Gd) until the join's synchronisation reads the exact count from total_host.
  bool deferred = false;
  unsigned long long* total_host = nullptr;  // pinned
  ~sigmo_gmcr() {
    if (total_host) (void)hipHostFree(total_host);
    dev_free(wide_pairs);
    dev_free(deep_pairs);
    dev_free(offsets);
    dev_free(counts);
    dev_free(indices);
    dev_free(pair_graph);
    dev_free(pair_bits);
    dev_free(scan_tmp);
    dev_free(total);
    dev_free(pair_counts);
  }
};

namespace {

int bind(sigmo_ctx* ctx) {
  if (!ctx) return sigmo_err::set(SIGMO_ERR_INVALID, "null context");
  SIGMO_HIP_TRY(hipSetDevice(ctx->device));
  return SIGMO_OK;
}

int check_scope(int scope) {
  if (scope != SIGMO_SCOPE_DATA && scope != SIGMO_SCOPE_QUERY) return sigmo_err::set(SIGMO_ERR_INVALID, "scope must be DATA or QUERY");
  return SIGMO_OK;
}

template <typename T>
int grow(T** p, uint64_t& cap, uint64_t need) {
  if (*p && cap >= need) return SIGMO_OK;
  dev_free(*p);
  cap = 0;
  SIGMO_TRY(dev_alloc(p, need));
  cap = need;
  return SIGMO_OK;
}

// The same for arrays sized by a result (GMCR pairs): a regrowth takes a quarter more, so a run
// over successive chunks reallocates rarely (hipFree waits for the whole device, including the
// next chunk's upload on the loader's stream).
template <typename T>
int grow_slack(T** p, uint64_t& cap, uint64_t need) {
  if (*p && cap >= need) return SIGMO_OK;
  return grow(p, cap, need + need / 4);
}

}  // namespace

// Accessors used by the batched form (batched.hip).
namespace sigmo_internal {
int ctx_info(sigmo_ctx* ctx, int* device, hipStream_t* stream, int* num_cus, unsigned long long** stats,
             unsigned long long** stats_host) {
  SIGMO_TRY(bind(ctx));
  *device = ctx->device;
  *stream = ctx->stream;
  *num_cus = ctx->num_cus;
  *stats = ctx->stats;
  *stats_host = ctx->stats_host;
  return SIGMO_OK;
}
// begin: start a timed region for (phase, kernel); end: close the region of `token`.
int ctx_timed(sigmo_ctx* ctx, int phase, int kernel, double bytes, bool begin, void** token) {
  if (begin) {
    auto* ps = new PhaseScope(ctx, phase, kernel, bytes);
    if (ps->status != SIGMO_OK) {
      const int st = ps->status;
      delete ps;
      return sigmo_err::set(st, "event record failed");
    }
    *token = ps;
    return SIGMO_OK;
  }
  auto* ps = static_cast<PhaseScope*>(*token);
  *token = nullptr;
  if (!ps) return SIGMO_OK;
  const int st = ps->close();
  delete ps;
  return st;
}
// Bytes known only once a kernel has finished (data-dependent outputs) join its accumulator.
void ctx_add_bytes(sigmo_ctx* ctx, int kernel, double bytes) {
  if (kernel >= 0 && kernel < SIGMO_KERNEL_COUNT) ctx->k_bytes[kernel] += bytes;
}
const sigmo_k::GraphDev graph_dev(const sigmo_graph* g) { return g->dev(); }
const std::vector<uint32_t>& graph_host_offsets(const sigmo_graph* g) { return g->host_go; }
sigmo_ctx* graph_ctx(const sigmo_graph* g) { return g->ctx; }
uint64_t graph_edges(const sigmo_graph* g) { return g->E; }
}  // namespace sigmo_internal

// ---- static device scratch of the allocation-free join ---------------------------------------
//
// sigmo_join_noalloc keeps its counters, per-pair counts and list of wide-molecule pairs in this
// file-scope array instead of allocated buffers.  Layout (64-bit words): [0, 16) join counters,
// then SIGMO_JOIN_NOALLOC_MAX_PAIRS int32 counts, then as many uint32 list entries: 16 MiB +
// 128 B in all.  Each device holds its own copy for the life of the process; host code reaches it
// only through the address hipGetSymbolAddress gives on that device (kept in g_scratch below),
// never through the array's name, which in host code is a stand-in, not a device address.
namespace {
constexpr uint64_t kScratchStatWords = 16;
constexpr uint64_t kScratchPairs = SIGMO_JOIN_NOALLOC_MAX_PAIRS;
constexpr uint64_t kScratchWords = kScratchStatWords + kScratchPairs;  // two 4-byte arrays of kScratchPairs
static __device__ __align__(256) uint64_t scratch[kScratchWords];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint64_t* g_scratch[kMaxDevices] = {};

// Once per device, with that device current: the device address of `scratch`.
int scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return SIGMO_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    (void)hipGetLastError();
    return sigmo_err::set(SIGMO_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  }
  g_scratch[device] = static_cast<uint64_t*>(p);
  return SIGMO_OK;
}

uint64_t* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  return g_scratch[device];
}
}  // namespace

extern "C" {

// ---- context ---------------------------------------------------------------------------------

int sigmo_ctx_create(int device, sigmo_ctx** out) {
  if (!out) return sigmo_err::set(SIGMO_ERR_INVALID, "null out");
  int n = 0;
  SIGMO_HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return sigmo_err::set(SIGMO_ERR_INVALID, "device index out of range");
  SIGMO_HIP_TRY(hipSetDevice(device));
  auto* c = new sigmo_ctx();
  c->device = device;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDeviceProperties(&prop, device);
  if (e == hipSuccess) {
    c->num_cus = prop.multiProcessorCount;
    c->name = prop.name;
    c->global_mem = prop.totalGlobalMem;
  }
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(reinterpret_cast<void**>(&c->stats), sigmo_k::kJoinStatWords * sizeof(unsigned long long));
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&c->stats_host), sigmo_k::kJoinStatWords * sizeof(unsigned long long), hipHostMallocDefault);
  if (e != hipSuccess) {
    sigmo_ctx_destroy(c);
    return sigmo_err::set(SIGMO_ERR_HIP, std::string("context setup: ") + hipGetErrorString(e));
  }
  const int st = scratch_setup(device);  // the allocation-free join's scratch on this device
  if (st != SIGMO_OK) {
    const std::string msg = sigmo_last_error();
    sigmo_ctx_destroy(c);
    return sigmo_err::set(st, msg);
  }
  *out = c;
  return SIGMO_OK;
}

int sigmo_ctx_destroy(sigmo_ctx* ctx) {
  if (!ctx) return SIGMO_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& t : ctx->pending) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : ctx->pool) (void)hipEventDestroy(e);
  if (ctx->stats) (void)hipFree(ctx->stats);
  if (ctx->stats_host) (void)hipHostFree(ctx->stats_host);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SIGMO_OK;
}

int sigmo_ctx_sync(sigmo_ctx* ctx) {
  SIGMO_TRY(bind(ctx));
  SIGMO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return SIGMO_OK;
}

int sigmo_phase_ms(sigmo_ctx* ctx, int phase, double* ms) {
  SIGMO_TRY(bind(ctx));
  if (phase < 0 || phase >= SIGMO_PHASE_COUNT || !ms) return sigmo_err::set(SIGMO_ERR_INVALID, "bad phase");
  SIGMO_TRY(ctx->collect());
  *ms = ctx->accum_ms[phase];
  return SIGMO_OK;
}

int sigmo_phase_reset(sigmo_ctx* ctx) {
  SIGMO_TRY(bind(ctx));
  SIGMO_TRY(ctx->collect());
  for (double& a : ctx->accum_ms) a = 0;
  for (int k = 0; k < SIGMO_KERNEL_COUNT; ++k) {
    ctx->k_launches[k] = 0;
    ctx->k_ms[k] = 0;
    ctx->k_bytes[k] = 0;
  }
  return SIGMO_OK;
}

int sigmo_kernel_stats(sigmo_ctx* ctx, int kernel, uint64_t* launches, double* ms, double* bytes) {
  SIGMO_TRY(bind(ctx));
  if (kernel < 0 || kernel >= SIGMO_KERNEL_COUNT) return sigmo_err::set(SIGMO_ERR_INVALID, "bad kernel id");
  SIGMO_TRY(ctx->collect());
  if (launches) *launches = ctx->k_launches[kernel];
  if (ms) *ms = ctx->k_ms[kernel];
  if (bytes) *bytes = ctx->k_bytes[kernel];
  return SIGMO_OK;
}

int sigmo_event_mark(sigmo_ctx* ctx, uint64_t* mark) {
  if (!ctx || !mark) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  *mark = ctx->launch_seq;
  return SIGMO_OK;
}

int sigmo_event_elapsed(sigmo_ctx* ctx, uint64_t from, uint64_t to, double* ms) {
  SIGMO_TRY(bind(ctx));
  if (!ms || to < from || to > ctx->launch_seq) return sigmo_err::set(SIGMO_ERR_INVALID, "bad launch range");
  SIGMO_TRY(ctx->collect());
  if (from < ctx->launch_base) return sigmo_err::set(SIGMO_ERR_INVALID, "launch range no longer retained");
  double t = 0;
  for (uint64_t i = from; i < to; ++i) t += ctx->launch_ms[i - ctx->launch_base];
  *ms = t;
  return SIGMO_OK;
}

int sigmo_device_info(sigmo_ctx* ctx, char* name, size_t name_len, uint64_t* global_mem) {
  if (!ctx) return sigmo_err::set(SIGMO_ERR_INVALID, "null context");
  if (name && name_len) {
    std::strncpy(name, ctx->name.c_str(), name_len - 1);
    name[name_len - 1] = '\0';
  }
  if (global_mem) *global_mem = ctx->global_mem;
  return SIGMO_OK;
}

// ---- graphs ----------------------------------------------------------------------------------

}  // extern "C"

namespace {

// A run of consecutive graphs of a host batch: its arrays are sub-ranges of the batch's, with
// ids still global to the batch (node ids from node_base, entry offsets from col_base).
struct HostView {
  const uint32_t* go;  // G + 1 entries
  uint32_t G;
  const uint32_t* ro;  // N + 1 entries
  uint64_t N;
  const uint32_t* ci;  // E entries
  uint64_t E;
  const uint8_t* nl;
  const uint8_t* el;
  uint32_t node_base, col_base;
};

HostView whole(const sigmo_host_graphs* h) {
  return HostView{h->graph_offsets.data(), h->num_graphs(), h->row_offsets.data(), h->total_nodes(), h->column_indices.data(),
                  h->total_edges(), h->node_labels.data(), h->edge_labels.data(), 0, 0};
}

// Fill `g` from a host view: validate, grow the device buffers where the new batch exceeds them,
// copy the CSR arrays (rebasing ids on the device), derive the adjacency words and label planes.
int graph_fill(sigmo_graph* g, const HostView& h, uint32_t reserve_G = 0, uint64_t reserve_N = 0, uint64_t reserve_E = 0) {
  const uint32_t G = h.G;
  const uint64_t N = h.N, E = h.E;
  if (G == 0) return sigmo_err::set(SIGMO_ERR_INVALID, "empty graph batch");
  if (N >= (1ull << 32) || E >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "batch exceeds u32 node/edge ids");
  uint32_t mx = 0, mn = UINT32_MAX;
  for (uint32_t i = 0; i < G; ++i) {
    const uint32_t n = h.go[i + 1] - h.go[i];
    mx = std::max(mx, n);
    mn = std::min(mn, n);
  }
  if (mn == 0) return sigmo_err::set(SIGMO_ERR_INVALID, "graphs with zero nodes are not supported");
  if (mx > SIGMO_MAX_DATA_GRAPH_NODES)
    return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "graph with " + std::to_string(mx) + " nodes exceeds the 128-node limit");
  int st = SIGMO_OK;
  if (G > g->cap_G) {
    dev_free(g->go);
    dev_free(g->lmask);
    g->cap_G = 0;
    const uint32_t cG = std::max(G, reserve_G);
    if (st == SIGMO_OK) st = dev_alloc(&g->go, uint64_t(cG) + 1);
    if (st == SIGMO_OK) st = dev_alloc(&g->lmask, 16 * uint64_t(cG));
    if (st != SIGMO_OK) return st;
    g->cap_G = cG;
  }
  if (N > g->cap_N) {
    dev_free(g->ro);
    dev_free(g->nl);
    dev_free(g->node_graph);
    dev_free(g->adj);
    dev_free(g->fold);
    dev_free(g->lplanes);
    g->cap_N = 0;
    const uint64_t cN = std::max(N, reserve_N), W64 = (cN + 63) / 64;
    if (st == SIGMO_OK) st = dev_alloc(&g->ro, cN + 1);
    if (st == SIGMO_OK) st = dev_alloc(&g->nl, cN);
    if (st == SIGMO_OK) st = dev_alloc(&g->node_graph, cN);
    if (st == SIGMO_OK) st = dev_alloc(&g->adj, 2 * cN);
    if (st == SIGMO_OK) st = dev_alloc(&g->fold, cN);
    if (st == SIGMO_OK) st = dev_alloc(&g->lplanes, 8 * W64);
    if (st != SIGMO_OK) return st;
    g->cap_N = cN;
  }
  if (E > g->cap_E || !g->ci) {
    dev_free(g->ci);
    dev_free(g->el);
    g->cap_E = 0;
    const uint64_t cE = std::max(E, reserve_E);
    if (st == SIGMO_OK) st = dev_alloc(&g->ci, cE);
    if (st == SIGMO_OK) st = dev_alloc(&g->el, cE);
    if (st != SIGMO_OK) return st;
    g->cap_E = cE;
  }
  g->G = G;
  g->N = N;
  g->E = E;
  g->max_nodes = mx;
  g->min_nodes = mn;
  g->host_go.resize(uint64_t(G) + 1);
  for (uint32_t i = 0; i <= G; ++i) g->host_go[i] = h.go[i] - h.node_base;
  bool present[16] = {false};
  for (uint64_t v = 0; v < N; ++v)
    if (h.nl[v] < 16) present[h.nl[v]] = true;
  g->slots.count = 0;
  for (int l = 0; l < 16; ++l)
    if (present[l]) g->slots.label[g->slots.count++] = static_cast<uint8_t>(l);
  hipStream_t s = g->ctx->stream;
  hipError_t e = hipMemcpyAsync(g->go, h.go, (uint64_t(G) + 1) * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(g->ro, h.ro, (N + 1) * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && E) e = hipMemcpyAsync(g->ci, h.ci, E * 4, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipMemcpyAsync(g->nl, h.nl, N, hipMemcpyHostToDevice, s);
  if (e == hipSuccess && E) e = hipMemcpyAsync(g->el, h.el, E, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = sigmo_k::sub_u32(s, g->go, uint64_t(G) + 1, h.node_base);
  if (e == hipSuccess) e = sigmo_k::sub_u32(s, g->ci, E, h.node_base);
  if (e == hipSuccess) e = sigmo_k::sub_u32(s, g->ro, N + 1, h.col_base);
  if (e == hipSuccess) e = sigmo_k::derive_graph(s, g->go, G, g->ro, g->ci, g->nl, N, g->node_graph, g->adj, g->fold, g->lplanes, g->lmask);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return sigmo_err::set(SIGMO_ERR_HIP, std::string("graph upload: ") + hipGetErrorString(e));
  return SIGMO_OK;
}

}  // namespace

namespace sigmo_internal {
// Upload `host` into *g, reusing its device buffers when they are large enough (a serving loop
// uploads one query batch per iteration without allocating once capacities have settled).
int graph_reupload(sigmo_ctx* ctx, const sigmo_host_graphs* host, sigmo_graph** g) {
  if (!*g) return sigmo_graph_upload(ctx, host, g);
  SIGMO_TRY(bind(ctx));
  if ((*g)->ctx != ctx) return sigmo_err::set(SIGMO_ERR_INVALID, "graph belongs to another context");
  return graph_fill(*g, whole(host));
}
// Upload graphs [first, first + count) of `host` into *g (created on first use), reusing its
// buffers: the sub-arrays go up as they are and their ids are rebased on the device.  Buffers
// that must grow are sized for at least reserve_{G,N,E} (the largest chunk a caller will load),
// so a stream of chunks allocates once.
int graph_upload_range(sigmo_ctx* ctx, const sigmo_host_graphs* host, uint32_t first, uint32_t count, uint32_t reserve_G, uint64_t reserve_N,
                       uint64_t reserve_E, sigmo_graph** g) {
  SIGMO_TRY(bind(ctx));
  if (!host || !g) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (count == 0 || uint64_t(first) + count > host->num_graphs()) return sigmo_err::set(SIGMO_ERR_INVALID, "graph range out of bounds");
  if (*g && (*g)->ctx != ctx) return sigmo_err::set(SIGMO_ERR_INVALID, "graph belongs to another context");
  const uint32_t node_lo = host->graph_offsets[first], node_hi = host->graph_offsets[first + count];
  const uint32_t col_lo = host->row_offsets[node_lo], col_hi = host->row_offsets[node_hi];
  const HostView v{host->graph_offsets.data() + first, count, host->row_offsets.data() + node_lo, uint64_t(node_hi - node_lo),
                   host->column_indices.data() + col_lo, uint64_t(col_hi - col_lo), host->node_labels.data() + node_lo,
                   host->edge_labels.data() + col_lo, node_lo, col_lo};
  const bool fresh = !*g;
  if (fresh) {
    *g = new sigmo_graph();
    (*g)->ctx = ctx;
  }
  const int st = graph_fill(*g, v, reserve_G, reserve_N, reserve_E);
  if (st != SIGMO_OK && fresh) {
    delete *g;
    *g = nullptr;
  }
  return st;
}
}  // namespace sigmo_internal

extern "C" {

int sigmo_graph_upload(sigmo_ctx* ctx, const sigmo_host_graphs* host, sigmo_graph** out) {
  SIGMO_TRY(bind(ctx));
  if (!host || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  auto* g = new sigmo_graph();
  g->ctx = ctx;
  const int st = graph_fill(g, whole(host));
  if (st != SIGMO_OK) {
    delete g;
    return st;
  }
  *out = g;
  return SIGMO_OK;
}

int sigmo_graph_free(sigmo_graph* g) {
  if (!g) return SIGMO_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  delete g;
  return SIGMO_OK;
}

int sigmo_graph_info(const sigmo_graph* g, uint32_t* num_graphs, uint64_t* total_nodes, uint64_t* total_edges) {
  if (!g) return sigmo_err::set(SIGMO_ERR_INVALID, "null graph");
  if (num_graphs) *num_graphs = g->G;
  if (total_nodes) *total_nodes = g->N;
  if (total_edges) *total_edges = g->E;
  return SIGMO_OK;
}

int sigmo_graph_alloc_size(const sigmo_graph* g, uint64_t* csr_bytes, uint64_t* device_bytes) {
  if (!g) return sigmo_err::set(SIGMO_ERR_INVALID, "null graph");
  if (csr_bytes) *csr_bytes = g->csr_bytes();
  if (device_bytes) *device_bytes = g->device_bytes();
  return SIGMO_OK;
}

// ---- adjacency-matrix graphs ------------------------------------------------------------------

namespace {
// Word count and row width of an n-node matrix: (uint8)((n*n + 63) / 64) (utils.hpp:17-21) and
// (uint16)sqrt(float(64 * words)) (utils.hpp:25-28).
uint32_t am_words(uint32_t n) { return ((n * n + 63) / 64) & 255u; }
uint32_t am_width(uint32_t words) { return static_cast<uint16_t>(std::sqrt(static_cast<float>(64 * words))); }
}  // namespace

int sigmo_am_graph_upload(sigmo_ctx* ctx, const sigmo_host_graphs* host, sigmo_am_graph** out) {
  SIGMO_TRY(bind(ctx));
  if (!host || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  const uint32_t G = host->num_graphs();
  const uint64_t N = host->total_nodes();
  if (G == 0) return sigmo_err::set(SIGMO_ERR_INVALID, "empty graph batch");
  // toAMGraph (graph.hpp:208-214) per graph, laid out as createDeviceAMGraph (graph.hpp:345-392)
  std::vector<uint32_t> go(G), nn(G);
  uint64_t words = 0;
  for (uint32_t g = 0; g < G; ++g) {
    const uint32_t n = host->graph_offsets[g + 1] - host->graph_offsets[g];
    if (n > 64) return sigmo_err::set(SIGMO_ERR_INVALID, "adjacency-matrix graphs hold at most 64 nodes (graph " + std::to_string(g) + " has " + std::to_string(n) + ")");
    go[g] = static_cast<uint32_t>(words);
    nn[g] = host->graph_offsets[g + 1];
    words += am_words(n);
  }
  std::vector<uint64_t> adj(words, 0);
  for (uint32_t g = 0; g < G; ++g) {
    const uint32_t base = host->graph_offsets[g], n = host->graph_offsets[g + 1] - base, W = am_width(am_words(n));
    uint64_t* m = adj.data() + go[g];
    for (uint32_t v = base; v < base + n; ++v)
      for (uint32_t i = host->row_offsets[v]; i < host->row_offsets[v + 1]; ++i) {
        const uint32_t a = v - base, b = host->column_indices[i] - base;
        const uint32_t iu = a * W + b, iv = b * W + a;
        m[iu / 64] |= 1ull << (iu % 64);
        m[iv / 64] |= 1ull << (iv % 64);
      }
  }
  auto* g = new sigmo_am_graph();
  g->ctx = ctx;
  g->G = G;
  g->N = N;
  g->words = words;
  int st = dev_alloc(&g->go, G);
  if (st == SIGMO_OK) st = dev_alloc(&g->nn, G);
  if (st == SIGMO_OK) st = dev_alloc(&g->adj, words);
  if (st == SIGMO_OK) st = dev_alloc(&g->nl, N);
  hipError_t e = hipSuccess;
  hipStream_t s = ctx->stream;
  if (st == SIGMO_OK) e = hipMemcpyAsync(g->go, go.data(), uint64_t(G) * 4, hipMemcpyHostToDevice, s);
  if (st == SIGMO_OK && e == hipSuccess) e = hipMemcpyAsync(g->nn, nn.data(), uint64_t(G) * 4, hipMemcpyHostToDevice, s);
  if (st == SIGMO_OK && e == hipSuccess && words) e = hipMemcpyAsync(g->adj, adj.data(), words * 8, hipMemcpyHostToDevice, s);
  if (st == SIGMO_OK && e == hipSuccess && N) e = hipMemcpyAsync(g->nl, host->node_labels.data(), N, hipMemcpyHostToDevice, s);
  if (st == SIGMO_OK && e == hipSuccess) e = hipStreamSynchronize(s);  // the host vectors go out of scope
  if (st == SIGMO_OK && e != hipSuccess) st = sigmo_err::set(SIGMO_ERR_HIP, std::string("AM graph upload: ") + hipGetErrorString(e));
  if (st != SIGMO_OK) {
    delete g;
    return st;
  }
  *out = g;
  return SIGMO_OK;
}

int sigmo_am_graph_free(sigmo_am_graph* g) {
  if (!g) return SIGMO_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  delete g;
  return SIGMO_OK;
}

int sigmo_am_graph_info(const sigmo_am_graph* g, uint32_t* num_graphs, uint64_t* total_nodes, uint64_t* adjacency_words) {
  if (!g) return sigmo_err::set(SIGMO_ERR_INVALID, "null graph");
  if (num_graphs) *num_graphs = g->G;
  if (total_nodes) *total_nodes = g->N;
  if (adjacency_words) *adjacency_words = g->words;
  return SIGMO_OK;
}

int sigmo_am_graph_download(sigmo_ctx* ctx, const sigmo_am_graph* g, uint32_t* graph_offsets, uint32_t* num_nodes, uint64_t* adjacency,
                            uint8_t* node_labels) {
  SIGMO_TRY(bind(ctx));
  if (!g) return sigmo_err::set(SIGMO_ERR_INVALID, "null graph");
  hipStream_t s = ctx->stream;
  hipError_t e = hipSuccess;
  if (graph_offsets) e = hipMemcpyAsync(graph_offsets, g->go, uint64_t(g->G) * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && num_nodes) e = hipMemcpyAsync(num_nodes, g->nn, uint64_t(g->G) * 4, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && adjacency && g->words) e = hipMemcpyAsync(adjacency, g->adj, g->words * 8, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess && node_labels && g->N) e = hipMemcpyAsync(node_labels, g->nl, g->N, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) return sigmo_err::set(SIGMO_ERR_HIP, std::string("AM graph download: ") + hipGetErrorString(e));
  return SIGMO_OK;
}

// ---- signatures -------------------------------------------------------------------------------

int sigmo_signatures_create(sigmo_ctx* ctx, uint64_t data_nodes, uint64_t query_nodes, sigmo_signatures** out) {
  SIGMO_TRY(bind(ctx));
  if (!out) return sigmo_err::set(SIGMO_ERR_INVALID, "null out");
  auto* s = new sigmo_signatures();
  s->ctx = ctx;
  s->D = data_nodes;
  s->capD = data_nodes;
  s->Q = query_nodes;
  const uint64_t W64 = (data_nodes + 63) / 64;
  int st = dev_alloc(&s->dsig, data_nodes);
  if (st == SIGMO_OK) st = dev_alloc(&s->qsig, query_nodes);
  if (st == SIGMO_OK) st = dev_alloc(&s->dreach, data_nodes);
  if (st == SIGMO_OK) st = dev_alloc(&s->qreach, query_nodes);
  if (st == SIGMO_OK) st = dev_alloc(&s->dplanes, 64 * W64);
  if (st == SIGMO_OK) st = dev_alloc(&s->dfront, data_nodes);
  if (st == SIGMO_OK) st = dev_alloc(&s->qfront, query_nodes);
  hipError_t e = hipSuccess;
  // signatures filled with 0 (signature.hpp:166-167); reachables start empty (zero)
  if (st == SIGMO_OK) e = sigmo_k::fill_u64(ctx->stream, s->dsig, data_nodes, 0);
  if (st == SIGMO_OK && e == hipSuccess) e = sigmo_k::fill_u64(ctx->stream, s->qsig, query_nodes, 0);
  if (st == SIGMO_OK && e == hipSuccess) e = sigmo_k::fill_u64(ctx->stream, s->dreach, data_nodes, 0);
  if (st == SIGMO_OK && e == hipSuccess) e = sigmo_k::fill_u64(ctx->stream, s->qreach, query_nodes, 0);
  if (st == SIGMO_OK && e == hipSuccess) e = sigmo_k::fill_u64(ctx->stream, s->dplanes, 64 * W64, 0);
  if (st == SIGMO_OK && e != hipSuccess) st = sigmo_err::set(SIGMO_ERR_HIP, std::string("signature init: ") + hipGetErrorString(e));
  if (st != SIGMO_OK) {
    delete s;
    return st;
  }
  *out = s;
  return SIGMO_OK;
}

int sigmo_signatures_free(sigmo_signatures* s) {
  if (!s) return SIGMO_OK;
  (void)hipSetDevice(s->ctx->device);
  (void)hipStreamSynchronize(s->ctx->stream);
  delete s;
  return SIGMO_OK;
}

static int sig_target(sigmo_signatures* s, const sigmo_graph* g, int scope, uint64_t** sigs, uint64_t** reach, uint64_t** planes) {
  SIGMO_TRY(check_scope(scope));
  if (!s || !g) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  const uint64_t n = scope == SIGMO_SCOPE_DATA ? s->D : s->Q;
  if (g->N != n) return sigmo_err::set(SIGMO_ERR_INVALID, "graph node count does not match the signature allocation");
  *sigs = scope == SIGMO_SCOPE_DATA ? s->dsig : s->qsig;
  *reach = scope == SIGMO_SCOPE_DATA ? s->dreach : s->qreach;
  *planes = scope == SIGMO_SCOPE_DATA ? s->dplanes : nullptr;
  return SIGMO_OK;
}

// Algorithmic HBM bytes of one signature pass: CSR + labels read, 8-byte signature written.
static double sig_bytes(const sigmo_graph* g) { return 4.0 * (g->N + 1) + 4.0 * g->E + g->N + 8.0 * g->N; }

// `fresh` (the pipeline's first calls): the signatures, resp. the reachables, are taken as
// zero without a clearing pass -- what the reference's constructor fill leaves
// (signature.hpp:166-167; reachables start empty, A3).  `planes_only` (the pipeline's data
// side, whose packed signatures nothing reads): only the bit-sliced planes are written.
static int signatures_generate(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_graph* g, int scope, bool fresh, bool planes_only = false) {
  SIGMO_TRY(bind(ctx));
  uint64_t *sigs, *reach, *planes;
  SIGMO_TRY(sig_target(s, g, scope, &sigs, &reach, &planes));
  if (planes_only && planes && fresh) sigs = nullptr;
  PhaseScope ps(ctx, scope == SIGMO_SCOPE_DATA ? SIGMO_PHASE_SIG_DATA : SIGMO_PHASE_SIG_QUERY,
                scope == SIGMO_SCOPE_DATA ? SIGMO_KERNEL_SIG_GENERATE : -1, sig_bytes(g));
  SIGMO_HIP_TRY(sigmo_k::sig_generate(ctx->stream, g->dev(), sigs, planes, fresh));
  return ps.close();
}

int sigmo_signatures_generate(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_graph* g, int scope) {
  return signatures_generate(ctx, s, g, scope, false);
}

static int signatures_refine(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_graph* g, int scope, uint32_t view_size, bool fresh_reach,
                             bool planes_only = false);

int sigmo_signatures_refine(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_graph* g, int scope, uint32_t view_size) {
  return signatures_refine(ctx, s, g, scope, view_size, false);
}

static int signatures_refine(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_graph* g, int scope, uint32_t view_size, bool fresh_reach,
                             bool planes_only) {
  SIGMO_TRY(bind(ctx));
  uint64_t *sigs, *reach, *planes;
  SIGMO_TRY(sig_target(s, g, scope, &sigs, &reach, &planes));
  if (planes_only && planes) sigs = nullptr;
  RefineState& st = scope == SIGMO_SCOPE_DATA ? s->dstate : s->qstate;
  uint64_t* front = scope == SIGMO_SCOPE_DATA ? s->dfront : s->qfront;
  if (st.graph && st.graph != g) {
    st.tainted = true;
    st.frontier_valid = false;
  }
  st.graph = g;
  PhaseScope ps(ctx, scope == SIGMO_SCOPE_DATA ? SIGMO_PHASE_SIG_DATA : SIGMO_PHASE_SIG_QUERY,
                scope == SIGMO_SCOPE_DATA ? SIGMO_KERNEL_SIG_REFINE : -1, sig_bytes(g));
  if (!st.tainted && st.frontier_valid && view_size == st.max_view + 1) {
    SIGMO_HIP_TRY(sigmo_k::sig_refine_next(ctx->stream, g->dev(), sigs, reach, front, planes));
    st.max_view = view_size;
  } else {
    const bool keep = !st.tainted && view_size > st.max_view;
    SIGMO_HIP_TRY(sigmo_k::sig_refine(ctx->stream, g->dev(), view_size, sigs, reach, keep ? front : nullptr, planes, fresh_reach));
    if (view_size > st.max_view) {
      st.max_view = view_size;
      st.frontier_valid = keep;
    }
  }
  return ps.close();
}

int sigmo_signatures_refine_view_based(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_graph* g, int scope, uint32_t view_size) {
  SIGMO_TRY(bind(ctx));
  uint64_t *sigs, *reach, *planes;
  SIGMO_TRY(sig_target(s, g, scope, &sigs, &reach, &planes));
  SIGMO_TRY(grow(&s->tmp, s->cap_tmp, g->N));
  PhaseScope ps(ctx, scope == SIGMO_SCOPE_DATA ? SIGMO_PHASE_SIG_DATA : SIGMO_PHASE_SIG_QUERY, -1, 0.0);
  if (g->N) SIGMO_HIP_TRY(hipMemcpyAsync(s->tmp, sigs, g->N * 8, hipMemcpyDeviceToDevice, ctx->stream));  // tmp_buff (:306-309)
  SIGMO_HIP_TRY(sigmo_k::sig_refine_view_based(ctx->stream, g->dev(), view_size, s->tmp, sigs, planes));
  return ps.close();
}

static int am_sig_target(sigmo_signatures* s, const sigmo_am_graph* g, int scope, uint64_t** sigs, uint64_t** planes) {
  SIGMO_TRY(check_scope(scope));
  if (!s || !g) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (g->N != (scope == SIGMO_SCOPE_DATA ? s->D : s->Q))
    return sigmo_err::set(SIGMO_ERR_INVALID, "graph node count does not match the signature allocation");
  *sigs = scope == SIGMO_SCOPE_DATA ? s->dsig : s->qsig;
  *planes = scope == SIGMO_SCOPE_DATA ? s->dplanes : nullptr;
  return SIGMO_OK;
}

int sigmo_signatures_generate_am(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_am_graph* g, int scope) {
  SIGMO_TRY(bind(ctx));
  uint64_t *sigs, *planes;
  SIGMO_TRY(am_sig_target(s, g, scope, &sigs, &planes));
  PhaseScope ps(ctx, scope == SIGMO_SCOPE_DATA ? SIGMO_PHASE_SIG_DATA : SIGMO_PHASE_SIG_QUERY, -1, 0.0);
  SIGMO_HIP_TRY(sigmo_k::am_generate(ctx->stream, g->dev(), sigs, planes));
  return ps.close();
}

int sigmo_signatures_refine_am(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_am_graph* g, int scope, uint32_t view_size) {
  SIGMO_TRY(bind(ctx));
  uint64_t *sigs, *planes;
  SIGMO_TRY(am_sig_target(s, g, scope, &sigs, &planes));
  PhaseScope ps(ctx, scope == SIGMO_SCOPE_DATA ? SIGMO_PHASE_SIG_DATA : SIGMO_PHASE_SIG_QUERY, -1, 0.0);
  SIGMO_HIP_TRY(sigmo_k::am_refine_power_graph(ctx->stream, g->dev(), view_size, sigs, planes));
  return ps.close();
}

int sigmo_signatures_refine_am_view_based(sigmo_ctx* ctx, sigmo_signatures* s, const sigmo_am_graph* g, int scope, uint32_t view_size) {
  SIGMO_TRY(bind(ctx));
  uint64_t *sigs, *planes;
  SIGMO_TRY(am_sig_target(s, g, scope, &sigs, &planes));
  SIGMO_TRY(grow(&s->tmp, s->cap_tmp, g->N));
  PhaseScope ps(ctx, scope == SIGMO_SCOPE_DATA ? SIGMO_PHASE_SIG_DATA : SIGMO_PHASE_SIG_QUERY, -1, 0.0);
  if (g->N) SIGMO_HIP_TRY(hipMemcpyAsync(s->tmp, sigs, g->N * 8, hipMemcpyDeviceToDevice, ctx->stream));  // tmp_buff (signature.hpp:218-223)
  SIGMO_HIP_TRY(sigmo_k::am_refine_view_based(ctx->stream, g->dev(), view_size, s->tmp, sigs, planes));
  return ps.close();
}

int sigmo_signatures_download(sigmo_ctx* ctx, const sigmo_signatures* s, int scope, uint64_t* out, uint64_t count) {
  SIGMO_TRY(bind(ctx));
  SIGMO_TRY(check_scope(scope));
  if (!s || (!out && count)) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  const uint64_t n = scope == SIGMO_SCOPE_DATA ? s->D : s->Q;
  if (count > n) return sigmo_err::set(SIGMO_ERR_INVALID, "count exceeds the signature allocation");
  if (!count) return SIGMO_OK;
  SIGMO_HIP_TRY(hipMemcpyAsync(out, scope == SIGMO_SCOPE_DATA ? s->dsig : s->qsig, count * 8, hipMemcpyDeviceToHost, ctx->stream));
  SIGMO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return SIGMO_OK;
}

int sigmo_signatures_alloc_size(const sigmo_signatures* s, int scope, uint64_t* bytes) {
  SIGMO_TRY(check_scope(scope));
  if (!s || !bytes) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  // signature.hpp:181-198: signatures plus the 16-byte Bitset<uint64_t> reachables
  *bytes = (scope == SIGMO_SCOPE_DATA ? s->D : s->Q) * (8 + 16);
  return SIGMO_OK;
}

// ---- candidates ---------------------------------------------------------------------------------

int sigmo_candidates_create(sigmo_ctx* ctx, uint64_t source_nodes, uint64_t target_nodes, sigmo_candidates** out) {
  SIGMO_TRY(bind(ctx));
  if (!out) return sigmo_err::set(SIGMO_ERR_INVALID, "null out");
  auto* c = new sigmo_candidates();
  c->ctx = ctx;
  c->Q = source_nodes;
  c->D = target_nodes;
  c->W64 = (target_nodes + 63) / 64;
  c->cap_words = c->Q * c->W64;
  // one spare word past the rows: the join reads a row's words in pairs
  int st = dev_alloc(&c->bits, c->Q * c->W64 + 1);
  if (st == SIGMO_OK) {
    hipError_t e = sigmo_k::fill_u64(ctx->stream, c->bits, c->Q * c->W64 + 1, 0);  // candidates.hpp:169-178
    if (e != hipSuccess) st = sigmo_err::set(SIGMO_ERR_HIP, hipGetErrorString(e));
  }
  if (st != SIGMO_OK) {
    delete c;
    return st;
  }
  *out = c;
  return SIGMO_OK;
}

int sigmo_candidates_free(sigmo_candidates* c) {
  if (!c) return SIGMO_OK;
  (void)hipSetDevice(c->ctx->device);
  (void)hipStreamSynchronize(c->ctx->stream);
  delete c;
  return SIGMO_OK;
}

static int check_cand(const sigmo_candidates* c, const sigmo_graph* q, const sigmo_graph* d) {
  if (!c || !q || !d) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (c->Q != q->N || c->D != d->N) return sigmo_err::set(SIGMO_ERR_INVALID, "candidate shape does not match the graphs");
  return SIGMO_OK;
}

int sigmo_candidates_filter(sigmo_ctx* ctx, sigmo_candidates* c, const sigmo_graph* query, const sigmo_graph* data,
                            const sigmo_signatures* s) {
  SIGMO_TRY(bind(ctx));
  SIGMO_TRY(check_cand(c, query, data));
  (void)s;  // signatures are captured but unused by the reference filter (isomorphism.hpp:39-44)
  PhaseScope ps(ctx, SIGMO_PHASE_FILTER, SIGMO_KERNEL_FILTER, 8.0 * c->Q * c->W64 + c->D + c->Q);
  SIGMO_HIP_TRY(sigmo_k::filter(ctx->stream, query->nl, c->Q, data->lplanes, c->D, c->bits, ctx->num_cus));
  return ps.close();
}

int sigmo_candidates_filter_data(sigmo_ctx* ctx, sigmo_candidates* c, const sigmo_graph* query, const sigmo_graph* data) {
  SIGMO_TRY(bind(ctx));
  if (!c || !query || !data) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (c->Q != data->N || c->D != query->N)
    return sigmo_err::set(SIGMO_ERR_INVALID, "data-domain candidates need data-node rows and query-node columns");
  PhaseScope ps(ctx, SIGMO_PHASE_FILTER, -1, 0.0);
  SIGMO_HIP_TRY(sigmo_k::filter_data(ctx->stream, query->nl, query->N, data->nl, data->N, c->bits));
  return ps.close();
}

int sigmo_candidates_refine(sigmo_ctx* ctx, sigmo_candidates* c, const sigmo_graph* query, const sigmo_graph* data,
                            const sigmo_signatures* s) {
  SIGMO_TRY(bind(ctx));
  SIGMO_TRY(check_cand(c, query, data));
  if (!s || s->D != c->D || s->Q != c->Q) return sigmo_err::set(SIGMO_ERR_INVALID, "signature shape does not match the candidates");
  PhaseScope ps(ctx, SIGMO_PHASE_FILTER, SIGMO_KERNEL_REFINE, 16.0 * c->Q * c->W64 + 8.0 * c->D + 8.0 * c->Q);
  SIGMO_HIP_TRY(sigmo_k::refine(ctx->stream, s->qsig, c->Q, s->dplanes, c->D, query->slots, c->bits, ctx->num_cus));
  return ps.close();
}

int sigmo_candidates_download(sigmo_ctx* ctx, const sigmo_candidates* c, uint32_t* words, uint64_t n_words) {
  SIGMO_TRY(bind(ctx));
  if (!c || (!words && n_words)) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  const uint64_t S32 = (c->D + 31) / 32;
  if (n_words != c->Q * S32) return sigmo_err::set(SIGMO_ERR_INVALID, "n_words must equal rows * ceil(target/32)");
  if (!n_words) return SIGMO_OK;
  // Row q of the reference image is the first 4*S32 bytes of device row q.
  SIGMO_HIP_TRY(hipMemcpy2DAsync(words, S32 * 4, c->bits, c->W64 * 8, S32 * 4, c->Q, hipMemcpyDeviceToHost, ctx->stream));
  SIGMO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return SIGMO_OK;
}

int sigmo_candidates_counts(sigmo_ctx* ctx, const sigmo_candidates* c, uint32_t* counts, uint64_t n) {
  SIGMO_TRY(bind(ctx));
  if (!c || (!counts && n)) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (n != c->Q) return sigmo_err::set(SIGMO_ERR_INVALID, "n must equal the number of source rows");
  if (!n) return SIGMO_OK;
  uint32_t* d = nullptr;
  SIGMO_TRY(dev_alloc(&d, n));
  hipError_t e = sigmo_k::row_counts(ctx->stream, c->bits, c->Q, c->W64, d);
  if (e == hipSuccess) e = hipMemcpyAsync(counts, d, n * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  dev_free(d);
  if (e != hipSuccess) return sigmo_err::set(SIGMO_ERR_HIP, hipGetErrorString(e));
  return SIGMO_OK;
}

int sigmo_candidates_alloc_size(const sigmo_candidates* c, uint64_t* ref_bytes, uint64_t* device_bytes) {
  if (!c) return sigmo_err::set(SIGMO_ERR_INVALID, "null candidates");
  if (ref_bytes) *ref_bytes = c->Q * ((c->D + 31) / 32) * 4;
  if (device_bytes) *device_bytes = c->Q * c->W64 * 8;
  return SIGMO_OK;
}

// ---- GMCR --------------------------------------------------------------------------------------

int sigmo_gmcr_create(sigmo_ctx* ctx, sigmo_gmcr** out) {
  SIGMO_TRY(bind(ctx));
  if (!out) return sigmo_err::set(SIGMO_ERR_INVALID, "null out");
  auto* m = new sigmo_gmcr();
  m->ctx = ctx;
  int st = dev_alloc(&m->total, 1);
  if (st != SIGMO_OK) {
    delete m;
    return st;
  }
  *out = m;
  return SIGMO_OK;
}

int sigmo_gmcr_free(sigmo_gmcr* m) {
  if (!m) return SIGMO_OK;
  (void)hipSetDevice(m->ctx->device);
  (void)hipStreamSynchronize(m->ctx->stream);
  delete m;
  return SIGMO_OK;
}

static int check_queries(const sigmo_graph* q) {
  for (uint32_t g = 0; g < q->G; ++g) {
    const uint32_t k = q->host_go[g + 1] - q->host_go[g];
    if (k < 2) return sigmo_err::set(SIGMO_ERR_INVALID, "query graph " + std::to_string(g) + " has fewer than 2 nodes");
    if (k > SIGMO_MAX_QUERY_NODES)
      return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "query graph " + std::to_string(g) + " exceeds 32 nodes");
  }
  return SIGMO_OK;
}

// Pair spaces up to this size let the plain pipeline skip the host round trip for the GMCR
// total: the index arrays are sized for every pair (4 x 4 B per pair at most).
constexpr uint64_t kDeferredPairs = 1ull << 25;

static int gmcr_generate(sigmo_ctx* ctx, sigmo_gmcr* m, const sigmo_graph* query, const sigmo_graph* data, const sigmo_candidates* c,
                         bool defer);

int sigmo_gmcr_generate(sigmo_ctx* ctx, sigmo_gmcr* m, const sigmo_graph* query, const sigmo_graph* data, const sigmo_candidates* c) {
  return gmcr_generate(ctx, m, query, data, c, false);
}

// The exact pair count of a deferred GMCR (a no-op otherwise).
static int gmcr_resolve(sigmo_gmcr* m, hipStream_t s) {
  if (!m->deferred) return SIGMO_OK;
  SIGMO_HIP_TRY(hipStreamSynchronize(s));
  m->P = *m->total_host;
  m->deferred = false;
  if (m->P >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "GMCR pair count exceeds u32");
  return SIGMO_OK;
}

static int gmcr_generate(sigmo_ctx* ctx, sigmo_gmcr* m, const sigmo_graph* query, const sigmo_graph* data, const sigmo_candidates* c,
                         bool defer) {
  SIGMO_TRY(bind(ctx));
  if (!m) return sigmo_err::set(SIGMO_ERR_INVALID, "null gmcr");
  SIGMO_TRY(check_cand(c, query, data));
  SIGMO_TRY(check_queries(query));
  const uint32_t Gq = query->G, Gd = data->G;
  const uint64_t strips = (uint64_t(Gd) + 63) / 64;
  SIGMO_TRY(grow(&m->pair_bits, m->cap_pair_bits, uint64_t(Gq) * strips));
  SIGMO_TRY(grow(&m->counts, m->cap_counts_g, uint64_t(Gd) + 1));
  SIGMO_TRY(grow(&m->offsets, m->cap_offsets, uint64_t(Gd) + 1));
  SIGMO_TRY(grow(&m->scan_tmp, m->cap_scan, (uint64_t(Gd) + 1023) / 1024 + 1));
  hipStream_t s = ctx->stream;
  PhaseScope ps(ctx, SIGMO_PHASE_GMCR, SIGMO_KERNEL_GMCR, 8.0 * c->Q * c->W64 + 4.0 * (uint64_t(Gd) + 1));
  SIGMO_HIP_TRY(sigmo_k::gmcr_mark(s, query->dev(), data->dev(), c->bits, c->W64, m->pair_bits));
  SIGMO_HIP_TRY(sigmo_k::gmcr_count(s, m->pair_bits, Gq, Gd, m->counts));
  SIGMO_HIP_TRY(sigmo_k::exclusive_scan(s, m->counts, Gd, m->offsets, m->scan_tmp, m->total));
  // gmcr.hpp:79-84 reads the total back to size the index array; a graph being recorded keeps
  // the previous run's count and sizes (the replay checks the total).  A deferred GMCR sizes the
  // arrays for every pair instead and leaves the total on the device for the join.
  uint64_t P = m->P;
  const uint64_t bound = uint64_t(Gq) * Gd;
  m->deferred = defer && !ctx->capturing && data->max_nodes <= 64 && bound <= kDeferredPairs;
  if (m->deferred) {
    SIGMO_TRY(grow(&m->indices, m->cap_indices, bound));
    SIGMO_TRY(grow(&m->pair_graph, m->cap_pair_graph, bound));
    if (!m->total_host) SIGMO_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&m->total_host), sizeof(unsigned long long)));
    SIGMO_HIP_TRY(hipMemcpyAsync(m->total_host, m->total, 8, hipMemcpyDeviceToHost, s));  // read after the join's sync
    P = bound;
  } else if (!ctx->capturing) {
    SIGMO_HIP_TRY(hipMemcpyAsync(&P, m->total, 8, hipMemcpyDeviceToHost, s));
    SIGMO_HIP_TRY(hipStreamSynchronize(s));
    if (P >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "GMCR pair count exceeds u32");
    SIGMO_TRY(grow_slack(&m->indices, m->cap_indices, P));
    SIGMO_TRY(grow_slack(&m->pair_graph, m->cap_pair_graph, P));
  }
  SIGMO_HIP_TRY(sigmo_k::gmcr_scatter(s, m->pair_bits, Gq, Gd, m->offsets, m->indices, m->pair_graph,
                                      std::min(m->cap_indices, m->cap_pair_graph)));
  SIGMO_TRY(ps.close());
  m->Gq = Gq;
  m->Gd = Gd;
  m->P = P;
  m->built = true;
  return SIGMO_OK;
}

int sigmo_gmcr_total(const sigmo_gmcr* m, uint64_t* total_query_indices) {
  if (m && m->deferred) SIGMO_TRY(gmcr_resolve(const_cast<sigmo_gmcr*>(m), m->ctx->stream));
  if (!m || !total_query_indices) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  *total_query_indices = m->P;
  return SIGMO_OK;
}

int sigmo_gmcr_download(sigmo_ctx* ctx, const sigmo_gmcr* m, uint32_t* data_graph_offsets, uint64_t n_offsets,
                        uint32_t* query_graph_indices, uint64_t n_indices) {
  SIGMO_TRY(bind(ctx));
  if (!m || !m->built) return sigmo_err::set(SIGMO_ERR_INVALID, "GMCR not generated");
  SIGMO_TRY(gmcr_resolve(const_cast<sigmo_gmcr*>(m), ctx->stream));
  if (data_graph_offsets) {
    if (n_offsets != uint64_t(m->Gd) + 1) return sigmo_err::set(SIGMO_ERR_INVALID, "n_offsets must be Gd+1");
    SIGMO_HIP_TRY(hipMemcpyAsync(data_graph_offsets, m->offsets, n_offsets * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  if (query_graph_indices) {
    if (n_indices != m->P) return sigmo_err::set(SIGMO_ERR_INVALID, "n_indices must be the pair count");
    if (n_indices)
      SIGMO_HIP_TRY(hipMemcpyAsync(query_graph_indices, m->indices, n_indices * 4, hipMemcpyDeviceToHost, ctx->stream));
  }
  SIGMO_HIP_TRY(hipStreamSynchronize(ctx->stream));
  return SIGMO_OK;
}

// ---- join ----------------------------------------------------------------------------------------

int sigmo_join(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, const sigmo_candidates* c, const sigmo_gmcr* m,
               int find_first, uint64_t* num_matches, int32_t* per_pair_counts) {
  SIGMO_TRY(bind(ctx));
  SIGMO_TRY(check_cand(c, query, data));
  if (!m || !m->built) return sigmo_err::set(SIGMO_ERR_INVALID, "GMCR not generated");
  if (m->Gq != query->G || m->Gd != data->G) return sigmo_err::set(SIGMO_ERR_INVALID, "GMCR built for other graphs");
  SIGMO_TRY(check_queries(query));
  auto* mm = const_cast<sigmo_gmcr*>(m);
  SIGMO_TRY(grow_slack(&mm->pair_counts, mm->cap_counts, m->P));
  hipStream_t s = ctx->stream;
  PhaseScope ps(ctx, SIGMO_PHASE_JOIN, SIGMO_KERNEL_JOIN, 0.0);
  SIGMO_HIP_TRY(sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(ctx->stats), sigmo_k::kJoinStatWords, 0));
  SIGMO_TRY(grow_slack(&mm->wide_pairs, mm->cap_wide, m->P));
  SIGMO_TRY(grow_slack(&mm->deep_pairs, mm->cap_deep, m->P));
  // a deferred GMCR: P bounds the grid and the kernel reads the exact count on the device
  SIGMO_HIP_TRY(sigmo_k::join_lanes(s, query->dev(), data->dev(), c->bits, c->W64, m->indices, m->pair_graph, m->P, find_first,
                                    mm->pair_counts, ctx->stats, mm->wide_pairs, ctx->num_cus,
                                    m->deferred || ctx->capturing ? reinterpret_cast<const unsigned long long*>(m->total) : nullptr,
                                    mm->deep_pairs));
  if (data->max_nodes > 64 && m->P) {  // pairs of molecules with > 64 nodes: the two-word lane join
    unsigned long long nw = mm->last_wide;  // a graph being recorded keeps the previous run's count
    if (!ctx->capturing) {
      SIGMO_HIP_TRY(hipMemcpyAsync(&nw, ctx->stats + 5, sizeof(nw), hipMemcpyDeviceToHost, s));
      SIGMO_HIP_TRY(hipStreamSynchronize(s));
      mm->last_wide = nw;
    }
    if (nw)
      SIGMO_HIP_TRY(sigmo_k::join_lanes_wide(s, query->dev(), data->dev(), c->bits, c->W64, m->indices, m->pair_graph, mm->wide_pairs, nw,
                                             find_first, mm->pair_counts, ctx->stats, ctx->num_cus,
                                             ctx->capturing ? ctx->stats + 5 : nullptr, mm->deep_pairs));
  }
  SIGMO_TRY(ps.close());
  SIGMO_HIP_TRY(hipMemcpyAsync(ctx->stats_host, ctx->stats, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  if (ctx->capturing) return SIGMO_OK;  // the replay reads the pinned mirror after the graph
  SIGMO_TRY(gmcr_resolve(mm, s));  // the exact pair count, before any per-pair copy
  if (per_pair_counts && m->P)
    SIGMO_HIP_TRY(hipMemcpyAsync(per_pair_counts, mm->pair_counts, m->P * 4, hipMemcpyDeviceToHost, s));
  SIGMO_HIP_TRY(hipStreamSynchronize(s));
  if (ctx->stats_host[3]) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "a per-pair match count exceeded int32");
  if (num_matches) *num_matches += ctx->stats_host[0];
  return SIGMO_OK;
}

int sigmo_join_wildcard(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, const sigmo_candidates* c, const sigmo_gmcr* m,
                        int find_first, uint64_t* num_matches, int32_t* per_pair_counts) {
  SIGMO_TRY(bind(ctx));
  SIGMO_TRY(check_cand(c, query, data));
  if (!m || !m->built) return sigmo_err::set(SIGMO_ERR_INVALID, "GMCR not generated");
  if (m->Gq != query->G || m->Gd != data->G) return sigmo_err::set(SIGMO_ERR_INVALID, "GMCR built for other graphs");
  SIGMO_TRY(check_queries(query));
  auto* mm = const_cast<sigmo_gmcr*>(m);
  SIGMO_TRY(gmcr_resolve(mm, ctx->stream));
  SIGMO_TRY(grow(&mm->pair_counts, mm->cap_counts, m->P));
  hipStream_t s = ctx->stream;
  PhaseScope ps(ctx, SIGMO_PHASE_JOIN, SIGMO_KERNEL_JOIN, 0.0);
  SIGMO_HIP_TRY(sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(ctx->stats), sigmo_k::kJoinStatWords, 0));
  SIGMO_HIP_TRY(sigmo_k::join_wildcard(s, query->dev(), data->dev(), query->el, data->el, c->bits, c->W64, m->indices, m->pair_graph, m->P,
                                       find_first, mm->pair_counts, ctx->stats, ctx->num_cus));
  SIGMO_TRY(ps.close());
  SIGMO_HIP_TRY(hipMemcpyAsync(ctx->stats_host, ctx->stats, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  if (per_pair_counts && m->P)
    SIGMO_HIP_TRY(hipMemcpyAsync(per_pair_counts, mm->pair_counts, m->P * 4, hipMemcpyDeviceToHost, s));
  SIGMO_HIP_TRY(hipStreamSynchronize(s));
  if (ctx->stats_host[3]) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "a per-pair match count exceeded int32");
  if (num_matches) *num_matches += ctx->stats_host[0];
  return SIGMO_OK;
}

int sigmo_join_noalloc(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, const sigmo_candidates* c, const sigmo_gmcr* m,
                       int find_first, uint64_t* num_matches, int32_t* per_pair_counts) {
  SIGMO_TRY(bind(ctx));
  SIGMO_TRY(check_cand(c, query, data));
  if (!m || !m->built) return sigmo_err::set(SIGMO_ERR_INVALID, "GMCR not generated");
  if (m->Gq != query->G || m->Gd != data->G) return sigmo_err::set(SIGMO_ERR_INVALID, "GMCR built for other graphs");
  SIGMO_TRY(check_queries(query));
  auto* mm = const_cast<sigmo_gmcr*>(m);
  hipStream_t s = ctx->stream;
  SIGMO_TRY(gmcr_resolve(mm, s));  // the exact pair count (a deferred GMCR left it on the device)
  if (m->P > kScratchPairs)
    return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "the allocation-free join takes at most " + std::to_string(kScratchPairs) +
                                                     " GMCR pairs (this GMCR has " + std::to_string(m->P) + "): use sigmo_join");
  uint64_t* base = scratch_on(ctx->device);
  if (!base) return sigmo_err::set(SIGMO_ERR_INVALID, "no scratch on this device (context set-up did not run)");
  auto* stats = reinterpret_cast<unsigned long long*>(base);
  auto* counts = reinterpret_cast<int32_t*>(base + kScratchStatWords);
  auto* wide = reinterpret_cast<uint32_t*>(counts + kScratchPairs);
  // the scratch keeps the previous call's contents: the counters (work cursors included) start at 0
  SIGMO_HIP_TRY(sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(stats), kScratchStatWords, 0));
  SIGMO_HIP_TRY(sigmo_k::join_lanes(s, query->dev(), data->dev(), c->bits, c->W64, m->indices, m->pair_graph, m->P, find_first, counts, stats,
                                    wide, ctx->num_cus));
  if (data->max_nodes > 64 && m->P) {  // pairs of molecules with > 64 nodes: the two-word lane join
    SIGMO_HIP_TRY(hipMemcpyAsync(ctx->stats_host, stats, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    SIGMO_HIP_TRY(hipStreamSynchronize(s));
    const uint64_t nw = ctx->stats_host[5];
    if (nw)
      SIGMO_HIP_TRY(sigmo_k::join_lanes_wide(s, query->dev(), data->dev(), c->bits, c->W64, m->indices, m->pair_graph, wide, nw,
                                             find_first, counts, stats, ctx->num_cus));
  }
  SIGMO_HIP_TRY(hipMemcpyAsync(ctx->stats_host, stats, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  if (per_pair_counts && m->P) SIGMO_HIP_TRY(hipMemcpyAsync(per_pair_counts, counts, m->P * 4, hipMemcpyDeviceToHost, s));
  SIGMO_HIP_TRY(hipStreamSynchronize(s));
  if (ctx->stats_host[3]) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "a per-pair match count exceeded int32");
  if (num_matches) *num_matches += ctx->stats_host[0];
  return SIGMO_OK;
}

// ---- plain pipeline --------------------------------------------------------------------------------

}  // extern "C"

struct sigmo_pipeline {
  sigmo_ctx* ctx = nullptr;
  const sigmo_graph* q = nullptr;
  const sigmo_graph* d = nullptr;
  sigmo_signatures* sig = nullptr;
  sigmo_candidates* cand = nullptr;
  sigmo_gmcr* gm = nullptr;
  bool has_counts = false;
  // a captured execute (sigmo_pipeline_capture)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  uint64_t* replay_host = nullptr;  // pinned: the GMCR total of the last replay
  uint64_t graph_P = 0, graph_nw = 0;
  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }
  ~sigmo_pipeline() {
    drop_graph();
    if (replay_host) (void)hipHostFree(replay_host);
    sigmo_signatures_free(sig);
    sigmo_candidates_free(cand);
    sigmo_gmcr_free(gm);
  }
};

namespace {
// The pipeline's launches in the driver's order (sigmo.cpp:102-159), from fresh signature state
// (signature.hpp:166-167; reachables empty).  Also what a capture records.
// sigmo_candidates_filter, then sigmo_candidates_refine against the current signatures, as one
// kernel (the pipeline's first refinement).
int candidates_filter_refine(sigmo_ctx* ctx, sigmo_candidates* c, const sigmo_graph* query, const sigmo_graph* data,
                             const sigmo_signatures* s) {
  SIGMO_TRY(check_cand(c, query, data));
  if (!s || s->D != c->D || s->Q != c->Q) return sigmo_err::set(SIGMO_ERR_INVALID, "signature shape does not match the candidates");
  PhaseScope ps(ctx, SIGMO_PHASE_FILTER, SIGMO_KERNEL_FILTER,
                8.0 * c->Q * c->W64 + c->D + 0.5 * query->slots.count * c->D + 9.0 * c->Q);
  SIGMO_HIP_TRY(sigmo_k::filter_refine(ctx->stream, query->nl, s->qsig, c->Q, data->lplanes, s->dplanes, c->D, query->slots, c->bits,
                                       ctx->num_cus));
  return ps.close();
}

int pipeline_enqueue(sigmo_pipeline* p, uint32_t refinement_steps, int find_first, uint64_t* matches) {
  sigmo_ctx* ctx = p->ctx;
  sigmo_signatures* sig = p->sig;
  struct Chain {  // the phases below run back to back: chain their timing events
    sigmo_ctx* c;
    explicit Chain(sigmo_ctx* x) : c(x) {
      c->chain = true;
      c->chain_end = nullptr;
    }
    ~Chain() {
      c->chain = false;
      c->chain_end = nullptr;
    }
  } chain(ctx);
  // every run starts from the constructor's state: zero signatures and empty reachables, taken
  // as such by the first generate and the first refinement (no clearing pass)
  sig->dstate.reset();
  sig->qstate.reset();
  // the data side's packed signatures are never read here (the refine reads its planes)
  SIGMO_TRY(signatures_generate(ctx, sig, p->d, SIGMO_SCOPE_DATA, true, true));
  SIGMO_TRY(signatures_generate(ctx, sig, p->q, SIGMO_SCOPE_QUERY, true));
  for (uint32_t r = 1; r <= refinement_steps; ++r) {
    SIGMO_TRY(signatures_refine(ctx, sig, p->d, SIGMO_SCOPE_DATA, r, r == 1, true));
    SIGMO_TRY(signatures_refine(ctx, sig, p->q, SIGMO_SCOPE_QUERY, r, r == 1));
    // the filter reads no signatures and the signature passes read no candidates: the filter
    // runs fused with the first candidate refinement, after that level's signatures
    SIGMO_TRY(r == 1 ? candidates_filter_refine(ctx, p->cand, p->q, p->d, sig) : sigmo_candidates_refine(ctx, p->cand, p->q, p->d, sig));
  }
  if (refinement_steps == 0) SIGMO_TRY(sigmo_candidates_filter(ctx, p->cand, p->q, p->d, sig));
  SIGMO_TRY(gmcr_generate(ctx, p->gm, p->q, p->d, p->cand, true));
  return sigmo_join(ctx, p->q, p->d, p->cand, p->gm, find_first, matches, nullptr);
}
}  // namespace

extern "C" {

int sigmo_pipeline_create(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, sigmo_pipeline** out) {
  SIGMO_TRY(bind(ctx));
  if (!query || !data || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  SIGMO_TRY(check_queries(query));
  auto* p = new sigmo_pipeline();
  p->ctx = ctx;
  p->q = query;
  p->d = data;
  int st = sigmo_signatures_create(ctx, data->N, query->N, &p->sig);
  if (st == SIGMO_OK) st = sigmo_candidates_create(ctx, query->N, data->N, &p->cand);
  if (st == SIGMO_OK) st = sigmo_gmcr_create(ctx, &p->gm);
  if (st == SIGMO_OK) st = sigmo_ctx_sync(ctx);
  if (st != SIGMO_OK) {
    delete p;
    return st;
  }
  *out = p;
  return SIGMO_OK;
}

int sigmo_pipeline_execute(sigmo_pipeline* p, uint32_t refinement_steps, int find_first, sigmo_pipeline_result* result) {
  if (!p || !result) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  sigmo_ctx* ctx = p->ctx;
  SIGMO_TRY(bind(ctx));
  std::memset(result, 0, sizeof(*result));
  SIGMO_TRY(ctx->collect());
  double before[SIGMO_PHASE_COUNT];
  std::memcpy(before, ctx->accum_ms, sizeof(before));
  hipEvent_t t0, t1;
  SIGMO_TRY(ctx->event(&t0));
  SIGMO_TRY(ctx->event(&t1));
  SIGMO_HIP_TRY(hipEventRecord(t0, ctx->stream));
  uint64_t matches = 0;
  SIGMO_TRY(pipeline_enqueue(p, refinement_steps, find_first, &matches));
  SIGMO_HIP_TRY(hipEventRecord(t1, ctx->stream));
  SIGMO_HIP_TRY(hipStreamSynchronize(ctx->stream));  // timing events carry no fence: wait on the stream
  float total = 0;
  SIGMO_HIP_TRY(hipEventElapsedTime(&total, t0, t1));
  ctx->pool.push_back(t0);
  ctx->pool.push_back(t1);
  p->has_counts = true;
  result->num_matches = matches;
  result->gmcr_pairs = p->gm->P;
  result->matched_pairs = ctx->stats_host[1];
  result->dfs_steps = ctx->stats_host[4];
  result->join_wave_steps = ctx->stats_host[7];
  result->total_ms = total;
  SIGMO_TRY(ctx->collect());
  for (int ph = 0; ph < SIGMO_PHASE_COUNT; ++ph) result->phase_ms[ph] = ctx->accum_ms[ph] - before[ph];
  return SIGMO_OK;
}

int sigmo_pipeline_capture(sigmo_pipeline* p, uint32_t refinement_steps, int find_first) {
  if (!p) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  sigmo_ctx* ctx = p->ctx;
  SIGMO_TRY(bind(ctx));
  sigmo_pipeline_result r;
  SIGMO_TRY(sigmo_pipeline_execute(p, refinement_steps, find_first, &r));  // sizes every buffer
  p->drop_graph();
  if (!p->replay_host) SIGMO_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&p->replay_host), 64));
  SIGMO_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  uint64_t unused = 0;
  int st = pipeline_enqueue(p, refinement_steps, find_first, &unused);
  if (st == SIGMO_OK && hipMemcpyAsync(p->replay_host, p->gm->total, 8, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess)
    st = sigmo_err::set(SIGMO_ERR_HIP, "graph capture: total readback");
  ctx->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(ctx->stream, &g);
  if (st != SIGMO_OK || end != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    if (st != SIGMO_OK) return st;
    return sigmo_err::set(SIGMO_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(end));
  }
  p->graph = g;
  SIGMO_HIP_TRY(hipGraphInstantiate(&p->graph_exec, g, nullptr, nullptr, 0));
  p->graph_P = p->gm->P;
  p->graph_nw = p->gm->last_wide;
  return SIGMO_OK;
}

int sigmo_pipeline_replay(sigmo_pipeline* p, sigmo_pipeline_result* result) {
  if (!p || !result) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (!p->graph_exec) return sigmo_err::set(SIGMO_ERR_INVALID, "no captured graph (call sigmo_pipeline_capture)");
  sigmo_ctx* ctx = p->ctx;
  SIGMO_TRY(bind(ctx));
  std::memset(result, 0, sizeof(*result));
  hipEvent_t t0, t1;
  SIGMO_TRY(ctx->event(&t0));
  SIGMO_TRY(ctx->event(&t1));
  SIGMO_HIP_TRY(hipEventRecord(t0, ctx->stream));
  SIGMO_HIP_TRY(hipGraphLaunch(p->graph_exec, ctx->stream));
  SIGMO_HIP_TRY(hipEventRecord(t1, ctx->stream));
  SIGMO_HIP_TRY(hipStreamSynchronize(ctx->stream));  // timing events carry no fence: wait on the stream
  float total = 0;
  SIGMO_HIP_TRY(hipEventElapsedTime(&total, t0, t1));
  ctx->pool.push_back(t0);
  ctx->pool.push_back(t1);
  const unsigned long long* st = ctx->stats_host;
  if (p->replay_host[0] != p->graph_P || (p->d->max_nodes > 64 && st[5] != p->graph_nw)) {
    p->has_counts = false;
    return sigmo_err::set(SIGMO_ERR_INVALID, "the captured graph no longer matches the data (pair counts changed): capture again");
  }
  if (st[3]) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "a per-pair match count exceeded int32");
  p->has_counts = true;
  result->num_matches = st[0];
  result->gmcr_pairs = p->graph_P;
  result->matched_pairs = st[1];
  result->dfs_steps = st[4];
  result->join_wave_steps = st[7];
  result->total_ms = total;
  return SIGMO_OK;
}

int sigmo_pipeline_pairs(sigmo_pipeline* p, uint32_t* data_graph_offsets, uint64_t n_offsets, uint32_t* query_graph_indices,
                         int32_t* counts, uint64_t n_pairs) {
  if (!p || !p->has_counts) return sigmo_err::set(SIGMO_ERR_INVALID, "pipeline has not executed");
  SIGMO_TRY(bind(p->ctx));
  SIGMO_TRY(sigmo_gmcr_download(p->ctx, p->gm, data_graph_offsets, n_offsets, query_graph_indices, n_pairs));
  if (counts) {
    if (n_pairs != p->gm->P) return sigmo_err::set(SIGMO_ERR_INVALID, "n_pairs must be the pair count");
    if (n_pairs) {
      SIGMO_HIP_TRY(hipMemcpyAsync(counts, p->gm->pair_counts, n_pairs * 4, hipMemcpyDeviceToHost, p->ctx->stream));
      SIGMO_HIP_TRY(hipStreamSynchronize(p->ctx->stream));
    }
  }
  return SIGMO_OK;
}

}  // extern "C"

namespace sigmo_internal {
// Point a pipeline at another data graph over the same queries (a streaming chunk): the
// data-side signature and candidate buffers are reused when the new graph fits them, and grown
// to at least reserve_N data nodes otherwise.  With *p null the pipeline is created that way.
int pipeline_rebind(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, uint64_t reserve_N, sigmo_pipeline** pp) {
  if (!pp || !query || !data) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  SIGMO_TRY(bind(ctx));
  if (!*pp) {
    SIGMO_TRY(check_queries(query));
    auto* np = new sigmo_pipeline();
    np->ctx = ctx;
    np->q = query;
    np->d = data;
    int st = sigmo_gmcr_create(ctx, &np->gm);
    if (st != SIGMO_OK) {
      delete np;
      return st;
    }
    *pp = np;
  }
  sigmo_pipeline* p = *pp;
  if (p->ctx != ctx || p->q != query) return sigmo_err::set(SIGMO_ERR_INVALID, "pipeline bound to other queries or context");
  p->drop_graph();
  const uint64_t N = data->N, W64 = (N + 63) / 64, cN = std::max(N, reserve_N);
  if (!p->sig || N > p->sig->capD) {
    sigmo_signatures_free(p->sig);
    p->sig = nullptr;
    SIGMO_TRY(sigmo_signatures_create(ctx, cN, query->N, &p->sig));
  }
  p->sig->D = N;
  SIGMO_HIP_TRY(sigmo_k::fill_u64(ctx->stream, p->sig->dplanes, 64 * W64, 0));
  if (!p->cand || query->N * W64 > p->cand->cap_words) {
    sigmo_candidates_free(p->cand);
    p->cand = nullptr;
    SIGMO_TRY(sigmo_candidates_create(ctx, query->N, cN, &p->cand));
  }
  p->cand->D = N;
  p->cand->W64 = W64;
  p->d = data;
  p->has_counts = false;
  return SIGMO_OK;
}
int ctx_device(const sigmo_ctx* ctx) { return ctx->device; }

// Add src's phase and kernel timings to dst's and clear them in src (a streaming pass whose chunks
// ran on several contexts reports them on the caller's).  Both contexts are synchronised first.
int ctx_merge_timings(sigmo_ctx* dst, sigmo_ctx* src) {
  SIGMO_TRY(bind(src));
  SIGMO_TRY(src->collect());
  SIGMO_TRY(bind(dst));
  SIGMO_TRY(dst->collect());
  for (int p = 0; p < SIGMO_PHASE_COUNT; ++p) {
    dst->accum_ms[p] += src->accum_ms[p];
    src->accum_ms[p] = 0;
  }
  for (int k = 0; k < SIGMO_KERNEL_COUNT; ++k) {
    dst->k_launches[k] += src->k_launches[k];
    dst->k_ms[k] += src->k_ms[k];
    dst->k_bytes[k] += src->k_bytes[k];
    src->k_launches[k] = 0;
    src->k_ms[k] = 0;
    src->k_bytes[k] = 0;
  }
  return SIGMO_OK;
}
}  // namespace sigmo_internal

extern "C" {

int sigmo_pipeline_free(sigmo_pipeline* p) {
  if (!p) return SIGMO_OK;
  (void)hipSetDevice(p->ctx->device);
  delete p;
  return SIGMO_OK;
}

int sigmo_pipeline_run(sigmo_ctx* ctx, const sigmo_graph* query, const sigmo_graph* data, uint32_t refinement_steps, int find_first,
                       sigmo_pipeline_result* result) {
  SIGMO_TRY(bind(ctx));
  if (!query || !data || !result) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  SIGMO_TRY(check_queries(query));
  std::memset(result, 0, sizeof(*result));
  // setup (outside the timed region, as sigmo.cpp:75-91)
  sigmo_signatures* sig = nullptr;
  sigmo_candidates* cand = nullptr;
  sigmo_gmcr* gm = nullptr;
  struct Guard {
    sigmo_signatures*& s;
    sigmo_candidates*& c;
    sigmo_gmcr*& g;
    ~Guard() {
      sigmo_signatures_free(s);
      sigmo_candidates_free(c);
      sigmo_gmcr_free(g);
    }
  } guard{sig, cand, gm};
  SIGMO_TRY(sigmo_signatures_create(ctx, data->N, query->N, &sig));
  SIGMO_TRY(sigmo_candidates_create(ctx, query->N, data->N, &cand));
  SIGMO_TRY(sigmo_gmcr_create(ctx, &gm));
  SIGMO_TRY(sigmo_phase_reset(ctx));
  hipEvent_t t0, t1;
  SIGMO_TRY(ctx->event(&t0));
  SIGMO_TRY(ctx->event(&t1));
  SIGMO_HIP_TRY(hipEventRecord(t0, ctx->stream));
  SIGMO_TRY(sigmo_signatures_generate(ctx, sig, data, SIGMO_SCOPE_DATA));
  SIGMO_TRY(sigmo_signatures_generate(ctx, sig, query, SIGMO_SCOPE_QUERY));
  SIGMO_TRY(sigmo_candidates_filter(ctx, cand, query, data, sig));
  for (uint32_t r = 1; r <= refinement_steps; ++r) {
    SIGMO_TRY(sigmo_signatures_refine(ctx, sig, data, SIGMO_SCOPE_DATA, r));
    SIGMO_TRY(sigmo_signatures_refine(ctx, sig, query, SIGMO_SCOPE_QUERY, r));
    SIGMO_TRY(sigmo_candidates_refine(ctx, cand, query, data, sig));
  }
  SIGMO_TRY(sigmo_gmcr_generate(ctx, gm, query, data, cand));
  uint64_t matches = 0;
  SIGMO_TRY(sigmo_join(ctx, query, data, cand, gm, find_first, &matches, nullptr));
  SIGMO_HIP_TRY(hipEventRecord(t1, ctx->stream));
  SIGMO_HIP_TRY(hipStreamSynchronize(ctx->stream));  // timing events carry no fence: wait on the stream
  float total = 0;
  SIGMO_HIP_TRY(hipEventElapsedTime(&total, t0, t1));
  ctx->pool.push_back(t0);
  ctx->pool.push_back(t1);
  result->num_matches = matches;
  result->gmcr_pairs = gm->P;
  result->matched_pairs = ctx->stats_host[1];
  result->dfs_steps = ctx->stats_host[4];
  result->join_wave_steps = ctx->stats_host[7];
  result->total_ms = total;
  for (int p = 0; p < SIGMO_PHASE_COUNT; ++p) SIGMO_TRY(sigmo_phase_ms(ctx, p, &result->phase_ms[p]));
  return SIGMO_OK;
}

}  // extern "C"
