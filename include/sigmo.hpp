// sigmo.hpp -- C++ API with the reference's namespaces, class and function names, over the
// C ABI of sigmo_hip.h (gfx950 kernels).  A reference driver such as library/src/sigmo.cpp
// compiles against it after replacing `sycl::queue` with `sigmo::Queue` (paths relative to
// the SIGMo repository).  Errors surface as std::runtime_error, as in the reference.
//
// Reference headers mirrored:
//   types.hpp:14-28            sigmo::types
//   device.hpp:14-21           sigmo::CandidatesDomain, sigmo::device::deviceOptions
//   utils.hpp:127-161          sigmo::utils::BatchedEvent
//   graph.hpp:37-112,255-432   sigmo::CSRGraph, DeviceBatchedCSRGraph, createDeviceCSRGraph, ...
//   graph.hpp:20-35,114-148,345-421  sigmo::AMGraph, DeviceBatchedAMGraph, createDeviceAMGraph, ...
//   utils.hpp:17-62            sigmo::utils::getNumOfAdjacencyIntegers, utils::adjacency_matrix (host)
//   io.hpp:17-55               sigmo::io::loadCSRGraphsFromFile / FromLines, loadAMGraphsFromFile / FromLines
//   signature.hpp:17-212       sigmo::signature::Signature<PowerGraph, 4>
//   candidates.hpp:14-204      sigmo::candidates::Candidates
//   isomorphism.hpp:22-127     sigmo::isomorphism::filter::{filterCandidates, refineCandidates}
//   gmcr.hpp:18-135            sigmo::isomorphism::mapping::GMCR
//   isomorphism.hpp:270-381    sigmo::isomorphism::join::joinCandidates
//   isomorphism.hpp:383-556    sigmo::isomorphism::join::joinWildcardCandidates
#pragma once

#include <chrono>
#include <cmath>
#include <cstdint>
#include <fstream>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "sigmo_hip.h"

namespace sigmo {

namespace types {
using row_offset_t = uint32_t;
using col_index_t = uint32_t;
using label_t = uint8_t;
using node_t = uint32_t;
using candidates_t = uint32_t;
using adjacency_t = uint64_t;
constexpr label_t WILDCARD_NODE = 255;
constexpr size_t MAX_NEIGHBORS = 4;
constexpr node_t NULL_NODE = 0xFFFFFFFFu;
}  // namespace types

enum class CandidatesDomain { Query, Data };

namespace detail {
inline void check(int status) {
  if (status != SIGMO_OK) throw std::runtime_error(sigmo_last_error());
}
}  // namespace detail

// The device + stream every reference class holds a sycl::queue for.
class Queue {
 public:
  explicit Queue(int device = 0) { detail::check(sigmo_ctx_create(device, &ctx_)); }
  ~Queue() { sigmo_ctx_destroy(ctx_); }
  Queue(const Queue&) = delete;
  Queue& operator=(const Queue&) = delete;
  sigmo_ctx* get() const { return ctx_; }
  void wait() { detail::check(sigmo_ctx_sync(ctx_)); }
  void wait_and_throw() { wait(); }
  std::string deviceName() const {
    char buf[256];
    detail::check(sigmo_device_info(ctx_, buf, sizeof(buf), nullptr));
    return buf;
  }
  uint64_t globalMemSize() const {
    uint64_t m = 0;
    detail::check(sigmo_device_info(ctx_, nullptr, 0, &m));
    return m;
  }

 private:
  sigmo_ctx* ctx_ = nullptr;
};

namespace device {
// Accepted for CLI compatibility (device.hpp:17-21); gfx950 launch shapes are chosen by the
// library.
struct DeviceOptions {
  size_t join_work_group_size = 128;
  size_t filter_work_group_size = 512;
};
inline DeviceOptions deviceOptions;
}  // namespace device

namespace utils {
// The kernels launched by one API call; getProfilingInfo sums their device durations.
class BatchedEvent {
 public:
  BatchedEvent() = default;
  BatchedEvent(sigmo_ctx* ctx, uint64_t from, uint64_t to) : ctx_(ctx), from_(from), to_(to) {}
  void wait() {
    if (ctx_) detail::check(sigmo_ctx_sync(ctx_));
  }
  size_t numEvents() const { return static_cast<size_t>(to_ - from_); }
  std::chrono::duration<double> getProfilingInfo() {
    double ms = 0;
    if (ctx_) detail::check(sigmo_event_elapsed(ctx_, from_, to_, &ms));
    return std::chrono::duration<double>(ms / 1e3);
  }

 private:
  sigmo_ctx* ctx_ = nullptr;
  uint64_t from_ = 0, to_ = 0;
};

// Words of an n-node adjacency matrix, (n*n + 63) / 64 as uint8 (utils.hpp:17-21).
inline uint8_t getNumOfAdjacencyIntegers(uint8_t num_nodes) {
  return static_cast<uint8_t>((uint32_t(num_nodes) * num_nodes + 63) / 64);
}

// Host forms of the matrix accessors (utils.hpp:23-62): the row width is
// (uint16)sqrt(float(64 * words)) and bit u*W + v marks the edge u-v.
namespace adjacency_matrix {
inline uint16_t rowWidth(uint8_t adjacency_matrix_size) {
  return static_cast<uint16_t>(std::sqrt(static_cast<float>(64 * uint32_t(adjacency_matrix_size))));
}
inline void setBit(types::adjacency_t* m, uint8_t adjacency_matrix_size, uint8_t u, uint8_t v) {
  const uint32_t W = rowWidth(adjacency_matrix_size), iu = u * W + v, iv = v * W + u;
  m[iu / 64] |= types::adjacency_t(1) << (iu % 64);
  m[iv / 64] |= types::adjacency_t(1) << (iv % 64);
}
inline bool isNeighbor(const types::adjacency_t* m, uint8_t adjacency_matrix_size, types::node_t node_id, types::node_t neighbor_id,
                       types::node_t previous_nodes = 0) {
  const uint32_t W = rowWidth(adjacency_matrix_size);
  const uint32_t idx = uint8_t(node_id - previous_nodes) * W + uint8_t(neighbor_id - previous_nodes);
  return (m[idx / 64] >> (idx % 64)) & 1u;
}
// The row's neighbours in ascending order (global ids), at most MAX_NEIGHBORS of them (the
// reference's callers read no more), NULL_NODE-terminated when fewer.
inline void getNeighbors(const types::adjacency_t* m, uint8_t adjacency_matrix_size, types::node_t node_id, types::node_t* neighbors,
                         types::node_t previous_nodes = 0) {
  const uint32_t W = rowWidth(adjacency_matrix_size), u = uint8_t(node_id - previous_nodes);
  size_t cnt = 0;
  for (uint32_t v = 0; v < W && cnt < types::MAX_NEIGHBORS; ++v) {
    const uint32_t idx = u * W + v;
    if ((m[idx / 64] >> (idx % 64)) & 1u) neighbors[cnt++] = v + previous_nodes;
  }
  if (cnt < types::MAX_NEIGHBORS) neighbors[cnt] = types::NULL_NODE;
}
}  // namespace adjacency_matrix

// Records the launch range of one call.
class EventScope {
 public:
  explicit EventScope(sigmo_ctx* ctx) : ctx_(ctx) { detail::check(sigmo_event_mark(ctx_, &from_)); }
  BatchedEvent done() {
    uint64_t to = 0;
    detail::check(sigmo_event_mark(ctx_, &to));
    return BatchedEvent(ctx_, from_, to);
  }

 private:
  sigmo_ctx* ctx_;
  uint64_t from_ = 0;
};
}  // namespace utils

// Host graph (graph.hpp:37-60).
class CSRGraph {
 public:
  CSRGraph(std::vector<types::row_offset_t> row_offsets, std::vector<types::col_index_t> column_indices,
           std::vector<types::label_t> node_labels, std::vector<types::label_t> edge_labels, size_t num_nodes)
      : row_offsets_(std::move(row_offsets)),
        column_indices_(std::move(column_indices)),
        node_labels_(std::move(node_labels)),
        edge_labels_(std::move(edge_labels)),
        num_nodes_(num_nodes) {}
  const types::row_offset_t* getRowOffsets() const { return row_offsets_.data(); }
  const types::col_index_t* getColumnIndices() const { return column_indices_.data(); }
  const types::label_t* getNodeLabels() const { return node_labels_.data(); }
  const types::label_t* getEdgeLabels() const { return edge_labels_.data(); }
  size_t getNumNodes() const { return num_nodes_; }

 private:
  std::vector<types::row_offset_t> row_offsets_;
  std::vector<types::col_index_t> column_indices_;
  std::vector<types::label_t> node_labels_;
  std::vector<types::label_t> edge_labels_;
  size_t num_nodes_;
};

// Device batched CSR (graph.hpp:62-70); the arrays live behind `handle`.
struct DeviceBatchedCSRGraph {
  sigmo_graph* handle = nullptr;
  uint32_t num_graphs = 0;
  size_t total_nodes = 0;
  size_t total_edges = 0;
};

// Host adjacency-matrix graph (graph.hpp:20-35) of at most 64 nodes.
class AMGraph {
 public:
  AMGraph(std::vector<types::adjacency_t> adjacency, std::vector<types::label_t> node_labels, uint8_t num_nodes)
      : adjacency_(std::move(adjacency)), node_labels_(std::move(node_labels)), num_nodes_(num_nodes) {}
  types::adjacency_t* getAdjacencyMatrix() { return adjacency_.data(); }
  const types::adjacency_t* getAdjacencyMatrix() const { return adjacency_.data(); }
  types::label_t* getNodeLabels() { return node_labels_.data(); }
  const types::label_t* getNodeLabels() const { return node_labels_.data(); }
  int getNumNodes() const { return num_nodes_; }

 private:
  std::vector<types::adjacency_t> adjacency_;
  std::vector<types::label_t> node_labels_;
  uint8_t num_nodes_;
};

// Device batched adjacency-matrix graphs (graph.hpp:114-121); the arrays live behind `handle`.
struct DeviceBatchedAMGraph {
  sigmo_am_graph* handle = nullptr;
  uint32_t num_graphs = 0;
  size_t total_nodes = 0;
};

namespace detail {
// Owning wrapper of a host batch.
struct HostBatch {
  sigmo_host_graphs* h = nullptr;
  ~HostBatch() { sigmo_host_graphs_free(h); }
};

inline std::vector<CSRGraph> split(const sigmo_host_graphs* b) {
  uint32_t G = 0;
  uint64_t N = 0, E = 0;
  check(sigmo_host_graphs_info(b, &G, &N, &E));
  const uint32_t *go, *ro, *ci;
  const uint8_t *nl, *el;
  check(sigmo_host_graphs_arrays(b, &go, &ro, &ci, &nl, &el));
  std::vector<CSRGraph> out;
  out.reserve(G);
  for (uint32_t g = 0; g < G; ++g) {
    const uint32_t lo = go[g], hi = go[g + 1], n = hi - lo;
    const uint32_t e0 = ro[lo], e1 = ro[hi];
    std::vector<uint32_t> rows(n + 1), cols(e1 - e0);
    for (uint32_t i = 0; i <= n; ++i) rows[i] = ro[lo + i] - e0;
    for (uint32_t e = e0; e < e1; ++e) cols[e - e0] = ci[e] - lo;
    out.emplace_back(std::move(rows), std::move(cols), std::vector<uint8_t>(nl + lo, nl + hi),
                     std::vector<uint8_t>(el + e0, el + e1), n);
  }
  return out;
}
}  // namespace detail

namespace detail {
// toAMGraph (graph.hpp:208-214) of every graph of a host batch.
inline std::vector<AMGraph> to_am(const sigmo_host_graphs* b) {
  std::vector<AMGraph> out;
  for (const CSRGraph& g : split(b)) {
    const size_t n = g.getNumNodes();
    if (n > 64) throw std::runtime_error("adjacency-matrix graphs hold at most 64 nodes");
    const uint8_t words = utils::getNumOfAdjacencyIntegers(static_cast<uint8_t>(n));
    std::vector<types::adjacency_t> m(words, 0);
    for (size_t u = 0; u < n; ++u)
      for (uint32_t e = g.getRowOffsets()[u]; e < g.getRowOffsets()[u + 1]; ++e)
        utils::adjacency_matrix::setBit(m.data(), words, static_cast<uint8_t>(u), static_cast<uint8_t>(g.getColumnIndices()[e]));
    out.emplace_back(std::move(m), std::vector<types::label_t>(g.getNodeLabels(), g.getNodeLabels() + n), static_cast<uint8_t>(n));
  }
  return out;
}
}  // namespace detail

namespace io {
// io.hpp:17-35: query graphs in the adjacency-matrix form.
inline std::vector<AMGraph> loadAMGraphsFromFile(const std::string& filename) {
  detail::HostBatch b;
  detail::check(sigmo_host_graphs_load_dat(filename.c_str(), &b.h));
  return detail::to_am(b.h);
}
inline std::vector<AMGraph> loadAMGraphsFromLines(const std::vector<std::string>& lines) {
  std::string text;
  for (const auto& l : lines) text += l + "\n";
  detail::HostBatch b;
  detail::check(sigmo_host_graphs_parse_dat(text.data(), text.size(), &b.h));
  return detail::to_am(b.h);
}
inline std::vector<CSRGraph> loadCSRGraphsFromFile(const std::string& filename) {
  detail::HostBatch b;
  detail::check(sigmo_host_graphs_load_dat(filename.c_str(), &b.h));
  return detail::split(b.h);
}
inline std::vector<CSRGraph> loadCSRGraphsFromLines(const std::vector<std::string>& lines) {
  std::string text;
  for (const auto& l : lines) text += l + "\n";
  detail::HostBatch b;
  detail::check(sigmo_host_graphs_parse_dat(text.data(), text.size(), &b.h));
  return detail::split(b.h);
}
// loadFileLinesMPI (src/sigmo_mpi.cpp:18-80) with the rank and the rank count passed in instead
// of read from MPI_COMM_WORLD: the lines this rank's byte range of the file yields, filtered as
// the reference filters them (sigmo_dat_shard_lines).
inline std::vector<std::string> loadFileLinesShard(const std::string& filename, int max_lines, int rank, int nprocs) {
  if (max_lines < 0 || rank < 0 || nprocs <= 0) throw std::runtime_error("negative line count, rank or rank count");
  size_t len = 0;
  uint32_t kept = 0;
  detail::check(sigmo_dat_shard_lines(filename.c_str(), static_cast<uint32_t>(rank), static_cast<uint32_t>(nprocs),
                                      static_cast<uint32_t>(max_lines), nullptr, 0, &len, &kept));
  std::string text(len, '\0');
  detail::check(sigmo_dat_shard_lines(filename.c_str(), static_cast<uint32_t>(rank), static_cast<uint32_t>(nprocs),
                                      static_cast<uint32_t>(max_lines), text.empty() ? nullptr : &text[0], text.size(), &len, &kept));
  if (len != text.size()) throw std::runtime_error(filename + " changed between the two reads of its shard");
  std::vector<std::string> lines;
  size_t at = 0;
  while (at < text.size()) {
    const size_t e = text.find('\n', at);
    lines.push_back(text.substr(at, e - at));
    at = e + 1;
  }
  return lines;
}
}  // namespace io

// createDeviceCSRGraph (graph.hpp:255-314): concatenate with global ids, upload.
inline DeviceBatchedCSRGraph createDeviceCSRGraph(Queue& queue, std::vector<CSRGraph>& graphs) {
  std::vector<uint32_t> go{0}, ro{0}, ci;
  std::vector<uint8_t> nl, el;
  for (const auto& g : graphs) {
    const uint32_t base = go.back(), cbase = static_cast<uint32_t>(ci.size());
    const size_t n = g.getNumNodes();
    const uint32_t m = g.getRowOffsets()[n];
    for (size_t i = 1; i <= n; ++i) ro.push_back(g.getRowOffsets()[i] + cbase);
    for (uint32_t e = 0; e < m; ++e) {
      ci.push_back(g.getColumnIndices()[e] + base);
      el.push_back(g.getEdgeLabels()[e]);
    }
    nl.insert(nl.end(), g.getNodeLabels(), g.getNodeLabels() + n);
    go.push_back(base + static_cast<uint32_t>(n));
  }
  detail::HostBatch b;
  detail::check(sigmo_host_graphs_from_arrays(static_cast<uint32_t>(graphs.size()), go.data(), ro.data(), ci.data(), nl.data(),
                                              el.data(), &b.h));
  DeviceBatchedCSRGraph d;
  detail::check(sigmo_graph_upload(queue.get(), b.h, &d.handle));
  uint64_t N = 0, E = 0;
  detail::check(sigmo_graph_info(d.handle, &d.num_graphs, &N, &E));
  d.total_nodes = N;
  d.total_edges = E;
  return d;
}

inline void destroyDeviceCSRGraph(DeviceBatchedCSRGraph& g, Queue&) {
  sigmo_graph_free(g.handle);
  g.handle = nullptr;
}

// createDeviceAMGraph (graph.hpp:345-392): the matrices' edges (each row's set bits) go up as a
// host batch and the library lays the device image out again from them; for matrices made
// by toAMGraph (symmetric) the image holds the same bits.
inline DeviceBatchedAMGraph createDeviceAMGraph(Queue& queue, std::vector<AMGraph>& graphs) {
  std::vector<uint32_t> go{0}, ro{0}, ci;
  std::vector<uint8_t> nl;
  for (const auto& g : graphs) {
    const uint32_t base = go.back(), n = static_cast<uint32_t>(g.getNumNodes());
    const uint8_t words = utils::getNumOfAdjacencyIntegers(static_cast<uint8_t>(n));
    const uint32_t W = utils::adjacency_matrix::rowWidth(words);
    for (uint32_t u = 0; u < n; ++u) {
      for (uint32_t v = 0; v < n; ++v) {
        const uint32_t idx = u * W + v;
        if ((g.getAdjacencyMatrix()[idx / 64] >> (idx % 64)) & 1u) ci.push_back(base + v);
      }
      ro.push_back(static_cast<uint32_t>(ci.size()));
    }
    nl.insert(nl.end(), g.getNodeLabels(), g.getNodeLabels() + n);
    go.push_back(base + n);
  }
  std::vector<uint8_t> el(ci.size(), 0);
  detail::HostBatch b;
  detail::check(sigmo_host_graphs_from_arrays(static_cast<uint32_t>(graphs.size()), go.data(), ro.data(), ci.data(), nl.data(),
                                              el.data(), &b.h));
  DeviceBatchedAMGraph d;
  detail::check(sigmo_am_graph_upload(queue.get(), b.h, &d.handle));
  uint64_t N = 0;
  detail::check(sigmo_am_graph_info(d.handle, &d.num_graphs, &N, nullptr));
  d.total_nodes = N;
  return d;
}

// destroyDeviceAMGraph (graph.hpp:394-399).
inline void destroyDeviceAMGraph(DeviceBatchedAMGraph& g, Queue&) {
  sigmo_am_graph_free(g.handle);
  g.handle = nullptr;
}

// Bytes of the reference CSR arrays (graph.hpp:324-328).
inline size_t getDeviceGraphAllocSize(const DeviceBatchedCSRGraph& g) {
  uint64_t csr = 0;
  detail::check(sigmo_graph_alloc_size(g.handle, &csr, nullptr));
  return csr;
}

namespace signature {
enum class Algorithm { ViewBased, PowerGraph };
enum class SignatureScope { Data, Query };

template <Algorithm A = Algorithm::PowerGraph, size_t Bits = 4>
class Signature {
  static_assert(Bits == 4, "signatures are 16 x 4-bit counts");

 public:
  Signature(Queue& queue, size_t data_nodes, size_t query_nodes) : queue_(queue) {
    detail::check(sigmo_signatures_create(queue.get(), data_nodes, query_nodes, &s_));
    data_nodes_ = data_nodes;
    query_nodes_ = query_nodes;
  }
  ~Signature() { sigmo_signatures_free(s_); }
  Signature(const Signature&) = delete;

  utils::BatchedEvent generateDataSignatures(DeviceBatchedCSRGraph& g) { return gen(g, SIGMO_SCOPE_DATA); }
  utils::BatchedEvent generateQuerySignatures(DeviceBatchedCSRGraph& g) { return gen(g, SIGMO_SCOPE_QUERY); }
  utils::BatchedEvent refineDataSignatures(DeviceBatchedCSRGraph& g, size_t view_size = 1) { return ref(g, SIGMO_SCOPE_DATA, view_size); }
  utils::BatchedEvent refineQuerySignatures(DeviceBatchedCSRGraph& g, size_t view_size = 1) { return ref(g, SIGMO_SCOPE_QUERY, view_size); }
  // The adjacency-matrix overloads (signature.hpp:56-96 dispatch on the graph type).
  utils::BatchedEvent generateDataSignatures(DeviceBatchedAMGraph& g) { return gen(g, SIGMO_SCOPE_DATA); }
  utils::BatchedEvent generateQuerySignatures(DeviceBatchedAMGraph& g) { return gen(g, SIGMO_SCOPE_QUERY); }
  utils::BatchedEvent refineDataSignatures(DeviceBatchedAMGraph& g, size_t view_size = 1) { return ref(g, SIGMO_SCOPE_DATA, view_size); }
  utils::BatchedEvent refineQuerySignatures(DeviceBatchedAMGraph& g, size_t view_size = 1) { return ref(g, SIGMO_SCOPE_QUERY, view_size); }

  std::vector<uint64_t> getDataSignatures() const { return download(SIGMO_SCOPE_DATA, data_nodes_); }
  std::vector<uint64_t> getQuerySignatures() const { return download(SIGMO_SCOPE_QUERY, query_nodes_); }
  size_t getMaxLabels() const { return 64 / Bits; }
  size_t getDataSignatureAllocationSize() const { return alloc(SIGMO_SCOPE_DATA); }
  size_t getQuerySignatureAllocationSize() const { return alloc(SIGMO_SCOPE_QUERY); }
  sigmo_signatures* handle() const { return s_; }

 private:
  utils::BatchedEvent gen(DeviceBatchedCSRGraph& g, int scope) {
    utils::EventScope ev(queue_.get());
    detail::check(sigmo_signatures_generate(queue_.get(), s_, g.handle, scope));
    return ev.done();
  }
  utils::BatchedEvent ref(DeviceBatchedCSRGraph& g, int scope, size_t view) {
    utils::EventScope ev(queue_.get());
    if constexpr (A == Algorithm::PowerGraph)  // signature.hpp:336-390
      detail::check(sigmo_signatures_refine(queue_.get(), s_, g.handle, scope, static_cast<uint32_t>(view)));
    else  // signature.hpp:298-334
      detail::check(sigmo_signatures_refine_view_based(queue_.get(), s_, g.handle, scope, static_cast<uint32_t>(view)));
    return ev.done();
  }
  utils::BatchedEvent gen(DeviceBatchedAMGraph& g, int scope) {  // signature.hpp:100-120
    utils::EventScope ev(queue_.get());
    detail::check(sigmo_signatures_generate_am(queue_.get(), s_, g.handle, scope));
    return ev.done();
  }
  utils::BatchedEvent ref(DeviceBatchedAMGraph& g, int scope, size_t view) {
    utils::EventScope ev(queue_.get());
    if constexpr (A == Algorithm::PowerGraph)  // signature.hpp:250-296
      detail::check(sigmo_signatures_refine_am(queue_.get(), s_, g.handle, scope, static_cast<uint32_t>(view)));
    else  // signature.hpp:214-248
      detail::check(sigmo_signatures_refine_am_view_based(queue_.get(), s_, g.handle, scope, static_cast<uint32_t>(view)));
    return ev.done();
  }
  std::vector<uint64_t> download(int scope, size_t n) const {
    std::vector<uint64_t> out(n);
    detail::check(sigmo_signatures_download(queue_.get(), s_, scope, out.data(), n));
    return out;
  }
  size_t alloc(int scope) const {
    uint64_t b = 0;
    detail::check(sigmo_signatures_alloc_size(s_, scope, &b));
    return b;
  }
  Queue& queue_;
  sigmo_signatures* s_ = nullptr;
  size_t data_nodes_ = 0, query_nodes_ = 0;
};
}  // namespace signature

namespace candidates {
class Candidates {
 public:
  // Host copy in the reference layout (candidates.hpp:16-28,77-104).
  struct CandidatesDevice {
    std::vector<types::candidates_t> words;
    types::candidates_t* candidates = nullptr;
    size_t source_nodes = 0, target_nodes = 0, single_node_size = 0;
    static constexpr types::candidates_t num_bits = 32;
    uint32_t getCandidatesCount(types::node_t s) const {
      uint32_t c = 0;
      for (size_t i = 0; i < single_node_size; ++i) c += static_cast<uint32_t>(__builtin_popcount(candidates[s * single_node_size + i]));
      return c;
    }
    bool contains(types::node_t s, types::node_t c) const {
      return (candidates[s * single_node_size + c / 32] >> (c % 32)) & 1u;
    }
  };

  Candidates(Queue& queue, size_t source_nodes, size_t target_nodes) : queue_(queue) {
    detail::check(sigmo_candidates_create(queue.get(), source_nodes, target_nodes, &c_));
    source_ = source_nodes;
    target_ = target_nodes;
  }
  ~Candidates() { sigmo_candidates_free(c_); }
  Candidates(const Candidates&) = delete;

  CandidatesDevice getHostCandidates() const {
    CandidatesDevice h;
    h.source_nodes = source_;
    h.target_nodes = target_;
    h.single_node_size = (target_ + 31) / 32;
    h.words.resize(source_ * h.single_node_size);
    detail::check(sigmo_candidates_download(queue_.get(), c_, h.words.data(), h.words.size()));
    h.candidates = h.words.data();
    return h;
  }
  std::vector<uint32_t> getCandidatesCounts() const {
    std::vector<uint32_t> out(source_);
    detail::check(sigmo_candidates_counts(queue_.get(), c_, out.data(), out.size()));
    return out;
  }
  size_t getCandidatesCount(types::node_t source_node) const { return getCandidatesCounts().at(source_node); }
  size_t getAllocationSize() const {
    uint64_t ref = 0;
    detail::check(sigmo_candidates_alloc_size(c_, &ref, nullptr));
    return ref;
  }
  sigmo_candidates* handle() const { return c_; }

 private:
  Queue& queue_;
  sigmo_candidates* c_ = nullptr;
  size_t source_ = 0, target_ = 0;
};
}  // namespace candidates

namespace isomorphism {
namespace filter {
// CandidatesDomain::Data stores the transposed bitset: construct the candidates with the data
// node count as source_nodes and the query node count as target_nodes (isomorphism.hpp:49-52).
template <CandidatesDomain D = CandidatesDomain::Query>
utils::BatchedEvent filterCandidates(Queue& queue, DeviceBatchedCSRGraph& query_graph, DeviceBatchedCSRGraph& data_graph,
                                     signature::Signature<>& signatures, candidates::Candidates& candidates) {
  utils::EventScope ev(queue.get());
  if constexpr (D == CandidatesDomain::Query)
    detail::check(sigmo_candidates_filter(queue.get(), candidates.handle(), query_graph.handle, data_graph.handle, signatures.handle()));
  else
    detail::check(sigmo_candidates_filter_data(queue.get(), candidates.handle(), query_graph.handle, data_graph.handle));
  return ev.done();
}

template <CandidatesDomain D = CandidatesDomain::Query>
utils::BatchedEvent refineCandidates(Queue& queue, DeviceBatchedCSRGraph& query_graph, DeviceBatchedCSRGraph& data_graph,
                                     signature::Signature<>& signatures, candidates::Candidates& candidates) {
  // The reference's Data instantiation runs the Query-layout loop on a transposed bitset
  // (isomorphism.hpp:64-127 never reads D); only the Query domain is meaningful here.
  static_assert(D == CandidatesDomain::Query, "refineCandidates works on Query-domain candidates");
  utils::EventScope ev(queue.get());
  detail::check(sigmo_candidates_refine(queue.get(), candidates.handle(), query_graph.handle, data_graph.handle, signatures.handle()));
  return ev.done();
}
}  // namespace filter

namespace mapping {
class GMCR {
 public:
  struct GMCRHost {
    std::vector<uint32_t> data_graph_offsets;
    std::vector<uint32_t> query_graph_indices;
    size_t total_query_indices = 0;
  };
  explicit GMCR(Queue& queue) : queue_(queue) { detail::check(sigmo_gmcr_create(queue.get(), &m_)); }
  ~GMCR() { sigmo_gmcr_free(m_); }
  GMCR(const GMCR&) = delete;

  utils::BatchedEvent generateGMCR(DeviceBatchedCSRGraph& query_graphs, DeviceBatchedCSRGraph& data_graphs,
                                   candidates::Candidates& candidates) {
    utils::EventScope ev(queue_.get());
    detail::check(sigmo_gmcr_generate(queue_.get(), m_, query_graphs.handle, data_graphs.handle, candidates.handle()));
    num_data_graphs_ = data_graphs.num_graphs;
    return ev.done();
  }
  size_t getTotalQueryIndices() const {
    uint64_t p = 0;
    detail::check(sigmo_gmcr_total(m_, &p));
    return p;
  }
  GMCRHost getGMCR() const {
    GMCRHost h;
    h.total_query_indices = getTotalQueryIndices();
    h.data_graph_offsets.resize(num_data_graphs_ + 1);
    h.query_graph_indices.resize(h.total_query_indices);
    detail::check(sigmo_gmcr_download(queue_.get(), m_, h.data_graph_offsets.data(), h.data_graph_offsets.size(),
                                      h.query_graph_indices.data(), h.query_graph_indices.size()));
    return h;
  }
  sigmo_gmcr* handle() const { return m_; }

 private:
  Queue& queue_;
  sigmo_gmcr* m_ = nullptr;
  uint32_t num_data_graphs_ = 0;
};
}  // namespace mapping

namespace join {
// Adds the number of matches to *num_matches (isomorphism.hpp:374-375 accumulate).
inline utils::BatchedEvent joinCandidates(Queue& queue, DeviceBatchedCSRGraph& query_graphs, DeviceBatchedCSRGraph& data_graphs,
                                          candidates::Candidates& candidates, mapping::GMCR& gmcr, size_t* num_matches,
                                          bool find_first = true, std::vector<int32_t>* per_pair_counts = nullptr) {
  utils::EventScope ev(queue.get());
  uint64_t total = 0;
  if (per_pair_counts) per_pair_counts->assign(gmcr.getTotalQueryIndices(), 0);
  detail::check(sigmo_join(queue.get(), query_graphs.handle, data_graphs.handle, candidates.handle(), gmcr.handle(), find_first ? 1 : 0,
                           &total, per_pair_counts ? per_pair_counts->data() : nullptr));
  if (num_matches) *num_matches += total;
  return ev.done();
}

// isomorphism.hpp:428-556 (with defineMatchingOrder :383-408 and the edge-label isValidMapping
// :410-426).
inline utils::BatchedEvent joinWildcardCandidates(Queue& queue, DeviceBatchedCSRGraph& query_graphs, DeviceBatchedCSRGraph& data_graphs,
                                                  candidates::Candidates& candidates, mapping::GMCR& gmcr, size_t* num_matches,
                                                  bool find_first = true, std::vector<int32_t>* per_pair_counts = nullptr) {
  utils::EventScope ev(queue.get());
  uint64_t total = 0;
  if (per_pair_counts) per_pair_counts->assign(gmcr.getTotalQueryIndices(), 0);
  detail::check(sigmo_join_wildcard(queue.get(), query_graphs.handle, data_graphs.handle, candidates.handle(), gmcr.handle(),
                                    find_first ? 1 : 0, &total, per_pair_counts ? per_pair_counts->data() : nullptr));
  if (num_matches) *num_matches += total;
  return ev.done();
}
}  // namespace join
}  // namespace isomorphism
}  // namespace sigmo
