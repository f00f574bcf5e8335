// Host-side batched CSR container and helpers shared by the generator, the .dat reader
// and the device upload path.
//
// Layout follows the reference's concatenated device image (library/include/graph.hpp:62-70,
// 264-294): graph_offsets[G+1] prefix of node counts, row_offsets[N+1] global prefix of
// directed entries, column_indices[E] global node ids, node_labels[N], edge_labels[E].
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sigmo_hip.h"

struct sigmo_host_graphs {
  std::vector<uint32_t> graph_offsets{0};
  std::vector<uint32_t> row_offsets{0};
  std::vector<uint32_t> column_indices;
  std::vector<uint8_t> node_labels;
  std::vector<uint8_t> edge_labels;

  uint32_t num_graphs() const { return static_cast<uint32_t>(graph_offsets.size() - 1); }
  uint64_t total_nodes() const { return node_labels.size(); }
  uint64_t total_edges() const { return column_indices.size(); }
};

namespace sigmo_host {

// One graph as an undirected edge list in input order (graph.hpp:184-205 IntermediateGraph).
struct EdgeListGraph {
  std::vector<uint8_t> labels;
  std::vector<uint32_t> eu, ev;
  std::vector<uint8_t> el;
};

// Append one graph to a batch, building its CSR by the counting sort of
// graph.hpp:216-247 (each undirected edge appended to both rows in input order) and
// rebasing ids to global ones as createDeviceCSRGraph does (graph.hpp:282-289).
void append_graph(sigmo_host_graphs& batch, const EdgeListGraph& g);

// Append all graphs of `src` to `dst` (ids rebased).
void append_batch(sigmo_host_graphs& dst, const sigmo_host_graphs& src, uint32_t first, uint32_t count);

// Structural validation: monotone offsets, ids inside their own graph.  Returns "" if ok.
std::string validate(const sigmo_host_graphs& g);

}  // namespace sigmo_host

// Thread-local error recording shared by every translation unit of the library.
namespace sigmo_err {
int set(int code, const std::string& msg);
}
