// Launch interface of the gfx950 kernels (sigmo_amd/csrc/kernels.hip).
//
// Device data layout (all arrays in HBM, 64-bit words little-endian):
//   candidates   bits[q * W64 + w], W64 = ceil(D/64); bit j of word w <=> data node 64w+j is
//                a candidate of query node q.  Byte-identical to the reference's uint32
//                rows (candidates.hpp:18-36) whenever ceil(D/32) is even.
//   adjacency    adj[2v], adj[2v+1]: bits of v's neighbours by local id (graph-relative,
//                < 128), low and high half; fold[v] = adj[2v] | adj[2v+1], the 64-position
//                view the signature BFS expands through.
//   label planes lplanes[w * 8 + b]: bit j = bit b of node_labels[64w + j].
//   label masks  lmask[g * 16 + l]: bit b = (node_labels[go[g] + b] == l), b < min(n, 64),
//                the 64-position view the reference's Bitset<uint64_t> BFS reads labels
//                through (signature.hpp:378-383).
//   sig planes   planes[p * W64 + w] (p = 4*label + bit): bit-sliced packed signatures.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sigmo_k {

struct GraphDev {
  const uint32_t* go;  // graph_offsets [G+1]
  const uint32_t* ro;  // row_offsets   [N+1]
  const uint32_t* ci;  // column_indices[E]
  const uint8_t* nl;   // node_labels   [N]
  const uint32_t* node_graph;  // [N]
  const uint64_t* adj;         // [2N]
  const uint64_t* fold;        // [N]  adj[2v] | adj[2v+1]
  const uint64_t* lplanes;     // [8 * ceil(N/64)]
  const uint64_t* lmask;       // [16 G]
  uint32_t G;
  uint64_t N;
};

// A batch of adjacency-matrix graphs (DeviceBatchedAMGraph, graph.hpp:114-148).
struct AMGraphDev {
  const uint32_t* go;   // graph_offsets [G]: first adjacency word of each graph
  const uint32_t* nn;   // num_nodes     [G]: cumulative node counts
  const uint64_t* adj;  // adjacency bit matrices
  const uint8_t* nl;    // node_labels   [N]
  uint32_t G;
  uint64_t N;
};

// Labels < 16 that may carry a nonzero signature count, compacted into slots.
struct LabelSlots {
  uint8_t label[16];
  int count;
};

hipError_t fill_u64(hipStream_t s, uint64_t* p, uint64_t n, uint64_t v);
hipError_t fill_u32(hipStream_t s, uint32_t* p, uint64_t n, uint32_t v);
// p[i] -= v (rebasing ids of a sub-range of a host batch)
hipError_t sub_u32(hipStream_t s, uint32_t* p, uint64_t n, uint32_t v);

// Derived per-node / per-graph arrays of an uploaded batch.
hipError_t derive_graph(hipStream_t s, const uint32_t* go, uint32_t G, const uint32_t* ro, const uint32_t* ci, const uint8_t* nl,
                        uint64_t N, uint32_t* node_graph, uint64_t* adj, uint64_t* fold, uint64_t* lplanes, uint64_t* lmask);

// generateCSRSignatures: increments the neighbour-label histogram onto sigs (onto zero with
// `fresh`, whatever sigs holds).  For the signature kernels below, sigs may be null when planes
// is not (and, for sig_generate, `fresh` is set): only the bit-sliced planes are written.
hipError_t sig_generate(hipStream_t s, const GraphDev& g, uint64_t* sigs, uint64_t* planes, bool fresh = false);
// refineCSRSignatures<PowerGraph>(view_size); frontier_out (nullable) receives the last BFS layer.
// With `fresh_reach` the reachables are taken as empty whatever `reach` holds.
hipError_t sig_refine(hipStream_t s, const GraphDev& g, uint32_t view_size, uint64_t* sigs, uint64_t* reach, uint64_t* frontier_out,
                      uint64_t* planes, bool fresh_reach = false);
// The same for view_size = M+1 when reach/frontier hold the state left by view_size = M.
hipError_t sig_refine_next(hipStream_t s, const GraphDev& g, uint64_t* sigs, uint64_t* reach, uint64_t* frontier, uint64_t* planes);

// refineCSRSignatures<ViewBased>(view_size): tmp holds the signatures before the pass.
hipError_t sig_refine_view_based(hipStream_t s, const GraphDev& g, uint32_t view_size, const uint64_t* tmp, uint64_t* sigs, uint64_t* planes);
// generateAMSignatures / refineAMSignatures<ViewBased> / refineAMSignatures<PowerGraph> over an
// AM batch; `tmp` holds the signatures before a ViewBased pass.
hipError_t am_generate(hipStream_t s, const AMGraphDev& g, uint64_t* sigs, uint64_t* planes);
hipError_t am_refine_view_based(hipStream_t s, const AMGraphDev& g, uint32_t view_size, const uint64_t* tmp, uint64_t* sigs, uint64_t* planes);
hipError_t am_refine_power_graph(hipStream_t s, const AMGraphDev& g, uint32_t view_size, uint64_t* sigs, uint64_t* planes);
// filterCandidates<Data>: the D x ceil(Q/64) transposed bitset (rows = data nodes).
hipError_t filter_data(hipStream_t s, const uint8_t* qlabels, uint64_t Q, const uint8_t* dlabels, uint64_t D, uint64_t* bits);

// filterCandidates<Query>: overwrites every word of the Q x W64 bitset.
hipError_t filter(hipStream_t s, const uint8_t* qlabels, uint64_t Q, const uint64_t* lplanes, uint64_t D, uint64_t* bits,
                  int num_cus);
// refineCandidates<Query> against bit-sliced data planes.
hipError_t refine(hipStream_t s, const uint64_t* qsig, uint64_t Q, const uint64_t* planes, uint64_t D, const LabelSlots& slots,
                  uint64_t* bits, int num_cus);
// filter followed by one refine against the given level's signatures, each word written once.
hipError_t filter_refine(hipStream_t s, const uint8_t* qlabels, const uint64_t* qsig, uint64_t Q, const uint64_t* lplanes,
                         const uint64_t* planes, uint64_t D, const LabelSlots& slots, uint64_t* bits, int num_cus);
// Per-row popcounts.
hipError_t row_counts(hipStream_t s, const uint64_t* bits, uint64_t Q, uint64_t W64, uint32_t* counts);
// Copy into the reference layout (uint32 words, stride ceil(D/32)).
hipError_t to_reference_layout(hipStream_t s, const uint64_t* bits, uint64_t Q, uint64_t D, uint32_t* out);

// GMCR.  pair_bits[qg * ceil(Gd/64) + strip]; counts[Gd]; offsets[Gd+1]; indices[P].
hipError_t gmcr_mark(hipStream_t s, const GraphDev& q, const GraphDev& d, const uint64_t* bits, uint64_t W64,
                     uint64_t* pair_bits);
hipError_t gmcr_count(hipStream_t s, const uint64_t* pair_bits, uint32_t Gq, uint32_t Gd, uint32_t* counts);
// Exclusive scan of counts[0..n) into offsets[0..n]; offsets[n] = total; block_tmp holds
// ceil(n/1024)+1 uint64 words.
hipError_t exclusive_scan(hipStream_t s, const uint32_t* counts, uint64_t n, uint32_t* offsets, uint64_t* block_tmp,
                          uint64_t* total_dev);
hipError_t gmcr_scatter(hipStream_t s, const uint64_t* pair_bits, uint32_t Gq, uint32_t Gd, const uint32_t* offsets,
                        uint32_t* indices, uint32_t* pair_graph, uint64_t cap);

// Slots of the join's statistics array (kJoinStatWords 64-bit words, zeroed before a join).
enum : uint32_t {
  kStatSum = 0,        // sum of per-pair counts
  kStatMatched = 1,    // pairs with count > 0
  kStatCursor = 2,     // work cursor of the lane join
  kStatOverflow = 3,   // 1 if a count exceeded int32
  kStatSteps = 4,      // DFS steps
  kStatWide = 5,       // pairs listed for the two-word join (molecules of > 64 nodes)
  kStatWideCursor = 6, // its work cursor
  kStatWaveIters = 7,  // loop iterations summed over wavefronts
  kStatDeep = 8,       // pairs listed for the full-depth pass of the lane join
  kStatDeepCursor = 9, // its work cursor
  kStatWideDeep = 10,  // pairs listed for the full-depth pass of the two-word join
  kStatWideDeepCursor = 11  // its work cursor
};
constexpr uint32_t kJoinStatWords = 16;

// Join.  stats[0] += sum of counts, stats[1] += pairs with count > 0, stats[2] = work cursor
// (zeroed by the caller), stats[3] |= 1 on int32 overflow.  With `deep` (capacity P) the join
// runs in two passes: a first that holds kJoinShallow query levels per lane in LDS and lists
// the pairs of deeper queries in `deep` (count in stats[8]), then a full-depth pass over that
// list; without it, one full-depth pass.
// Lane-per-pair join of the pairs whose molecule has <= 64 nodes; the others are appended to
// `wide` (count in stats[5]) for join_lanes_wide.  With P_dev the kernel reads the pair count
// from the device and P only bounds the grid.
hipError_t join_lanes(hipStream_t s, const GraphDev& q, const GraphDev& d, const uint64_t* bits, uint64_t W64, const uint32_t* indices,
                      const uint32_t* pair_graph, uint64_t P, int find_first, int32_t* counts, unsigned long long* stats, uint32_t* wide,
                      int num_cus, const unsigned long long* P_dev = nullptr, uint32_t* deep = nullptr);
// joinWildcardCandidates: matching order by candidate count, edge labels compared, one DFS per
// candidate of the first node; one lane per pair (qel/del: edge labels of the CSR entries).
hipError_t join_wildcard(hipStream_t s, const GraphDev& q, const GraphDev& d, const uint8_t* qel, const uint8_t* del, const uint64_t* bits,
                         uint64_t W64, const uint32_t* indices, const uint32_t* pair_graph, uint64_t P, int find_first, int32_t* counts,
                         unsigned long long* stats, int num_cus);
// The lane-per-pair join over the `nw` listed pairs of molecules with 65..128 nodes
// (stats[kStatWideCursor] is the work counter).  With nw_dev the kernel reads the list's length on
// the device, bounded by nw.  With `deep` (capacity nw) it runs in two passes as join_lanes does.
hipError_t join_lanes_wide(hipStream_t s, const GraphDev& q, const GraphDev& d, const uint64_t* bits, uint64_t W64, const uint32_t* indices,
                           const uint32_t* pair_graph, const uint32_t* list, uint64_t nw, int find_first, int32_t* counts,
                           unsigned long long* stats, int num_cus, const unsigned long long* nw_dev = nullptr, uint32_t* deep = nullptr);

}  // namespace sigmo_k
