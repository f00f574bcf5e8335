// gfx950 kernels for SIGMo's hot path: signatures, candidate filter/refine, GMCR, join.
//
// Wave64 throughout.  The bitset kernels are streaming (one pass over 64-bit word strips,
// coalesced across lanes) and bound by HBM; the join (join_lanes.hpp) is a persistent
// lane-per-pair DFS: each lane runs one GMCR pair's search, its mapping and the adjacency rows of
// its mapped positions in LDS, and takes the next pair when it finishes.  Reference semantics
// are cited per kernel (paths relative to the SIGMo repository); the integer conventions for the
// reference's undefined behaviour are: shift counts masked to width-1, Bitset<uint64_t> positions taken mod 64,
// PowerGraph reachables starting at zero, per-row refine, ascending query ids inside GMCR
// segments, fresh join state per pair with mapping[] = local node 0.
#include <algorithm>

#include "kernels.hpp"
#include "device_common.hpp"
#include "join_lanes.hpp"
#include "join_wildcard.hpp"

namespace sigmo_k {
namespace {


// Largest g with go[g] <= v: the upper-bound search of graph.hpp:100-111.
__device__ __forceinline__ uint32_t graph_of(const uint32_t* go, uint32_t G, uint32_t v) {
  uint32_t lo = 0, hi = G;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (v >= go[mid]) lo = mid + 1;
    else hi = mid;
  }
  return lo - 1;
}

__global__ void k_fill_u64(uint64_t* p, uint64_t n, uint64_t v) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}
__global__ void k_sub_u32(uint32_t* p, uint64_t n, uint32_t v) {
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] -= v;
}

inline unsigned grid_for(uint64_t n, unsigned block, unsigned cap = 1u << 20) {
  uint64_t g = (n + block - 1) / block;
  if (g > cap) g = cap;
  if (g == 0) g = 1;
  return static_cast<unsigned>(g);
}

// ---- derived graph arrays ------------------------------------------------------------------

// node_graph, 128-bit local adjacency, 8 label bit-planes per 64-node strip and the
// per-graph 64-position label masks (lmask must be zero on entry).  A wavefront covers 64
// consecutive nodes: one binary search finds the graph of its first node, the next 64 graph
// starts are loaded one per lane, and each node's graph follows from six shuffle steps over
// them.  The first lane of each graph in the wavefront ORs that graph's label masks in with one
// atomic per label present.
__global__ __launch_bounds__(256) void k_derive_nodes(const uint32_t* __restrict__ go, uint32_t G, const uint32_t* __restrict__ ro,
                                                      const uint32_t* __restrict__ ci, const uint8_t* __restrict__ nl, uint64_t N,
                                                      uint32_t* __restrict__ node_graph, uint64_t* __restrict__ adj,
                                                      uint64_t* __restrict__ fold, uint64_t* __restrict__ lplanes, unsigned long long* __restrict__ lmask) {
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint64_t v0 = v - lane;
  if (v0 >= N) return;  // wave-uniform; no block barrier in this kernel
  const bool valid = v < N;
  const uint32_t g0 = graph_of(go, G, static_cast<uint32_t>(v0));
  // starts of graphs g0+1 .. g0+64 (go[G] = N closes the last graph; past it, none)
  const uint32_t gi = g0 + 1 + lane;
  const uint32_t st = gi <= G ? go[gi] : 0xFFFFFFFFu;
  // graph of node v: g0 + (number of those starts <= v), the starts being ascending
  uint32_t pos = 0;
#pragma unroll
  for (uint32_t step = 32; step; step >>= 1) {
    const uint32_t cand = static_cast<uint32_t>(__shfl(static_cast<int>(st), static_cast<int>(pos + step - 1)));
    if (valid && cand <= static_cast<uint32_t>(v)) pos += step;
  }
  const uint32_t first = go[g0];
  const uint32_t before = static_cast<uint32_t>(__shfl(static_cast<int>(st), static_cast<int>(pos ? pos - 1 : 0)));
  const uint32_t g = g0 + pos, prev = pos ? before : first;
  uint32_t lab = 0;
  if (valid) {
    node_graph[v] = g;
    uint64_t lo = 0, hi = 0;
    for (uint32_t i = ro[v]; i < ro[v + 1]; ++i) {
      const uint32_t w = ci[i] - prev;  // < 128, checked at upload
      if (w < 64) lo |= 1ull << w;
      else hi |= 1ull << (w & 63);
    }
    adj[2 * v] = lo;
    adj[2 * v + 1] = hi;
    fold[v] = lo | hi;
    lab = nl[v];
  }
  // label masks: lanes of one graph are consecutive; the first of them owns the graph's bits
  const uint64_t vm = __ballot(valid);
  const bool head = valid && (lane == 0 || static_cast<uint32_t>(v) == prev);
  const uint64_t heads = __ballot(head);
  const uint64_t after = lane < 63 ? heads & ~((2ull << lane) - 1) : 0ull;
  const uint32_t stop = after ? static_cast<uint32_t>(__builtin_ctzll(after)) : 64u;
  const uint64_t seg = (stop == 64 ? ~0ull : ((1ull << stop) - 1)) & ~((1ull << lane) - 1) & vm;
  const uint32_t local = static_cast<uint32_t>(v) - prev;
  uint32_t present = valid && lab < 16 ? 1u << lab : 0u;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) present |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(present), o));
  for (uint32_t ls = present; ls; ls &= ls - 1) {  // wave-uniform
    const uint32_t l = static_cast<uint32_t>(__builtin_ctz(ls));
    const uint64_t b = __ballot(valid && lab == l) & seg;
    if (head && b && local < 64) atomicOr(&lmask[static_cast<uint64_t>(g) * 16 + l], (b >> lane) << local);
  }
  uint64_t mine = 0;
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint64_t m = __ballot((lab >> b) & 1u);
    if (lane == static_cast<uint32_t>(b)) mine = m;
  }
  const uint64_t w = v >> 6;
  if (lane < 8) lplanes[w * 8 + lane] = mine;
}

// ---- signatures --------------------------------------------------------------------------

// SignatureDevice::incrementLabelCount(label, 1), signature.hpp:44-51: +1 below 15 only,
// labels >= 16 ignored.
__device__ __forceinline__ uint64_t sig_inc(uint64_t sig, uint32_t label) {
  if (label < 16) {
    const uint32_t sh = 4 * label;
    const uint32_t c = static_cast<uint32_t>(sig >> sh) & 15u;
    if (c < 15) sig = (sig & ~(15ull << sh)) | (static_cast<uint64_t>(c + 1) << sh);
  }
  return sig;
}

// Bit-slice the packed signatures of the block's 64-node strips: plane p (= 4*label+bit)
// of strip w holds bit p of the signatures of the wave that holds the strip (a 64 x 64 bit
// transpose).  Staged through LDS
// so each plane row is written in contiguous runs of (BLOCK/64) words.
template <int BLOCK>
__device__ __forceinline__ void write_planes(uint64_t sig, uint64_t W64, uint64_t* __restrict__ planes, uint64_t* tile) {
  constexpr int WPB = BLOCK / kWave;
  const int wave = threadIdx.x / kWave, lane = lane_id();
  tile[lane * WPB + wave] = transpose64(sig);  // lane p: bit j = bit p of lane j's signature
  __syncthreads();
  const uint64_t w0 = static_cast<uint64_t>(blockIdx.x) * WPB;
  const int p = threadIdx.x / WPB, j = threadIdx.x % WPB;
  if (w0 + j < W64) planes[static_cast<uint64_t>(p) * W64 + w0 + j] = tile[p * WPB + j];
}

constexpr int kSigBlock = 512;

// Union of the folded adjacency fold[prev + b] over the set bits b of f.  Four positions per
// round, their loads issued together (a missing one repeats the first).
__device__ __forceinline__ uint64_t expand_frontier(const uint64_t* __restrict__ fold, uint64_t prev, uint64_t f) {
  uint64_t next = 0;
  while (f) {
    uint32_t b[4];
    b[0] = __builtin_ctzll(f);
    f &= f - 1;
#pragma unroll
    for (int t = 1; t < 4; ++t) {
      b[t] = f ? static_cast<uint32_t>(__builtin_ctzll(f)) : b[0];
      f &= f - 1;
    }
    uint64_t w[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) w[t] = fold[prev + b[t]];
#pragma unroll
    for (int t = 0; t < 4; ++t) next |= w[t];
  }
  return next;
}

// Saturated 16-label histogram of a position set, labels read through the graph's 64-position
// label masks (the +1-below-15 rule of signature.hpp:44-51 applied count times).
template <class P>
__device__ __forceinline__ uint64_t label_histogram(uint64_t r, P lm) {
  uint64_t s = 0;
#pragma unroll
  for (int l = 0; l < 16; ++l) {
    uint32_t c = __popcll(r & lm[l]);
    c = c < 15 ? c : 15;
    s |= static_cast<uint64_t>(c) << (4 * l);
  }
  return s;
}

// A signature block's nodes belong to a run of consecutive graphs.  When the run has at most
// kStagedGraphs graphs and kStagedRows nodes, its label masks and the folded adjacency
// (fold = adj_lo | adj_hi) of its nodes are staged in LDS: the BFS expansions and the histograms then
// read LDS instead of global memory.  Block-uniform result; g0 = the run's first graph, a0 its
// first node.
constexpr uint32_t kStagedGraphs = 64;
constexpr uint32_t kStagedRows = 1024;
__device__ __forceinline__ bool stage_block(const GraphDev& g, uint64_t* lms, uint64_t* fold, uint32_t& g0, uint64_t& a0) {
  const uint64_t v0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x;
  if (v0 >= g.N) return false;
  const uint64_t vl = (v0 + blockDim.x < g.N ? v0 + blockDim.x : g.N) - 1;
  g0 = g.node_graph[v0];
  const uint32_t g1 = g.node_graph[vl];
  a0 = g.go[g0];
  const uint64_t na = static_cast<uint64_t>(g.go[g1 + 1]) - a0;
  const bool staged = g1 - g0 + 1 <= kStagedGraphs && na <= kStagedRows;
  if (staged) {
    for (uint32_t t = threadIdx.x; t < (g1 - g0 + 1) * 16; t += blockDim.x) lms[t] = g.lmask[static_cast<uint64_t>(g0) * 16 + t];
    for (uint32_t t = threadIdx.x; t < na; t += blockDim.x) fold[t] = g.fold[a0 + t];
  }
  __syncthreads();
  return staged;
}

// expand_frontier over the staged folded adjacency; `at` = the graph's first node - a0.
__device__ __forceinline__ uint64_t expand_frontier_lds(const uint64_t* fold, uint64_t at, uint64_t f) {
  uint64_t next = 0;
  for (; f; f &= f - 1) next |= fold[at + __builtin_ctzll(f)];
  return next;
}

// generateCSRSignatures, signature.hpp:125-152: one thread per node, neighbour labels
// incremented onto the existing value (onto zero with `fresh`: the constructor's fill,
// signature.hpp:166-167, without a separate pass).
__global__ __launch_bounds__(kSigBlock) void k_sig_generate(const uint32_t* __restrict__ ro, const uint32_t* __restrict__ ci,
                                                            const uint8_t* __restrict__ nl, uint64_t N, uint64_t* __restrict__ sigs,
                                                            uint64_t* __restrict__ planes, int fresh) {
  __shared__ uint64_t tile[64 * (kSigBlock / kWave)];
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  uint64_t s = 0;
  if (v < N) {
    s = fresh ? 0ull : sigs[v];
    for (uint32_t i = ro[v]; i < ro[v + 1]; ++i) s = sig_inc(s, nl[ci[i]]);
    if (sigs) sigs[v] = s;  // null: the planes are the only output
  }
  if (planes) write_planes<kSigBlock>(s, (N + 63) / 64, planes, tile);
}

// incrementLabelCount(label, add) with the reference's uint8 arithmetic (signature.hpp:44-51,
// :32-37): below 15 only, and the new count (cur + add) mod 256 is stored only if < 16, so an
// add that wrapped below zero acts as a decrement.
__device__ __forceinline__ uint64_t sig_add(uint64_t sig, uint32_t label, uint32_t add) {
  const uint32_t sh = 4 * label;
  const uint32_t c = static_cast<uint32_t>(sig >> sh) & 15u;
  const uint32_t nv = (c + add) & 255u;
  if (c < 15 && nv < 16) sig = (sig & ~(15ull << sh)) | (static_cast<uint64_t>(nv) << sh);
  return sig;
}

// refineCSRSignatures<ViewBased>(view_size), signature.hpp:298-334: for each CSR entry to a
// neighbour u and each label l < 16, c = (uint8)(tmp[u][l] - (l == label(v) ? view_size : 0))
// is added when nonzero; `tmp` is the signatures before this pass.
__global__ __launch_bounds__(kSigBlock) void k_sig_refine_view_based(const uint32_t* __restrict__ ro, const uint32_t* __restrict__ ci,
                                                                     const uint8_t* __restrict__ nl, uint64_t N, uint32_t view_size,
                                                                     const uint64_t* __restrict__ tmp, uint64_t* __restrict__ sigs,
                                                                     uint64_t* __restrict__ planes) {
  __shared__ uint64_t tile[64 * (kSigBlock / kWave)];
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  uint64_t s = 0;
  if (v < N) {
    s = sigs[v];
    const uint32_t own = nl[v];
    for (uint32_t i = ro[v]; i < ro[v + 1]; ++i) {
      const uint64_t t = tmp[ci[i]];
      for (uint32_t l = 0; l < 16; ++l) {
        uint32_t c = static_cast<uint32_t>(t >> (4 * l)) & 15u;
        if (l == own) c = (c - view_size) & 255u;
        if (c) s = sig_add(s, l, c);
      }
    }
    sigs[v] = s;
  }
  if (planes) write_planes<kSigBlock>(s, (N + 63) / 64, planes, tile);
}

// ---- adjacency-matrix (AM) graphs ------------------------------------------------------------
// DeviceBatchedAMGraph, graph.hpp:114-148: per graph an n x n bit matrix, bit u*W + v set for
// an edge u-v, starting at word go[g]; nn[] holds the cumulative node counts.

// Row width of an n-node graph's matrix: words = (uint8)((n*n + 63) / 64) (utils.hpp:17-21),
// W = (uint16)sqrt(float(64 * words)) (utils.hpp:25-28).  For 64 * words <= 16320 the float
// square root is within 2^-16 of the real one while sqrt(k*k - 1) sits at least 1/256 below k,
// so its floor is the integer square root computed here.
__device__ __forceinline__ uint32_t am_width(uint32_t n) {
  const uint32_t x = 64u * (((n * n + 63u) / 64u) & 255u);
  uint32_t r = static_cast<uint32_t>(__builtin_sqrtf(static_cast<float>(x)));
  while (r * r > x) --r;
  while ((r + 1) * (r + 1) <= x) ++r;
  return r;
}

// binaryAMSearch, utils.hpp:100-112: the graph holding node v (first g with nn[g] > v).
__device__ __forceinline__ uint32_t am_graph_of(const uint32_t* __restrict__ nn, uint32_t G, uint32_t v) {
  uint32_t lo = 0, hi = G - 1;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) / 2;
    if (nn[mid] <= v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Row u (u < n) of a matrix of row width W <= 64: bits [u*W, u*W + W) as a mask.  The second
// word is read only when the row crosses into it, which keeps the read inside the graph's
// ceil(n*n/64) words (u*W + W <= n*W <= W*W <= 64 * words).
__device__ __forceinline__ uint64_t am_row(const uint64_t* __restrict__ m, uint32_t u, uint32_t W) {
  const uint32_t b = u * W, w = b / 64, sh = b % 64;
  uint64_t r = m[w] >> sh;
  if (sh && sh + W > 64) r |= m[w + 1] << (64 - sh);
  return W >= 64 ? r : r & ((1ull << W) - 1);
}

struct AMNode {
  uint32_t prev, n, W;
  const uint64_t* m;
};

__device__ __forceinline__ AMNode am_node(const AMGraphDev& g, uint32_t v) {
  const uint32_t gi = am_graph_of(g.nn, g.G, v);
  AMNode a;
  a.prev = gi ? g.nn[gi - 1] : 0;
  a.n = g.nn[gi] - a.prev;
  a.W = am_width(a.n);
  a.m = g.adj + g.go[gi];
  return a;
}

// The first four set bits of a row: getNeighbors (utils.hpp:34-48) lists the row's neighbours in
// ascending order into a MAX_NEIGHBORS = 4 array (types.hpp:28) whose readers stop at four.
__device__ __forceinline__ uint64_t am_first4(uint64_t row) {
  uint64_t keep = 0;
  for (int k = 0; k < 4 && row; ++k) {
    keep |= row & (0 - row);
    row &= row - 1;
  }
  return keep;
}

// generateAMSignatures, signature.hpp:100-120: thread per node, the labels of its first four
// neighbours incremented onto the existing value.
__global__ __launch_bounds__(kSigBlock) void k_am_generate(AMGraphDev g, uint64_t* __restrict__ sigs, uint64_t* __restrict__ planes) {
  __shared__ uint64_t tile[64 * (kSigBlock / kWave)];
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  uint64_t s = 0;
  if (v < g.N) {
    const AMNode a = am_node(g, static_cast<uint32_t>(v));
    s = sigs[v];
    for (uint64_t r = am_first4(am_row(a.m, static_cast<uint32_t>(v) - a.prev, a.W)); r; r &= r - 1)
      s = sig_inc(s, g.nl[a.prev + __builtin_ctzll(r)]);
    sigs[v] = s;
  }
  if (planes) write_planes<kSigBlock>(s, (g.N + 63) / 64, planes, tile);
}

// refineAMSignatures<ViewBased>(view_size), signature.hpp:214-248: the CSR ViewBased rule
// (k_sig_refine_view_based) over the node's first four AM neighbours; `tmp` holds the
// signatures before the pass.
__global__ __launch_bounds__(kSigBlock) void k_am_refine_view_based(AMGraphDev g, uint32_t view_size, const uint64_t* __restrict__ tmp,
                                                                    uint64_t* __restrict__ sigs, uint64_t* __restrict__ planes) {
  __shared__ uint64_t tile[64 * (kSigBlock / kWave)];
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  uint64_t s = 0;
  if (v < g.N) {
    const AMNode a = am_node(g, static_cast<uint32_t>(v));
    s = sigs[v];
    const uint32_t own = g.nl[v];
    for (uint64_t r = am_first4(am_row(a.m, static_cast<uint32_t>(v) - a.prev, a.W)); r; r &= r - 1) {
      const uint64_t t = tmp[a.prev + __builtin_ctzll(r)];
      for (uint32_t l = 0; l < 16; ++l) {
        uint32_t c = static_cast<uint32_t>(t >> (4 * l)) & 15u;
        if (l == own) c = (c - view_size) & 255u;
        if (c) s = sig_add(s, l, c);
      }
    }
    sigs[v] = s;
  }
  if (planes) write_planes<kSigBlock>(s, (g.N + 63) / 64, planes, tile);
}

// refineAMSignatures<PowerGraph>(view_size), signature.hpp:250-296: a BFS of view_size layers
// over 32-bit position sets (Bitset<uint32_t>: a position p is bit (uint8)p & 31, so positions
// >= 32 alias), expanding each set position's first four neighbours; the signature is then
// cleared and rebuilt as the label histogram of the reached positions but the node's own.
// Every set position is below n (for n > 32 because it is below 32), so its row lies inside
// the matrix.
__global__ __launch_bounds__(kSigBlock) void k_am_refine_power_graph(AMGraphDev g, uint32_t view_size, uint64_t* __restrict__ sigs,
                                                                     uint64_t* __restrict__ planes) {
  __shared__ uint64_t tile[64 * (kSigBlock / kWave)];
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  uint64_t s = 0;
  if (v < g.N) {
    const AMNode a = am_node(g, static_cast<uint32_t>(v));
    const uint32_t self = 1u << ((static_cast<uint32_t>(v) - a.prev) & 31u);
    uint32_t frontier = self, reach = self;
    for (uint32_t it = 0; it < view_size && frontier; ++it) {
      uint32_t next = 0;
      for (uint32_t f = frontier; f; f &= f - 1) {
        const uint32_t u = static_cast<uint32_t>(__builtin_ctz(f));
        for (uint64_t r = am_first4(am_row(a.m, u, a.W)); r; r &= r - 1) {
          const uint32_t b = 1u << (static_cast<uint32_t>(__builtin_ctzll(r)) & 31u);
          next |= b & ~reach;
          reach |= b;
        }
      }
      frontier = next;
    }
    reach &= ~self;
    for (; reach; reach &= reach - 1) s = sig_inc(s, g.nl[a.prev + __builtin_ctz(reach)]);
    sigs[v] = s;
  }
  if (planes) write_planes<kSigBlock>(s, (g.N + 63) / 64, planes, tile);
}

// filterCandidates<Data>, isomorphism.hpp:49-52: row = data node, word w of the row = query
// nodes 64w..64w+63 whose label equals the data node's or is the wildcard.
__global__ __launch_bounds__(256) void k_filter_data(const uint8_t* __restrict__ qlabels, uint64_t Q, const uint8_t* __restrict__ dlabels,
                                                     uint64_t D, uint64_t* __restrict__ bits) {
  const uint64_t WQ = (Q + 63) / 64;
  const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (i >= D * WQ) return;
  const uint64_t dn = i / WQ, w = i % WQ;
  const uint8_t dl = dlabels[dn];
  uint64_t word = 0;
  for (uint32_t b = 0; b < 64 && w * 64 + b < Q; ++b) {
    const uint8_t ql = qlabels[w * 64 + b];
    if (ql == dl || ql == 255u) word |= 1ull << b;
  }
  bits[i] = word;
}

// refineCSRSignatures<PowerGraph>(view_size), signature.hpp:336-390.  The reference's
// Bitset<uint64_t> holds local ids mod 64 and expands bit b as node go[g]+b, so the BFS
// runs over the folded adjacency adj_lo|adj_hi of node go[g]+b.  The level signature is the
// saturated label histogram of ball(view_size) minus the start bit minus the old
// reachables; the old reachables then absorb it.
// With `frontier` non-null the last BFS layer (the sphere at view_size) is stored for
// k_sig_refine_next.
__global__ __launch_bounds__(kSigBlock) void k_sig_refine(GraphDev g, uint32_t view_size, uint64_t* __restrict__ sigs,
                                                          uint64_t* __restrict__ reach, uint64_t* __restrict__ frontier_out,
                                                          uint64_t* __restrict__ planes, int fresh_reach) {
  __shared__ uint64_t tile[64 * (kSigBlock / kWave)];
  __shared__ uint64_t lms[kStagedGraphs * 16];
  __shared__ uint64_t fold[kStagedRows];
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  // the node's own loads go out before the block's staging barrier
  const bool live = v < g.N;
  const uint32_t gid = live ? g.node_graph[v] : 0u;
  const uint32_t prev = live ? g.go[gid] : 0u;
  const uint64_t old = live && !fresh_reach ? reach[v] : 0ull;  // fresh: the reachables start empty
  uint32_t g0 = 0;
  uint64_t a0 = 0;
  const bool staged = stage_block(g, lms, fold, g0, a0);
  uint64_t s = 0;
  if (live) {
    const uint64_t self = 1ull << ((static_cast<uint32_t>(v) - prev) & 63u);
    uint64_t ball = self, frontier = self;
    for (uint32_t it = 0; it < view_size && frontier; ++it) {
      uint64_t next = staged ? expand_frontier_lds(fold, prev - a0, frontier) : expand_frontier(g.fold, prev, frontier);
      next &= ~ball;
      ball |= next;
      frontier = next;
    }
    const uint64_t r = ball & ~self & ~old;
    s = staged ? label_histogram(r, lms + (gid - g0) * 16) : label_histogram(r, g.lmask + static_cast<uint64_t>(gid) * 16);
    if (sigs) sigs[v] = s;  // null: the planes are the only output
    reach[v] = r | old;
    if (frontier_out) frontier_out[v] = frontier;
  }
  if (planes) write_planes<kSigBlock>(s, (g.N + 63) / 64, planes, tile);
}

// One more level after a call at view_size M, when reach = ball(M) minus the start position
// and frontier = the sphere at M (the state every call sequence 1, 2, ..., M leaves): the
// reference's ball(M+1) minus old reachables is exactly the next BFS layer.
__global__ __launch_bounds__(kSigBlock) void k_sig_refine_next(GraphDev g, uint64_t* __restrict__ sigs, uint64_t* __restrict__ reach,
                                                               uint64_t* __restrict__ frontier, uint64_t* __restrict__ planes) {
  __shared__ uint64_t tile[64 * (kSigBlock / kWave)];
  __shared__ uint64_t lms[kStagedGraphs * 16];
  __shared__ uint64_t fold[kStagedRows];
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  // the node's own loads go out before the block's staging barrier
  const bool live = v < g.N;
  const uint32_t gid = live ? g.node_graph[v] : 0u;
  const uint32_t prev = live ? g.go[gid] : 0u;
  const uint64_t old = live ? reach[v] : 0ull;
  const uint64_t f = live ? frontier[v] : 0ull;
  uint32_t g0 = 0;
  uint64_t a0 = 0;
  const bool staged = stage_block(g, lms, fold, g0, a0);
  uint64_t s = 0;
  if (live) {
    const uint64_t self = 1ull << ((static_cast<uint32_t>(v) - prev) & 63u);
    uint64_t next = staged ? expand_frontier_lds(fold, prev - a0, f) : expand_frontier(g.fold, prev, f);
    next &= ~(old | self);
    s = staged ? label_histogram(next, lms + (gid - g0) * 16) : label_histogram(next, g.lmask + static_cast<uint64_t>(gid) * 16);
    if (sigs) sigs[v] = s;  // null: the planes are the only output
    reach[v] = old | next;
    frontier[v] = next;
  }
  if (planes) write_planes<kSigBlock>(s, (g.N + 63) / 64, planes, tile);
}

// ---- candidate filter / refine -----------------------------------------------------------

constexpr int kBitsBlock = 256;

// Rows per block for the word-strip kernels.  Each thread loads its strip's planes once per
// block, so a block should cover at least `min_rows` rows (planes then cost <= ~10% of the
// row traffic); beyond that, enough row chunks to give every CU ~4 blocks of 4 waves.
// Blocks of a (column strip group, row chunk) grid are placed so that the row chunks of one strip
// group run on one XCD: the hardware hands blocks to its 8 XCDs round-robin (block b to XCD
// b mod 8), so XCD x walks the strip groups x, x+8, ... each with its row chunks in a row, and a
// strip's planes come from HBM into that XCD's L2 once instead of once per chunk.
constexpr uint64_t kXcds = 8;
__device__ __forceinline__ bool xcd_tile(uint64_t b, uint64_t cols, uint64_t chunks, uint64_t& col, uint64_t& chunk) {
  const uint64_t x = b % kXcds, t = b / kXcds;
  col = x + kXcds * (t / chunks);
  chunk = t % chunks;
  return col < cols;
}
inline uint64_t xcd_grid(uint64_t cols, uint64_t chunks) { return kXcds * ((cols + kXcds - 1) / kXcds) * chunks; }

inline unsigned rows_per_block(uint64_t Q, uint64_t col_blocks, int num_cus, uint64_t min_rows, uint64_t blocks_per_cu = 4) {
  const uint64_t target_blocks = static_cast<uint64_t>(num_cus > 0 ? num_cus : 256) * blocks_per_cu;
  uint64_t chunks = (target_blocks + col_blocks - 1) / col_blocks;
  if (chunks < 1) chunks = 1;
  uint64_t rpb = (Q + chunks - 1) / chunks;
  if (rpb < min_rows) rpb = min_rows;
  if (rpb > Q) rpb = Q;
  return static_cast<unsigned>(rpb > 0 ? rpb : 1);
}

// filterCandidates<Query>, isomorphism.hpp:22-62: bit (q,d) <=> label(q) == label(d) or
// label(q) == 255.  Thread = one 64-bit word of a row; the strip's 8 label bit-planes stay in
// registers across the block's rows; every word is written once (no atomics).
__global__ __launch_bounds__(kBitsBlock) void k_filter(const uint8_t* __restrict__ qlabels, uint64_t Q, const uint64_t* __restrict__ lplanes,
                                                       uint64_t D, uint64_t* __restrict__ bits, unsigned rpb) {
  const uint64_t W64 = (D + 63) / 64;
  const uint64_t w = blockIdx.x * static_cast<uint64_t>(kBitsBlock) + threadIdx.x;
  if (w >= W64) return;
  uint64_t P[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) P[b] = lplanes[w * 8 + b];
  const uint64_t valid = (w == W64 - 1 && (D & 63)) ? (~0ull >> (64 - (D & 63))) : ~0ull;
  const uint64_t q0 = static_cast<uint64_t>(blockIdx.y) * rpb;
  const uint64_t q1 = q0 + rpb < Q ? q0 + rpb : Q;
  for (uint64_t q = q0; q < q1; ++q) {
    const uint32_t ql = qlabels[q];
    bits[q * W64 + w] = ql != 255u ? label_match(P, ql, valid) : valid;
  }
}

// refineCandidates<Query>, isomorphism.hpp:64-127: a set bit (q,d) survives iff for every
// label l < 16, count_q(l) <= count_d(l) at the current level.  Only labels that occur as
// query node labels can have nonzero query counts; their data planes (4 per label) sit in
// registers, and each query row costs one bit-sliced >= per nonzero query count.
template <int U>
__global__ __launch_bounds__(kBitsBlock) void k_refine(const uint64_t* __restrict__ qsig, uint64_t Q, const uint64_t* __restrict__ planes,
                                                       uint64_t D, LabelSlots slots, uint64_t* __restrict__ bits, unsigned rpb,
                                                       unsigned chunks) {
  const uint64_t W64 = (D + 63) / 64;
  uint64_t col, chunk;
  if (!xcd_tile(blockIdx.x, (W64 + kBitsBlock - 1) / kBitsBlock, chunks, col, chunk)) return;
  const uint64_t w = col * kBitsBlock + threadIdx.x;
  if (w >= W64) return;
  uint64_t P[U][4];
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int b = 0; b < 4; ++b) P[u][b] = planes[(static_cast<uint64_t>(slots.label[u]) * 4 + b) * W64 + w];
  const uint64_t q0 = chunk * rpb;
  const uint64_t q1 = q0 + rpb < Q ? q0 + rpb : Q;
  auto keep_of = [&](uint64_t qs) {
    uint64_t keep = ~0ull;
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const uint32_t c = static_cast<uint32_t>(qs >> (4 * slots.label[u])) & 15u;
      if (c) keep &= ge4(P[u], c);
    }
    return keep;
  };
  constexpr int kRows = 4;  // loads in flight per thread
  uint64_t q = q0;
  for (; q + kRows <= q1; q += kRows) {
    uint64_t old[kRows];
#pragma unroll
    for (int i = 0; i < kRows; ++i) old[i] = bits[(q + i) * W64 + w];
#pragma unroll
    for (int i = 0; i < kRows; ++i) {
      const uint64_t nw = old[i] & keep_of(qsig[q + i]);
      if (nw != old[i]) bits[(q + i) * W64 + w] = nw;
    }
  }
  for (; q < q1; ++q) {
    const uint64_t old = bits[q * W64 + w];
    const uint64_t nw = old & keep_of(qsig[q]);
    if (nw != old) bits[q * W64 + w] = nw;
  }
}

// filterCandidates<Query> (isomorphism.hpp:22-62) and the first refineCandidates<Query> pass
// (:64-127) in one sweep: word = label match of row q AND the signature test at the level that
// pass reads.  The filter reads no signatures and the signature refinement reads no candidates,
// so computing that level's signatures first and then running this gives the same bits as
// filter, signatures, refine -- with each word written once and never re-read.
template <int U>
__global__ __launch_bounds__(kBitsBlock) void k_filter_refine(const uint8_t* __restrict__ qlabels, const uint64_t* __restrict__ qsig,
                                                              uint64_t Q, const uint64_t* __restrict__ lplanes,
                                                              const uint64_t* __restrict__ planes, uint64_t D, LabelSlots slots,
                                                              uint64_t* __restrict__ bits, unsigned rpb, unsigned chunks) {
  const uint64_t W64 = (D + 63) / 64;
  uint64_t col, chunk;
  if (!xcd_tile(blockIdx.x, (W64 + kBitsBlock - 1) / kBitsBlock, chunks, col, chunk)) return;
  const uint64_t w = col * kBitsBlock + threadIdx.x;
  if (w >= W64) return;
  uint64_t L[8];
#pragma unroll
  for (int b = 0; b < 8; ++b) L[b] = lplanes[w * 8 + b];
  uint64_t P[U][4];
#pragma unroll
  for (int u = 0; u < U; ++u)
#pragma unroll
    for (int b = 0; b < 4; ++b) P[u][b] = planes[(static_cast<uint64_t>(slots.label[u]) * 4 + b) * W64 + w];
  const uint64_t valid = (w == W64 - 1 && (D & 63)) ? (~0ull >> (64 - (D & 63))) : ~0ull;
  const uint64_t q0 = chunk * rpb;
  const uint64_t q1 = q0 + rpb < Q ? q0 + rpb : Q;
  constexpr int kRows = 4;  // the rows' label and signature loads go out together
  for (uint64_t q = q0; q < q1; q += kRows) {
    uint32_t ql[kRows];
    uint64_t qs[kRows];
#pragma unroll
    for (int i = 0; i < kRows; ++i) {
      const uint64_t qi = q + i < q1 ? q + i : q1 - 1;
      ql[i] = qlabels[qi];
      qs[i] = qsig[qi];
    }
#pragma unroll
    for (int i = 0; i < kRows; ++i) {
      uint64_t m = ql[i] != 255u ? label_match(L, ql[i], valid) : valid;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t c = static_cast<uint32_t>(qs[i] >> (4 * slots.label[u])) & 15u;
        if (c) m &= ge4(P[u], c);
      }
      if (q + i < q1) bits[(q + i) * W64 + w] = m;
    }
  }
}

__global__ __launch_bounds__(256) void k_row_counts(const uint64_t* __restrict__ bits, uint64_t W64, uint32_t* __restrict__ counts) {
  __shared__ uint32_t part[256 / kWave];
  const uint64_t q = blockIdx.x;
  uint32_t c = 0;
  for (uint64_t w = threadIdx.x; w < W64; w += blockDim.x) c += __popcll(bits[q * W64 + w]);
  for (int off = kWave / 2; off > 0; off >>= 1) c += __shfl_xor(c, off);
  if (lane_id() == 0) part[threadIdx.x / kWave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int i = 0; i < 256 / kWave; ++i) t += part[i];
    counts[q] = t;
  }
}

// ---- GMCR ----------------------------------------------------------------------------------

// A molecule's node range [s, e) as bitset words ws..we, the end words masked by ms / me.  A
// molecule has 1..128 nodes, so the range spans at most three words.
struct RowRange {
  uint32_t ws, we;
  uint64_t ms, me;
};
__device__ __forceinline__ RowRange row_range(uint32_t s, uint32_t e) {
  RowRange r;
  r.ws = s >> 6;
  r.we = (e - 1) >> 6;
  r.ms = ~0ull << (s & 63);
  r.me = ~0ull >> (63 - ((e - 1) & 63));
  if (r.ws == r.we) r.ms = r.me = r.ms & r.me;
  return r;
}
// The bits of a bitset row inside the range (nonzero <=> the row has a candidate there).
__device__ __forceinline__ uint64_t row_bits(const uint64_t* __restrict__ row, const RowRange& r) {
  uint64_t x = (row[r.ws] & r.ms) | (row[r.we] & r.me);
  if (r.we > r.ws + 1) x |= row[r.ws + 1];
  return x;
}

// gmcr.hpp:47-68 / :94-115 predicate for (query graph = blockIdx.y, data graph = lane):
// every query node has a candidate inside the molecule.  One ballot word per 64 molecules.
// The query nodes' rows are read four at a time (past the last node the last row repeats), so
// a wave keeps several loads in flight; it stops once none of its molecules is alive.
__global__ __launch_bounds__(256) void k_gmcr_mark(GraphDev q, GraphDev d, const uint64_t* __restrict__ bits, uint64_t W64,
                                                   uint64_t* __restrict__ pair_bits) {
  const uint32_t qg = blockIdx.y;
  const uint64_t dg = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  const uint64_t strips = (static_cast<uint64_t>(d.G) + 63) / 64;
  bool alive = dg < d.G;
  RowRange rr{};
  if (alive) rr = row_range(d.go[dg], d.go[dg + 1]);
  const uint32_t qoff = q.go[qg], k = q.go[qg + 1] - qoff;  // k >= 1
  constexpr uint32_t kRows = 4;
  for (uint32_t i = 0; i < k; i += kRows) {
    if (__ballot(alive) == 0) break;
    if (alive) {
      uint64_t x[kRows];
#pragma unroll
      for (uint32_t t = 0; t < kRows; ++t) {
        const uint32_t qi = qoff + (i + t < k ? i + t : k - 1);
        x[t] = row_bits(bits + static_cast<uint64_t>(qi) * W64, rr);
      }
#pragma unroll
      for (uint32_t t = 0; t < kRows; ++t) alive = alive && x[t] != 0;
    }
  }
  const uint64_t res = __ballot(alive);
  if (lane_id() == 0 && (dg >> 6) < strips) pair_bits[static_cast<uint64_t>(qg) * strips + (dg >> 6)] = res;
}

__global__ __launch_bounds__(256) void k_gmcr_count(const uint64_t* __restrict__ pair_bits, uint32_t Gq, uint32_t Gd,
                                                    uint32_t* __restrict__ counts) {
  const uint64_t dg = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (dg >= Gd) return;
  const uint64_t strips = (static_cast<uint64_t>(Gd) + 63) / 64;
  const uint64_t strip = dg >> 6;
  const uint32_t bit = dg & 63;
  uint32_t c = 0, qg = 0;
  constexpr uint32_t kBatch = 8;  // loads in flight per thread
  for (; qg + kBatch <= Gq; qg += kBatch) {
    uint64_t w[kBatch];
#pragma unroll
    for (uint32_t i = 0; i < kBatch; ++i) w[i] = pair_bits[(qg + i) * strips + strip];
#pragma unroll
    for (uint32_t i = 0; i < kBatch; ++i) c += (w[i] >> bit) & 1ull;
  }
  for (; qg < Gq; ++qg) c += (pair_bits[qg * strips + strip] >> bit) & 1ull;
  counts[dg] = c;
}

// Kernel 3 of gmcr.hpp (:94-115) with a deterministic order: each molecule's segment lists its
// query graphs in ascending id.  Also records the molecule of each pair for the join.
__global__ __launch_bounds__(256) void k_gmcr_scatter(const uint64_t* __restrict__ pair_bits, uint32_t Gq, uint32_t Gd,
                                                      const uint32_t* __restrict__ offsets, uint32_t* __restrict__ indices,
                                                      uint32_t* __restrict__ pair_graph, uint64_t cap) {
  // writes stop at `cap` (the arrays' size): a graph replay sized by an earlier run cannot write
  // past them, and reports the changed pair count instead
  const uint64_t dg = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (dg >= Gd) return;
  const uint64_t strips = (static_cast<uint64_t>(Gd) + 63) / 64;
  const uint64_t strip = dg >> 6;
  const uint32_t bit = dg & 63;
  uint32_t pos = offsets[dg];
  auto emit = [&](uint32_t qg, uint64_t w) {
    if ((w >> bit) & 1ull) {
      if (pos < cap) {
        indices[pos] = qg;
        pair_graph[pos] = static_cast<uint32_t>(dg);
      }
      ++pos;
    }
  };
  uint32_t qg = 0;
  constexpr uint32_t kBatch = 8;  // loads in flight per thread
  for (; qg + kBatch <= Gq; qg += kBatch) {
    uint64_t w[kBatch];
#pragma unroll
    for (uint32_t i = 0; i < kBatch; ++i) w[i] = pair_bits[(qg + i) * strips + strip];
#pragma unroll
    for (uint32_t i = 0; i < kBatch; ++i) emit(qg + i, w[i]);
  }
  for (; qg < Gq; ++qg) emit(qg, pair_bits[qg * strips + strip]);
}

// ---- exclusive scan (u32 counts, u64 carries) -------------------------------------------------

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__device__ __forceinline__ uint64_t wave_inclusive(uint64_t x) {
  const int lane = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    const uint64_t y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  return x;
}

// Exclusive scan of one value per thread across the block; returns the block total.
__device__ __forceinline__ uint64_t block_exclusive(uint64_t x, uint64_t& excl, uint64_t* wave_tot) {
  const int wave = threadIdx.x / kWave, lane = lane_id();
  const uint64_t inc = wave_inclusive(x);
  if (lane == kWave - 1) wave_tot[wave] = inc;
  __syncthreads();
  uint64_t base = 0, total = 0;
  for (int i = 0; i < kScanBlock / kWave; ++i) {
    if (i < wave) base += wave_tot[i];
    total += wave_tot[i];
  }
  excl = base + inc - x;
  __syncthreads();
  return total;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* __restrict__ in, uint64_t n, uint64_t* __restrict__ block_sums) {
  __shared__ uint64_t wave_tot[kScanBlock / kWave];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  uint64_t x = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) x += in[base + i];
  uint64_t excl;
  const uint64_t total = block_exclusive(x, excl, wave_tot);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// Single block: exclusive scan of the block sums in place, total to total_dev.
__global__ __launch_bounds__(kScanBlock) void k_scan_top(uint64_t* __restrict__ block_sums, uint64_t nb, uint64_t* __restrict__ total_dev) {
  __shared__ uint64_t wave_tot[kScanBlock / kWave];
  uint64_t carry = 0;
  for (uint64_t base = 0; base < nb; base += kScanBlock) {
    const uint64_t i = base + threadIdx.x;
    const uint64_t x = i < nb ? block_sums[i] : 0;
    uint64_t excl;
    const uint64_t total = block_exclusive(x, excl, wave_tot);
    if (i < nb) block_sums[i] = carry + excl;
    carry += total;
  }
  if (threadIdx.x == 0) *total_dev = carry;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_final(const uint32_t* __restrict__ in, uint64_t n, const uint64_t* __restrict__ block_sums,
                                                           const uint64_t* __restrict__ total_dev, uint32_t* __restrict__ out) {
  __shared__ uint64_t wave_tot[kScanBlock / kWave];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint64_t x = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    v[i] = base + i < n ? in[base + i] : 0u;
    x += v[i];
  }
  uint64_t excl;
  block_exclusive(x, excl, wave_tot);
  uint64_t run = block_sums[blockIdx.x] + excl;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = static_cast<uint32_t>(run);
    run += v[i];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = static_cast<uint32_t>(*total_dev);
}

}  // namespace

// ---- launchers -------------------------------------------------------------------------------

hipError_t fill_u64(hipStream_t s, uint64_t* p, uint64_t n, uint64_t v) {
  if (!n) return hipSuccess;
  k_fill_u64<<<grid_for(n, 256, 16384), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}
hipError_t fill_u32(hipStream_t s, uint32_t* p, uint64_t n, uint32_t v) {
  if (!n) return hipSuccess;
  k_fill_u32<<<grid_for(n, 256, 16384), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}
hipError_t sub_u32(hipStream_t s, uint32_t* p, uint64_t n, uint32_t v) {
  if (!n || !v) return hipSuccess;
  k_sub_u32<<<grid_for(n, 256, 16384), 256, 0, s>>>(p, n, v);
  return hipGetLastError();
}

hipError_t derive_graph(hipStream_t s, const uint32_t* go, uint32_t G, const uint32_t* ro, const uint32_t* ci, const uint8_t* nl,
                        uint64_t N, uint32_t* node_graph, uint64_t* adj, uint64_t* fold, uint64_t* lplanes, uint64_t* lmask) {
  hipError_t err = fill_u64(s, lmask, static_cast<uint64_t>(G) * 16, 0);
  if (err != hipSuccess || !N) return err;
  const unsigned blocks = static_cast<unsigned>((N + 255) / 256);
  k_derive_nodes<<<blocks, 256, 0, s>>>(go, G, ro, ci, nl, N, node_graph, adj, fold, lplanes,
                                        reinterpret_cast<unsigned long long*>(lmask));
  return hipGetLastError();
}

hipError_t sig_generate(hipStream_t s, const GraphDev& g, uint64_t* sigs, uint64_t* planes, bool fresh) {
  if (!g.N) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((g.N + kSigBlock - 1) / kSigBlock);
  k_sig_generate<<<blocks, kSigBlock, 0, s>>>(g.ro, g.ci, g.nl, g.N, sigs, planes, fresh ? 1 : 0);
  return hipGetLastError();
}

hipError_t sig_refine_view_based(hipStream_t s, const GraphDev& g, uint32_t view_size, const uint64_t* tmp, uint64_t* sigs, uint64_t* planes) {
  if (!g.N) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((g.N + kSigBlock - 1) / kSigBlock);
  k_sig_refine_view_based<<<blocks, kSigBlock, 0, s>>>(g.ro, g.ci, g.nl, g.N, view_size, tmp, sigs, planes);
  return hipGetLastError();
}

hipError_t am_generate(hipStream_t s, const AMGraphDev& g, uint64_t* sigs, uint64_t* planes) {
  if (!g.N) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((g.N + kSigBlock - 1) / kSigBlock);
  k_am_generate<<<blocks, kSigBlock, 0, s>>>(g, sigs, planes);
  return hipGetLastError();
}

hipError_t am_refine_view_based(hipStream_t s, const AMGraphDev& g, uint32_t view_size, const uint64_t* tmp, uint64_t* sigs, uint64_t* planes) {
  if (!g.N) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((g.N + kSigBlock - 1) / kSigBlock);
  k_am_refine_view_based<<<blocks, kSigBlock, 0, s>>>(g, view_size, tmp, sigs, planes);
  return hipGetLastError();
}

hipError_t am_refine_power_graph(hipStream_t s, const AMGraphDev& g, uint32_t view_size, uint64_t* sigs, uint64_t* planes) {
  if (!g.N) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((g.N + kSigBlock - 1) / kSigBlock);
  k_am_refine_power_graph<<<blocks, kSigBlock, 0, s>>>(g, view_size, sigs, planes);
  return hipGetLastError();
}

hipError_t filter_data(hipStream_t s, const uint8_t* qlabels, uint64_t Q, const uint8_t* dlabels, uint64_t D, uint64_t* bits) {
  const uint64_t n = D * ((Q + 63) / 64);
  if (!n) return hipSuccess;
  k_filter_data<<<static_cast<unsigned>((n + 255) / 256), 256, 0, s>>>(qlabels, Q, dlabels, D, bits);
  return hipGetLastError();
}

hipError_t sig_refine(hipStream_t s, const GraphDev& g, uint32_t view_size, uint64_t* sigs, uint64_t* reach, uint64_t* frontier_out,
                      uint64_t* planes, bool fresh_reach) {
  if (!g.N) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((g.N + kSigBlock - 1) / kSigBlock);
  k_sig_refine<<<blocks, kSigBlock, 0, s>>>(g, view_size, sigs, reach, frontier_out, planes, fresh_reach ? 1 : 0);
  return hipGetLastError();
}

hipError_t sig_refine_next(hipStream_t s, const GraphDev& g, uint64_t* sigs, uint64_t* reach, uint64_t* frontier, uint64_t* planes) {
  if (!g.N) return hipSuccess;
  const unsigned blocks = static_cast<unsigned>((g.N + kSigBlock - 1) / kSigBlock);
  k_sig_refine_next<<<blocks, kSigBlock, 0, s>>>(g, sigs, reach, frontier, planes);
  return hipGetLastError();
}

hipError_t filter(hipStream_t s, const uint8_t* qlabels, uint64_t Q, const uint64_t* lplanes, uint64_t D, uint64_t* bits, int num_cus) {
  const uint64_t W64 = (D + 63) / 64;
  if (!Q || !W64) return hipSuccess;
  const uint64_t cols = (W64 + kBitsBlock - 1) / kBitsBlock;
  const unsigned rpb = rows_per_block(Q, cols, num_cus, 32, 16);  // store-only: favour parallelism
  const dim3 grid(static_cast<unsigned>(cols), static_cast<unsigned>((Q + rpb - 1) / rpb));
  k_filter<<<grid, kBitsBlock, 0, s>>>(qlabels, Q, lplanes, D, bits, rpb);
  return hipGetLastError();
}

template <int U>
static hipError_t launch_refine(hipStream_t s, dim3 grid, const uint64_t* qsig, uint64_t Q, const uint64_t* planes, uint64_t D,
                                const LabelSlots& slots, uint64_t* bits, unsigned rpb, unsigned chunks) {
  k_refine<U><<<grid, kBitsBlock, 0, s>>>(qsig, Q, planes, D, slots, bits, rpb, chunks);
  return hipGetLastError();
}

hipError_t refine(hipStream_t s, const uint64_t* qsig, uint64_t Q, const uint64_t* planes, uint64_t D, const LabelSlots& slots,
                  uint64_t* bits, int num_cus) {
  const uint64_t W64 = (D + 63) / 64;
  if (!Q || !W64 || slots.count == 0) return hipSuccess;
  const uint64_t cols = (W64 + kBitsBlock - 1) / kBitsBlock;
  // 32 B of planes per label and word against 16 B of row traffic per row.  When the column
  // strips alone give every CU 4 blocks (bitsets of tens of GB), split the rows in two as well:
  // 8 blocks per CU keep more loads in flight (1,000,000 molecules x 1,000 queries: 27.7 -> 24.2
  // ms per pass)
  const uint64_t cus = static_cast<uint64_t>(num_cus > 0 ? num_cus : 256);
  const unsigned rpb = rows_per_block(Q, cols, num_cus, 16 * static_cast<uint64_t>(slots.count), cols >= cus * 4 ? 8 : 4);
  // one block per (word strip, row chunk) in a 1-D grid placed by xcd_tile: a strip's chunks run
  // on one XCD, one after another, and share its planes through that XCD's L2
  const unsigned chunks = static_cast<unsigned>((Q + rpb - 1) / rpb);
  if (xcd_grid(cols, chunks) > 0x7fffffffull) return hipErrorInvalidValue;
  const dim3 grid(static_cast<unsigned>(xcd_grid(cols, chunks)));
  switch (slots.count) {
#define SIGMO_REFINE_CASE(u) \
  case u: return launch_refine<u>(s, grid, qsig, Q, planes, D, slots, bits, rpb, chunks);
    SIGMO_REFINE_CASE(1) SIGMO_REFINE_CASE(2) SIGMO_REFINE_CASE(3) SIGMO_REFINE_CASE(4)
    SIGMO_REFINE_CASE(5) SIGMO_REFINE_CASE(6) SIGMO_REFINE_CASE(7) SIGMO_REFINE_CASE(8)
    SIGMO_REFINE_CASE(9) SIGMO_REFINE_CASE(10) SIGMO_REFINE_CASE(11) SIGMO_REFINE_CASE(12)
    SIGMO_REFINE_CASE(13) SIGMO_REFINE_CASE(14) SIGMO_REFINE_CASE(15) SIGMO_REFINE_CASE(16)
#undef SIGMO_REFINE_CASE
    default: return hipErrorInvalidValue;
  }
}

template <int U>
static hipError_t launch_filter_refine(hipStream_t s, dim3 grid, const uint8_t* qlabels, const uint64_t* qsig, uint64_t Q,
                                       const uint64_t* lplanes, const uint64_t* planes, uint64_t D, const LabelSlots& slots,
                                       uint64_t* bits, unsigned rpb, unsigned chunks) {
  k_filter_refine<U><<<grid, kBitsBlock, 0, s>>>(qlabels, qsig, Q, lplanes, planes, D, slots, bits, rpb, chunks);
  return hipGetLastError();
}

hipError_t filter_refine(hipStream_t s, const uint8_t* qlabels, const uint64_t* qsig, uint64_t Q, const uint64_t* lplanes,
                         const uint64_t* planes, uint64_t D, const LabelSlots& slots, uint64_t* bits, int num_cus) {
  if (slots.count == 0) return filter(s, qlabels, Q, lplanes, D, bits, num_cus);  // the refine keeps every bit
  const uint64_t W64 = (D + 63) / 64;
  if (!Q || !W64) return hipSuccess;
  const uint64_t cols = (W64 + kBitsBlock - 1) / kBitsBlock;
  // store-only like the filter, with 64 + 32 U bytes of planes per word and block.  Measured on
  // 100,000 molecules x 100 queries (U = 8): 64 rows and 16 blocks per CU 128 us, 128 rows and
  // 8 blocks 140 us, 41 rows 135 us
  const unsigned rpb = rows_per_block(Q, cols, num_cus, std::max<uint64_t>(64, 8 * static_cast<uint64_t>(slots.count)), 16);
  const unsigned chunks = static_cast<unsigned>((Q + rpb - 1) / rpb);  // row chunk fastest, as in refine
  if (xcd_grid(cols, chunks) > 0x7fffffffull) return hipErrorInvalidValue;
  const dim3 grid(static_cast<unsigned>(xcd_grid(cols, chunks)));
  switch (slots.count) {
#define SIGMO_FILTER_REFINE_CASE(u) \
  case u: return launch_filter_refine<u>(s, grid, qlabels, qsig, Q, lplanes, planes, D, slots, bits, rpb, chunks);
    SIGMO_FILTER_REFINE_CASE(1) SIGMO_FILTER_REFINE_CASE(2) SIGMO_FILTER_REFINE_CASE(3) SIGMO_FILTER_REFINE_CASE(4)
    SIGMO_FILTER_REFINE_CASE(5) SIGMO_FILTER_REFINE_CASE(6) SIGMO_FILTER_REFINE_CASE(7) SIGMO_FILTER_REFINE_CASE(8)
    SIGMO_FILTER_REFINE_CASE(9) SIGMO_FILTER_REFINE_CASE(10) SIGMO_FILTER_REFINE_CASE(11) SIGMO_FILTER_REFINE_CASE(12)
    SIGMO_FILTER_REFINE_CASE(13) SIGMO_FILTER_REFINE_CASE(14) SIGMO_FILTER_REFINE_CASE(15) SIGMO_FILTER_REFINE_CASE(16)
#undef SIGMO_FILTER_REFINE_CASE
    default: return hipErrorInvalidValue;
  }
}

hipError_t row_counts(hipStream_t s, const uint64_t* bits, uint64_t Q, uint64_t W64, uint32_t* counts) {
  if (!Q) return hipSuccess;
  k_row_counts<<<static_cast<unsigned>(Q), 256, 0, s>>>(bits, W64, counts);
  return hipGetLastError();
}

hipError_t gmcr_mark(hipStream_t s, const GraphDev& q, const GraphDev& d, const uint64_t* bits, uint64_t W64, uint64_t* pair_bits) {
  if (!q.G || !d.G) return hipSuccess;
  const dim3 grid(static_cast<unsigned>((d.G + 255) / 256), q.G);
  k_gmcr_mark<<<grid, 256, 0, s>>>(q, d, bits, W64, pair_bits);
  return hipGetLastError();
}

hipError_t gmcr_count(hipStream_t s, const uint64_t* pair_bits, uint32_t Gq, uint32_t Gd, uint32_t* counts) {
  if (!Gd) return hipSuccess;
  k_gmcr_count<<<(Gd + 255) / 256, 256, 0, s>>>(pair_bits, Gq, Gd, counts);
  return hipGetLastError();
}

hipError_t exclusive_scan(hipStream_t s, const uint32_t* counts, uint64_t n, uint32_t* offsets, uint64_t* block_tmp, uint64_t* total_dev) {
  const uint64_t nb = (n + kScanTile - 1) / kScanTile;
  if (nb == 0) {
    k_fill_u32<<<1, 64, 0, s>>>(offsets, 1, 0);
    k_fill_u64<<<1, 64, 0, s>>>(total_dev, 1, 0);
    return hipGetLastError();
  }
  k_scan_reduce<<<static_cast<unsigned>(nb), kScanBlock, 0, s>>>(counts, n, block_tmp);
  k_scan_top<<<1, kScanBlock, 0, s>>>(block_tmp, nb, total_dev);
  k_scan_final<<<static_cast<unsigned>(nb), kScanBlock, 0, s>>>(counts, n, block_tmp, total_dev, offsets);
  return hipGetLastError();
}

hipError_t gmcr_scatter(hipStream_t s, const uint64_t* pair_bits, uint32_t Gq, uint32_t Gd, const uint32_t* offsets, uint32_t* indices,
                        uint32_t* pair_graph, uint64_t cap) {
  if (!Gd) return hipSuccess;
  k_gmcr_scatter<<<(Gd + 255) / 256, 256, 0, s>>>(pair_bits, Gq, Gd, offsets, indices, pair_graph, cap);
  return hipGetLastError();
}

hipError_t join_lanes(hipStream_t s, const GraphDev& q, const GraphDev& d, const uint64_t* bits, uint64_t W64, const uint32_t* indices,
                      const uint32_t* pair_graph, uint64_t P, int find_first, int32_t* counts, unsigned long long* stats, uint32_t* wide,
                      int num_cus, const unsigned long long* P_dev, uint32_t* deep) {
  if (!P) return hipSuccess;
  const BitsetMasks src{bits, W64};
  return launch_join_lanes(s, q, d, src, indices, pair_graph, P, find_first, counts, stats, wide, num_cus, P_dev, deep, 4, 0);
}

hipError_t join_wildcard(hipStream_t s, const GraphDev& q, const GraphDev& d, const uint8_t* qel, const uint8_t* del, const uint64_t* bits,
                         uint64_t W64, const uint32_t* indices, const uint32_t* pair_graph, uint64_t P, int find_first, int32_t* counts,
                         unsigned long long* stats, int num_cus) {
  if (!P) return hipSuccess;
  const uint64_t cus = static_cast<uint64_t>(num_cus > 0 ? num_cus : 256);
  const unsigned blocks = static_cast<unsigned>(std::max<uint64_t>(1, std::min<uint64_t>((P + 255) / 256, cus * 8)));
  if (find_first)
    k_join_wildcard<true><<<blocks, kWildcardWaves * 64, 0, s>>>(q, d, qel, del, bits, W64, indices, pair_graph, P, counts, stats);
  else
    k_join_wildcard<false><<<blocks, kWildcardWaves * 64, 0, s>>>(q, d, qel, del, bits, W64, indices, pair_graph, P, counts, stats);
  return hipGetLastError();
}

hipError_t join_lanes_wide(hipStream_t s, const GraphDev& q, const GraphDev& d, const uint64_t* bits, uint64_t W64, const uint32_t* indices,
                           const uint32_t* pair_graph, const uint32_t* list, uint64_t nw, int find_first, int32_t* counts,
                           unsigned long long* stats, int num_cus, const unsigned long long* nw_dev, uint32_t* deep) {
  return launch_join_wide(s, q, d, BitsetMasks{bits, W64}, indices, pair_graph, list, nw, find_first, counts, stats, num_cus, nw_dev, deep);
}

}  // namespace sigmo_k
