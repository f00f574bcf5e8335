// The reference's wildcard/edge-label join, joinWildcardCandidates (isomorphism.hpp:428-556),
// one lane per GMCR pair.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "kernels.hpp"

namespace sigmo_k {
namespace {

constexpr uint32_t kWildcardWaves = 4;  // wavefronts per block

// Label of the first CSR entry of row a that points at b (graph.hpp:88-93), 255 if none.
__device__ __forceinline__ uint8_t first_edge_label(const uint32_t* __restrict__ ro, const uint32_t* __restrict__ ci,
                                                    const uint8_t* __restrict__ el, uint32_t a, uint32_t b) {
  for (uint32_t i = ro[a], e = ro[a + 1]; i < e; ++i)
    if (ci[i] == b) return el[i];
  return 0xFFu;
}

// Position of the idx-th set bit (ascending) of the 128-bit set (lo, hi).
__device__ __forceinline__ uint32_t nth_bit128(uint64_t lo, uint64_t hi, uint32_t idx) {
  const uint32_t cl = static_cast<uint32_t>(__popcll(lo));
  uint64_t x = idx < cl ? lo : hi;
  uint32_t r = idx < cl ? idx : idx - cl;
  for (; r; --r) x &= x - 1;
  return (idx < cl ? 0u : 64u) + static_cast<uint32_t>(__builtin_ctzll(x));
}

// Per pair, as the reference kernel does per sub-group:
//  * matching order (defineMatchingOrder, :383-408): the query node with most candidates in
//    the molecule first (ties: the larger index; every k <= 32 node is considered, as with a
//    sub-group of >= 32 lanes), then the other nodes ascending;
//  * one DFS per candidate of that node (:495-548), each from fresh state (every mapping
//    entry = local node 0), summed into the pair's count; find-first ends a DFS at its first
//    leaf;
//  * a candidate is valid when, for every earlier node of the order, query and data adjacency
//    agree and adjacent pairs carry equal edge labels (isValidMapping, :410-426): adjacency by
//    set algebra on the mapped positions' rows, labels by the first CSR entry of each pair;
//  * an exhausted frame unsets its node's (possibly stale) mapping, and a pop that leaves one
//    frame resets `visited` to the first node's candidate (:521-526); `visited` is one word
//    indexed by position mod 64 (utils.hpp:71-73).
// A frame resumes after its last accepted candidate, which is what advancing the reference's
// candidate index past it does.
template <bool FIND_FIRST>
__global__ __launch_bounds__(kWildcardWaves * 64) void k_join_wildcard(GraphDev q, GraphDev d, const uint8_t* __restrict__ qel,
                                                                       const uint8_t* __restrict__ del, const uint64_t* __restrict__ bits,
                                                                       uint64_t W64, const uint32_t* __restrict__ indices,
                                                                       const uint32_t* __restrict__ pair_graph, uint64_t P,
                                                                       int32_t* __restrict__ counts, unsigned long long* __restrict__ stats) {
  __shared__ uint8_t maps_all[kWildcardWaves][32][kWave];
  __shared__ uint8_t order_all[kWildcardWaves][32][kWave];
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint32_t wave = threadIdx.x / kWave;
  uint8_t (*maps)[kWave] = maps_all[wave];
  uint8_t (*order)[kWave] = order_all[wave];
  uint64_t sum = 0, matched = 0, overflow = 0;
  // candidates of query node t of the pair inside the molecule, as (lo, hi) position words
  auto mask2 = [&](uint32_t qrow, uint32_t s, uint32_t n, uint64_t& lo, uint64_t& hi) {
    const uint64_t* row = bits + static_cast<uint64_t>(qrow) * W64;
    const uint32_t w0 = s >> 6, last = (s + n - 1) >> 6, sh = s & 63;
    const uint64_t x0 = row[w0];
    const uint64_t x1 = w0 + 1 <= last ? row[w0 + 1] : 0ull;
    const uint64_t x2 = w0 + 2 <= last ? row[w0 + 2] : 0ull;
    lo = sh ? (x0 >> sh) | (x1 << (64 - sh)) : x0;
    hi = sh ? (x1 >> sh) | (x2 << (64 - sh)) : x1;
    if (n < 64) {
      lo &= (1ull << n) - 1;
      hi = 0;
    } else if (n == 64) {
      hi = 0;
    } else if (n < 128) {
      hi &= (1ull << (n - 64)) - 1;
    }
  };
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; j < P; j += stride) {
    const uint32_t dg = pair_graph[j], qg = indices[j];
    const uint32_t s = d.go[dg], n = d.go[dg + 1] - s;
    const uint32_t qoff = q.go[qg], k = q.go[qg + 1] - qoff;
    // matching order
    uint32_t first = 0, maxc = 0;
    for (uint32_t t = 0; t < k; ++t) {
      uint64_t lo, hi;
      mask2(qoff + t, s, n, lo, hi);
      const uint32_t c = static_cast<uint32_t>(__popcll(lo) + __popcll(hi));
      if (c >= maxc) {
        maxc = c;
        first = t;
      }
    }
    order[0][lane] = static_cast<uint8_t>(first);
    for (uint32_t t = 0, o = 1; t < k; ++t)
      if (t != first) order[o++][lane] = static_cast<uint8_t>(t);
    uint64_t flo, fhi;
    mask2(qoff + first, s, n, flo, fhi);
    uint64_t count = 0;
    for (uint32_t start = 0; start < maxc; ++start) {
      for (uint32_t t = 0; t < k; ++t) maps[t][lane] = 0;  // every mapping entry starts at local node 0
      const uint32_t root = nth_bit128(flo, fhi, start);
      maps[first][lane] = static_cast<uint8_t>(root);
      uint64_t visited = 1ull << (root & 63);
      uint32_t depth = 1, pos = 0;
      while (depth > 0) {
        if (depth == k) {  // leaf
          ++count;
          if (FIND_FIRST) break;
          depth = k - 1;
          pos = static_cast<uint32_t>(maps[order[depth][lane]][lane]) + 1;
          continue;
        }
        const uint32_t qn = order[depth][lane];
        uint64_t clo, chi;
        mask2(qoff + qn, s, n, clo, chi);
        clo &= ~visited & bits_from(pos);
        chi &= ~visited & (pos <= 64 ? ~0ull : bits_from(pos - 64));
        // adjacency agreement with every earlier node of the order
        uint64_t must_lo = ~0ull, must_hi = ~0ull, not_lo = 0, not_hi = 0;
        uint32_t nbr = 0;  // orders i < depth whose node is a query neighbour of qn
        for (uint32_t i = 0; i < depth; ++i) {
          const uint32_t qi = order[i][lane];
          const uint32_t m = maps[qi][lane];
          const uint64_t rl = d.adj[2 * (static_cast<uint64_t>(s) + m)], rh = d.adj[2 * (static_cast<uint64_t>(s) + m) + 1];
          if ((static_cast<uint32_t>(q.adj[2 * (static_cast<uint64_t>(qoff) + qi)]) >> qn) & 1u) {
            must_lo &= rl;
            must_hi &= rh;
            nbr |= 1u << i;
          } else {
            not_lo |= rl;
            not_hi |= rh;
          }
        }
        uint64_t el = clo & must_lo & ~not_lo, eh = chi & must_hi & ~not_hi;
        int accepted = -1;
        while (el | eh) {  // ascending; the first whose edge labels all agree
          const uint32_t c = el ? static_cast<uint32_t>(__builtin_ctzll(el)) : 64u + static_cast<uint32_t>(__builtin_ctzll(eh));
          if (el) el &= el - 1;
          else eh &= eh - 1;
          bool ok = true;
          for (uint32_t b = nbr; b && ok; b &= b - 1) {
            const uint32_t qi = order[__builtin_ctz(b)][lane];
            const uint32_t m = maps[qi][lane];
            ok = first_edge_label(d.ro, d.ci, del, s + m, s + c) == first_edge_label(q.ro, q.ci, qel, qoff + qi, qoff + qn);
          }
          if (ok) {
            accepted = static_cast<int>(c);
            break;
          }
        }
        if (accepted >= 0) {
          maps[qn][lane] = static_cast<uint8_t>(accepted);
          visited |= 1ull << (accepted & 63);
          ++depth;
          pos = 0;
        } else {  // exhausted frame: pop
          visited &= ~(1ull << (maps[qn][lane] & 63u));
          --depth;
          if (depth == 1) visited = 1ull << (maps[first][lane] & 63u);
          if (depth > 0) pos = static_cast<uint32_t>(maps[order[depth][lane]][lane]) + 1;
        }
      }
    }
    if (count > 0x7FFFFFFFull) overflow = 1;
    counts[j] = static_cast<int32_t>(count > 0x7FFFFFFFull ? 0x7FFFFFFFull : count);
    sum += count;
    matched += count > 0;
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o);
    matched += __shfl_xor(matched, o);
    overflow |= __shfl_xor(overflow, o);
  }
  if (lane == 0) {
    if (sum) atomicAdd(&stats[0], static_cast<unsigned long long>(sum));
    if (matched) atomicAdd(&stats[1], static_cast<unsigned long long>(matched));
    if (overflow) atomicOr(&stats[3], 1ull);
  }
}

}  // namespace
}  // namespace sigmo_k
