// The wavefront-per-pair DFS of joinCandidates (isomorphism.hpp:270-381), shared by the plain
// and batched pipelines: per-pair operands and the reference DFS on lane-distributed state.
#pragma once

#include "device_common.hpp"
#include "kernels.hpp"

namespace sigmo_k {

// Per-pair operands of the wavefront join: molecule range, query range, the lane-distributed
// candidate masks / query adjacency (lane d = query node d) and molecule adjacency (lane p =
// node p, and node p+64 for molecules of more than 64 nodes).
struct JoinPair {
  uint32_t dg, s, e, n, qoff, k;
  uint64_t clo, chi;  // candidates of query node `lane` at molecule positions 0-63 / 64-127
  uint32_t qa;        // adjacency of query node `lane` by query-local id
  uint64_t alo, ahi, blo, bhi;
};

__device__ __forceinline__ void join_load(const GraphDev& q, const GraphDev& d, const uint64_t* __restrict__ bits, uint64_t W64,
                                          const uint32_t* __restrict__ indices, const uint32_t* __restrict__ pair_graph, uint64_t j,
                                          const JoinPair* prev, JoinPair& o) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint32_t dg = pair_graph[j], qg = indices[j];
  o.dg = dg;
  o.s = d.go[dg];
  o.e = d.go[dg + 1];
  o.n = o.e - o.s;
  o.qoff = q.go[qg];
  o.k = q.go[qg + 1] - o.qoff;
  const uint32_t s = o.s, e = o.e, n = o.n, k = o.k;
  if (prev && prev->dg == dg) {
    o.alo = prev->alo;
    o.ahi = prev->ahi;
    o.blo = prev->blo;
    o.bhi = prev->bhi;
  } else {
    o.alo = o.ahi = o.blo = o.bhi = 0;
    if (lane < n) {
      o.alo = d.adj[2 * (static_cast<uint64_t>(s) + lane)];
      o.ahi = d.adj[2 * (static_cast<uint64_t>(s) + lane) + 1];
    }
    if (lane + 64 < n) {
      o.blo = d.adj[2 * (static_cast<uint64_t>(s) + lane + 64)];
      o.bhi = d.adj[2 * (static_cast<uint64_t>(s) + lane + 64) + 1];
    }
  }
  o.clo = o.chi = 0;
  o.qa = 0;
  if (lane < k) {
    const uint64_t* row = bits + static_cast<uint64_t>(o.qoff + lane) * W64;
    const uint32_t w0 = s >> 6, last = (e - 1) >> 6;
    o.clo = row[w0];
    o.chi = w0 + 1 <= last ? row[w0 + 1] : 0;
    o.qa = static_cast<uint32_t>(q.adj[2 * static_cast<uint64_t>(o.qoff + lane)]);
    // a third word (molecules straddling three words) is read in join_finish
  }
}

// Turn the raw words of join_load into molecule-local candidate masks (funnel shift by s mod 64).
__device__ __forceinline__ void join_finish(const uint64_t* __restrict__ bits, uint64_t W64, JoinPair& o) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  if (lane < o.k) {
    const uint32_t sh = o.s & 63, w0 = o.s >> 6, last = (o.e - 1) >> 6;
    uint64_t x0 = o.clo, x1 = o.chi, x2 = 0;
    if (w0 + 2 <= last) x2 = bits[static_cast<uint64_t>(o.qoff + lane) * W64 + w0 + 2];
    uint64_t lo = sh ? (x0 >> sh) | (x1 << (64 - sh)) : x0;
    uint64_t hi = sh ? (x1 >> sh) | (x2 << (64 - sh)) : x1;
    if (o.n < 64) lo &= (1ull << o.n) - 1;
    if (o.n <= 64) hi = 0;
    else if (o.n < 128) hi &= (1ull << (o.n - 64)) - 1;
    o.clo = lo;
    o.chi = hi;
  }
}

// The reference DFS of one pair on the lane-distributed operands (see k_join_wave).  WIDE
// handles molecules of 65..128 nodes (positions 64..127 in the `hi` words).
//
// The set of candidates of depth d that pass isValidMapping depends only on mapping[0..d-1],
// which cannot change while the search is at depth >= d.  It is computed once when the search
// descends to d (ballot over the molecule's positions) and kept in lane d; a revisit after
// backtracking reads it back and only applies the current `visited` set and the resume
// position, which is what the reference re-evaluates per candidate.
template <bool FIND_FIRST, bool WIDE>
__device__ __forceinline__ uint64_t join_dfs(const JoinPair& o, uint64_t& steps) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint32_t k = uniform32(o.k);
  uint32_t map = 0;             // lane d: mapping[d] (local id; every entry starts at node 0)
  uint64_t mslo = 0, mshi = 0;  // lane d: positions mapped at depths < d on the current path
  uint64_t elo = 0, ehi = 0;    // lane d: valid candidates of depth d for the current mapping[0..d-1]
  uint32_t depth = 0, pos = 0;
  bool fresh = true;            // first visit of this depth since mapping[0..depth-1] last changed
  uint64_t visited = 0, count = 0;
  while (true) {
    ++steps;
    if (depth == k) {  // leaf (:336-344)
      ++count;
      if (FIND_FIRST) break;
      depth = k - 1;
      pos = rl32(map, depth) + 1;
      fresh = false;
      continue;
    }
    uint64_t El, Eh;
    if (fresh) {
      El = rl64(o.clo, depth);
      Eh = WIDE ? rl64(o.chi, depth) : 0ull;
      if (depth > 0 && (El | Eh)) {
        // A: positions that must be adjacent (mapped query neighbours), N: must not be
        const uint32_t qad = rl32(o.qa, depth), low = (1u << depth) - 1;
        uint64_t Al = 0, Ah = 0, Nl, Nh;
        for (uint32_t t = qad & low; t; t &= t - 1) {
          const uint32_t m = rl32(map, static_cast<uint32_t>(__builtin_ctz(t)));
          if (!WIDE || m < 64) Al |= 1ull << (m & 63);
          else Ah |= 1ull << (m & 63);
        }
        const uint64_t Ml = rl64(mslo, depth), Mh = WIDE ? rl64(mshi, depth) : 0ull;
        if (static_cast<uint32_t>(__popcll(Ml) + __popcll(Mh)) == depth) {  // injective: N = mapped \ A
          Nl = Ml & ~Al;
          Nh = Mh & ~Ah;
        } else {  // a stale unset let a node be mapped twice: collect N exactly
          Nl = Nh = 0;
          for (uint32_t t = ~qad & low; t; t &= t - 1) {
            const uint32_t m = rl32(map, static_cast<uint32_t>(__builtin_ctz(t)));
            if (!WIDE || m < 64) Nl |= 1ull << (m & 63);
            else Nh |= 1ull << (m & 63);
          }
        }
        if (El) {
          const bool ok = (o.alo & Al) == Al && !(o.alo & Nl) && (!WIDE || ((o.ahi & Ah) == Ah && !(o.ahi & Nh)));
          El &= __ballot(ok);
        }
        if (WIDE && Eh) {
          const bool ok = (o.blo & Al) == Al && !(o.blo & Nl) && (o.bhi & Ah) == Ah && !(o.bhi & Nh);
          Eh &= __ballot(ok);
        }
      }
      if (lane == depth) {
        elo = El;
        ehi = Eh;
      }
      fresh = false;
    } else {
      El = rl64(elo, depth);
      Eh = WIDE ? rl64(ehi, depth) : 0ull;
    }
    const uint64_t el = El & ~visited & bits_from(pos);
    const uint64_t eh = WIDE ? Eh & ~visited & (pos <= 64 ? ~0ull : bits_from(pos - 64)) : 0ull;
    if (el | eh) {  // accept (:366-370)
      const uint32_t c = el ? static_cast<uint32_t>(__builtin_ctzll(el)) : 64u + static_cast<uint32_t>(__builtin_ctzll(eh));
      const uint64_t Ml = rl64(mslo, depth), Mh = WIDE ? rl64(mshi, depth) : 0ull;
      if (lane == depth) map = c;
      if (lane == depth + 1) {
        mslo = Ml | (c < 64 ? 1ull << c : 0ull);
        if (WIDE) mshi = Mh | (c >= 64 ? 1ull << (c & 63) : 0ull);
      }
      visited |= 1ull << (c & 63);
      ++depth;
      pos = 0;
      fresh = true;
    } else {  // exhausted frame (:346-354)
      visited &= ~(1ull << (rl32(map, depth) & 63));
      if (depth == 0) break;
      if (depth == 1) visited = 0;
      --depth;
      pos = rl32(map, depth) + 1;
      fresh = false;
    }
  }
  return count;
}

}  // namespace sigmo_k
