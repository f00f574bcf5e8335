// Lane-per-pair join shared by the plain form (masks cut from the candidate bitset) and the
// batched form (masks from the survivor arena).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "device_common.hpp"
#include "kernels.hpp"

namespace sigmo_k {
namespace {

// Candidate masks of the survivor arena: k consecutive words per pair (molecules of <= 64 nodes).
// `overflow` is the pair evaluation's arena-overflow flag: once set, offsets past the arena were
// handed out and no mask may be read, so the join does nothing and the host repeats the run with
// a larger arena.
struct ArenaMasks {
  const uint64_t* arena;
  const uint32_t* mask_off;
  const unsigned long long* overflow;
  __device__ __forceinline__ bool aborted() const { return uniform32(static_cast<uint32_t>(*overflow != 0ull)) != 0u; }
  __device__ __forceinline__ uint32_t base(uint64_t j) const { return mask_off[j]; }
  __device__ __forceinline__ uint64_t mask(uint32_t base, uint32_t, uint32_t depth, uint32_t, uint32_t) const {
    return arena[static_cast<uint64_t>(base) + depth];
  }
  // molecules of 65..128 nodes: two words per query node (positions 0-63, 64-127)
  __device__ __forceinline__ void mask2(uint32_t base, uint32_t, uint32_t depth, uint32_t, uint32_t, uint64_t& lo, uint64_t& hi) const {
    lo = arena[static_cast<uint64_t>(base) + 2 * depth];
    hi = arena[static_cast<uint64_t>(base) + 2 * depth + 1];
  }
  // The masks of query nodes 0..k-1 (0 from k on), two per 16-byte load: a pair's block starts on a
  // 4-word boundary and spans a multiple of 4 words, so a load that starts below k stays inside it.
  template <int N>
  __device__ __forceinline__ void masks(uint32_t base, uint32_t, uint32_t k, uint32_t, uint32_t, uint64_t (&m)[N]) const {
    static_assert(N % 2 == 0, "masks come in pairs");
#pragma unroll
    for (uint32_t t = 0; t < static_cast<uint32_t>(N); t += 2) {
      uint4 v = make_uint4(0, 0, 0, 0);
      if (t < k) v = *reinterpret_cast<const uint4*>(arena + static_cast<uint64_t>(base) + t);
      m[t] = static_cast<uint64_t>(v.x) | (static_cast<uint64_t>(v.y) << 32);
      m[t + 1] = t + 1 < k ? (static_cast<uint64_t>(v.z) | (static_cast<uint64_t>(v.w) << 32)) : 0ull;
    }
  }
  // The two-word masks of query nodes 0..k-1, one 16-byte load per node.
  template <int N>
  __device__ __forceinline__ void masks2(uint32_t base, uint32_t, uint32_t k, uint32_t, uint32_t, uint64_t (&lo)[N], uint64_t (&hi)[N]) const {
#pragma unroll
    for (uint32_t t = 0; t < static_cast<uint32_t>(N); ++t) {
      uint4 v = make_uint4(0, 0, 0, 0);
      if (t < k) v = *reinterpret_cast<const uint4*>(arena + static_cast<uint64_t>(base) + 2 * t);
      lo[t] = static_cast<uint64_t>(v.x) | (static_cast<uint64_t>(v.y) << 32);
      hi[t] = static_cast<uint64_t>(v.z) | (static_cast<uint64_t>(v.w) << 32);
    }
  }
};

// Candidate masks cut from the bitset row of the query node: positions s..s+n-1 (n <= 64 spans
// at most two 64-bit words), shifted to bit 0.
struct BitsetMasks {
  const uint64_t* bits;
  uint64_t W64;
  __device__ __forceinline__ bool aborted() const { return false; }
  __device__ __forceinline__ uint32_t base(uint64_t) const { return 0; }
  // Both words come from one 16-byte load.  The second may be the next row's first word, or the
  // word past the last row (the bitset allocation carries one spare word): it only fills bits
  // at positions >= 64 - sh, which n <= 64 - sh then masks off.
  __device__ __forceinline__ uint64_t mask(uint32_t, uint32_t qrow, uint32_t, uint32_t s, uint32_t n) const {
    typedef uint64_t u64x2 __attribute__((ext_vector_type(2))) __attribute__((aligned(8)));
    const uint64_t* row = bits + static_cast<uint64_t>(qrow) * W64;
    const uint32_t w0 = s >> 6, sh = s & 63;
    const u64x2 v = *reinterpret_cast<const u64x2*>(row + w0);
    const uint64_t lo = v.x, hi = v.y;
    uint64_t m = sh ? (lo >> sh) | (hi << (64 - sh)) : lo;
    if (n < 64) m &= (1ull << n) - 1;
    return m;
  }
  // positions s..s+n-1 with 64 < n <= 128: up to three words of the row, funnel-shifted
  __device__ __forceinline__ void mask2(uint32_t, uint32_t qrow, uint32_t, uint32_t s, uint32_t n, uint64_t& lo, uint64_t& hi) const {
    const uint64_t* row = bits + static_cast<uint64_t>(qrow) * W64;
    const uint32_t w0 = s >> 6, last = (s + n - 1) >> 6, sh = s & 63;
    const uint64_t x0 = row[w0];
    const uint64_t x1 = w0 + 1 <= last ? row[w0 + 1] : 0ull;
    const uint64_t x2 = w0 + 2 <= last ? row[w0 + 2] : 0ull;
    lo = sh ? (x0 >> sh) | (x1 << (64 - sh)) : x0;
    hi = sh ? (x1 >> sh) | (x2 << (64 - sh)) : x1;
    if (n < 128) hi &= (1ull << (n - 64)) - 1;
  }
  template <int N>
  __device__ __forceinline__ void masks(uint32_t base, uint32_t qoff, uint32_t k, uint32_t s, uint32_t n, uint64_t (&m)[N]) const {
#pragma unroll
    for (uint32_t t = 0; t < static_cast<uint32_t>(N); ++t) m[t] = t < k ? mask(base, qoff + t, t, s, n) : 0ull;
  }
  template <int N>
  __device__ __forceinline__ void masks2(uint32_t base, uint32_t qoff, uint32_t k, uint32_t s, uint32_t n, uint64_t (&lo)[N], uint64_t (&hi)[N]) const {
#pragma unroll
    for (uint32_t t = 0; t < static_cast<uint32_t>(N); ++t) {
      lo[t] = hi[t] = 0;
      if (t < k) mask2(base, qoff + t, t, s, n, lo[t], hi[t]);
    }
  }
};

constexpr uint32_t kJoinLaneWaves = 2;  // wavefronts per block of the lane-per-pair join

// Query depth the first lane-join pass holds in LDS.  Surviving pairs are mostly small queries
// (on 20 M molecules x 2,000 queries of U[4,32] nodes: 89% have k <= 8, 98.5% k <= 12), so the
// first pass keeps 12 levels per lane (6.75 KB per wavefront instead of 18 KB, 2.5x the resident
// wavefronts) and lists the deeper pairs for a second pass that keeps all 32.
constexpr int kJoinShallow = 12;
constexpr uint64_t kJoinTwoPassPairs = 1ull << 24;  // pair counts from which the lane join runs in two passes

// Join with one lane per GMCR pair, for molecules of <= 64 nodes (the rest are listed in
// `wide` for the two-word kernel) and queries of <= KMAX nodes (the rest are listed in `deep`
// for a second pass of this kernel with KMAX = 32 over that list).  Each lane runs the reference
// DFS of isomorphism.hpp:270-381 for its own pair with the reference's bookkeeping: every
// mapping entry starts at local node 0, an exhausted frame unsets the visited bit of its
// (possibly stale) mapping entry, `visited` is cleared when depth 1 is exhausted, and a frame
// resumes at mapping[depth] + 1.  The validity test of isomorphism.hpp:150-168 -- for every
// j < depth, qAdj(j, depth) == dAdj(mapping[j], c) -- becomes set algebra on adjacency rows:
// adjacency is symmetric, so the candidates that pass are those in the row of every mapped query
// neighbour and in the row of no other mapped node.  The rows of the mapped positions and the
// mapping sit in LDS (depth-major, one column per lane).  With `list` the pairs are list[0..P),
// claimed through stats[cursor]; without it they are 0..P through stats[kStatCursor].
template <bool FIND_FIRST, class Src, int KMAX = 32>
__global__ __launch_bounds__(kJoinLaneWaves * 64) void k_join_lanes(GraphDev q, GraphDev d, Src src, const uint32_t* __restrict__ indices,
                                                                   const uint32_t* __restrict__ pair_graph, uint64_t P, uint32_t chunk,
                                                                   int32_t* __restrict__ counts, unsigned long long* __restrict__ stats,
                                                                   uint32_t* __restrict__ wide, const unsigned long long* __restrict__ P_dev,
                                                                   const uint32_t* __restrict__ list = nullptr, uint32_t* __restrict__ deep = nullptr,
                                                                   uint32_t cursor = kStatCursor) {
  static_assert(KMAX >= 2 && KMAX <= 32, "query depth");
  __shared__ uint64_t rows_all[kJoinLaneWaves][KMAX][kWave];
  if (src.aborted()) return;  // block-uniform: no barrier follows
  // the pair (or list) count left on the device; P, what the arrays were sized for, bounds it
  if (P_dev) P = min(P, static_cast<uint64_t>(*P_dev));
  __shared__ uint8_t maps_all[kJoinLaneWaves][KMAX][kWave];
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint32_t wave = threadIdx.x / kWave;
  uint64_t (*rows)[kWave] = rows_all[wave];
  uint8_t (*maps)[kWave] = maps_all[wave];
  const uint64_t below = (1ull << lane) - 1;
  uint64_t next = 0, end = 0;  // claimed pairs [next, end) (wave-uniform)
  bool exhausted = false;
  uint64_t sum = 0, matched = 0, steps = 0, overflow = 0, wave_iters = 0;

  bool have = false;
  uint64_t j = 0, visited = 0, count = 0;
  uint32_t s = 0, n = 0, qoff = 0, k = 0, base = 0, depth = 0, pos = 0;
  // A shallow pass keeps its pair's candidate masks in registers, loaded together when the pair
  // is taken: the DFS then reads no mask from memory (each of them used to be one dependent load
  // per step, and the pairs' masks do not stay in L2 meanwhile).
  constexpr bool kRegMasks = KMAX <= 16;
  uint64_t mk[kRegMasks ? KMAX : 1];
#pragma unroll
  for (int t = 0; t < (kRegMasks ? KMAX : 1); ++t) mk[t] = 0;
  // Loads a step needs are issued a step ahead: `qad` is the query adjacency row of `depth`, and
  // after an accept the new mapped node's row waits in `nrow` (`fresh`) until the next step, which
  // uses it last and then stores it to LDS.
  uint32_t qad = 0;
  uint64_t nrow = 0;
  bool fresh = false;
  // Each lane prefetches its next pair while it runs the current one: nst = 1 once the pair's
  // index is claimed and its molecule, query graph and mask base are loading; nst = 2 once its
  // offsets and sizes are loading too.  A lane that finishes its pair takes the prefetched one
  // without waiting on that chain of dependent loads.
  uint32_t nst = 0, nj = 0, ndg = 0, nqg = 0, nbase = 0, ns = 0, nn = 0, nqoff = 0, nk = 0;
  while (true) {
    if (!have && nst == 2) {  // take the prefetched pair
      j = nj;
      s = ns;
      n = nn;
      qoff = nqoff;
      k = nk;
      base = nbase;
      for (uint32_t t = 0; t < k; ++t) maps[t][lane] = 0;  // every mapping entry starts at local node 0
      if constexpr (kRegMasks) src.masks(base, qoff, k, s, n, mk);
      qad = static_cast<uint32_t>(q.fold[qoff]);
      fresh = false;
      depth = 0;
      pos = 0;
      visited = 0;
      count = 0;
      have = true;
      nst = 0;
    }
    bool wide_pair = false, deep_pair = false;
    if (nst == 1) {  // the prefetched pair's offsets and sizes
      ns = d.go[ndg];
      nn = d.go[ndg + 1] - ns;
      nqoff = q.go[nqg];
      nk = q.go[nqg + 1] - nqoff;
      wide_pair = nn > 64;
      deep_pair = KMAX < 32 && !wide_pair && nk > static_cast<uint32_t>(KMAX);
      nst = wide_pair || deep_pair ? 0u : 2u;
    }
    const uint64_t wm = __ballot(wide_pair);  // molecules of > 64 nodes go to the two-word kernel
    if (wm) {
      unsigned long long b = 0;
      if (lane == 0) b = atomicAdd(&stats[kStatWide], static_cast<unsigned long long>(__popcll(wm)));
      const uint64_t wb = uniform64(__shfl(b, 0));
      if (wide_pair) wide[wb + __popcll(wm & below)] = nj;
    }
    const uint64_t dm = KMAX < 32 ? __ballot(deep_pair) : 0ull;  // queries deeper than KMAX go to the second pass
    if (dm) {
      unsigned long long b = 0;
      if (lane == 0) b = atomicAdd(&stats[kStatDeep], static_cast<unsigned long long>(__popcll(dm)));
      const uint64_t db = uniform64(__shfl(b, 0));
      if (deep_pair) deep[db + __popcll(dm & below)] = nj;
    }
    uint64_t need = __ballot(nst == 0);  // lanes with nothing prefetched claim a pair
    while (need && !exhausted) {
      if (next >= end) {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(&stats[cursor], static_cast<unsigned long long>(chunk));
        next = uniform64(__shfl(b, 0));
        end = next + chunk < P ? next + chunk : P;
        if (next >= P) {
          exhausted = true;
          break;
        }
      }
      const uint32_t cnt = min(static_cast<uint32_t>(end - next), static_cast<uint32_t>(__popcll(need)));
      const uint32_t rn = __popcll(need & below);
      const bool take = ((need >> lane) & 1ull) && rn < cnt;
      if (take) {
        nj = list ? list[next + rn] : static_cast<uint32_t>(next + rn);
        ndg = pair_graph[nj];
        nqg = indices[nj];
        nbase = src.base(nj);
        nst = 1;
      }
      next += cnt;
      need &= ~__ballot(take);
    }
    if (__ballot(have || nst != 0) == 0) break;
    ++wave_iters;
    if (have) {
      ++steps;
      bool done = false;
      if (depth == k) {  // leaf (:336-344)
        ++count;
        if (FIND_FIRST) {
          done = true;
        } else {
          // the pending row is the leaf's last node's, which no step reads
          fresh = false;
          depth = k - 1;
          pos = static_cast<uint32_t>(maps[depth][lane]) + 1;
          qad = static_cast<uint32_t>(q.fold[static_cast<uint64_t>(qoff) + depth]);
        }
      } else {
        // adjacency rows as 8-byte folded words: query graphs (<= 32 nodes) and these molecules
        // (<= 64 nodes) have no neighbour at positions >= 64, so fold = the low row word
        // candidates in the row of every mapped query neighbour and in no row of another mapped
        // node: AND over t of row_t, complemented where t is not adjacent to `depth`
        uint64_t md = 0;
        if (kRegMasks) {
#pragma unroll
          for (uint32_t t = 0; t < (kRegMasks ? KMAX : 1); ++t) md = t == depth ? mk[t] : md;
        } else {
          md = src.mask(base, qoff + depth, depth, s, n);
        }
        uint64_t el = md & ~visited & bits_from(pos);
        const uint32_t lim = fresh ? depth - 1 : depth;
        if (kRegMasks) {  // unrolled over KMAX; the compiler branches around the rows at t >= lim
#pragma unroll
          for (uint32_t t = 0; t < (kRegMasks ? KMAX : 1); ++t) {
            const uint64_t rw = rows[t][lane];
            if (t < lim) el &= rw ^ (static_cast<uint64_t>(((qad >> t) & 1u)) - 1ull);
          }
        } else {
          for (uint32_t t = 0; t < lim; ++t) el &= rows[t][lane] ^ (static_cast<uint64_t>(((qad >> t) & 1u)) - 1ull);
        }
        if (fresh) {
          el &= nrow ^ (static_cast<uint64_t>(((qad >> lim) & 1u)) - 1ull);
          rows[lim][lane] = nrow;
          fresh = false;
        }
        if (el) {  // accept (:366-370)
          const uint32_t c = static_cast<uint32_t>(__builtin_ctzll(el));
          maps[depth][lane] = static_cast<uint8_t>(c);
          nrow = d.fold[static_cast<uint64_t>(s) + c];
          fresh = true;
          visited |= 1ull << c;
          ++depth;
          pos = 0;
          if (depth < k) qad = static_cast<uint32_t>(q.fold[static_cast<uint64_t>(qoff) + depth]);
        } else {  // exhausted frame (:346-354)
          visited &= ~(1ull << (maps[depth][lane] & 63u));
          if (depth == 0) {
            done = true;
          } else {
            if (depth == 1) visited = 0;
            --depth;
            pos = static_cast<uint32_t>(maps[depth][lane]) + 1;
            qad = static_cast<uint32_t>(q.fold[static_cast<uint64_t>(qoff) + depth]);
          }
        }
      }
      if (done) {
        if (count > 0x7FFFFFFFull) overflow = 1;
        counts[j] = static_cast<int32_t>(count > 0x7FFFFFFFull ? 0x7FFFFFFFull : count);
        sum += count;
        matched += count > 0;
        have = false;
      }
    }
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o);
    matched += __shfl_xor(matched, o);
    steps += __shfl_xor(steps, o);
    overflow |= __shfl_xor(overflow, o);
  }
  if (lane == 0) {
    if (sum) atomicAdd(&stats[0], static_cast<unsigned long long>(sum));
    if (matched) atomicAdd(&stats[1], static_cast<unsigned long long>(matched));
    if (overflow) atomicOr(&stats[3], 1ull);
    atomicAdd(&stats[4], static_cast<unsigned long long>(steps));
    atomicAdd(&stats[7], static_cast<unsigned long long>(wave_iters));
  }
}

// The same lane-per-pair DFS for the pairs listed in `list` (molecules of 65..128 nodes).  The
// positions span two words: candidates and the adjacency rows of mapped positions are (lo, hi)
// pairs, while `visited` stays one word addressed by position mod 64, the reference's Bitset
// aliasing (utils.hpp:71-73), exactly as the wavefront join does for these molecules.  LDS holds
// two row words per depth and lane: with KMAX = kJoinShallow (12.75 KB per wavefront) the pass
// lists the pairs of deeper queries in `deep` (count in stats[kStatWideDeep]) for a second pass
// with KMAX = 32 (34.8 KB).  The shallow pass keeps its pair's masks in registers and, as the
// narrow lane join, prefetches its next pair, issues each step's loads a step ahead and unrolls
// its row loop.
template <bool FIND_FIRST, class Src, int KMAX = 32>
__global__ __launch_bounds__(64) void k_join_lanes_wide(GraphDev q, GraphDev d, Src src, const uint32_t* __restrict__ indices,
                                                        const uint32_t* __restrict__ pair_graph, const uint32_t* __restrict__ list, uint64_t P,
                                                        uint32_t chunk, uint32_t cursor, int32_t* __restrict__ counts,
                                                        unsigned long long* __restrict__ stats,
                                                        const unsigned long long* __restrict__ P_dev = nullptr,
                                                        uint32_t* __restrict__ deep = nullptr) {
  static_assert(KMAX >= 2 && KMAX <= 32, "query depth");
  __shared__ uint64_t rows_lo[KMAX][kWave];
  __shared__ uint64_t rows_hi[KMAX][kWave];
  __shared__ uint8_t maps[KMAX][kWave];
  if (src.aborted()) return;
  // the list's length as the lane join left it on the device, bounded by what the caller sized
  if (P_dev) P = min(P, static_cast<uint64_t>(*P_dev));
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint64_t below = (1ull << lane) - 1;
  uint64_t next = 0, end = 0;
  bool exhausted = false;
  uint64_t sum = 0, matched = 0, steps = 0, overflow = 0, wave_iters = 0;

  bool have = false;
  uint64_t j = 0, visited = 0, count = 0;
  uint32_t s = 0, n = 0, qoff = 0, k = 0, base = 0, depth = 0, pos = 0;
  constexpr bool kReg = KMAX <= 16;
  uint64_t mlo[kReg ? KMAX : 1], mhi[kReg ? KMAX : 1];
#pragma unroll
  for (int t = 0; t < (kReg ? KMAX : 1); ++t) mlo[t] = mhi[t] = 0;
  uint32_t qad = 0;
  uint64_t nlo = 0, nhi = 0;  // the accepted node's row words, stored to LDS by the next step
  bool fresh = false;
  uint32_t nst = 0, nj = 0, ndg = 0, nqg = 0, nbase = 0, ns = 0, nn = 0, nqoff = 0, nk = 0;
  while (true) {
    if (!have && nst == 2) {  // take the prefetched pair
      j = nj;
      s = ns;
      n = nn;
      qoff = nqoff;
      k = nk;
      base = nbase;
      for (uint32_t t = 0; t < k; ++t) maps[t][lane] = 0;  // every mapping entry starts at local node 0
      if constexpr (kReg) src.masks2(base, qoff, k, s, n, mlo, mhi);
      qad = static_cast<uint32_t>(q.adj[2 * static_cast<uint64_t>(qoff)]);
      fresh = false;
      depth = 0;
      pos = 0;
      visited = 0;
      count = 0;
      have = true;
      nst = 0;
    }
    bool deep_pair = false;
    if (nst == 1) {  // the prefetched pair's offsets and sizes
      ns = d.go[ndg];
      nn = d.go[ndg + 1] - ns;
      nqoff = q.go[nqg];
      nk = q.go[nqg + 1] - nqoff;
      deep_pair = KMAX < 32 && nk > static_cast<uint32_t>(KMAX);
      nst = deep_pair ? 0u : 2u;
    }
    const uint64_t dm = KMAX < 32 ? __ballot(deep_pair) : 0ull;  // queries deeper than KMAX go to the second pass
    if (dm) {
      unsigned long long b = 0;
      if (lane == 0) b = atomicAdd(&stats[kStatWideDeep], static_cast<unsigned long long>(__popcll(dm)));
      const uint64_t db = uniform64(__shfl(b, 0));
      if (deep_pair) deep[db + __popcll(dm & below)] = nj;
    }
    uint64_t need = __ballot(nst == 0);  // lanes with nothing prefetched claim a pair
    while (need && !exhausted) {
      if (next >= end) {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(&stats[cursor], static_cast<unsigned long long>(chunk));
        next = uniform64(__shfl(b, 0));
        end = next + chunk < P ? next + chunk : P;
        if (next >= P) {
          exhausted = true;
          break;
        }
      }
      const uint32_t cnt = min(static_cast<uint32_t>(end - next), static_cast<uint32_t>(__popcll(need)));
      const uint32_t rn = __popcll(need & below);
      const bool take = ((need >> lane) & 1ull) && rn < cnt;
      if (take) {
        nj = list[next + rn];
        ndg = pair_graph[nj];
        nqg = indices[nj];
        nbase = src.base(nj);
        nst = 1;
      }
      next += cnt;
      need &= ~__ballot(take);
    }
    if (__ballot(have || nst != 0) == 0) break;
    ++wave_iters;
    if (have) {
      ++steps;
      bool done = false;
      if (depth == k) {  // leaf (:336-344)
        ++count;
        if (FIND_FIRST) {
          done = true;
        } else {
          fresh = false;  // the pending rows are the leaf's last node's, which no step reads
          depth = k - 1;
          pos = static_cast<uint32_t>(maps[depth][lane]) + 1;
          qad = static_cast<uint32_t>(q.adj[2 * (static_cast<uint64_t>(qoff) + depth)]);
        }
      } else {
        uint64_t clo = 0, chi = 0;
        if (kReg) {
#pragma unroll
          for (uint32_t t = 0; t < (kReg ? KMAX : 1); ++t) {
            clo = t == depth ? mlo[t] : clo;
            chi = t == depth ? mhi[t] : chi;
          }
        } else {
          src.mask2(base, qoff + depth, depth, s, n, clo, chi);
        }
        clo &= ~visited & bits_from(pos);
        chi &= ~visited & (pos <= 64 ? ~0ull : bits_from(pos - 64));
        uint64_t el = clo, eh = chi;
        const uint32_t lim = fresh ? depth - 1 : depth;
        if (kReg) {  // unrolled over KMAX; the compiler branches around the rows at t >= lim
#pragma unroll
          for (uint32_t t = 0; t < (kReg ? KMAX : 1); ++t) {
            const uint64_t rl = rows_lo[t][lane], rh = rows_hi[t][lane];
            if (t < lim) {
              const uint64_t flip = static_cast<uint64_t>(((qad >> t) & 1u)) - 1ull;  // all ones where t is not adjacent
              el &= rl ^ flip;
              eh &= rh ^ flip;
            }
          }
        } else {
          for (uint32_t t = 0; t < lim; ++t) {
            const uint64_t flip = static_cast<uint64_t>(((qad >> t) & 1u)) - 1ull;
            el &= rows_lo[t][lane] ^ flip;
            eh &= rows_hi[t][lane] ^ flip;
          }
        }
        if (fresh) {
          const uint64_t flip = static_cast<uint64_t>(((qad >> lim) & 1u)) - 1ull;
          el &= nlo ^ flip;
          eh &= nhi ^ flip;
          rows_lo[lim][lane] = nlo;
          rows_hi[lim][lane] = nhi;
          fresh = false;
        }
        if (el | eh) {  // accept (:366-370)
          const uint32_t c = el ? static_cast<uint32_t>(__builtin_ctzll(el)) : 64u + static_cast<uint32_t>(__builtin_ctzll(eh));
          maps[depth][lane] = static_cast<uint8_t>(c);
          nlo = d.adj[2 * (static_cast<uint64_t>(s) + c)];
          nhi = d.adj[2 * (static_cast<uint64_t>(s) + c) + 1];
          fresh = true;
          visited |= 1ull << (c & 63);
          ++depth;
          pos = 0;
          if (depth < k) qad = static_cast<uint32_t>(q.adj[2 * (static_cast<uint64_t>(qoff) + depth)]);
        } else {  // exhausted frame (:346-354)
          visited &= ~(1ull << (maps[depth][lane] & 63u));
          if (depth == 0) {
            done = true;
          } else {
            if (depth == 1) visited = 0;
            --depth;
            pos = static_cast<uint32_t>(maps[depth][lane]) + 1;
            qad = static_cast<uint32_t>(q.adj[2 * (static_cast<uint64_t>(qoff) + depth)]);
          }
        }
      }
      if (done) {
        if (count > 0x7FFFFFFFull) overflow = 1;
        counts[j] = static_cast<int32_t>(count > 0x7FFFFFFFull ? 0x7FFFFFFFull : count);
        sum += count;
        matched += count > 0;
        have = false;
      }
    }
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) {
    sum += __shfl_xor(sum, o);
    matched += __shfl_xor(matched, o);
    steps += __shfl_xor(steps, o);
    overflow |= __shfl_xor(overflow, o);
  }
  if (lane == 0) {
    if (sum) atomicAdd(&stats[0], static_cast<unsigned long long>(sum));
    if (matched) atomicAdd(&stats[1], static_cast<unsigned long long>(matched));
    if (overflow) atomicOr(&stats[3], 1ull);
    atomicAdd(&stats[4], static_cast<unsigned long long>(steps));
    atomicAdd(&stats[7], static_cast<unsigned long long>(wave_iters));
  }
}

// The lane join over P pairs: with `deep`, a kJoinShallow-level pass over all of them that lists
// the deeper queries, then a full-depth pass over that list (its length read on the device);
// without it, one full-depth pass.  blocks_per_cu (<= 4 by LDS at full depth) and chunk (0: ~16
// claims per wavefront) shape the first pass.
template <class Src>
hipError_t launch_join_lanes(hipStream_t s, const GraphDev& q, const GraphDev& d, const Src& src, const uint32_t* indices,
                             const uint32_t* pair_graph, uint64_t P, int find_first, int32_t* counts, unsigned long long* stats,
                             uint32_t* wide, int num_cus, const unsigned long long* P_dev, uint32_t* deep, uint32_t blocks_per_cu,
                             uint32_t chunk) {
  if (!P) return hipSuccess;
  // Two passes pay off once the pairs fill the device many times over (1.4 x 10^8 pairs of a
  // 1 M-molecule x 2,000-query chunk: 50.6 -> 45.4 ms); on 1.5 x 10^6 pairs the second pass's
  // serial tail costs more than the first pass gains (0.53 -> 0.83 ms), so small joins run one
  // full-depth pass.
  if (P < kJoinTwoPassPairs) deep = nullptr;
  const uint64_t cus = static_cast<uint64_t>(num_cus > 0 ? num_cus : 256);
  // the shallow pass keeps 2.75x the wavefronts resident: up to 11 blocks per CU (13.5 KB of LDS
  // each, 80 VGPRs), which the LDS holds at once
  const uint64_t per_cu = deep ? static_cast<uint64_t>(blocks_per_cu) * 11 / 4 : blocks_per_cu;
  const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((P + 127) / 128, cus * per_cu));
  const uint32_t ch = chunk ? chunk
                            : static_cast<uint32_t>(std::min<uint64_t>(1024, std::max<uint64_t>(64, P / (blocks * kJoinLaneWaves * 16))));
  const dim3 grid(static_cast<unsigned>(blocks)), block(kJoinLaneWaves * 64);
  if (!deep) {
    if (find_first)
      hipLaunchKernelGGL((k_join_lanes<true, Src, 32>), grid, block, 0, s, q, d, src, indices, pair_graph, P, ch, counts, stats, wide, P_dev,
                         nullptr, nullptr, kStatCursor);
    else
      hipLaunchKernelGGL((k_join_lanes<false, Src, 32>), grid, block, 0, s, q, d, src, indices, pair_graph, P, ch, counts, stats, wide, P_dev,
                         nullptr, nullptr, kStatCursor);
    return hipGetLastError();
  }
  if (find_first)
    hipLaunchKernelGGL((k_join_lanes<true, Src, kJoinShallow>), grid, block, 0, s, q, d, src, indices, pair_graph, P, ch, counts, stats, wide,
                       P_dev, nullptr, deep, kStatCursor);
  else
    hipLaunchKernelGGL((k_join_lanes<false, Src, kJoinShallow>), grid, block, 0, s, q, d, src, indices, pair_graph, P, ch, counts, stats, wide,
                       P_dev, nullptr, deep, kStatCursor);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // the deep list is short (queries of more than kJoinShallow nodes): small claims, full depth
  const uint64_t dblocks = std::max<uint64_t>(1, std::min<uint64_t>((P + 127) / 128, cus * 4));
  const dim3 dgrid(static_cast<unsigned>(dblocks));
  if (find_first)
    hipLaunchKernelGGL((k_join_lanes<true, Src, 32>), dgrid, block, 0, s, q, d, src, indices, pair_graph, P, 16u, counts, stats, wide,
                       stats + kStatDeep, deep, nullptr, kStatDeepCursor);
  else
    hipLaunchKernelGGL((k_join_lanes<false, Src, 32>), dgrid, block, 0, s, q, d, src, indices, pair_graph, P, 16u, counts, stats, wide,
                       stats + kStatDeep, deep, nullptr, kStatDeepCursor);
  return hipGetLastError();
}

// The two-word lane join over the nw pairs listed in `list` (their count also at nw_dev when
// given, nw then only bounding the grid).  With `deep` (capacity nw) a kJoinShallow-level pass
// (12 resident one-wavefront blocks per CU) lists the deeper queries for a full-depth pass over
// that list (4 blocks per CU); without it, one full-depth pass.
template <class Src>
hipError_t launch_join_wide(hipStream_t s, const GraphDev& q, const GraphDev& d, const Src& src, const uint32_t* indices,
                            const uint32_t* pair_graph, const uint32_t* list, uint64_t nw, int find_first, int32_t* counts,
                            unsigned long long* stats, int num_cus, const unsigned long long* nw_dev, uint32_t* deep) {
  if (!nw) return hipSuccess;
  const uint64_t cus = static_cast<uint64_t>(num_cus > 0 ? num_cus : 256);
  const uint64_t per_cu = deep ? 12 : 4;
  const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>((nw + 63) / 64, cus * per_cu));
  const uint32_t chunk = static_cast<uint32_t>(std::min<uint64_t>(1024, std::max<uint64_t>(64, nw / (blocks * 16))));
  if (!deep) {
    if (find_first)
      hipLaunchKernelGGL((k_join_lanes_wide<true, Src, 32>), dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s, q, d, src, indices, pair_graph,
                         list, nw, chunk, static_cast<uint32_t>(kStatWideCursor), counts, stats, nw_dev, nullptr);
    else
      hipLaunchKernelGGL((k_join_lanes_wide<false, Src, 32>), dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s, q, d, src, indices, pair_graph,
                         list, nw, chunk, static_cast<uint32_t>(kStatWideCursor), counts, stats, nw_dev, nullptr);
    return hipGetLastError();
  }
  if (find_first)
    hipLaunchKernelGGL((k_join_lanes_wide<true, Src, kJoinShallow>), dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s, q, d, src, indices,
                       pair_graph, list, nw, chunk, static_cast<uint32_t>(kStatWideCursor), counts, stats, nw_dev, deep);
  else
    hipLaunchKernelGGL((k_join_lanes_wide<false, Src, kJoinShallow>), dim3(static_cast<unsigned>(blocks)), dim3(64), 0, s, q, d, src, indices,
                       pair_graph, list, nw, chunk, static_cast<uint32_t>(kStatWideCursor), counts, stats, nw_dev, deep);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const uint64_t dblocks = std::max<uint64_t>(1, std::min<uint64_t>((nw + 63) / 64, cus * 4));
  if (find_first)
    hipLaunchKernelGGL((k_join_lanes_wide<true, Src, 32>), dim3(static_cast<unsigned>(dblocks)), dim3(64), 0, s, q, d, src, indices, pair_graph,
                       deep, nw, 16u, static_cast<uint32_t>(kStatWideDeepCursor), counts, stats, stats + kStatWideDeep, nullptr);
  else
    hipLaunchKernelGGL((k_join_lanes_wide<false, Src, 32>), dim3(static_cast<unsigned>(dblocks)), dim3(64), 0, s, q, d, src, indices, pair_graph,
                       deep, nw, 16u, static_cast<uint32_t>(kStatWideDeepCursor), counts, stats, stats + kStatWideDeep, nullptr);
  return hipGetLastError();
}

}  // namespace
}  // namespace sigmo_k
