// Batched (non-materialising) form of SIGMo's filter -> refine -> GMCR -> join.
//
// The plain form writes a Q x D candidate bitset and passes over it 2R+2 times; after R = 6
// levels ~1% of its bits are set and ~8% of the (query graph, molecule) pairs survive GMCR.
// The batched form computes, for each (query graph, molecule) pair, the refined candidate
// mask of each query node directly inside the molecule, from level signatures prepared once
// per molecule, and stops at the first query node with no candidate -- the GMCR predicate of
// gmcr.hpp:56-61 (every query node has a candidate in the molecule).  Survivors keep their
// per-node masks, which are exactly the molecule-range slices of the plain form's rows, so
// GMCR and join results are identical.
//
// Semantics per query node q and molecule position p (paths relative to the SIGMo repository):
//   label(q) == label(p) or label(q) == 255                         isomorphism.hpp:49
//   for r = 1..R, for l < 16: sig_r(q)[l] <= sig_r(p)[l]             isomorphism.hpp:103-110
// with sig_r the PowerGraph level-r signature of signature.hpp:336-390 (sphere at BFS distance
// r over positions mod 64, labels read through the 64-position view).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sigmo_hip.h"
#include "device_common.hpp"
#include "host_graphs.hpp"
#include "join_dfs.hpp"
#include "kernels.hpp"

namespace sigmo_k {
namespace {

constexpr uint32_t kMaxLevels = 8;     // requirement encoding keeps 3 bits of level
constexpr uint32_t kReqStride = kMaxLevels * 16;
constexpr uint32_t kNoWide = 0xFFFFFFFFu;

// Level-r signature of every node's sphere, one BFS layer per level (the reachables of the
// reference equal the BFS ball of the previous level when levels are refined in order).
__device__ __forceinline__ uint64_t next_layer(const GraphDev& g, uint32_t prev, uint64_t& ball, uint64_t& frontier) {
  uint64_t next = 0, f = frontier;
  while (f) {
    const uint32_t b = __builtin_ctzll(f);
    f &= f - 1;
    const uint64_t u = static_cast<uint64_t>(prev) + b;
    next |= g.adj[2 * u] | g.adj[2 * u + 1];
  }
  next &= ~ball;
  ball |= next;
  frontier = next;
  return next;
}

__device__ __forceinline__ uint64_t histogram16(uint64_t r, const uint64_t* __restrict__ lm) {
  uint64_t s = 0;
#pragma unroll
  for (int l = 0; l < 16; ++l) {
    uint32_t c = __popcll(r & lm[l]);
    c = c < 15 ? c : 15;
    s |= static_cast<uint64_t>(c) << (4 * l);
  }
  return s;
}

// One wavefront per molecule: lane p is position p (and p+64 for molecules of 65..128 nodes).
// planes[(m*R + r)*64 + i]: bit p = bit i of the level-(r+1) signature of position p, i.e.
// plane 4*label + bit; planes_hi the same for positions 64..127; lab_hi[w*16+l]: positions
// 64..127 with label l.  Positions 0..63 use the graph's label masks.
__global__ __launch_bounds__(256) void k_db_prepare(GraphDev g, uint32_t R, const uint32_t* __restrict__ wide_idx,
                                                    uint64_t* __restrict__ planes, uint64_t* __restrict__ planes_hi,
                                                    uint64_t* __restrict__ lab_hi) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint64_t m = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (m >= g.G) return;  // wave-uniform
  const uint32_t s = g.go[m], n = g.go[m + 1] - s;
  const uint64_t* lm = g.lmask + m * 16;
  const uint32_t halves = n > 64 ? 2 : 1;
  const uint32_t wi = wide_idx[m];
  for (uint32_t h = 0; h < halves; ++h) {
    const uint32_t pos = lane + 64 * h;
    const bool valid = pos < n;
    uint64_t self = 1ull << (pos & 63), ball = self, frontier = self;
    for (uint32_t r = 0; r < R; ++r) {
      uint64_t sig = 0;
      if (valid) sig = histogram16(next_layer(g, s, ball, frontier), lm);
      uint64_t mine = 0;
#pragma unroll
      for (int i = 0; i < 64; ++i) {
        const uint64_t b = __ballot((sig >> i) & 1ull);
        if (lane == static_cast<uint32_t>(i)) mine = b;
      }
      if (h == 0) planes[(m * R + r) * 64 + lane] = mine;
      else planes_hi[(static_cast<uint64_t>(wi) * R + r) * 64 + lane] = mine;
    }
    if (h == 1) {
      const uint32_t lab = valid ? g.nl[s + pos] : 0xFFu;
      uint64_t mine = 0;
#pragma unroll
      for (int l = 0; l < 16; ++l) {
        const uint64_t b = __ballot(lab == static_cast<uint32_t>(l));
        if (lane == static_cast<uint32_t>(l)) mine = b;
      }
      if (lane < 16) lab_hi[static_cast<uint64_t>(wi) * 16 + lane] = mine;
    }
  }
}

// Query side: all R level signatures of every query node (query graphs have <= 32 nodes, so
// positions never alias), turned into requirement lists: (level, label, count) with count > 0,
// encoded as (r) | (l << 3) | (c << 7), levels ascending.
__global__ __launch_bounds__(256) void k_query_requirements(GraphDev q, uint32_t R, uint16_t* __restrict__ reqs,
                                                            uint8_t* __restrict__ nreq) {
  const uint64_t v = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (v >= q.N) return;
  const uint32_t gid = q.node_graph[v];
  const uint32_t prev = q.go[gid];
  uint64_t self = 1ull << ((static_cast<uint32_t>(v) - prev) & 63u), ball = self, frontier = self;
  uint32_t cnt = 0;
  for (uint32_t r = 0; r < R; ++r) {
    const uint64_t sig = histogram16(next_layer(q, prev, ball, frontier), q.lmask + static_cast<uint64_t>(gid) * 16);
    for (uint32_t l = 0; l < 16; ++l) {
      const uint32_t c = static_cast<uint32_t>(sig >> (4 * l)) & 15u;
      if (c) reqs[v * kReqStride + cnt++] = static_cast<uint16_t>(r | (l << 3) | (c << 7));
    }
  }
  nreq[v] = static_cast<uint8_t>(cnt);
}

struct PairArgs {
  GraphDev q, d;
  uint32_t R;
  const uint64_t* planes;
  const uint64_t* planes_hi;
  const uint64_t* lab_hi;
  const uint32_t* wide_idx;
  const uint16_t* reqs;
  const uint8_t* nreq;
  uint64_t* scratch;      // [wave][lane][32][2]
  uint64_t* arena;        // survivor masks: k words (2k for molecules of > 64 nodes)
  uint64_t arena_cap;
  uint32_t* rec_qg;       // records, per molecule a contiguous block in query-graph order
  uint32_t* rec_off;
  uint64_t rec_cap;
  uint32_t* mol_block;    // [Gd] first record of the molecule
  uint32_t* mol_count;    // [Gd]
  unsigned long long* counters;  // [0] molecule cursor, [1] arena top, [2] record top, [3] overflow flag
  uint32_t mol_chunk;
};

// Candidate mask of query node q at molecule m restricted to the label test (filter).
__device__ __forceinline__ void label_mask(const PairArgs& a, uint64_t m, uint32_t s, uint32_t n, uint32_t wi, uint32_t ql,
                                           uint64_t& lo, uint64_t& hi) {
  const uint64_t all_lo = n >= 64 ? ~0ull : ((1ull << n) - 1);
  const uint64_t all_hi = n <= 64 ? 0ull : (n >= 128 ? ~0ull : ((1ull << (n - 64)) - 1));
  if (ql == 255u) {
    lo = all_lo;
    hi = all_hi;
  } else if (ql < 16u) {
    lo = a.d.lmask[m * 16 + ql];
    hi = n > 64 ? a.lab_hi[static_cast<uint64_t>(wi) * 16 + ql] : 0ull;
  } else {
    lo = hi = 0;
    for (uint32_t p = 0; p < n; ++p)
      if (a.d.nl[s + p] == ql) {
        if (p < 64) lo |= 1ull << p;
        else hi |= 1ull << (p - 64);
      }
  }
}

// Persistent pair filter.  A wavefront takes molecules from a cursor; its lanes take query
// graphs of that molecule in ascending order, each lane running one pair as a small state
// machine (start node / apply one requirement / finish node) so that lanes whose pair died
// take the next query graph at once.  Per molecule, surviving query graphs are collected in
// LDS (bitmask + arena offset) and emitted as one block of records in query-graph order.
__global__ __launch_bounds__(256) void k_pair_filter(PairArgs a) {
  extern __shared__ uint8_t smem[];
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t Gq = a.q.G;
  const uint32_t words = (Gq + 63) / 64;
  // per-wave LDS: survivor bits [words] u64, arena offsets [Gq] u32
  const size_t per_wave = static_cast<size_t>(words) * 8 + static_cast<size_t>(Gq) * 4;
  uint64_t* surv = reinterpret_cast<uint64_t*>(smem + wave * ((per_wave + 15) & ~size_t(15)));
  uint32_t* slot = reinterpret_cast<uint32_t*>(surv + words);
  const uint64_t gwave = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + wave;
  uint64_t* scratch = a.scratch + (gwave * 64 + lane) * 64;
  const uint32_t R = a.R;

  uint64_t mnext = 0, mend = 0;
  while (true) {
    if (mnext >= mend) {
      unsigned long long b = 0;
      if (lane == 0) b = atomicAdd(&a.counters[0], static_cast<unsigned long long>(a.mol_chunk));
      mnext = uniform64(__shfl(b, 0));
      mend = mnext + a.mol_chunk < a.d.G ? mnext + a.mol_chunk : a.d.G;
      if (mnext >= a.d.G) break;
    }
    const uint64_t m = mnext++;
    const uint32_t s = a.d.go[m], n = a.d.go[m + 1] - s;
    const bool wide = n > 64;
    const uint32_t wi = a.wide_idx[m];
    for (uint32_t w = lane; w < words; w += kWave) surv[w] = 0;
    __builtin_amdgcn_wave_barrier();

    bool have = false, fresh = false;
    uint32_t qg = 0, i = 0, k = 0, qoff = 0, ri = 0, nr = 0;
    uint64_t lo = 0, hi = 0;
    uint32_t next_qg = 0;  // wave-uniform
    while (true) {
      const uint64_t need = __ballot(!have);
      if (need && next_qg < Gq) {
        const uint32_t rank = __popcll(need & ((1ull << lane) - 1));
        if (!have && next_qg + rank < Gq) {
          qg = next_qg + rank;
          qoff = a.q.go[qg];
          k = a.q.go[qg + 1] - qoff;
          i = 0;
          have = true;
          fresh = true;
        }
        next_qg = next_qg + __popcll(need) < Gq ? next_qg + __popcll(need) : Gq;
      }
      if (__ballot(have) == 0) break;
      if (have) {
        const uint32_t q = qoff + i;
        if (fresh) {  // start query node i of the pair: the filter's label test
          label_mask(a, m, s, n, wi, a.q.nl[q], lo, hi);
          ri = 0;
          nr = a.nreq[q];
          fresh = false;
          if (!(lo | hi)) have = false;
        } else if (ri < nr) {  // one refine requirement: count_q(l) <= count_p(l) at level r
          const uint32_t req = a.reqs[static_cast<uint64_t>(q) * kReqStride + ri];
          const uint32_t r = req & 7u, l = (req >> 3) & 15u, c = (req >> 7) & 15u;
          const uint64_t* pl = a.planes + (m * R + r) * 64 + l * 4;
          lo &= ge4_lane(pl[0], pl[1], pl[2], pl[3], c);
          if (wide) {
            const uint64_t* ph = a.planes_hi + (static_cast<uint64_t>(wi) * R + r) * 64 + l * 4;
            hi &= ge4_lane(ph[0], ph[1], ph[2], ph[3], c);
          }
          ++ri;
          if (!(lo | hi)) have = false;
        } else {  // node done: keep its mask
          scratch[2 * i] = lo;
          scratch[2 * i + 1] = hi;
          ++i;
          if (i < k) {
            fresh = true;
          } else {  // every query node has a candidate: the pair survives GMCR
            const uint32_t stride = wide ? 2 : 1;
            const unsigned long long off = atomicAdd(&a.counters[1], static_cast<unsigned long long>(k * stride));
            if (off + k * stride <= a.arena_cap) {
              for (uint32_t t = 0; t < k; ++t) {
                a.arena[off + t * stride] = scratch[2 * t];
                if (wide) a.arena[off + t * stride + 1] = scratch[2 * t + 1];
              }
            } else {
              atomicOr(&a.counters[3], 1ull);
            }
            slot[qg] = static_cast<uint32_t>(off);
            atomicOr(reinterpret_cast<unsigned long long*>(&surv[qg >> 6]), 1ull << (qg & 63));
            have = false;
          }
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
    // emit this molecule's survivors as one block of records, ascending query graph
    uint32_t mine = 0;
    for (uint32_t w = lane; w < words; w += kWave) mine += __popcll(surv[w]);
    uint32_t total = mine;
    for (int o = kWave / 2; o > 0; o >>= 1) total += __shfl_xor(total, o);
    unsigned long long blk = 0;
    if (lane == 0 && total) blk = atomicAdd(&a.counters[2], static_cast<unsigned long long>(total));
    blk = uniform64(__shfl(blk, 0));
    if (lane == 0) {
      a.mol_block[m] = static_cast<uint32_t>(blk);
      a.mol_count[m] = total;
    }
    if (total && blk + total > a.rec_cap) {
      if (lane == 0) atomicOr(&a.counters[3], 2ull);
    } else if (total) {
      // lane-strided words; exclusive prefix of the popcounts in word order
      uint64_t base = 0;
      for (uint32_t w0 = 0; w0 < words; w0 += kWave) {
        const uint32_t w = w0 + lane;
        const uint64_t x = w < words ? surv[w] : 0ull;
        uint32_t c = __popcll(x), incl = c;
        for (int o = 1; o < kWave; o <<= 1) {
          const uint32_t y = __shfl_up(incl, o);
          if (lane >= static_cast<uint32_t>(o)) incl += y;
        }
        uint64_t pos = blk + base + incl - c;
        for (uint64_t t = x; t; t &= t - 1) {
          const uint32_t g = w * 64 + __builtin_ctzll(t);
          a.rec_qg[pos] = g;
          a.rec_off[pos] = slot[g];
          ++pos;
        }
        base += __shfl(incl, kWave - 1);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// Canonical GMCR arrays from the per-molecule record blocks.
__global__ __launch_bounds__(256) void k_pairs_scatter(uint32_t Gd, const uint32_t* __restrict__ offsets,
                                                       const uint32_t* __restrict__ mol_block, const uint32_t* __restrict__ rec_qg,
                                                       const uint32_t* __restrict__ rec_off, uint32_t* __restrict__ indices,
                                                       uint32_t* __restrict__ pair_graph, uint32_t* __restrict__ mask_off) {
  const uint64_t m = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (m >= Gd) return;
  const uint32_t o = offsets[m], c = offsets[m + 1] - o, b = mol_block[m];
  for (uint32_t i = 0; i < c; ++i) {
    indices[o + i] = rec_qg[b + i];
    pair_graph[o + i] = static_cast<uint32_t>(m);
    mask_off[o + i] = rec_off[b + i];
  }
}

// Join over the canonical pairs with candidate masks from the arena (the molecule-local masks
// the plain form extracts from the bitset rows).
__device__ __forceinline__ void join_load_arena(const GraphDev& q, const GraphDev& d, const uint64_t* __restrict__ arena,
                                                const uint32_t* __restrict__ mask_off, const uint32_t* __restrict__ indices,
                                                const uint32_t* __restrict__ pair_graph, uint64_t j, const JoinPair* prev, JoinPair& o) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint32_t dg = pair_graph[j], qg = indices[j];
  o.dg = dg;
  o.s = d.go[dg];
  o.e = d.go[dg + 1];
  o.n = o.e - o.s;
  o.qoff = q.go[qg];
  o.k = q.go[qg + 1] - o.qoff;
  if (prev && prev->dg == dg) {
    o.alo = prev->alo;
    o.ahi = prev->ahi;
    o.blo = prev->blo;
    o.bhi = prev->bhi;
  } else {
    o.alo = o.ahi = o.blo = o.bhi = 0;
    if (lane < o.n) {
      o.alo = d.adj[2 * (static_cast<uint64_t>(o.s) + lane)];
      o.ahi = d.adj[2 * (static_cast<uint64_t>(o.s) + lane) + 1];
    }
    if (lane + 64 < o.n) {
      o.blo = d.adj[2 * (static_cast<uint64_t>(o.s) + lane + 64)];
      o.bhi = d.adj[2 * (static_cast<uint64_t>(o.s) + lane + 64) + 1];
    }
  }
  o.clo = o.chi = 0;
  o.qa = 0;
  if (lane < o.k) {
    const uint64_t base = mask_off[j];
    if (o.n > 64) {
      o.clo = arena[base + 2 * lane];
      o.chi = arena[base + 2 * lane + 1];
    } else {
      o.clo = arena[base + lane];
    }
    o.qa = static_cast<uint32_t>(q.adj[2 * static_cast<uint64_t>(o.qoff + lane)]);
  }
}

template <bool FIND_FIRST>
__global__ __launch_bounds__(256) void k_join_arena(GraphDev q, GraphDev d, const uint64_t* __restrict__ arena,
                                                    const uint32_t* __restrict__ mask_off, const uint32_t* __restrict__ indices,
                                                    const uint32_t* __restrict__ pair_graph, uint64_t P, uint32_t chunk,
                                                    int32_t* __restrict__ counts, unsigned long long* __restrict__ stats) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  uint64_t sum = 0, matched = 0, steps = 0, overflow = 0;
  uint64_t next = 0, end = 0;
  auto grab = [&]() {
    unsigned long long jj = 0;
    if (lane == 0) jj = atomicAdd(&stats[2], static_cast<unsigned long long>(chunk));
    next = uniform64(__shfl(jj, 0));
    end = next + chunk < P ? next + chunk : P;
  };
  grab();
  if (next < P) {
    JoinPair cur, nxt;
    join_load_arena(q, d, arena, mask_off, indices, pair_graph, next, nullptr, cur);
    while (true) {
      const uint64_t j = next++;
      if (next >= end) grab();
      const bool more = next < P;
      if (more) join_load_arena(q, d, arena, mask_off, indices, pair_graph, next, &cur, nxt);
      const uint64_t count = cur.n > 64 ? join_dfs<FIND_FIRST, true>(cur, steps) : join_dfs<FIND_FIRST, false>(cur, steps);
      if (lane == 0) {
        if (count > 0x7FFFFFFFull) overflow = 1;
        counts[j] = static_cast<int32_t>(count > 0x7FFFFFFFull ? 0x7FFFFFFFull : count);
      }
      sum += count;
      matched += count > 0;
      if (!more) break;
      cur = nxt;
    }
  }
  if (lane == 0) {
    if (sum) atomicAdd(&stats[0], static_cast<unsigned long long>(sum));
    if (matched) atomicAdd(&stats[1], static_cast<unsigned long long>(matched));
    if (overflow) atomicOr(&stats[3], 1ull);
    atomicAdd(&stats[4], static_cast<unsigned long long>(steps));
  }
}

}  // namespace

hipError_t db_prepare(hipStream_t s, const GraphDev& g, uint32_t R, const uint32_t* wide_idx, uint64_t* planes, uint64_t* planes_hi,
                      uint64_t* lab_hi) {
  if (!g.G) return hipSuccess;
  k_db_prepare<<<(g.G + 3) / 4, 256, 0, s>>>(g, R, wide_idx, planes, planes_hi, lab_hi);
  return hipGetLastError();
}

hipError_t query_requirements(hipStream_t s, const GraphDev& q, uint32_t R, uint16_t* reqs, uint8_t* nreq) {
  if (!q.N) return hipSuccess;
  k_query_requirements<<<static_cast<unsigned>((q.N + 255) / 256), 256, 0, s>>>(q, R, reqs, nreq);
  return hipGetLastError();
}

}  // namespace sigmo_k

// ---- host objects and C ABI -----------------------------------------------------------------------

namespace {

#define BATCH_HIP_TRY(expr)                                                                        \
  do {                                                                                             \
    hipError_t _e = (expr);                                                                        \
    if (_e != hipSuccess)                                                                          \
      return sigmo_err::set(_e == hipErrorOutOfMemory ? SIGMO_ERR_NOMEM : SIGMO_ERR_HIP,           \
                            std::string(#expr) + ": " + hipGetErrorString(_e));                    \
  } while (0)

#define BATCH_TRY(expr)            \
  do {                             \
    int _s = (expr);               \
    if (_s != SIGMO_OK) return _s; \
  } while (0)

template <typename T>
int balloc(T** p, uint64_t n) {
  *p = nullptr;
  hipError_t e = hipMalloc(reinterpret_cast<void**>(p), (n ? n : 1) * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return sigmo_err::set(SIGMO_ERR_NOMEM, "hipMalloc of " + std::to_string(n * sizeof(T)) + " bytes failed");
  }
  return SIGMO_OK;
}
template <typename T>
void bfree(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}
template <typename T>
int bgrow(T** p, uint64_t& cap, uint64_t need) {
  if (*p && cap >= need) return SIGMO_OK;
  bfree(*p);
  cap = 0;
  BATCH_TRY(balloc(p, need));
  cap = need;
  return SIGMO_OK;
}

}  // namespace

// Internal views of the objects defined in capi.hip.
namespace sigmo_internal {
int ctx_info(sigmo_ctx* ctx, int* device, hipStream_t* stream, int* num_cus, unsigned long long** stats, unsigned long long** stats_host);
int ctx_timed(sigmo_ctx* ctx, int phase, int kernel, double bytes, bool begin, void** token);
void ctx_add_bytes(sigmo_ctx* ctx, int kernel, double bytes);
const sigmo_k::GraphDev graph_dev(const sigmo_graph* g);
const std::vector<uint32_t>& graph_host_offsets(const sigmo_graph* g);
sigmo_ctx* graph_ctx(const sigmo_graph* g);
}  // namespace sigmo_internal

using namespace sigmo_internal;

struct sigmo_db {
  sigmo_ctx* ctx = nullptr;
  const sigmo_graph* data = nullptr;
  uint32_t R = 0;
  uint32_t num_wide = 0;
  uint32_t* wide_idx = nullptr;
  uint64_t *planes = nullptr, *planes_hi = nullptr, *lab_hi = nullptr;
  uint64_t bytes = 0;
  ~sigmo_db() {
    bfree(wide_idx);
    bfree(planes);
    bfree(planes_hi);
    bfree(lab_hi);
  }
};

struct sigmo_batch {
  sigmo_ctx* ctx = nullptr;
  const sigmo_db* db = nullptr;
  // query side
  uint16_t* reqs = nullptr;
  uint8_t* nreq = nullptr;
  uint64_t cap_reqs = 0, cap_nreq = 0;
  // pair filter
  uint64_t* scratch = nullptr;
  uint64_t cap_scratch = 0;
  uint64_t* arena = nullptr;
  uint64_t cap_arena = 0;
  uint32_t *rec_qg = nullptr, *rec_off = nullptr;
  uint64_t cap_rec = 0, cap_rec2 = 0;
  uint32_t *mol_block = nullptr, *mol_count = nullptr;
  uint64_t cap_mb = 0, cap_mc = 0;
  unsigned long long* counters = nullptr;
  // canonical GMCR and join
  uint32_t *offsets = nullptr, *indices = nullptr, *pair_graph = nullptr, *mask_off = nullptr;
  uint64_t cap_off = 0, cap_idx = 0, cap_pg = 0, cap_mo = 0;
  uint64_t* scan_tmp = nullptr;
  uint64_t cap_scan = 0;
  uint64_t* total = nullptr;
  int32_t* counts = nullptr;
  uint64_t cap_counts = 0;
  uint64_t P = 0;
  uint32_t Gq = 0;
  bool ran = false;
  ~sigmo_batch() {
    bfree(reqs);
    bfree(nreq);
    bfree(scratch);
    bfree(arena);
    bfree(rec_qg);
    bfree(rec_off);
    bfree(mol_block);
    bfree(mol_count);
    bfree(counters);
    bfree(offsets);
    bfree(indices);
    bfree(pair_graph);
    bfree(mask_off);
    bfree(scan_tmp);
    bfree(total);
    bfree(counts);
  }
};

extern "C" {

int sigmo_db_prepare(sigmo_ctx* ctx, const sigmo_graph* data, uint32_t refinement_steps, sigmo_db** out) {
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *st, *sh;
  BATCH_TRY(ctx_info(ctx, &device, &s, &cus, &st, &sh));
  if (!data || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (refinement_steps == 0 || refinement_steps > sigmo_k::kMaxLevels)
    return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "the batched form supports 1..8 refinement levels");
  const auto& go = graph_host_offsets(data);
  const uint32_t G = static_cast<uint32_t>(go.size() - 1);
  std::vector<uint32_t> widx(G, sigmo_k::kNoWide);
  uint32_t nw = 0;
  for (uint32_t g = 0; g < G; ++g)
    if (go[g + 1] - go[g] > 64) widx[g] = nw++;
  auto* db = new sigmo_db();
  db->ctx = ctx;
  db->data = data;
  db->R = refinement_steps;
  db->num_wide = nw;
  const uint64_t R = refinement_steps;
  int stt = balloc(&db->wide_idx, G);
  if (stt == SIGMO_OK) stt = balloc(&db->planes, uint64_t(G) * R * 64);
  if (stt == SIGMO_OK) stt = balloc(&db->planes_hi, uint64_t(nw) * R * 64);
  if (stt == SIGMO_OK) stt = balloc(&db->lab_hi, uint64_t(nw) * 16);
  if (stt != SIGMO_OK) {
    delete db;
    return stt;
  }
  db->bytes = G * 4ull + (uint64_t(G) + nw) * R * 512 + nw * 128ull;
  hipError_t e = hipMemcpyAsync(db->wide_idx, widx.data(), G * 4ull, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    delete db;
    return sigmo_err::set(SIGMO_ERR_HIP, std::string("db prepare: ") + hipGetErrorString(e));
  }
  stt = sigmo_db_refresh(db);
  if (stt == SIGMO_OK && hipStreamSynchronize(s) != hipSuccess) stt = sigmo_err::set(SIGMO_ERR_HIP, "db prepare failed");
  if (stt != SIGMO_OK) {
    delete db;
    return stt;
  }
  *out = db;
  return SIGMO_OK;
}

int sigmo_db_refresh(sigmo_db* db) {
  if (!db) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *st, *sh;
  BATCH_TRY(ctx_info(db->ctx, &device, &s, &cus, &st, &sh));
  const sigmo_k::GraphDev g = graph_dev(db->data);
  // algorithmic bytes: CSR adjacency words and label masks read, planes written
  const double bytes = double(g.N) * 16.0 + double(g.G) * 128.0 + double(db->bytes);
  void* tok = nullptr;
  BATCH_TRY(ctx_timed(db->ctx, SIGMO_PHASE_SIG_DATA, SIGMO_KERNEL_DB_PREPARE, bytes, true, &tok));
  BATCH_HIP_TRY(sigmo_k::db_prepare(s, g, db->R, db->wide_idx, db->planes, db->planes_hi, db->lab_hi));
  BATCH_TRY(ctx_timed(db->ctx, 0, 0, 0.0, false, &tok));
  return SIGMO_OK;
}

int sigmo_db_free(sigmo_db* db) {
  if (!db) return SIGMO_OK;
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *st, *sh;
  if (ctx_info(db->ctx, &device, &s, &cus, &st, &sh) == SIGMO_OK) (void)hipStreamSynchronize(s);
  delete db;
  return SIGMO_OK;
}

int sigmo_db_alloc_size(const sigmo_db* db, uint64_t* bytes) {
  if (!db || !bytes) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  *bytes = db->bytes;
  return SIGMO_OK;
}

int sigmo_batch_create(sigmo_ctx* ctx, const sigmo_db* db, sigmo_batch** out) {
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *st, *sh;
  BATCH_TRY(ctx_info(ctx, &device, &s, &cus, &st, &sh));
  if (!db || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  auto* b = new sigmo_batch();
  b->ctx = ctx;
  b->db = db;
  int stt = balloc(&b->counters, 8);
  if (stt == SIGMO_OK) stt = balloc(&b->total, 1);
  if (stt != SIGMO_OK) {
    delete b;
    return stt;
  }
  *out = b;
  return SIGMO_OK;
}

int sigmo_batch_free(sigmo_batch* b) {
  if (!b) return SIGMO_OK;
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *st, *sh;
  if (ctx_info(b->ctx, &device, &s, &cus, &st, &sh) == SIGMO_OK) (void)hipStreamSynchronize(s);
  delete b;
  return SIGMO_OK;
}

int sigmo_batch_run(sigmo_batch* b, const sigmo_graph* query, int find_first, sigmo_batch_result* res) {
  if (!b || !query || !res) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *stats, *stats_host;
  BATCH_TRY(ctx_info(b->ctx, &device, &s, &cus, &stats, &stats_host));
  std::memset(res, 0, sizeof(*res));
  const sigmo_db* db = b->db;
  const sigmo_k::GraphDev qd = graph_dev(query), dd = graph_dev(db->data);
  const auto& qgo = graph_host_offsets(query);
  const uint32_t Gq = qd.G, Gd = dd.G;
  uint64_t ksum = 0;
  for (uint32_t g = 0; g < Gq; ++g) {
    const uint32_t k = qgo[g + 1] - qgo[g];
    if (k < 2) return sigmo_err::set(SIGMO_ERR_INVALID, "query graph " + std::to_string(g) + " has fewer than 2 nodes");
    if (k > SIGMO_MAX_QUERY_NODES) return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "query graph exceeds 32 nodes");
    ksum += k;
  }
  const size_t lds_per_wave = ((static_cast<size_t>((Gq + 63) / 64) * 8 + static_cast<size_t>(Gq) * 4) + 15) & ~size_t(15);
  if (lds_per_wave * 4 > 64 * 1024)
    return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "the batched form takes at most 3968 query graphs per run");
  // buffers
  const uint64_t waves = static_cast<uint64_t>(cus) * 8 * 4;  // persistent pair-filter grid: 8 blocks of 4 waves per CU
  BATCH_TRY(bgrow(&b->reqs, b->cap_reqs, qd.N * sigmo_k::kReqStride));
  BATCH_TRY(bgrow(&b->nreq, b->cap_nreq, qd.N));
  BATCH_TRY(bgrow(&b->scratch, b->cap_scratch, waves * 64 * 64));
  BATCH_TRY(bgrow(&b->mol_block, b->cap_mb, Gd));
  BATCH_TRY(bgrow(&b->mol_count, b->cap_mc, uint64_t(Gd) + 1));
  BATCH_TRY(bgrow(&b->offsets, b->cap_off, uint64_t(Gd) + 1));
  BATCH_TRY(bgrow(&b->scan_tmp, b->cap_scan, (uint64_t(Gd) + 1023) / 1024 + 1));
  // record / arena capacity: start from 1/8 of the pair space (grown and replayed on overflow)
  const double mean_k = Gq ? double(ksum) / Gq : 0.0;
  uint64_t want_rec = std::max<uint64_t>(uint64_t(Gq) * Gd / 8, 1 << 16);
  if (b->cap_rec > want_rec) want_rec = b->cap_rec;
  {
    void* tok = nullptr;
    BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_SIG_QUERY, -1, 0.0, true, &tok));
    BATCH_HIP_TRY(sigmo_k::query_requirements(s, qd, db->R, b->reqs, b->nreq));
    BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
  }
  for (int attempt = 0; attempt < 6; ++attempt) {
    const uint64_t want_arena = static_cast<uint64_t>(want_rec * (mean_k + 1.0) * 1.1) + 64;
    BATCH_TRY(bgrow(&b->rec_qg, b->cap_rec, want_rec));
    BATCH_TRY(bgrow(&b->rec_off, b->cap_rec2, want_rec));
    BATCH_TRY(bgrow(&b->arena, b->cap_arena, want_arena));
    if (b->cap_arena >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "candidate arena exceeds 2^32 words");

    // algorithmic reads: each molecule's level planes (every plane a query node may test), label
    // masks and offsets, and the query requirement lists; writes are added once they are known
    const double reads = double(Gd) * db->R * 512.0 + double(db->num_wide) * db->R * 512.0 + double(Gd) * (128.0 + 8.0) +
                         double(qd.N) * (1.0 + 32.0 * db->R);
    void* tok = nullptr;
    BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_FILTER, SIGMO_KERNEL_PAIR_FILTER, reads, true, &tok));
    BATCH_HIP_TRY(sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(b->counters), 8, 0));
    sigmo_k::PairArgs a;
    a.q = qd;
    a.d = dd;
    a.R = db->R;
    a.planes = db->planes;
    a.planes_hi = db->planes_hi;
    a.lab_hi = db->lab_hi;
    a.wide_idx = db->wide_idx;
    a.reqs = b->reqs;
    a.nreq = b->nreq;
    a.scratch = b->scratch;
    a.arena = b->arena;
    a.arena_cap = b->cap_arena;
    a.rec_qg = b->rec_qg;
    a.rec_off = b->rec_off;
    a.rec_cap = b->cap_rec < b->cap_rec2 ? b->cap_rec : b->cap_rec2;
    a.mol_block = b->mol_block;
    a.mol_count = b->mol_count;
    a.counters = b->counters;
    a.mol_chunk = 4;
    const uint64_t blocks = std::min<uint64_t>(waves / 4, (uint64_t(Gd) + 3) / 4);
    hipLaunchKernelGGL(sigmo_k::k_pair_filter, dim3(static_cast<unsigned>(blocks)), dim3(256), lds_per_wave * 4, s, a);
    BATCH_HIP_TRY(hipGetLastError());
    BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
    unsigned long long cnt[4];
    BATCH_HIP_TRY(hipMemcpyAsync(cnt, b->counters, sizeof(cnt), hipMemcpyDeviceToHost, s));
    BATCH_HIP_TRY(hipStreamSynchronize(s));
    // writes: candidate masks, (query graph, arena offset) records, per-molecule block and count
    ctx_add_bytes(b->ctx, SIGMO_KERNEL_PAIR_FILTER, double(cnt[1]) * 8.0 + double(cnt[2]) * 8.0 + double(Gd) * 8.0);
    if (!cnt[3]) break;
    if (attempt == 5) return sigmo_err::set(SIGMO_ERR_NOMEM, "pair buffers overflowed repeatedly");
    want_rec = std::max<uint64_t>(want_rec * 2, cnt[2] + 1);  // replay with larger buffers
  }
  // canonical GMCR: scan the per-molecule survivor counts, scatter the blocks
  void* tok = nullptr;
  BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_GMCR, -1, 0.0, true, &tok));
  BATCH_HIP_TRY(sigmo_k::exclusive_scan(s, b->mol_count, Gd, b->offsets, b->scan_tmp, b->total));
  uint64_t P = 0;
  BATCH_HIP_TRY(hipMemcpyAsync(&P, b->total, 8, hipMemcpyDeviceToHost, s));
  BATCH_HIP_TRY(hipStreamSynchronize(s));
  if (P >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "GMCR pair count exceeds u32");
  BATCH_TRY(bgrow(&b->indices, b->cap_idx, P));
  BATCH_TRY(bgrow(&b->pair_graph, b->cap_pg, P));
  BATCH_TRY(bgrow(&b->mask_off, b->cap_mo, P));
  BATCH_TRY(bgrow(&b->counts, b->cap_counts, P));
  if (Gd) {
    hipLaunchKernelGGL(sigmo_k::k_pairs_scatter, dim3((Gd + 255) / 256), dim3(256), 0, s, Gd, b->offsets, b->mol_block, b->rec_qg,
                       b->rec_off, b->indices, b->pair_graph, b->mask_off);
    BATCH_HIP_TRY(hipGetLastError());
  }
  BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
  // join
  BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_JOIN, SIGMO_KERNEL_JOIN, 0.0, true, &tok));
  BATCH_HIP_TRY(sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(stats), 8, 0));
  if (P) {
    uint64_t blocks = std::min<uint64_t>((P + 3) / 4, static_cast<uint64_t>(cus) * 8);
    uint64_t chunk = P / (blocks * 4 * 8);
    chunk = chunk < 1 ? 1 : (chunk > 64 ? 64 : chunk);
    if (find_first)
      hipLaunchKernelGGL(sigmo_k::k_join_arena<true>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, qd, dd, b->arena, b->mask_off,
                         b->indices, b->pair_graph, P, static_cast<uint32_t>(chunk), b->counts, stats);
    else
      hipLaunchKernelGGL(sigmo_k::k_join_arena<false>, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, qd, dd, b->arena, b->mask_off,
                         b->indices, b->pair_graph, P, static_cast<uint32_t>(chunk), b->counts, stats);
    BATCH_HIP_TRY(hipGetLastError());
  }
  BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
  BATCH_HIP_TRY(hipMemcpyAsync(stats_host, stats, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
  BATCH_HIP_TRY(hipStreamSynchronize(s));
  if (stats_host[3]) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "a per-pair match count exceeded int32");
  b->P = P;
  b->Gq = Gq;
  b->ran = true;
  res->num_matches = stats_host[0];
  res->matched_pairs = stats_host[1];
  res->gmcr_pairs = P;
  res->dfs_steps = stats_host[4];
  return SIGMO_OK;
}

int sigmo_batch_pairs(sigmo_batch* b, uint32_t* data_graph_offsets, uint64_t n_offsets, uint32_t* query_graph_indices, int32_t* counts,
                      uint64_t n_pairs) {
  if (!b || !b->ran) return sigmo_err::set(SIGMO_ERR_INVALID, "batch has not run");
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *st, *sh;
  BATCH_TRY(ctx_info(b->ctx, &device, &s, &cus, &st, &sh));
  const uint64_t Gd = graph_dev(b->db->data).G;
  if (data_graph_offsets) {
    if (n_offsets != Gd + 1) return sigmo_err::set(SIGMO_ERR_INVALID, "n_offsets must be Gd+1");
    BATCH_HIP_TRY(hipMemcpyAsync(data_graph_offsets, b->offsets, n_offsets * 4, hipMemcpyDeviceToHost, s));
  }
  if ((query_graph_indices || counts) && n_pairs != b->P) return sigmo_err::set(SIGMO_ERR_INVALID, "n_pairs must be the pair count");
  if (query_graph_indices && n_pairs)
    BATCH_HIP_TRY(hipMemcpyAsync(query_graph_indices, b->indices, n_pairs * 4, hipMemcpyDeviceToHost, s));
  if (counts && n_pairs) BATCH_HIP_TRY(hipMemcpyAsync(counts, b->counts, n_pairs * 4, hipMemcpyDeviceToHost, s));
  BATCH_HIP_TRY(hipStreamSynchronize(s));
  return SIGMO_OK;
}

}  // extern "C"
