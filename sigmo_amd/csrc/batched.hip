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

constexpr uint32_t kMaxLevels = 8;        // token encoding keeps 3 bits of level
constexpr uint32_t kNoWide = 0xFFFFFFFFu;
constexpr uint16_t kEndToken = 0xFFFF;    // closes a query node's token segment
constexpr uint32_t kArenaChunk = 1024;    // arena words a wavefront claims at a time
constexpr uint32_t kRecordChunk = 256;    // pair records a wavefront claims at a time

// Tokens of one query node: [label] [requirement]* [end], requirement = r | l << 3 | c << 7
// (level r < 8, label l < 16, count 1 <= c <= 15).  Segments are 4-token aligned with one spare
// word, so the reader's one-word lookahead stays inside the segment.
__host__ __device__ constexpr uint32_t token_segment(uint32_t R) { return ((2 + 16 * R + 3) & ~3u) + 4; }

// Orders this wavefront's LDS accesses across lanes (the compiler may not move them across).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// Level-r sphere of every node, one BFS layer per call (the reachables of the reference are
// the BFS ball of the previous level when levels are refined in order).  Positions are taken
// mod 64: bit b of the frontier expands node prev + b, as in signature.hpp:336-390.
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

// 16 four-bit label counts of the nodes in r (labels >= 16 are not counted; counts saturate
// at 15), signature.hpp:44-51.
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

// Nibble-wise a <= b for all 16 nibbles (even and odd nibbles in 8-bit fields, borrow-free).
__device__ __forceinline__ bool nib_le_all(uint64_t a, uint64_t b) {
  constexpr uint64_t L = 0x0F0F0F0F0F0F0F0Full, H = 0x1010101010101010ull;
  const uint64_t te = ((b & L) | H) - (a & L);
  const uint64_t to = (((b >> 4) & L) | H) - ((a >> 4) & L);
  return (te & to & H) == H;
}

// Nibble-wise max.
__device__ __forceinline__ uint64_t nib_max(uint64_t a, uint64_t b) {
  constexpr uint64_t L = 0x0F0F0F0F0F0F0F0Full, H = 0x1010101010101010ull;
  const uint64_t ae = a & L, be = b & L, ao = (a >> 4) & L, bo = (b >> 4) & L;
  const uint64_t me = ((((ae | H) - be) & H) >> 4) * 0xFull;  // 0x0F in the fields where a >= b
  const uint64_t mo = ((((ao | H) - bo) & H) >> 4) * 0xFull;
  return ((ae & me) | (be & ~me)) | (((ao & mo) | (bo & ~mo)) << 4);
}

__device__ __forceinline__ uint64_t wave_nib_max(uint64_t x) {
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) x = nib_max(x, static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(x), o)));
  return x;
}

// 64 x 64 bit-matrix transpose across the wavefront: lane p holds row p on entry, lane i
// holds column i (bit p = bit i of lane p's row) on exit.  Six butterfly stages, each
// exchanging the off-diagonal j x j blocks between lanes p and p ^ j.
__device__ __forceinline__ uint64_t transpose64(uint64_t x) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  constexpr uint64_t keep[6] = {0x00000000FFFFFFFFull, 0x0000FFFF0000FFFFull, 0x00FF00FF00FF00FFull,
                                0x0F0F0F0F0F0F0F0Full, 0x3333333333333333ull, 0x5555555555555555ull};
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const uint32_t j = 32u >> t;
    const uint64_t L = keep[t];
    const uint64_t y = static_cast<uint64_t>(__shfl_xor(static_cast<unsigned long long>(x), static_cast<int>(j)));
    x = (lane & j) ? ((x & ~L) | ((y & ~L) >> j)) : ((x & L) | ((y & L) << j));
  }
  return x;
}

// One wavefront per molecule: lane p is position p (and p+64 for molecules of 65..128 nodes).
// planes[(m*R + r)*64 + i]: bit p = bit i of the level-(r+1) signature of position p, i.e.
// plane 4*label + bit; planes_hi the same for positions 64..127; lab_hi[w*16+l]: positions
// 64..127 with label l (positions 0..63 use the graph's label masks); dmax[m*R + r]: the
// nibble-wise max of the level-(r+1) signatures over the molecule's positions.
__global__ __launch_bounds__(256) void k_db_prepare(GraphDev g, uint32_t R, const uint32_t* __restrict__ wide_idx,
                                                    uint64_t* __restrict__ planes, uint64_t* __restrict__ planes_hi,
                                                    uint64_t* __restrict__ lab_hi, uint64_t* __restrict__ dmax) {
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
      const uint64_t col = transpose64(sig);
      if (h == 0) planes[(m * R + r) * 64 + lane] = col;
      else planes_hi[(static_cast<uint64_t>(wi) * R + r) * 64 + lane] = col;
      const uint64_t mx = wave_nib_max(sig);
      if (lane == 0) dmax[m * R + r] = h == 0 ? mx : nib_max(dmax[m * R + r], mx);
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

// One wavefront per query graph, lane i = node i (query graphs have <= 32 nodes, so
// positions never alias): writes each node's token segment, the graph's label set (labels
// < 16) and, per level, the nibble-wise max of its nodes' signatures.
__global__ __launch_bounds__(256) void k_query_prepare(GraphDev q, uint32_t R, uint16_t* __restrict__ tokens,
                                                       uint32_t* __restrict__ qlab, uint64_t* __restrict__ qmax) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint64_t g = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (g >= q.G) return;  // wave-uniform
  const uint32_t prev = q.go[g], k = q.go[g + 1] - prev;
  const bool valid = lane < k;
  const uint32_t seg = token_segment(R);
  uint16_t* t = tokens + (static_cast<uint64_t>(prev) + lane) * seg;
  const uint32_t label = valid ? q.nl[prev + lane] : 0u;
  if (valid) t[0] = static_cast<uint16_t>(label);
  uint64_t self = 1ull << (lane & 63), ball = self, frontier = self;
  uint32_t cnt = 1;
  for (uint32_t r = 0; r < R; ++r) {
    uint64_t sig = 0;
    if (valid) {
      sig = histogram16(next_layer(q, prev, ball, frontier), q.lmask + g * 16);
      for (uint32_t l = 0; l < 16; ++l) {
        const uint32_t c = static_cast<uint32_t>(sig >> (4 * l)) & 15u;
        if (c) t[cnt++] = static_cast<uint16_t>(r | (l << 3) | (c << 7));
      }
    }
    const uint64_t mx = wave_nib_max(sig);
    if (lane == 0) qmax[g * R + r] = mx;
  }
  if (valid) t[cnt] = kEndToken;
  uint32_t bits = valid && label < 16 ? 1u << label : 0u;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) bits |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(bits), o));
  if (lane == 0) qlab[g] = bits;
}

struct PairArgs {
  GraphDev q, d;
  uint32_t R, has_wide, mol_chunk;
  const uint64_t* planes;
  const uint64_t* planes_hi;
  const uint64_t* lab_hi;
  const uint32_t* wide_idx;
  const uint64_t* dmax;
  const uint64_t* tok64;  // token segments as 64-bit words
  const uint32_t* qlab;
  const uint64_t* qmax;
  uint64_t* scratch;      // per wavefront: masks [32][64] (low words), then [32][64] (high words)
  uint32_t* slot;         // per wavefront: arena offset of each surviving query graph [Gq]
  uint64_t* arena;        // survivor masks: k words (2k, interleaved, for molecules of > 64 nodes)
  uint64_t arena_cap;
  uint32_t* rec_qg;       // records: per molecule a contiguous block in query-graph order
  uint32_t* rec_off;
  uint64_t rec_cap;
  uint32_t* mol_block;    // [Gd] first record of the molecule
  uint32_t* mol_count;    // [Gd]
  unsigned long long* counters;  // [0] molecule cursor, [1] arena top, [2] record top, [3] overflow flags
};

__host__ __device__ inline uint64_t pair_lds_words(uint32_t R, uint32_t has_wide, uint32_t Gq) {
  return static_cast<uint64_t>(R) * 64 * (has_wide ? 2 : 1) + 32 + (Gq + 63) / 64 + 32;
}

// Persistent pair filter.  A wavefront takes molecules from a cursor and stages the molecule's
// level planes and label masks in LDS.  Query graphs pass two necessary tests first, 64 at a
// time: label set (labels < 16) contained in the molecule's, and per level the graph's
// nibble-wise signature max <= the molecule's (a node fits nowhere if it exceeds the max).
// Passing graphs are handed to idle lanes; a lane runs its pair as a small state machine over
// the query's token segments (label test, one requirement per step, node end), so lanes whose
// pair died take the next graph at once, and the molecule-wide early exit is the GMCR predicate
// of gmcr.hpp:56-61.  Survivors' masks go to the arena, per molecule one block of records in
// query-graph order.  Arena and record space is claimed per wavefront in chunks.
__global__ __launch_bounds__(256) void k_pair_filter(PairArgs a) {
  extern __shared__ uint64_t lds[];
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t Gq = a.q.G;
  const uint32_t words = (Gq + 63) / 64;
  const uint32_t R = a.R;
  uint64_t* lp = lds + wave * pair_lds_words(R, a.has_wide, Gq);
  uint64_t* lph = lp + R * 64;
  uint64_t* llm = lph + (a.has_wide ? R * 64 : 0);
  uint64_t* llh = llm + 16;
  uint64_t* surv = llh + 16;
  uint32_t* xfer = reinterpret_cast<uint32_t*>(surv + words);
  const uint64_t gwave = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + wave;
  uint64_t* sc_lo = a.scratch + gwave * (a.has_wide ? 2 * 32 * 64 : 32 * 64);
  uint64_t* sc_hi = sc_lo + 32 * 64;
  uint32_t* slot = a.slot + gwave * Gq;
  const uint32_t seg = token_segment(R);
  const uint64_t below = (1ull << lane) - 1;
  uint64_t ck_next = 0, ck_end = 0, rk_next = 0, rk_end = 0;  // claimed arena / record chunks

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
    {
      const uint64_t* src = a.planes + m * R * 64;
      for (uint32_t t = lane; t < R * 64; t += kWave) lp[t] = src[t];
      if (wide) {
        const uint64_t* srch = a.planes_hi + static_cast<uint64_t>(wi) * R * 64;
        for (uint32_t t = lane; t < R * 64; t += kWave) lph[t] = srch[t];
      }
    }
    uint64_t lmv = 0, lhv = 0;
    if (lane < 16) {
      lmv = a.d.lmask[m * 16 + lane];
      lhv = wide ? a.lab_hi[static_cast<uint64_t>(wi) * 16 + lane] : 0ull;
      llm[lane] = lmv;
      llh[lane] = lhv;
    }
    const uint32_t mlab = static_cast<uint32_t>(__ballot((lmv | lhv) != 0)) & 0xFFFFu;
    for (uint32_t w = lane; w < words; w += kWave) surv[w] = 0;
    const uint64_t* mmax = a.dmax + m * R;
    const uint64_t all_lo = n >= 64 ? ~0ull : ((1ull << n) - 1);
    const uint64_t all_hi = n <= 64 ? 0ull : (n >= 128 ? ~0ull : ((1ull << (n - 64)) - 1));
    wave_sync();

    bool have = false, fresh = false;
    uint32_t qg = 0, i = 0, k = 0, tp = 0;
    uint64_t nodebase = 0, tw = 0, tn = 0, pf0 = 0, pf1 = 0, lo = 0, hi = 0;
    uint32_t scan = 0, pend_base = 0;  // wave-uniform
    uint64_t pend = 0;                 // wave-uniform: graphs pend_base + j that passed the tests
    while (true) {
      // hand passing query graphs to idle lanes
      uint64_t need = __ballot(!have);
      while (need) {
        if (!pend) {
          if (scan >= Gq) break;
          const uint32_t g = scan + lane;
          bool ok = g < Gq;
          if (ok) ok = (a.qlab[g] & ~mlab) == 0u;
          for (uint32_t r = 0; r < R && ok; ++r) ok = nib_le_all(a.qmax[static_cast<uint64_t>(g) * R + r], mmax[r]);
          pend = __ballot(ok);
          pend_base = scan;
          scan += kWave;
          continue;
        }
        const uint32_t cnt = min(__popcll(pend), __popcll(need));
        const bool give = ((pend >> lane) & 1ull) && static_cast<uint32_t>(__popcll(pend & below)) < cnt;
        const uint32_t rn = __popcll(need & below);
        const bool take = ((need >> lane) & 1ull) && rn < cnt;
        if (give) xfer[__popcll(pend & below)] = pend_base + lane;
        wave_sync();
        if (take) {
          qg = xfer[rn];
          const uint32_t qoff = a.q.go[qg];
          k = a.q.go[qg + 1] - qoff;
          i = 0;
          nodebase = static_cast<uint64_t>(qoff) * seg;
          tw = a.tok64[nodebase / 4];
          tn = a.tok64[nodebase / 4 + 1];
          if (k > 1) {
            pf0 = a.tok64[(nodebase + seg) / 4];
            pf1 = a.tok64[(nodebase + seg) / 4 + 1];
          }
          have = true;
          fresh = true;
        }
        wave_sync();
        pend &= ~__ballot(give);
        need &= ~__ballot(take);
      }
      if (__ballot(have) == 0) break;
      bool done = false;
      if (have) {
        if (fresh) {  // node start: the filter's label test (isomorphism.hpp:49)
          const uint32_t ql = static_cast<uint32_t>(tw & 0xFFFFu);
          if (ql == 255u) {
            lo = all_lo;
            hi = all_hi;
          } else if (ql < 16u) {
            lo = llm[ql];
            hi = llh[ql];
          } else {
            lo = hi = 0;
            for (uint32_t p = 0; p < n; ++p)
              if (a.d.nl[s + p] == ql) {
                if (p < 64) lo |= 1ull << p;
                else hi |= 1ull << (p - 64);
              }
          }
          tp = 1;
          fresh = false;
          if (!(lo | hi)) have = false;
        } else {
          const uint32_t tok = static_cast<uint32_t>(tw >> (16 * (tp & 3))) & 0xFFFFu;
          if (tok == kEndToken) {  // node done: keep its mask
            sc_lo[i * 64 + lane] = lo;
            if (wide) sc_hi[i * 64 + lane] = hi;
            ++i;
            if (i == k) {  // every query node has a candidate: the pair survives GMCR
              done = true;
              have = false;
            } else {
              nodebase += seg;
              tw = pf0;
              tn = pf1;
              tp = 0;
              fresh = true;
              if (i + 1 < k) {
                pf0 = a.tok64[(nodebase + seg) / 4];
                pf1 = a.tok64[(nodebase + seg) / 4 + 1];
              }
            }
          } else {  // requirement: count_q(l) <= count_p(l) at level r (isomorphism.hpp:103-110)
            const uint32_t r = tok & 7u, l = (tok >> 3) & 15u, c = (tok >> 7) & 15u;
            const uint64_t* pl = lp + r * 64 + l * 4;
            lo &= ge4_lane(pl[0], pl[1], pl[2], pl[3], c);
            if (wide) {
              const uint64_t* ph = lph + r * 64 + l * 4;
              hi &= ge4_lane(ph[0], ph[1], ph[2], ph[3], c);
            }
            if (!(lo | hi)) have = false;
            ++tp;
            if ((tp & 3) == 0) {
              tw = tn;
              tn = a.tok64[(nodebase + tp) / 4 + 1];
            }
          }
        }
      }
      // survivors of this step claim arena space from the wavefront's chunk
      if (__ballot(done)) {
        const uint32_t stride = wide ? 2 : 1;
        const uint32_t want = done ? k * stride : 0u;
        uint32_t incl = want;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
          const uint32_t y = __shfl_up(incl, o);
          if (lane >= static_cast<uint32_t>(o)) incl += y;
        }
        const uint32_t total = uniform32(__shfl(incl, kWave - 1));
        if (ck_next + total > ck_end) {
          const uint64_t grab = total > kArenaChunk ? total : kArenaChunk;
          unsigned long long b = 0;
          if (lane == 0) b = atomicAdd(&a.counters[1], static_cast<unsigned long long>(grab));
          ck_next = uniform64(__shfl(b, 0));
          ck_end = ck_next + grab;
        }
        const uint64_t off = ck_next + incl - want;
        ck_next += total;
        if (done) {
          if (off + want <= a.arena_cap) {
            for (uint32_t t = 0; t < k; ++t) {
              a.arena[off + t * stride] = sc_lo[t * 64 + lane];
              if (wide) a.arena[off + t * stride + 1] = sc_hi[t * 64 + lane];
            }
          } else {
            atomicOr(&a.counters[3], 1ull);
          }
          slot[qg] = static_cast<uint32_t>(off);
          atomicOr(reinterpret_cast<unsigned long long*>(&surv[qg >> 6]), 1ull << (qg & 63));
        }
      }
    }
    wave_sync();
    // emit this molecule's survivors as one block of records, ascending query graph
    uint32_t mine = 0;
    for (uint32_t w = lane; w < words; w += kWave) mine += __popcll(surv[w]);
    uint32_t total = mine;
    for (int o = kWave / 2; o > 0; o >>= 1) total += __shfl_xor(total, o);
    total = uniform32(total);
    uint64_t blk = 0;
    if (total) {
      if (rk_next + total > rk_end) {
        const uint64_t grab = total > kRecordChunk ? total : kRecordChunk;
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(&a.counters[2], static_cast<unsigned long long>(grab));
        rk_next = uniform64(__shfl(b, 0));
        rk_end = rk_next + grab;
      }
      blk = rk_next;
      rk_next += total;
    }
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
        const uint32_t c = __popcll(x);
        uint32_t incl = c;
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
    wave_sync();
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
                      uint64_t* lab_hi, uint64_t* dmax) {
  if (!g.G) return hipSuccess;
  k_db_prepare<<<(g.G + 3) / 4, 256, 0, s>>>(g, R, wide_idx, planes, planes_hi, lab_hi, dmax);
  return hipGetLastError();
}

hipError_t query_prepare(hipStream_t s, const GraphDev& q, uint32_t R, uint16_t* tokens, uint32_t* qlab, uint64_t* qmax) {
  if (!q.G) return hipSuccess;
  k_query_prepare<<<(q.G + 3) / 4, 256, 0, s>>>(q, R, tokens, qlab, qmax);
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
uint64_t graph_edges(const sigmo_graph* g);
}  // namespace sigmo_internal

using namespace sigmo_internal;

struct sigmo_db {
  sigmo_ctx* ctx = nullptr;
  const sigmo_graph* data = nullptr;
  uint32_t R = 0;
  uint32_t num_wide = 0;
  uint32_t* wide_idx = nullptr;
  uint64_t *planes = nullptr, *planes_hi = nullptr, *lab_hi = nullptr, *dmax = nullptr;
  uint64_t bytes = 0;
  ~sigmo_db() {
    bfree(wide_idx);
    bfree(planes);
    bfree(planes_hi);
    bfree(lab_hi);
    bfree(dmax);
  }
};

struct sigmo_batch {
  sigmo_ctx* ctx = nullptr;
  const sigmo_db* db = nullptr;
  // query side
  uint16_t* tokens = nullptr;
  uint32_t* qlab = nullptr;
  uint64_t* qmax = nullptr;
  uint64_t cap_tokens = 0, cap_qlab = 0, cap_qmax = 0;
  // pair filter
  uint64_t* scratch = nullptr;
  uint64_t cap_scratch = 0;
  uint32_t* slot = nullptr;
  uint64_t cap_slot = 0;
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
    bfree(tokens);
    bfree(qlab);
    bfree(qmax);
    bfree(scratch);
    bfree(slot);
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
  if (stt == SIGMO_OK) stt = balloc(&db->dmax, uint64_t(G) * R);
  if (stt != SIGMO_OK) {
    delete db;
    return stt;
  }
  db->bytes = G * 4ull + (uint64_t(G) + nw) * R * 512 + nw * 128ull + uint64_t(G) * R * 8;
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
  // algorithmic bytes: the CSR and labels read once, R level signatures per data node written:
  // 4(D+1) + 4E + D + 8 R D
  const double bytes = 4.0 * (double(g.N) + 1.0) + 4.0 * double(graph_edges(db->data)) + double(g.N) + 8.0 * db->R * double(g.N);
  void* tok = nullptr;
  BATCH_TRY(ctx_timed(db->ctx, SIGMO_PHASE_SIG_DATA, SIGMO_KERNEL_DB_PREPARE, bytes, true, &tok));
  BATCH_HIP_TRY(sigmo_k::db_prepare(s, g, db->R, db->wide_idx, db->planes, db->planes_hi, db->lab_hi, db->dmax));
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
  const uint32_t R = db->R;
  const uint32_t has_wide = db->num_wide ? 1u : 0u;
  const size_t lds_block = sigmo_k::pair_lds_words(R, has_wide, Gq) * 8 * 4;
  if (lds_block > 64 * 1024) return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "too many query graphs for one batched run");
  // persistent pair-filter grid: as many 4-wave blocks as registers and LDS keep resident
  int resident = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, sigmo_k::k_pair_filter, 256, lds_block) != hipSuccess) {
    (void)hipGetLastError();
    resident = 4;
  }
  const uint64_t per_cu = std::max<uint64_t>(1, std::min<uint64_t>(8, static_cast<uint64_t>(resident)));
  const uint64_t blocks = std::max<uint64_t>(1, std::min<uint64_t>(uint64_t(cus) * per_cu, (uint64_t(Gd) + 3) / 4));
  const uint64_t waves = blocks * 4;
  const uint64_t seg = sigmo_k::token_segment(R);
  BATCH_TRY(bgrow(&b->tokens, b->cap_tokens, qd.N * seg));
  BATCH_TRY(bgrow(&b->qlab, b->cap_qlab, Gq));
  BATCH_TRY(bgrow(&b->qmax, b->cap_qmax, uint64_t(Gq) * R));
  BATCH_TRY(bgrow(&b->scratch, b->cap_scratch, waves * 32 * 64 * (has_wide ? 2 : 1)));
  BATCH_TRY(bgrow(&b->slot, b->cap_slot, waves * Gq));
  BATCH_TRY(bgrow(&b->mol_block, b->cap_mb, Gd));
  BATCH_TRY(bgrow(&b->mol_count, b->cap_mc, uint64_t(Gd) + 1));
  BATCH_TRY(bgrow(&b->offsets, b->cap_off, uint64_t(Gd) + 1));
  BATCH_TRY(bgrow(&b->scan_tmp, b->cap_scan, (uint64_t(Gd) + 1023) / 1024 + 1));
  {
    void* tok = nullptr;
    BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_SIG_QUERY, -1, 0.0, true, &tok));
    BATCH_HIP_TRY(sigmo_k::query_prepare(s, qd, R, b->tokens, b->qlab, b->qmax));
    BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
  }
  // record / arena capacity: start from 1/8 of the pair space (at most 2^26 pairs) plus every
  // wavefront's open chunk.  On overflow the counters still hold the exact demand, so one replay
  // with that size suffices; capacities are kept for later runs.
  const double mean_k = Gq ? double(ksum) / Gq : 0.0;
  const uint64_t slack_rec = waves * sigmo_k::kRecordChunk, slack_arena = waves * sigmo_k::kArenaChunk;
  const uint64_t guess = std::min<uint64_t>(std::max<uint64_t>(uint64_t(Gq) * Gd / 8, 1 << 16), 1ull << 26);
  uint64_t want_rec = guess + slack_rec;
  uint64_t want_arena = static_cast<uint64_t>(double(guess) * (mean_k + 1.0) * (has_wide ? 2.0 : 1.0)) + slack_arena;
  want_rec = std::max(want_rec, std::min(b->cap_rec, b->cap_rec2));
  want_arena = std::max(want_arena, b->cap_arena);
  for (int attempt = 0; attempt < 6; ++attempt) {
    BATCH_TRY(bgrow(&b->rec_qg, b->cap_rec, want_rec));
    BATCH_TRY(bgrow(&b->rec_off, b->cap_rec2, want_rec));
    BATCH_TRY(bgrow(&b->arena, b->cap_arena, want_arena));
    if (b->cap_arena >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "candidate arena exceeds 2^32 words");

    // algorithmic reads: the data side once per run -- a label byte and R level signatures per
    // data node, (1 + 8R) D -- plus the query side; writes are added once they are known.
    const double reads = (1.0 + 8.0 * R) * double(dd.N) + double(qd.N) * (1.0 + 2.0 * 16.0 * R) + double(Gq) * (4.0 + 8.0 * R);
    void* tok = nullptr;
    BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_FILTER, SIGMO_KERNEL_PAIR_FILTER, reads, true, &tok));
    BATCH_HIP_TRY(sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(b->counters), 8, 0));
    sigmo_k::PairArgs a;
    a.q = qd;
    a.d = dd;
    a.R = R;
    a.has_wide = has_wide;
    a.mol_chunk = 4;
    a.planes = db->planes;
    a.planes_hi = db->planes_hi;
    a.lab_hi = db->lab_hi;
    a.wide_idx = db->wide_idx;
    a.dmax = db->dmax;
    a.tok64 = reinterpret_cast<const uint64_t*>(b->tokens);
    a.qlab = b->qlab;
    a.qmax = b->qmax;
    a.scratch = b->scratch;
    a.slot = b->slot;
    a.arena = b->arena;
    a.arena_cap = b->cap_arena;
    a.rec_qg = b->rec_qg;
    a.rec_off = b->rec_off;
    a.rec_cap = std::min(b->cap_rec, b->cap_rec2);
    a.mol_block = b->mol_block;
    a.mol_count = b->mol_count;
    a.counters = b->counters;
    hipLaunchKernelGGL(sigmo_k::k_pair_filter, dim3(static_cast<unsigned>(blocks)), dim3(256), lds_block, s, a);
    BATCH_HIP_TRY(hipGetLastError());
    BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
    unsigned long long cnt[4];
    BATCH_HIP_TRY(hipMemcpyAsync(cnt, b->counters, sizeof(cnt), hipMemcpyDeviceToHost, s));
    BATCH_HIP_TRY(hipStreamSynchronize(s));
    // writes: candidate masks, (query graph, arena offset) records, per-molecule block and count
    ctx_add_bytes(b->ctx, SIGMO_KERNEL_PAIR_FILTER, double(cnt[1]) * 8.0 + double(cnt[2]) * 8.0 + double(Gd) * 8.0);
    if (!cnt[3]) break;
    if (attempt == 5) return sigmo_err::set(SIGMO_ERR_NOMEM, "pair buffers overflowed repeatedly");
    // replay with larger buffers
    if (cnt[3] & 2ull) want_rec = std::max<uint64_t>(want_rec * 2, cnt[2] + slack_rec);
    if (cnt[3] & 1ull) want_arena = std::max<uint64_t>(want_arena * 2, cnt[1] + slack_arena);
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
