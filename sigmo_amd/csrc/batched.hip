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
#include <type_traits>
#include <chrono>
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/iterator/counting_iterator.hpp>

#include "../../include/sigmo_hip.h"
#include "device_common.hpp"
#include "host_graphs.hpp"
#include "join_lanes.hpp"
#include "kernels.hpp"

namespace sigmo_k {
namespace {

constexpr uint32_t kMaxLevels = 8;        // token encoding keeps 3 bits of level
constexpr uint32_t kNoWide = 0xFFFFFFFFu;
constexpr uint16_t kEndToken = 0xFFFF;    // closes a query node's token segment

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
    next |= g.fold[u];
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


// Per-molecule facts the pair pre-test reads, stored in width-class order.
struct MolInfo {
  uint32_t m, s, n, x, wi, mlab;  // molecule, first node, nodes, plane index, wide index, labels < 16 present
};

// Per-query-graph facts the pair pre-test reads.
struct QueryInfo {
  uint32_t qoff, k, qlab, qcls;  // first node, nodes, labels < 16 present, signature classes present (bit c)
  uint64_t first;                // first token word of the graph's first node
};

// The database's signature classes: up to three labels (< 16), the database's most frequent, for
// which the pre-test keeps per-label maxima (l[c] = 0xFFFFFFFF: class c unused).
constexpr uint32_t kClasses = 3;
struct ClassLabels {
  uint32_t l[kClasses];
};

// Nibble-wise max over the 64 positions of a level signature, from its transposed planes (lane i
// holds plane i = bit i%4 of the count of label i/4): lane 4l reads the label's other three
// planes and takes the largest 4-bit value greedily from the top bit; four ballots then place
// bit b of label l's max at bit 4l + b.
__device__ __forceinline__ uint64_t planes_nib_max(uint64_t col) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint64_t p1 = static_cast<uint64_t>(__shfl_down(static_cast<unsigned long long>(col), 1));
  const uint64_t p2 = static_cast<uint64_t>(__shfl_down(static_cast<unsigned long long>(col), 2));
  const uint64_t p3 = static_cast<uint64_t>(__shfl_down(static_cast<unsigned long long>(col), 3));
  const uint64_t planes[4] = {col, p1, p2, p3};
  uint64_t cand = ~0ull;
  uint32_t v = 0;
#pragma unroll
  for (int b = 3; b >= 0; --b) {
    const uint64_t t = cand & planes[b];
    if (t) {
      v |= 1u << b;
      cand = t;
    }
  }
  const bool leader = (lane & 3u) == 0;
  uint64_t mx = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) mx |= __ballot(leader && ((v >> b) & 1u)) << b;
  return mx;
}

// One wavefront per molecule: lane p is position p (and p+64 for molecules of 65..128 nodes).
// The sort key of the width-class order: class (<= 32, 33..64, 65..128 nodes) << 23 | the
// molecule's label set (labels < 16) << 7 | (nodes - 1).  Sorted, molecules of one label set and
// similar size share wavefronts, so the pre-test's label test (and, mostly, its maxima tests)
// passes or fails for whole wavefronts and the tests after a failing one are skipped.
__global__ __launch_bounds__(256) void k_db_order_keys(uint32_t G, const uint32_t* __restrict__ go, const uint8_t* __restrict__ nl,
                                                       uint32_t* __restrict__ keys) {
  const uint64_t m = blockIdx.x * 256ull + threadIdx.x;
  if (m >= G) return;
  const uint32_t s = go[m], n = go[m + 1] - s;
  uint32_t labels = 0;
  for (uint32_t p = 0; p < n; ++p) {
    const uint32_t l = nl[s + p];
    if (l < 16) labels |= 1u << l;
  }
  const uint32_t cls = n <= 32 ? 0u : (n <= 64 ? 1u : 2u);
  keys[m] = (cls << 23) | (labels << 7) | ((n - 1) & 127u);
}

// Level planes, bit p = bit i of the level-(r+1) signature of position p for plane i =
// 4*label + bit, go to one of two arrays by the molecule's width class: planes32[(x*R + r)*64
// + i] (u32) for molecules of <= 32 nodes, planes64[(x*R + r)*64 + i] (u64) otherwise, x =
// pidx[m]; planes_hi holds positions 64..127 and lab_hi[w*16+l] the positions 64..127 with
// label l (w = wide_idx[m]; positions 0..63 use the graph's label masks).  Wavefront c handles
// molecule m = cls_list[c] (width-class order, grouped by k_db_order_keys) and writes minfo[c], dmax[c*R + r], the
// nibble-wise max of the level-(r+1) signatures over the molecule's positions, and
// cmax[(k*R + r)*G + c], the same max over the positions labelled cl.l[k] (0 if none).
__global__ __launch_bounds__(256) void k_db_prepare(GraphDev g, uint32_t R, const uint32_t* __restrict__ cls_list,
                                                    const uint32_t* __restrict__ pidx, const uint32_t* __restrict__ wide_idx,
                                                    uint32_t* __restrict__ planes32, uint64_t* __restrict__ planes64,
                                                    uint64_t* __restrict__ planes_hi, uint64_t* __restrict__ lab_hi,
                                                    uint64_t* __restrict__ dmax, MolInfo* __restrict__ minfo, ClassLabels cl,
                                                    uint64_t* __restrict__ cmax, uint32_t c0) {
  __shared__ uint64_t fold_all[4][kWave];
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint64_t c = c0 + static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (c >= g.G) return;  // wave-uniform
  const uint64_t m = cls_list[c];
  const uint32_t s = g.go[m], n = g.go[m + 1] - s;
  const uint64_t* lm = g.lmask + m * 16;
  // the label masks of positions 0..63 and the labels that occur there (wave-uniform): the
  // histogram only visits those labels
  uint64_t lmv[16];
  uint32_t present = 0;
#pragma unroll
  for (int l = 0; l < 16; ++l) {
    lmv[l] = lm[l];
    if (lmv[l]) present |= 1u << l;
  }
  const uint32_t halves = n > 64 ? 2 : 1;
  const uint64_t x = pidx[m];
  const uint32_t wi = wide_idx[m];
  // Only positions 0..63 are ever expanded (frontier bit b expands node s + b), each by its
  // neighbours folded mod 64: stage those 64 words in LDS once.
  uint64_t* fold = fold_all[threadIdx.x / kWave];
  fold[lane] = lane < n ? g.fold[static_cast<uint64_t>(s) + lane] : 0ull;
  wave_sync();
  uint32_t labels = 0;
  for (uint32_t h = 0; h < halves; ++h) {
    const uint32_t pos = lane + 64 * h;
    const bool valid = pos < n;
    const uint32_t mylab = valid ? g.nl[s + pos] : 0x100u;
    if (mylab < 16) labels |= 1u << mylab;
    uint64_t cpos[kClasses];  // this half's positions of each class label
#pragma unroll
    for (uint32_t k = 0; k < kClasses; ++k) cpos[k] = __ballot(mylab == cl.l[k]);
    uint64_t self = 1ull << (pos & 63), ball = self, frontier = self;
    for (uint32_t r = 0; r < R; ++r) {
      uint64_t sig = 0;
      if (valid) {
        uint64_t next = 0;
        for (uint64_t f = frontier; f; f &= f - 1) next |= fold[__builtin_ctzll(f)];
        next &= ~ball;
        ball |= next;
        frontier = next;
#pragma unroll
        for (int l = 0; l < 16; ++l)
          if ((present >> l) & 1u) {  // counts saturate at 15 (signature.hpp:44-51)
            uint32_t cnt = __popcll(next & lmv[l]);
            cnt = cnt < 15 ? cnt : 15;
            sig |= static_cast<uint64_t>(cnt) << (4 * l);
          }
      }
      const uint64_t col = transpose64(sig);
      if (h == 1) planes_hi[(static_cast<uint64_t>(wi) * R + r) * 64 + lane] = col;
      else if (n <= 32) planes32[(x * R + r) * 64 + lane] = static_cast<uint32_t>(col);
      else planes64[(x * R + r) * 64 + lane] = col;
      const uint64_t mx = planes_nib_max(col);
      if (lane == 0) dmax[c * R + r] = h == 0 ? mx : nib_max(dmax[c * R + r], mx);
#pragma unroll
      for (uint32_t k = 0; k < kClasses; ++k) {
        const uint64_t cm = planes_nib_max(col & cpos[k]);
        uint64_t* dst = cmax + (static_cast<uint64_t>(k) * R + r) * g.G + c;
        if (lane == 0) *dst = h == 0 ? cm : nib_max(*dst, cm);
      }
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
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) labels |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(labels), o));
  if (lane == 0) {
    MolInfo info;
    info.m = static_cast<uint32_t>(m);
    info.s = s;
    info.n = n;
    info.x = static_cast<uint32_t>(x);
    info.wi = wi;
    info.mlab = labels;
    minfo[c] = info;
  }
}

// planes_nib_max of the two 32-position halves of a plane column at once (positions 0..31 and
// 32..63 of `col`): the nibble-wise maxima over each half's positions.
__device__ __forceinline__ void planes_nib_max2(uint64_t col, uint64_t& lo, uint64_t& hi) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint64_t p1 = static_cast<uint64_t>(__shfl_down(static_cast<unsigned long long>(col), 1));
  const uint64_t p2 = static_cast<uint64_t>(__shfl_down(static_cast<unsigned long long>(col), 2));
  const uint64_t p3 = static_cast<uint64_t>(__shfl_down(static_cast<unsigned long long>(col), 3));
  const uint64_t planes[4] = {col, p1, p2, p3};
  uint32_t ca = ~0u, cb = ~0u, va = 0, vb = 0;
#pragma unroll
  for (int b = 3; b >= 0; --b) {
    const uint32_t ta = ca & static_cast<uint32_t>(planes[b]), tb = cb & static_cast<uint32_t>(planes[b] >> 32);
    if (ta) {
      va |= 1u << b;
      ca = ta;
    }
    if (tb) {
      vb |= 1u << b;
      cb = tb;
    }
  }
  const bool leader = (lane & 3u) == 0;
  lo = hi = 0;
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    lo |= __ballot(leader && ((va >> b) & 1u)) << b;
    hi |= __ballot(leader && ((vb >> b) & 1u)) << b;
  }
}

// k_db_prepare for molecules of <= 32 nodes, two per wavefront: lanes 0..31 hold the positions of
// width-class entry 2v, lanes 32..63 those of 2v+1 (none past `count`).  Each lane's BFS runs over
// its own molecule's positions; one 64 x 64 transpose gives both molecules' level planes (low and
// high 32 bits of each plane word), so the transposes, and the lanes, serve two molecules.
__global__ __launch_bounds__(256) void k_db_prepare_pairs(GraphDev g, uint32_t R, uint32_t count, const uint32_t* __restrict__ cls_list,
                                                          const uint32_t* __restrict__ pidx, uint32_t* __restrict__ planes32,
                                                          uint64_t* __restrict__ dmax, MolInfo* __restrict__ minfo, ClassLabels cl,
                                                          uint64_t* __restrict__ cmax) {
  __shared__ uint64_t fold_all[4][kWave];
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint64_t v = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  const uint64_t cA = 2 * v, cB = cA + 1;
  if (cA >= count) return;  // wave-uniform
  const bool hasB = cB < count;
  const uint32_t half = lane >> 5, p = lane & 31u;
  const uint64_t c = half ? cB : cA;
  const bool live = !half || hasB;
  const uint64_t m = live ? cls_list[c] : 0;
  const uint32_t s = live ? g.go[m] : 0, n = live ? g.go[m + 1] - s : 0;
  const uint64_t* lm = g.lmask + m * 16;
  uint64_t lmv[16];
  uint32_t present = 0;
#pragma unroll
  for (int l = 0; l < 16; ++l) {
    lmv[l] = live ? lm[l] : 0ull;
    if (lmv[l]) present |= 1u << l;
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) present |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(present), o));
  uint64_t* fold = fold_all[threadIdx.x / kWave];
  fold[lane] = p < n ? g.fold[static_cast<uint64_t>(s) + p] : 0ull;
  wave_sync();
  const uint64_t* myfold = fold + 32 * half;
  const bool valid = p < n;
  const uint32_t mylab = valid ? g.nl[s + p] : 0x100u;
  uint32_t labels = mylab < 16 ? 1u << mylab : 0u;
  uint64_t cpos[kClasses];
#pragma unroll
  for (uint32_t k = 0; k < kClasses; ++k) cpos[k] = __ballot(mylab == cl.l[k]);
  const uint64_t xA = pidx[cls_list[cA]], xB = hasB ? pidx[cls_list[cB]] : 0;
  uint64_t self = 1ull << p, ball = self, frontier = self;
  for (uint32_t r = 0; r < R; ++r) {
    uint64_t sig = 0;
    if (valid) {
      uint64_t next = 0;
      for (uint64_t f = frontier; f; f &= f - 1) next |= myfold[__builtin_ctzll(f)];
      next &= ~ball;
      ball |= next;
      frontier = next;
#pragma unroll
      for (int l = 0; l < 16; ++l)
        if ((present >> l) & 1u) {  // counts saturate at 15 (signature.hpp:44-51)
          uint32_t cnt = __popcll(next & lmv[l]);
          cnt = cnt < 15 ? cnt : 15;
          sig |= static_cast<uint64_t>(cnt) << (4 * l);
        }
    }
    const uint64_t col = transpose64(sig);
    planes32[(xA * R + r) * 64 + lane] = static_cast<uint32_t>(col);
    if (hasB) planes32[(xB * R + r) * 64 + lane] = static_cast<uint32_t>(col >> 32);
    uint64_t ma, mb;
    planes_nib_max2(col, ma, mb);
    if (lane == 0) dmax[cA * R + r] = ma;
    if (lane == 1 && hasB) dmax[cB * R + r] = mb;
#pragma unroll
    for (uint32_t k = 0; k < kClasses; ++k) {
      planes_nib_max2(col & cpos[k], ma, mb);
      if (lane == 0) cmax[(static_cast<uint64_t>(k) * R + r) * g.G + cA] = ma;
      if (lane == 1 && hasB) cmax[(static_cast<uint64_t>(k) * R + r) * g.G + cB] = mb;
    }
  }
#pragma unroll
  for (int o = 16; o > 0; o >>= 1) labels |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(labels), o));  // within each half
  if (p == 0 && live) {
    MolInfo info;
    info.m = static_cast<uint32_t>(m);
    info.s = s;
    info.n = n;
    info.x = static_cast<uint32_t>(half ? xB : xA);
    info.wi = kNoWide;
    info.mlab = labels;
    minfo[c] = info;
  }
}

// One wavefront per query graph, lane i = node i (query graphs have <= 32 nodes, so
// positions never alias): writes each node's tokens, the graph's label set (labels < 16) and,
// per level, the nibble-wise max of its nodes' signatures.  The graph's nodes' token streams
// are packed one after another from the start of its region (token qoff * segment), each padded
// to a whole token word (4 tokens), so a pair walks one contiguous run of words: the next node
// starts at the word after the one holding the current node's end token.
__global__ __launch_bounds__(256) void k_query_prepare(GraphDev q, uint32_t R, uint16_t* __restrict__ tokens,
                                                       QueryInfo* __restrict__ qinfo, uint64_t* __restrict__ qmax) {
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint64_t g = static_cast<uint64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave;
  if (g >= q.G) return;  // wave-uniform
  const uint32_t prev = q.go[g], k = q.go[g + 1] - prev;
  const bool valid = lane < k;
  const uint32_t seg = token_segment(R);
  const uint32_t label = valid ? q.nl[prev + lane] : 0u;
  // the node's level signatures, and its token count: label + one per nonzero nibble + end
  uint64_t sig[kMaxLevels];
  uint64_t self = 1ull << (lane & 63), ball = self, frontier = self;
  uint32_t ntok = 2;
#pragma unroll
  for (uint32_t r = 0; r < kMaxLevels; ++r) {
    sig[r] = 0;
    if (r < R) {
      if (valid) {
        sig[r] = histogram16(next_layer(q, prev, ball, frontier), q.lmask + g * 16);
        for (uint32_t l = 0; l < 16; ++l) ntok += ((sig[r] >> (4 * l)) & 15u) != 0;
      }
      const uint64_t mx = wave_nib_max(sig[r]);
      if (lane == 0) qmax[g * R + r] = mx;
    }
  }
  // exclusive scan of the padded lengths over the graph's nodes
  const uint32_t len = valid ? (ntok + 3) & ~3u : 0u;
  uint32_t incl = len;
#pragma unroll
  for (int o = 1; o < kWave; o <<= 1) {
    const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), o));
    if (lane >= static_cast<uint32_t>(o)) incl += y;
  }
  uint16_t* t = tokens + static_cast<uint64_t>(prev) * seg + (incl - len);
  uint64_t first = label;  // lane 0: the first four tokens of node 0
  uint32_t cnt = 1;
  if (valid) {
    t[0] = static_cast<uint16_t>(label);
#pragma unroll
    for (uint32_t r = 0; r < kMaxLevels; ++r)
      if (r < R)
        for (uint32_t l = 0; l < 16; ++l) {
          const uint32_t c = static_cast<uint32_t>(sig[r] >> (4 * l)) & 15u;
          if (c) {
            const uint32_t tk = r | (l << 3) | (c << 7);
            if (cnt < 4) first |= static_cast<uint64_t>(tk) << (16 * cnt);
            t[cnt++] = static_cast<uint16_t>(tk);
          }
        }
    t[cnt] = kEndToken;
  }
  if (cnt < 4) first |= static_cast<uint64_t>(kEndToken) << (16 * cnt);
  uint32_t bits = valid && label < 16 ? 1u << label : 0u;
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) bits |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(bits), o));
  if (lane == 0) {
    QueryInfo info;
    info.qoff = prev;
    info.k = k;
    info.qlab = bits;
    info.qcls = 0;  // set by k_query_classes
    info.first = first;
    qinfo[g] = info;
  }
}

// The per-class signature maxima the pre-test reads, from the query graphs' token streams
// (k_query_prepare's output: each requirement token carries one nonzero count, of label l at
// level r, of its node's signature, and every (r, l) occurs at most once per node, so the max over
// a class's nodes is the max over their tokens).  qcmax[(g*3 + c)*R + r] is the nibble-wise max
// of the level-(r+1) signatures of graph g's nodes labelled cl.l[c], and bit c of qinfo[g].qcls is
// set when there is such a node.  One wavefront per graph reads the graph's token words 64 at a
// time, a word per lane: a node's stream starts on a word boundary with its label token and its
// last word holds the end token (which no label or requirement token equals), so the words that
// hold an end token delimit the nodes and give each word its node's label.  Counts go to per-class
// LDS maxima by atomic max; lanes then pack them into nibbles.
__global__ __launch_bounds__(256) void k_query_classes(uint32_t Gq, uint32_t R, const uint64_t* __restrict__ tok64, ClassLabels cl,
                                                       QueryInfo* __restrict__ qinfo, uint64_t* __restrict__ qcmax) {
  constexpr uint32_t kSlots = kClasses * kMaxLevels * 16;
  __shared__ uint32_t cnt_all[4][kSlots];
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint32_t wave = threadIdx.x / kWave;
  const uint64_t g = static_cast<uint64_t>(blockIdx.x) * 4 + wave;
  if (g >= Gq) return;  // wave-uniform; no block barrier follows
  uint32_t* cnt = cnt_all[wave];
  for (uint32_t t = lane; t < kSlots; t += kWave) cnt[t] = 0;
  wave_sync();
  const QueryInfo qi = qinfo[g];
  const uint64_t seg = token_segment(R);
  const uint64_t below = (1ull << lane) - 1;
  const uint64_t end = (static_cast<uint64_t>(qi.qoff) + qi.k) * seg / 4;
  uint32_t nodes = 0, carry_label = 0, qcls = 0;
  bool starts_new = true;  // the chunk's first word starts a node
  for (uint64_t base = static_cast<uint64_t>(qi.qoff) * seg / 4; base < end && nodes < qi.k; base += kWave) {  // wave-uniform
    const uint64_t idx = base + lane;
    const bool in = idx < end;
    const uint64_t w = in ? tok64[idx] : ~0ull;
    bool has_end = false;
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) has_end |= ((w >> (16 * j)) & 0xFFFFull) == kEndToken;
    const uint64_t ends = __ballot(in && has_end);
    const uint64_t prev = ends & below;  // end words before this lane's in the chunk
    const uint32_t start_lane = prev ? 64u - static_cast<uint32_t>(__builtin_clzll(prev)) : 0u;
    const uint32_t first = static_cast<uint32_t>(w & 0xFFFFull);
    const uint32_t lab_here = static_cast<uint32_t>(__shfl(static_cast<int>(first), static_cast<int>(start_lane)));
    const uint32_t label = prev || starts_new ? lab_here : carry_label;
    const bool is_start = prev ? lane == start_lane : (lane == 0 && starts_new);
    const bool valid = in && nodes + static_cast<uint32_t>(__popcll(prev)) < qi.k;
    uint32_t c = kClasses;
#pragma unroll
    for (uint32_t cc = 0; cc < kClasses; ++cc)
      if (c == kClasses && label == cl.l[cc]) c = cc;
    if (valid && c < kClasses) {
      if (is_start) qcls |= 1u << c;
      for (uint32_t j = is_start ? 1u : 0u; j < 4; ++j) {
        const uint32_t t = static_cast<uint32_t>(w >> (16 * j)) & 0xFFFFu;
        if (t == kEndToken) break;
        atomicMax(&cnt[(c * kMaxLevels + (t & 7u)) * 16 + ((t >> 3) & 15u)], (t >> 7) & 15u);
      }
    }
    nodes += static_cast<uint32_t>(__popcll(ends));
    starts_new = (ends >> 63) & 1ull;
    carry_label = static_cast<uint32_t>(__shfl(static_cast<int>(label), kWave - 1));
  }
#pragma unroll
  for (int o = kWave / 2; o > 0; o >>= 1) qcls |= static_cast<uint32_t>(__shfl_xor(static_cast<int>(qcls), o));
  wave_sync();
  if (lane < kClasses * R) {
    const uint32_t c = lane / R, r = lane - c * R;
    uint64_t v = 0;
    for (uint32_t l = 0; l < 16; ++l) v |= static_cast<uint64_t>(cnt[(c * kMaxLevels + r) * 16 + l]) << (4 * l);
    qcmax[(g * kClasses + c) * R + r] = v;
  }
  if (lane == 0) qinfo[g].qcls = qcls;
}

// ---- pair pre-test ------------------------------------------------------------------------------
//
// Two exact necessary conditions for a (query graph, molecule) pair to survive GMCR, tested for
// every pair before any candidate mask is built: (1) the graph's label set (labels < 16) is
// contained in the molecule's, and at every level the graph's nibble-wise signature max is <=
// the molecule's (a node whose signature exceeds the max fits no position); (2) for each class
// label present in the graph, the same holds between the max over the graph's nodes of that label
// and the max over the molecule's positions of that label (a node's candidates carry its label,
// isomorphism.hpp:49).  Query nodes of label 255 only enter (1).

struct PretestArgs {
  uint32_t R, Gq, W, G;          // W: bitmask words per molecule; G: molecules
  const MolInfo* minfo;          // width-class order
  const uint64_t* dmax;          // [c * R + r]
  const uint64_t* cmax;          // [(k * R + r) * G + c]
  const QueryInfo* qinfo;
  const uint64_t* qmax;          // [g * R + r]
  const uint64_t* qcmax;         // [(g * 3 + k) * R + r]
  unsigned long long* surv;      // [m * W + w] out: bit g of word w = pair (g, m) passes
  uint32_t* pcnt;                // [c] out: passing pairs of the molecule
};

// Nibble-wise a <= b for the 16 four-bit fields of several words, bit-sliced: bit b of every
// nibble of a word is one plane (bits 4l), and four words share a set of planes, word j's
// fields at bits 4l + j.  Over the planes, b >= a is the lowest-bit-up chain of ge4_bits with the
// query's planes as per-position thresholds, one v_bitop3_b32 per plane and 32 positions: a
// group of four words costs 8 operations, where splitting each word into even and odd nibbles
// took two 64-bit subtractions and two ANDs per word.  Positions of an unused word hold 0 on both
// sides and pass.
constexpr uint64_t kNibBit = 0x1111111111111111ull;
__device__ __forceinline__ uint64_t nib_plane(uint64_t w, uint32_t b, uint32_t j) { return ((w >> b) & kNibBit) << j; }

// acc &= (x >= m) position-wise over four planes (lo and hi halves of 64 positions)
__device__ __forceinline__ void planes_ge(const uint64_t x[4], const uint64_t* m, uint32_t& acc_lo, uint32_t& acc_hi) {
  uint32_t lo = ge_first(static_cast<uint32_t>(x[0]), static_cast<uint32_t>(m[0]));
  uint32_t hi = ge_first(static_cast<uint32_t>(x[0] >> 32), static_cast<uint32_t>(m[0] >> 32));
#pragma unroll
  for (int b = 1; b < 4; ++b) {
    lo = ge_stage(static_cast<uint32_t>(x[b]), static_cast<uint32_t>(m[b]), lo);
    hi = ge_stage(static_cast<uint32_t>(x[b] >> 32), static_cast<uint32_t>(m[b] >> 32), hi);
  }
  acc_lo &= lo;
  acc_hi &= hi;
}

// One thread per molecule (width-class order) and the query graphs 64 at a time from an LDS tile
// (every lane reads the same graph: broadcast reads), for RR levels in GA groups of four.  Tile
// words of a graph: [0] label set | classes << 32, then the all-node maxima's planes (4 words per
// group), then each class's (4 words per group).  A graph whose label test, or whose all-node
// test, fails on all 64 molecules of a wavefront skips the rest.
template <uint32_t RR>
__global__ __launch_bounds__(256) void k_pair_pretest(PretestArgs a) {
  constexpr uint32_t GA = (RR + 3) / 4;
  constexpr uint32_t S = 1 + 4 * GA * (1 + kClasses);
  __shared__ uint64_t tile[kWave * S];
  const uint64_t c = blockIdx.x * 256ull + threadIdx.x;
  const bool valid = c < a.G;
  uint32_t mlab = 0, m = 0;
  uint64_t md[GA][4], mc[kClasses][GA][4];  // the molecule's maxima as planes
#pragma unroll
  for (uint32_t g = 0; g < GA; ++g)
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) {
      md[g][b] = 0;
#pragma unroll
      for (uint32_t k = 0; k < kClasses; ++k) mc[k][g][b] = 0;
    }
  if (valid) {
    const MolInfo mi = a.minfo[c];
    mlab = mi.mlab;
    m = mi.m;
  }
#pragma unroll
  for (uint32_t r = 0; r < RR; ++r) {
    const uint64_t d = valid ? a.dmax[c * RR + r] : 0ull;
#pragma unroll
    for (uint32_t b = 0; b < 4; ++b) md[r / 4][b] |= nib_plane(d, b, r % 4);
#pragma unroll
    for (uint32_t k = 0; k < kClasses; ++k) {
      const uint64_t x = valid ? a.cmax[(static_cast<uint64_t>(k) * RR + r) * a.G + c] : 0ull;
#pragma unroll
      for (uint32_t b = 0; b < 4; ++b) mc[k][r / 4][b] |= nib_plane(x, b, r % 4);
    }
  }
  uint32_t cnt = 0;
  for (uint32_t w = 0; w < a.W; ++w) {
    __syncthreads();  // the previous tile has been read
    for (uint32_t t = threadIdx.x; t < kWave * S; t += 256) {
      const uint32_t j = t / S, f = t - j * S;
      const uint32_t g = w * kWave + j;
      uint64_t v = 0;
      if (g < a.Gq) {
        if (f == 0) {
          const QueryInfo qi = a.qinfo[g];
          v = qi.qlab | (static_cast<uint64_t>(qi.qcls) << 32);
        } else {
          const uint32_t i = f - 1;                   // (class slot, group, plane)
          const uint32_t cs = i / (4 * GA), grp = (i / 4) % GA, b = i % 4;  // cs 0: all nodes, 1 + k: class k
          const uint64_t* src = cs == 0 ? a.qmax + static_cast<uint64_t>(g) * RR
                                        : a.qcmax + (static_cast<uint64_t>(g) * kClasses + (cs - 1)) * RR;
          for (uint32_t jj = 0; jj < 4 && 4 * grp + jj < RR; ++jj) v |= nib_plane(src[4 * grp + jj], b, jj);
        }
      }
      tile[j * S + f] = v;
    }
    __syncthreads();
    const uint32_t nj = min(static_cast<uint32_t>(kWave), a.Gq - w * kWave);
    uint64_t word = 0;
    for (uint32_t j = 0; j < nj; ++j) {
      const uint64_t* e = tile + j * S;
      const uint64_t h = e[0];
      const bool lab = valid && (static_cast<uint32_t>(h) & ~mlab) == 0u;
      if (!__ballot(lab)) continue;  // wave-uniform
      uint32_t acc_lo = ~0u, acc_hi = ~0u;
#pragma unroll
      for (uint32_t g = 0; g < GA; ++g) planes_ge(md[g], e + 1 + 4 * g, acc_lo, acc_hi);
      bool fits = lab && (acc_lo & acc_hi) == ~0u;
      const uint32_t qcls = static_cast<uint32_t>(h >> 32);
      if (qcls && __ballot(fits)) {  // wave-uniform
#pragma unroll
        for (uint32_t k = 0; k < kClasses; ++k)
          if ((qcls >> k) & 1u) {
#pragma unroll
            for (uint32_t g = 0; g < GA; ++g) planes_ge(mc[k][g], e + 1 + 4 * GA * (1 + k) + 4 * g, acc_lo, acc_hi);
          }
        fits = fits && (acc_lo & acc_hi) == ~0u;
      }
      word |= static_cast<uint64_t>(fits) << j;
    }
    if (valid) a.surv[static_cast<uint64_t>(m) * a.W + w] = word;
    cnt += static_cast<uint32_t>(__popcll(word));
  }
  if (valid) a.pcnt[c] = cnt;
}

// The pre-tested pairs listed in (width-class order, query graph) order: entry = c | g << 32,
// molecule c's pairs from poff[c] (the exclusive scan of pcnt).  Entries past `cap` are not
// written; a list longer than `cap` sets the overflow flag (counters[3]) and the host repeats the
// run with the demand (the scan's total, counters[7]).
__global__ __launch_bounds__(256) void k_pair_expand(uint32_t G, uint32_t W, const MolInfo* __restrict__ minfo,
                                                     const unsigned long long* __restrict__ surv, const uint32_t* __restrict__ poff,
                                                     uint64_t* __restrict__ list, uint64_t cap, unsigned long long* __restrict__ counters) {
  const uint64_t c = blockIdx.x * 256ull + threadIdx.x;
  if (c == 0 && counters[7] > cap) atomicOr(&counters[3], 1ull);
  if (c >= G) return;
  uint64_t pos = poff[c];
  const uint64_t end = poff[c + 1];
  if (pos == end) return;
  const uint64_t m = minfo[c].m;
  for (uint32_t w = 0; w < W && pos < end; ++w) {
    uint64_t x = surv[m * W + w];
    while (x) {
      const uint32_t g = w * kWave + static_cast<uint32_t>(__builtin_ctzll(x));
      x &= x - 1;
      if (pos < cap) list[pos] = c | (static_cast<uint64_t>(g) << 32);
      ++pos;
    }
  }
}

// ---- pair evaluation ----------------------------------------------------------------------------

// Arena words of a pair's block: its k node masks (two words each for molecules of > 64 nodes),
// rounded up to a whole 32-byte group.
__host__ __device__ __forceinline__ uint32_t pair_words(uint32_t k, uint32_t stride) { return (k * stride + 3) & ~3u; }

constexpr uint32_t kSurviveWaves = 4;     // wavefronts per block
constexpr uint32_t kLaneChunk = 64;       // arena words a lane owns at a time (>= the largest pair, 2 x 32)
constexpr uint32_t kWaveChunk = 64 * kLaneChunk;

// Candidate masks of one width class: NW = 1 (<= 32 positions, u32), 2 (<= 64, u64) or
// 4 (<= 128, two u64 words).
template <int NW>
struct Mask {
  using T = typename std::conditional<NW == 1, uint32_t, uint64_t>::type;
  T lo = 0;
  uint64_t hi = 0;  // NW == 4 only
  __device__ __forceinline__ bool any() const { return NW == 4 ? ((static_cast<uint64_t>(lo) | hi) != 0) : (lo != 0); }
};

// One lane's walk over a pair's token words.
template <int NW>
struct PairWalk {
  Mask<NW> mask;
  uint64_t tw = 0;        // the current token word
  uint64_t nodebase = 0;  // token index of the current node's stream
  uint64_t base = 0;      // the pair's first arena word
  uint32_t i = 0, k = 0, tp = 0;  // current node, nodes, token offset within the node's stream
  // node masks are held four words at a time and stored as one aligned 32-byte group (pair
  // blocks start on a 4-word boundary and take a multiple of 4 words): whole sectors go out
  // together, and a pair that dies never stores its open group
  uint64_t buf[4] = {0, 0, 0, 0};
};

enum : int { kWalkOn = 0, kWalkDied = 1, kWalkSurvived = 2 };

// One token word of a pair: a node's label token sets its mask to the filter's label test
// (isomorphism.hpp:49), each requirement token ANDs one bit-sliced refine test
// (isomorphism.hpp:103-110), and the end token stores the node's mask.  The pair dies at the
// first node without candidates (the GMCR predicate, gmcr.hpp:56-61) and survives when its last
// node ends.  `P` gives the four plane words of (level r, label l) and the position masks of a
// label; `ns()` gives the molecule's node count and first node (read only for labels >= 16).  All
// four tokens' plane reads and the next word's load are issued together; requirements of a node
// commute, so their order within the word is free.
template <int NW, class Planes, class NS>
__device__ __forceinline__ int walk_step(const Planes& P, NS ns, const uint8_t* __restrict__ nl, const uint64_t* __restrict__ tok64,
                                         uint32_t R, PairWalk<NW>& w, uint64_t* __restrict__ arena, uint64_t arena_cap) {
  using T = typename Mask<NW>::T;
  constexpr uint32_t stride = NW == 4 ? 2u : 1u;
  uint32_t tok[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) tok[j] = static_cast<uint32_t>(w.tw >> (16 * j)) & 0xFFFFu;
  const uint32_t e = tok[0] == kEndToken ? 0u : tok[1] == kEndToken ? 1u : tok[2] == kEndToken ? 2u : tok[3] == kEndToken ? 3u : 4u;
  const bool last = e < 4 && w.i + 1 == w.k;  // the pair's last node ends in this word
  uint64_t tw_next = 0;
  if (!last) tw_next = tok64[(w.nodebase + w.tp + 4) / 4];  // this node's next word, or the next node's first
  uint64_t pv[4][4], ph[4][4];
#pragma unroll
  for (int j = 0; j < 4; ++j) P.load(min(tok[j] & 7u, R - 1), (tok[j] >> 3) & 15u, pv[j], ph[j]);  // clamped, masked below
  const uint32_t ql = tok[0];
  if (w.tp == 0) {  // the node's label token: the filter's label test (isomorphism.hpp:49)
    if (ql < 16u) {
      uint64_t lo = 0, hi = 0;
      P.lab(ql, lo, hi);
      w.mask.lo = static_cast<T>(lo);
      if (NW == 4) w.mask.hi = hi;
    } else {
      const uint2 mns = ns();  // (n, s)
      if (ql == 255u) {
        w.mask.lo = static_cast<T>(mns.x >= 64 ? ~0ull : ((1ull << mns.x) - 1));
        if (NW == 4) w.mask.hi = mns.x >= 128 ? ~0ull : ((1ull << (mns.x - 64)) - 1);
      } else {
        uint64_t l0 = 0, l1 = 0;
        for (uint32_t p = 0; p < mns.x; ++p)
          if (nl[mns.y + p] == ql) {
            if (p < 64) l0 |= 1ull << p;
            else l1 |= 1ull << (p - 64);
          }
        w.mask.lo = static_cast<T>(l0);
        if (NW == 4) w.mask.hi = l1;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {  // count_q(l) <= count_p(l) at level r (isomorphism.hpp:103-110)
    const bool req = static_cast<uint32_t>(j) < e && (w.tp != 0 || j != 0);
    const uint32_t c = req ? (tok[j] >> 7) & 15u : 0u;  // c = 0: always true
    w.mask.lo &= static_cast<T>(ge4_bits<T>(static_cast<T>(pv[j][0]), static_cast<T>(pv[j][1]), static_cast<T>(pv[j][2]),
                                            static_cast<T>(pv[j][3]), c));
    if (NW == 4) w.mask.hi &= ge4_bits<uint64_t>(ph[j][0], ph[j][1], ph[j][2], ph[j][3], c);
  }
  int out = kWalkOn;
  if (!w.mask.any()) {
    out = kWalkDied;
  } else if (e < 4) {  // node done: its mask joins the register group, stored when full or last
    const uint32_t wd = w.i * stride;  // the node's first word in the pair's block
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      if (j == (wd & 3u)) w.buf[j] = static_cast<uint64_t>(w.mask.lo);
      if (NW == 4 && j == (wd & 3u) + 1) w.buf[j] = w.mask.hi;
    }
    if ((wd & 3u) + stride == 4u || w.i + 1 == w.k) {
      const uint64_t at = w.base + (wd & ~3u);
      if (at + 4 <= arena_cap) {  // 32-byte aligned: two 16-byte stores
        uint4* dst = reinterpret_cast<uint4*>(arena + at);
        dst[0] = make_uint4(static_cast<uint32_t>(w.buf[0]), static_cast<uint32_t>(w.buf[0] >> 32), static_cast<uint32_t>(w.buf[1]),
                            static_cast<uint32_t>(w.buf[1] >> 32));
        dst[1] = make_uint4(static_cast<uint32_t>(w.buf[2]), static_cast<uint32_t>(w.buf[2] >> 32), static_cast<uint32_t>(w.buf[3]),
                            static_cast<uint32_t>(w.buf[3] >> 32));
      }
    }
    ++w.i;
    if (w.i == w.k) {
      out = kWalkSurvived;
    } else {
      w.nodebase += w.tp + 4;  // the next node's stream starts at the next word
      w.tp = 0;
    }
  } else {
    w.tp += 4;
  }
  w.tw = tw_next;
  return out;
}

// A lane's arena sub-chunk and its wavefront's chunk: lanes whose sub-chunk cannot hold a new
// pair of `need` words get a fresh one from the wavefront's chunk, which claims kWaveChunk words
// of the arena at a time (setting the overflow flag, counters[3], past the arena's end).
struct ArenaCursor {
  uint64_t wc_next = 0, wc_end = 0;  // the wavefront's chunk (wave-uniform)
  uint64_t lc_next = 0, lc_end = 0;  // the lane's sub-chunk
  __device__ __forceinline__ void reserve(bool take, uint32_t need, uint64_t below, unsigned long long* counters, uint64_t arena_cap) {
    const bool fresh = take && lc_next + need > lc_end;
    const uint64_t fm = __ballot(fresh);
    if (fm) {
      const uint32_t nf = __popcll(fm);
      if (wc_next + static_cast<uint64_t>(nf) * kLaneChunk > wc_end) {
        unsigned long long b = 0;
        if (lane_id() == 0) {
          b = atomicAdd(&counters[1], static_cast<unsigned long long>(kWaveChunk));
          if (b + kWaveChunk > arena_cap) atomicOr(&counters[3], 1ull);
        }
        wc_next = uniform64(__shfl(b, 0));
        wc_end = wc_next + kWaveChunk;
      }
      if (fresh) {
        lc_next = wc_next + static_cast<uint64_t>(__popcll(fm & below)) * kLaneChunk;
        lc_end = lc_next + kLaneChunk;
      }
      wc_next += static_cast<uint64_t>(nf) * kLaneChunk;
    }
  }
};

// A pre-tested (query graph, molecule) pair waiting for a lane.
struct PairEntry {
  uint32_t slot, qg, qoff, k;
  uint64_t tw;  // first token word of the pair's first query node
};

struct SurviveArgs {
  GraphDev q, d;
  uint32_t R, Gq, W;       // W: survival-bitmask words per molecule
  uint32_t count;          // molecules of the width class
  uint32_t claim;          // molecules a wavefront claims at a time
  const MolInfo* minfo;    // the width class's molecules, ascending
  const uint32_t* pcnt;    // the width class's pre-tested pair counts
  const uint32_t* planes32;
  const uint64_t* planes64;
  const uint64_t* planes_hi;
  const uint64_t* lab_hi;
  const uint64_t* tok64;   // token segments as 64-bit words
  const QueryInfo* qinfo;
  uint64_t* arena;         // survivor masks: k words (2k, interleaved, for molecules of > 64 nodes)
  uint64_t arena_cap;
  unsigned long long* surv;      // [Gd][W] in: the pre-test's pairs; out: bit qg of row m = the pair survives GMCR
  uint32_t* pair_off;            // [Gd * Gq] arena offset of each surviving pair
  unsigned long long* counters;  // [0] molecule cursor, [1] arena top, [3] overflow, [4..5] statistics
};

// Molecule slots per wavefront and LDS words (u64) per slot, by width class.
__host__ __device__ constexpr uint32_t ring_slots(int nw) { return nw == 1 ? 4u : (nw == 2 ? 3u : 2u); }
__host__ __device__ constexpr uint32_t slot_words(uint32_t R, int nw) {
  return nw == 1 ? R * 32 + 8 : (nw == 2 ? R * 64 + 16 : 2 * R * 64 + 32);
}
__host__ __device__ constexpr uint32_t plist_words() { return kWave * sizeof(PairEntry) / 8; }
__host__ __device__ constexpr uint32_t survive_lds_bytes(uint32_t R, int nw) {
  return kSurviveWaves * (ring_slots(nw) * slot_words(R, nw) + plist_words()) * 8;
}

// The plane words of a molecule staged in an LDS slot (layout of k_pair_survive).
template <int NW>
struct SlotPlanes {
  const uint64_t* slot;
  uint32_t R;
  __device__ __forceinline__ void load(uint32_t r, uint32_t l, uint64_t pv[4], uint64_t ph[4]) const {
    if (NW == 1) {
      const uint4 v = *reinterpret_cast<const uint4*>(reinterpret_cast<const uint32_t*>(slot) + r * 64 + l * 4);
      pv[0] = v.x;
      pv[1] = v.y;
      pv[2] = v.z;
      pv[3] = v.w;
    } else {
#pragma unroll
      for (int b = 0; b < 4; ++b) pv[b] = slot[r * 64 + l * 4 + b];
      if (NW == 4) {
#pragma unroll
        for (int b = 0; b < 4; ++b) ph[b] = slot[R * 64 + r * 64 + l * 4 + b];
      }
    }
  }
  __device__ __forceinline__ void lab(uint32_t l, uint64_t& lo, uint64_t& hi) const {
    const uint32_t lm_at = NW == 1 ? 2 * R * 32 : (NW == 4 ? 2 * R * 64 : R * 64);  // u32 index for NW == 1
    lo = NW == 1 ? reinterpret_cast<const uint32_t*>(slot)[lm_at + l] : slot[lm_at + l];
    hi = NW == 4 ? slot[lm_at + 16 + l] : 0ull;
  }
};

// The plane words of a molecule read from the database in device memory.
template <int NW>
struct GlobalPlanes {
  const uint32_t* p32;  // the molecule's planes (NW == 1)
  const uint64_t* p64;  // the molecule's planes (NW >= 2)
  const uint64_t* phi;  // positions 64..127 (NW == 4)
  const uint64_t* lm;   // label masks of positions 0..63
  const uint64_t* lmh;  // label masks of positions 64..127 (NW == 4)
  __device__ __forceinline__ void load(uint32_t r, uint32_t l, uint64_t pv[4], uint64_t ph[4]) const {
    if (NW == 1) {
      const uint4 v = *reinterpret_cast<const uint4*>(p32 + r * 64 + l * 4);
      pv[0] = v.x;
      pv[1] = v.y;
      pv[2] = v.z;
      pv[3] = v.w;
    } else {
      const uint4* s = reinterpret_cast<const uint4*>(p64 + r * 64 + l * 4);
      const uint4 v0 = s[0], v1 = s[1];
      pv[0] = v0.x | (static_cast<uint64_t>(v0.y) << 32);
      pv[1] = v0.z | (static_cast<uint64_t>(v0.w) << 32);
      pv[2] = v1.x | (static_cast<uint64_t>(v1.y) << 32);
      pv[3] = v1.z | (static_cast<uint64_t>(v1.w) << 32);
      if (NW == 4) {
        const uint4* t = reinterpret_cast<const uint4*>(phi + r * 64 + l * 4);
        const uint4 u0 = t[0], u1 = t[1];
        ph[0] = u0.x | (static_cast<uint64_t>(u0.y) << 32);
        ph[1] = u0.z | (static_cast<uint64_t>(u0.w) << 32);
        ph[2] = u1.x | (static_cast<uint64_t>(u1.y) << 32);
        ph[3] = u1.z | (static_cast<uint64_t>(u1.w) << 32);
      }
    }
  }
  __device__ __forceinline__ void lab(uint32_t l, uint64_t& lo, uint64_t& hi) const {
    lo = lm[l];
    hi = NW == 4 ? lmh[l] : 0ull;
  }
};

// Persistent pair evaluation of one width class, molecule by molecule.  Each wavefront keeps a
// ring of molecule slots in LDS -- the molecule's level planes and label masks, staged with
// coalesced loads when the wavefront claims it (molecules with no pre-tested pair are skipped) --
// and lists the open molecule's pre-tested pairs, 64 query graphs at a time, in an LDS list from
// which idle lanes take pairs of any slot.  A slot is reused once its molecule's list is done and
// none of its pairs is pending.  Each lane owns an arena sub-chunk: node masks go straight to the
// pair's final place, and a pair that dies leaves its space to the lane's next pair (and clears
// its pre-test bit).  Suited to many pairs per molecule (large query batches).
template <int NW>
__global__ __launch_bounds__(256) void k_pair_survive(SurviveArgs a) {
  constexpr uint32_t stride = NW == 4 ? 2u : 1u;
  constexpr uint32_t S = ring_slots(NW);
  extern __shared__ uint64_t lds[];
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint32_t wave = threadIdx.x / kWave;
  const uint32_t R = a.R, Gq = a.Gq;
  const uint32_t sw = slot_words(R, NW);
  uint64_t* ring = lds + static_cast<uint64_t>(wave) * (S * sw + plist_words());
  PairEntry* plist = reinterpret_cast<PairEntry*>(ring + S * sw);
  const uint32_t seg = token_segment(R);
  const uint64_t below = (1ull << lane) - 1;

  // wave-uniform state
  ArenaCursor ac;
  uint32_t lhead = 0, lcount = 0;             // pending list entries [lhead, lcount)
  uint32_t pending[S];                        // listed pairs of the slot's molecule not yet finished
  MolInfo smol[S];                            // the slot's molecule
  uint32_t open = S, scan_qg = 0;             // slot being listed (S: none) and its next query graph
  // With at most 64 bitmask words per molecule the open molecule's words sit in registers (lane i:
  // word i, listed bits cleared), loaded with one coalesced load when it is opened.
  const bool reg_words = a.W <= kWave;
  uint64_t mw = 0;
  uint32_t om = 0;                            // the open molecule
  bool exhausted = false;
  uint64_t mc_next = 0, mc_end = 0;           // claimed molecules [mc_next, mc_end)
#pragma unroll
  for (uint32_t sl = 0; sl < S; ++sl) pending[sl] = 0;
  uint64_t wave_steps = 0, lane_steps = 0;    // statistics

  // per-lane pair state
  bool have = false;
  uint32_t sl_of = 0, qg = 0;
  PairWalk<NW> pw;
  while (true) {
    uint64_t need = __ballot(!have);
    while (need) {
      if (lhead < lcount) {  // hand waiting pairs to idle lanes
        const uint32_t cnt = min(lcount - lhead, static_cast<uint32_t>(__popcll(need)));
        const uint32_t rn = __popcll(need & below);
        const bool take = ((need >> lane) & 1ull) && rn < cnt;
        if (take) {
          const PairEntry& e = plist[lhead + rn];
          sl_of = e.slot;
          qg = e.qg;
          pw.k = e.k;
          pw.tw = e.tw;
          pw.nodebase = static_cast<uint64_t>(e.qoff) * seg;
          pw.i = 0;
          pw.tp = 0;
          have = true;
        }
        wave_sync();
        lhead += cnt;
        need &= ~__ballot(take);
        ac.reserve(take, pair_words(pw.k, stride), below, a.counters, a.arena_cap);
        if (take) pw.base = ac.lc_next;
        continue;
      }
      if (reg_words && open < S && __ballot(mw != 0ull)) {
        // list the open molecule's next pre-tested pairs: the words from the lowest unlisted one up
        // whose pairs fit the 64-entry list (the first always does), then each entry's query graph
        // fields, one load per entry, all in flight together
        const uint32_t pc = static_cast<uint32_t>(__popcll(mw));
        uint32_t incl = pc;
#pragma unroll
        for (int d = 1; d < kWave; d <<= 1) {
          const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), d));
          if (lane >= static_cast<uint32_t>(d)) incl += y;
        }
        const bool sel = pc != 0 && incl <= kWave;
        const uint64_t selm = __ballot(sel);
        for (uint64_t f = selm; f; f &= f - 1) {  // wave-uniform
          const int w = static_cast<int>(__builtin_ctzll(f));
          const uint64_t x = __shfl(mw, w);
          const uint32_t at = static_cast<uint32_t>(__shfl(static_cast<int>(incl - pc), w));
          if ((x >> lane) & 1ull) {
            PairEntry& e = plist[at + __popcll(x & below)];
            e.slot = open;
            e.qg = static_cast<uint32_t>(w) * kWave + lane;
          }
        }
        const uint32_t cnt = static_cast<uint32_t>(__shfl(static_cast<int>(incl), 63 - static_cast<int>(__builtin_clzll(selm))));
        if (sel) mw = 0;
        wave_sync();
        if (lane < cnt) {
          PairEntry& e = plist[lane];
          const QueryInfo qi = a.qinfo[e.qg];
          e.qoff = qi.qoff;
          e.k = qi.k;
          e.tw = qi.first;
        }
        wave_sync();
        lhead = 0;
        lcount = cnt;
#pragma unroll
        for (uint32_t sl = 0; sl < S; ++sl)
          if (sl == open) pending[sl] += lcount;
        continue;
      }
      if (!reg_words && open < S && scan_qg < Gq) {  // list the open molecule's pre-tested pairs of the next 64 query graphs
        const uint64_t word = a.surv[static_cast<uint64_t>(om) * a.W + scan_qg / kWave];
        const uint32_t g = scan_qg + lane;
        const bool ok = (word >> lane) & 1ull;
        if (ok) {
          const QueryInfo qi = a.qinfo[g];
          PairEntry& e = plist[__popcll(word & below)];
          e.slot = open;
          e.qg = g;
          e.qoff = qi.qoff;
          e.k = qi.k;
          e.tw = qi.first;
        }
        wave_sync();
        lhead = 0;
        lcount = __popcll(word);
#pragma unroll
        for (uint32_t sl = 0; sl < S; ++sl)
          if (sl == open) pending[sl] += lcount;
        scan_qg += kWave;
        continue;
      }
      // the open molecule is fully listed: open the next one with pairs in a free slot
      if (exhausted) break;
      uint32_t free_slot = S;
#pragma unroll
      for (uint32_t sl = 0; sl < S; ++sl)
        if (free_slot == S && pending[sl] == 0) free_slot = sl;
      if (free_slot == S) break;  // every slot still has pairs in flight
      if (mc_next >= mc_end) {  // claim the next few molecules (one atomic per claim)
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(&a.counters[0], static_cast<unsigned long long>(a.claim));
        mc_next = uniform64(__shfl(b, 0));
        mc_end = mc_next + a.claim < a.count ? mc_next + a.claim : a.count;
      }
      if (mc_next >= a.count) {
        exhausted = true;
        open = S;
        break;
      }
      const uint64_t c = mc_next++;
      if (a.pcnt[c] == 0) continue;  // nothing passed the pre-test: not staged
      const MolInfo mi = a.minfo[c];
      uint64_t* slot = ring + free_slot * sw;
      if (NW == 1) {
        const uint64_t* src = reinterpret_cast<const uint64_t*>(a.planes32 + static_cast<uint64_t>(mi.x) * R * 64);
        for (uint32_t t = lane; t < R * 32; t += kWave) slot[t] = src[t];
        if (lane < 16) reinterpret_cast<uint32_t*>(slot + R * 32)[lane] = static_cast<uint32_t>(a.d.lmask[static_cast<uint64_t>(mi.m) * 16 + lane]);
      } else {
        const uint64_t* src = a.planes64 + static_cast<uint64_t>(mi.x) * R * 64;
        for (uint32_t t = lane; t < R * 64; t += kWave) slot[t] = src[t];
        const uint32_t lm_at = NW == 4 ? 2 * R * 64 : R * 64;
        if (lane < 16) slot[lm_at + lane] = a.d.lmask[static_cast<uint64_t>(mi.m) * 16 + lane];
        if (NW == 4) {
          const uint64_t* srch = a.planes_hi + static_cast<uint64_t>(mi.wi) * R * 64;
          for (uint32_t t = lane; t < R * 64; t += kWave) slot[R * 64 + t] = srch[t];
          if (lane < 16) slot[lm_at + 16 + lane] = a.lab_hi[static_cast<uint64_t>(mi.wi) * 16 + lane];
        }
      }
      wave_sync();
#pragma unroll
      for (uint32_t sl = 0; sl < S; ++sl)
        if (sl == free_slot) smol[sl] = mi;
      open = free_slot;
      om = mi.m;
      scan_qg = 0;
      if (reg_words) mw = lane < a.W ? a.surv[static_cast<uint64_t>(om) * a.W + lane] : 0ull;
    }
    const uint64_t active = __ballot(have);
    if (active == 0) break;
    ++wave_steps;
    lane_steps += __popcll(active);
    bool finished = false;
    if (have) {
      const SlotPlanes<NW> P{ring + sl_of * sw, R};
      auto ns = [&]() {
        MolInfo mi = {};
#pragma unroll
        for (uint32_t sl = 0; sl < S; ++sl)
          if (sl == sl_of) mi = smol[sl];
        return make_uint2(mi.n, mi.s);
      };
      const int st = walk_step<NW>(P, ns, a.d.nl, a.tok64, R, pw, a.arena, a.arena_cap);
      if (st != kWalkOn) {
        uint32_t m = 0;
#pragma unroll
        for (uint32_t sl = 0; sl < S; ++sl)
          if (sl == sl_of) m = smol[sl].m;
        if (st == kWalkSurvived) {  // every query node has a candidate: the pair survives GMCR
          a.pair_off[static_cast<uint64_t>(m) * Gq + qg] = static_cast<uint32_t>(pw.base);
          ac.lc_next = pw.base + pair_words(pw.k, stride);
        } else {
          atomicAnd(&a.surv[static_cast<uint64_t>(m) * a.W + (qg >> 6)], ~(1ull << (qg & 63)));
        }
        have = false;
        finished = true;
      }
    }
    // pairs finished this step release their slot's count
#pragma unroll
    for (uint32_t sl = 0; sl < S; ++sl) pending[sl] -= __popcll(__ballot(finished && sl_of == sl));
  }
  if (lane == 0) {
    atomicAdd(&a.counters[4], static_cast<unsigned long long>(wave_steps));
    atomicAdd(&a.counters[5], static_cast<unsigned long long>(lane_steps));
  }
}

struct EvalArgs {
  GraphDev d;
  uint32_t R, Gq, W;
  const MolInfo* minfo;    // width-class order (all classes)
  const uint32_t* planes32;
  const uint64_t* planes64;
  const uint64_t* planes_hi;
  const uint64_t* lab_hi;
  const uint64_t* tok64;
  const QueryInfo* qinfo;
  const uint64_t* list;    // pre-tested pairs, c | g << 32
  const uint32_t* poff;    // list offset of each molecule (width-class order)
  uint64_t list_cap;
  uint32_t first, count;   // the width class's molecules [first, first + count)
  uint32_t chunk;          // list entries a wavefront claims at a time
  uint64_t* arena;
  uint64_t arena_cap;
  unsigned long long* surv;
  uint32_t* pair_off;
  unsigned long long* counters;  // [0] list cursor, [1] arena top, [3] overflow, [4..5] statistics
};

// Persistent pair evaluation of one width class, pair by pair: lanes take the class's entries of
// the pre-tested pair list as they become idle and read the molecule's planes from device memory
// (neighbouring entries mostly share a molecule, so the reads share cache lines).  No molecule is
// staged, so every lane has a pair however few pairs a molecule has (small query batches).
// Survivors and deaths are recorded as in k_pair_survive.
template <int NW>
__global__ __launch_bounds__(256) void k_pair_eval(EvalArgs a) {
  constexpr uint32_t stride = NW == 4 ? 2u : 1u;
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint32_t R = a.R, Gq = a.Gq;
  const uint32_t seg = token_segment(R);
  const uint64_t below = (1ull << lane) - 1;
  const uint64_t lo = a.poff[a.first];
  const uint64_t hi = min(static_cast<uint64_t>(a.poff[a.first + a.count]), a.list_cap);
  ArenaCursor ac;
  uint64_t next = 0, end = 0;  // claimed entries [next, end) (wave-uniform)
  bool exhausted = lo >= hi;
  uint64_t wave_steps = 0, lane_steps = 0;

  bool have = false;
  uint32_t qg = 0, m = 0, n = 0, s = 0;
  GlobalPlanes<NW> P{};
  PairWalk<NW> pw;
  while (true) {
    uint64_t need = __ballot(!have);
    while (need && !exhausted) {
      if (next >= end) {
        unsigned long long b = 0;
        if (lane == 0) b = atomicAdd(&a.counters[0], static_cast<unsigned long long>(a.chunk));
        next = lo + uniform64(__shfl(b, 0));
        end = next + a.chunk < hi ? next + a.chunk : hi;
        if (next >= hi) {
          exhausted = true;
          break;
        }
      }
      const uint32_t cnt = min(static_cast<uint32_t>(end - next), static_cast<uint32_t>(__popcll(need)));
      const uint32_t rn = __popcll(need & below);
      const bool take = ((need >> lane) & 1ull) && rn < cnt;
      if (take) {
        const uint64_t e = a.list[next + rn];
        const MolInfo mi = a.minfo[static_cast<uint32_t>(e)];
        qg = static_cast<uint32_t>(e >> 32);
        const QueryInfo qi = a.qinfo[qg];
        m = mi.m;
        n = mi.n;
        s = mi.s;
        P.p32 = a.planes32 + static_cast<uint64_t>(mi.x) * R * 64;
        P.p64 = a.planes64 + static_cast<uint64_t>(mi.x) * R * 64;
        P.phi = a.planes_hi + static_cast<uint64_t>(NW == 4 ? mi.wi : 0u) * R * 64;
        P.lm = a.d.lmask + static_cast<uint64_t>(mi.m) * 16;
        P.lmh = a.lab_hi + static_cast<uint64_t>(NW == 4 ? mi.wi : 0u) * 16;
        pw.k = qi.k;
        pw.tw = qi.first;
        pw.nodebase = static_cast<uint64_t>(qi.qoff) * seg;
        pw.i = 0;
        pw.tp = 0;
        have = true;
      }
      next += cnt;
      need &= ~__ballot(take);
      ac.reserve(take, pair_words(pw.k, stride), below, a.counters, a.arena_cap);
      if (take) pw.base = ac.lc_next;
    }
    const uint64_t active = __ballot(have);
    if (active == 0) break;
    ++wave_steps;
    lane_steps += __popcll(active);
    if (have) {
      auto ns = [&]() { return make_uint2(n, s); };
      const int st = walk_step<NW>(P, ns, a.d.nl, a.tok64, R, pw, a.arena, a.arena_cap);
      if (st == kWalkSurvived) {
        a.pair_off[static_cast<uint64_t>(m) * Gq + qg] = static_cast<uint32_t>(pw.base);
        ac.lc_next = pw.base + pair_words(pw.k, stride);
        have = false;
      } else if (st == kWalkDied) {
        atomicAnd(&a.surv[static_cast<uint64_t>(m) * a.W + (qg >> 6)], ~(1ull << (qg & 63)));
        have = false;
      }
    }
  }
  if (lane == 0) {
    atomicAdd(&a.counters[4], static_cast<unsigned long long>(wave_steps));
    atomicAdd(&a.counters[5], static_cast<unsigned long long>(lane_steps));
  }
}

// Survivors per molecule: the popcount of its row of the survival bitmask.
__global__ __launch_bounds__(256) void k_surv_count(uint32_t Gd, uint32_t W, const unsigned long long* __restrict__ surv,
                                                    uint32_t* __restrict__ counts) {
  const uint64_t m = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (m >= Gd) return;
  uint32_t c = 0;
  for (uint32_t w = 0; w < W; ++w) c += __popcll(surv[m * W + w]);
  counts[m] = c;
}

// Canonical GMCR arrays (gmcr.hpp:73-115 order: molecules ascending, query graphs ascending
// within a molecule) from the survival bitmask, with each pair's arena offset.  One wavefront
// per molecule: lane i loads bitmask word i of each run of 64 (one coalesced load), the lanes'
// popcounts are scanned, and the wavefront visits only the nonzero words, lane l writing query
// graph 64w+l of word w, so each word's survivors go out in one coalesced store per array.
__global__ __launch_bounds__(256) void k_surv_scatter(uint32_t Gd, uint32_t Gq, uint32_t W, const unsigned long long* __restrict__ surv,
                                                      const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ pair_off,
                                                      uint32_t* __restrict__ indices, uint32_t* __restrict__ pair_graph,
                                                      uint32_t* __restrict__ mask_off, uint64_t cap) {
  // `cap` bounds the writes: a replayed graph sized by an earlier run never writes past its
  // arrays (the replay then reports the changed pair count)
  const uint64_t m = (blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x) / kWave;
  if (m >= Gd) return;  // wave-uniform
  const uint32_t lane = static_cast<uint32_t>(lane_id());
  const uint64_t below = (1ull << lane) - 1;
  uint64_t o = offsets[m];
  if (offsets[m + 1] == o) return;  // no survivors (wave-uniform)
  for (uint32_t w0 = 0; w0 < W; w0 += kWave) {  // wave-uniform
    const uint64_t xl = w0 + lane < W ? surv[m * W + w0 + lane] : 0ull;
    const uint32_t pc = static_cast<uint32_t>(__popcll(xl));
    uint32_t incl = pc;
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
      const uint32_t y = static_cast<uint32_t>(__shfl_up(static_cast<int>(incl), d));
      if (lane >= static_cast<uint32_t>(d)) incl += y;
    }
    for (uint64_t nz = __ballot(xl != 0ull); nz; nz &= nz - 1) {  // wave-uniform
      const int wl = static_cast<int>(__builtin_ctzll(nz));
      const uint64_t x = __shfl(xl, wl);
      const uint32_t ob = static_cast<uint32_t>(__shfl(static_cast<int>(incl - pc), wl));
      if ((x >> lane) & 1ull) {
        const uint64_t r = o + ob + static_cast<uint64_t>(__popcll(x & below));
        const uint32_t g = (w0 + static_cast<uint32_t>(wl)) * 64 + lane;
        if (r < cap) {
          indices[r] = g;
          pair_graph[r] = static_cast<uint32_t>(m);
          mask_off[r] = pair_off[static_cast<uint64_t>(m) * Gq + g];
        }
      }
    }
    o += static_cast<uint32_t>(__shfl(static_cast<int>(incl), kWave - 1));
  }
}

// k_surv_scatter with one thread per molecule instead of one wavefront, for batches of at most 64
// query graphs (one bitmask word per molecule): most molecules then have no or a few survivors,
// and a wavefront per molecule would leave almost all of its lanes idle.
__global__ __launch_bounds__(256) void k_surv_scatter_thread(uint32_t Gd, uint32_t Gq, uint32_t W, const unsigned long long* __restrict__ surv,
                                                             const uint32_t* __restrict__ offsets, const uint32_t* __restrict__ pair_off,
                                                             uint32_t* __restrict__ indices, uint32_t* __restrict__ pair_graph,
                                                             uint32_t* __restrict__ mask_off, uint64_t cap) {
  const uint64_t m = blockIdx.x * 256ull + threadIdx.x;
  if (m >= Gd) return;
  uint64_t o = offsets[m];
  if (offsets[m + 1] == o) return;
  for (uint32_t w = 0; w < W; ++w) {
    uint64_t x = surv[m * W + w];
    while (x) {
      const uint32_t g = w * 64 + static_cast<uint32_t>(__builtin_ctzll(x));
      x &= x - 1;
      if (o < cap) {
        indices[o] = g;
        pair_graph[o] = static_cast<uint32_t>(m);
        mask_off[o] = pair_off[m * Gq + g];
      }
      ++o;
    }
  }
}

// Per-request statistics of a batch run.  Request r owns the batch's query graphs
// [rg[r], rg[r+1]); each GMCR pair (molecule m, query graph g, count c) adds its match flag,
// c, and mix64(mix64(((g - rg[r]) << 32) | m) + c) to its request (sums mod 2^64, mix64 being
// the splitmix64 finaliser).  Sums do not depend on the order pairs are visited, so a request's
// hash is the same whether it ran alone or batched with others.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

constexpr uint32_t kMaxBatchRequests = 256;

__global__ __launch_bounds__(256) void k_request_stats(uint64_t P, const uint32_t* __restrict__ indices, const uint32_t* __restrict__ pair_graph,
                                                       const int32_t* __restrict__ counts, const uint32_t* __restrict__ rg, uint32_t nreq,
                                                       unsigned long long* __restrict__ out) {
  __shared__ uint32_t srg[kMaxBatchRequests + 1];
  __shared__ unsigned long long acc[kMaxBatchRequests][3];
  for (uint32_t t = threadIdx.x; t <= nreq; t += blockDim.x) srg[t] = rg[t];
  for (uint32_t t = threadIdx.x; t < nreq; t += blockDim.x) acc[t][0] = acc[t][1] = acc[t][2] = 0;
  __syncthreads();
  for (uint64_t j = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; j < P; j += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t g = indices[j];
    uint32_t lo = 0, hi = nreq;  // last r with srg[r] <= g
    while (hi - lo > 1) {
      const uint32_t mid = (lo + hi) / 2;
      if (srg[mid] <= g) lo = mid;
      else hi = mid;
    }
    const int32_t c = counts[j];
    if (c > 0) atomicAdd(&acc[lo][0], 1ull);
    atomicAdd(&acc[lo][1], static_cast<unsigned long long>(c));
    const uint64_t key = (static_cast<uint64_t>(g - srg[lo]) << 32) | pair_graph[j];
    atomicAdd(&acc[lo][2], static_cast<unsigned long long>(mix64(mix64(key) + static_cast<uint32_t>(c))));
  }
  __syncthreads();
  for (uint32_t t = threadIdx.x; t < nreq; t += blockDim.x)
    for (int f = 0; f < 3; ++f)
      if (acc[t][f]) atomicAdd(&out[3 * t + f], acc[t][f]);
}

}  // namespace

// The first `small` width-class entries (molecules of <= 32 nodes) two per wavefront, the rest one
// per wavefront.
hipError_t db_prepare(hipStream_t s, const GraphDev& g, uint32_t R, const uint32_t* cls_list, const uint32_t* pidx, const uint32_t* wide_idx,
                      uint32_t* planes32, uint64_t* planes64, uint64_t* planes_hi, uint64_t* lab_hi, uint64_t* dmax, MolInfo* minfo,
                      const ClassLabels& cl, uint64_t* cmax, uint32_t small) {
  if (!g.G) return hipSuccess;
  if (small) {
    const uint32_t waves = (small + 1) / 2;
    k_db_prepare_pairs<<<(waves + 3) / 4, 256, 0, s>>>(g, R, small, cls_list, pidx, planes32, dmax, minfo, cl, cmax);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (g.G > small)
    k_db_prepare<<<(g.G - small + 3) / 4, 256, 0, s>>>(g, R, cls_list, pidx, wide_idx, planes32, planes64, planes_hi, lab_hi, dmax, minfo, cl,
                                                       cmax, small);
  return hipGetLastError();
}

hipError_t query_prepare(hipStream_t s, const GraphDev& q, uint32_t R, uint16_t* tokens, QueryInfo* qinfo, uint64_t* qmax, const ClassLabels& cl,
                         uint64_t* qcmax) {
  if (!q.G) return hipSuccess;
  k_query_prepare<<<(q.G + 3) / 4, 256, 0, s>>>(q, R, tokens, qinfo, qmax);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_query_classes<<<(q.G + 3) / 4, 256, 0, s>>>(q.G, R, reinterpret_cast<const uint64_t*>(tokens), cl, qinfo, qcmax);
  return hipGetLastError();
}

// Node-label histogram (256 bins) of a graph batch.
__global__ __launch_bounds__(256) void k_label_hist(const uint8_t* __restrict__ nl, uint64_t N, unsigned long long* __restrict__ hist) {
  __shared__ unsigned int bins[256];
  bins[threadIdx.x] = 0;
  __syncthreads();
  for (uint64_t v = blockIdx.x * 256ull + threadIdx.x; v < N; v += static_cast<uint64_t>(gridDim.x) * 256) atomicAdd(&bins[nl[v]], 1u);
  __syncthreads();
  if (bins[threadIdx.x]) atomicAdd(&hist[threadIdx.x], static_cast<unsigned long long>(bins[threadIdx.x]));
}

hipError_t label_hist(hipStream_t s, const uint8_t* nl, uint64_t N, unsigned long long* hist) {
  hipError_t e = fill_u64(s, reinterpret_cast<uint64_t*>(hist), 256, 0);
  if (e != hipSuccess || !N) return e;
  const uint64_t blocks = std::min<uint64_t>(1024, (N + 255) / 256);
  k_label_hist<<<static_cast<unsigned>(blocks), 256, 0, s>>>(nl, N, hist);
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
// Growth with an eighth to spare, for buffers whose size follows the data of a streaming chunk:
// chunks of similar size then rarely reallocate (hipFree waits for the whole device).
template <typename T>
int bgrow_slack(T** p, uint64_t& cap, uint64_t need) {
  if (*p && cap >= need) return SIGMO_OK;
  return bgrow(p, cap, need + need / 8);
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

// Pair spaces up to this size run with one host synchronisation (sigmo_batch_run).
constexpr uint64_t kOneSyncPairs = 1ull << 25;
// Query graphs per batch below which the pair evaluation runs pair by pair from the list: with
// fewer, a molecule has too few pre-tested pairs to keep a wavefront's lanes busy from the few
// molecules its LDS ring holds.
constexpr uint32_t kListModeGraphs = 256;

struct sigmo_batch;
namespace sigmo_internal {
sigmo_ctx* batch_ctx(const sigmo_batch* b);
}

struct sigmo_db {
  sigmo_ctx* ctx = nullptr;
  const sigmo_graph* data = nullptr;
  uint32_t R = 0;
  uint32_t num_wide = 0;
  uint32_t cls_count[3] = {0, 0, 0};  // molecules of <= 32, 33..64 and 65..128 nodes
  uint32_t* cls_list = nullptr;       // the three classes' molecules one after another, grouped by label set and size
  uint32_t* pidx = nullptr;           // index of the molecule's planes in planes32 or planes64
  uint32_t* wide_idx = nullptr;
  uint32_t* planes32 = nullptr;
  uint64_t *planes64 = nullptr, *planes_hi = nullptr, *lab_hi = nullptr, *dmax = nullptr;  // dmax in width-class order
  uint64_t* cmax = nullptr;                                                                // per-class maxima, [(k*R + r)*G + c]
  sigmo_k::MolInfo* minfo = nullptr;                                                       // width-class order
  sigmo_k::ClassLabels cls{{0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu}};                        // the pre-test's class labels
  uint64_t bytes = 0;
  uint64_t cap_G = 0, cap_32 = 0, cap_64 = 0, cap_hi = 0, cap_dmax = 0, cap_lab = 0, cap_cmax = 0;  // allocated elements
  uint64_t generation = 0;  // bumped whenever a rebind changes the data or any buffer address
  uint32_t* order_keys = nullptr;  // 2G: the sort keys and their sorted copy
  uint8_t* sort_tmp = nullptr;
  uint64_t cap_keys = 0, cap_sort = 0;
  ~sigmo_db() {
    bfree(order_keys);
    bfree(sort_tmp);
    bfree(cls_list);
    bfree(pidx);
    bfree(wide_idx);
    bfree(planes32);
    bfree(planes64);
    bfree(planes_hi);
    bfree(lab_hi);
    bfree(dmax);
    bfree(cmax);
    bfree(minfo);
  }
};

struct sigmo_batch {
  sigmo_ctx* ctx = nullptr;
  const sigmo_db* db = nullptr;
  // query side
  uint16_t* tokens = nullptr;
  sigmo_k::QueryInfo* qinfo = nullptr;
  uint64_t* qmax = nullptr;
  uint64_t* qcmax = nullptr;  // per-class maxima, [(g*3 + k)*R + r]
  uint64_t cap_tokens = 0, cap_qinfo = 0, cap_qmax = 0, cap_qcmax = 0;
  // pre-test: passing pairs per molecule (width-class order), their scan and the pair list
  uint32_t *pcnt = nullptr, *poff = nullptr;
  uint64_t cap_pcnt = 0, cap_poff = 0;
  uint64_t* list = nullptr;
  uint64_t cap_list = 0;
  // pair evaluation
  uint64_t* arena = nullptr;
  uint64_t cap_arena = 0;
  unsigned long long* surv = nullptr;
  uint64_t cap_surv = 0;
  uint32_t* pair_off = nullptr;
  uint64_t cap_pair_off = 0;
  uint32_t* mol_count = nullptr;
  uint64_t cap_mc = 0;
  unsigned long long* counters = nullptr;
  // canonical GMCR and join
  uint32_t *offsets = nullptr, *indices = nullptr, *pair_graph = nullptr, *mask_off = nullptr;
  uint64_t cap_off = 0, cap_idx = 0, cap_pg = 0, cap_mo = 0;
  uint64_t* scan_tmp = nullptr;
  uint64_t cap_scan = 0;
  uint64_t* total = nullptr;
  int32_t* counts = nullptr;
  uint64_t cap_counts = 0;
  uint32_t* wide_pairs = nullptr;  // pairs of molecules with > 64 nodes, for the two-word lane join
  uint64_t cap_wide = 0;
  uint32_t* deep_pairs = nullptr;  // pairs of queries deeper than the lane join's first pass holds
  uint64_t cap_deep = 0;
  unsigned long long* run_host = nullptr;  // pinned: a one-sync run's filter counters [0..7] and GMCR total [8]
  uint32_t* req_offsets = nullptr;        // per-request statistics
  unsigned long long* req_out = nullptr;
  sigmo_batch_tuning tune{};               // 0 fields: the built-in heuristics
  // a captured run (sigmo_batch_capture): the graph, its sizes and its pinned readback
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  unsigned long long* graph_host = nullptr;  // [0..7] filter counters, [8] GMCR total, [9..16] join stats
  uint64_t graph_P = 0, graph_nw = 0, graph_db_generation = 0;
  uint32_t graph_Gq = 0;
  int graph_find_first = 0;
  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
  }
  // bgrow, dropping a captured graph whose buffer addresses a reallocation would invalidate
  template <typename T>
  int grow(T** p, uint64_t& cap, uint64_t need, bool slack = false) {
    if (*p && cap >= need) return SIGMO_OK;
    drop_graph();
    return bgrow(p, cap, slack ? need + need / 4 : need);
  }
  uint64_t P = 0;
  uint32_t Gq = 0;
  bool ran = false;
  ~sigmo_batch() {
    bfree(tokens);
    bfree(qinfo);
    bfree(qmax);
    bfree(qcmax);
    bfree(pcnt);
    bfree(poff);
    bfree(list);
    bfree(arena);
    bfree(surv);
    bfree(pair_off);
    bfree(mol_count);
    bfree(counters);
    bfree(offsets);
    bfree(indices);
    bfree(pair_graph);
    bfree(mask_off);
    bfree(scan_tmp);
    bfree(total);
    bfree(counts);
    bfree(wide_pairs);
    bfree(deep_pairs);
    bfree(req_offsets);
    bfree(req_out);
    drop_graph();
    if (graph_host) (void)hipHostFree(graph_host);
    if (run_host) (void)hipHostFree(run_host);
  }
};

sigmo_ctx* sigmo_internal::batch_ctx(const sigmo_batch* b) { return b->ctx; }

namespace {
// The pre-test's class labels: the three most frequent node labels < 16 of the data (ties to the
// smaller label), unused slots where fewer occur.  Any choice gives the same results; frequent
// labels prune most.
int choose_classes(sigmo_db* db, const sigmo_graph* data, hipStream_t s) {
  const sigmo_k::GraphDev g = graph_dev(data);
  unsigned long long* hist = nullptr;
  BATCH_TRY(balloc(&hist, 256));
  std::vector<unsigned long long> h(256, 0);
  hipError_t e = sigmo_k::label_hist(s, g.nl, g.N, hist);
  if (e == hipSuccess) e = hipMemcpyAsync(h.data(), hist, 256 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  bfree(hist);
  if (e != hipSuccess) return sigmo_err::set(SIGMO_ERR_HIP, std::string("label histogram: ") + hipGetErrorString(e));
  for (uint32_t k = 0; k < sigmo_k::kClasses; ++k) {
    uint32_t best = 0xFFFFFFFFu;
    for (uint32_t l = 0; l < 16; ++l)
      if (h[l] && (best == 0xFFFFFFFFu || h[l] > h[best])) best = l;
    db->cls.l[k] = best;
    if (best != 0xFFFFFFFFu) h[best] = 0;
  }
  return SIGMO_OK;
}

// (Re)build the database's class lists and level planes for `data`; buffers are reused when
// the new graph fits them (a streaming chunk), grown otherwise.
int db_fill(sigmo_db* db, const sigmo_graph* data) {
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *st, *sh;
  BATCH_TRY(ctx_info(db->ctx, &device, &s, &cus, &st, &sh));
  const auto& go = graph_host_offsets(data);
  const uint32_t G = static_cast<uint32_t>(go.size() - 1);
  std::vector<uint32_t> widx(G, sigmo_k::kNoWide), pidx(G, 0);
  uint32_t nw = 0, n32 = 0, n64 = 0, count[3] = {0, 0, 0};
  for (uint32_t g = 0; g < G; ++g) {
    const uint32_t n = go[g + 1] - go[g];
    const int c = n <= 32 ? 0 : (n <= 64 ? 1 : 2);
    ++count[c];
    pidx[g] = c == 0 ? n32++ : n64++;
    if (c == 2) widx[g] = nw++;
  }
  const uint64_t R = db->R;
  uint64_t cap_g = db->cap_G;
  BATCH_TRY(bgrow(&db->cls_list, cap_g, G));
  cap_g = db->cap_G;
  BATCH_TRY(bgrow(&db->pidx, cap_g, G));
  cap_g = db->cap_G;
  BATCH_TRY(bgrow(&db->wide_idx, cap_g, G));
  cap_g = db->cap_G;
  BATCH_TRY(bgrow(&db->minfo, cap_g, G));
  db->cap_G = cap_g;
  BATCH_TRY(bgrow(&db->dmax, db->cap_dmax, uint64_t(G) * R));
  BATCH_TRY(bgrow(&db->cmax, db->cap_cmax, uint64_t(G) * R * sigmo_k::kClasses));
  BATCH_TRY(bgrow_slack(&db->planes32, db->cap_32, uint64_t(n32) * R * 64));
  BATCH_TRY(bgrow_slack(&db->planes64, db->cap_64, uint64_t(n64) * R * 64));
  BATCH_TRY(bgrow_slack(&db->planes_hi, db->cap_hi, uint64_t(nw) * R * 64));
  BATCH_TRY(bgrow_slack(&db->lab_hi, db->cap_lab, uint64_t(nw) * 16));
  db->data = data;
  db->generation++;
  db->num_wide = nw;
  for (int c = 0; c < 3; ++c) db->cls_count[c] = count[c];
  db->bytes = G * (12ull + sizeof(sigmo_k::MolInfo)) + (uint64_t(n32) * 256 + (uint64_t(n64) + nw) * 512) * R + nw * 128ull +
              uint64_t(G) * R * 8 * (1 + sigmo_k::kClasses);
  BATCH_TRY(choose_classes(db, data, s));
  BATCH_HIP_TRY(hipMemcpyAsync(db->wide_idx, widx.data(), G * 4ull, hipMemcpyHostToDevice, s));
  BATCH_HIP_TRY(hipMemcpyAsync(db->pidx, pidx.data(), G * 4ull, hipMemcpyHostToDevice, s));
  // the width-class order: molecules sorted by (class, label set, size), stably (by index within a key)
  BATCH_TRY(bgrow_slack(&db->order_keys, db->cap_keys, 2ull * G));
  const sigmo_k::GraphDev gd = graph_dev(data);
  hipLaunchKernelGGL(sigmo_k::k_db_order_keys, dim3((G + 255) / 256), dim3(256), 0, s, G, gd.go, gd.nl, db->order_keys);
  BATCH_HIP_TRY(hipGetLastError());
  size_t tmp_bytes = 0;
  const rocprim::counting_iterator<uint32_t> ids(0u);
  BATCH_HIP_TRY(rocprim::radix_sort_pairs(nullptr, tmp_bytes, db->order_keys, db->order_keys + G, ids, db->cls_list, G, 0u, 25u, s));
  BATCH_TRY(bgrow_slack(&db->sort_tmp, db->cap_sort, std::max<uint64_t>(tmp_bytes, 1)));
  BATCH_HIP_TRY(rocprim::radix_sort_pairs(db->sort_tmp, tmp_bytes, db->order_keys, db->order_keys + G, ids, db->cls_list, G, 0u, 25u, s));
  BATCH_HIP_TRY(hipStreamSynchronize(s));  // the host vectors go out of scope
  BATCH_TRY(sigmo_db_refresh(db));
  BATCH_HIP_TRY(hipStreamSynchronize(s));
  return SIGMO_OK;
}
}  // namespace

namespace sigmo_internal {
int db_rebind(sigmo_db* db, const sigmo_graph* data) {
  if (!db || !data) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  return db_fill(db, data);
}
}  // namespace sigmo_internal

extern "C" {

int sigmo_db_prepare(sigmo_ctx* ctx, const sigmo_graph* data, uint32_t refinement_steps, sigmo_db** out) {
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *st, *sh;
  BATCH_TRY(ctx_info(ctx, &device, &s, &cus, &st, &sh));
  if (!data || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (refinement_steps == 0 || refinement_steps > sigmo_k::kMaxLevels)
    return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "the batched form supports 1..8 refinement levels");
  auto* db = new sigmo_db();
  db->ctx = ctx;
  db->R = refinement_steps;
  const int stt = db_fill(db, data);
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
  BATCH_HIP_TRY(sigmo_k::db_prepare(s, g, db->R, db->cls_list, db->pidx, db->wide_idx, db->planes32, db->planes64, db->planes_hi, db->lab_hi,
                                    db->dmax, db->minfo, db->cls, db->cmax, db->cls_count[0]));
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
  int stt = balloc(&b->counters, 16);
  if (stt == SIGMO_OK) stt = balloc(&b->total, 1);
  if (stt != SIGMO_OK) {
    delete b;
    return stt;
  }
  *out = b;
  return SIGMO_OK;
}

int sigmo_batch_workspace_bytes(const sigmo_batch* b, uint32_t num_query_graphs, uint64_t query_nodes, uint64_t* bytes) {
  if (!b || !bytes) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  const sigmo_db* db = b->db;
  const uint64_t Gd = graph_dev(db->data).G, Gq = num_query_graphs, R = db->R;
  const uint64_t pairs = Gq * Gd, W = (Gq + 63) / 64;
  const double mean_k = Gq ? double(query_nodes) / double(Gq) : 0.0;
  // the arena's first sizing (see sigmo_batch_run) without the per-wavefront slack.  The GMCR
  // arrays (indices, molecule, arena offset, count: 16 B per pair) are sized for every pair on
  // the one-synchronisation path (pair space <= kOneSyncPairs, no molecule over 64 nodes), and
  // for the pairs that survive otherwise, priced at the arena's guess of 1/8 of the pairs (plus
  // 4 B per pair for the list of wide-molecule pairs).  A run whose survivors exceed the guess
  // grows the arena and those arrays to the measured demand; a batch keeps its buffers at their
  // largest size, so a caller admitting by this figure should leave room for that growth.  In
  // list mode (below 256 query graphs by default) the pre-tested pair list takes 8 B per listed
  // pair, priced at its first sizing (a quarter of the pairs; every pair on the one-sync path).
  const uint64_t guess = std::min<uint64_t>(std::max<uint64_t>(pairs / 8, 1 << 16), 1ull << 26);
  const uint64_t arena_words = static_cast<uint64_t>(double(guess) * (mean_k + 1.0) * (db->num_wide ? 2.0 : 1.0));
  const bool one_sync = !db->num_wide && pairs && pairs <= kOneSyncPairs;
  const uint64_t gmcr_pairs = one_sync ? pairs : guess;
  const bool list_mode = b->tune.filter_mode == 2 || (b->tune.filter_mode == 0 && Gq < kListModeGraphs);
  const uint64_t list = !list_mode ? 0 : (one_sync ? pairs : std::min<uint64_t>(pairs, pairs / 4 + 65536));
  *bytes = query_nodes * sigmo_k::token_segment(static_cast<uint32_t>(R)) * 2 + Gq * (sizeof(sigmo_k::QueryInfo) + 8 * R * (1 + sigmo_k::kClasses)) +
           Gd * W * 8 + pairs * 4 + arena_words * 8 + gmcr_pairs * 16 + (db->num_wide ? guess * 4 : 0) + (Gd + 1) * 8 + list * 8;
  return SIGMO_OK;
}

int sigmo_batch_request_stats(sigmo_batch* b, const uint32_t* request_graph_offsets, uint32_t num_requests, sigmo_request_stats* out) {
  if (!b || !b->ran) return sigmo_err::set(SIGMO_ERR_INVALID, "batch has not run");
  if (!request_graph_offsets || !out || num_requests == 0) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (num_requests > sigmo_k::kMaxBatchRequests) return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "at most 256 requests per batch");
  if (request_graph_offsets[0] != 0 || request_graph_offsets[num_requests] != b->Gq)
    return sigmo_err::set(SIGMO_ERR_INVALID, "request graph offsets must cover the batch's query graphs");
  for (uint32_t r = 0; r < num_requests; ++r)
    if (request_graph_offsets[r + 1] < request_graph_offsets[r]) return sigmo_err::set(SIGMO_ERR_INVALID, "request graph offsets must ascend");
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *st, *sh;
  BATCH_TRY(ctx_info(b->ctx, &device, &s, &cus, &st, &sh));
  if (!b->req_offsets) BATCH_TRY(balloc(&b->req_offsets, sigmo_k::kMaxBatchRequests + 1));
  if (!b->req_out) BATCH_TRY(balloc(&b->req_out, 3 * sigmo_k::kMaxBatchRequests));
  BATCH_HIP_TRY(hipMemcpyAsync(b->req_offsets, request_graph_offsets, (num_requests + 1) * 4ull, hipMemcpyHostToDevice, s));
  BATCH_HIP_TRY(sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(b->req_out), 3ull * num_requests, 0));
  if (b->P) {
    const uint64_t blocks = std::min<uint64_t>((b->P + 255) / 256, static_cast<uint64_t>(cus) * 8);
    hipLaunchKernelGGL(sigmo_k::k_request_stats, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, b->P, b->indices, b->pair_graph,
                       b->counts, b->req_offsets, num_requests, b->req_out);
    BATCH_HIP_TRY(hipGetLastError());
  }
  std::vector<unsigned long long> h(3ull * num_requests);
  BATCH_HIP_TRY(hipMemcpyAsync(h.data(), b->req_out, h.size() * 8, hipMemcpyDeviceToHost, s));
  BATCH_HIP_TRY(hipStreamSynchronize(s));
  for (uint32_t r = 0; r < num_requests; ++r) {
    out[r].matched_pairs = h[3 * r];
    out[r].embeddings = h[3 * r + 1];
    out[r].hash = h[3 * r + 2];
  }
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

}  // extern "C"

namespace {

// Launch shape and sizes of one batched run, fixed before any launch.
struct RunShape {
  sigmo_k::GraphDev qd, dd;
  uint32_t R = 0, Gq = 0, Gd = 0, W = 0;
  uint64_t pairs = 0, ksum = 0;
  int cus = 0;
  bool list_mode = false;  // pair evaluation from the pre-tested pair list (k_pair_eval)
  uint32_t lds_bytes[3] = {0, 0, 0};
  uint64_t cls_blocks[3] = {0, 0, 0}, waves_total = 0;
};


int shape_run(const sigmo_batch* b, const sigmo_graph* query, int cus, RunShape& L) {
  const sigmo_db* db = b->db;
  L.qd = graph_dev(query);
  L.dd = graph_dev(db->data);
  const auto& qgo = graph_host_offsets(query);
  L.Gq = L.qd.G;
  L.Gd = L.dd.G;
  L.cus = cus;
  for (uint32_t g = 0; g < L.Gq; ++g) {
    const uint32_t k = qgo[g + 1] - qgo[g];
    if (k < 2) return sigmo_err::set(SIGMO_ERR_INVALID, "query graph " + std::to_string(g) + " has fewer than 2 nodes");
    if (k > SIGMO_MAX_QUERY_NODES) return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "query graph exceeds 32 nodes");
    L.ksum += k;
  }
  L.R = db->R;
  L.W = (L.Gq + 63) / 64;
  L.pairs = uint64_t(L.Gq) * L.Gd;
  L.list_mode = b->tune.filter_mode == 2 || (b->tune.filter_mode == 0 && L.Gq < kListModeGraphs);
  // one persistent launch per width class, as many 4-wave blocks as registers and LDS keep
  // resident (or the tuned count, capped by residency)
  const void* kern[3] = {reinterpret_cast<const void*>(sigmo_k::k_pair_survive<1>), reinterpret_cast<const void*>(sigmo_k::k_pair_survive<2>),
                         reinterpret_cast<const void*>(sigmo_k::k_pair_survive<4>)};
  const void* kern_list[3] = {reinterpret_cast<const void*>(sigmo_k::k_pair_eval<1>), reinterpret_cast<const void*>(sigmo_k::k_pair_eval<2>),
                              reinterpret_cast<const void*>(sigmo_k::k_pair_eval<4>)};
  for (int c = 0; c < 3; ++c) L.lds_bytes[c] = L.list_mode ? 0u : sigmo_k::survive_lds_bytes(L.R, c == 0 ? 1 : (c == 1 ? 2 : 4));
  for (int c = 0; c < 3; ++c) {
    if (!db->cls_count[c]) continue;
    if (L.list_mode) {
      int resident = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kern_list[c], 256, 0) != hipSuccess) {
        (void)hipGetLastError();
        resident = 4;
      }
      // two blocks per CU unless tuned: each lane reads its pair's molecule planes from device
      // memory, and fewer molecules in flight per XCD keep more of them in L2 (10 M molecules x 25
      // query graphs: 17.6 -> 16.2 ms with 2 blocks instead of 8; x 64: 24.2 -> 22.4 ms)
      uint64_t per_cu = std::max<uint64_t>(1, std::min<uint64_t>(2, static_cast<uint64_t>(resident)));
      if (b->tune.filter_blocks_per_cu) per_cu = std::min<uint64_t>(static_cast<uint64_t>(std::max(1, resident)), b->tune.filter_blocks_per_cu);
      L.cls_blocks[c] = uint64_t(cus) * per_cu;
      L.waves_total += L.cls_blocks[c] * sigmo_k::kSurviveWaves;
      continue;
    }
    if (L.lds_bytes[c] > 64 * 1024) return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "molecule slots exceed 64 KB of LDS per block");
    int resident = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&resident, kern[c], 256, L.lds_bytes[c]) != hipSuccess) {
      (void)hipGetLastError();
      resident = 2;
    }
    uint64_t per_cu = std::max<uint64_t>(1, std::min<uint64_t>(8, static_cast<uint64_t>(resident)));
    if (b->tune.filter_blocks_per_cu) per_cu = std::min<uint64_t>(per_cu, b->tune.filter_blocks_per_cu);
    L.cls_blocks[c] = std::max<uint64_t>(1, std::min<uint64_t>(uint64_t(cus) * per_cu, (uint64_t(db->cls_count[c]) + 3) / 4));
    L.waves_total += L.cls_blocks[c] * sigmo_k::kSurviveWaves;
  }
  return SIGMO_OK;
}

// Pair evaluation: the pre-test over every pair (its bits in surv), the scan of its per-molecule
// counts (total in counters[7]) and, per width class, the evaluation of the passing pairs --
// molecule by molecule from LDS-staged planes (k_pair_survive) or, in list mode, pair by pair
// from the pre-tested pair list (k_pair_expand, then k_pair_eval).  Counters are cleared by a
// kernel.  Deaths clear their pre-test bit, so surv ends as the GMCR survival bitmask.
hipError_t launch_filter(sigmo_batch* b, const RunShape& L, hipStream_t s) {
  const sigmo_db* db = b->db;
  hipError_t e = sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(b->counters), 16, 0);
  if (e != hipSuccess || !L.Gd || !L.Gq) return e;
  sigmo_k::PretestArgs p;
  p.R = L.R;
  p.Gq = L.Gq;
  p.W = L.W;
  p.G = L.Gd;
  p.minfo = db->minfo;
  p.dmax = db->dmax;
  p.cmax = db->cmax;
  p.qinfo = b->qinfo;
  p.qmax = b->qmax;
  p.qcmax = b->qcmax;
  p.surv = b->surv;
  p.pcnt = b->pcnt;
  const dim3 pgrid((L.Gd + 255) / 256);
  switch (L.R) {  // one instantiation per level count keeps the molecule's maxima in registers
    case 1: hipLaunchKernelGGL(sigmo_k::k_pair_pretest<1>, pgrid, dim3(256), 0, s, p); break;
    case 2: hipLaunchKernelGGL(sigmo_k::k_pair_pretest<2>, pgrid, dim3(256), 0, s, p); break;
    case 3: hipLaunchKernelGGL(sigmo_k::k_pair_pretest<3>, pgrid, dim3(256), 0, s, p); break;
    case 4: hipLaunchKernelGGL(sigmo_k::k_pair_pretest<4>, pgrid, dim3(256), 0, s, p); break;
    case 5: hipLaunchKernelGGL(sigmo_k::k_pair_pretest<5>, pgrid, dim3(256), 0, s, p); break;
    case 6: hipLaunchKernelGGL(sigmo_k::k_pair_pretest<6>, pgrid, dim3(256), 0, s, p); break;
    case 7: hipLaunchKernelGGL(sigmo_k::k_pair_pretest<7>, pgrid, dim3(256), 0, s, p); break;
    default: hipLaunchKernelGGL(sigmo_k::k_pair_pretest<8>, pgrid, dim3(256), 0, s, p); break;
  }
  e = hipGetLastError();
  if (e == hipSuccess) e = sigmo_k::exclusive_scan(s, b->pcnt, L.Gd, b->poff, b->scan_tmp, reinterpret_cast<uint64_t*>(b->counters + 7));
  if (e == hipSuccess && L.list_mode) {
    hipLaunchKernelGGL(sigmo_k::k_pair_expand, dim3((L.Gd + 255) / 256), dim3(256), 0, s, L.Gd, L.W, db->minfo, b->surv, b->poff, b->list,
                       b->cap_list, b->counters);
    e = hipGetLastError();
  }
  uint64_t first = 0;
  for (int c = 0; c < 3 && e == hipSuccess; ++c) {
    const uint32_t count = db->cls_count[c];
    const uint64_t at = first;
    first += count;
    if (!count) continue;
    const int nw = c == 0 ? 1 : (c == 1 ? 2 : 4);
    e = sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(b->counters), 1, 0);  // each launch starts its own cursor
    if (e != hipSuccess) break;
    const dim3 grid(static_cast<unsigned>(L.cls_blocks[c]));
    if (L.list_mode) {
      sigmo_k::EvalArgs a;
      a.d = L.dd;
      a.R = L.R;
      a.Gq = L.Gq;
      a.W = L.W;
      a.minfo = db->minfo;
      a.planes32 = db->planes32;
      a.planes64 = db->planes64;
      a.planes_hi = db->planes_hi;
      a.lab_hi = db->lab_hi;
      a.tok64 = reinterpret_cast<const uint64_t*>(b->tokens);
      a.qinfo = b->qinfo;
      a.list = b->list;
      a.poff = b->poff;
      a.list_cap = b->cap_list;
      a.first = static_cast<uint32_t>(at);
      a.count = count;
      a.chunk = 64u * (b->tune.filter_claim ? b->tune.filter_claim : 2u);
      a.arena = b->arena;
      a.arena_cap = b->cap_arena;
      a.surv = b->surv;
      a.pair_off = b->pair_off;
      a.counters = b->counters;
      if (nw == 1) hipLaunchKernelGGL(sigmo_k::k_pair_eval<1>, grid, dim3(256), 0, s, a);
      else if (nw == 2) hipLaunchKernelGGL(sigmo_k::k_pair_eval<2>, grid, dim3(256), 0, s, a);
      else hipLaunchKernelGGL(sigmo_k::k_pair_eval<4>, grid, dim3(256), 0, s, a);
    } else {
      sigmo_k::SurviveArgs a;
      a.q = L.qd;
      a.d = L.dd;
      a.R = L.R;
      a.Gq = L.Gq;
      a.W = L.W;
      a.count = count;
      a.minfo = db->minfo + at;
      a.pcnt = b->pcnt + at;
      a.planes32 = db->planes32;
      a.planes64 = db->planes64;
      a.planes_hi = db->planes_hi;
      a.lab_hi = db->lab_hi;
      a.tok64 = reinterpret_cast<const uint64_t*>(b->tokens);
      a.qinfo = b->qinfo;
      a.arena = b->arena;
      a.arena_cap = b->cap_arena;
      a.surv = b->surv;
      a.pair_off = b->pair_off;
      a.counters = b->counters;
      const uint64_t auto_claim = std::max<uint64_t>(1, std::min<uint64_t>(16, count / (L.cls_blocks[c] * sigmo_k::kSurviveWaves * 4)));
      a.claim = static_cast<uint32_t>(b->tune.filter_claim ? std::min<uint32_t>(b->tune.filter_claim, 64) : auto_claim);
      if (nw == 1) hipLaunchKernelGGL(sigmo_k::k_pair_survive<1>, grid, dim3(256), L.lds_bytes[0], s, a);
      else if (nw == 2) hipLaunchKernelGGL(sigmo_k::k_pair_survive<2>, grid, dim3(256), L.lds_bytes[1], s, a);
      else hipLaunchKernelGGL(sigmo_k::k_pair_survive<4>, grid, dim3(256), L.lds_bytes[2], s, a);
    }
    e = hipGetLastError();
  }
  return e;
}

// GMCR: survivors per molecule and their exclusive scan (total in b->total).
hipError_t launch_count_scan(sigmo_batch* b, const RunShape& L, hipStream_t s) {
  if (L.Gd) {
    hipLaunchKernelGGL(sigmo_k::k_surv_count, dim3((L.Gd + 255) / 256), dim3(256), 0, s, L.Gd, L.W, b->surv, b->mol_count);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return sigmo_k::exclusive_scan(s, b->mol_count, L.Gd, b->offsets, b->scan_tmp, b->total);
}

hipError_t launch_scatter(sigmo_batch* b, const RunShape& L, hipStream_t s) {
  if (!L.Gd) return hipSuccess;
  const uint64_t cap = std::min(b->cap_idx, std::min(b->cap_pg, b->cap_mo));
  if (L.W == 1) {  // at most 64 query graphs: a thread per molecule
    const dim3 grid(static_cast<unsigned>((uint64_t(L.Gd) + 255) / 256));
    hipLaunchKernelGGL(sigmo_k::k_surv_scatter_thread, grid, dim3(256), 0, s, L.Gd, L.Gq, L.W, b->surv, b->offsets, b->pair_off, b->indices,
                       b->pair_graph, b->mask_off, cap);
    return hipGetLastError();
  }
  const uint64_t blocks = (static_cast<uint64_t>(L.Gd) * sigmo_k::kWave + 255) / 256;  // a wavefront per molecule
  hipLaunchKernelGGL(sigmo_k::k_surv_scatter, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, L.Gd, L.Gq, L.W, b->surv, b->offsets,
                     b->pair_off, b->indices, b->pair_graph, b->mask_off, cap);
  return hipGetLastError();
}

// Lane join over P pairs (molecules of <= 64 nodes; the others are listed in wide_pairs): a
// shallow first pass and a full-depth pass over the pairs of deeper queries (deep_pairs).
hipError_t launch_join(sigmo_batch* b, const RunShape& L, hipStream_t s, unsigned long long* stats, uint64_t P, int find_first,
                       const unsigned long long* P_dev = nullptr) {
  if (!P) return hipSuccess;
  const uint32_t per_cu = b->tune.join_blocks_per_cu ? std::min<uint32_t>(b->tune.join_blocks_per_cu, 4) : 4;
  return sigmo_k::launch_join_lanes(s, L.qd, L.dd, sigmo_k::ArenaMasks{b->arena, b->mask_off, b->counters + 3}, b->indices, b->pair_graph, P,
                                    find_first, b->counts, stats, b->wide_pairs, L.cus, P_dev, b->deep_pairs, per_cu, b->tune.join_chunk);
}

// Lane join over the nw listed pairs of 65..128-node molecules (two-word positions).
hipError_t launch_join_wide(sigmo_batch* b, const RunShape& L, hipStream_t s, unsigned long long* stats, uint64_t nw, int find_first,
                            const unsigned long long* nw_dev = nullptr) {
  return sigmo_k::launch_join_wide(s, L.qd, L.dd, sigmo_k::ArenaMasks{b->arena, b->mask_off, b->counters + 3}, b->indices, b->pair_graph,
                                   b->wide_pairs, nw, find_first, b->counts, stats, L.cus, nw_dev, b->deep_pairs);
}

// Buffers of a run that do not depend on its results.
int size_fixed_buffers(sigmo_batch* b, const RunShape& L) {
  const uint64_t seg = sigmo_k::token_segment(L.R);
  if (L.qd.N * seg >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "query batch too large for one batched run");
  BATCH_TRY(b->grow(&b->tokens, b->cap_tokens, L.qd.N * seg));
  BATCH_TRY(b->grow(&b->qinfo, b->cap_qinfo, L.Gq));
  BATCH_TRY(b->grow(&b->qmax, b->cap_qmax, uint64_t(L.Gq) * L.R));
  BATCH_TRY(b->grow(&b->qcmax, b->cap_qcmax, uint64_t(L.Gq) * L.R * sigmo_k::kClasses));
  BATCH_TRY(b->grow(&b->pcnt, b->cap_pcnt, uint64_t(L.Gd) + 1));
  BATCH_TRY(b->grow(&b->poff, b->cap_poff, uint64_t(L.Gd) + 1));
  BATCH_TRY(b->grow(&b->surv, b->cap_surv, uint64_t(L.Gd) * L.W));
  BATCH_TRY(b->grow(&b->pair_off, b->cap_pair_off, L.pairs));
  BATCH_TRY(b->grow(&b->mol_count, b->cap_mc, uint64_t(L.Gd) + 1));
  BATCH_TRY(b->grow(&b->offsets, b->cap_off, uint64_t(L.Gd) + 1));
  BATCH_TRY(b->grow(&b->scan_tmp, b->cap_scan, (uint64_t(L.Gd) + 1023) / 1024 + 1));
  return SIGMO_OK;
}

}  // namespace

extern "C" {

int sigmo_batch_run(sigmo_batch* b, const sigmo_graph* query, int find_first, sigmo_batch_result* res) {
  if (!b || !query || !res) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *stats, *stats_host;
  BATCH_TRY(ctx_info(b->ctx, &device, &s, &cus, &stats, &stats_host));
  std::memset(res, 0, sizeof(*res));
  const sigmo_db* db = b->db;
  RunShape L;
  BATCH_TRY(shape_run(b, query, cus, L));
  const uint32_t R = L.R, Gq = L.Gq, Gd = L.Gd, W = L.W;
  // From here on the result arrays are overwritten: until this run succeeds, no earlier run's
  // results may be read (sigmo_batch_pairs, sigmo_batch_request_stats), so `ran` is cleared and
  // set again only on success.  `sized` remembers that an earlier run sized the arena.
  const bool sized = b->ran;
  b->ran = false;
  BATCH_TRY(size_fixed_buffers(b, L));
  {
    void* tok = nullptr;
    BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_SIG_QUERY, -1, 0.0, true, &tok));
    BATCH_HIP_TRY(sigmo_k::query_prepare(s, L.qd, R, b->tokens, b->qinfo, b->qmax, db->cls, b->qcmax));
    BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
  }
  // arena capacity: start from 1/8 of the pairs (at most 2^26) plus every wavefront's open chunk.
  // On overflow the counter still holds the exact demand, so one replay with that size
  // suffices; the capacity is kept for later runs.
  const double mean_k = Gq ? double(L.ksum) / Gq : 0.0;
  // every wavefront may leave its chunk's tail and each lane its sub-chunk's tail unused
  const uint64_t slack_arena = L.waves_total * 2 * sigmo_k::kWaveChunk;
  const uint64_t guess = std::min<uint64_t>(std::max<uint64_t>(L.pairs / 8, 1 << 16), 1ull << 26);
  uint64_t want_arena = static_cast<uint64_t>(double(guess) * (mean_k + 1.0) * (db->num_wide ? 2.0 : 1.0)) + slack_arena;
  want_arena = std::max(want_arena, b->cap_arena);
  // One synchronisation per run once an earlier run has sized the arena: with no molecule of
  // more than 64 nodes and a pair space of at most kOneSyncPairs, the GMCR arrays are sized for
  // every pair, the join reads the pair count on the device, and the filter's counters and the
  // total are read after the join.  An arena overflow (data that changed) falls back to the
  // sized path below, which repeats the run.
  const uint64_t bound = uint64_t(Gq) * Gd;
  if (sized && !db->num_wide && bound && bound <= kOneSyncPairs && b->arena && b->cap_arena >= want_arena) {
    if (!b->run_host) BATCH_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&b->run_host), 16 * sizeof(unsigned long long)));
    BATCH_TRY(b->grow(&b->indices, b->cap_idx, bound));
    BATCH_TRY(b->grow(&b->pair_graph, b->cap_pg, bound));
    BATCH_TRY(b->grow(&b->mask_off, b->cap_mo, bound));
    BATCH_TRY(b->grow(&b->counts, b->cap_counts, bound));
    BATCH_TRY(b->grow(&b->deep_pairs, b->cap_deep, bound));
    if (L.list_mode) BATCH_TRY(b->grow(&b->list, b->cap_list, bound));
    const double reads = (1.0 + 8.0 * R) * double(L.dd.N) + double(L.qd.N) * (1.0 + 2.0 * 16.0 * R) + double(Gq) * (4.0 + 8.0 * R);
    void* tok = nullptr;
    BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_FILTER, SIGMO_KERNEL_PAIR_FILTER, reads, true, &tok));
    BATCH_HIP_TRY(launch_filter(b, L, s));
    BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
    res->one_sync = 1;
    BATCH_HIP_TRY(hipMemcpyAsync(b->run_host, b->counters, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_GMCR, -1, 0.0, true, &tok));
    BATCH_HIP_TRY(launch_count_scan(b, L, s));
    BATCH_HIP_TRY(launch_scatter(b, L, s));
    BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
    BATCH_HIP_TRY(hipMemcpyAsync(b->run_host + 8, b->total, sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_JOIN, SIGMO_KERNEL_JOIN, 0.0, true, &tok));
    BATCH_HIP_TRY(sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(stats), sigmo_k::kJoinStatWords, 0));
    BATCH_HIP_TRY(launch_join(b, L, s, stats, bound, find_first, reinterpret_cast<const unsigned long long*>(b->total)));
    BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
    BATCH_HIP_TRY(hipMemcpyAsync(stats_host, stats, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
    BATCH_HIP_TRY(hipStreamSynchronize(s));
    const unsigned long long* cnt = b->run_host;
    if (!cnt[3]) {
      const uint64_t P = b->run_host[8];
      if (P >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "GMCR pair count exceeds u32");
      if (stats_host[3]) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "a per-pair match count exceeded int32");
      ctx_add_bytes(b->ctx, SIGMO_KERNEL_PAIR_FILTER, double(cnt[1]) * 8.0 + double(Gd) * W * 8.0);
      res->filter_wave_steps = cnt[4];
      res->filter_lane_steps = cnt[5];
      res->filter_scans = cnt[7];
      b->P = P;
      b->Gq = Gq;
      b->ran = true;
      res->num_matches = stats_host[0];
      res->matched_pairs = stats_host[1];
      res->gmcr_pairs = P;
      res->dfs_steps = stats_host[4];
      res->join_wave_steps = stats_host[7];
      return SIGMO_OK;
    }
    want_arena = std::max<uint64_t>(want_arena + 1, cnt[1] + cnt[1] / 8 + slack_arena);  // overflowed: size and repeat
    res->arena_repeats = 1;
  }
  // the pre-tested pair list (list mode): a quarter of the pairs at first, then the measured demand
  uint64_t want_list = L.list_mode ? std::max<uint64_t>(b->cap_list, std::min<uint64_t>(L.pairs, L.pairs / 4 + 65536)) : 0;
  for (int attempt = 0; attempt < 4; ++attempt) {
    BATCH_TRY(b->grow(&b->arena, b->cap_arena, want_arena));
    if (L.list_mode) BATCH_TRY(b->grow(&b->list, b->cap_list, want_list));
    if (b->cap_arena >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "candidate arena exceeds 2^32 words");
    // algorithmic reads: the data side once per run -- a label byte and R level signatures per
    // data node, (1 + 8R) D -- plus the query side; writes are added once they are known.
    const double reads = (1.0 + 8.0 * R) * double(L.dd.N) + double(L.qd.N) * (1.0 + 2.0 * 16.0 * R) + double(Gq) * (4.0 + 8.0 * R);
    void* tok = nullptr;
    BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_FILTER, SIGMO_KERNEL_PAIR_FILTER, reads, true, &tok));
    BATCH_HIP_TRY(launch_filter(b, L, s));
    BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
    unsigned long long cnt[8];
    BATCH_HIP_TRY(hipMemcpyAsync(cnt, b->counters, sizeof(cnt), hipMemcpyDeviceToHost, s));
    BATCH_HIP_TRY(hipStreamSynchronize(s));
    res->filter_wave_steps = cnt[4];
    res->filter_lane_steps = cnt[5];
    res->filter_scans = cnt[7];
    if (cnt[7] >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "pre-tested pairs of one run exceed 2^32");
    // writes: candidate masks, the survival bits and arena offsets
    ctx_add_bytes(b->ctx, SIGMO_KERNEL_PAIR_FILTER, double(cnt[1]) * 8.0 + double(Gd) * W * 8.0);
    if (!cnt[3]) break;
    if (attempt == 3) return sigmo_err::set(SIGMO_ERR_NOMEM, "candidate arena or pair list overflowed repeatedly");
    res->arena_repeats++;
    // replay with the demand just measured; the claim order (and so the space left unused at
    // sub-chunk ends) varies between runs, hence the 1/8 margin.  The list's demand is exact.
    if (L.list_mode && cnt[7] > b->cap_list) want_list = cnt[7];
    if (cnt[1] > b->cap_arena) want_arena = std::max<uint64_t>(want_arena + 1, cnt[1] + cnt[1] / 8 + slack_arena);
  }
  // canonical GMCR: count survivors per molecule, scan, scatter in (molecule, query graph) order
  void* tok = nullptr;
  BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_GMCR, -1, 0.0, true, &tok));
  BATCH_HIP_TRY(launch_count_scan(b, L, s));
  uint64_t P = 0;
  BATCH_HIP_TRY(hipMemcpyAsync(&P, b->total, 8, hipMemcpyDeviceToHost, s));
  BATCH_HIP_TRY(hipStreamSynchronize(s));
  if (P >= (1ull << 32)) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "GMCR pair count exceeds u32");
  // result-sized arrays regrow with a quarter to spare: each hipFree waits for the whole device
  BATCH_TRY(b->grow(&b->indices, b->cap_idx, P, true));
  BATCH_TRY(b->grow(&b->pair_graph, b->cap_pg, P, true));
  BATCH_TRY(b->grow(&b->mask_off, b->cap_mo, P, true));
  BATCH_TRY(b->grow(&b->counts, b->cap_counts, P, true));
  BATCH_TRY(b->grow(&b->wide_pairs, b->cap_wide, db->num_wide ? P : 1, true));
  BATCH_TRY(b->grow(&b->deep_pairs, b->cap_deep, P, true));
  BATCH_HIP_TRY(launch_scatter(b, L, s));
  BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
  // join: lane per pair; pairs of molecules with > 64 nodes are listed for the two-word kernel
  BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_JOIN, SIGMO_KERNEL_JOIN, 0.0, true, &tok));
  BATCH_HIP_TRY(sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(stats), sigmo_k::kJoinStatWords, 0));
  BATCH_HIP_TRY(launch_join(b, L, s, stats, P, find_first));
  if (P && db->num_wide) {
    unsigned long long nw = 0;
    BATCH_HIP_TRY(hipMemcpyAsync(&nw, stats + 5, sizeof(nw), hipMemcpyDeviceToHost, s));
    BATCH_HIP_TRY(hipStreamSynchronize(s));
    BATCH_HIP_TRY(launch_join_wide(b, L, s, stats, nw, find_first));
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
  res->join_wave_steps = stats_host[7];
  return SIGMO_OK;
}

int sigmo_batch_set_tuning(sigmo_batch* b, const sigmo_batch_tuning* t) {
  if (!b || !t) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (t->join_blocks_per_cu > 4) return sigmo_err::set(SIGMO_ERR_INVALID, "join_blocks_per_cu must be 0..4 (LDS keeps 4 resident)");
  if (t->filter_claim > 64) return sigmo_err::set(SIGMO_ERR_INVALID, "filter_claim must be 0..64");
  if (t->filter_mode > 2) return sigmo_err::set(SIGMO_ERR_INVALID, "filter_mode must be 0 (automatic), 1 (molecule) or 2 (list)");
  b->tune = *t;
  return SIGMO_OK;
}

int sigmo_batch_get_tuning(const sigmo_batch* b, sigmo_batch_tuning* t) {
  if (!b || !t) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  *t = b->tune;
  return SIGMO_OK;
}

int sigmo_batch_autotune(sigmo_batch* b, const sigmo_graph* query, int find_first, uint32_t repeats, sigmo_batch_tuning* chosen,
                         double* best_ms) {
  if (!b || !query) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (repeats == 0) repeats = 2;
  sigmo_batch_result r;
  BATCH_TRY(sigmo_batch_run(b, query, find_first, &r));  // sizes every buffer
  auto time_with = [&](const sigmo_batch_tuning& t, double* ms) -> int {
    b->tune = t;
    double best = 1e300;
    for (uint32_t i = 0; i < repeats; ++i) {
      const auto t0 = std::chrono::steady_clock::now();
      BATCH_TRY(sigmo_batch_run(b, query, find_first, &r));
      best = std::min(best, std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    }
    *ms = best;
    return SIGMO_OK;
  };
  sigmo_batch_tuning best_t{};
  double best = 0;
  BATCH_TRY(time_with(best_t, &best));
  // coordinate search: pair evaluation mode and shape, then the join's
  const uint32_t filter_modes[] = {1, 2};
  const uint32_t filter_blocks[] = {1, 2, 4, 0};
  const uint32_t filter_claims[] = {2, 8, 32, 0};
  const uint32_t join_blocks[] = {2, 3, 4};
  const uint32_t join_chunks[] = {16, 64, 256, 0};
  auto sweep = [&](uint32_t sigmo_batch_tuning::*field, const uint32_t* vals, size_t n) -> int {
    for (size_t i = 0; i < n; ++i) {
      sigmo_batch_tuning t = best_t;
      t.*field = vals[i];
      if (t.*field == best_t.*field) continue;
      double ms = 0;
      BATCH_TRY(time_with(t, &ms));
      if (ms < best * 0.98) {  // a 2% margin against timing noise
        best = ms;
        best_t = t;
      }
    }
    return SIGMO_OK;
  };
  BATCH_TRY(sweep(&sigmo_batch_tuning::filter_mode, filter_modes, 2));
  BATCH_TRY(sweep(&sigmo_batch_tuning::filter_blocks_per_cu, filter_blocks, 4));
  BATCH_TRY(sweep(&sigmo_batch_tuning::filter_claim, filter_claims, 4));
  BATCH_TRY(sweep(&sigmo_batch_tuning::join_blocks_per_cu, join_blocks, 3));
  BATCH_TRY(sweep(&sigmo_batch_tuning::join_chunk, join_chunks, 4));
  b->tune = best_t;
  if (chosen) *chosen = best_t;
  if (best_ms) *best_ms = best;
  return SIGMO_OK;
}

int sigmo_batch_capture(sigmo_batch* b, const sigmo_graph* query, int find_first, int refresh_db) {
  if (!b || !query) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *stats, *stats_host;
  BATCH_TRY(ctx_info(b->ctx, &device, &s, &cus, &stats, &stats_host));
  // a plain run first: it sizes the arena and the GMCR arrays and gives the pair counts
  sigmo_batch_result r;
  BATCH_TRY(sigmo_batch_run(b, query, find_first, &r));
  const uint64_t nw = stats_host[5];
  RunShape L;
  BATCH_TRY(shape_run(b, query, cus, L));
  if (!b->graph_host) BATCH_HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(&b->graph_host), 32 * sizeof(unsigned long long)));
  b->drop_graph();
  const sigmo_db* db = b->db;
  BATCH_HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  hipError_t e = hipSuccess;
  if (refresh_db)
    e = sigmo_k::db_prepare(s, L.dd, db->R, db->cls_list, db->pidx, db->wide_idx, db->planes32, db->planes64, db->planes_hi, db->lab_hi,
                            db->dmax, db->minfo, db->cls, db->cmax, db->cls_count[0]);
  if (e == hipSuccess) e = sigmo_k::query_prepare(s, L.qd, L.R, b->tokens, b->qinfo, b->qmax, db->cls, b->qcmax);
  if (e == hipSuccess) e = launch_filter(b, L, s);
  if (e == hipSuccess) e = launch_count_scan(b, L, s);
  if (e == hipSuccess) e = launch_scatter(b, L, s);
  if (e == hipSuccess) e = sigmo_k::fill_u64(s, reinterpret_cast<uint64_t*>(stats), sigmo_k::kJoinStatWords, 0);
  // the joins read this replay's pair counts on the device (bounded by the captured sizes), so a
  // replay on changed data reads only entries the replay itself wrote
  if (e == hipSuccess) e = launch_join(b, L, s, stats, b->P, find_first, reinterpret_cast<const unsigned long long*>(b->total));
  if (e == hipSuccess && db->num_wide) e = launch_join_wide(b, L, s, stats, nw, find_first, stats + 5);
  if (e == hipSuccess) e = hipMemcpyAsync(b->graph_host, b->counters, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(b->graph_host + 8, b->total, sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipMemcpyAsync(b->graph_host + 9, stats, 8 * sizeof(unsigned long long), hipMemcpyDeviceToHost, s);
  hipGraph_t g = nullptr;
  const hipError_t end = hipStreamEndCapture(s, &g);
  if (e != hipSuccess || end != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return sigmo_err::set(SIGMO_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e != hipSuccess ? e : end));
  }
  b->graph = g;
  BATCH_HIP_TRY(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
  b->graph_P = b->P;
  b->graph_nw = nw;
  b->graph_Gq = L.Gq;
  b->graph_find_first = find_first;
  b->graph_db_generation = db->generation;
  return SIGMO_OK;
}

int sigmo_batch_replay(sigmo_batch* b, sigmo_batch_result* res) {
  if (!b || !res) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (!b->graph_exec)
    return sigmo_err::set(SIGMO_ERR_INVALID, "no captured graph (never captured, or a later run reallocated its buffers): capture again");
  if (b->db->generation != b->graph_db_generation)
    return sigmo_err::set(SIGMO_ERR_INVALID, "the database was rebound after the capture: capture again");
  int device = 0, cus = 0;
  hipStream_t s;
  unsigned long long *stats, *stats_host;
  BATCH_TRY(ctx_info(b->ctx, &device, &s, &cus, &stats, &stats_host));
  std::memset(res, 0, sizeof(*res));
  b->ran = false;  // the replay overwrites the result arrays
  void* tok = nullptr;
  BATCH_TRY(ctx_timed(b->ctx, SIGMO_PHASE_FILTER, -1, 0.0, true, &tok));  // the whole graph, as one timed region
  BATCH_HIP_TRY(hipGraphLaunch(b->graph_exec, s));
  BATCH_TRY(ctx_timed(b->ctx, 0, 0, 0.0, false, &tok));
  BATCH_HIP_TRY(hipStreamSynchronize(s));
  const unsigned long long* h = b->graph_host;
  if (h[3] || h[8] != b->graph_P || (b->db->num_wide && h[9 + 5] != b->graph_nw)) {
    return sigmo_err::set(SIGMO_ERR_INVALID, "the captured graph no longer matches the data (arena or pair counts changed): capture again");
  }
  if (h[9 + 3]) return sigmo_err::set(SIGMO_ERR_OVERFLOW, "a per-pair match count exceeded int32");
  b->P = b->graph_P;
  b->Gq = b->graph_Gq;
  b->ran = true;
  res->num_matches = h[9 + 0];
  res->matched_pairs = h[9 + 1];
  res->gmcr_pairs = b->graph_P;
  res->dfs_steps = h[9 + 4];
  res->join_wave_steps = h[9 + 7];
  res->filter_wave_steps = h[4];
  res->filter_lane_steps = h[5];
  res->filter_scans = h[7];
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
