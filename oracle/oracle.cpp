// oracle.cpp -- CPU restatement of SIGMo's hot path.  TEST INFRASTRUCTURE ONLY.
//
// This file is the parity checker for the gfx950 implementation.  Only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only to check or
// time; no product code path calls it.
//
// It restates the reference SYCL kernels literally, in the reference's own data layout
// (uint32 candidate words, row stride ceil(D/32)), one loop per reference loop, with the
// conventions that make the reference's undefined behaviour definite:
//   * every variable shift count is masked to width-1 (uint64 & 63, uint32 & 31);
//   * Bitset<uint64_t> indices pass through uint8_t and subtract the offset in uint32;
//   * PowerGraph reachables start at zero;
//   * refineCandidates is applied per row (no cross-row race);
//   * GMCR segments list query graphs in ascending id; query graphs with < 2 nodes are
//     rejected;
//   * the join starts every (query, molecule) pair from fresh state with every mapping[]
//     entry equal to the molecule's first node id (local id 0).
// Reference files are cited per function (paths relative to the SIGMo repository).
//
// Planted-mistake switches (`mutation` arguments) exist so tests can show that the
// pinning checks reject plausible errors; 0 is the faithful restatement.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

using u8 = uint8_t;
using u16 = uint16_t;
using u32 = uint32_t;
using u64 = uint64_t;

constexpr u8 kWildcard = 255;  // types.hpp:22
constexpr int kMaxQueryNodes = 33;  // reference: 30 (isomorphism.hpp:288); k+1 frames needed

inline int ctz64(u64 x) { return x ? __builtin_ctzll(x) : 64; }
inline int ctz32(u32 x) { return x ? __builtin_ctz(x) : 32; }

// ---- SignatureDevice, signature.hpp:23-54 (Bits = 4, 16 labels) -------------------------
inline u8 sig_get(u64 sig, u8 label) {
  if (label < 16) return (sig >> (label * 4)) & 15;  // :39-42
  return 0;
}
inline void sig_set(u64& sig, u8 label, u8 count) {  // :32-37
  if (label < 16 && count < 16) {
    sig &= ~(u64(15) << (label * 4));
    sig |= u64(count) << (label * 4);
  }
}
inline void sig_inc(u64& sig, u8 label, u8 add = 1) {  // :44-51
  if (label < 16) {
    u8 count = sig_get(sig, label);
    if (count < 15) sig_set(sig, label, static_cast<u8>(count + add));
  }
}

// ---- Bitset<uint64_t>, utils.hpp:64-96 -------------------------------------------------
struct Bitset64 {
  u64 data = 0;
  u32 offset = 0;
  Bitset64() = default;
  explicit Bitset64(u32 off) : offset(off) {}
  static u32 amount(u8 idx, u32 off) { return (static_cast<u32>(idx) - off) & 63u; }  // A1/A2
  void set(u8 idx) { data |= u64(1) << amount(idx, offset); }
  void unset(u8 idx) { data &= ~(u64(1) << amount(idx, offset)); }
  bool get(u8 idx) const { return (data & (u64(1) << amount(idx, offset))) != 0; }
  void clear() { data = 0; }
  void merge(const Bitset64& o) { data |= o.data; }
  void difference(const Bitset64& o) { data &= ~o.data; }
  bool empty() const { return data == 0; }
  u16 size() const { return static_cast<u16>(__builtin_popcountll(data)); }
  u16 getSetBit(u8 idx) const {  // :85-95
    idx = static_cast<u8>(idx - offset);
    u16 ret = 0xFFFF;
    u64 tmp = data;
    for (int count = 0; count <= idx; ++count) {
      ret = static_cast<u16>(ctz64(tmp));
      tmp &= ~(u64(1) << (ret & 63));
    }
    return ret;
  }
};

// ---- DeviceBatchedCSRGraph accessors, graph.hpp:62-112 ---------------------------------
struct Graph {
  u32 G;
  const u32* go;  // graph_offsets
  const u32* ro;  // row_offsets
  const u32* ci;  // column_indices
  const u8* nl;   // node_labels
  u64 N() const { return go[G]; }
  bool isNeighbor(u32 a, u32 b) const {  // :72-77
    for (u32 i = ro[a]; i < ro[a + 1]; ++i)
      if (ci[i] == b) return true;
    return false;
  }
  u32 nodes(u32 g) const { return go[g + 1] - go[g]; }  // :84
  u32 prev(u32 g) const { return go[g]; }              // :86
  u64 graphID(u32 v) const {                           // :100-111
    u64 lo = 0, hi = G;
    while (lo < hi) {
      u64 mid = lo + (hi - lo) / 2;
      if (v >= go[mid]) lo = mid + 1;
      else hi = mid;
    }
    return lo > 0 ? lo - 1 : static_cast<u64>(-1);
  }
};

// ---- CandidatesDevice, candidates.hpp:16-162 ------------------------------------------
struct Cands {
  u32* w;
  u64 stride;  // single_node_size = ceil(target/32)
  void insert(u32 s, u32 c) const { w[s * stride + c / 32] |= u32(1) << (c % 32); }      // :32-36
  void remove(u32 s, u32 c) const { w[s * stride + c / 32] &= ~(u32(1) << (c % 32)); }   // :60-64
  bool contains(u32 s, u32 c) const { return (w[s * stride + c / 32] & (u32(1) << (c % 32))) != 0; }
  u32 count(u32 s) const {  // :77-81
    u32 n = 0;
    for (u64 i = 0; i < stride; ++i) n += __builtin_popcount(w[s * stride + i]);
    return n;
  }
  u32 mask_for(u64 i, u32 si, u32 ei, u32 gs, u32 ge) const {  // :92-99 / :138-145, A1
    u32 mask = ~u32(0);
    if (si == ei - 1) mask = (mask >> ((32u - (ge - gs)) & 31u)) << ((gs % 32u) & 31u);
    else if (i == si) mask <<= (gs % 32u);
    else if (i == ei - 1) mask >>= ((32u - (ge % 32u)) & 31u);
    return mask;
  }
  u32 count(u32 s, u32 gs, u32 ge) const {  // :86-104
    u32 n = 0;
    const u32 si = gs / 32, ei = (ge + 31) / 32;
    for (u64 i = si; i < ei; ++i) n += __builtin_popcount(w[s * stride + i] & mask_for(i, si, ei, gs, ge));
    return n;
  }
  u32 at(u32 s, u32 idx, u32 gs, u32 ge) const {  // :132-162
    u32 n = 0;
    const u32 si = gs / 32, ei = (ge + 31) / 32;
    for (u64 i = si; i < ei; ++i) {
      u32 block = w[s * stride + i] & mask_for(i, si, ei, gs, ge);
      u32 bc = __builtin_popcount(block);
      if (n + bc > idx) {
        while (block != 0) {
          u32 t = ctz32(block);
          if (n == idx) return static_cast<u32>(i * 32 + t);
          block ^= u32(1) << t;
          ++n;
        }
      } else {
        n += bc;
      }
    }
    return 0xFFFFFFFFu;
  }
};

Graph make_graph(u32 G, const u32* go, const u32* ro, const u32* ci, const u8* nl) { return Graph{G, go, ro, ci, nl}; }

}  // namespace

extern "C" {

void oracle_set_threads(int n) {
  if (n > 0) omp_set_num_threads(n);
}
int oracle_get_threads(void) { return omp_get_max_threads(); }

// SignatureDevice known-answer hooks (signature.hpp:32-51).
uint64_t oracle_sig_set(uint64_t sig, uint8_t label, uint8_t count) {
  sig_set(sig, label, count);
  return sig;
}
uint64_t oracle_sig_inc(uint64_t sig, uint8_t label, uint8_t add) {
  sig_inc(sig, label, add);
  return sig;
}
uint8_t oracle_sig_get(uint64_t sig, uint8_t label) { return sig_get(sig, label); }

// CandidatesDevice single-word hooks (candidates.hpp:32-64,77-162).
void oracle_cand_insert(uint32_t* words, uint64_t stride, uint32_t s, uint32_t c) { Cands{words, stride}.insert(s, c); }
void oracle_cand_remove(uint32_t* words, uint64_t stride, uint32_t s, uint32_t c) { Cands{words, stride}.remove(s, c); }
uint32_t oracle_cand_count(uint32_t* words, uint64_t stride, uint32_t s) { return Cands{words, stride}.count(s); }
uint32_t oracle_cand_count_range(uint32_t* words, uint64_t stride, uint32_t s, uint32_t gs, uint32_t ge) {
  return Cands{words, stride}.count(s, gs, ge);
}
uint32_t oracle_cand_at(uint32_t* words, uint64_t stride, uint32_t s, uint32_t idx, uint32_t gs, uint32_t ge) {
  return Cands{words, stride}.at(s, idx, gs, ge);
}

// generateCSRSignatures, signature.hpp:125-152 (range over every node, increments onto the
// existing values; the constructor zero-fills them, :166-167).
void oracle_generate_signatures(uint32_t G, const uint32_t* go, const uint32_t* ro, const uint32_t* ci, const uint8_t* nl,
                                uint64_t* sigs) {
  const Graph g = make_graph(G, go, ro, ci, nl);
  const int64_t N = static_cast<int64_t>(g.N());
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < N; ++v) {
    for (u32 i = ro[v]; i < ro[v + 1]; ++i) sig_inc(sigs[v], nl[ci[i]]);
  }
}

// refineCSRSignatures<PowerGraph>, signature.hpp:336-390.  `reach` holds the per-node
// old_reachables data words (zero before the first call).
// mutation 1: BFS bits use the true local id without the 64-bit wrap (breaks n > 64).
// mutation 2: the start node is not removed from the sphere.
void oracle_refine_signatures(uint32_t G, const uint32_t* go, const uint32_t* ro, const uint32_t* ci, const uint8_t* nl,
                              uint32_t view_size, uint64_t* sigs, uint64_t* reach, int mutation) {
  const Graph g = make_graph(G, go, ro, ci, nl);
  const int64_t N = static_cast<int64_t>(g.N());
#pragma omp parallel for schedule(dynamic, 4096)
  for (int64_t v = 0; v < N; ++v) {
    const u64 gid = g.graphID(static_cast<u32>(v));
    const u32 prev = g.prev(static_cast<u32>(gid));
    Bitset64 frontier, reachable;
    Bitset64 old;
    old.data = reach[v];
    const u32 local = static_cast<u32>(v) - prev;
    if (mutation == 1 && local >= 64) {
      // the mistaken variant simply drops nodes it cannot represent
      sigs[v] = 0;
      continue;
    }
    frontier.set(static_cast<u8>(local));
    reachable.set(static_cast<u8>(local));
    for (u32 it = 0; it < view_size && !frontier.empty(); it++) {
      Bitset64 next;
      for (u32 idx = 0; idx < frontier.size(); idx++) {
        const u32 u = frontier.getSetBit(static_cast<u8>(idx)) + prev;
        for (u32 i = ro[u]; i < ro[u + 1]; ++i) {
          const u32 nb = ci[i] - prev;
          if (!reachable.get(static_cast<u8>(nb))) {
            reachable.set(static_cast<u8>(nb));
            next.set(static_cast<u8>(nb));
          }
        }
      }
      frontier = next;
    }
    if (mutation != 2) reachable.unset(static_cast<u8>(local));
    reachable.difference(old);
    u64 s = 0;
    for (u32 idx = 0; idx < reachable.size(); idx++) {
      const u32 u = reachable.getSetBit(static_cast<u8>(idx)) + prev;
      sig_inc(s, nl[u]);
    }
    sigs[v] = s;
    reachable.merge(old);
    reach[v] = reachable.data;
  }
}

// refineCSRSignatures<ViewBased>, signature.hpp:298-334: signatures are copied (tmp_buff), then
// every node adds, for each CSR entry to a neighbour u and each label l < 16, the count
// c = (uint8)(tmp[u][l] - (l == label(v) ? view_size : 0)) when c > 0 -- through
// incrementLabelCount, whose uint8 sum makes a wrapped c act as a decrement (:44-51).
// mutation 1: the neighbour's count is not reduced for the node's own label.
void oracle_refine_view_based(uint32_t G, const uint32_t* go, const uint32_t* ro, const uint32_t* ci, const uint8_t* nl,
                              uint64_t view_size, uint64_t* sigs, int mutation) {
  const Graph g = make_graph(G, go, ro, ci, nl);
  const int64_t N = static_cast<int64_t>(g.N());
  std::vector<u64> tmp(sigs, sigs + N);
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < N; ++v) {
    const u8 node_label = nl[v];
    for (u32 i = ro[v]; i < ro[v + 1]; ++i) {
      const u32 nb = ci[i];
      for (u8 l = 0; l < 16; l++) {
        u8 count = sig_get(tmp[nb], l);
        if (l == node_label && mutation != 1) count = static_cast<u8>(count - view_size);
        if (count > 0) sig_inc(sigs[v], l, count);
      }
    }
  }
}

// filterCandidates<Data>, isomorphism.hpp:49-52: the same label test, stored transposed --
// row = data node, column = query node (Candidates constructed with D rows and Q targets).
void oracle_filter_data(uint64_t Q, const uint8_t* qlabels, uint64_t D, const uint8_t* dlabels, uint32_t* words) {
  const u64 stride = (Q + 31) / 32;
  const Cands c{words, stride};
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t d = 0; d < static_cast<int64_t>(D); ++d)
    for (u64 q = 0; q < Q; ++q) {
      if (qlabels[q] != dlabels[d] && qlabels[q] != kWildcard) continue;
      c.insert(static_cast<u32>(d), static_cast<u32>(q));
    }
}

// filterCandidates<Query>, isomorphism.hpp:22-62: d in C(q) iff the labels are equal or the
// query label is the wildcard.  ORs into `words` (zeroed by the Candidates constructor,
// candidates.hpp:165-179).  Rows are independent, so the loop runs over rows.
void oracle_filter(uint64_t Q, const uint8_t* qlabels, uint64_t D, const uint8_t* dlabels, uint32_t* words) {
  const u64 stride = (D + 31) / 32;
  const Cands c{words, stride};
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < static_cast<int64_t>(Q); ++q)
    for (u64 d = 0; d < D; ++d) {
      if (qlabels[q] != dlabels[d] && qlabels[q] != kWildcard) continue;
      c.insert(static_cast<u32>(q), static_cast<u32>(d));
    }
}

// refineCandidates<Query>, isomorphism.hpp:64-127: a set bit (q,d) survives iff
// query count <= data count on all 16 labels (current-level signatures).
// mutation 1: compare only labels 0..7.
void oracle_refine_candidates(uint64_t Q, const uint64_t* qsig, uint64_t D, const uint64_t* dsig, uint32_t* words,
                              int mutation) {
  const u64 stride = (D + 31) / 32;
  const Cands c{words, stride};
  const int max_labels = mutation == 1 ? 8 : 16;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t q = 0; q < static_cast<int64_t>(Q); ++q) {
    const u64 qs = qsig[q];
    for (u64 d = 0; d < D; ++d) {
      if (!c.contains(static_cast<u32>(q), static_cast<u32>(d))) continue;
      const u64 ds = dsig[d];
      bool keep = true;
      for (int l = 0; l < max_labels && keep; l++) keep = keep && (sig_get(qs, static_cast<u8>(l)) <= sig_get(ds, static_cast<u8>(l)));
      if (!keep) c.remove(static_cast<u32>(q), static_cast<u32>(d));
    }
  }
}

// Per-row popcounts, candidates.hpp:77-81 (the driver's candidate statistics, sigmo.cpp:201-214).
void oracle_candidate_counts(uint64_t Q, uint64_t D, const uint32_t* words, uint32_t* counts) {
  const u64 stride = (D + 31) / 32;
  const Cands c{const_cast<u32*>(words), stride};
#pragma omp parallel for
  for (int64_t q = 0; q < static_cast<int64_t>(Q); ++q) counts[q] = c.count(static_cast<u32>(q));
}

// GMCR::generateGMCR, gmcr.hpp:34-128, kernel 1 + serial scan (offsets[Gd+1]).
// Returns the total (query-graph, data-graph) pairs or -1 if a query graph has < 2 nodes.
int64_t oracle_gmcr_offsets(uint32_t Gq, const uint32_t* qgo, uint32_t Gd, const uint32_t* dgo, uint64_t D,
                            const uint32_t* words, uint32_t* offsets) {
  for (u32 qg = 0; qg < Gq; ++qg)
    if (qgo[qg + 1] - qgo[qg] < 2) return -1;
  const Cands c{const_cast<u32*>(words), (D + 31) / 32};
  std::memset(offsets, 0, sizeof(u32) * (Gd + 1));
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t dg = 0; dg < static_cast<int64_t>(Gd); ++dg) {
    const u32 s = dgo[dg], e = dgo[dg + 1];
    u32 cnt = 0;
    for (u32 qg = 0; qg < Gq; ++qg) {  // kernel 1, gmcr.hpp:52-68
      const u32 k = qgo[qg + 1] - qgo[qg], off = qgo[qg];
      bool add = true;
      for (u32 i = 0; i < k && add; i++) add = add && (c.count(off + i, s, e) > 0);
      if (add) cnt++;
    }
    offsets[dg + 1] = cnt;
  }
  for (u64 i = 1; i <= Gd; i++) offsets[i] += offsets[i - 1];  // kernel 2, gmcr.hpp:73-75
  return offsets[Gd];
}

// Kernel 3 (gmcr.hpp:94-115) with the within-segment order fixed to ascending query id.
void oracle_gmcr_indices(uint32_t Gq, const uint32_t* qgo, uint32_t Gd, const uint32_t* dgo, uint64_t D,
                         const uint32_t* words, const uint32_t* offsets, uint32_t* indices) {
  const Cands c{const_cast<u32*>(words), (D + 31) / 32};
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t dg = 0; dg < static_cast<int64_t>(Gd); ++dg) {
    const u32 s = dgo[dg], e = dgo[dg + 1];
    u32 cur = offsets[dg];
    for (u32 qg = 0; qg < Gq; ++qg) {
      const u32 k = qgo[qg + 1] - qgo[qg], off = qgo[qg];
      if (k <= 1) continue;
      bool add = true;
      for (u32 i = 0; i < k && add; i++) add = add && (c.count(off + i, s, e) > 0);
      if (add) indices[cur++] = qg;
    }
  }
}

// joinCandidates, isomorphism.hpp:270-381 (DFS control flow :319-371, isValidMapping
// :150-168).  Writes each pair's count (GMCR order) and returns the sum.
// mutation 1: textbook backtracking: whenever a frame is popped (leaf or exhausted) the
//             parent's current candidate is released; for molecules of <= 64 nodes this
//             counts true induced embeddings.
// mutation 2: mapping[] starts at global node 0 instead of the molecule's first node.
// mutation 3: the induced test checks query edges only (non-edges not required).
// (Indexing visited by global id instead of id - offset is not a mistake: it rotates every
// position by the same constant, so the mod-64 classes and all results are unchanged.)
int64_t oracle_join(uint32_t Gq, const uint32_t* qgo, const uint32_t* qro, const uint32_t* qci, uint32_t Gd,
                    const uint32_t* dgo, const uint32_t* dro, const uint32_t* dci, uint64_t D, const uint32_t* words,
                    const uint32_t* offsets, const uint32_t* indices, int find_first, int32_t* counts, int mutation) {
  const Graph qgr = make_graph(Gq, qgo, qro, qci, nullptr);
  const Graph dgr = make_graph(Gd, dgo, dro, dci, nullptr);
  const Cands c{const_cast<u32*>(words), (D + 31) / 32};
  int64_t total = 0;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total) reduction(| : bad)
  for (int64_t dg = 0; dg < static_cast<int64_t>(Gd); ++dg) {
    const u32 s = dgo[dg], e = dgo[dg + 1];
    for (u32 j = offsets[dg]; j < offsets[dg + 1]; ++j) {
      const u32 qg = indices[j];
      const u32 qoff = qgr.prev(qg);
      const u32 k = qgr.nodes(qg);
      if (k + 1 > static_cast<u32>(kMaxQueryNodes)) {
        bad = 1;
        counts[j] = -1;
        continue;
      }
      struct Frame {
        u32 depth;
        u64 idx;
      } stack[kMaxQueryNodes];
      u32 mapping[kMaxQueryNodes];
      for (auto& m : mapping) m = mutation == 2 ? 0u : s;
      Bitset64 visited{s};
      u32 top = 0;
      stack[top++] = {0, 0};
      int64_t count = 0;
      while (top > 0) {
        const Frame f = stack[top - 1];
        const u32 qn = f.depth;
        if (f.depth == k) {  // :336-344
          count++;
          top--;
          if (mutation == 1 && top > 0) visited.unset(static_cast<u8>(mapping[top - 1]));
          if (find_first) break;
          continue;
        }
        if (f.idx >= c.count(qn + qoff, s, e)) {  // :346-354
          top--;
          if (mutation != 1) {
            visited.unset(static_cast<u8>(mapping[qn]));
            if (top == 1) visited.clear();
          } else if (top > 0) {
            // textbook: release the candidate held by the frame we return to
            visited.unset(static_cast<u8>(mapping[top - 1]));
          }
          continue;
        }
        const u32 cand = c.at(qn + qoff, static_cast<u32>(f.idx), s, e);  // :357
        stack[top - 1].idx++;
        if (visited.get(static_cast<u8>(cand))) continue;  // :363
        bool valid = true;
        if (f.depth != 0) {  // isValidMapping, :150-168
          for (u32 i = 0; i < f.depth; i++) {
            const bool qa = qgr.isNeighbor(i + qoff, f.depth + qoff);
            const bool da = dgr.isNeighbor(mapping[i], cand);
            if (mutation == 3 ? (qa && !da) : (qa != da)) {
              valid = false;
              break;
            }
          }
        }
        if (valid) {  // :366-370
          mapping[f.depth] = cand;
          visited.set(static_cast<u8>(cand));
          stack[top++] = {f.depth + 1, 0};
        }
      }
      if (count > INT32_MAX) bad = 1;
      counts[j] = static_cast<int32_t>(std::min<int64_t>(count, INT32_MAX));
      total += count;
    }
  }
  return bad ? -1 : total;
}

// joinWildcardCandidates, isomorphism.hpp:428-556, with defineMatchingOrder (:383-408) and the
// 8-argument isValidMapping (:410-426).  Per pair:
//   * the matching order starts with the query node of most candidates in the molecule among
//     the first `sg_size` nodes (ties: the larger index, :397-401), then every other node in
//     ascending order (:404-407);
//   * each candidate of that first node starts its own DFS (:413-416; the reference splits
//     them over the sub-group's lanes), from fresh state: every mapping[] entry equal to the
//     molecule's first node id (local id 0);
//   * a candidate is valid when, for every earlier node of the order, query adjacency equals
//     data adjacency and adjacent pairs carry equal edge labels (first CSR entry, graph.hpp:88-93);
//   * an exhausted frame unsets its node's (possibly stale) mapping; when one frame remains,
//     visited is reset to the first node's candidate (:521-526); find-first stops a DFS at its
//     first leaf, so the pair's count is the number of first-node candidates that extend.
// mutation 1: edge labels not compared.  mutation 2: matching order = query order.
// mutation 3: ties of the first node go to the smaller index.
int64_t oracle_join_wildcard(uint32_t Gq, const uint32_t* qgo, const uint32_t* qro, const uint32_t* qci, const uint8_t* qel,
                             uint32_t Gd, const uint32_t* dgo, const uint32_t* dro, const uint32_t* dci, const uint8_t* del, uint64_t D,
                             const uint32_t* words, const uint32_t* offsets, const uint32_t* indices, int find_first, uint32_t sg_size,
                             int32_t* counts, int mutation) {
  const Graph qgr = make_graph(Gq, qgo, qro, qci, nullptr);
  const Graph dgr = make_graph(Gd, dgo, dro, dci, nullptr);
  const Cands c{const_cast<u32*>(words), (D + 31) / 32};
  auto edge_label = [](const u32* ro, const u32* ci, const u8* el, u32 a, u32 b) -> u8 {  // graph.hpp:88-93
    for (u32 i = ro[a]; i < ro[a + 1]; ++i)
      if (ci[i] == b) return el[i];
    return static_cast<u8>(-1);
  };
  int64_t total = 0;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : total) reduction(| : bad)
  for (int64_t dg = 0; dg < static_cast<int64_t>(Gd); ++dg) {
    const u32 s = dgo[dg], e = dgo[dg + 1];
    for (u32 j = offsets[dg]; j < offsets[dg + 1]; ++j) {
      const u32 qg = indices[j];
      const u32 qoff = qgr.prev(qg);
      const u32 k = qgr.nodes(qg);
      if (k + 1 > static_cast<u32>(kMaxQueryNodes)) {
        bad = 1;
        counts[j] = -1;
        continue;
      }
      // defineMatchingOrder
      u32 order[kMaxQueryNodes];
      u32 max_candidates = 0;
      int first = -1;
      const u32 considered = std::min(k, sg_size);
      for (u32 i = 0; i < considered; ++i) max_candidates = std::max(max_candidates, c.count(qoff + i, s, e));
      for (u32 i = 0; i < considered; ++i)
        if (c.count(qoff + i, s, e) == max_candidates && (first < 0 || mutation != 3)) first = static_cast<int>(i);
      if (mutation == 2) {
        first = 0;
        max_candidates = c.count(qoff, s, e);
      }
      order[0] = static_cast<u32>(first);
      for (u32 i = 0, jj = 1; i < k; ++i)
        if (static_cast<int>(i) != first) order[jj++] = i;
      int64_t count = 0;
      for (u32 start = 0; start < max_candidates; ++start) {
        struct Frame {
          u32 depth;
          u64 idx;
        } stack[kMaxQueryNodes];
        u32 mapping[kMaxQueryNodes];
        for (auto& m : mapping) m = s;
        Bitset64 visited{s};
        u32 top = 0;
        mapping[order[0]] = c.at(order[0] + qoff, start, s, e);
        visited.set(static_cast<u8>(mapping[order[0]]));
        stack[top++] = {1, 0};
        while (top > 0) {
          const Frame f = stack[top - 1];
          if (f.depth == k) {
            count++;
            top--;
            if (find_first) break;
            continue;
          }
          const u32 qn = order[f.depth];
          if (f.idx >= c.count(qn + qoff, s, e)) {
            top--;
            visited.unset(static_cast<u8>(mapping[qn]));
            if (top == 1) {
              visited.clear();
              visited.set(static_cast<u8>(mapping[order[0]]));
            }
            continue;
          }
          const u32 cand = c.at(qn + qoff, static_cast<u32>(f.idx), s, e);
          stack[top - 1].idx++;
          if (visited.get(static_cast<u8>(cand))) continue;
          bool valid = true;
          for (u32 i = 0; i < f.depth && valid; i++) {  // isValidMapping (8 args)
            const u32 qa_n = order[i] + qoff, qb_n = qn + qoff;
            const bool qa = qgr.isNeighbor(qa_n, qb_n);
            const bool da = dgr.isNeighbor(mapping[order[i]], cand);
            if (qa != da) valid = false;
            else if (qa && mutation != 1 && edge_label(dro, dci, del, mapping[order[i]], cand) != edge_label(qro, qci, qel, qa_n, qb_n))
              valid = false;
          }
          if (valid) {
            mapping[qn] = cand;
            visited.set(static_cast<u8>(cand));
            stack[top++] = {f.depth + 1, 0};
          }
        }
      }
      if (count > INT32_MAX) bad = 1;
      counts[j] = static_cast<int32_t>(std::min<int64_t>(count, INT32_MAX));
      total += count;
    }
  }
  return bad ? -1 : total;
}

// ---- adjacency-matrix (AM) query graphs -------------------------------------------------------
// AMGraph / DeviceBatchedAMGraph, graph.hpp:20-36, 114-148, 208-214, 345-392: per graph an
// n x n bit matrix in uint64 words, bit u*W + v set for every edge (both directions), where the
// word count is utils.hpp:17-21's (n*n + 63) / 64 (as uint8) and the row width W is
// utils.hpp:25-28's (uint16) sqrt(float(64 * words)).  num_nodes[] is cumulative; graph g's matrix
// starts at word graph_offsets[g].  Graphs of more than 64 nodes are not supported.

static u8 am_words(u32 n) { return static_cast<u8>((n * n + 63) / 64); }
static u16 am_width(u8 words) { return static_cast<u16>(std::sqrt(static_cast<float>(64 * static_cast<u32>(words)))); }

// utils.hpp:100-112 binaryAMSearch over the cumulative node counts.
static u32 am_search(const u32* num_nodes, u32 total_graphs, u32 node_id) {
  u32 low = 0, high = total_graphs - 1;
  while (low < high) {
    const u32 mid = (low + high) / 2;
    if (num_nodes[mid] <= node_id) low = mid + 1;
    else high = mid;
  }
  return low;
}

// utils.hpp:34-48 getNeighbors: the row's set bits in ascending order.  The reference writes
// every neighbour into a MAX_NEIGHBORS = 4 array and its readers stop at 4 (types.hpp:28,
// signature.hpp:110-113): the first four are the neighbours that count.  Returns how many.
static u32 am_neighbors(const u64* adj, u32 prev, u32 n, u32 u, u32 out[4]) {
  const u16 W = am_width(am_words(n));
  u32 cnt = 0;
  for (u32 v = 0; v < W && cnt < 4; ++v) {
    const u32 idx = u * W + v;
    if ((adj[idx / 64] >> (idx % 64)) & 1ull) out[cnt++] = v + prev;
  }
  return cnt;
}

// Total adjacency words of the AM form of a CSR batch (graph.hpp:208-214, 345-392); 0 if a
// graph has more than 64 nodes.
uint64_t oracle_am_words(uint32_t G, const uint32_t* go) {
  u64 total = 0;
  for (u32 g = 0; g < G; ++g) {
    const u32 n = go[g + 1] - go[g];
    if (n > 64) return 0;
    total += am_words(n);
  }
  return total;
}

// toAMGraph (graph.hpp:208-214, setBit utils.hpp:24-32) for every graph of a CSR batch, laid
// out as createDeviceAMGraph does: graph_offsets[g] (words), num_nodes[g] (cumulative).
void oracle_am_build(uint32_t G, const uint32_t* go, const uint32_t* ro, const uint32_t* ci, uint32_t* graph_offsets, uint32_t* num_nodes,
                     uint64_t* adjacency) {
  u64 off = 0;
  for (u32 g = 0; g < G; ++g) {
    const u32 s = go[g], n = go[g + 1] - s;
    const u8 words = am_words(n);
    const u16 W = am_width(words);
    graph_offsets[g] = static_cast<u32>(off);
    num_nodes[g] = go[g + 1];
    u64* m = adjacency + off;
    std::fill(m, m + words, 0ull);
    for (u32 v = s; v < s + n; ++v)
      for (u32 i = ro[v]; i < ro[v + 1]; ++i) {
        const u8 a = static_cast<u8>(v - s), b = static_cast<u8>(ci[i] - s);
        const u32 iu = a * W + b, iv = b * W + a;
        m[iu / 64] |= 1ull << (iu % 64);
        m[iv / 64] |= 1ull << (iv % 64);
      }
    off += words;
  }
}

// generateAMSignatures, signature.hpp:100-120: each node's (first four) neighbours' labels are
// counted onto its signature.
void oracle_am_generate(uint32_t G, const uint32_t* num_nodes, const uint32_t* graph_offsets, const uint64_t* adjacency,
                        const uint8_t* labels, uint64_t total_nodes, uint64_t* sigs) {
  for (u64 v = 0; v < total_nodes; ++v) {
    const u32 g = am_search(num_nodes, G, static_cast<u32>(v));
    const u32 prev = g ? num_nodes[g - 1] : 0;
    u32 nb[4];
    const u32 c = am_neighbors(adjacency + graph_offsets[g], prev, num_nodes[g] - prev, static_cast<u32>(v) - prev, nb);
    for (u32 i = 0; i < c; ++i) sig_inc(sigs[v], labels[nb[i]]);
  }
}

// refineAMSignatures<PowerGraph>(view_size), signature.hpp:250-296: a BFS of view_size layers
// over uint32 position sets (Bitset<uint32_t>, positions through uint8, shifts masked & 31),
// then the signature is cleared and rebuilt as the label histogram of every reached node but
// the node itself.  mutation 1: the histogram of the last layer only (the sphere, not the ball).
void oracle_am_refine_pg(uint32_t G, const uint32_t* num_nodes, const uint32_t* graph_offsets, const uint64_t* adjacency,
                         const uint8_t* labels, uint64_t total_nodes, uint64_t view_size, uint64_t* sigs, int mutation) {
  auto bit = [](u32 idx) { return 1u << (static_cast<u8>(idx) & 31u); };
  for (u64 v = 0; v < total_nodes; ++v) {
    const u32 g = am_search(num_nodes, G, static_cast<u32>(v));
    const u32 prev = g ? num_nodes[g - 1] : 0, n = num_nodes[g] - prev;
    const u64* adj = adjacency + graph_offsets[g];
    const u32 self = static_cast<u32>(v) - prev;
    u32 frontier = bit(self), reachable = bit(self);
    for (u64 it = 0; it < view_size && frontier; ++it) {
      u32 next = 0;
      for (u32 f = frontier; f; f &= f - 1) {
        const u32 u = static_cast<u32>(ctz32(f));
        u32 nb[4];
        const u32 c = am_neighbors(adj, prev, n, u, nb);
        for (u32 i = 0; i < c; ++i) {
          const u32 w = nb[i] - prev;
          if (!(reachable & bit(w))) {
            reachable |= bit(w);
            next |= bit(w);
          }
        }
      }
      frontier = next;
    }
    if (mutation == 1) reachable = frontier;  // planted: the last layer only
    reachable &= ~bit(self);
    sigs[v] = 0;
    for (u32 r = reachable; r; r &= r - 1) sig_inc(sigs[v], labels[prev + static_cast<u32>(ctz32(r))]);
  }
}

// refineAMSignatures<ViewBased>(view_size), signature.hpp:214-248: as the CSR form
// (oracle_refine_view_based) over the node's (first four) AM neighbours.
void oracle_am_refine_vb(uint32_t G, const uint32_t* num_nodes, const uint32_t* graph_offsets, const uint64_t* adjacency,
                         const uint8_t* labels, uint64_t total_nodes, uint64_t view_size, uint64_t* sigs) {
  std::vector<u64> tmp(sigs, sigs + total_nodes);
  for (u64 v = 0; v < total_nodes; ++v) {
    const u32 g = am_search(num_nodes, G, static_cast<u32>(v));
    const u32 prev = g ? num_nodes[g - 1] : 0;
    u32 nb[4];
    const u32 c = am_neighbors(adjacency + graph_offsets[g], prev, num_nodes[g] - prev, static_cast<u32>(v) - prev, nb);
    const u8 own = labels[v];
    for (u32 i = 0; i < c; ++i)
      for (u8 l = 0; l < 16; l++) {
        u8 count = sig_get(tmp[nb[i]], l);
        if (l == own) count = static_cast<u8>(count - view_size);
        if (count > 0) sig_inc(sigs[v], l, count);
      }
  }
}

}  // extern "C"
