// Deterministic synthetic molecule / query / arrival generator.
//
// The reference ships no generator: its inputs are ZINC molecules converted by
// scripts/smile2graph.py.  This generator produces molecule-like graphs with the
// reference's own label and bond code tables (scripts/smile2graph.py:13-14) so the same
// seeded input feeds the GPU path and the CPU oracle.  Every choice is fixed here:
//
//  * PRNG: xoshiro256** seeded through SplitMix64.  Each role (molecules, queries,
//    perturbation, ring selection, arrivals) has its own stream seed
//    splitmix64(seed ^ ROLE); molecule i draws from its own stream
//    splitmix64(role_seed + i * 0x9E3779B97F4A7C15), so any slice of a sequence equals the
//    same slice of a longer one and generation parallelises.
//  * Normal by Box-Muller (cosine branch), LogNormal = exp(Normal), Poisson by inversion,
//    uniform integers by 128-bit multiply-high.  No std::*_distribution.
//  * Molecule: n from the size law; atom labels C .72 (8), N .11 (0), O .11 (7), S .02 (9),
//    F .015 (10), Cl .01 (1), Br .005 (3), P .01 (5); degree caps C4 N3 O2 S2 P3 halogen 1.
//    Atom i attaches to a uniform earlier atom with free degree; if none is free the
//    molecule stops early.  Poisson(2.5) ring closures, each between a uniform pair at
//    spanning-tree distance 4 or 5 with free degree on both ends and no edge yet.  Bonds:
//    single/double/triple .85/.13/.02 per edge, then each 6-ring is aromatic (4) with p .7
//    (one draw per ring, applied to its 6 edges).
//  * Query: a fresh molecule from the query's own stream, k ~ U[kmin,kmax] (molecule
//    redrawn while n < k), BFS from a uniform start visiting neighbours in ascending id,
//    first k visited nodes, induced subgraph, node ids in BFS order, labels kept.
//    Perturbation relabels one uniform node of an exact fraction of queries (chosen by a
//    seeded permutation) with a label drawn from the atom law conditioned on != original.
//    With a ring fraction, that exact fraction of queries (seeded permutation) must contain
//    a cycle and the others must be acyclic (BFS sets are resampled until they agree).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "host_graphs.hpp"

namespace sigmo_gen {

static inline uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static inline uint64_t mix(uint64_t x) { return splitmix64(x); }

struct Rng {
  uint64_t s[4];
  explicit Rng(uint64_t seed) {
    uint64_t x = seed;
    for (auto& w : s) w = splitmix64(x);
  }
  static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t next() {
    const uint64_t r = rotl(s[1] * 5, 7) * 9;
    const uint64_t t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return r;
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }  // [0,1)
  uint64_t below(uint64_t n) { return static_cast<uint64_t>((static_cast<unsigned __int128>(next()) * n) >> 64); }
  double normal() {
    const double u1 = 1.0 - uniform();  // (0,1]
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  int poisson(double lambda) {
    const double u = uniform();
    double p = std::exp(-lambda), F = p;
    int k = 0;
    while (u >= F && k < 1000) {
      ++k;
      p *= lambda / k;
      F += p;
    }
    return k;
  }
};

enum Role : uint64_t {
  ROLE_MOLECULES = 0x4d4f4c45434c4553ull,
  ROLE_QUERIES = 0x5155455249455321ull,
  ROLE_PERTURB = 0x5045525455524221ull,
  ROLE_RINGS = 0x52494e4753454c21ull,
  ROLE_ARRIVALS = 0x4152524956414c53ull,
};

static inline uint64_t role_seed(uint64_t seed, Role r) { return mix(seed ^ static_cast<uint64_t>(r)); }
static inline uint64_t item_seed(uint64_t rs, uint64_t i) { return mix(rs + i * 0x9E3779B97F4A7C15ull); }

// Atom law (scripts/smile2graph.py:13 label codes).
static const uint8_t kAtomLabel[8] = {8, 0, 7, 9, 10, 1, 3, 5};          // C N O S F Cl Br P
static const double kAtomCdf[8] = {0.72, 0.83, 0.94, 0.96, 0.975, 0.985, 0.99, 1.0};
static inline uint8_t cap_of(uint8_t label) {
  switch (label) {
    case 8: return 4;  // C
    case 0: return 3;  // N
    case 7: return 2;  // O
    case 9: return 2;  // S
    case 5: return 3;  // P
    default: return 1; // F, Cl, Br
  }
}
static inline uint8_t draw_atom(Rng& r) {
  const double u = r.uniform();
  for (int i = 0; i < 8; ++i)
    if (u < kAtomCdf[i]) return kAtomLabel[i];
  return kAtomLabel[0];
}

static uint32_t draw_size(Rng& r, int dist) {
  double x;
  int lo, hi;
  if (dist == SIGMO_SIZE_NORMAL) {
    x = 24.0 + 8.0 * r.normal();
    lo = 6;
    hi = 64;
  } else {
    x = std::exp(std::log(24.0) + 0.45 * r.normal());
    lo = 6;
    hi = 128;
  }
  double y = std::nearbyint(x);
  if (y < lo) y = lo;
  if (y > hi) y = hi;
  return static_cast<uint32_t>(y);
}

struct Molecule {
  sigmo_host::EdgeListGraph g;
  std::vector<std::vector<uint32_t>> adj;  // sorted neighbour lists
};

static void build_molecule(Rng& r, int dist, Molecule& m) {
  const uint32_t target = draw_size(r, dist);
  auto& g = m.g;
  g.labels.clear();
  g.eu.clear();
  g.ev.clear();
  g.el.clear();
  std::vector<uint8_t> deg, cap;
  std::vector<uint32_t> parent;
  std::vector<uint32_t> freeset;
  g.labels.push_back(draw_atom(r));
  deg.push_back(0);
  cap.push_back(cap_of(g.labels[0]));
  parent.push_back(UINT32_MAX);
  for (uint32_t i = 1; i < target; ++i) {
    const uint8_t lab = draw_atom(r);
    freeset.clear();
    for (uint32_t j = 0; j < i; ++j)
      if (deg[j] < cap[j]) freeset.push_back(j);
    if (freeset.empty()) break;
    const uint32_t p = freeset[r.below(freeset.size())];
    g.labels.push_back(lab);
    deg.push_back(1);
    cap.push_back(cap_of(lab));
    parent.push_back(p);
    deg[p]++;
    g.eu.push_back(p);
    g.ev.push_back(i);
  }
  const uint32_t n = static_cast<uint32_t>(g.labels.size());
  // spanning-tree depths for path reconstruction
  std::vector<uint32_t> depth(n, 0);
  for (uint32_t i = 1; i < n; ++i) depth[i] = depth[parent[i]] + 1;
  // pairs at tree distance 4 or 5 (BFS over the tree, bounded at depth 5)
  std::vector<std::vector<uint32_t>> tadj(n);
  for (size_t e = 0; e < g.eu.size(); ++e) {
    tadj[g.eu[e]].push_back(g.ev[e]);
    tadj[g.ev[e]].push_back(g.eu[e]);
  }
  std::vector<std::pair<uint32_t, uint32_t>> pairs;  // (a<b) with dist
  std::vector<uint8_t> pdist;
  std::vector<int> dist_from(n, -1);
  std::vector<uint32_t> q;
  for (uint32_t a = 0; a < n; ++a) {
    q.clear();
    q.push_back(a);
    dist_from[a] = 0;
    for (size_t h = 0; h < q.size(); ++h) {
      const uint32_t u = q[h];
      if (dist_from[u] == 5) continue;
      for (uint32_t w : tadj[u])
        if (dist_from[w] < 0) {
          dist_from[w] = dist_from[u] + 1;
          q.push_back(w);
        }
    }
    for (uint32_t u : q) {
      if (u > a && (dist_from[u] == 4 || dist_from[u] == 5)) {
        pairs.emplace_back(a, u);
        pdist.push_back(static_cast<uint8_t>(dist_from[u]));
      }
      dist_from[u] = -1;
    }
  }
  const int closures = r.poisson(2.5);
  std::vector<uint32_t> ok;
  std::vector<std::pair<uint32_t, uint32_t>> six_rings;
  std::vector<uint8_t> used(pairs.size(), 0);
  for (int c = 0; c < closures; ++c) {
    ok.clear();
    for (uint32_t i = 0; i < pairs.size(); ++i) {
      const auto& pr = pairs[i];
      if (!used[i] && deg[pr.first] < cap[pr.first] && deg[pr.second] < cap[pr.second]) ok.push_back(i);
    }
    if (ok.empty()) break;
    const uint32_t pick = ok[r.below(ok.size())];
    used[pick] = 1;
    const auto pr = pairs[pick];
    deg[pr.first]++;
    deg[pr.second]++;
    g.eu.push_back(pr.first);
    g.ev.push_back(pr.second);
    if (pdist[pick] == 5) six_rings.push_back(pr);
  }
  // bond labels (scripts/smile2graph.py:14 codes)
  g.el.resize(g.eu.size());
  for (size_t e = 0; e < g.eu.size(); ++e) {
    const double u = r.uniform();
    g.el[e] = u < 0.85 ? 1 : (u < 0.98 ? 2 : 3);
  }
  if (!six_rings.empty()) {
    // edge index of tree edge (parent[i], i) is i-1
    for (size_t k = 0; k < six_rings.size(); ++k) {
      const bool aromatic = r.uniform() < 0.7;
      if (!aromatic) continue;
      uint32_t a = six_rings[k].first, b = six_rings[k].second;
      // closure edge index
      for (size_t e = n - 1; e < g.eu.size(); ++e)
        if ((g.eu[e] == a && g.ev[e] == b)) g.el[e] = 4;
      while (a != b) {
        if (depth[a] >= depth[b]) {
          g.el[a - 1] = 4;
          a = parent[a];
        } else {
          g.el[b - 1] = 4;
          b = parent[b];
        }
      }
    }
  }
  m.adj.assign(n, {});
  for (size_t e = 0; e < g.eu.size(); ++e) {
    m.adj[g.eu[e]].push_back(g.ev[e]);
    m.adj[g.ev[e]].push_back(g.eu[e]);
  }
  for (auto& l : m.adj) std::sort(l.begin(), l.end());
}

static void gen_range(uint64_t rs, uint64_t first, uint64_t count, int dist, sigmo_host_graphs& out) {
  Molecule m;
  for (uint64_t i = first; i < first + count; ++i) {
    Rng r(item_seed(rs, i));
    build_molecule(r, dist, m);
    sigmo_host::append_graph(out, m.g);
  }
}

// Fisher-Yates permutation of [0,n) from a role stream.
static std::vector<uint32_t> permutation(uint64_t seed, uint32_t n) {
  std::vector<uint32_t> p(n);
  for (uint32_t i = 0; i < n; ++i) p[i] = i;
  Rng r(seed);
  for (uint32_t i = n; i > 1; --i) std::swap(p[i - 1], p[r.below(i)]);
  return p;
}

}  // namespace sigmo_gen

using namespace sigmo_gen;

extern "C" {

int sigmo_gen_molecules(uint64_t seed, uint64_t first, uint64_t count, int size_dist, int threads, sigmo_host_graphs** out) {
  if (!out) return sigmo_err::set(SIGMO_ERR_INVALID, "null out");
  if (size_dist != SIGMO_SIZE_NORMAL && size_dist != SIGMO_SIZE_LOGNORMAL)
    return sigmo_err::set(SIGMO_ERR_INVALID, "unknown size distribution");
  if (count > (1ull << 32) / 4) return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "too many molecules for one batch");
  const uint64_t rs = role_seed(seed, ROLE_MOLECULES);
  if (threads <= 0) threads = static_cast<int>(std::max(1u, std::thread::hardware_concurrency()));
  if (count < 4096) threads = 1;
  const uint64_t T = static_cast<uint64_t>(threads);
  std::vector<sigmo_host_graphs> parts(T);
  std::vector<std::thread> pool;
  for (uint64_t t = 0; t < T; ++t) {
    const uint64_t lo = count * t / T, hi = count * (t + 1) / T;
    pool.emplace_back([&, t, lo, hi] { gen_range(rs, first + lo, hi - lo, size_dist, parts[t]); });
  }
  for (auto& th : pool) th.join();
  auto* b = new sigmo_host_graphs();
  uint64_t N = 0, E = 0;
  for (auto& p : parts) {
    N += p.total_nodes();
    E += p.total_edges();
  }
  if (N >= (1ull << 32) || E >= (1ull << 32)) {
    delete b;
    return sigmo_err::set(SIGMO_ERR_OVERFLOW, "batch exceeds u32 node/edge ids");
  }
  b->graph_offsets.reserve(count + 1);
  b->row_offsets.reserve(N + 1);
  b->column_indices.reserve(E);
  b->edge_labels.reserve(E);
  b->node_labels.reserve(N);
  for (auto& p : parts) {
    sigmo_host::append_batch(*b, p, 0, p.num_graphs());
    p = sigmo_host_graphs();
  }
  *out = b;
  return SIGMO_OK;
}

int sigmo_gen_queries(uint64_t seed, uint32_t count, uint32_t kmin, uint32_t kmax, int size_dist, double perturb_frac,
                      double ring_frac, sigmo_host_graphs** out) {
  if (!out) return sigmo_err::set(SIGMO_ERR_INVALID, "null out");
  if (kmin < 2 || kmax < kmin || kmax > SIGMO_MAX_QUERY_NODES)
    return sigmo_err::set(SIGMO_ERR_INVALID, "query size range must satisfy 2 <= kmin <= kmax <= 32");
  if (size_dist != SIGMO_SIZE_NORMAL && size_dist != SIGMO_SIZE_LOGNORMAL)
    return sigmo_err::set(SIGMO_ERR_INVALID, "unknown size distribution");
  const uint32_t max_n = size_dist == SIGMO_SIZE_NORMAL ? 64 : 128;
  if (kmax > max_n) return sigmo_err::set(SIGMO_ERR_INVALID, "kmax above molecule size range");
  const uint64_t rs = role_seed(seed, ROLE_QUERIES);
  std::vector<uint8_t> want_ring(count, 0), perturb(count, 0);
  if (ring_frac >= 0) {
    auto p = permutation(role_seed(seed, ROLE_RINGS), count);
    const uint32_t nr = static_cast<uint32_t>(std::floor(ring_frac * count + 1e-9));
    for (uint32_t i = 0; i < nr && i < count; ++i) want_ring[p[i]] = 1;
  }
  if (perturb_frac > 0) {
    auto p = permutation(role_seed(seed, ROLE_PERTURB), count);
    const uint32_t np = static_cast<uint32_t>(std::floor(perturb_frac * count + 1e-9));
    for (uint32_t i = 0; i < np && i < count; ++i) perturb[p[i]] = 1;
  }
  auto* b = new sigmo_host_graphs();
  Molecule m;
  std::vector<uint32_t> order, pos;
  for (uint32_t qi = 0; qi < count; ++qi) {
    Rng r(item_seed(rs, qi));
    const uint32_t k = kmin + static_cast<uint32_t>(r.below(kmax - kmin + 1));
    sigmo_host::EdgeListGraph qg;
    bool done = false;
    for (int attempt = 0; attempt < 100000 && !done; ++attempt) {
      build_molecule(r, size_dist, m);
      const uint32_t n = static_cast<uint32_t>(m.g.labels.size());
      if (n < k) continue;
      const uint32_t start = static_cast<uint32_t>(r.below(n));
      order.clear();
      pos.assign(n, UINT32_MAX);
      order.push_back(start);
      pos[start] = 0;
      for (size_t h = 0; h < order.size() && order.size() < k; ++h)
        for (uint32_t w : m.adj[order[h]])
          if (pos[w] == UINT32_MAX && order.size() < k) {
            pos[w] = static_cast<uint32_t>(order.size());
            order.push_back(w);
          }
      if (order.size() < k) continue;
      qg.labels.assign(k, 0);
      qg.eu.clear();
      qg.ev.clear();
      qg.el.clear();
      for (uint32_t i = 0; i < k; ++i) qg.labels[i] = m.g.labels[order[i]];
      // induced edges, sorted by (new u, new v)
      std::vector<std::pair<std::pair<uint32_t, uint32_t>, uint8_t>> es;
      for (size_t e = 0; e < m.g.eu.size(); ++e) {
        uint32_t a = pos[m.g.eu[e]], c = pos[m.g.ev[e]];
        if (a == UINT32_MAX || c == UINT32_MAX) continue;
        if (a > c) std::swap(a, c);
        es.push_back({{a, c}, m.g.el[e]});
      }
      std::sort(es.begin(), es.end());
      const bool cyclic = es.size() >= k;
      if (ring_frac >= 0 && cyclic != static_cast<bool>(want_ring[qi])) continue;
      for (auto& e : es) {
        qg.eu.push_back(e.first.first);
        qg.ev.push_back(e.first.second);
        qg.el.push_back(e.second);
      }
      done = true;
    }
    if (!done) {
      delete b;
      return sigmo_err::set(SIGMO_ERR_UNSUPPORTED, "query generation did not converge");
    }
    if (perturb[qi]) {
      const uint32_t node = static_cast<uint32_t>(r.below(k));
      uint8_t lab;
      do lab = draw_atom(r);
      while (lab == qg.labels[node]);
      qg.labels[node] = lab;
    }
    sigmo_host::append_graph(*b, qg);
  }
  *out = b;
  return SIGMO_OK;
}

int sigmo_gen_arrivals(uint64_t seed, uint32_t count, double qps, uint32_t qmin, uint32_t qmax, double* arrival_s,
                       uint32_t* queries_per_request) {
  if (!arrival_s || !queries_per_request || qps <= 0 || qmin == 0 || qmax < qmin)
    return sigmo_err::set(SIGMO_ERR_INVALID, "bad arrival arguments");
  Rng r(role_seed(seed, ROLE_ARRIVALS));
  double t = 0;
  for (uint32_t i = 0; i < count; ++i) {
    t += -std::log(1.0 - r.uniform()) / qps;
    arrival_s[i] = t;
    queries_per_request[i] = qmin + static_cast<uint32_t>(r.below(qmax - qmin + 1));
  }
  return SIGMO_OK;
}

}  // extern "C"
