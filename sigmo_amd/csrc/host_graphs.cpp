// Host batched-CSR helpers, the .dat reader/writer and the host-batch C ABI.
//
// The .dat grammar is the reference's (library/include/graph.hpp:154-206):
//   "n#N l#L" then N pairs "id label", then "e#E" then E triples "u v label" or E pairs
//   "u v" (edge label 0); anything else is "Invalid graph format".
#include "host_graphs.hpp"

#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <sstream>
#include <stdexcept>

namespace sigmo_err {
static thread_local std::string g_last_error;
int set(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}
const char* get() { return g_last_error.c_str(); }
void clear() { g_last_error.clear(); }
}  // namespace sigmo_err

extern "C" const char* sigmo_last_error(void) { return sigmo_err::get(); }

namespace sigmo_host {

void append_graph(sigmo_host_graphs& b, const EdgeListGraph& g) {
  const uint32_t n = static_cast<uint32_t>(g.labels.size());
  const uint32_t node_base = static_cast<uint32_t>(b.node_labels.size());
  const uint32_t col_base = static_cast<uint32_t>(b.column_indices.size());
  // counting sort of the directed entries (graph.hpp:221-244)
  std::vector<uint32_t> rows(n + 1, 0);
  for (size_t i = 0; i < g.eu.size(); ++i) {
    rows[g.eu[i] + 1]++;
    rows[g.ev[i] + 1]++;
  }
  for (uint32_t i = 1; i <= n; ++i) rows[i] += rows[i - 1];
  const uint32_t m = rows[n];
  b.column_indices.resize(col_base + m);
  b.edge_labels.resize(col_base + m);
  std::vector<uint32_t> cur(n, 0);
  for (size_t i = 0; i < g.eu.size(); ++i) {
    const uint32_t a = g.eu[i], c = g.ev[i];
    b.column_indices[col_base + rows[a] + cur[a]] = c + node_base;
    b.edge_labels[col_base + rows[a] + cur[a]] = g.el[i];
    b.column_indices[col_base + rows[c] + cur[c]] = a + node_base;
    b.edge_labels[col_base + rows[c] + cur[c]] = g.el[i];
    cur[a]++;
    cur[c]++;
  }
  for (uint32_t i = 1; i <= n; ++i) b.row_offsets.push_back(col_base + rows[i]);
  b.node_labels.insert(b.node_labels.end(), g.labels.begin(), g.labels.end());
  b.graph_offsets.push_back(node_base + n);
}

void append_batch(sigmo_host_graphs& dst, const sigmo_host_graphs& src, uint32_t first, uint32_t count) {
  const uint32_t node_lo = src.graph_offsets[first];
  const uint32_t node_hi = src.graph_offsets[first + count];
  const uint32_t col_lo = src.row_offsets[node_lo];
  const uint32_t col_hi = src.row_offsets[node_hi];
  const uint32_t node_base = static_cast<uint32_t>(dst.node_labels.size());
  const uint32_t col_base = static_cast<uint32_t>(dst.column_indices.size());
  for (uint32_t g = first; g < first + count; ++g) dst.graph_offsets.push_back(src.graph_offsets[g + 1] - node_lo + node_base);
  for (uint32_t v = node_lo; v < node_hi; ++v) dst.row_offsets.push_back(src.row_offsets[v + 1] - col_lo + col_base);
  if (dst.column_indices.capacity() < col_base + (col_hi - col_lo))  // grow geometrically: appends repeat
    dst.column_indices.reserve(std::max<size_t>(col_base + (col_hi - col_lo), 2 * dst.column_indices.capacity()));
  for (uint32_t e = col_lo; e < col_hi; ++e) dst.column_indices.push_back(src.column_indices[e] - node_lo + node_base);
  dst.edge_labels.insert(dst.edge_labels.end(), src.edge_labels.begin() + col_lo, src.edge_labels.begin() + col_hi);
  dst.node_labels.insert(dst.node_labels.end(), src.node_labels.begin() + node_lo, src.node_labels.begin() + node_hi);
}

std::string validate(const sigmo_host_graphs& g) {
  const uint64_t G = g.num_graphs(), N = g.total_nodes(), E = g.total_edges();
  if (g.graph_offsets.empty() || g.graph_offsets[0] != 0) return "graph_offsets[0] must be 0";
  if (g.graph_offsets.back() != N) return "graph_offsets[G] != total nodes";
  if (g.row_offsets.size() != N + 1 || g.row_offsets[0] != 0 || g.row_offsets.back() != E)
    return "row_offsets must have N+1 entries from 0 to E";
  if (g.edge_labels.size() != E) return "edge_labels must have E entries";
  for (uint64_t i = 0; i < G; ++i)
    if (g.graph_offsets[i + 1] < g.graph_offsets[i]) return "graph_offsets not monotone";
  for (uint64_t v = 0; v < N; ++v)
    if (g.row_offsets[v + 1] < g.row_offsets[v]) return "row_offsets not monotone";
  for (uint64_t gi = 0; gi < G; ++gi) {
    const uint32_t lo = g.graph_offsets[gi], hi = g.graph_offsets[gi + 1];
    for (uint32_t v = lo; v < hi; ++v)
      for (uint32_t e = g.row_offsets[v]; e < g.row_offsets[v + 1]; ++e)
        if (g.column_indices[e] < lo || g.column_indices[e] >= hi) return "edge leaves its graph (graph " + std::to_string(gi) + ")";
  }
  // Adjacency must be symmetric, as the reference's CSR always is (toCSRGraph appends every
  // undirected edge in both directions, graph.hpp:216-247): the validity test of
  // isomorphism.hpp:150-168 reads query row i at column `depth` and data row mapping[i] at column
  // c, and the kernels read each pair of rows in whichever direction suits their layout.
  for (uint64_t v = 0; v < N; ++v)
    for (uint32_t e = g.row_offsets[v]; e < g.row_offsets[v + 1]; ++e) {
      const uint32_t u = g.column_indices[e];
      bool back = false;
      for (uint32_t f = g.row_offsets[u]; f < g.row_offsets[u + 1] && !back; ++f) back = g.column_indices[f] == v;
      if (!back) return "adjacency is not symmetric: node " + std::to_string(u) + " does not list neighbour " + std::to_string(v);
    }
  return "";
}

// ---- .dat parsing (graph.hpp:154-206) ---------------------------------------------------

static bool parse_u64(const std::string& t, uint64_t& v) {
  if (t.empty()) return false;
  char* end = nullptr;
  errno = 0;
  unsigned long long x = std::strtoull(t.c_str(), &end, 10);
  if (errno || *end != '\0') return false;
  v = x;
  return true;
}

static bool parse_i64(const std::string& t, long long& v) {
  if (t.empty()) return false;
  char* end = nullptr;
  errno = 0;
  long long x = std::strtoll(t.c_str(), &end, 10);
  if (errno || *end != '\0') return false;
  v = x;
  return true;
}

static bool parse_header(const std::string& tok, const char* prefix, uint64_t& v) {
  if (tok.size() < 3 || tok.compare(0, 2, prefix) != 0) return false;
  return parse_u64(tok.substr(2), v);
}

// Parse one line into an edge-list graph; throws std::runtime_error on malformed input.
static EdgeListGraph parse_line(const std::string& line) {
  std::istringstream iss(line);
  std::string tok;
  uint64_t n = 0, nl = 0, ne = 0;
  if (!(iss >> tok) || !parse_header(tok, "n#", n)) throw std::runtime_error("Invalid graph format: expected n#N");
  if (!(iss >> tok) || !parse_header(tok, "l#", nl)) throw std::runtime_error("Invalid graph format: expected l#L");
  EdgeListGraph g;
  g.labels.assign(n, 0);
  for (uint64_t i = 0; i < n; ++i) {
    std::string a, b;
    uint64_t id;
    long long lab;
    if (!(iss >> a >> b) || !parse_u64(a, id) || !parse_i64(b, lab) || id >= n)
      throw std::runtime_error("Invalid graph format: bad node record");
    g.labels[id] = static_cast<uint8_t>(lab);  // graph.hpp:174 stores through an int -> uint8
  }
  if (!(iss >> tok) || !parse_header(tok, "e#", ne)) throw std::runtime_error("Invalid graph format: expected e#E");
  std::vector<std::string> toks;
  while (iss >> tok) toks.push_back(tok);
  const bool labelled = toks.size() == ne * 3;
  if (!labelled && toks.size() != ne * 2) throw std::runtime_error("Invalid graph format");
  const size_t stride = labelled ? 3 : 2;
  for (uint64_t i = 0; i < ne; ++i) {
    uint64_t u, v;
    long long l = 0;
    if (!parse_u64(toks[i * stride], u) || !parse_u64(toks[i * stride + 1], v) || (labelled && !parse_i64(toks[i * stride + 2], l)))
      throw std::runtime_error("Invalid graph format: bad edge record");
    if (u >= n || v >= n) throw std::runtime_error("Invalid graph format: edge endpoint out of range");
    g.eu.push_back(static_cast<uint32_t>(u));
    g.ev.push_back(static_cast<uint32_t>(v));
    g.el.push_back(static_cast<uint8_t>(l));
  }
  return g;
}

static bool blank(const std::string& s) {
  for (char c : s)
    if (!std::isspace(static_cast<unsigned char>(c))) return false;
  return true;
}

static int parse_stream(std::istream& in, sigmo_host_graphs** out) {
  auto* b = new sigmo_host_graphs();
  std::string line;
  size_t lineno = 0;
  try {
    while (std::getline(in, line)) {
      ++lineno;
      if (blank(line)) continue;
      append_graph(*b, parse_line(line));
    }
  } catch (const std::exception& e) {
    delete b;
    return sigmo_err::set(SIGMO_ERR_INVALID, std::string(e.what()) + " (line " + std::to_string(lineno) + ")");
  }
  *out = b;
  return SIGMO_OK;
}

// ---- one rank's share of a .dat file (library/src/sigmo_mpi.cpp:18-80) ------------------
//
// Rank r of p reads the bytes [r * (size / p), r * (size / p) + 2048 * max_lines + 1024) of the
// file, the bytes past its end reading as NUL (:32-44).  A rank other than the first drops
// everything up to and including its first '\n' (:49-54), even when its range starts on a line;
// a rank other than the last drops its last '\n' and what follows (:57-62).  The rest splits at
// '\n' as std::getline does (a final '\n' ends the last line, an empty segment between two '\n'
// is a line) into at most max_lines lines (:65-70), and only then is a line kept if it holds
// exactly one 'n', one 'e' and one 'l' character (:72-77).  A run of NULs past the end of the
// file is kept as one NUL: the parser splits tokens at whitespace and a number's digits end at
// a NUL, so a run of any length yields the same tokens and values.  The byte count is computed
// in 64 bits (the reference multiplies max_lines by 2048 in int).
static int shard_lines(const char* path, uint32_t rank, uint32_t nprocs, uint32_t max_lines, std::string& text, uint32_t& kept) {
  if (!path) return sigmo_err::set(SIGMO_ERR_INVALID, "null path");
  if (nprocs == 0 || rank >= nprocs) return sigmo_err::set(SIGMO_ERR_INVALID, "rank must be below nprocs");
  FILE* f = std::fopen(path, "rb");
  if (!f) return sigmo_err::set(SIGMO_ERR_IO, std::string("cannot open ") + path);
  if (fseeko(f, 0, SEEK_END) != 0) {
    std::fclose(f);
    return sigmo_err::set(SIGMO_ERR_IO, std::string("cannot size ") + path);
  }
  const off_t end = ftello(f);
  if (end < 0) {
    std::fclose(f);
    return sigmo_err::set(SIGMO_ERR_IO, std::string("cannot size ") + path);
  }
  const uint64_t size = static_cast<uint64_t>(end);
  const uint64_t start = static_cast<uint64_t>(rank) * (size / nprocs);
  const uint64_t span = 2048ull * max_lines + 1024ull;
  const uint64_t n = start < size ? std::min(span, size - start) : 0;
  std::string data(n, '\0');
  if (n && (fseeko(f, static_cast<off_t>(start), SEEK_SET) != 0 || std::fread(&data[0], 1, n, f) != n)) {
    std::fclose(f);
    return sigmo_err::set(SIGMO_ERR_IO, std::string("short read from ") + path);
  }
  std::fclose(f);
  if (n < span) data.push_back('\0');  // the NUL run past the end of the file
  if (rank != 0) {
    const size_t p = data.find('\n');
    if (p != std::string::npos) data.erase(0, p + 1);
  }
  if (rank != nprocs - 1) {
    const size_t p = data.rfind('\n');
    if (p != std::string::npos) data.resize(p);
  }
  text.clear();
  kept = 0;
  uint32_t taken = 0;
  size_t at = 0;
  while (at < data.size() && taken < max_lines) {
    size_t e = data.find('\n', at);
    if (e == std::string::npos) e = data.size();
    ++taken;
    const char* b = data.data() + at;
    const char* z = data.data() + e;
    if (std::count(b, z, 'n') == 1 && std::count(b, z, 'e') == 1 && std::count(b, z, 'l') == 1) {
      text.append(b, z);
      text.push_back('\n');
      ++kept;
    }
    at = e + 1;
  }
  return SIGMO_OK;
}

}  // namespace sigmo_host

using namespace sigmo_host;

extern "C" {

int sigmo_host_graphs_from_arrays(uint32_t num_graphs, const uint32_t* graph_offsets, const uint32_t* row_offsets,
                                  const uint32_t* column_indices, const uint8_t* node_labels, const uint8_t* edge_labels,
                                  sigmo_host_graphs** out) {
  if (!out || !graph_offsets || !row_offsets) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  auto* b = new sigmo_host_graphs();
  const uint32_t N = graph_offsets[num_graphs];
  const uint32_t E = row_offsets[N];
  if ((N && !node_labels) || (E && (!column_indices || !edge_labels))) {
    delete b;
    return sigmo_err::set(SIGMO_ERR_INVALID, "null array with nonzero size");
  }
  b->graph_offsets.assign(graph_offsets, graph_offsets + num_graphs + 1);
  b->row_offsets.assign(row_offsets, row_offsets + N + 1);
  b->column_indices.assign(column_indices, column_indices + E);
  b->node_labels.assign(node_labels, node_labels + N);
  b->edge_labels.assign(edge_labels, edge_labels + E);
  std::string err = validate(*b);
  if (!err.empty()) {
    delete b;
    return sigmo_err::set(SIGMO_ERR_INVALID, err);
  }
  *out = b;
  return SIGMO_OK;
}

int sigmo_host_graphs_load_dat(const char* path, sigmo_host_graphs** out) {
  if (!path || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  std::ifstream f(path);
  if (!f) return sigmo_err::set(SIGMO_ERR_IO, std::string("cannot open ") + path);
  return parse_stream(f, out);
}

int sigmo_host_graphs_parse_dat(const char* text, size_t len, sigmo_host_graphs** out) {
  if (!out || (!text && len)) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  std::istringstream s(std::string(text ? text : "", len));
  return parse_stream(s, out);
}

int sigmo_dat_shard_lines(const char* path, uint32_t rank, uint32_t nprocs, uint32_t max_lines, char* text, size_t cap, size_t* len,
                          uint32_t* num_lines) {
  if (!len) return sigmo_err::set(SIGMO_ERR_INVALID, "null length");
  std::string t;
  uint32_t kept = 0;
  int st;
  try {
    st = shard_lines(path, rank, nprocs, max_lines, t, kept);
  } catch (const std::bad_alloc&) {
    return sigmo_err::set(SIGMO_ERR_NOMEM, "shard does not fit host memory");
  }
  if (st != SIGMO_OK) return st;
  *len = t.size();
  if (num_lines) *num_lines = kept;
  if (text && cap >= t.size()) std::memcpy(text, t.data(), t.size());
  return SIGMO_OK;
}

int sigmo_host_graphs_load_dat_shard(const char* path, uint32_t rank, uint32_t nprocs, uint32_t max_lines, sigmo_host_graphs** out) {
  if (!out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  std::string t;
  uint32_t kept = 0;
  int st;
  try {
    st = shard_lines(path, rank, nprocs, max_lines, t, kept);
  } catch (const std::bad_alloc&) {
    return sigmo_err::set(SIGMO_ERR_NOMEM, "shard does not fit host memory");
  }
  if (st != SIGMO_OK) return st;
  std::istringstream s(t);
  return parse_stream(s, out);
}

int sigmo_host_graphs_save_dat(const sigmo_host_graphs* g, const char* path) {
  if (!g || !path) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  FILE* f = std::fopen(path, "w");
  if (!f) return sigmo_err::set(SIGMO_ERR_IO, std::string("cannot open ") + path);
  std::string line;
  for (uint32_t gi = 0; gi < g->num_graphs(); ++gi) {
    const uint32_t lo = g->graph_offsets[gi], hi = g->graph_offsets[gi + 1];
    line = "n#" + std::to_string(hi - lo) + " l#19";
    for (uint32_t v = lo; v < hi; ++v) line += " " + std::to_string(v - lo) + " " + std::to_string(g->node_labels[v]);
    // each undirected edge once, from its smaller endpoint, in row order
    std::string edges;
    uint32_t ne = 0;
    for (uint32_t v = lo; v < hi; ++v)
      for (uint32_t e = g->row_offsets[v]; e < g->row_offsets[v + 1]; ++e) {
        const uint32_t w = g->column_indices[e];
        if (w > v) {
          edges += " " + std::to_string(v - lo) + " " + std::to_string(w - lo) + " " + std::to_string(g->edge_labels[e]);
          ++ne;
        }
      }
    line += " e#" + std::to_string(ne) + edges + "\n";
    if (std::fputs(line.c_str(), f) < 0) {
      std::fclose(f);
      return sigmo_err::set(SIGMO_ERR_IO, "write failed");
    }
  }
  std::fclose(f);
  return SIGMO_OK;
}

int sigmo_host_graphs_info(const sigmo_host_graphs* g, uint32_t* num_graphs, uint64_t* total_nodes, uint64_t* total_edges) {
  if (!g) return sigmo_err::set(SIGMO_ERR_INVALID, "null graphs");
  if (num_graphs) *num_graphs = g->num_graphs();
  if (total_nodes) *total_nodes = g->total_nodes();
  if (total_edges) *total_edges = g->total_edges();
  return SIGMO_OK;
}

int sigmo_host_graphs_arrays(const sigmo_host_graphs* g, const uint32_t** graph_offsets, const uint32_t** row_offsets,
                             const uint32_t** column_indices, const uint8_t** node_labels, const uint8_t** edge_labels) {
  if (!g) return sigmo_err::set(SIGMO_ERR_INVALID, "null graphs");
  if (graph_offsets) *graph_offsets = g->graph_offsets.data();
  if (row_offsets) *row_offsets = g->row_offsets.data();
  if (column_indices) *column_indices = g->column_indices.data();
  if (node_labels) *node_labels = g->node_labels.data();
  if (edge_labels) *edge_labels = g->edge_labels.data();
  return SIGMO_OK;
}

int sigmo_host_graphs_slice(const sigmo_host_graphs* g, uint32_t first, uint32_t count, sigmo_host_graphs** out) {
  if (!g || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (static_cast<uint64_t>(first) + count > g->num_graphs()) return sigmo_err::set(SIGMO_ERR_INVALID, "slice out of range");
  auto* b = new sigmo_host_graphs();
  append_batch(*b, *g, first, count);
  *out = b;
  return SIGMO_OK;
}

int sigmo_host_graphs_concat(const sigmo_host_graphs* a, const sigmo_host_graphs* b, sigmo_host_graphs** out) {
  if (!a || !b || !out) return sigmo_err::set(SIGMO_ERR_INVALID, "null argument");
  if (a->total_nodes() + b->total_nodes() >= (1ull << 32) || a->total_edges() + b->total_edges() >= (1ull << 32))
    return sigmo_err::set(SIGMO_ERR_OVERFLOW, "concatenation exceeds u32 ids");
  auto* r = new sigmo_host_graphs();
  append_batch(*r, *a, 0, a->num_graphs());
  append_batch(*r, *b, 0, b->num_graphs());
  *out = r;
  return SIGMO_OK;
}

int sigmo_host_graphs_free(sigmo_host_graphs* g) {
  delete g;
  return SIGMO_OK;
}

}  // extern "C"
