// sigmo_cli -- command-line driver with the reference's flags and report lines
// (library/src/sigmo.cpp:11-220, flags library/src/utils.hpp:41-87), running the gfx950
// implementation through include/sigmo.hpp.  Its stdout keeps the lines that
// scripts/utils/output_analyzer.py parses ("# Matches:", "Total time:", ...).
//
// Extra flag (no reference counterpart): --generate <seed>:<molecules>[:lognormal] and
// --generate-queries <seed>:<count>:<kmin>:<kmax> build the seeded synthetic inputs instead
// of reading .dat files.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <iostream>
#include <locale>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "sigmo.hpp"

namespace {

struct Args {
  bool print_candidates = false, skip_join = false, query_data = false, find_all = false, skip_print_candidates = false;
  int refinement_steps = 0;
  std::string query_file, data_file, candidates_domain = "query";
  size_t multiply_factor_query = 1, multiply_factor_data = 1;
  size_t max_data_graphs = 1000000, max_query_graphs = 1000;
  bool filter_active = false;
  size_t filter_min = 0, filter_max = static_cast<size_t>(-1);
  std::string gen_data, gen_queries;
};

[[noreturn]] void usage(const char* argv0, int code) {
  std::cout << "Usage: " << argv0 << " -Q <query.dat> -D <data.dat> [options]\n"
            << "  -i, --iterations N          refinement iterations\n"
            << "  -c, --candidates-domain S   query (data is parsed and ignored, as in the reference driver)\n"
            << "  -m, --multiply N            multiply query and data graphs\n"
            << "  -d, --mul-data N            multiply data graphs\n"
            << "  -q, --mul-query N           multiply query graphs\n"
            << "      --skip-join             stop after filtering\n"
            << "      --find-all              count every embedding\n"
            << "      --query-filter min[:max]\n"
            << "      --skip-candidates-analysis\n"
            << "      --max-data-graphs N     (default 1000000)\n"
            << "      --max-query-graphs N    (default 1000)\n"
            << "      --join-work-group N     accepted; launch shapes are chosen by the library\n"
            << "      --filter-work-group N   accepted; launch shapes are chosen by the library\n"
            << "  -p, --print-candidates      per-query-node candidate counts to stderr\n"
            << "      --generate seed:count[:lognormal]       synthetic molecules instead of -D\n"
            << "      --generate-queries seed:count:kmin:kmax  synthetic queries instead of -Q\n";
  std::exit(code);
}

Args parse(int argc, char** argv) {
  Args a;
  auto value = [&](int& i) -> std::string {
    std::string arg = argv[i];
    auto eq = arg.find('=');
    if (eq != std::string::npos && arg.rfind("--", 0) == 0) return arg.substr(eq + 1);
    if (i + 1 >= argc) throw std::runtime_error("missing value for " + arg);
    return argv[++i];
  };
  for (int i = 1; i < argc; ++i) {
    std::string arg = argv[i];
    std::string key = arg.substr(0, arg.find('='));
    if (key == "-h" || key == "--help") usage(argv[0], 0);
    else if (key == "-p" || key == "--print-candidates") a.print_candidates = true;
    else if (key == "-i" || key == "--iterations") a.refinement_steps = std::stoi(value(i));
    else if (key == "-Q") a.query_file = value(i);
    else if (key == "-D") a.data_file = value(i);
    else if (key == "-c" || key == "--candidates-domain") a.candidates_domain = value(i);
    else if (key == "-m" || key == "--multiply") a.multiply_factor_data = a.multiply_factor_query = std::stoul(value(i));
    else if (key == "-d" || key == "--mul-data") a.multiply_factor_data = std::stoul(value(i));
    else if (key == "-q" || key == "--mul-query") a.multiply_factor_query = std::stoul(value(i));
    else if (key == "--skip-join") a.skip_join = true;
    else if (key == "--find-all") a.find_all = true;
    else if (key == "--skip-candidates-analysis") a.skip_print_candidates = true;
    else if (key == "--max-data-graphs") a.max_data_graphs = std::stoul(value(i));
    else if (key == "--max-query-graphs") a.max_query_graphs = std::stoul(value(i));
    else if (key == "--join-work-group") sigmo::device::deviceOptions.join_work_group_size = std::stoul(value(i));
    else if (key == "--filter-work-group") sigmo::device::deviceOptions.filter_work_group_size = std::stoul(value(i));
    else if (key == "--query-filter") {
      a.filter_active = true;
      std::string f = value(i);
      auto c = f.find(':');
      a.filter_min = std::stoul(f.substr(0, c));
      if (c != std::string::npos) a.filter_max = std::stoul(f.substr(c + 1));
    } else if (key == "--generate") a.gen_data = value(i);
    else if (key == "--generate-queries") a.gen_queries = value(i);
    else throw std::runtime_error("unknown option " + arg);
  }
  const bool have_q = !a.query_file.empty() || !a.gen_queries.empty();
  const bool have_d = !a.data_file.empty() || !a.gen_data.empty();
  if (have_q && have_d) a.query_data = true;
  else if (have_q || have_d) throw std::runtime_error("Both query and data files must be provided");
  return a;
}

// src/utils.hpp:127-138: '.' as the thousands separator.
std::string formatNumber(size_t number) {
  std::string s = std::to_string(number), out;
  int n = 0;
  for (auto it = s.rbegin(); it != s.rend(); ++it) {
    if (n && n % 3 == 0) out.push_back('.');
    out.push_back(*it);
    ++n;
  }
  return std::string(out.rbegin(), out.rend());
}

// src/utils.hpp:140-157.
std::string getBytesSize(size_t num_bytes, bool round = true) {
  double bytes = static_cast<double>(num_bytes);
  std::string unit = "B";
  if (num_bytes >= (1ull << 30)) {
    bytes /= double(1ull << 30);
    unit = "GB";
  } else if (num_bytes >= (1ull << 20)) {
    bytes /= double(1ull << 20);
    unit = "MB";
  } else if (num_bytes >= 1024) {
    bytes /= 1024;
    unit = "KB";
  }
  if (round) return std::to_string(static_cast<int>(std::lround(bytes))) + " " + unit;
  return std::to_string(bytes) + " " + unit;
}

std::vector<std::string> split(const std::string& s, char sep) {
  std::vector<std::string> out;
  std::stringstream ss(s);
  std::string t;
  while (std::getline(ss, t, sep)) out.push_back(t);
  return out;
}

std::vector<sigmo::CSRGraph> generated(const std::string& spec, bool queries) {
  auto f = split(spec, ':');
  sigmo::detail::HostBatch b;
  if (!queries) {
    if (f.size() < 2) throw std::runtime_error("--generate expects seed:count[:lognormal]");
    const int dist = f.size() > 2 && f[2] == "lognormal" ? SIGMO_SIZE_LOGNORMAL : SIGMO_SIZE_NORMAL;
    sigmo::detail::check(sigmo_gen_molecules(std::stoull(f[0]), 0, std::stoull(f[1]), dist, 0, &b.h));
  } else {
    if (f.size() < 4) throw std::runtime_error("--generate-queries expects seed:count:kmin:kmax");
    const uint32_t kmax = std::stoul(f[3]);
    const int dist = kmax > 20 ? SIGMO_SIZE_LOGNORMAL : SIGMO_SIZE_NORMAL;
    sigmo::detail::check(sigmo_gen_queries(std::stoull(f[0]), std::stoul(f[1]), std::stoul(f[2]), kmax, dist, 0.0, -1.0, &b.h));
  }
  return sigmo::detail::split(b.h);
}

using ms = std::chrono::milliseconds;
using clk = std::chrono::high_resolution_clock;
long long to_ms(std::chrono::duration<double> d) { return std::chrono::duration_cast<ms>(d).count(); }

}  // namespace

int main(int argc, char** argv) try {
  Args args = parse(argc, argv);
  if (!args.query_data) throw std::runtime_error("Specify input data");
  sigmo::Queue queue{0};

  auto query_graphs = args.gen_queries.empty() ? sigmo::io::loadCSRGraphsFromFile(args.query_file) : generated(args.gen_queries, true);
  auto data_graphs = args.gen_data.empty() ? sigmo::io::loadCSRGraphsFromFile(args.data_file) : generated(args.gen_data, false);
  if (args.filter_active) {  // sigmo.cpp:28-35
    query_graphs.erase(std::remove_if(query_graphs.begin(), query_graphs.end(),
                                      [&](const sigmo::CSRGraph& g) {
                                        return g.getNumNodes() > args.filter_max || g.getNumNodes() < args.filter_min;
                                      }),
                       query_graphs.end());
  }
  const size_t nq = query_graphs.size(), nd = data_graphs.size();
  for (size_t i = 1; i < args.multiply_factor_query; ++i) query_graphs.insert(query_graphs.end(), query_graphs.begin(), query_graphs.begin() + nq);
  for (size_t i = 1; i < args.multiply_factor_data; ++i) data_graphs.insert(data_graphs.end(), data_graphs.begin(), data_graphs.begin() + nd);
  if (query_graphs.size() > args.max_query_graphs) query_graphs.erase(query_graphs.begin() + args.max_query_graphs, query_graphs.end());
  if (data_graphs.size() > args.max_data_graphs) data_graphs.erase(data_graphs.begin() + args.max_data_graphs, data_graphs.end());

  auto device_query_graph = sigmo::createDeviceCSRGraph(queue, query_graphs);
  auto device_data_graph = sigmo::createDeviceCSRGraph(queue, data_graphs);
  const size_t query_nodes = device_query_graph.total_nodes, data_nodes = device_data_graph.total_nodes;
  const size_t data_graph_bytes = sigmo::getDeviceGraphAllocSize(device_data_graph);
  const size_t query_graph_bytes = sigmo::getDeviceGraphAllocSize(device_query_graph);

  std::cout << "------------- Input Data -------------\n"
            << "Reed data graph and query graph\n"
            << "# Query Nodes " << query_nodes << "\n# Query Graphs " << device_query_graph.num_graphs << "\n# Data Nodes " << data_nodes
            << "\n# Data Graphs " << device_data_graph.num_graphs << "\n";
  std::cout << "------------- Configs -------------\n"
            << "Filter domain: " << args.candidates_domain << "\n"
            << "Filter Work Group Size: " << sigmo::device::deviceOptions.filter_work_group_size << "\n"
            << "Join Work Group Size: " << sigmo::device::deviceOptions.join_work_group_size << "\n"
            << "Find all: " << (args.find_all ? "Yes" : "No") << "\n";

  auto setup_start = clk::now();
  std::cout << "------------- Setup Data -------------\n"
            << "Allocated " << getBytesSize(data_graph_bytes) << " for graph data\n"
            << "Allocated " << getBytesSize(query_graph_bytes) << " for query data\n";
  sigmo::candidates::Candidates candidates{queue, query_nodes, data_nodes};
  const size_t candidates_bytes = candidates.getAllocationSize();
  std::cout << "Allocated " << getBytesSize(candidates_bytes) << " for candidates\n";
  sigmo::signature::Signature<> signatures{queue, data_nodes, query_nodes};
  const size_t dsb = signatures.getDataSignatureAllocationSize(), qsb = signatures.getQuerySignatureAllocationSize();
  std::cout << "Allocated " << getBytesSize(dsb) << " for data signatures\n"
            << "Allocated " << getBytesSize(qsb) << " for query signatures\n"
            << "Allocated " << getBytesSize(std::max(dsb, qsb)) << " for temporary buffer\n";
  queue.wait();
  auto setup_end = clk::now();
  std::cout << "Total allocated memory: "
            << getBytesSize(dsb + qsb + candidates_bytes + std::max(dsb, qsb) + data_graph_bytes + query_graph_bytes, false) << " out of "
            << getBytesSize(queue.globalMemSize()) << " available on " << queue.deviceName() << "\n";

  std::vector<std::chrono::duration<double>> data_sig_times, query_sig_times, filter_times;
  std::cout << "------------- Runtime Filter Phase -------------\n";
  auto filter_start = clk::now();
  std::cout << "[*] Initialization Step:\n";
  auto e1 = signatures.generateDataSignatures(device_data_graph);
  queue.wait_and_throw();
  data_sig_times.push_back(e1.getProfilingInfo());
  std::cout << "- Data signatures generated in " << to_ms(data_sig_times.back()) << " ms\n";
  auto e2 = signatures.generateQuerySignatures(device_query_graph);
  queue.wait_and_throw();
  query_sig_times.push_back(e2.getProfilingInfo());
  std::cout << "- Query signatures generated in " << to_ms(query_sig_times.back()) << " ms\n";
  auto e3 = sigmo::isomorphism::filter::filterCandidates(queue, device_query_graph, device_data_graph, signatures, candidates);
  queue.wait_and_throw();
  filter_times.push_back(e3.getProfilingInfo());
  std::cout << "- Candidates filtered in " << to_ms(filter_times.back()) << " ms\n";
  for (int ref_step = 1; ref_step <= args.refinement_steps; ++ref_step) {
    std::cout << "[*] Refinement step " << ref_step << ":\n";
    auto r1 = signatures.refineDataSignatures(device_data_graph, ref_step);
    queue.wait_and_throw();
    data_sig_times.push_back(r1.getProfilingInfo());
    std::cout << "- Data signatures refined in " << to_ms(data_sig_times.back()) << " ms\n";
    auto r2 = signatures.refineQuerySignatures(device_query_graph, ref_step);
    queue.wait_and_throw();
    query_sig_times.push_back(r2.getProfilingInfo());
    std::cout << "- Query signatures refined in " << to_ms(query_sig_times.back()) << " ms\n";
    auto r3 = sigmo::isomorphism::filter::refineCandidates(queue, device_query_graph, device_data_graph, signatures, candidates);
    queue.wait_and_throw();
    filter_times.push_back(r3.getProfilingInfo());
    std::cout << "- Candidates refined in " << to_ms(filter_times.back()) << " ms\n";
  }
  auto filter_end = clk::now();

  std::chrono::duration<double> join_time{0};
  size_t num_matches = 0;
  auto mapping_start = clk::now(), mapping_end = mapping_start, join_start = mapping_start, join_end = mapping_start;
  if (!args.skip_join) {
    std::cout << "[*] Generating DQCR\n";
    mapping_start = clk::now();
    sigmo::isomorphism::mapping::GMCR gmcr{queue};
    gmcr.generateGMCR(device_query_graph, device_data_graph, candidates);
    mapping_end = clk::now();
    std::cout << "[*] Starting Join\n";
    join_start = clk::now();
    auto je = sigmo::isomorphism::join::joinCandidates(queue, device_query_graph, device_data_graph, candidates, gmcr, &num_matches,
                                                       !args.find_all);
    je.wait();
    join_time = je.getProfilingInfo();
    join_end = clk::now();
  }
  std::cout << "[!] End\n";

  auto sum = [](const std::vector<std::chrono::duration<double>>& v) {
    return std::accumulate(v.begin(), v.end(), std::chrono::duration<double>(0));
  };
  const auto tsd = sum(data_sig_times), tsq = sum(query_sig_times), tf = sum(filter_times);
  std::cout << "------------- Overall GPU Stats -------------\n"
            << "Data signature time: " << to_ms(tsd) << " ms\n"
            << "Query signature time: " << to_ms(tsq) << " ms\n"
            << "Filter time: " << to_ms(tf) << " ms\n";
  if (args.skip_join) std::cout << "Join time: skipped\n";
  else std::cout << "Join time: " << to_ms(join_time) << " ms\n";
  std::cout << "Total time: " << to_ms(tsd + tf + tsq + join_time) << " ms\n";
  std::cout << "------------- Overall Host Stats -------------\n"
            << "Setup Data time: " << to_ms(setup_end - setup_start) << " ms (not included in total)\n"
            << "Filter time: " << to_ms(filter_end - filter_start) << " ms\n";
  if (args.skip_join) {
    std::cout << "Mapping time: skipped\nJoin time: skipped\n";
  } else {
    std::cout << "Mapping time: " << to_ms(mapping_end - mapping_start) << " ms\n"
              << "Join time: " << to_ms(join_end - join_start) << " ms\n";
  }
  std::cout << "Total time: " << to_ms(clk::now() - setup_end) << " ms\n";
  std::cout << "------------- Results -------------\n";
  if (!args.skip_print_candidates) {
    auto counts = candidates.getCandidatesCounts();
    std::vector<size_t> sizes(counts.begin(), counts.end());
    size_t total = 0, zero = 0;
    for (size_t i = 0; i < sizes.size(); ++i) {
      total += sizes[i];
      zero += sizes[i] == 0;
      if (args.print_candidates) std::cerr << "Node " << i << ": " << sizes[i] << "\n";
    }
    std::sort(sizes.begin(), sizes.end());
    std::cout << "# Total candidates: " << formatNumber(total) << "\n"
              << "# Average candidates: " << formatNumber(sizes.empty() ? 0 : total / sizes.size()) << "\n"
              << "# Median candidates: " << formatNumber(sizes.empty() ? 0 : sizes[sizes.size() / 2]) << "\n"
              << "# Zero candidates: " << formatNumber(zero) << "\n";
  }
  if (!args.skip_join) std::cout << "# Matches: " << formatNumber(num_matches) << std::endl;
  sigmo::destroyDeviceCSRGraph(device_data_graph, queue);
  sigmo::destroyDeviceCSRGraph(device_query_graph, queue);
  return 0;
} catch (const std::exception& e) {
  std::cerr << "sigmo_cli: " << e.what() << std::endl;
  return 1;
}
