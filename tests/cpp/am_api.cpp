// Host half of the adjacency-matrix API of include/sigmo.hpp (io::loadAMGraphsFromFile,
// utils::getNumOfAdjacencyIntegers, utils::adjacency_matrix::{getNeighbors, isNeighbor}).
// Prints, per graph: "g <nodes> <words> <word hex>..." and per node "n <neighbours, NULL
// terminated>" plus "a <u> <v>" for every pair isNeighbor reports; no device is used.
#include <cstdio>

#include "sigmo.hpp"

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  auto graphs = sigmo::io::loadAMGraphsFromFile(argv[1]);
  for (auto& g : graphs) {
    const int n = g.getNumNodes();
    const uint8_t words = sigmo::utils::getNumOfAdjacencyIntegers(static_cast<uint8_t>(n));
    std::printf("g %d %u", n, unsigned(words));
    for (uint8_t w = 0; w < words; ++w) std::printf(" %llx", static_cast<unsigned long long>(g.getAdjacencyMatrix()[w]));
    std::printf("\n");
    for (int u = 0; u < n; ++u) {
      sigmo::types::node_t nb[sigmo::types::MAX_NEIGHBORS];
      sigmo::utils::adjacency_matrix::getNeighbors(g.getAdjacencyMatrix(), words, u + 100, nb, 100);
      std::printf("n");
      for (size_t i = 0; i < sigmo::types::MAX_NEIGHBORS && nb[i] != sigmo::types::NULL_NODE; ++i) std::printf(" %u", nb[i] - 100);
      std::printf("\n");
      for (int v = 0; v < n; ++v)
        if (sigmo::utils::adjacency_matrix::isNeighbor(g.getAdjacencyMatrix(), words, u, v)) std::printf("a %d %d\n", u, v);
    }
  }
  return 0;
}
