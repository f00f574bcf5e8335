// Prints the lines io::loadFileLinesShard returns for one rank (a NUL byte as "<NUL>"), one a
// line, then "graphs N" with N the graphs io::loadCSRGraphsFromLines builds from them (-1 when
// it rejects them).  Usage: shard_lines FILE MAX_LINES RANK NPROCS
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>

#include "sigmo.hpp"

int main(int argc, char** argv) {
  if (argc != 5) {
    std::fprintf(stderr, "usage: %s FILE MAX_LINES RANK NPROCS\n", argv[0]);
    return 2;
  }
  const auto lines = sigmo::io::loadFileLinesShard(argv[1], std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]));
  for (const std::string& l : lines) {
    std::string out;
    for (char c : l) out += c == '\0' ? std::string("<NUL>") : std::string(1, c);
    std::printf("%s\n", out.c_str());
  }
  long n = -1;
  try {
    n = static_cast<long>(sigmo::io::loadCSRGraphsFromLines(lines).size());
  } catch (const std::exception&) {
    n = -1;
  }
  std::printf("graphs %ld\n", n);
  return 0;
}
