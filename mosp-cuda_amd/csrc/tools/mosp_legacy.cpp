// mosp_legacy -- the reference's file-based wrappers as commands:
//
//   mosp_legacy sosp <csrPrefix> <distancesIn> <treeIn> <insert.txt> <delete.txt>
//                    <objective> <source> <distancesOut> <treeOut>
//       parallelSOSPUpdate (headers/parallelSOSPUpdate.cuh:32-39 of the
//       reference): the update of one objective's tree.
//   mosp_legacy combined <csrPrefix> <K> <source> <distancesOut> <treeOut>
//                        <pref: p1,..,pK or -> <tree_0> .. <tree_{K-1}>
//       parallelCombinedGraph (headers/parallelCombinedGraph.cuh:46-53): the
//       combined graph of K trees and its SOSP tree.
// Exit status 0 on success, 1 on an error (message on stdout), 2 on usage.
#include <cstdlib>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "csrGraph.h"
#include "parallelCombinedGraph.h"
#include "parallelSOSPUpdate.h"

namespace {

int usage() {
  std::cerr << "usage: mosp_legacy sosp <csrPrefix> <distancesIn> <treeIn> <insert.txt> "
               "<delete.txt> <objective> <source> <distancesOut> <treeOut>\n"
               "       mosp_legacy combined <csrPrefix> <K> <source> <distancesOut> "
               "<treeOut> <p1,..,pK|-> <tree_0> .. <tree_{K-1}>\n";
  return 2;
}

bool integer(const char *text, long long lo, long long hi, int &out) {
  long long v = 0;
  if (!parseInteger(text, lo, hi, v)) {
    std::cerr << "not an integer in [" << lo << ", " << hi << "]: '" << text << "'\n";
    return false;
  }
  out = static_cast<int>(v);
  return true;
}

}  // namespace

int main(int argc, char **argv) {
  if (argc < 2) return usage();
  const std::string cmd = argv[1];
  if (cmd == "sosp") {
    if (argc != 11) return usage();
    int k = 0, source = 0;
    if (!integer(argv[7], 0, 31, k) || !integer(argv[8], 0, 2147483647, source)) return 2;
    return parallelSOSPUpdate(argv[2], argv[3], argv[4], argv[5], argv[6], k, source, argv[9],
                              argv[10])
               ? 0
               : 1;
  }
  if (cmd == "combined") {
    if (argc < 8) return usage();
    int K = 0, source = 0;
    if (!integer(argv[3], 1, 32, K) || !integer(argv[4], 0, 2147483647, source)) return 2;
    if (argc != 8 + K) return usage();
    std::vector<int> pref;
    if (std::string(argv[7]) != "-") {
      std::istringstream list(argv[7]);
      std::string item;
      while (std::getline(list, item, ',')) {
        int p = 0;
        if (!integer(item.c_str(), -2147483648LL, 2147483647, p)) return 2;  // checked by the call
        pref.push_back(p);
      }
    }
    std::vector<std::string> trees(argv + 8, argv + 8 + K);
    return parallelCombinedGraph(argv[2], trees, K, source, "output/combinedGraph", argv[5],
                                 argv[6], pref)
               ? 0
               : 1;
  }
  return usage();
}
