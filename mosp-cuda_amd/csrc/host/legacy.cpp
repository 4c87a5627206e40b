// legacy.cpp -- the reference's file-based wrappers over this library:
// parallelSOSPUpdate (src/parallelSOSPUpdate.cu:216-536) and
// parallelCombinedGraph (src/parallelCombinedGraph.cu:79-164), with their
// signatures, input files and output formats.  They read with the
// library's readers, apply the batch with applyChangeBatch (deletions
// first, then insertions overwrite the first remaining (u,v) or append; a
// weight counts as raised against the original first occurrence) and run
// the same GPU engine as mospUpdate.  Trees are canonical (lowest-id
// parents), so the outputs equal the reference's for the same inputs.
// With instrumentation enabled they record the reference's stage and
// counter names (src/parallelSOSPUpdate.cu:229-497,
// src/parallelCombinedGraph.cu:100-150).
#include <algorithm>
#include <iostream>
#include <string>
#include <vector>

#include "combinedGraphGpu.h"
#include "csrGraph.h"
#include "deviceArray.h"
#include "mosp_hip.h"
#include "deviceGraph.h"
#include "parallelCombinedGraph.h"
#include "parallelSOSPUpdate.h"
#include "sospUpdateGpu.h"
#include "stageTimer.h"

namespace {

// Vertex count = lines of <prefix>RowPtr.txt - 1 (a last value without a
// newline counts as a line); -1 if the file cannot be read.
int count_csr_nodes(const std::string &prefix) {
  FILE *f = std::fopen((prefix + "RowPtr.txt").c_str(), "rb");
  if (f == nullptr) {
    return -1;
  }
  std::vector<char> buf(1 << 20);
  long long lines = 0;
  char last = '\n';
  size_t got;
  while ((got = std::fread(buf.data(), 1, buf.size(), f)) > 0) {
    lines += std::count(buf.begin(), buf.begin() + static_cast<std::ptrdiff_t>(got), '\n');
    last = buf[got - 1];
  }
  std::fclose(f);
  if (last != '\n') {
    ++lines;
  }
  return static_cast<int>(lines - 1);
}

}  // namespace

bool parallelSOSPUpdate(const std::string &originalCsrPrefix,
                        const std::string &distancesInputPath, const std::string &treeInputPath,
                        const std::string &insertPath, const std::string &deletePath,
                        int objectiveIndex, int source, const std::string &distancesOutputPath,
                        const std::string &treeOutputPath) {
  CsrGraph original;
  ScopedStage readStage("sosp/0a_read_csr_text");
  if (!readCsrGraph(originalCsrPrefix, original)) {
    std::cout << "Error: Could not read original CSR graph.\n";
    return false;
  }
  readStage.stop();
  const int n = original.numberOfNodes, K = original.numberOfObjectives;
  if (n == 0) {
    std::cout << "Error: Graph has no vertices.\n";
    return false;
  }
  if (objectiveIndex < 0 || objectiveIndex >= K) {
    std::cout << "Error: objectiveIndex out of range.\n";
    return false;
  }
  if (source < 0 || source >= n) {
    std::cout << "Error: source vertex out of range.\n";
    return false;
  }
  std::vector<long long> distances;
  std::vector<int> parent;
  ChangeBatch batch;
  {
    ScopedStage treeStage("sosp/0c_read_tree_text");
    if (!readDistances(distancesInputPath, n, distances) ||
        !readParents(treeInputPath, n, parent)) {
      return false;
    }
  }
  {
    ScopedStage changesStage("sosp/0d_read_changes_text");
    if (!readChangeBatch(insertPath, deletePath, K, n, batch)) {
      return false;
    }
  }
  CsrGraph updated;
  {
    ScopedStage applyStage("sosp/0e_apply_changes_host");
    if (!applyChangeBatch(original, batch, updated)) {
      return false;
    }
  }
  // Changed edges of this objective: every deletion and every insertion that
  // raised its weight; heads: every insertion's head.
  const int k = objectiveIndex;
  std::vector<int> changedFrom(batch.deleteFrom), changedTo(batch.deleteTo);
  for (int i = 0; i < batch.numberOfInserts(); ++i) {
    if ((batch.weightIncreaseMask[i] >> k) & 1u) {
      changedFrom.push_back(batch.insertFrom[i]);
      changedTo.push_back(batch.insertTo[i]);
    }
  }
  // The bound covers the original and the inserted weights; the width
  // follows the updated graph (src/parallelSOSPUpdate.cu:262-268, 413-420).
  long long maxWeight = 1;
  for (int e = 0; e < original.numberOfEdges(); ++e) {
    maxWeight = std::max<long long>(maxWeight, original.weight(e, k));
  }
  long long weightSum = 0;
  for (int e = 0; e < updated.numberOfEdges(); ++e) {
    weightSum += updated.weight(e, k);
    maxWeight = std::max<long long>(maxWeight, updated.weight(e, k));
  }
  const long long delta = defaultDelta(updated.numberOfEdges(), n, weightSum);

  DeviceGraph device;
  SospWorkspace workspace;
  DeviceArray<int> dFrom, dTo, dHeads, dParent;
  DeviceArray<long long> dDistances;
  ScopedStage uploadStage("sosp/2b_alloc_h2d", true);
  if (!uploadDeviceGraph(updated, device) || !workspace.reserve(n) ||
      !dFrom.upload(changedFrom) || !dTo.upload(changedTo) || !dHeads.upload(batch.insertTo) ||
      !dDistances.upload(distances) || !dParent.upload(parent)) {
    std::cout << "Error: could not copy the graph to the GPU.\n";
    return false;
  }
  DeviceChanges changes;
  changes.changedFrom = dFrom.data();
  changes.changedTo = dTo.data();
  changes.numberOfChanged = static_cast<int>(changedFrom.size());
  changes.insertHeads = dHeads.data();
  changes.numberOfInsertHeads = batch.numberOfInserts();
  uploadStage.stop();
  SospStats stats;
  {
    ScopedStage updateStage("sosp/update_gpu", true);
    if (!sospUpdateGpu(device.out(k), device.in(k), changes, source, delta, maxWeight, workspace,
                       dDistances.data(), dParent.data(), &stats)) {
      std::cout << "Error: SOSP update failed on the GPU.\n";
      return false;
    }
  }
  recordCounter("sosp/invalidated", stats.invalidated);
  recordCounter("sosp/jump_rounds", stats.jumpRounds);
  recordCounter("sosp/iterations", stats.iterations);
  recordCounter("sosp/epochs", stats.epochs);
  recordCounter("sosp/pushes", static_cast<double>(stats.pushes));
  {
    ScopedStage downloadStage("sosp/4_d2h", true);
    if (!dDistances.download(distances) || !dParent.download(parent)) {
      std::cout << "Error: could not copy the result from the GPU.\n";
      return false;
    }
  }
  ScopedStage writeStage("sosp/5_write_text");
  return writeDistances(distancesOutputPath, distances) &&
         writeParents(treeOutputPath, parent);
}

bool parallelCombinedGraph(const std::string &originalCsrPrefix,
                           const std::vector<std::string> &treeInputPaths, int K, int source,
                           const std::string & /*workDir*/,
                           const std::string &distancesOutputPath,
                           const std::string &treeOutputPath,
                           const std::vector<int> &preferences) {
  if (K <= 0 || static_cast<int>(treeInputPaths.size()) < K) {
    std::cout << "Error: K=" << K << " but only " << treeInputPaths.size()
              << " tree paths supplied.\n";
    return false;
  }
  if (K > 32) {
    std::cout << "Error: at most 32 objectives are supported.\n";
    return false;
  }
  if (preferenceScale(preferences, K) == 0) {
    std::cout << "Error: invalid preference vector (need K values >= 1).\n";
    return false;
  }
  const int n = count_csr_nodes(originalCsrPrefix);
  if (n <= 0) {
    std::cout << "Error: Could not read the vertex count of " << originalCsrPrefix << "\n";
    return false;
  }
  if (source < 0 || source >= n) {
    std::cout << "Error: source vertex " << source << " out of range.\n";
    return false;
  }
  std::vector<int> parents(static_cast<size_t>(n) * K);
  ScopedStage readStage("comb/0_read_trees_text");
  for (int k = 0; k < K; ++k) {
    std::vector<int> parent;
    if (!readParents(treeInputPaths[k], n, parent)) {
      return false;
    }
    std::copy(parent.begin(), parent.end(), parents.begin() + static_cast<size_t>(k) * n);
  }
  readStage.stop();
  ScopedStage uploadStage("comb/1_alloc_h2d", true);
  DeviceArray<int> dParents, dParent(n);
  DeviceArray<long long> dDistances(n);
  CombineWorkspace combineWorkspace;
  SospWorkspace sospWorkspace;
  if (!dParents.upload(parents) || !combineWorkspace.reserve(n, K) ||
      !sospWorkspace.reserve(n)) {
    std::cout << "Error: could not allocate the combined graph on the GPU.\n";
    return false;
  }
  uploadStage.stop();
  CombineStats stats;
  {
    ScopedStage gpuStage("comb/steps2_3_gpu", true);
    if (!combinedGraphSospGpu(dParents.data(), n, K, preferences, source, 0, combineWorkspace,
                              sospWorkspace, dDistances.data(), dParent.data(), &stats)) {
      std::cout << "Error: combined graph failed on the GPU.\n";
      return false;
    }
  }
  recordCounter("comb/edges", stats.numberOfEdges);
  recordCounter("comb/iterations", stats.search.iterations);
  recordCounter("comb/pushes", static_cast<double>(stats.search.pushes));
  ScopedStage writeStage("comb/2_d2h_write_text");
  std::vector<long long> distances;
  std::vector<int> parent;
  if (!dDistances.download(distances) || !dParent.download(parent) ||
      !writeDistances(distancesOutputPath, distances) || !writeParents(treeOutputPath, parent)) {
    std::cout << "Error: could not write the combined graph result.\n";
    return false;
  }
  std::cout << "parallelCombinedGraph: " << stats.numberOfEdges << " edges, " << n
            << " vertices; output written to:\n"
            << "  Distances: " << distancesOutputPath << "\n"
            << "  SSSP Tree: " << treeOutputPath << "\n";
  return true;
}

extern "C" int mosp_parallel_sosp_update(const char *csr_prefix, const char *distances_in,
                                         const char *tree_in, const char *insert_path,
                                         const char *delete_path, int32_t objective,
                                         int32_t source, const char *distances_out,
                                         const char *tree_out) {
  if (csr_prefix == nullptr || distances_in == nullptr || tree_in == nullptr ||
      insert_path == nullptr || delete_path == nullptr || distances_out == nullptr ||
      tree_out == nullptr) {
    std::cout << "Error: a path is missing.\n";
    return 0;
  }
  return parallelSOSPUpdate(csr_prefix, distances_in, tree_in, insert_path, delete_path,
                            objective, source, distances_out, tree_out)
             ? 1
             : 0;
}

extern "C" int mosp_parallel_combined_graph(const char *csr_prefix,
                                            const char *const *tree_paths, int32_t k,
                                            int32_t source, const char *distances_out,
                                            const char *tree_out, const int32_t *pref,
                                            int32_t n_pref) {
  if (csr_prefix == nullptr || distances_out == nullptr || tree_out == nullptr ||
      (k > 0 && tree_paths == nullptr) || (n_pref > 0 && pref == nullptr)) {
    std::cout << "Error: a path or the preference vector is missing.\n";
    return 0;
  }
  std::vector<std::string> trees;
  for (int i = 0; i < std::max(k, 0); ++i) {
    if (tree_paths[i] == nullptr) {
      std::cout << "Error: tree path " << i << " is missing.\n";
      return 0;
    }
    trees.emplace_back(tree_paths[i]);
  }
  const std::vector<int> preferences(pref, pref + std::max(n_pref, 0));
  return parallelCombinedGraph(csr_prefix, trees, k, source, "output/combinedGraph",
                               distances_out, tree_out, preferences)
             ? 1
             : 0;
}
