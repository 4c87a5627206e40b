// parallelSOSPUpdate.h -- the file-based single-objective update.  Drop-in
// for the reference's headers/parallelSOSPUpdate.cuh:32-39 (same name,
// signature and defaults): reads the original CSR, the objective's old tree
// and the batch files, applies the batch, runs the SOSP update of
// `objectiveIndex` on the GPU and writes the new tree in the tree file format.
#ifndef MOSP_PARALLEL_SOSP_UPDATE_H
#define MOSP_PARALLEL_SOSP_UPDATE_H

#include <string>

bool parallelSOSPUpdate(
    const std::string &originalCsrPrefix, const std::string &distancesInputPath,
    const std::string &treeInputPath, const std::string &insertPath,
    const std::string &deletePath, int objectiveIndex, int source = 0,
    const std::string &distancesOutputPath =
        "output/parallelSospUpdateDistancesTrees/distancesCsr.txt",
    const std::string &treeOutputPath =
        "output/parallelSospUpdateDistancesTrees/SSSPTreeCsr.txt");

#endif  // MOSP_PARALLEL_SOSP_UPDATE_H
