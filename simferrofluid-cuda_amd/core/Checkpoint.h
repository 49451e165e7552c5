// The reference's checkpoint file `<frame>.sav` (demo/Driver.cpp:95-100,
// 108-118): Simulation::Save (core/Simulation.cpp:118-125) followed by
// StopWatch::Save (core/StopWatch.cpp:18-26). Little-endian, no header:
//   velocity face arrays, axis 0, 1, 2   FP64, (n + e_a) values each, z-fastest
//   level set                            FP64, n^3 values
//   initial volume, cumulative volume error, volume error   FP64 x 3
//   maximum mesh size                    int32
//   stopwatch count c                    uint64 (size_t)
//   c names, each NUL-terminated
//   c call counts uint32, c mean times FP64, c maximum times FP64
#pragma once
#include "GridData.h"
#include "SGridData.h"

#include <cstdint>
#include <istream>
#include <ostream>
#include <string>
#include <vector>

namespace Pivot {

// The simulation state a checkpoint carries besides the two grids.
struct CheckpointScalars {
    double initVolume = 0;
    double cumulVolError = 0;
    double volError = 0;
    std::int32_t maxMeshSize = 0;
};

// StopWatch's statistics section.
struct StopWatchStats {
    std::vector<std::string> names;
    std::vector<std::uint32_t> counts;
    std::vector<double> avgTime, maxTime;
};

void SaveCheckpoint(std::ostream &out, SGridData<double> const &velocity, GridData<double> const &levelSet,
                    CheckpointScalars const &scalars, StopWatchStats const &stats);

// Reads into velocity and levelSet, which must have the checkpoint's grids
// (the file has no header). Returns false when the stream ends early.
bool LoadCheckpoint(std::istream &in, SGridData<double> &velocity, GridData<double> &levelSet,
                    CheckpointScalars &scalars, StopWatchStats &stats);

}  // namespace Pivot
