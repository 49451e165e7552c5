// CPU-only round trip of the checkpoint format through the drop-in
// containers: `checkpoint_test write FILE` writes a seeded checkpoint of a
// 5x4x3 grid; `checkpoint_test read FILE` reads one and prints its sums.
#include "Checkpoint.h"
#include "StaggeredGrid.h"

#include <cstdio>
#include <cstring>
#include <fstream>

using namespace Pivot;

int main(int argc, char **argv) {
    if (argc != 3) return 2;
    StaggeredGrid const sgrid(0, 0.25, Vector3i(5, 4, 3), Vector3d(0.1, 0.2, 0.3));
    SGridData<double> vel(sgrid.GetFaceGrids());
    GridData<double> phi(sgrid.GetCellGrid());
    CheckpointScalars s;
    StopWatchStats st;
    if (std::strcmp(argv[1], "write") == 0) {
        for (int a = 0; a < 3; a++)
            for (size_t i = 0; i < vel[a].GetData().size(); i++) vel[a][(int)i] = 1000.0 * a + 0.5 * (double)i;
        for (size_t i = 0; i < phi.GetData().size(); i++) phi[(int)i] = -1.0 + 0.125 * (double)i;
        s.initVolume = 0.3125;
        s.cumulVolError = -1.5e-3;
        s.volError = 2.5e-4;
        s.maxMeshSize = 4242;
        st.names = {"Advect", "Project"};
        st.counts = {3, 7};
        st.avgTime = {0.25, 1.5};
        st.maxTime = {0.5, 2.0};
        std::ofstream out(argv[2], std::ios::binary);
        SaveCheckpoint(out, vel, phi, s, st);
        return out ? 0 : 1;
    }
    std::ifstream in(argv[2], std::ios::binary);
    if (!LoadCheckpoint(in, vel, phi, s, st)) return 1;
    double sv = 0, sp = 0;
    for (int a = 0; a < 3; a++)
        for (double v : vel[a].GetData()) sv += v;
    for (double v : phi.GetData()) sp += v;
    std::printf("%.17g %.17g %.17g %.17g %.17g %d %zu", sv, sp, s.initVolume, s.cumulVolError, s.volError,
                s.maxMeshSize, st.names.size());
    for (size_t i = 0; i < st.names.size(); i++)
        std::printf(" %s %u %.17g %.17g", st.names[i].c_str(), st.counts[i], st.avgTime[i], st.maxTime[i]);
    std::printf("\n");
    return 0;
}
