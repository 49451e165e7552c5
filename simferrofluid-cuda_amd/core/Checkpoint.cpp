#include "Checkpoint.h"

namespace Pivot {

namespace {
template <class T>
void put(std::ostream &out, T const &v) {
    out.write(reinterpret_cast<char const *>(&v), sizeof(T));
}
template <class T>
void put(std::ostream &out, std::vector<T> const &v) {
    out.write(reinterpret_cast<char const *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}
template <class T>
void get(std::istream &in, T &v) {
    in.read(reinterpret_cast<char *>(&v), sizeof(T));
}
template <class T>
void get(std::istream &in, std::vector<T> &v) {
    in.read(reinterpret_cast<char *>(v.data()), (std::streamsize)(v.size() * sizeof(T)));
}
}  // namespace

void SaveCheckpoint(std::ostream &out, SGridData<double> const &velocity, GridData<double> const &levelSet,
                    CheckpointScalars const &s, StopWatchStats const &st) {
    velocity.Save(out);
    levelSet.Save(out);
    put(out, s.initVolume);
    put(out, s.cumulVolError);
    put(out, s.volError);
    put(out, s.maxMeshSize);
    put(out, static_cast<std::uint64_t>(st.names.size()));
    for (auto const &name : st.names) out << name << std::ends;
    put(out, st.counts);
    put(out, st.avgTime);
    put(out, st.maxTime);
}

bool LoadCheckpoint(std::istream &in, SGridData<double> &velocity, GridData<double> &levelSet,
                    CheckpointScalars &s, StopWatchStats &st) {
    velocity.Load(in);
    levelSet.Load(in);
    get(in, s.initVolume);
    get(in, s.cumulVolError);
    get(in, s.volError);
    get(in, s.maxMeshSize);
    std::uint64_t count = 0;
    get(in, count);
    if (!in) return false;
    st.names.assign(count, std::string());
    for (auto &name : st.names) std::getline(in, name, '\0');
    st.counts.assign(count, 0);
    st.avgTime.assign(count, 0.0);
    st.maxTime.assign(count, 0.0);
    get(in, st.counts);
    get(in, st.avgTime);
    get(in, st.maxTime);
    return static_cast<bool>(in);
}

}  // namespace Pivot
