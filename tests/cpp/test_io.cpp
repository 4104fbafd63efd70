// Drives the on-disk formats of include/ettk_io.hpp (volume lists, MRC, "matlab" arrays) for tests/test_io_cpu.py.
// Usage: test_io <dir>.  Reads dir/vol.raw (float32, extents in dir/dims.txt), dir/in_short.mrc, dir/in_list.txt;
// writes dir/out.mrc, dir/in_short.raw, dir/out_list.txt, dir/out_list17.txt, dir/out_wedge.raw, dir/out.matlab,
// dir/roundtrip.raw.  test_io <dir> state: the classification outputs (state_mode below).
#include <cstdio>
#include <fstream>
#include <iostream>

#include "ettk_io.hpp"

static void write_raw(const std::string& f, const std::vector<float>& v) {
    std::ofstream o(f, std::ios::binary);
    o.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(float)));
}

static std::vector<float> read_raw(const std::string& f, size_t n) {
    std::vector<float> v(n);
    std::ifstream in(f, std::ios::binary);
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(float)));
    if (!in) throw std::runtime_error("short raw file " + f);
    return v;
}

// Mode "state": reads dir/state_in.txt ("K n cols N" and the class of each volume), dir/D.raw (K x n x cols, cols
// fastest), dir/avg.raw (K volumes of N^3) and dir/vols.txt (a volume list); writes the classification outputs under
// dir/run (iteration 7) and dir/run_state (level 3), and dir/run_idx_state with reference indexes 4, 9, ...
static int state_mode(const std::string& dir) {
    std::ifstream in(dir + "/state_in.txt");
    int K, n, cols, N;
    in >> K >> n >> cols >> N;
    std::vector<int> cls(n);
    for (int& c : cls) in >> c;
    const std::vector<float> D = read_raw(dir + "/D.raw", (size_t)K * n * cols);
    const std::vector<float> A = read_raw(dir + "/avg.raw", (size_t)K * N * N * N);
    std::vector<ettk::VolumeRecord> vols = ettk::readVolumeList(dir + "/vols.txt");
    const std::vector<ettk::VolumeRecord> given = vols;
    ettk::saveMraIteration(dir + "/run", 7, D.data(), K, n, cols, vols);
    std::vector<ettk::Volume> avgs(K);
    for (int k = 0; k < K; k++) {
        avgs[k].nx = avgs[k].ny = avgs[k].nz = N;
        avgs[k].data.assign(A.begin() + (size_t)k * N * N * N, A.begin() + (size_t)(k + 1) * N * N * N);
    }
    ettk::saveState(dir + "/run_state", given, cls, avgs, 3);
    std::vector<int> idx;
    for (int k = 0; k < K; k++) idx.push_back(4 + 5 * k);
    ettk::saveState(dir + "/run_idx_state", given, cls, avgs, 3, idx);
    std::printf("state ok: %d references, %d volumes\n", K, n);
    return 0;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    try {
        if (argc > 2 && std::string(argv[2]) == "state") return state_mode(dir);
        ettk::Volume v;
        std::ifstream dims(dir + "/dims.txt");
        dims >> v.nx >> v.ny >> v.nz;
        v.data.resize((size_t)v.nx * v.ny * v.nz);
        std::ifstream raw(dir + "/vol.raw", std::ios::binary);
        raw.read(reinterpret_cast<char*>(v.data.data()), (std::streamsize)(v.data.size() * 4));
        ettk::writeMrc(dir + "/out.mrc", v);
        write_raw(dir + "/roundtrip.raw", ettk::readMrc(dir + "/out.mrc").data);
        write_raw(dir + "/in_short.raw", ettk::readMrc(dir + "/in_short.mrc").data);

        std::vector<ettk::VolumeRecord> list = ettk::readVolumeList(dir + "/in_list.txt");
        ettk::writeVolumeList(dir + "/out_list.txt", list);
        ettk::writeVolumeList(dir + "/out_list17.txt", list, 17);
        std::vector<float> w;
        for (const auto& r : list)
            if (!r.wedge.data.empty()) w.insert(w.end(), r.wedge.data.begin(), r.wedge.data.end());
        write_raw(dir + "/out_wedge.raw", w);

        std::vector<float> a(2 * 3 * 17 * 1);
        for (size_t i = 0; i < a.size(); i++) a[i] = 0.5f * (float)i - 3.0f;
        ettk::writeMatlab(dir + "/out.matlab", {2, 3, 17, 1}, a.data());
        std::vector<int> d;
        std::vector<float> b = ettk::readMatlab(dir + "/out.matlab", 4, d);
        if (b != a || d != std::vector<int>({2, 3, 17, 1})) throw std::runtime_error("matlab round trip");
        std::printf("io ok: %zu records\n", list.size());
    } catch (const std::exception& e) {
        std::fprintf(stderr, "test_io: %s\n", e.what());
        return 1;
    }
    return 0;
}
