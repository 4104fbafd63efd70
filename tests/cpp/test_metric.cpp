// Drives the C++ drop-in (include/ettk_metric.hpp) the way MPI_Classification.cpp and nbfLoopClustering.h
// drive nbfCorrelationImageMetric: getDistances with resume entries, getImage, multi-reference assignment, the
// symmetric sub-volume matrix.
// Usage: test_metric <dir>; reads dir/in.txt and the raw float32 volumes, writes dir/out_*.txt / .bin.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <vector>

#include "ettk_metric.hpp"

static std::vector<float> read_raw(const std::string& f, size_t n) {
    std::vector<float> v(n);
    std::ifstream in(f, std::ios::binary);
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(float)));
    if (!in) throw std::runtime_error("cannot read " + f);
    return v;
}

int main(int argc, char** argv) {
    if (argc < 2) return 2;
    const std::string dir = argv[1];
    try {
        std::ifstream meta(dir + "/in.txt");
        int N, B, n, R;
        meta >> N >> B >> n >> R;
        const size_t n3 = (size_t)N * N * N;
        std::vector<float> h_refs = read_raw(dir + "/refs.bin", R * n3), h_parts = read_raw(dir + "/parts.bin", n * n3);
        float *d_refs = nullptr, *d_parts = nullptr, *d_avg = nullptr;
        ettk::check_hip(hipMalloc(&d_refs, R * n3 * sizeof(float)), "hipMalloc");
        ettk::check_hip(hipMalloc(&d_parts, n * n3 * sizeof(float)), "hipMalloc");
        ettk::check_hip(hipMalloc(&d_avg, n3 * sizeof(float)), "hipMalloc");
        ettk::check_hip(hipMemcpy(d_refs, h_refs.data(), R * n3 * sizeof(float), hipMemcpyHostToDevice), "copy");
        ettk::check_hip(hipMemcpy(d_parts, h_parts.data(), n * n3 * sizeof(float), hipMemcpyHostToDevice), "copy");

        std::vector<ettk::WedgedSubImage3D> vols(n);
        for (int i = 0; i < n; i++) {
            float lw, uw;
            meta >> lw >> uw;
            vols[i].d_volume = d_parts + i * n3;
            vols[i].setWedge(lw, uw);
        }
        std::vector<ettk::WedgedAverageImage3D> refs(R);
        for (int r = 0; r < R; r++) refs[r].d_average = d_refs + r * n3;

        ettk::CorrelationImageMetric metric(N, B);
        metric.bandPassOn(0.05f, 0.5f, 0.01f, 0.01f);
        metric.maskOn(N / 2.0f - 5.0f, 2.0f);
        metric.setSquareMaskSize(5);
        metric.setToUseMutualCorrelation(true);
        metric.setToComputeOverlapNormalizedDistances(true);
        metric.setMissingWedgeCompensation(true);
        metric.setTranslationSearchRestriction(8.0f);
        metric.setExecutionForm(1);

        // getDistances type 0 with one resumed entry
        ettk::DistanceArray D(R, n, 17, 1);
        D(0, 1, 0, 0) = 5.0f;
        metric.getDistances(refs, vols, D, 0);
        {
            std::ofstream o(dir + "/out_D0.txt");
            o.precision(9);
            for (int r = 0; r < R; r++)
                for (int i = 0; i < n; i++) {
                    for (int c = 0; c < 17; c++) o << D(r, i, c, 0) << (c < 16 ? " " : "\n");
                }
        }
        // the 20-column row of type 2: distance, correlation scale, wedge overlap, M16, one spare column
        {
            ettk::DistanceArray D20(R, n, 20, 1);
            metric.getDistances(refs, vols, D20, 2);
            std::ofstream o(dir + "/out_D20.txt");
            o.precision(9);
            for (int r = 0; r < R; r++)
                for (int i = 0; i < n; i++)
                    for (int c = 0; c < 20; c++) o << D20(r, i, c, 0) << (c < 19 ? " " : "\n");
        }
        // average of the volumes under their type-0 alignments to reference 0, unit weights
        refs[0].setVolumes(&vols);
        std::vector<float> al((size_t)n * 17, 0.0f);
        for (int i = 0; i < n; i++) {
            al[(size_t)i * 17] = 1.0f;
            for (int c = 0; c < 16; c++) al[(size_t)i * 17 + 1 + c] = D(0, i, 1 + c, 0);
        }
        al[(size_t)1 * 17] = 0.0f;  // the resumed entry holds no matrix: weight 0
        refs[0].setAlignments(al, 1);
        std::vector<float> keep(n3);
        ettk::check_hip(hipMemcpy(keep.data(), d_refs, n3 * sizeof(float), hipMemcpyDeviceToHost), "copy");
        refs[0].d_average = d_avg;
        metric.getImage(refs[0]);
        std::vector<float> avg(n3);
        ettk::check_hip(hipMemcpy(avg.data(), d_avg, n3 * sizeof(float), hipMemcpyDeviceToHost), "copy");
        {
            std::ofstream o(dir + "/out_avg.bin", std::ios::binary);
            o.write(reinterpret_cast<const char*>(avg.data()), (std::streamsize)(n3 * sizeof(float)));
        }
        refs[0].d_average = d_refs;
        // multi-reference assignment on the type-0 alignments (all entries recomputed)
        ettk::DistanceArray A(R, n, 17, 1);
        metric.getDistances(refs, vols, A, 0);
        std::vector<int> best = metric.mraAssignment(refs, vols, A);
        {
            std::ofstream o(dir + "/out_mra.txt");
            o.precision(9);
            for (int i = 0; i < n; i++) o << best[i] << "\n";
            for (int r = 0; r < R; r++)
                for (int i = 0; i < n; i++)
                    for (int c = 0; c < 17; c++) o << A(r, i, c, 0) << (c < 16 ? " " : "\n");
        }
        // the symmetric sub-volume matrix of the search without refinement (type 3), 19-column rows, one entry
        // resumed; makeDistanceMatrix completes the lower triangle and the diagonal
        {
            ettk::DistanceArray S(n, n, 19, 1);
            S(0, 2, 0, 0) = 7.0f;
            metric.getDistances(vols, S, 3);
            std::ofstream o(dir + "/out_sym.txt");
            o.precision(9);
            for (int i = 0; i < n; i++)
                for (int j = 0; j < n; j++)
                    for (int c = 0; c < 19; c++) o << S(i, j, c, 0) << (c < 18 ? " " : "\n");
        }
        (void)hipFree(d_refs);
        (void)hipFree(d_parts);
        (void)hipFree(d_avg);
    } catch (const ettk::MetricError& e) {
        std::cerr << "MetricError " << e.code << ": " << e.what() << std::endl;
        return 1;
    } catch (const std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 1;
    }
    std::cout << "ok" << std::endl;
    return 0;
}
