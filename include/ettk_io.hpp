/*
 * ettk_io.hpp -- the on-disk formats a caller of the alignment path reads and writes (host only, no device code):
 *
 *   - volume lists: the tab-separated text of nbfWedgedSubImage3D::write / read
 *     (inc/em/nbfWedgedSubImage3D.h:343-600): a header of 33 column names, then one row per sub-volume:
 *     number, lwedge, uwedge, position (3), geometry (3), normal (3), matrix (16), magnification (3), cutOffset,
 *     file name.  Rows are written with the default stream formatting the reference uses (6 significant
 *     digits); rows whose wedge limits are both 0 take their wedge from the "<file>.wedge.mrc" and
 *     "<file>.spherical.wedge.mrc" sidecars when those exist (WSub.h:553-590).
 *   - MRC volumes: nbfMrcWriter::write (inc/io/nbfMrcWriter.h) and nbfMrcReader::read (inc/io/nbfMrcReader.h):
 *     a 1024-byte header (nx, ny, nz, mode, start 0 0 0, grid = nx ny nz, cell 410 410 410, angles 90 90 90,
 *     axes 1 2 3, min / max / mean of the central z section, zeros); data with x fastest and the rows of each
 *     section in reverse y order (the writer's default `reverse`; the reader's vtkImageReader default of a file
 *     origin at the upper left undoes it).  Reading accepts modes 0 (uchar), 1 (short) and 2 (float) and takes
 *     the header size as the file size minus the data size, so extended headers are skipped.
 *   - "matlab" arrays: nbfMatlabWriter::write (inc/io/nbfMatlabWriter.h:100-300): the extents as 32-bit ints,
 *     then the values as 4-byte floats with the first index fastest.
 *   - the classification's outputs built from those: the per-iteration distance array and alignment lists of
 *     nbfLoopClustering::doMra, the state files of nbfLoopClustering::saveState, and averages written negated.
 * Volumes in memory are float32 arrays with x fastest, v[(z * ny + y) * nx + x], the order the C ABI takes.
 * Errors throw std::runtime_error (the reference prints to cerr and returns false).
 */
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace ettk {

struct Volume {
    int nx = 0, ny = 0, nz = 0;
    std::vector<float> data;  // x fastest
    float& at(int x, int y, int z) { return data[((size_t)z * ny + y) * nx + x]; }
    float at(int x, int y, int z) const { return data[((size_t)z * ny + y) * nx + x]; }
};

// ------------------------------------------------------------------------------------------ MRC
inline void writeMrc(const std::string& file, const Volume& v) {
    if ((size_t)v.nx * v.ny * v.nz != v.data.size()) throw std::runtime_error("writeMrc: extents do not match data");
    std::ofstream out(file, std::ios::binary);
    if (!out) throw std::runtime_error("writeMrc: cannot open " + file);
    auto w32 = [&](int32_t x) { out.write(reinterpret_cast<const char*>(&x), 4); };
    auto wf = [&](float x) { out.write(reinterpret_cast<const char*>(&x), 4); };
    w32(v.nx);
    w32(v.ny);
    w32(v.nz);
    w32(2);  // mode 2: float32
    w32(0);
    w32(0);
    w32(0);
    w32(v.nx);
    w32(v.ny);
    w32(v.nz);
    for (int i = 0; i < 3; i++) wf(410.0f);
    for (int i = 0; i < 3; i++) wf(90.0f);
    w32(1);
    w32(2);
    w32(3);
    // min, max, mean of the central section (nbfMrcWriter::header: A(all, all, Nz / 2))
    const int zc = v.nz / 2;
    float mn = v.at(0, 0, zc), mx = mn;
    double sum = 0;
    for (int y = 0; y < v.ny; y++)
        for (int x = 0; x < v.nx; x++) {
            const float a = v.at(x, y, zc);
            mn = std::min(mn, a);
            mx = std::max(mx, a);
            sum += a;
        }
    wf(mn);
    wf(mx);
    wf((float)(sum / ((double)v.nx * v.ny)));
    const std::vector<char> zeros(1024 - 22 * 4, 0);
    out.write(zeros.data(), (std::streamsize)zeros.size());
    for (int z = 0; z < v.nz; z++)
        for (int y = v.ny - 1; y >= 0; y--) out.write(reinterpret_cast<const char*>(&v.data[((size_t)z * v.ny + y) * v.nx]),
                                                      (std::streamsize)v.nx * 4);
    if (!out) throw std::runtime_error("writeMrc: write failed for " + file);
}

inline Volume readMrc(const std::string& file) {
    std::ifstream in(file, std::ios::binary);
    if (!in) throw std::runtime_error("readMrc: cannot open " + file);
    int32_t h[4];
    in.read(reinterpret_cast<char*>(h), sizeof(h));
    if (!in) throw std::runtime_error("readMrc: short header in " + file);
    Volume v;
    v.nx = h[0];
    v.ny = h[1];
    v.nz = h[2];
    const int mode = h[3];
    const int bpv = mode == 0 ? 1 : mode == 1 ? 2 : mode == 2 ? 4 : 0;
    if (bpv == 0 || v.nx <= 0 || v.ny <= 0 || v.nz <= 0) throw std::runtime_error("readMrc: unsupported header in " + file);
    in.seekg(0, std::ios::end);
    const long long size = (long long)in.tellg();
    const long long nvox = (long long)v.nx * v.ny * v.nz;
    const long long header = size - nvox * bpv;
    if (header < 0) throw std::runtime_error("readMrc: file shorter than its data in " + file);
    in.seekg(header, std::ios::beg);
    std::vector<char> raw((size_t)(nvox * bpv));
    in.read(raw.data(), (std::streamsize)raw.size());
    if (!in) throw std::runtime_error("readMrc: short data in " + file);
    v.data.resize((size_t)nvox);
    for (int z = 0; z < v.nz; z++)
        for (int fy = 0; fy < v.ny; fy++) {
            const int y = v.ny - 1 - fy;  // the file's first row is the top row
            for (int x = 0; x < v.nx; x++) {
                const size_t f = ((size_t)z * v.ny + fy) * v.nx + x;
                float a;
                if (mode == 0) a = (float)(unsigned char)raw[f];
                else if (mode == 1) {
                    int16_t s;
                    std::memcpy(&s, &raw[f * 2], 2);
                    a = (float)s;
                } else {
                    std::memcpy(&a, &raw[f * 4], 4);
                }
                v.at(x, y, z) = a;
            }
        }
    return v;
}

inline bool fileExists(const std::string& f) {
    std::ifstream in(f, std::ios::binary);
    return (bool)in;
}

// ------------------------------------------------------------------------------------------ volume lists
struct VolumeRecord {
    float lwedge = -90.0f, uwedge = 90.0f;
    double position[3] = {0, 0, 0};
    double geometry[3] = {0, 0, 0};
    double normal[3] = {0, 0, 0};
    double matrix[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double magnification[3] = {1, 1, 1};
    double cutOffset = 0;
    std::string fileName = "NA";
    // stored wedges read from the sidecars (empty when the limits are used): full grid, x fastest, and the
    // (2B x 2B) spherical image, row-major
    Volume wedge, sphericalWedge;
};

inline const char* const kVolumeListColumns[33] = {
    "number", "lwedge", "uwedge", "posX", "posY", "posZ", "geomX", "geomY", "geomZ", "normalX", "normalY", "normalZ",
    "matrix[0]", "matrix[1]", "matrix[2]", "matrix[3]", "matrix[4]", "matrix[5]", "matrix[6]", "matrix[7]",
    "matrix[8]", "matrix[9]", "matrix[10]", "matrix[11]", "matrix[12]", "matrix[13]", "matrix[14]", "matrix[15]",
    "magnification[0]", "magnification[1]", "magnification[2]", "cutOffset", "filename"};

/* nbfWedgedSubImage3D::write: `precision` 6 reproduces the reference's text exactly; 17 keeps doubles exact. */
inline void writeVolumeList(const std::string& file, const std::vector<VolumeRecord>& vols, int precision = 6) {
    std::ostringstream s;
    s << std::setprecision(precision);
    for (int c = 0; c < 33; c++) s << kVolumeListColumns[c] << (c < 32 ? "\t" : "\n");
    for (size_t i = 0; i < vols.size(); i++) {
        const VolumeRecord& v = vols[i];
        s << i + 1 << "\t" << v.lwedge << "\t" << v.uwedge << "\t";
        for (double x : v.position) s << x << "\t";
        for (double x : v.geometry) s << x << "\t";
        for (double x : v.normal) s << x << "\t";
        for (double x : v.matrix) s << x << "\t";
        for (double x : v.magnification) s << x << "\t";
        s << v.cutOffset << "\t" << v.fileName << "\n";
    }
    std::ofstream out(file);
    if (!out) throw std::runtime_error("writeVolumeList: cannot open " + file);
    out << s.str();
}

/* nbfWedgedSubImage3D::read: rows until the first that does not parse; the sidecar wedges of rows whose limits
   are both 0 are loaded when present. */
inline std::vector<VolumeRecord> readVolumeList(const std::string& file) {
    std::ifstream in(file);
    if (!in) throw std::runtime_error("readVolumeList: cannot open " + file);
    std::stringstream s;
    s << in.rdbuf();
    std::string tok;
    for (int c = 0; c < 33; c++) s >> tok;
    std::vector<VolumeRecord> out;
    while (true) {
        VolumeRecord v;
        int number;
        s >> number >> v.lwedge >> v.uwedge;
        for (double& x : v.position) s >> x;
        for (double& x : v.geometry) s >> x;
        for (double& x : v.normal) s >> x;
        for (double& x : v.matrix) s >> x;
        for (double& x : v.magnification) s >> x;
        s >> v.cutOffset >> v.fileName;
        if (s.fail()) break;
        if (std::abs(v.lwedge) + std::abs(v.uwedge) == 0) {
            if (fileExists(v.fileName + ".wedge.mrc")) v.wedge = readMrc(v.fileName + ".wedge.mrc");
            if (fileExists(v.fileName + ".spherical.wedge.mrc"))
                v.sphericalWedge = readMrc(v.fileName + ".spherical.wedge.mrc");
        }
        out.push_back(std::move(v));
    }
    return out;
}

// ------------------------------------------------------------------------------------------ "matlab" arrays
/* nbfMatlabWriter::write for 1 to 4 dimensions: dims[0] is the fastest index of `values`. */
inline void writeMatlab(const std::string& file, const std::vector<int>& dims, const float* values) {
    if (dims.empty() || dims.size() > 4) throw std::runtime_error("writeMatlab: 1 to 4 dimensions");
    std::ofstream out(file, std::ios::binary);
    if (!out) throw std::runtime_error("writeMatlab: cannot open " + file);
    size_t n = 1;
    for (int d : dims) {
        const int32_t x = d;
        out.write(reinterpret_cast<const char*>(&x), 4);
        n *= (size_t)d;
    }
    out.write(reinterpret_cast<const char*>(values), (std::streamsize)(n * sizeof(float)));
    if (!out) throw std::runtime_error("writeMatlab: write failed for " + file);
}

inline std::vector<float> readMatlab(const std::string& file, int ndims, std::vector<int>& dims) {
    std::ifstream in(file, std::ios::binary);
    if (!in) throw std::runtime_error("readMatlab: cannot open " + file);
    dims.assign(ndims, 0);
    size_t n = 1;
    for (int& d : dims) {
        int32_t x;
        in.read(reinterpret_cast<char*>(&x), 4);
        d = x;
        n *= (size_t)std::max(0, x);
    }
    std::vector<float> v(n);
    in.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(float)));
    if (!in) throw std::runtime_error("readMatlab: short data in " + file);
    return v;
}

// ------------------------------------------------------------------------------------------ classification outputs
/* "_iteration_" and the iteration number padded to three digits, the tag of the classification's per-iteration
   files (MPI_Classification.cpp:673-677, nbfLoopClustering.h:1597-1602 and 1636-1642). */
inline std::string iterationTag(int iteration) {
    std::ostringstream s;
    s << "_iteration_" << (iteration < 10 ? "00" : iteration < 100 ? "0" : "") << iteration;
    return s.str();
}

/* An average written as the classification writes its averages and global averages (MPI_Classification.cpp:
   954-977, nbfLoopClustering::saveState at nbfLoopClustering.h:5610-5622): the values negated, float MRC.  The
   variance map is written as it is (writeMrc). */
inline void writeAverageMrc(const std::string& file, const Volume& average) {
    Volume v = average;
    for (float& x : v.data) x = -x;
    writeMrc(file, v);
}

/* The per-iteration outputs of nbfLoopClustering::doMra (nbfLoopClustering.h:1595-1645), from the distance rows
   `D` of K references x n volumes x `cols` columns, [k][j][c] with c fastest (the rows of getDistances:
   distance, [correlation scale], [wedge overlap], M00..M33):
     "<header><tag>_alignmentToReferences.matlab": D as a (K, n, cols, 1) array, the reference index fastest;
     "<header><tag>_alignments_to_reference_<k>.txt" for every reference k: `volumes` with each cutOffset set to
     |D(k, j, 0)| and each matrix to the 16 values from column 1 (rows of at most 17 columns) or column 2 (longer
     rows), the columns the reference reads; for rows of 19 or 20 columns that is not where M00 sits, and the
     files carry what the reference's would.
   The volumes keep the alignment to the last reference afterwards, as the reference's input list does. */
inline void saveMraIteration(const std::string& header, int iteration, const float* D, int K, int n, int cols,
                             std::vector<VolumeRecord>& volumes) {
    if (K < 1 || n < 1 || cols < 17 || (int)volumes.size() != n)
        throw std::runtime_error("saveMraIteration: D must be K x n x (>= 17) with one row per volume");
    const std::string tag = iterationTag(iteration);
    std::vector<float> a((size_t)K * n * cols);
    for (int k = 0; k < K; k++)
        for (int j = 0; j < n; j++)
            for (int c = 0; c < cols; c++) a[k + (size_t)K * (j + (size_t)n * c)] = D[((size_t)k * n + j) * cols + c];
    writeMatlab(header + tag + "_alignmentToReferences.matlab", {K, n, cols, 1}, a.data());
    const int offset = cols > 17 ? 2 : 1;
    for (int k = 0; k < K; k++) {
        for (int j = 0; j < n; j++) {
            const float* row = D + ((size_t)k * n + j) * cols;
            volumes[j].cutOffset = std::fabs(row[0]);
            for (int e = 0; e < 16; e++) volumes[j].matrix[e] = row[offset + e];
        }
        writeVolumeList(header + tag + "_alignments_to_reference_" + std::to_string(k) + ".txt", volumes);
    }
}

/* nbfLoopClustering::saveState (nbfLoopClustering.h:5542-5637) for classes without symmetry:
     "<file>_volumes.txt": the volumes as given (their current alignments);
     "<file>_classes.txt": the same list with each cutOffset replaced by the volume's class (0 for a volume in no
     class, as the reference leaves it);
     one MRC per average, negated (writeAverageMrc): "<file>_level_<level>_average_<i, three digits>.mrc", or
     "<file>_average_<level>_<referenceIndexes[i]>.mrc" when reference indexes are given;
     "<file>_averages.txt": the list of those files, each at position N/2 with the default wedge and identity.
   `classOf[j]` is the class of volume j or -1.  The reference's binary dump of its average objects
   ("_averages.bin", the serialisation of nbfWedgedAverageImage3D for its own resume) is not written. */
inline void saveState(const std::string& file, const std::vector<VolumeRecord>& volumes, const std::vector<int>& classOf,
                      const std::vector<Volume>& averages, int level, const std::vector<int>& referenceIndexes = {}) {
    if (classOf.size() != volumes.size()) throw std::runtime_error("saveState: one class entry per volume");
    if (!referenceIndexes.empty() && referenceIndexes.size() != averages.size())
        throw std::runtime_error("saveState: one reference index per average");
    writeVolumeList(file + "_volumes.txt", volumes);
    std::vector<VolumeRecord> classes = volumes;
    for (size_t j = 0; j < classes.size(); j++) classes[j].cutOffset = classOf[j] >= 0 ? classOf[j] : 0;
    writeVolumeList(file + "_classes.txt", classes);
    std::vector<VolumeRecord> list;
    for (size_t i = 0; i < averages.size(); i++) {
        std::ostringstream name;
        if (!referenceIndexes.empty()) {
            name << file << "_average_" << level << "_" << referenceIndexes[i] << ".mrc";
        } else {
            name << file << "_level_" << level << "_average_" << (i < 10 ? "00" : i < 100 ? "0" : "") << i << ".mrc";
        }
        writeAverageMrc(name.str(), averages[i]);
        VolumeRecord r;
        r.position[0] = averages[i].nx / 2.0;
        r.position[1] = averages[i].ny / 2.0;
        r.position[2] = averages[i].nz / 2.0;
        r.fileName = name.str();
        list.push_back(r);
    }
    writeVolumeList(file + "_averages.txt", list);
}

}  // namespace ettk
