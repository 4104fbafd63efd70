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
 * Volumes in memory are float32 arrays with x fastest, v[(z * ny + y) * nx + x], the order the C ABI takes.
 * Errors throw std::runtime_error (the reference prints to cerr and returns false).
 */
#pragma once

#include <algorithm>
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

}  // namespace ettk
