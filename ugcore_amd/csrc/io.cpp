// File formats for fixtures and interop, host side (no GPU calls):
//  * ConnectionViewer text files (.mat, .vec): ConnectionViewer::WriteMatrix
//    and WriteVector (lib_algebra/common/connection_viewer_output.h:74-141,
//    402-422), which ugcore's SaveMatrixForConnectionViewer /
//    SaveVectorForConnectionViewer (lib_disc/function_spaces/
//    grid_function_util.h) call with the DoF positions. Layout: "1", the
//    dimension, the number of rows, one position per row (default stream
//    precision, 6 significant digits), "1", then one "i j value" line per
//    stored entry. Matrix values keep the default precision; vector values use
//    digits10 + 1 = 16 significant digits; a zero entry is written as "0".
//    (The reference ends each line with std::endl; the bytes are the same.)
//  * The binary vector file of SaveToFile / ReadFromFile in a serial build
//    (bridge/algebra_bridges/restart_bridge.cpp:49-75, 146-158): int32 number
//    of processes (1), int32 offset of the end of the payload, then the payload
//    Serialize(Vector<double>) writes (lib_algebra/cpu_algebra/
//    vector_impl.h:347-361): the size as a size_t, then the raw doubles.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <limits>
#include <string>
#include <vector>

#include "ug_gpu/gpu_algebra.hpp"

namespace ug {
namespace gpu {

namespace {

void write_grid_header(std::ostream& f, int64_t n, const double* pos, int dim) {
    f << 1 << "\n";
    f << dim << "\n";
    f << n << "\n";
    for (int64_t i = 0; i < n; ++i) {
        const double* p = pos + i * dim;
        if (dim == 1)
            f << p[0] << " 0.0\n";
        else if (dim == 2)
            f << p[0] << " " << p[1] << "\n";
        else
            f << p[0] << " " << p[1] << " " << p[2] << "\n";
    }
    f << 1 << "\n";
}

}  // namespace

void cv_write_matrix(const std::string& path, int64_t n, const int32_t* rowptr, const int32_t* col, const double* val,
                     const double* pos, int dim) {
    if (dim < 1 || dim > 3) throw UGError("ConnectionViewer: dimension must be 1, 2 or 3");
    std::ofstream f(path.c_str(), std::ios::out);
    if (!f) throw UGError("ConnectionViewer: cannot open " + path);
    write_grid_header(f, n, pos, dim);
    for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rowptr[i]; k < rowptr[i + 1]; ++k) {
            if (val[k] != 0.0)
                f << i << " " << col[k] << " " << val[k] << "\n";
            else
                f << i << " " << col[k] << " 0\n";
        }
    if (!f) throw UGError("ConnectionViewer: write failed for " + path);
}

void cv_write_vector(const std::string& path, int64_t n, const double* v, const double* pos, int dim) {
    if (dim < 1 || dim > 3) throw UGError("ConnectionViewer: dimension must be 1, 2 or 3");
    std::ofstream f(path.c_str(), std::ios::out);
    if (!f) throw UGError("ConnectionViewer: cannot open " + path);
    write_grid_header(f, n, pos, dim);
    for (int64_t i = 0; i < n; ++i)
        f << i << " " << i << " " << std::setprecision(std::numeric_limits<double>::digits10 + 1) << v[i] << "\n";
    if (!f) throw UGError("ConnectionViewer: write failed for " + path);
}

void vector_file_write(const std::string& path, int64_t n, const double* v) {
    const uint64_t count = (uint64_t)n;
    const int64_t payload = (int64_t)sizeof(count) + n * (int64_t)sizeof(double);
    if (2 * (int64_t)sizeof(int32_t) + payload > std::numeric_limits<int32_t>::max())
        throw UGError("SaveToFile: the serial file format stores a 32-bit offset (vector too large)");
    const int32_t num_procs = 1;
    const int32_t next_offset = (int32_t)(2 * sizeof(int32_t) + payload);
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) throw UGError("SaveToFile: cannot open " + path);
    bool ok = std::fwrite(&num_procs, sizeof(num_procs), 1, f) == 1;
    ok = ok && std::fwrite(&next_offset, sizeof(next_offset), 1, f) == 1;
    ok = ok && std::fwrite(&count, sizeof(count), 1, f) == 1;
    ok = ok && (n == 0 || std::fwrite(v, sizeof(double), (size_t)n, f) == (size_t)n);
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) throw UGError("SaveToFile: write failed for " + path);
}

// reads the header; returns the vector length and leaves the file positioned
// at the first value
static int64_t vector_file_open(const std::string& path, FILE*& f) {
    f = std::fopen(path.c_str(), "rb");
    if (!f) throw UGError("ReadFromFile: cannot open " + path);
    int32_t num_procs = 0, next_offset = 0;
    uint64_t count = 0;
    bool ok = std::fread(&num_procs, sizeof(num_procs), 1, f) == 1;
    ok = ok && std::fread(&next_offset, sizeof(next_offset), 1, f) == 1;
    ok = ok && std::fread(&count, sizeof(count), 1, f) == 1;
    if (!ok) {
        std::fclose(f);
        throw UGError("ReadFromFile: truncated header in " + path);
    }
    if (num_procs != 1) {
        std::fclose(f);
        throw UGError("ReadFromFile: the file holds " + std::to_string(num_procs) +
                      " process parts, a serial run reads exactly 1");
    }
    const int64_t payload = (int64_t)next_offset - 2 * (int64_t)sizeof(int32_t);
    if (payload != (int64_t)sizeof(count) + (int64_t)count * (int64_t)sizeof(double)) {
        std::fclose(f);
        throw UGError("ReadFromFile: payload size does not match the vector length in " + path);
    }
    return (int64_t)count;
}

int64_t vector_file_length(const std::string& path) {
    FILE* f = nullptr;
    const int64_t n = vector_file_open(path, f);
    std::fclose(f);
    return n;
}

void vector_file_read(const std::string& path, int64_t n, double* v) {
    FILE* f = nullptr;
    const int64_t count = vector_file_open(path, f);
    if (count != n) {
        std::fclose(f);
        throw UGError("ReadFromFile: the file holds " + std::to_string(count) + " values, expected " +
                      std::to_string(n));
    }
    const bool ok = n == 0 || std::fread(v, sizeof(double), (size_t)n, f) == (size_t)n;
    std::fclose(f);
    if (!ok) throw UGError("ReadFromFile: truncated values in " + path);
}

void SaveToFile(const GPUVector& v, const std::string& filename) {
    std::vector<double> h((size_t)v.size());
    if (v.size()) v.download(h.data());
    vector_file_write(filename, v.size(), h.data());
}

void ReadFromFile(GPUVector& v, const std::string& filename) {
    const int64_t n = vector_file_length(filename);
    std::vector<double> h((size_t)n);
    vector_file_read(filename, n, h.data());
    v.resize(n);
    if (n) v.upload(h.data());
}

}  // namespace gpu
}  // namespace ug
