// Binary grid I/O of ripley domains at one rank (host code):
// readBinaryGrid (ripley/src/ripleycpp.cpp:63-86 with
// Rectangle::readBinaryGridImpl, Rectangle.cpp:572-690, and
// Brick::readBinaryGridImpl, Brick.cpp:669-818) and writeBinaryGrid
// (Rectangle.cpp / Brick.cpp:964-1050 writeBinaryGridImpl).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <string>
#include <type_traits>
#include <vector>

#include "rg_internal.h"

namespace rg {

namespace {

void swap_bytes(char* p, size_t n)
{
    for (size_t i = 0; i < n / 2; ++i) std::swap(p[i], p[n - 1 - i]);
}

template <typename T>
bool is_nan_value(T v)
{
    if constexpr (std::is_floating_point<T>::value) return std::isnan(v);
    return false;
}

// The file holds nv[0] x nv[1] (x nv[2]) values, x fastest, each value
// num_comp components of type T.  Along axis d the values are placed from
// sample max(first[d], 0) on, starting with value max(-first[d], 0), and each
// value covers multiplier[d] consecutive samples; along z a reversed file
// runs backwards from its last plane.  So every output sample s along an axis
// has at most one source value, and the placement is a per-axis map
//   src_d(s) = base_d + (s - start_d) / mul_d  (or none)
// with start_d = max(first_d, 0), base_d = max(-first_d, 0), valid while
// s >= start_d and (s - start_d) / mul_d < min(nv_d - base_d, n_d - start_d)
// -- the count of values the reference's reader takes along each axis
// (ripley/src/Brick.cpp:733-748 at one rank, offset 0).  The reader below
// walks the output samples, maps each to its source value, and reads each
// file line (one y row of one z plane) once, when the first sample that needs
// it comes up.
struct AxisMap {
    std::vector<int64_t> src;   // file index of the value feeding sample s, or -1
    int64_t lo = -1, hi = -1;   // smallest / largest file index used
};

AxisMap axis_map(int64_t n, int64_t first, int64_t nv, int64_t mul, bool reversed)
{
    AxisMap a;
    a.src.assign((size_t)n, -1);
    const int64_t start = std::max<int64_t>(first, 0);
    const int64_t base = std::max<int64_t>(-first, 0);
    const int64_t count = std::min(nv - base, n - start);
    for (int64_t s = start; s < n; ++s) {
        const int64_t j = (s - start) / mul;
        if (j >= count) break;
        const int64_t v = base + j;
        const int64_t f = reversed ? nv - 1 - v : v;
        a.src[(size_t)s] = f;
        a.lo = a.lo < 0 ? f : std::min(a.lo, f);
        a.hi = std::max(a.hi, f);
    }
    return a;
}

template <typename T>
T decode(const unsigned char* p, bool swap)
{
    unsigned char b[sizeof(T)];
    for (size_t k = 0; k < sizeof(T); ++k) b[k] = swap ? p[sizeof(T) - 1 - k] : p[k];
    T v;
    std::memcpy(&v, b, sizeof(T));
    return v;
}

template <typename T>
void read_grid(const std::string& filename, int dim, const int64_t* grid, int num_comp, int dpp,
               const int64_t* first, const int64_t* num_values, const int32_t* multiplier, const int32_t* reverse,
               bool swap, double* out)
{
    const int64_t n[3] = {grid[0], grid[1], dim == 3 ? grid[2] : 1};
    const int64_t nv[3] = {num_values[0], num_values[1], dim == 3 ? num_values[2] : 1};
    for (int d = 0; d < dim; ++d)
        RG_CHECK(multiplier[d] >= 1, RG_VALUE_ERROR, "readBinaryGrid(): all multipliers must be positive");
    if (dim == 2)
        RG_CHECK(reverse[0] == 0 && reverse[1] == 0, RG_VALUE_ERROR, "readBinaryGrid(): reversing not supported yet");
    else
        RG_CHECK(reverse[0] == 0 && reverse[1] == 0, RG_VALUE_ERROR,
                 "readBinaryGrid(): reversing only supported in Z-direction currently");

    std::ifstream f(filename, std::ifstream::binary);
    RG_CHECK(!f.fail(), RG_VALUE_ERROR, "readBinaryGrid(): cannot open file " + filename);
    f.seekg(0, std::ios::end);
    const int64_t value_bytes = num_comp * (int64_t)sizeof(T);
    RG_CHECK((int64_t)f.tellg() >= nv[0] * nv[1] * nv[2] * value_bytes, RG_VALUE_ERROR,
             "readBinaryGrid(): not enough data in file");

    AxisMap ax[3];
    for (int d = 0; d < 3; ++d) {
        const bool active = d < dim;
        ax[d] = axis_map(n[d], active ? first[d] : 0, nv[d], active ? multiplier[d] : 1,
                         active && d == 2 && reverse[2] != 0);
    }
    if (ax[0].lo < 0 || ax[1].lo < 0 || ax[2].lo < 0) return;   // no sample receives a value

    // one file line: the x range [lo, hi] of row (fy, fz)
    const int64_t xlo = ax[0].lo, width = ax[0].hi - ax[0].lo + 1;
    std::vector<unsigned char> line((size_t)(width * value_bytes));
    int64_t line_y = -1, line_z = -1;
    const int64_t sample_size = (int64_t)dpp * num_comp;
    for (int64_t z = 0; z < n[2]; ++z) {
        const int64_t fz = ax[2].src[(size_t)z];
        if (fz < 0) continue;
        for (int64_t y = 0; y < n[1]; ++y) {
            const int64_t fy = ax[1].src[(size_t)y];
            if (fy < 0) continue;
            if (fy != line_y || fz != line_z) {
                f.seekg(((fz * nv[1] + fy) * nv[0] + xlo) * value_bytes);
                f.read(reinterpret_cast<char*>(line.data()), width * value_bytes);
                RG_CHECK(!f.fail(), RG_VALUE_ERROR, "readBinaryGrid(): read error in " + filename);
                line_y = fy;
                line_z = fz;
            }
            double* row = out + (z * n[1] + y) * n[0] * sample_size;
            for (int64_t x = 0; x < n[0]; ++x) {
                const int64_t fx = ax[0].src[(size_t)x];
                if (fx < 0) continue;
                // the sample's components in order, each copied to all of
                // its data points; a NaN component is dropped, so the
                // following components move up and the sample's last slots
                // keep what they held (the reference reader's behaviour)
                double* dest = row + x * sample_size;
                const unsigned char* src = line.data() + (fx - xlo) * value_bytes;
                for (int c = 0; c < num_comp; ++c) {
                    const T v = decode<T>(src + c * sizeof(T), swap);
                    if (is_nan_value(v)) continue;
                    for (int q = 0; q < dpp; ++q) *dest++ = static_cast<double>(v);
                }
            }
        }
    }
}

template <typename T>
void write_grid(const std::string& filename, int dim, const int64_t* grid, const double* in, bool swap)
{
    const int64_t total = grid[0] * grid[1] * (dim == 3 ? grid[2] : 1);
    std::vector<T> buf((size_t)total);
    for (int64_t i = 0; i < total; ++i) {
        T v = static_cast<T>(in[i]);
        if (swap) swap_bytes(reinterpret_cast<char*>(&v), sizeof(T));
        buf[i] = v;
    }
    std::ofstream f(filename, std::ofstream::binary | std::ofstream::trunc);
    RG_CHECK(!f.fail(), RG_VALUE_ERROR, "writeBinaryGrid(): cannot open file " + filename);
    f.write(reinterpret_cast<const char*>(buf.data()), total * (int64_t)sizeof(T));
    RG_CHECK(!f.fail(), RG_VALUE_ERROR, "writeBinaryGrid(): write error in " + filename);
}

bool needs_swap(int byte_order)
{
    RG_CHECK(byte_order == RG_BYTEORDER_LITTLE_ENDIAN || byte_order == RG_BYTEORDER_BIG_ENDIAN, RG_VALUE_ERROR,
             "binary grid: unknown byte order");
    // this library runs on little-endian hosts only
    return byte_order != RG_BYTEORDER_NATIVE;
}

}  // namespace

void read_binary_grid(const char* filename, int dim, const int64_t* grid, int num_comp, int dpp,
                      const int64_t* first, const int64_t* num_values, const int32_t* multiplier,
                      const int32_t* reverse, int byte_order, int data_type, double* out)
{
    RG_CHECK(filename && grid && first && num_values && multiplier && reverse && out, RG_VALUE_ERROR,
             "readBinaryGrid(): NULL argument");
    RG_CHECK(dim == 2 || dim == 3, RG_VALUE_ERROR, "readBinaryGrid(): dim must be 2 or 3");
    RG_CHECK(num_comp >= 1 && dpp >= 1, RG_VALUE_ERROR, "readBinaryGrid(): bad data point shape");
    const bool swap = needs_swap(byte_order);
    if (data_type == RG_DATATYPE_INT32)
        read_grid<int32_t>(filename, dim, grid, num_comp, dpp, first, num_values, multiplier, reverse, swap, out);
    else if (data_type == RG_DATATYPE_FLOAT32)
        read_grid<float>(filename, dim, grid, num_comp, dpp, first, num_values, multiplier, reverse, swap, out);
    else if (data_type == RG_DATATYPE_FLOAT64)
        read_grid<double>(filename, dim, grid, num_comp, dpp, first, num_values, multiplier, reverse, swap, out);
    else
        throw Error(RG_VALUE_ERROR, "readBinaryGrid(): invalid or unsupported datatype");
}

void write_binary_grid(const char* filename, int dim, const int64_t* grid, const double* in, int byte_order,
                       int data_type)
{
    RG_CHECK(filename && grid && in, RG_VALUE_ERROR, "writeBinaryGrid(): NULL argument");
    RG_CHECK(dim == 2 || dim == 3, RG_VALUE_ERROR, "writeBinaryGrid(): dim must be 2 or 3");
    const bool swap = needs_swap(byte_order);
    if (data_type == RG_DATATYPE_INT32)
        write_grid<int32_t>(filename, dim, grid, in, swap);
    else if (data_type == RG_DATATYPE_FLOAT32)
        write_grid<float>(filename, dim, grid, in, swap);
    else if (data_type == RG_DATATYPE_FLOAT64)
        write_grid<double>(filename, dim, grid, in, swap);
    else
        throw Error(RG_VALUE_ERROR, "writeBinaryGrid(): invalid or unsupported data type");
}

}  // namespace rg
