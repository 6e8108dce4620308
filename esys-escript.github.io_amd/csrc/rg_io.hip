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

// One rank's share is the whole grid here (offset 0), so the reference's
// per-rank offsets and "rest" counts are zero.
template <typename T>
void read_grid(const std::string& filename, int dim, const int64_t* grid, int num_comp, int dpp,
               const int64_t* first, const int64_t* num_values, const int32_t* multiplier, const int32_t* reverse,
               bool swap, double* out)
{
    int64_t n[3] = {grid[0], grid[1], dim == 3 ? grid[2] : 1};
    int64_t fst[3] = {first[0], first[1], dim == 3 ? first[2] : 0};
    int64_t nv[3] = {num_values[0], num_values[1], dim == 3 ? num_values[2] : 1};
    int64_t mul[3] = {multiplier[0], multiplier[1], dim == 3 ? multiplier[2] : 1};
    const bool rev_z = dim == 3 && reverse[2] != 0;
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
    const int64_t filesize = f.tellg();
    const int64_t reqsize = nv[0] * nv[1] * nv[2] * num_comp * (int64_t)sizeof(T);
    RG_CHECK(filesize >= reqsize, RG_VALUE_ERROR, "readBinaryGrid(): not enough data in file");

    for (int d = 0; d < dim; ++d)
        if (fst[d] >= n[d] || fst[d] + nv[d] * mul[d] <= 0) return;   // nothing lands in the grid

    int64_t first_w[3], idx[3], num[3];
    for (int d = 0; d < 3; ++d) {
        first_w[d] = std::max<int64_t>(0, fst[d]);       // first sample written
        idx[d] = std::max<int64_t>(0, -fst[d]);          // first value read
        num[d] = std::min(nv[d] - idx[d], n[d] - first_w[d]);
    }
    if (rev_z) idx[2] = nv[2] - idx[2] - 1;
    const int z_step = rev_z ? -1 : 1;

    std::vector<T> values((size_t)std::max<int64_t>(0, num[0]) * num_comp);
    for (int64_t z = 0; z < num[2]; ++z) {
        for (int64_t y = 0; y < num[1]; ++y) {
            const int64_t fileofs = num_comp * (idx[0] + (idx[1] + y) * nv[0] + (idx[2] + z_step * z) * nv[0] * nv[1]);
            f.seekg(fileofs * (int64_t)sizeof(T));
            f.read(reinterpret_cast<char*>(values.data()), num[0] * num_comp * (int64_t)sizeof(T));
            RG_CHECK(!f.fail(), RG_VALUE_ERROR, "readBinaryGrid(): read error in " + filename);
            for (int64_t x = 0; x < num[0]; ++x) {
                // a block of mul0 x mul1 x mul2 samples receives this value
                for (int64_t m2 = 0; m2 < mul[2]; ++m2) {
                    const int64_t dz = first_w[2] + z * mul[2] + m2;
                    if (dz >= n[2]) break;
                    for (int64_t m1 = 0; m1 < mul[1]; ++m1) {
                        const int64_t dy = first_w[1] + y * mul[1] + m1;
                        if (dy >= n[1]) break;
                        for (int64_t m0 = 0; m0 < mul[0]; ++m0) {
                            const int64_t dx = first_w[0] + x * mul[0] + m0;
                            if (dx >= n[0]) break;
                            double* dest = out + (dx + dy * n[0] + dz * n[0] * n[1]) * (int64_t)dpp * num_comp;
                            // as the reference: the components are written one
                            // after the other, each to all data points of the
                            // sample; a NaN component is skipped (the sample
                            // keeps its fill) and the next one moves up
                            for (int c = 0; c < num_comp; ++c) {
                                T raw = values[x * num_comp + c];
                                if (swap) swap_bytes(reinterpret_cast<char*>(&raw), sizeof(T));
                                if (is_nan_value(raw)) continue;
                                const double v = static_cast<double>(raw);
                                for (int q = 0; q < dpp; ++q) *dest++ = v;
                            }
                        }
                    }
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
