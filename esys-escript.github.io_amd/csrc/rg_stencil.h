// The diagonal-offset stencil row product shared by every SpMV-shaped kernel.
#pragma once

#include "rg_device.h"

namespace rg {

template <int DIM>
struct StencilShape {
    static constexpr int NS = (DIM == 3) ? 27 : 9;   // full stencil slots
    static constexpr int CENTER = NS / 2;            // slot of the main diagonal
};

// acc[r] += sum over slots s of block A(i, i+off_s) x[(i+off_s)*BS + :]
// in ascending column order (the order of PASO's CSR row,
// paso/src/SparseMatrix_MatrixVector.cpp:262-320).  Columns outside [0, N)
// are skipped, as the reference stencil SpMV does
// (cusplibrary/cusp/detail/host/spmv.h:240-437).  With SYM only the main and
// upper slots are stored; slot s < CENTER of row i is the transpose of the
// stored slot CENTER-s... of row i+off_s.
template <int DIM, int BS, bool SYM, typename VT, typename XT>
__device__ __forceinline__ void stencil_row(const VT* __restrict__ val, int64_t ld, int64_t N,
                                            int64_t s1, int64_t s2, int64_t i,
                                            const XT* __restrict__ x, double (&acc)[BS])
{
    constexpr int NS = StencilShape<DIM>::NS;
    constexpr int C = StencilShape<DIM>::CENTER;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        int dx, dy, dz;
        slot_delta<DIM>(s, dx, dy, dz);
        const int64_t j = i + dx + dy * s1 + dz * s2;
        if ((uint64_t)j < (uint64_t)N) {
            double xv[BS];
#pragma unroll
            for (int c = 0; c < BS; ++c) xv[c] = (double)x[j * BS + c];
            if (!SYM || s >= C) {
                const int t = SYM ? s - C : s;
#pragma unroll
                for (int c = 0; c < BS; ++c)
#pragma unroll
                    for (int r = 0; r < BS; ++r)
                        acc[r] = fma((double)val[(int64_t)(t * BS * BS + r + BS * c) * ld + i], xv[c], acc[r]);
            } else {
                const int t = C - s;   // stored slot of the mirrored offset
#pragma unroll
                for (int c = 0; c < BS; ++c)
#pragma unroll
                    for (int r = 0; r < BS; ++r)
                        acc[r] = fma((double)val[(int64_t)(t * BS * BS + c + BS * r) * ld + j], xv[c], acc[r]);
            }
        }
    }
}

// Number of stored slots.
template <int DIM, bool SYM>
__host__ __device__ constexpr int stored_slots()
{
    return SYM ? StencilShape<DIM>::CENTER + 1 : StencilShape<DIM>::NS;
}

// Whether the neighbour of node (x,y,z) at full slot s lies in the grid.
template <int DIM>
__device__ __forceinline__ bool slot_in_grid(int s, int64_t x, int64_t y, int64_t z, int64_t n0,
                                             int64_t n1, int64_t n2)
{
    int dx, dy, dz;
    slot_delta<DIM>(s, dx, dy, dz);
    const int64_t xx = x + dx, yy = y + dy, zz = z + dz;
    bool ok = (xx >= 0 && xx < n0 && yy >= 0 && yy < n1);
    if (DIM == 3) ok = ok && (zz >= 0 && zz < n2);
    return ok;
}

}  // namespace rg
