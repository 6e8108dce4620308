// The diagonal-offset stencil row product shared by every SpMV-shaped kernel.
//
// Rows are processed R at a time per thread (R = 16 bytes / value size, so
// one dwordx4 load fetches a slot's values for all R rows), and x is read
// through 16-byte windows: for each (dy, dz) pair the R+2 values
// x[b-1 .. b+R] serve all 3*R products of the three dx slots.  Every load of
// a row group is issued before the first multiply; there is no per-slot
// branch.  Blocks whose rows lie within one stencil reach of either end of
// the vector take a checked, branch-free path (clamped indices + select).
#pragma once

#include "rg_device.h"

namespace rg {

template <int DIM>
struct StencilShape {
    static constexpr int NS = (DIM == 3) ? 27 : 9;   // full stencil slots
    static constexpr int CENTER = NS / 2;            // slot of the main diagonal
};

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

// Between the z-groups of a row group's loads: with ON the loads of one
// z-plane group are not hoisted above the multiplies of the previous one
// (fewer live registers, higher occupancy).  Measured on the 128^3 Brick:
// 2% faster for the full-storage vectorised path; for the symmetric path
// neutral in FP64 and, with FP32 values, 256 -> ~195 VGPRs (1 -> 2 waves per
// SIMD: the FP32-SpMV PCG 0.240 -> 0.170 s), so both paths set it.
template <bool ON = false>
__device__ __forceinline__ void zgroup_fence()
{
    if (ON) __builtin_amdgcn_sched_barrier(0);
}

template <typename T>
__device__ __forceinline__ T ldg(const T* p)
{
    return __builtin_nontemporal_load(p);
}

// Stored values of symmetric storage are read twice: as the upper slots of
// their own row and, transposed, as lower slots of the neighbour rows, so
// they are loaded through the caches normally (a non-temporal hint here made
// the transposed re-reads miss L2: 67.0 vs 52.5 us for the 128^3 SpMV).
// Full storage reads every value once and keeps the streaming hint.
template <bool SYM, typename T>
__device__ __forceinline__ T ld_val(const T* p)
{
    if constexpr (SYM) return *p;
    else return __builtin_nontemporal_load(p);
}

// acc[r] += sum over slots s of block A(i, i+off_s) x[(i+off_s)*BS + :] in
// ascending column order (the order of PASO's CSR row,
// paso/src/SparseMatrix_MatrixVector.cpp:262-320).  Columns outside [0, N)
// contribute nothing, as in the reference stencil SpMV
// (cusplibrary/cusp/detail/host/spmv.h:240-437); with CHECK the indices are
// clamped and the products selected away (branch free), without CHECK the
// caller guarantees all columns are in range.  With SYM only the main and
// upper slots are stored; slot s < CENTER of row i is the transpose of
// stored slot CENTER-s of row i+off_s.
template <int DIM, int BS, bool SYM, bool CHECK, typename VT, typename XT>
__device__ __forceinline__ void stencil_row(const VT* __restrict__ val, int64_t ld, int64_t N,
                                            int64_t s1, int64_t s2, int64_t i,
                                            const XT* __restrict__ x, double (&acc)[BS])
{
    constexpr int NS = StencilShape<DIM>::NS;
    constexpr int C = StencilShape<DIM>::CENTER;
    // slots per group of loads: a z-plane, or one (dy, dz) line of three
    // blocks for symmetric storage with block size > 1 (its transposed
    // loads otherwise need 256 VGPRs, one wave per SIMD; 111 and four waves
    // this way, while full storage is best left with z-plane groups).  x is
    // loaded value by value: 16-byte windows over a line's 3 BS values ran
    // the 96^3 3x3 SpMV at 338 instead of 332 us.
    constexpr bool LINE = SYM && BS > 1;
    constexpr int G = LINE ? 3 : 9;
#pragma unroll
    for (int g0 = 0; g0 < NS; g0 += G) {
        double xv[G][BS];
        double av[G][BS * BS];
        bool ok[G];
#pragma unroll
        for (int q = 0; q < G; ++q) {
            const int s = g0 + q;
            int dx, dy, dz;
            slot_delta<DIM>(s, dx, dy, dz);
            int64_t j = i + dx + dy * s1 + dz * s2;
            ok[q] = true;
            if (CHECK) {
                ok[q] = (uint64_t)j < (uint64_t)N;
                j = ok[q] ? j : i;
            }
#pragma unroll
            for (int c = 0; c < BS; ++c) xv[q][c] = (double)x[j * BS + c];
            if (!SYM || s >= C) {
                const int t = SYM ? s - C : s;
#pragma unroll
                for (int e = 0; e < BS * BS; ++e) av[q][e] = (double)ld_val<SYM>(&val[(int64_t)(t * BS * BS + e) * ld + i]);
            } else {
                const int t = C - s;
#pragma unroll
                for (int r = 0; r < BS; ++r)
#pragma unroll
                    for (int c = 0; c < BS; ++c)   // transposed block of row j
                        av[q][r + BS * c] = (double)val[(int64_t)(t * BS * BS + c + BS * r) * ld + j];
            }
        }
#pragma unroll
        for (int q = 0; q < G; ++q)
#pragma unroll
            for (int c = 0; c < BS; ++c)
#pragma unroll
                for (int r = 0; r < BS; ++r) {
                    const double t = fma(av[q][r + BS * c], xv[q][c], acc[r]);
                    acc[r] = (!CHECK || ok[q]) ? t : acc[r];
                }
        zgroup_fence<LINE>();
    }
}

typedef double rg_dbl2_t __attribute__((ext_vector_type(2)));

// x[b-1+k], k = 0..R+1, from (R+2)/2 16-byte loads at b-1, b+1, ... .  The
// loads are 8-byte aligned: the target reads unaligned vectors (the
// compiler emits one dwordx4 for the 16-byte copy), and this halves the
// cache lines each wave requests against aligned loads with a select
// (x: 3 loads of ~9 lines each per plane -> 2; 128^3 SpMV, symmetric storage
// 53.1 -> 47.0 us, full storage 88.3 -> 85.8 us).  Loading only each lane's
// own aligned pair and taking x[b-1], x[b+R] from the neighbour lanes by
// DPP wave shifts (one more load for the wave edges) requested fewer lines
// still but ran slower (47.4 vs 46.5 us).
template <int R>
__device__ __forceinline__ void x_window(const double* __restrict__ x, int64_t b, double (&w)[R + 2])
{
#pragma unroll
    for (int k = 0; k < (R + 2) / 2; ++k) {
        rg_dbl2_t v;
        __builtin_memcpy(&v, x + b - 1 + 2 * k, sizeof(v));
        w[2 * k] = v[0];
        w[2 * k + 1] = v[1];
    }
}

typedef double rg_dbl2 __attribute__((ext_vector_type(2)));
typedef double rg_dbl4 __attribute__((ext_vector_type(4)));
typedef float rg_flt4 __attribute__((ext_vector_type(4)));
typedef float rg_flt8 __attribute__((ext_vector_type(8)));

// Rows per thread of the vectorised scalar stencil: the bytes of one slot's
// values a thread loads at once (16: one dwordx4 per slot).
#ifndef RG_SPMV_BYTES
#define RG_SPMV_BYTES 16
#endif
template <typename VT>
__host__ __device__ constexpr int spmv_rows_per_thread()
{
    return RG_SPMV_BYTES / (int)sizeof(VT);
}

template <typename VT, int R>
struct VecOf;
template <>
struct VecOf<double, 2> {
    using T = rg_dbl2;
};
template <>
struct VecOf<double, 4> {
    using T = rg_dbl4;
};
template <>
struct VecOf<float, 4> {
    using T = rg_flt4;
};
template <>
struct VecOf<float, 8> {
    using T = rg_flt8;
};

template <typename V>
__device__ __forceinline__ double vget(const V& v, int k)
{
    return (double)v[k];
}

// Scalar (bs = 1) full-storage stencil for R consecutive rows i0..i0+R-1 of
// an interior row group (all columns in range, i0 % R == 0).
template <int DIM, typename VT, int R>
__device__ __forceinline__ void stencil_rows_s1(const VT* __restrict__ val, int64_t ld, int64_t s1, int64_t s2,
                                                int64_t i0, const double* __restrict__ x, double (&acc)[R])
{
    using V = typename VecOf<VT, R>::T;
    constexpr int NZ = (DIM == 3) ? 3 : 1;   // z-plane groups of three (dy) planes
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
#pragma unroll
    for (int zg = 0; zg < NZ; ++zg) {
        V av[3][3];
        double w[3][R + 2];
#pragma unroll
        for (int py = 0; py < 3; ++py) {
            const int pl = py + 3 * zg;
            const int dy = py - 1, dz = (DIM == 3) ? zg - 1 : 0;
            const int64_t b = i0 + dy * s1 + dz * s2;
            x_window<R>(x, b, w[py]);
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int s = dx + 3 * pl;
                av[py][dx] = ldg(reinterpret_cast<const V*>(val + (int64_t)s * ld + i0));
            }
        }
#pragma unroll
        for (int py = 0; py < 3; ++py)
#pragma unroll
            for (int dx = 0; dx < 3; ++dx)
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] = fma(vget(av[py][dx], r), w[py][r + dx], acc[r]);
        zgroup_fence<true>();
    }
}

// R consecutive values v[j0 .. j0+R-1] of one value stream by one
// (unaligned, see x_window) 16-byte load.
template <typename VT, int R>
__device__ __forceinline__ void value_window(const VT* __restrict__ v, int64_t j0, double (&w)[R])
{
    using V = typename VecOf<VT, R>::T;
    V a;
    __builtin_memcpy(&a, v + j0, sizeof(V));
#pragma unroll
    for (int k = 0; k < R; ++k) w[k] = (double)a[k];
}

// Scalar symmetric-storage stencil for R consecutive rows of an interior row
// group: the main and upper slots of row i are stored at row i; slot s below
// the main diagonal is the stored slot CENTER-s of row i+off_s.
template <int DIM, typename VT, int R>
__device__ __forceinline__ void stencil_rows_s1_sym(const VT* __restrict__ val, int64_t ld, int64_t s1, int64_t s2,
                                                    int64_t i0, const double* __restrict__ x, double (&acc)[R])
{
    using V = typename VecOf<VT, R>::T;
    constexpr int NPL = (DIM == 3) ? 9 : 3;
    constexpr int C = StencilShape<DIM>::CENTER;
#pragma unroll
    for (int r = 0; r < R; ++r) acc[r] = 0.0;
#pragma unroll
    for (int pl = 0; pl < NPL; ++pl) {
        const int dy = pl % 3 - 1, dz = (DIM == 3) ? pl / 3 - 1 : 0;
        const int64_t b = i0 + dy * s1 + dz * s2;
        double w[R + 2];
        x_window<R>(x, b, w);
        double av[3][R];
#pragma unroll
        for (int dx = 0; dx < 3; ++dx) {
            const int s = dx + 3 * pl;
            if (s >= C) {
                const V v = ld_val<true>(reinterpret_cast<const V*>(val + (int64_t)(s - C) * ld + i0));
#pragma unroll
                for (int r = 0; r < R; ++r) av[dx][r] = (double)v[r];
            } else {
                value_window<VT, R>(val + (int64_t)(C - s) * ld, b + dx - 1, av[dx]);
            }
        }
#pragma unroll
        for (int dx = 0; dx < 3; ++dx)
#pragma unroll
            for (int r = 0; r < R; ++r) acc[r] = fma(av[dx][r], w[r + dx], acc[r]);
        if (pl % 3 == 2) zgroup_fence<true>();
    }
}

// Rows of one block of a chunked SpMV-shaped kernel.  Calls epi(i, acc) for
// every row i of [r0, r1) (acc: BS values of row i).  For bs = 1 (full or
// symmetric storage, FP64 or FP32 values) interior row groups take the
// vectorised path.
template <int DIM, int BS, bool SYM, typename VT, class Epi>
__device__ __forceinline__ void stencil_block(const VT* __restrict__ val, int64_t ld, int64_t N, int64_t s1,
                                              int64_t s2, int64_t r0, int64_t r1, const double* __restrict__ x,
                                              Epi&& epi)
{
    constexpr int R = spmv_rows_per_thread<VT>();
    const int64_t reach = s2 * (DIM == 3) + s1 + R + 2;
    const bool fast = (BS == 1) && ((reinterpret_cast<uintptr_t>(x) & 7) == 0) &&
                      (r0 % R == 0) && ((r1 - r0) % R == 0) && (r0 >= reach) && (r1 + reach <= N);
    if (fast) {
        for (int64_t i0 = r0 + (int64_t)threadIdx.x * R; i0 < r1; i0 += (int64_t)kBlock * R) {
            double acc[R];
            if (SYM) stencil_rows_s1_sym<DIM, VT, R>(val, ld, s1, s2, i0, x, acc);
            else stencil_rows_s1<DIM, VT, R>(val, ld, s1, s2, i0, x, acc);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                double a1[BS];
                a1[0] = acc[r];
                epi(i0 + r, a1);
            }
        }
    } else {
        const bool interior = (r0 >= reach) && (r1 + reach <= N);
        for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
            double acc[BS];
#pragma unroll
            for (int r = 0; r < BS; ++r) acc[r] = 0.0;
            if (interior) stencil_row<DIM, BS, SYM, false>(val, ld, N, s1, s2, i, x, acc);
            else stencil_row<DIM, BS, SYM, true>(val, ld, N, s1, s2, i, x, acc);
            epi(i, acc);
        }
    }
}

// As stencil_block for scalar rows (bs = 1), handing the vectorised path's
// R consecutive rows to epi_run(i0, acc[R]) at once (the epilogue can then
// move its vector streams R rows per access) and the other rows one by one
// to epi_row(i, acc[1]).
template <int DIM, bool SYM, typename VT, class EpiRun, class EpiRow>
__device__ __forceinline__ void stencil_block_runs(const VT* __restrict__ val, int64_t ld, int64_t N, int64_t s1,
                                                   int64_t s2, int64_t r0, int64_t r1, const double* __restrict__ x,
                                                   EpiRun&& epi_run, EpiRow&& epi_row)
{
    constexpr int R = spmv_rows_per_thread<VT>();
    const int64_t reach = s2 * (DIM == 3) + s1 + R + 2;
    const bool fast = ((reinterpret_cast<uintptr_t>(x) & 7) == 0) && (r0 % R == 0) && ((r1 - r0) % R == 0) &&
                      (r0 >= reach) && (r1 + reach <= N);
    if (fast) {
        for (int64_t i0 = r0 + (int64_t)threadIdx.x * R; i0 < r1; i0 += (int64_t)kBlock * R) {
            double acc[R];
            if (SYM) stencil_rows_s1_sym<DIM, VT, R>(val, ld, s1, s2, i0, x, acc);
            else stencil_rows_s1<DIM, VT, R>(val, ld, s1, s2, i0, x, acc);
            epi_run(i0, acc);
        }
    } else {
        const bool interior = (r0 >= reach) && (r1 + reach <= N);
        for (int64_t i = r0 + threadIdx.x; i < r1; i += kBlock) {
            double acc[1] = {0.0};
            if (interior) stencil_row<DIM, 1, SYM, false>(val, ld, N, s1, s2, i, x, acc);
            else stencil_row<DIM, 1, SYM, true>(val, ld, N, s1, s2, i, x, acc);
            epi_row(i, acc);
        }
    }
}

// R consecutive doubles at p (p 16-byte aligned, R even) in 16-byte accesses
template <int R>
__device__ __forceinline__ void load_run(const double* __restrict__ p, double (&v)[R])
{
    static_assert(R % 2 == 0, "runs of an even number of rows");
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
        const double2 t = *reinterpret_cast<const double2*>(p + 2 * k);
        v[2 * k] = t.x;
        v[2 * k + 1] = t.y;
    }
}
template <int R>
__device__ __forceinline__ void store_run(double* __restrict__ p, const double (&v)[R])
{
#pragma unroll
    for (int k = 0; k < R / 2; ++k) *reinterpret_cast<double2*>(p + 2 * k) = double2{v[2 * k], v[2 * k + 1]};
}

}  // namespace rg
