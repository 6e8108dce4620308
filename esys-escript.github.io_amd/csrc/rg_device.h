// Device-side helpers shared by the kernels of libripley_hip (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "rg_internal.h"

namespace rg {

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Streaming kernels remap the hardware block id so that the blocks of
// one XCD cover one contiguous range of rows: neighbouring stencil planes of
// x are then re-read from that XCD's L2 instead of from other XCDs.
__device__ __forceinline__ int64_t xcd_block()
{
    const int64_t b = blockIdx.x, G = gridDim.x;
    if (G % kNumXcd) return b;
    return (b % kNumXcd) * (G / kNumXcd) + b / kNumXcd;
}

// Row range [r0, r1) of this block when n rows are split into gridDim.x
// contiguous chunks (chunk is a multiple of the block size).
struct RowRange {
    int64_t r0, r1;
};
__device__ __forceinline__ RowRange block_rows(int64_t n, int64_t chunk)
{
    const int64_t lb = xcd_block();
    const int64_t r0 = lb * chunk;
    int64_t r1 = r0 + chunk;
    if (r1 > n) r1 = n;
    return {r0, r1};
}

// Sum over the 64 lanes of a wavefront (butterfly; every lane gets the sum).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Sum of NV values over a 256-thread block, in a fixed order (deterministic);
// the result is valid in all threads.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV])
{
    __shared__ double red[NV][kBlock / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double s = wave_sum(v[k]);
        if (lane == 0) red[k][w] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < kBlock / 64; ++j) s += red[k][j];
        v[k] = s;
    }
    __syncthreads();
}

// Per-workgroup partial sums of a reducing kernel: the block sums of NV
// values go to partials[k*gridDim.x + blockIdx.x] with plain stores.  A
// one-workgroup stage kernel launched next on the stream reduces them in a
// fixed order (reduce_partials); the kernel boundary makes the stores
// visible, so no device-scope fence or atomic is needed in the hot kernel.
template <int NV>
__device__ __forceinline__ void publish_block_sums(double (&v)[NV], double* partials)
{
    block_sum<NV>(v);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = v[k];
    }
}

// Sum of the G partials of each of NV rows, by one 256-thread workgroup in a
// fixed order; valid in all threads.
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* partials, int G, double (&tot)[NV])
{
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double* p = partials + (int64_t)k * G;
        double s = 0.0;
        int b = threadIdx.x;
        // batches of independent loads, summed in index order
        for (; b + 15 * kBlock < G; b += 16 * kBlock) {
            double v[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) v[u] = p[b + u * kBlock];
#pragma unroll
            for (int u = 0; u < 16; ++u) s += v[u];
        }
        for (; b < G; b += kBlock) s += p[b];
        tot[k] = s;
    }
    block_sum<NV>(tot);
}

// Sum of the G partials of each of NV rows by one workgroup of NT threads
// (NT a multiple of 64): thread t adds partials t, t+NT, ... in order, then
// the waves' sums are combined in a fixed order; valid in all threads.
// With NT = 1024 a stage reads 8192 partials of every row in a single round
// trip: all those loads are issued before the first add, with indices past G
// clamped (G >= 1) and their values selected away -- a load behind a branch
// would be waited for before the next one is issued.
// after_issue() runs once those loads are in flight (more loads of the
// caller, served in the same round trip).
struct NoOp {
    __device__ void operator()() const {}
};
template <int NV, int NT, class F = NoOp>
__device__ __forceinline__ void reduce_partials_wide(const double* partials, int G, double (&tot)[NV],
                                                     F&& after_issue = F())
{
    __shared__ double red[NV][NT / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double v[NV][8];
#pragma unroll
    for (int k = 0; k < NV; ++k)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int b = threadIdx.x + u * NT;
            v[k][u] = partials[(int64_t)k * G + (b < G ? b : G - 1)];
        }
    after_issue();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        const double* p = partials + (int64_t)k * G;
        double s = 0.0;
#pragma unroll
        for (int u = 0; u < 8; ++u) s += ((int)threadIdx.x + u * NT < G) ? v[k][u] : 0.0;
        for (int b = threadIdx.x + 8 * NT; b < G; b += NT) s += p[b];
        s = wave_sum(s);
        if (lane == 0) red[k][w] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; ++k) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < NT / 64; ++j) s += red[k][j];
        tot[k] = s;
    }
}

// Full-stencil slot s -> (dx, dy, dz) in {-1,0,1}, getDiagonalIndices order.
template <int DIM>
__host__ __device__ __forceinline__ void slot_delta(int s, int& dx, int& dy, int& dz)
{
    dx = s % 3 - 1;
    dy = (s / 3) % 3 - 1;
    dz = (DIM == 3) ? s / 9 - 1 : 0;
}

}  // namespace rg
