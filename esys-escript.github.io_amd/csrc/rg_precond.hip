// PASO's multi-colour preconditioners on the ripley stencil.
//
//   * colouring: Pattern::borrowColoringPointer (paso/src/Pattern.cpp:
//     166-195) -- colour c is the maximal independent set that Pattern::mis
//     (paso/src/Pattern_mis.cpp:44-110) finds among the vertices left
//     uncoloured, with the fixed priorities fmod(0.4142135623730951 (i+1), 1);
//     computed here by one select and one mark kernel per MIS round, the
//     neighbours coming from the stencil geometry;
//   * Gauss-Seidel: Preconditioner_LocalSmoother_Sweep_colored
//     (paso/src/Smoother.cpp:331-507) with Preconditioner_Smoother_solve's
//     sweeps (Smoother.cpp:100-127);
//   * ILU0: Solver_getILU / Solver_solveILU (paso/src/ILU.cpp:43-462).
//
// Rows are stored in colour order (pos -> row = perm[pos]) in structure-of-
// arrays copies of the full stencil split by colour: plow holds the
// couplings of a row to neighbours of lower colours, pup those to higher
// colours (each zero elsewhere, also outside the grid), pdiag / pdinv the
// main block and its inverse, e.g. plow[(t*bs*bs + e)*ld + pos].  A
// substitution kernel of colour c then streams the contiguous range of its
// rows and needs no colour test: forward is b - (plow x), backward
// pdiag x - (pup x), both branch-free stencil products with the neighbour
// values x gathered in node order.  Rows of one colour are never neighbours,
// so a colour's rows read only values finished by earlier launches.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rg_dispatch.h"
#include "rg_stencil.h"

namespace rg {

namespace {

constexpr int kMisAvailable = -1, kMisInMis = -3, kMisConnected = -4;
constexpr double kMisSeed = .4142135623730951;

struct Grid {
    int n0, n1, n2;
    int64_t s1, s2, N;
};

Grid grid_of(const Geom& g)
{
    return {(int)g.nn[0], (int)g.nn[1], (int)g.nn[2], g.s1, g.s2, g.N};
}

__device__ __forceinline__ void node_coords(const Grid& g, int64_t i, int& x, int& y, int& z)
{
    const uint32_t u = (uint32_t)i;
    const uint32_t q = u / (uint32_t)g.n0;
    x = (int)(u - q * (uint32_t)g.n0);
    const uint32_t q2 = q / (uint32_t)g.n1;
    y = (int)(q - q2 * (uint32_t)g.n1);
    z = (int)q2;
}

// neighbour of node (x,y,z) = i at full slot s; false when outside the grid
template <int DIM>
__device__ __forceinline__ bool neighbour(const Grid& g, int s, int x, int y, int z, int64_t i, int64_t& k)
{
    int dx, dy, dz;
    slot_delta<DIM>(s, dx, dy, dz);
    const int xx = x + dx, yy = y + dy, zz = z + dz;
    bool ok = xx >= 0 && xx < g.n0 && yy >= 0 && yy < g.n1;
    if (DIM == 3) ok = ok && zz >= 0 && zz < g.n2;
    k = i + dx + dy * g.s1 + dz * g.s2;
    return ok;
}

__device__ __forceinline__ double mis_value(int64_t i)
{
    return fmod(kMisSeed * (double)(i + 1), 1.0);
}

// ------------------------------------------------------------ colouring
// One MIS round, select step: an available vertex joins when its value is
// below every neighbour's (Pattern_mis.cpp:71-85).  The values of all
// vertices are taken from the marker state at launch (sel is a separate
// array), as the reference computes them in a loop of their own.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_mis_select(Grid g, const int32_t* __restrict__ marker,
                                                       uint8_t* __restrict__ sel, int* __restrict__ any)
{
    constexpr int NS = StencilShape<DIM>::NS, C = StencilShape<DIM>::CENTER;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.N) return;
    if (marker[i] != kMisAvailable) {
        sel[i] = 0;
        return;
    }
    any[0] = 1;
    const double vi = mis_value(i);
    int x, y, z;
    node_coords(g, i, x, y, z);
    bool in = true;
    for (int s = 0; s < NS && in; ++s) {
        if (s == C) continue;
        int64_t k;
        if (!neighbour<DIM>(g, s, x, y, z, i, k)) continue;
        const double vk = (marker[k] == kMisAvailable) ? mis_value(k) : 2.0;
        in = !(vk <= vi);
    }
    sel[i] = in ? 1 : 0;
}

// mark step (Pattern_mis.cpp:88-100): the selected vertices enter the set and
// their neighbours are connected.  Selected vertices are never neighbours,
// so no store here races with another value for the same vertex.
template <int DIM>
__global__ __launch_bounds__(kBlock) void k_mis_mark(Grid g, int32_t* __restrict__ marker,
                                                     const uint8_t* __restrict__ sel)
{
    constexpr int NS = StencilShape<DIM>::NS, C = StencilShape<DIM>::CENTER;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= g.N || !sel[i]) return;
    int x, y, z;
    node_coords(g, i, x, y, z);
    for (int s = 0; s < NS; ++s) {
        if (s == C) continue;
        int64_t k;
        if (neighbour<DIM>(g, s, x, y, z, i, k)) marker[k] = kMisConnected;
    }
    marker[i] = kMisInMis;
}

// colour assignment after one MIS (Pattern.cpp:181-190): members get colour
// `out`; the marker becomes the colour, so uncoloured vertices are available
// again for the next MIS.
__global__ __launch_bounds__(kBlock) void k_colour_assign(int64_t n, int32_t* __restrict__ marker,
                                                          int32_t* __restrict__ colour, int out,
                                                          int* __restrict__ left)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    int32_t c = colour[i];
    if (marker[i] == kMisInMis) {
        c = out;
        colour[i] = out;
    }
    marker[i] = c;
    if (c < 0) left[0] = 1;
}

// ------------------------------------------------------ colour-ordered copies
// Block A(row, row + off_t) of full slot t: from the node-ordered stencil
// values (NODE, with the lower slots of symmetric storage as the transposed
// stored blocks of the neighbour) or from the colour-ordered full copy.
template <int DIM, int BS, bool SYM, bool NODE>
__device__ __forceinline__ void fetch_block(const double* src, int64_t ld, int t, int64_t row, int64_t k,
                                            int64_t pos, double (&v)[BS * BS])
{
    constexpr int C = StencilShape<DIM>::CENTER, B2 = BS * BS;
    if (!NODE) {
#pragma unroll
        for (int e = 0; e < B2; ++e) v[e] = src[(int64_t)(t * B2 + e) * ld + pos];
    } else if (!SYM || t >= C) {
        const int ts = SYM ? t - C : t;
#pragma unroll
        for (int e = 0; e < B2; ++e) v[e] = src[(int64_t)(ts * B2 + e) * ld + row];
    } else {
        const int ts = C - t;
#pragma unroll
        for (int r = 0; r < BS; ++r)
#pragma unroll
            for (int c = 0; c < BS; ++c) v[r + BS * c] = src[(int64_t)(ts * B2 + c + BS * r) * ld + k];
    }
}

// Colour-ordered copies of row pos: plow / pup get the blocks of the
// neighbours of lower / higher colour (zero otherwise and outside the grid),
// diag the main block; with NODE the source is the balanced node-ordered
// matrix (Gauss-Seidel), else the colour-ordered ILU0 factors, whose main
// slot holds the inverted pivot.
template <int DIM, int BS, bool SYM, bool NODE>
__global__ __launch_bounds__(kBlock) void k_split(const double* __restrict__ src, int64_t ld, Grid g,
                                                  const int32_t* __restrict__ perm,
                                                  const uint8_t* __restrict__ colour, double* __restrict__ plow,
                                                  double* __restrict__ pup, double* __restrict__ diag)
{
    constexpr int NS = StencilShape<DIM>::NS, C = StencilShape<DIM>::CENTER, B2 = BS * BS;
    const int64_t pos = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (pos >= g.N) return;
    const int64_t row = perm[pos];
    const int ci = colour[row];
    int x, y, z;
    node_coords(g, row, x, y, z);
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        int64_t k;
        const bool ok = neighbour<DIM>(g, t, x, y, z, row, k);
        double v[B2];
#pragma unroll
        for (int e = 0; e < B2; ++e) v[e] = 0.0;
        if (ok) fetch_block<DIM, BS, SYM, NODE>(src, ld, t, row, k, pos, v);
        if (t == C) {
#pragma unroll
            for (int e = 0; e < B2; ++e) diag[(int64_t)e * ld + pos] = v[e];
        }
        const int ck = ok ? (int)colour[k] : ci;
        const bool lo = ok && t != C && ck < ci, up = ok && t != C && ck > ci;
#pragma unroll
        for (int e = 0; e < B2; ++e) {
            plow[(int64_t)(t * B2 + e) * ld + pos] = lo ? v[e] : 0.0;
            pup[(int64_t)(t * B2 + e) * ld + pos] = up ? v[e] : 0.0;
        }
    }
}

// full colour-ordered copy (the ILU0 elimination works in place on it)
template <int DIM, int BS, bool SYM>
__global__ __launch_bounds__(kBlock) void k_permute_values(const double* __restrict__ val, int64_t ld, Grid g,
                                                           const int32_t* __restrict__ perm,
                                                           double* __restrict__ pval)
{
    constexpr int NS = StencilShape<DIM>::NS, B2 = BS * BS;
    const int64_t pos = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (pos >= g.N) return;
    const int64_t row = perm[pos];
    int x, y, z;
    node_coords(g, row, x, y, z);
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        int64_t k;
        const bool ok = neighbour<DIM>(g, t, x, y, z, row, k);
        double v[B2];
#pragma unroll
        for (int e = 0; e < B2; ++e) v[e] = 0.0;
        if (ok) fetch_block<DIM, BS, SYM, true>(val, ld, t, row, k, pos, v);
#pragma unroll
        for (int e = 0; e < B2; ++e) pval[(int64_t)(t * B2 + e) * ld + pos] = v[e];
    }
}

template <int BS>
__global__ __launch_bounds__(kBlock) void k_permute_diag(const double* __restrict__ dinv, int64_t ld, int64_t n,
                                                         const int32_t* __restrict__ perm,
                                                         double* __restrict__ pdinv)
{
    const int64_t pos = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (pos >= n) return;
    const int64_t row = perm[pos];
#pragma unroll
    for (int e = 0; e < BS * BS; ++e) pdinv[(int64_t)e * ld + pos] = dinv[(int64_t)e * ld + row];
}

// ------------------------------------------------------------ block algebra
// column-major bs x bs blocks stored with stride ld: M(r, c) = M[(r + bs c) ld]

// acc -= M v (BlockOps_SMV_2/3, paso/src/BlockOps.h:46-85)
template <int BS>
__device__ __forceinline__ void block_smv(double (&acc)[BS], const double* M, int64_t ld, const double (&v)[BS])
{
#pragma unroll
    for (int r = 0; r < BS; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < BS; ++c) s = fma(M[(int64_t)(r + BS * c) * ld], v[c], s);
        acc[r] -= s;
    }
}

// out = M v (BlockOps_MViP_2/3, BlockOps.h:187-217)
template <int BS>
__device__ __forceinline__ void block_mv(const double* M, int64_t ld, const double (&v)[BS], double (&out)[BS])
{
#pragma unroll
    for (int r = 0; r < BS; ++r) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < BS; ++c) s = fma(M[(int64_t)(r + BS * c) * ld], v[c], s);
        out[r] = s;
    }
}

// closed-form inverse of BlockOps_invM_2/3 (BlockOps.h:100-145), as
// Solver_getILU uses it; false for a zero determinant
template <int BS>
__device__ __forceinline__ bool block_inverse(const double (&A)[BS * BS], double (&I)[BS * BS])
{
    if (BS == 1) {
        I[0] = 1.0 / A[0];
        return fabs(A[0]) > 0.0;
    } else if (BS == 2) {
        const double A11 = A[0], A12 = A[2], A21 = A[1], A22 = A[3];
        double D = A11 * A22 - A12 * A21;
        const bool ok = fabs(D) > 0.0;
        D = 1.0 / D;
        I[0] = A22 * D; I[1] = -A21 * D; I[2] = -A12 * D; I[3] = A11 * D;
        return ok;
    } else {
        const double A11 = A[0], A21 = A[1], A31 = A[2], A12 = A[3], A22 = A[4], A32 = A[5], A13 = A[6],
                     A23 = A[7], A33 = A[8];
        double D = A11 * (A22 * A33 - A23 * A32) + A12 * (A31 * A23 - A21 * A33) + A13 * (A21 * A32 - A31 * A22);
        const bool ok = fabs(D) > 0.0;
        D = 1.0 / D;
        I[0] = (A22 * A33 - A23 * A32) * D;
        I[1] = (A31 * A23 - A21 * A33) * D;
        I[2] = (A21 * A32 - A31 * A22) * D;
        I[3] = (A13 * A32 - A12 * A33) * D;
        I[4] = (A11 * A33 - A31 * A13) * D;
        I[5] = (A12 * A31 - A11 * A32) * D;
        I[6] = (A12 * A23 - A13 * A22) * D;
        I[7] = (A13 * A21 - A11 * A23) * D;
        I[8] = (A11 * A22 - A12 * A21) * D;
        return ok;
    }
}

struct ColourArgs {
    Grid g;
    const int32_t* perm;
    const int32_t* ipos;
    const uint8_t* colour;
    double* pval;
    int64_t ld;
};

// ------------------------------------------------------------ sweeps
struct SweepArgs {
    const int32_t* perm;
    const double* plow;
    const double* pup;
    const double* pdiag;
    const double* pdinv;
    int64_t ld, N, s1, s2;
    const KrylovState* guard;
};

// acc -= sum over the stencil slots t != C of M(t) x[row + off_t], for the
// split copies: entries outside the grid or of the other colour side are
// zero, so only indices outside [0, N) need care (clamped and selected away).
template <int DIM, int BS>
__device__ __forceinline__ void split_product(const SweepArgs& a, const double* M, int64_t pos, int64_t row,
                                              const double* __restrict__ x, double (&acc)[BS])
{
    constexpr int NS = StencilShape<DIM>::NS, C = StencilShape<DIM>::CENTER, B2 = BS * BS;
#pragma unroll
    for (int t = 0; t < NS; ++t) {
        if (t == C) continue;
        int dx, dy, dz;
        slot_delta<DIM>(t, dx, dy, dz);
        int64_t k = row + dx + dy * a.s1 + dz * a.s2;
        const bool in = (uint64_t)k < (uint64_t)a.N;
        k = in ? k : row;
        double xk[BS], mv[B2];
#pragma unroll
        for (int c = 0; c < BS; ++c) xk[c] = x[k * BS + c];
#pragma unroll
        for (int e = 0; e < B2; ++e) mv[e] = M[(int64_t)(t * B2 + e) * a.ld + pos];
#pragma unroll
        for (int r = 0; r < BS; ++r) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < BS; ++c) s = fma(mv[r + BS * c], xk[c], s);
            acc[r] = in ? acc[r] - s : acc[r];
        }
    }
}

// Forward substitution for the rows of colour positions [p0, p1):
//   x_i = Dinv_i (b_i - sum_{c(k) < c(i)} A_ik x_k)
// with Dinv the inverse main-diagonal block (Gauss-Seidel) or the inverted
// ILU pivot (ILU).  Smoother.cpp:331-403, ILU.cpp:333-397.
template <int DIM, int BS>
__global__ __launch_bounds__(kBlock) void k_colour_forward(SweepArgs a, int64_t p0, int64_t p1,
                                                           const double* __restrict__ b, double* __restrict__ x)
{
    if (a.guard && a.guard->done) return;
    const int64_t pos = p0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (pos >= p1) return;
    const int64_t row = a.perm[pos];
    double acc[BS];
#pragma unroll
    for (int q = 0; q < BS; ++q) acc[q] = b[row * BS + q];
    split_product<DIM, BS>(a, a.plow, pos, row, x, acc);
    double out[BS];
    block_mv<BS>(a.pdinv + pos, a.ld, acc, out);
#pragma unroll
    for (int q = 0; q < BS; ++q) x[row * BS + q] = out[q];
}

// Backward substitution (in place):
//   Gauss-Seidel: x_i = Dinv_i (A_ii x_i - sum_{c(k) > c(i)} A_ik x_k)
//   ILU:          x_i = x_i - sum_{c(k) > c(i)} U_ik x_k
// (Smoother.cpp:405-473, ILU.cpp:399-458).
template <int DIM, int BS, bool ILU>
__global__ __launch_bounds__(kBlock) void k_colour_backward(SweepArgs a, int64_t p0, int64_t p1,
                                                            double* __restrict__ x)
{
    if (a.guard && a.guard->done) return;
    const int64_t pos = p0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (pos >= p1) return;
    const int64_t row = a.perm[pos];
    double xi[BS], acc[BS];
#pragma unroll
    for (int q = 0; q < BS; ++q) xi[q] = x[row * BS + q];
    if (ILU) {
#pragma unroll
        for (int q = 0; q < BS; ++q) acc[q] = xi[q];
    } else {
        block_mv<BS>(a.pdiag + pos, a.ld, xi, acc);
    }
    split_product<DIM, BS>(a, a.pup, pos, row, x, acc);
    if (ILU) {
#pragma unroll
        for (int q = 0; q < BS; ++q) x[row * BS + q] = acc[q];
    } else {
        double out[BS];
        block_mv<BS>(a.pdinv + pos, a.ld, acc, out);
#pragma unroll
        for (int q = 0; q < BS; ++q) x[row * BS + q] = out[q];
    }
}

// ------------------------------------------------------------- ILU0
// Elimination of the rows of colour `color` (Solver_getILU, ILU.cpp:66-297):
// for colours c2 < color, for each neighbour k of colour c2 in row order and
// each neighbour j of k with colour > c2 that row i also holds,
// F_ij -= F_ik F_kj; then F_ii <- inv(F_ii) and F_ik <- F_ii F_ik for the
// neighbours of higher colour.  Rows of lower colours are final.
template <int DIM, int BS>
__global__ __launch_bounds__(kBlock) void k_ilu0_colour(ColourArgs a, int color, int64_t p0, int64_t p1,
                                                        int* __restrict__ failed)
{
    constexpr int NS = StencilShape<DIM>::NS, C = StencilShape<DIM>::CENTER, B2 = BS * BS;
    const int64_t pos = p0 + (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (pos >= p1) return;
    const int64_t row = a.perm[pos];
    const int64_t ld = a.ld;
    double* F = a.pval;
    int x0, y0, z0;
    node_coords(a.g, row, x0, y0, z0);
    for (int c2 = 0; c2 < color; ++c2) {
#pragma unroll 1
        for (int t1 = 0; t1 < NS; ++t1) {
            int64_t k;
            if (!neighbour<DIM>(a.g, t1, x0, y0, z0, row, k)) continue;
            if (a.colour[k] != c2) continue;
            int dx1, dy1, dz1;
            slot_delta<DIM>(t1, dx1, dy1, dz1);
            double Aik[B2];
#pragma unroll
            for (int e = 0; e < B2; ++e) Aik[e] = F[(int64_t)(t1 * B2 + e) * ld + pos];
            const int64_t kpos = a.ipos[k];
            const int xk = x0 + dx1, yk = y0 + dy1, zk = z0 + dz1;
#pragma unroll 1
            for (int t2 = 0; t2 < NS; ++t2) {
                int64_t j;
                if (!neighbour<DIM>(a.g, t2, xk, yk, zk, k, j)) continue;
                int dx2, dy2, dz2;
                slot_delta<DIM>(t2, dx2, dy2, dz2);
                const int ex = dx1 + dx2, ey = dy1 + dy2, ez = dz1 + dz2;
                if (ex < -1 || ex > 1 || ey < -1 || ey > 1 || ez < -1 || ez > 1) continue;   // j not in row i
                if (a.colour[j] <= c2) continue;
                const int t3 = (ex + 1) + 3 * (ey + 1) + (DIM == 3 ? 9 * (ez + 1) : 0);
                double S[B2];
#pragma unroll
                for (int e = 0; e < B2; ++e) S[e] = F[(int64_t)(t2 * B2 + e) * ld + kpos];
#pragma unroll
                for (int c = 0; c < BS; ++c)
#pragma unroll
                    for (int r = 0; r < BS; ++r) {
                        double s = 0.0;
#pragma unroll
                        for (int mm = 0; mm < BS; ++mm) s = fma(Aik[r + BS * mm], S[mm + BS * c], s);
                        F[(int64_t)(t3 * B2 + r + BS * c) * ld + pos] -= s;
                    }
            }
        }
    }
    double D[B2], Di[B2];
#pragma unroll
    for (int e = 0; e < B2; ++e) D[e] = F[(int64_t)(C * B2 + e) * ld + pos];
    if (!block_inverse<BS>(D, Di)) {
        failed[0] = 1;
        return;
    }
#pragma unroll
    for (int e = 0; e < B2; ++e) F[(int64_t)(C * B2 + e) * ld + pos] = Di[e];
#pragma unroll 1
    for (int t = 0; t < NS; ++t) {
        if (t == C) continue;
        int64_t k;
        if (!neighbour<DIM>(a.g, t, x0, y0, z0, row, k)) continue;
        if (a.colour[k] <= color) continue;
        double T[B2];
#pragma unroll
        for (int e = 0; e < B2; ++e) T[e] = F[(int64_t)(t * B2 + e) * ld + pos];
#pragma unroll
        for (int c = 0; c < BS; ++c)
#pragma unroll
            for (int r = 0; r < BS; ++r) {
                double s = 0.0;
#pragma unroll
                for (int mm = 0; mm < BS; ++mm) s = fma(Di[r + BS * mm], T[mm + BS * c], s);
                F[(int64_t)(t * B2 + r + BS * c) * ld + pos] = s;
            }
    }
}

// ------------------------------------------------------- vector utilities
__global__ __launch_bounds__(kBlock) void k_pcopy(double* __restrict__ y, const double* __restrict__ x, int64_t n,
                                                  const KrylovState* guard)
{
    if (guard && guard->done) return;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock) y[k] = x[k];
}

__global__ __launch_bounds__(kBlock) void k_padd(double* __restrict__ y, const double* __restrict__ x, int64_t n,
                                                 const KrylovState* guard)
{
    if (guard && guard->done) return;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock) y[k] += x[k];
}

int blocks_for(int64_t n)
{
    return (int)std::max<int64_t>(1, (n + kBlock - 1) / kBlock);
}

ColourArgs colour_args(rg_matrix* m)
{
    const rg_colouring* c = m->colouring;
    ColourArgs a;
    a.g = grid_of(m->g);
    a.perm = c->perm;
    a.ipos = c->ipos;
    a.colour = c->colour;
    a.pval = m->pval;
    a.ld = m->g.ld;
    return a;
}

SweepArgs sweep_args(rg_matrix* m, const KrylovState* guard)
{
    SweepArgs a;
    a.perm = m->colouring->perm;
    a.plow = m->plow;
    a.pup = m->pup;
    a.pdiag = m->pdiag;
    a.pdinv = m->pdinv;
    a.ld = m->g.ld;
    a.N = m->g.N;
    a.s1 = m->g.s1;
    a.s2 = m->g.s2;
    a.guard = guard;
    return a;
}

// one symmetric sweep of colour substitutions: forward b -> x over all
// colours, backward in place over colours nc-2 .. 0 (the last colour's
// backward step changes nothing and is skipped, as in both references)
template <bool ILU>
void colour_sweep(rg_matrix* m, const double* b, double* x, hipStream_t st, const KrylovState* guard)
{
    const rg_colouring* c = m->colouring;
    const SweepArgs a = sweep_args(m, guard);
    dispatch(m->g, false, [&](auto D, auto B, auto) {
        if constexpr (!ILU || B.value <= 3) {
            for (int col = 0; col < c->ncolors; ++col) {
                const int64_t p0 = c->cstart[col], p1 = c->cstart[col + 1];
                if (p1 > p0)
                    hipLaunchKernelGGL((k_colour_forward<D.value, B.value>), dim3(blocks_for(p1 - p0)),
                                       dim3(kBlock), 0, st, a, p0, p1, b, x);
            }
            for (int col = c->ncolors - 2; col >= 0; --col) {
                const int64_t p0 = c->cstart[col], p1 = c->cstart[col + 1];
                if (p1 > p0)
                    hipLaunchKernelGGL((k_colour_backward<D.value, B.value, ILU>), dim3(blocks_for(p1 - p0)),
                                       dim3(kBlock), 0, st, a, p0, p1, x);
            }
        }
    });
    RG_HIP(hipGetLastError());
}

}  // namespace

// ============================================================== host side

void ensure_colouring(rg_matrix* m)
{
    if (m->colouring) return;
    const Geom& g = m->g;
    RG_CHECK(g.N < ((int64_t)1 << 31), RG_VALUE_ERROR, "colouring: too many rows");
    hipStream_t st = m->ctx->stream;
    const Grid gr = grid_of(g);
    const int64_t n = g.N;
    const int nb = blocks_for(n);
    int32_t *marker = nullptr, *colour = nullptr;
    uint8_t* sel = nullptr;
    int* flag = nullptr;
    rg_colouring* c = new rg_colouring();
    int* hflag = reinterpret_cast<int*>(m->ctx->pinned);
    auto read_flag = [&]() {
        RG_HIP(hipMemcpyAsync(hflag, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        RG_HIP(hipStreamSynchronize(st));
        return *hflag;
    };
    try {
        RG_HIP(hipMalloc(&marker, sizeof(int32_t) * n));
        RG_HIP(hipMalloc(&colour, sizeof(int32_t) * n));
        RG_HIP(hipMalloc(&sel, n));
        RG_HIP(hipMalloc(&flag, sizeof(int)));
        RG_HIP(hipMemsetAsync(marker, 0xff, sizeof(int32_t) * n, st));   // -1: available
        RG_HIP(hipMemsetAsync(colour, 0xff, sizeof(int32_t) * n, st));   // -1: uncoloured
        int out = 0, rounds = 0;
        for (;;) {
            for (;;) {   // Pattern::mis
                // priorities of neighbours never tie on these grids, so every
                // round selects at least the smallest available one
                RG_CHECK(++rounds < 100000, RG_VALUE_ERROR, "colouring: no progress");
                RG_HIP(hipMemsetAsync(flag, 0, sizeof(int), st));
                if (g.dim == 2)
                    hipLaunchKernelGGL(k_mis_select<2>, dim3(nb), dim3(kBlock), 0, st, gr, marker, sel, flag);
                else
                    hipLaunchKernelGGL(k_mis_select<3>, dim3(nb), dim3(kBlock), 0, st, gr, marker, sel, flag);
                RG_HIP(hipGetLastError());
                if (!read_flag()) break;
                if (g.dim == 2)
                    hipLaunchKernelGGL(k_mis_mark<2>, dim3(nb), dim3(kBlock), 0, st, gr, marker, sel);
                else
                    hipLaunchKernelGGL(k_mis_mark<3>, dim3(nb), dim3(kBlock), 0, st, gr, marker, sel);
            }
            RG_HIP(hipMemsetAsync(flag, 0, sizeof(int), st));
            hipLaunchKernelGGL(k_colour_assign, dim3(nb), dim3(kBlock), 0, st, n, marker, colour, out, flag);
            RG_HIP(hipGetLastError());
            ++out;
            if (!read_flag()) break;
            RG_CHECK(out < 255, RG_VALUE_ERROR, "colouring: too many colours");
        }
        c->ncolors = out;
        // counting sort of the rows by colour (stable: ascending row within a colour)
        std::vector<int32_t> hc(n);
        RG_HIP(hipMemcpy(hc.data(), colour, sizeof(int32_t) * n, hipMemcpyDeviceToHost));
        c->cstart.assign(out + 1, 0);
        for (int64_t i = 0; i < n; ++i) c->cstart[hc[i] + 1]++;
        for (int k = 0; k < out; ++k) c->cstart[k + 1] += c->cstart[k];
        std::vector<int64_t> next(c->cstart.begin(), c->cstart.end() - 1);
        std::vector<int32_t> perm(n), ipos(n);
        std::vector<uint8_t> c8(n);
        for (int64_t i = 0; i < n; ++i) {
            const int64_t p = next[hc[i]]++;
            perm[p] = (int32_t)i;
            ipos[i] = (int32_t)p;
            c8[i] = (uint8_t)hc[i];
        }
        RG_HIP(hipMalloc(&c->perm, sizeof(int32_t) * n));
        RG_HIP(hipMalloc(&c->ipos, sizeof(int32_t) * n));
        RG_HIP(hipMalloc(&c->colour, n));
        RG_HIP(hipMemcpy(c->perm, perm.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        RG_HIP(hipMemcpy(c->ipos, ipos.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        RG_HIP(hipMemcpy(c->colour, c8.data(), n, hipMemcpyHostToDevice));
    } catch (...) {
        hipFree(marker);
        hipFree(colour);
        hipFree(sel);
        hipFree(flag);
        if (c->perm) hipFree(c->perm);
        if (c->ipos) hipFree(c->ipos);
        if (c->colour) hipFree(c->colour);
        delete c;
        throw;
    }
    hipFree(marker);
    hipFree(colour);
    hipFree(sel);
    hipFree(flag);
    m->colouring = c;
}

void free_colouring(rg_matrix* m)
{
    if (m->colouring) {
        hipFree(m->colouring->perm);
        hipFree(m->colouring->ipos);
        hipFree(m->colouring->colour);
        delete m->colouring;
        m->colouring = nullptr;
    }
    double* bufs[] = {m->pval, m->plow, m->pup, m->pdiag, m->pdinv};
    for (double* b : bufs)
        if (b) hipFree(b);
    m->pval = m->plow = m->pup = m->pdiag = m->pdinv = nullptr;
    m->pval_state = 0;
}

// Preconditioner_alloc (paso/src/Preconditioner.cpp:46-107) for the
// colour-based preconditioners.  Gauss-Seidel reads the live (balanced)
// values and the inverse diagonal kept with the solver data, so its copy is
// refreshed whenever the matrix was re-balanced; ILU0 factors are solver
// data and are kept until the values are reset without preserving it.
void prec_setup(rg_matrix* m, int type)
{
    if (type != RG_PRECONDITIONER_GAUSS_SEIDEL && type != RG_PRECONDITIONER_ILU0) return;
    const Geom& g = m->g;
    RG_CHECK(type != RG_PRECONDITIONER_ILU0 || g.bs <= 3, RG_VALUE_ERROR,
             "Solver_getILU: block size greater than 3 is not supported.");
    const int want = (type == RG_PRECONDITIONER_GAUSS_SEIDEL) ? 1 : 2;
    if (m->pval_state == want) return;   // reset when the values are re-balanced or the solver data dropped
    ensure_colouring(m);
    hipStream_t st = m->ctx->stream;
    const int B2 = g.bs * g.bs;
    const size_t full = sizeof(double) * (size_t)g.nslot * B2 * g.ld, blk = sizeof(double) * (size_t)B2 * g.ld;
    if (!m->plow) RG_HIP(hipMalloc(&m->plow, full));
    if (!m->pup) RG_HIP(hipMalloc(&m->pup, full));
    if (!m->pdiag) RG_HIP(hipMalloc(&m->pdiag, blk));
    if (!m->pdinv) RG_HIP(hipMalloc(&m->pdinv, blk));
    m->pval_state = 0;
    const Grid gr = grid_of(g);
    const int nb = blocks_for(g.N);
    const rg_colouring* c = m->colouring;
    if (want == 1) {
        // Gauss-Seidel: split the balanced values, inverse diagonal from invMain
        dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
            hipLaunchKernelGGL((k_split<D.value, B.value, S.value, true>), dim3(nb), dim3(kBlock), 0, st, m->val,
                               g.ld, gr, c->perm, c->colour, m->plow, m->pup, m->pdiag);
        });
        dispatch_bs(g.bs, [&](auto B) {
            hipLaunchKernelGGL((k_permute_diag<B.value>), dim3(nb), dim3(kBlock), 0, st, m->dinv, g.ld, g.N, c->perm,
                               m->pdinv);
        });
        RG_HIP(hipGetLastError());
    } else {
        // ILU0: eliminate in a full colour-ordered copy, then split the
        // factors (the main slot holds the inverted pivot)
        if (!m->pval) RG_HIP(hipMalloc(&m->pval, full));
        dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
            hipLaunchKernelGGL((k_permute_values<D.value, B.value, S.value>), dim3(nb), dim3(kBlock), 0, st, m->val,
                               g.ld, gr, c->perm, m->pval);
        });
        RG_HIP(hipGetLastError());
        int* flag = m->work->flag;
        RG_HIP(hipMemsetAsync(flag, 0, sizeof(int), st));
        const ColourArgs a = colour_args(m);
        dispatch(g, false, [&](auto D, auto B, auto) {
            if constexpr (B.value <= 3) {
                for (int col = 0; col < c->ncolors; ++col) {
                    const int64_t p0 = c->cstart[col], p1 = c->cstart[col + 1];
                    if (p1 > p0)
                        hipLaunchKernelGGL((k_ilu0_colour<D.value, B.value>), dim3(blocks_for(p1 - p0)),
                                           dim3(kBlock), 0, st, a, col, p0, p1, flag);
                }
                hipLaunchKernelGGL((k_split<D.value, B.value, false, false>), dim3(nb), dim3(kBlock), 0, st, m->pval,
                                   g.ld, gr, c->perm, c->colour, m->plow, m->pup, m->pdinv);
            }
        });
        RG_HIP(hipGetLastError());
        int failed = 0;
        RG_HIP(hipMemcpyAsync(&failed, flag, sizeof(int), hipMemcpyDeviceToHost, st));
        RG_HIP(hipStreamSynchronize(st));
        RG_CHECK(!failed, RG_VALUE_ERROR, "Solver_getILU: non-regular main diagonal block.");
    }
    m->pval_state = want;
}

// Preconditioner_solve (Preconditioner.cpp:111-137) for Gauss-Seidel, ILU0
// and none; Jacobi stays with the solver's fused kernels.
void prec_apply(rg_matrix* m, int type, int sweeps, const double* b, double* x, hipStream_t st,
                const KrylovState* guard)
{
    const int64_t n = m->g.N * m->g.bs;
    const int vg = 1024;
    if (type == RG_PRECONDITIONER_NONE) {
        hipLaunchKernelGGL(k_pcopy, dim3(vg), dim3(kBlock), 0, st, x, b, n, guard);
        RG_HIP(hipGetLastError());
        return;
    }
    if (type == RG_PRECONDITIONER_ILU0) {
        RG_CHECK(m->pval_state == 2, RG_STATE_ERROR, "ILU0: factors are not available");
        colour_sweep<true>(m, b, x, st, guard);
        return;
    }
    RG_CHECK(type == RG_PRECONDITIONER_GAUSS_SEIDEL && m->pval_state == 1, RG_STATE_ERROR,
             "Gauss-Seidel: smoother data is not available");
    colour_sweep<false>(m, b, x, st, guard);
    if (sweeps <= 1) return;
    rg_solver_work* w = m->work;
    if (!w->pb1) RG_HIP(hipMalloc(&w->pb1, sizeof(double) * n));
    if (!w->pb2) RG_HIP(hipMalloc(&w->pb2, sizeof(double) * n));
    for (int s = 1; s < sweeps; ++s) {
        // x += S^-1 (b - A x)   (Smoother.cpp:116-125)
        hipLaunchKernelGGL(k_pcopy, dim3(vg), dim3(kBlock), 0, st, w->pb1, b, n, guard);
        launch_spmv_on(m, -1.0, x, 1.0, w->pb1, st);
        colour_sweep<false>(m, w->pb1, w->pb2, st, guard);
        hipLaunchKernelGGL(k_padd, dim3(vg), dim3(kBlock), 0, st, x, w->pb2, n, guard);
    }
    RG_HIP(hipGetLastError());
}

}  // namespace rg
