// Batched Jacobi-PCG: (1) many right-hand sides on one operator -- the
// stencil SpMV becomes an SpMM over an N x K row-major block, so the matrix
// is streamed once per iteration for all K columns; (2) many independent
// scalar systems of uneven sizes -- every kernel spans all systems through
// a tile table (tile -> system, row range), so one iteration is one set of
// launches for the whole batch.  Each column / system has its own Krylov
// scalars (a KrylovState each, rg_krylov.h), its own convergence and its
// own done flag; the iterations are captured once into a hipGraph and
// replayed check_interval iterations at a time.
//
// Both solve the unbalanced system from x = 0 with the textbook Jacobi-PCG
// (PCG.cpp:62-368 without PASO's residual smoothing), the SpMV in FP64 or
// FP32 values (RG_METHOD_PCG_PLAIN / RG_METHOD_PCG_MIXED), the vectors,
// dots and updates in FP64 (with FP32 values on the z-marching SpMM the
// directions p and q = A p in FP32), and an FP64 residual replacement
// r = b - A x at the end of every chunk; a column converges only on that
// FP64 residual: ||b - A x||_2 <= tol ||b||_2 per column.  Tile shapes (rows per thread of
// the SpMM, rows per tile of the batch) are picked by timing candidates on
// the actual operator (autotune).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "rg_device.h"
#include "rg_dispatch.h"
#include "rg_krylov.h"
#include "rg_stencil.h"

namespace rg {

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ============================================================ many RHS (SpMM)
// Thread t of a block handles column k = t % KP of rows (t / KP) * R ..
// + R - 1 of each pass; KP = K rounded up to a power of two (lanes k >= K
// idle).  Vectors are row-major N x K (column index fastest).
struct MArgs {
    const double* dinv;      // FP64 Jacobi inverse, ld stride (unused slots)
    const double* b;
    double* x;
    double* r;
    double* p;
    double* q;
    float* pf;               // p and q in FP32 (the FP32-value z-marching path)
    float* qf;
    double* p2;              // the other direction block: the previous direction, and the
    float* pf2;              // target of the next one
    KrylovState* st;         // K states
    double* partials;        // [quantity][K][G]
    int64_t N, ld, s1, s2, chunk;
    int K;
};

// the direction vector p or q = A p of element type PT (FP64, or FP32 on the
// FP32-value z-marching path)
template <class PT>
__device__ __forceinline__ PT* dir_vec(double* d, float* f)
{
    if constexpr (std::is_same_v<PT, float>) return f;
    else return d;
}

// Per-column block sums of NV values, thread t holding CPT consecutive
// columns k = (t % KPC) * CPT + c in v[q][c]; partials[(q*K + k)*gridDim.x +
// blockIdx.x] (fixed order).
template <int KPC, int CPT, int NV>
__device__ __forceinline__ void publish_column_sums(double (&v)[NV][CPT], double* partials, int K)
{
    __shared__ double red[NV][kBlock * CPT];
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
        for (int c = 0; c < CPT; ++c) red[q][threadIdx.x * CPT + c] = v[q][c];
    __syncthreads();
    if ((int)threadIdx.x < K) {
        const int kk = threadIdx.x / CPT, c = threadIdx.x % CPT;
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
            for (int t = kk; t < kBlock; t += KPC) s += red[q][t * CPT + c];
            partials[((int64_t)q * K + threadIdx.x) * gridDim.x + blockIdx.x] = s;
        }
    }
}

template <int CPT>
struct ColVec {
    double v[CPT];
};
template <int CPT>
__device__ __forceinline__ ColVec<CPT> load_cols(const double* p)
{
    ColVec<CPT> r;
    if constexpr (CPT == 1) {
        r.v[0] = p[0];
    } else if constexpr (CPT == 2) {
        const double2 t = *reinterpret_cast<const double2*>(p);
        r.v[0] = t.x;
        r.v[1] = t.y;
    } else {
        const double2 t0 = *reinterpret_cast<const double2*>(p);
        const double2 t1 = *reinterpret_cast<const double2*>(p + 2);
        r.v[0] = t0.x;
        r.v[1] = t0.y;
        r.v[2] = t1.x;
        r.v[3] = t1.y;
    }
    return r;
}

// Y = A X (stencil) for R consecutive rows and CPT consecutive columns per
// thread (KPC threads per row; x windows shared by the three dx slots of
// each (dy, dz) pair), then epi(row, first column, acc[CPT]).
template <int DIM, typename VT, int KPC, int R, int CPT, class Epi>
__device__ __forceinline__ void spmm_rows(const VT* __restrict__ val, const MArgs& a, const double* __restrict__ X,
                                          Epi&& epi)
{
    constexpr int NPL = (DIM == 3) ? 9 : 3;
    constexpr int RPP = (kBlock / KPC) * R;
    const int k0 = (threadIdx.x % KPC) * CPT;
    const bool active = k0 < a.K;
    const int64_t sub = (int64_t)(threadIdx.x / KPC) * R;
    const RowRange rr = block_rows(a.N, a.chunk);
    for (int64_t i0 = rr.r0 + sub; i0 < rr.r1; i0 += RPP) {
        double acc[R][CPT];
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int c = 0; c < CPT; ++c) acc[r][c] = 0.0;
        // all nine (dy, dz) planes unrolled: their x window loads are in
        // flight together (one wave per SIMD for 4 columns per thread, yet
        // faster than one plane at a time: 1.9 against 2.7 ms on the 128^3
        // Brick with 64 columns)
#pragma unroll
        for (int pl = 0; pl < NPL; ++pl) {
            const int dy = pl % 3 - 1, dz = (DIM == 3) ? pl / 3 - 1 : 0;
            const int64_t b = i0 + dy * a.s1 + dz * a.s2;
            ColVec<CPT> w[R + 2];
#pragma unroll
            for (int t = 0; t < R + 2; ++t) {
                const int64_t j = b - 1 + t;
                const bool ok = active && j >= 0 && j < a.N;
                w[t] = load_cols<CPT>(X + (ok ? j : 0) * a.K + (active ? k0 : 0));
                if (!ok)
#pragma unroll
                    for (int c = 0; c < CPT; ++c) w[t].v[c] = 0.0;
            }
#pragma unroll
            for (int dx = 0; dx < 3; ++dx) {
                const int s = dx + 3 * pl;
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const int64_t row = i0 + r;
                    const double v = (row < rr.r1) ? (double)val[(int64_t)s * a.ld + row] : 0.0;
#pragma unroll
                    for (int c = 0; c < CPT; ++c) acc[r][c] = fma(v, w[r + dx].v[c], acc[r][c]);
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (active && i0 + r < rr.r1) epi(i0 + r, k0, acc[r]);
    }
}

// ------------------------------------------- z-marching SpMM (64 columns)
// Y = A X for exactly 64 columns on a Brick: a workgroup owns a TX x 4
// (x, y) tile of nodes (TX = 16 or 8), all 64 columns (lane = column, wave = y line of the
// tile) and a segment of kZSeg z-planes, and marches up z.  X enters LDS
// one z-plane of the tile at a time (TX + 2 by 6 nodes with the halo),
// prefetched into registers while the previous plane is in use; each X
// plane contributes to the three output planes it touches (its dz = +1, 0,
// -1 slots), whose sums stay in registers, and the lowest of them is then
// complete and handed to the epilogue.  Each X value is read from L2 about
// 1.9 times (the halo) instead of once per stencil slot, and the matrix
// values -- the same for the 64 lanes of a wave -- are staged in LDS one
// plane ahead and read back as broadcasts.
constexpr int kZY = 4, kZK = 64, kZSeg = 16;
constexpr int kZHY = kZY + 2;
static_assert(kZY * 64 == kBlock, "z-march tile: one wave per y line");

struct ZArgs {
    int64_t n0, n1, n2, ld;
    int nbx, nby, nseg;
};

// TX: x nodes per thread (the tile is TX x 4 nodes)
inline ZArgs zmarch_args(const Geom& g, int TX)
{
    ZArgs z;
    z.n0 = g.nn[0]; z.n1 = g.nn[1]; z.n2 = g.nn[2]; z.ld = g.ld;
    z.nbx = (int)((z.n0 + TX - 1) / TX);
    z.nby = (int)((z.n1 + kZY - 1) / kZY);
    z.nseg = (int)((z.n2 + kZSeg - 1) / kZSeg);
    return z;
}

// epi(row, sum) for every output row of the workgroup; X holds FP64 or
// FP32 entries (XT), the sums are FP64
template <int kZX, typename VT, typename XT, class Epi>
__device__ __forceinline__ void spmm_zmarch(const VT* __restrict__ val, const ZArgs z, const XT* __restrict__ X,
                                            Epi&& epi)
{
    constexpr int kZHX = kZX + 2;
    constexpr int kZRows = kZHX * kZHY;                  // halo rows per plane
    constexpr int XPC = 16 / (int)sizeof(XT);            // X entries per 16-byte piece
    constexpr int TPR = kZK / XPC;                       // staging threads per halo row
    constexpr int kZPer = (kZRows * TPR + kBlock - 1) / kBlock;   // pieces staged per thread
    __shared__ __align__(16) XT xs[kZRows * kZK];
    const int c = threadIdx.x & 63;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int b = blockIdx.x;
    const int bx = b % z.nbx, by = (b / z.nbx) % z.nby, bs = b / (z.nbx * z.nby);
    const int64_t x0 = (int64_t)bx * kZX, y0 = (int64_t)by * kZY;
    const int64_t y = y0 + wv;
    const bool yok = y < z.n1;
    const int64_t za = (int64_t)bs * kZSeg, zb = min(za + (int64_t)kZSeg, z.n2);
    const int nx = (int)min((int64_t)kZX, z.n0 - x0);
    // staging: TPR threads per halo row, 16 bytes (XPC columns) each; the
    // row of a piece depends on the lane, so its address arithmetic stays in
    // vector registers
    uint4 pre[kZPer];
    const int ch = threadIdx.x % TPR;
    auto fetch = [&](int64_t zp) {
#pragma unroll
        for (int k = 0; k < kZPer; ++k) {
            const int r = (int)(threadIdx.x / TPR) + (kBlock / TPR) * k;   // halo row
            const int hy = r / kZHX, hx = r % kZHX;
            const int64_t gx = x0 - 1 + hx, gy = y0 - 1 + hy;
            const bool ok = r < kZRows && gx >= 0 && gx < z.n0 && gy >= 0 && gy < z.n1;
            pre[k] = ok ? *reinterpret_cast<const uint4*>(X + ((zp * z.n1 + gy) * z.n0 + gx) * kZK + XPC * ch)
                        : uint4{0u, 0u, 0u, 0u};
        }
    };
    // the matrix values of the tile's rows, one output plane at a time, in a
    // ring of three planes in LDS (plane zo in slot zo % 3): staged one X
    // plane ahead like X (16-byte pieces, the row of a piece from the lane),
    // read back as wave-uniform (broadcast) LDS reads.  A partial tile's
    // pieces run on past its line, into the next rows or the value arrays'
    // slack; those products are never written.
    constexpr int VPC = 16 / (int)sizeof(VT);              // values per piece
    constexpr int CPL = kZX / VPC;                         // pieces per line
    constexpr int NCH = 27 * kZY * CPL;                    // pieces per output plane
    constexpr int VPER = (NCH + kBlock - 1) / kBlock;
    static_assert(kZX % VPC == 0, "z-march tile width");
    __shared__ __align__(16) VT vs[3 * 27 * kZY * kZX];
    uint4 vpre[VPER];
    auto fetch_vals = [&](int64_t zo) {
#pragma unroll
        for (int k = 0; k < VPER; ++k) {
            const int e = (int)threadIdx.x + kBlock * k;
            const int slot = e / (kZY * CPL), rem = e % (kZY * CPL);
            const int ty = rem / CPL, pc = rem % CPL;
            const bool ok = e < NCH && zo >= za && zo < zb && y0 + ty < z.n1;
            uint4 v = uint4{0u, 0u, 0u, 0u};
            if (ok)
                __builtin_memcpy(&v, val + (int64_t)slot * z.ld + ((zo * z.n1 + y0 + ty) * z.n0 + x0) + pc * VPC,
                                 sizeof(v));
            vpre[k] = v;
        }
    };
    auto store_vals = [&](int64_t zo) {
        VT* dst = vs + (int)(zo % 3) * (27 * kZY * kZX);
#pragma unroll
        for (int k = 0; k < VPER; ++k) {
            const int e = (int)threadIdx.x + kBlock * k;
            if (e < NCH) {
                const int slot = e / (kZY * CPL), rem = e % (kZY * CPL);
                const int ty = rem / CPL, pc = rem % CPL;
                *reinterpret_cast<uint4*>(dst + (slot * kZY + ty) * kZX + pc * VPC) = vpre[k];
            }
        }
    };
    double am[kZX], ac[kZX], ap[kZX];   // output planes zp-1, zp, zp+1
#pragma unroll
    for (int x = 0; x < kZX; ++x) am[x] = ac[x] = ap[x] = 0.0;
    // the staged X plane's contributions to output plane zo (dz = zp - zo)
    auto add_slot = [&](int64_t zo, int dz, int dy, int dx, const double (&w)[kZHX], double (&A)[kZX]) {
        const VT* vp = vs + (int)(zo % 3) * (27 * kZY * kZX) +
                       (((dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)) * kZY + wv) * kZX;
        VT t[kZX];
#pragma unroll
        for (int k = 0; k < kZX / VPC; ++k)
            *reinterpret_cast<uint4*>(t + k * VPC) = *reinterpret_cast<const uint4*>(vp + k * VPC);
#pragma unroll
        for (int x = 0; x < kZX; ++x) A[x] = fma((double)t[x], w[x + 1 + dx], A[x]);
    };
    auto finish = [&](int64_t zo, const double (&A)[kZX]) {
        const int64_t rowb = (zo * z.n1 + y) * z.n0 + x0;
#pragma unroll
        for (int x = 0; x < kZX; ++x)
            if (x < nx) epi(rowb + x, A[x]);
    };
    const int64_t zfirst = za > 0 ? za - 1 : 0;
    const int64_t zlast = zb < z.n2 ? zb : z.n2 - 1;
    // the first two output planes' values, then X plane zfirst
    fetch_vals(zfirst);
    store_vals(zfirst);
    fetch_vals(zfirst + 1);
    store_vals(zfirst + 1);
    fetch(zfirst);
    for (int64_t zp = zfirst; zp <= zlast; ++zp) {
        __syncthreads();   // every wave is done with the previous plane
#pragma unroll
        for (int k = 0; k < kZPer; ++k) {
            const int r = (int)(threadIdx.x / TPR) + (kBlock / TPR) * k;
            if (r < kZRows) *reinterpret_cast<uint4*>(xs + r * kZK + XPC * ch) = pre[k];
        }
        // plane zp+1's values (fetched during the previous step) replace
        // plane zp-2's, finished then
        if (zp > zfirst) store_vals(zp + 1);
        __syncthreads();
        if (zp < zlast) {   // in flight while this plane is used
            fetch(zp + 1);
            fetch_vals(zp + 2);
        }
        if (yok) {
            const bool cm = zp - 1 >= za, cc = zp >= za && zp < zb, cp = zp + 1 < zb;
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy) {
                double w[kZHX];
#pragma unroll
                for (int hx = 0; hx < kZHX; ++hx) w[hx] = (double)xs[((wv + 1 + dy) * kZHX + hx) * kZK + c];
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    if (cm) add_slot(zp - 1, 1, dy, dx, w, am);
                    if (cc) add_slot(zp, 0, dy, dx, w, ac);
                    if (cp) add_slot(zp + 1, -1, dy, dx, w, ap);
                }
            }
            if (cm) finish(zp - 1, am);   // output plane zp-1 has all its X planes
        }
#pragma unroll
        for (int x = 0; x < kZX; ++x) {
            am[x] = ac[x];
            ac[x] = ap[x];
            ap[x] = 0.0;
        }
    }
    // the top plane of the Brick has no X plane above it
    if (yok && zlast == z.n2 - 1 && zlast >= za && zlast < zb) finish(zlast, am);
}

// ------------------------------- z-marching SpMM, one x node per lane
// The tile and the march of spmm_zmarch with an 8-node x run, but lane =
// g + 8 xl: column group g (columns 8g .. 8g+7) of the node at x offset xl.
// A lane reads its node's matrix value once per slot (eight distinct LDS
// addresses per wave, the values staged in LDS as FP64 so that none is
// converted per use) and the X entries of its 8 columns at each (dx, dy)
// as 16-byte LDS reads; it keeps 3 x 8 sums (output planes zp-1, zp, zp+1).
constexpr int kZN = 8;   // x nodes of the tile
constexpr int kZG = 8;   // columns per lane

// epi(row, g, sums[8]) for every output row of the workgroup
template <typename VT, typename XT, class Epi>
__device__ __forceinline__ void spmm_zmarch_n(const VT* __restrict__ val, const ZArgs z, const XT* __restrict__ X,
                                              Epi&& epi)
{
    constexpr int kZX = kZN;
    constexpr int kZHX = kZX + 2;
    constexpr int kZRows = kZHX * kZHY;
    constexpr int XPC = 16 / (int)sizeof(XT);
    constexpr int TPR = kZK / XPC;
    constexpr int kZPer = (kZRows * TPR + kBlock - 1) / kBlock;
    __shared__ __align__(16) XT xs[kZRows * kZK];
    const int lane = threadIdx.x & 63;
    const int g = lane % kZG, xl = lane / kZG;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int b = blockIdx.x;
    const int bx = b % z.nbx, by = (b / z.nbx) % z.nby, bs = b / (z.nbx * z.nby);
    const int64_t x0 = (int64_t)bx * kZX, y0 = (int64_t)by * kZY;
    const int64_t y = y0 + wv;
    const bool yok = y < z.n1;
    const int64_t za = (int64_t)bs * kZSeg, zb = min(za + (int64_t)kZSeg, z.n2);
    const int nx = (int)min((int64_t)kZX, z.n0 - x0);
    uint4 pre[kZPer];
    const int ch = threadIdx.x % TPR;
    auto fetch = [&](int64_t zp) {
#pragma unroll
        for (int k = 0; k < kZPer; ++k) {
            const int r = (int)(threadIdx.x / TPR) + (kBlock / TPR) * k;   // halo row
            const int hy = r / kZHX, hx = r % kZHX;
            const int64_t gx = x0 - 1 + hx, gy = y0 - 1 + hy;
            const bool ok = r < kZRows && gx >= 0 && gx < z.n0 && gy >= 0 && gy < z.n1;
            pre[k] = ok ? *reinterpret_cast<const uint4*>(X + ((zp * z.n1 + gy) * z.n0 + gx) * kZK + XPC * ch)
                        : uint4{0u, 0u, 0u, 0u};
        }
    };
    // matrix values: a ring of three output planes in LDS as FP64, staged
    // one X plane ahead (16-byte pieces of the value arrays; a partial
    // tile's pieces run on past its line into the next rows or the arrays'
    // slack, and those products are never written)
    constexpr int VPC = 16 / (int)sizeof(VT);
    constexpr int CPL = kZX / VPC;
    constexpr int NCH = 27 * kZY * CPL;
    constexpr int VPER = (NCH + kBlock - 1) / kBlock;
    constexpr int kPlaneV = 27 * kZY * kZX;
    __shared__ __align__(16) double vs[3 * kPlaneV];
    uint4 vpre[VPER];
    auto fetch_vals = [&](int64_t zo) {
#pragma unroll
        for (int k = 0; k < VPER; ++k) {
            const int e = (int)threadIdx.x + kBlock * k;
            const int slot = e / (kZY * CPL), rem = e % (kZY * CPL);
            const int ty = rem / CPL, pc = rem % CPL;
            const bool ok = e < NCH && zo >= za && zo < zb && y0 + ty < z.n1;
            uint4 v = uint4{0u, 0u, 0u, 0u};
            if (ok)
                __builtin_memcpy(&v, val + (int64_t)slot * z.ld + ((zo * z.n1 + y0 + ty) * z.n0 + x0) + pc * VPC,
                                 sizeof(v));
            vpre[k] = v;
        }
    };
    auto store_vals = [&](int64_t zo) {
        double* dst = vs + (int)(zo % 3) * kPlaneV;
#pragma unroll
        for (int k = 0; k < VPER; ++k) {
            const int e = (int)threadIdx.x + kBlock * k;
            if (e < NCH) {
                const int slot = e / (kZY * CPL), rem = e % (kZY * CPL);
                const int ty = rem / CPL, pc = rem % CPL;
                VT t[VPC];
                __builtin_memcpy(t, &vpre[k], sizeof(t));
                double* d = dst + (slot * kZY + ty) * kZX + pc * VPC;
#pragma unroll
                for (int j = 0; j < VPC; j += 2) *reinterpret_cast<double2*>(d + j) = double2{(double)t[j], (double)t[j + 1]};
            }
        }
    };
    double am[kZG], ac[kZG], ap[kZG];   // output planes zp-1, zp, zp+1
#pragma unroll
    for (int c = 0; c < kZG; ++c) am[c] = ac[c] = ap[c] = 0.0;
    auto finish = [&](int64_t zo, const double (&A)[kZG]) {
        if (xl < nx) epi((zo * z.n1 + y) * z.n0 + x0 + xl, g, A);
    };
    const int64_t zfirst = za > 0 ? za - 1 : 0;
    const int64_t zlast = zb < z.n2 ? zb : z.n2 - 1;
    fetch_vals(zfirst);
    store_vals(zfirst);
    fetch_vals(zfirst + 1);
    store_vals(zfirst + 1);
    fetch(zfirst);
    for (int64_t zp = zfirst; zp <= zlast; ++zp) {
        __syncthreads();   // every wave is done with the previous plane
#pragma unroll
        for (int k = 0; k < kZPer; ++k) {
            const int r = (int)(threadIdx.x / TPR) + (kBlock / TPR) * k;
            if (r < kZRows) *reinterpret_cast<uint4*>(xs + r * kZK + XPC * ch) = pre[k];
        }
        if (zp > zfirst) store_vals(zp + 1);
        __syncthreads();
        if (zp < zlast) {   // in flight while this plane is used
            fetch(zp + 1);
            fetch_vals(zp + 2);
        }
        if (yok) {
            const bool cm = zp - 1 >= za, cc = zp >= za && zp < zb, cp = zp + 1 < zb;
            const double* vm = vs + (int)((zp + 2) % 3) * kPlaneV + wv * kZX + xl;   // plane zp-1
            const double* vc = vs + (int)(zp % 3) * kPlaneV + wv * kZX + xl;
            const double* vp = vs + (int)((zp + 1) % 3) * kPlaneV + wv * kZX + xl;
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy) {
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    XT wt[kZG];
                    const XT* src = xs + ((wv + 1 + dy) * kZHX + (xl + 1 + dx)) * kZK + kZG * g;
#pragma unroll
                    for (int k = 0; k < kZG / XPC; ++k)
                        *reinterpret_cast<uint4*>(wt + k * XPC) = *reinterpret_cast<const uint4*>(src + k * XPC);
                    double w[kZG];
#pragma unroll
                    for (int c = 0; c < kZG; ++c) w[c] = (double)wt[c];
                    const int s0 = (dx + 1) + 3 * (dy + 1);   // slot of dz = -1
                    if (cm) {   // X plane zp is output plane zp-1's dz = +1 neighbour
                        const double v = vm[(s0 + 18) * kZY * kZX];
#pragma unroll
                        for (int c = 0; c < kZG; ++c) am[c] = fma(v, w[c], am[c]);
                    }
                    if (cc) {
                        const double v = vc[(s0 + 9) * kZY * kZX];
#pragma unroll
                        for (int c = 0; c < kZG; ++c) ac[c] = fma(v, w[c], ac[c]);
                    }
                    if (cp) {
                        const double v = vp[s0 * kZY * kZX];
#pragma unroll
                        for (int c = 0; c < kZG; ++c) ap[c] = fma(v, w[c], ap[c]);
                    }
                }
            }
            if (cm) finish(zp - 1, am);
        }
#pragma unroll
        for (int c = 0; c < kZG; ++c) {
            am[c] = ac[c];
            ac[c] = ap[c];
            ap[c] = 0.0;
        }
    }
    if (yok && zlast == z.n2 - 1 && zlast >= za && zlast < zb) finish(zlast, am);
}

// Per-column block sums of NV values held as v[q][c] for column kZG g + c
// by lane g + kZG xl of every wave; partials[(q*K + k)*gridDim.x +
// blockIdx.x] in a fixed order (64 columns)
template <int NV>
__device__ __forceinline__ void publish_group_sums(double (&v)[NV][kZG], double* partials, int K)
{
    __shared__ double red[NV][kBlock / 64][kZK];
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
        for (int c = 0; c < kZG; ++c)
#pragma unroll
            for (int o = kZG; o < 64; o <<= 1) v[q][c] += __shfl_xor(v[q][c], o, 64);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    if (lane < kZG)
#pragma unroll
        for (int q = 0; q < NV; ++q)
#pragma unroll
            for (int c = 0; c < kZG; ++c) red[q][wv][lane * kZG + c] = v[q][c];
    __syncthreads();
    if ((int)threadIdx.x < K) {
#pragma unroll
        for (int q = 0; q < NV; ++q) {
            double s = 0.0;
#pragma unroll
            for (int w = 0; w < kBlock / 64; ++w) s += red[q][w][threadIdx.x];
            partials[((int64_t)q * K + threadIdx.x) * gridDim.x + blockIdx.x] = s;
        }
    }
}

template <typename T>
__device__ __forceinline__ void load8(const T* p, T (&t)[kZG])
{
#pragma unroll
    for (int k = 0; k < kZG * (int)sizeof(T) / 16; ++k)
        *reinterpret_cast<uint4*>(t + k * 16 / sizeof(T)) = *reinterpret_cast<const uint4*>(p + k * 16 / sizeof(T));
}
template <typename T>
__device__ __forceinline__ void store8(T* p, const T (&t)[kZG])
{
#pragma unroll
    for (int k = 0; k < kZG * (int)sizeof(T) / 16; ++k)
        *reinterpret_cast<uint4*>(p + k * 16 / sizeof(T)) = *reinterpret_cast<const uint4*>(t + k * 16 / sizeof(T));
}

// q = A p and the partial p.q per column, one x node per lane
template <typename VT, typename PT>
__global__ __launch_bounds__(kBlock) void k_mspmv_dot_zn(MArgs a, ZArgs z, const VT* __restrict__ val)
{
    double s[1][kZG] = {};
    const PT* P = dir_vec<PT>(a.p, a.pf);
    PT* Q = dir_vec<PT>(a.q, a.qf);
    spmm_zmarch_n<VT>(val, z, P, [&](int64_t i, int g, const double (&A)[kZG]) {
        PT qs[kZG], pv[kZG];
#pragma unroll
        for (int c = 0; c < kZG; ++c) qs[c] = (PT)A[c];
        store8(Q + i * kZK + kZG * g, qs);
        load8(P + i * kZK + kZG * g, pv);
#pragma unroll
        for (int c = 0; c < kZG; ++c) s[0][c] = fma((double)qs[c], (double)pv[c], s[0][c]);
    });
    publish_group_sums<1>(s, a.partials, a.K);
}

// r = b - A x (FP64 values), partials of r.D^-1 r and r.r, one x node per lane
__global__ __launch_bounds__(kBlock) void k_mreplace_zn(MArgs a, ZArgs z, const double* __restrict__ val)
{
    double s[2][kZG] = {};
    const int g = (threadIdx.x & 63) % kZG;
    bool live[kZG];
#pragma unroll
    for (int c = 0; c < kZG; ++c) live[c] = !a.st[kZG * g + c].done;
    spmm_zmarch_n<double>(val, z, a.x, [&](int64_t i, int gg, const double (&A)[kZG]) {
        const double di = a.dinv[i];
#pragma unroll
        for (int c = 0; c < kZG; ++c) {
            if (!live[c]) continue;
            const int64_t e = i * kZK + kZG * gg + c;
            const double rv = a.b[e] - A[c];
            a.r[e] = rv;
            s[0][c] += rv * di * rv;
            s[1][c] += rv * rv;
        }
    });
    publish_group_sums<2>(s, a.partials, a.K);
}

// Y = A X, one x node per lane
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_mspmv_zn(ZArgs z, const VT* __restrict__ val, const double* __restrict__ X,
                                                     double* __restrict__ Y)
{
    spmm_zmarch_n<VT>(val, z, X, [&](int64_t i, int g, const double (&A)[kZG]) { store8(Y + i * kZK + kZG * g, A); });
}

// q = A p and the partial p.q per column (64 columns, z-marching); with
// PT = float q is rounded to FP32 and the dot taken of the stored values
template <int TX, typename VT, typename PT>
__global__ __launch_bounds__(kBlock) void k_mspmv_dot_z(MArgs a, ZArgs z, const VT* __restrict__ val)
{
    double s[1][1] = {{0.0}};
    const int c = threadIdx.x & 63;
    const PT* P = dir_vec<PT>(a.p, a.pf);
    PT* Q = dir_vec<PT>(a.q, a.qf);
    // p at the row from memory again (its plane was staged two planes
    // earlier, usually still in L2): holding it in registers until the
    // row's sum is complete cost occupancy
    spmm_zmarch<TX, VT>(val, z, P, [&](int64_t i, double q) {
        const PT qs = (PT)q;
        Q[i * kZK + c] = qs;
        s[0][0] = fma((double)qs, (double)P[i * kZK + c], s[0][0]);
    });
    publish_column_sums<kZK, 1, 1>(s, a.partials, a.K);
}

// r = b - A x (FP64 values) with the partials of r.D^-1 r and r.r per
// column (64 columns, z-marching)
template <int TX>
__global__ __launch_bounds__(kBlock) void k_mreplace_z(MArgs a, ZArgs z, const double* __restrict__ val)
{
    double s[2][1] = {{0.0}, {0.0}};
    const int c = threadIdx.x & 63;
    const bool live = !a.st[c].done;
    spmm_zmarch<TX, double>(val, z, a.x, [&](int64_t i, double ax) {
        if (!live) return;
        const int64_t e = i * kZK + c;
        const double rv = a.b[e] - ax;
        a.r[e] = rv;
        s[0][0] += rv * a.dinv[i] * rv;
        s[1][0] += rv * rv;
    });
    publish_column_sums<kZK, 1, 2>(s, a.partials, a.K);
}

// Y = A X (64 columns, z-marching)
template <int TX, typename VT>
__global__ __launch_bounds__(kBlock) void k_mspmv_z(ZArgs z, const VT* __restrict__ val, const double* __restrict__ X,
                                                    double* __restrict__ Y)
{
    const int c = threadIdx.x & 63;
    spmm_zmarch<TX, VT>(val, z, X, [&](int64_t i, double q) { Y[i * kZK + c] = q; });
}

// q = A p, partial p.q per column
template <int DIM, typename VT, int KPC, int R, int CPT>
__global__ __launch_bounds__(kBlock) void k_mspmv_dot(MArgs a, const VT* __restrict__ val)
{
    double s[1][CPT] = {};
    spmm_rows<DIM, VT, KPC, R, CPT>(val, a, a.p, [&](int64_t i, int k0, const double (&acc)[CPT]) {
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            const int64_t e = i * a.K + k0 + c;
            a.q[e] = acc[c];
            s[0][c] = fma(acc[c], a.p[e], s[0][c]);
        }
    });
    publish_column_sums<KPC, CPT, 1>(s, a.partials, a.K);
}

// y = A x (plain SpMM, for the autotuner and rg_matrix_vector_multi)
template <int DIM, typename VT, int KPC, int R, int CPT>
__global__ __launch_bounds__(kBlock) void k_mspmv(MArgs a, const VT* __restrict__ val, const double* __restrict__ X,
                                                  double* __restrict__ Y)
{
    spmm_rows<DIM, VT, KPC, R, CPT>(val, a, X, [&](int64_t i, int k0, const double (&acc)[CPT]) {
#pragma unroll
        for (int c = 0; c < CPT; ++c) Y[i * a.K + k0 + c] = acc[c];
    });
}

// r = b - A x (FP64 values), partial r.D^-1 r and r.r per column
template <int DIM, int KPC, int R, int CPT>
__global__ __launch_bounds__(kBlock) void k_mreplace(MArgs a, const double* __restrict__ val)
{
    double s[2][CPT] = {};
    spmm_rows<DIM, double, KPC, R, CPT>(val, a, a.x, [&](int64_t i, int k0, const double (&acc)[CPT]) {
#pragma unroll
        for (int c = 0; c < CPT; ++c) {
            if (a.st[k0 + c].done) continue;
            const int64_t e = i * a.K + k0 + c;
            const double rv = a.b[e] - acc[c];
            a.r[e] = rv;
            s[0][c] += rv * a.dinv[i] * rv;
            s[1][c] += rv * rv;
        }
    });
    publish_column_sums<KPC, CPT, 2>(s, a.partials, a.K);
}

// element loop of the vector kernels: (row i, column k) pairs of this block
template <int KP, class F>
__device__ __forceinline__ void column_loop(const MArgs& a, F&& f)
{
    const int k = threadIdx.x % KP;
    if (k >= a.K) return;
    const RowRange rr = block_rows(a.N, a.chunk);
    for (int64_t i = rr.r0 + threadIdx.x / KP; i < rr.r1; i += kBlock / KP) f(i, k);
}

// r = b, p = D^-1 r, partial r.D^-1 r and r.r (x = 0)
template <int KP, typename PT>
__global__ __launch_bounds__(kBlock) void k_minit(MArgs a)
{
    double s[2][1] = {};
    PT* P = dir_vec<PT>(a.p, a.pf);
    column_loop<KP>(a, [&](int64_t i, int k) {
        const int64_t e = i * a.K + k;
        const double rv = a.b[e];
        const double z = a.dinv[i] * rv;
        a.r[e] = rv;
        a.x[e] = 0.0;
        P[e] = (PT)z;
        s[0][0] += rv * z;
        s[1][0] += rv * rv;
    });
    publish_column_sums<KP, 1, 2>(s, a.partials, a.K);
}

// r -= alpha q, partial r.D^-1 r and r.r per column.  With WITHX (the
// last iteration of a chunk, before the residual replacement reads x) also
// x += alpha_prev p_prev (pending from the previous iteration, also for a
// column that ended there) + alpha p.
template <int KP, bool WITHX, typename PT>
__global__ __launch_bounds__(kBlock) void k_mupdate(MArgs a)
{
    double s[2][1] = {};
    const int k = threadIdx.x % KP;
    const int kc = k < a.K ? k : 0;
    const bool live = k < a.K && !a.st[kc].done;
    const bool xprev = WITHX && k < a.K && a.st[kc].early_prev;
    const double alpha = live ? a.st[kc].alpha : 0.0;
    const double alpha_prev = xprev ? a.st[kc].alpha_prev : 0.0;
    const PT* P = dir_vec<PT>(a.p, a.pf);
    const PT* P2 = dir_vec<PT>(a.p2, a.pf2);
    const PT* Q = dir_vec<PT>(a.q, a.qf);
    column_loop<KP>(a, [&](int64_t i, int kk) {
        if (!live && !xprev) return;
        const int64_t e = i * a.K + kk;
        if (WITHX) {   // only the pending directions are read (an ended column's may never have been written)
            double xv = a.x[e];
            if (live) xv = fma(alpha, (double)P[e], xv);
            if (xprev) xv = fma(alpha_prev, (double)P2[e], xv);
            a.x[e] = xv;
        }
        if (!live) return;
        const double rv = fma(-alpha, (double)Q[e], a.r[e]);
        a.r[e] = rv;
        s[0][0] += rv * a.dinv[i] * rv;
        s[1][0] += rv * rv;
    });
    publish_column_sums<KP, 1, 2>(s, a.partials, a.K);
}

// p_next = D^-1 r + beta p into the other direction block (a.p2) for the
// running columns.  x is updated every second iteration: with ODD also x +=
// alpha_prev p_prev + alpha p where pending (p_prev is read from the block
// that p_next then overwrites), so x and p are read once for two updates.
template <int KP, typename PT, bool ODD>
__global__ __launch_bounds__(kBlock) void k_mpupdate(MArgs a)
{
    const int k = threadIdx.x % KP;
    const int kc = k < a.K ? k : 0;
    const bool live = k < a.K && !a.st[kc].done;
    const bool xcur = ODD && k < a.K && a.st[kc].early;
    const bool xprev = ODD && k < a.K && a.st[kc].early_prev;
    const double beta = live ? a.st[kc].beta : 0.0;
    const double alpha = xcur ? a.st[kc].alpha : 0.0;
    const double alpha_prev = xprev ? a.st[kc].alpha_prev : 0.0;
    const PT* P = dir_vec<PT>(a.p, a.pf);
    PT* P2 = dir_vec<PT>(a.p2, a.pf2);
    column_loop<KP>(a, [&](int64_t i, int kk) {
        if (!live && !xcur && !xprev) return;
        const int64_t e = i * a.K + kk;
        // only the pending directions are read: the current one of a column
        // that ended before writing it may never have been written
        if (ODD && (xcur || xprev)) {
            double xv = a.x[e];
            if (xcur) xv = fma(alpha, (double)P[e], xv);
            if (xprev) xv = fma(alpha_prev, (double)P2[e], xv);
            a.x[e] = xv;
        }
        if (live) P2[e] = (PT)fma(beta, (double)P[e], a.dinv[i] * a.r[e]);
    });
}

// one workgroup per column: reduce its partials, apply the stage
template <int STAGE, int NV>
__global__ __launch_bounds__(kBlock) void k_mstage(KrylovState* sts, const double* __restrict__ partials, int G, int K)
{
    KrylovState* st = sts + blockIdx.x;
    if (!stage_starts<STAGE>() && st->done) return;
    double tot[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const double* p = partials + ((int64_t)q * K + blockIdx.x) * G;
        double s = 0.0;
        for (int b = threadIdx.x; b < G; b += kBlock) s += p[b];
        tot[q] = s;
    }
    block_sum<NV>(tot);
    if (threadIdx.x != 0) return;
    stage_apply<STAGE, NV>(st, tot);
}

// dinv[i] = 1 / a_ii of a scalar stencil matrix (main slot), unbalanced
__global__ void k_diag_inv(const double* __restrict__ val, int64_t ld, int64_t N, int center, double* __restrict__ d,
                           int* __restrict__ flag)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < N; i += (int64_t)gridDim.x * blockDim.x) {
        const double a = val[(int64_t)center * ld + i];
        if (a == 0.0) flag[0] = 1;
        d[i] = a != 0.0 ? 1.0 / a : 0.0;
    }
}

// ========================================================= batch of systems
struct Inst {
    const double* val;
    const float* val32;
    const double* dinv;
    int64_t N, ld, s1, s2;
    int64_t off;    // offset of the system's vectors in the batch vectors
};
struct Tile {
    int32_t inst;
    int32_t pad;
    int64_t r0, r1;
};
struct BArgs {
    const Inst* inst;
    const Tile* tiles;
    const double* b;
    double* x;
    double* r;
    double* p;
    double* q;
    KrylovState* st;
    double* partials;   // [quantity][G]
};

template <int NV>
__device__ __forceinline__ void publish_tile_sums(double (&v)[NV], double* partials)
{
    publish_block_sums<NV>(v, partials);
}

template <int DIM, typename VT>
__global__ __launch_bounds__(kBlock) void k_bspmv_dot(BArgs a)
{
    const Tile t = a.tiles[blockIdx.x];
    const Inst in = a.inst[t.inst];
    double s[1] = {0.0};
    if (!a.st[t.inst].done) {
        const VT* val = std::is_same<VT, float>::value ? (const VT*)in.val32 : (const VT*)in.val;
        const double* p = a.p + in.off;
        double* q = a.q + in.off;
        stencil_block<DIM, 1, false, VT>(val, in.ld, in.N, in.s1, in.s2, t.r0, t.r1, p,
                                         [&](int64_t i, const double (&acc)[1]) {
                                             q[i] = acc[0];
                                             s[0] = fma(acc[0], p[i], s[0]);
                                         });
    }
    publish_tile_sums<1>(s, a.partials);
}

template <int DIM, typename VT>
__global__ __launch_bounds__(kBlock) void k_bspmv(BArgs a, const double* __restrict__ X, double* __restrict__ Y)
{
    const Tile t = a.tiles[blockIdx.x];
    const Inst in = a.inst[t.inst];
    const VT* val = std::is_same<VT, float>::value ? (const VT*)in.val32 : (const VT*)in.val;
    const double* x = X + in.off;
    double* y = Y + in.off;
    stencil_block<DIM, 1, false, VT>(val, in.ld, in.N, in.s1, in.s2, t.r0, t.r1, x,
                                     [&](int64_t i, const double (&acc)[1]) { y[i] = acc[0]; });
}

template <int DIM>
__global__ __launch_bounds__(kBlock) void k_breplace(BArgs a)
{
    const Tile t = a.tiles[blockIdx.x];
    const Inst in = a.inst[t.inst];
    double s[2] = {0.0, 0.0};
    if (!a.st[t.inst].done) {
        const double* x = a.x + in.off;
        const double* b = a.b + in.off;
        double* r = a.r + in.off;
        stencil_block<DIM, 1, false, double>(in.val, in.ld, in.N, in.s1, in.s2, t.r0, t.r1, x,
                                             [&](int64_t i, const double (&acc)[1]) {
                                                 const double rv = b[i] - acc[0];
                                                 r[i] = rv;
                                                 s[0] += rv * in.dinv[i] * rv;
                                                 s[1] += rv * rv;
                                             });
    }
    publish_tile_sums<2>(s, a.partials);
}

__global__ __launch_bounds__(kBlock) void k_binit(BArgs a)
{
    const Tile t = a.tiles[blockIdx.x];
    const Inst in = a.inst[t.inst];
    double s[2] = {0.0, 0.0};
    for (int64_t i = t.r0 + threadIdx.x; i < t.r1; i += kBlock) {
        const int64_t e = in.off + i;
        const double rv = a.b[e];
        const double z = in.dinv[i] * rv;
        a.r[e] = rv;
        a.x[e] = 0.0;
        a.p[e] = z;
        s[0] += rv * z;
        s[1] += rv * rv;
    }
    publish_tile_sums<2>(s, a.partials);
}

__global__ __launch_bounds__(kBlock) void k_bupdate(BArgs a)
{
    const Tile t = a.tiles[blockIdx.x];
    const Inst in = a.inst[t.inst];
    double s[2] = {0.0, 0.0};
    if (!a.st[t.inst].done) {
        const double alpha = a.st[t.inst].alpha;
        for (int64_t i = t.r0 + threadIdx.x; i < t.r1; i += kBlock) {
            const int64_t e = in.off + i;
            a.x[e] = fma(alpha, a.p[e], a.x[e]);
            const double rv = fma(-alpha, a.q[e], a.r[e]);
            a.r[e] = rv;
            s[0] += rv * in.dinv[i] * rv;
            s[1] += rv * rv;
        }
    }
    publish_tile_sums<2>(s, a.partials);
}

__global__ __launch_bounds__(kBlock) void k_bpupdate(BArgs a)
{
    const Tile t = a.tiles[blockIdx.x];
    const Inst in = a.inst[t.inst];
    if (a.st[t.inst].done) return;
    const double beta = a.st[t.inst].beta;
    for (int64_t i = t.r0 + threadIdx.x; i < t.r1; i += kBlock) {
        const int64_t e = in.off + i;
        a.p[e] = fma(beta, a.p[e], in.dinv[i] * a.r[e]);
    }
}

// one workgroup per system: reduce the partials of its tiles [tb[j], tb[j+1])
template <int STAGE, int NV>
__global__ __launch_bounds__(kBlock) void k_bstage(KrylovState* sts, const double* __restrict__ partials, int G,
                                                   const int* __restrict__ tb)
{
    KrylovState* st = sts + blockIdx.x;
    if (!stage_starts<STAGE>() && st->done) return;
    const int b0 = tb[blockIdx.x], b1 = tb[blockIdx.x + 1];
    double tot[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        const double* p = partials + (int64_t)q * G;
        double s = 0.0;
        for (int b = b0 + threadIdx.x; b < b1; b += kBlock) s += p[b];
        tot[q] = s;
    }
    block_sum<NV>(tot);
    if (threadIdx.x != 0) return;
    stage_apply<STAGE, NV>(st, tot);
}

// ================================================================ host side
template <class T>
struct DevArray {
    T* p = nullptr;
    size_t n = 0;
    void alloc(size_t count)
    {
        if (count <= n && p) return;
        if (p) hipFree(p);
        p = nullptr;
        n = 0;
        RG_HIP(hipMalloc(&p, sizeof(T) * std::max<size_t>(count, 1)));
        n = count;
    }
    ~DevArray()
    {
        if (p) hipFree(p);
    }
};

// The chunked, graph-captured iteration loop shared by both batched
// solvers.  iteration(s, last) issues one iteration on stream s.
struct ChunkRunner {
    hipStream_t stream = nullptr;      // caller's stream
    hipStream_t capture = nullptr;     // internal capture stream
    hipGraphExec_t graph = nullptr;
    int graph_iters = -1;
    std::vector<unsigned char> graph_key;   // everything the captured launches depend on

    ~ChunkRunner()
    {
        if (graph) hipGraphExecDestroy(graph);
        if (capture) hipStreamDestroy(capture);
    }

    template <class F>
    void run(const std::vector<unsigned char>& key, int iters, F&& iteration)
    {
        if (!graph || graph_iters != iters || graph_key != key) {
            if (graph) {
                RG_HIP(hipGraphExecDestroy(graph));
                graph = nullptr;
            }
            if (!capture) RG_HIP(hipStreamCreateWithFlags(&capture, hipStreamNonBlocking));
            hipGraph_t g;
            RG_HIP(hipStreamBeginCapture(capture, hipStreamCaptureModeThreadLocal));
            try {
                for (int k = 0; k < iters; ++k) iteration(capture, k == iters - 1);
            } catch (...) {
                hipStreamEndCapture(capture, &g);
                throw;
            }
            RG_HIP(hipStreamEndCapture(capture, &g));
            RG_HIP(hipGraphInstantiate(&graph, g, nullptr, nullptr, 0));
            RG_HIP(hipGraphDestroy(g));
            graph_iters = iters;
            graph_key = key;
        }
        RG_HIP(hipGraphLaunch(graph, stream));
    }
};

// Polls `count` Krylov states every chunk (two chunks in flight) until all
// are done.
template <class Chunk>
void poll_until_done(hipStream_t st, KrylovState* d_states, int count, Chunk&& chunk, std::vector<KrylovState>& out)
{
    KrylovState* pin = nullptr;
    RG_HIP(hipHostMalloc(&pin, sizeof(KrylovState) * 2 * count));
    hipEvent_t evs[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        if (evs[0]) hipEventDestroy(evs[0]);
        if (evs[1]) hipEventDestroy(evs[1]);
        hipHostFree(pin);
    };
    try {
        RG_HIP(hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
        RG_HIP(hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
        int launched = 0;
        bool done = false;
        while (!done) {
            chunk();
            KrylovState* slot = pin + (launched % 2) * count;
            RG_HIP(hipMemcpyAsync(slot, d_states, sizeof(KrylovState) * count, hipMemcpyDeviceToHost, st));
            RG_HIP(hipEventRecord(evs[launched % 2], st));
            ++launched;
            if (launched >= 2) {
                RG_HIP(hipEventSynchronize(evs[(launched - 2) % 2]));
                const KrylovState* s = pin + ((launched - 2) % 2) * count;
                done = true;
                for (int j = 0; j < count; ++j) done = done && s[j].done;
            }
        }
        RG_HIP(hipStreamSynchronize(st));
        out.resize(count);
        RG_HIP(hipMemcpy(out.data(), d_states, sizeof(KrylovState) * count, hipMemcpyDeviceToHost));
    } catch (...) {
        cleanup();
        throw;
    }
    cleanup();
}

void init_states(hipStream_t st, KrylovState* d, int count, double rtol, int maxit)
{
    std::vector<KrylovState> h(count);
    for (auto& s : h) {
        std::memset(&s, 0, sizeof(s));
        s.tol = rtol;
        s.maxit = maxit;
        s.replace_mode = 1;
    }
    RG_HIP(hipMemcpyAsync(d, h.data(), sizeof(KrylovState) * count, hipMemcpyHostToDevice, st));
    RG_HIP(hipStreamSynchronize(st));
}

void fill_diagnostics(const std::vector<KrylovState>& s, double t, rg_diagnostics* dg, int count)
{
    for (int j = 0; j < count; ++j) {
        rg_diagnostics& d = dg[j];
        d.num_iter = s[j].num_iter;
        d.converged = s[j].status == KS_CONVERGED;
        d.status = s[j].status == KS_CONVERGED ? RG_OK
                   : s[j].status == KS_MAXITER ? RG_MAX_ITER_REACHED
                                               : RG_BREAKDOWN;
        d.outer_steps = 1;
        d.residual_norm = s[j].norm;
        d.set_up_time = 0.0;
        d.net_time = t;
        d.time = t;
        d.fallbacks = 0;
        d.reserved = 0;
    }
}

int batch_status(const rg_diagnostics* dg, int count, bool accept)
{
    int worst = RG_OK;
    for (int j = 0; j < count; ++j)
        if (dg[j].status != RG_OK && !(accept && dg[j].status == RG_MAX_ITER_REACHED)) worst = dg[j].status;
    return worst;
}

// Jacobi inverse (unbalanced) and, for FP32 values, the FP32 copy.
void prepare_matrix(rg_matrix* m, bool fp32, double* dinv)
{
    const Geom& g = m->g;
    RG_CHECK(!m->is_balanced, RG_STATE_ERROR,
             "batched solve: the matrix was balanced by an earlier solve; set or reset its values first");
    hipStream_t st = m->ctx->stream;
    DevArray<int> flag;
    flag.alloc(1);
    RG_HIP(hipMemsetAsync(flag.p, 0, sizeof(int), st));
    const int center = (g.dim == 3 ? 27 : 9) / 2;
    hipLaunchKernelGGL(k_diag_inv, dim3(std::min<int64_t>((g.N + 255) / 256, 4096)), dim3(256), 0, st, m->val, g.ld,
                       g.N, center, dinv, flag.p);
    RG_HIP(hipGetLastError());
    int failed = 0;
    RG_HIP(hipMemcpyAsync(&failed, flag.p, sizeof(int), hipMemcpyDeviceToHost, st));
    RG_HIP(hipStreamSynchronize(st));
    RG_CHECK(!failed, RG_VALUE_ERROR, "batched solve: zero on the main diagonal");
    if (fp32 && !m->val32_valid) {
        if (!m->val32) {
            RG_HIP(hipMalloc(&m->val32, value_count(m->g) * sizeof(float)));
        }
        launch_to_fp32(m);
        m->val32_valid = true;
    }
}

template <class F>
void dispatch_kp(int K, F&& f)
{
    if (K <= 1) f(IC<1>{});
    else if (K <= 2) f(IC<2>{});
    else if (K <= 4) f(IC<4>{});
    else if (K <= 8) f(IC<8>{});
    else if (K <= 16) f(IC<16>{});
    else if (K <= 32) f(IC<32>{});
    else f(IC<64>{});
}


}  // namespace

// ---------------------------------------------------------------- many RHS
// SpMM tile shapes: (rows per thread R, columns per thread CPT).  A shape
// applies when CPT divides the number of columns.  Shapes 4-6 are the
// z-marching LDS tiles (exactly 64 columns, Brick only): 4 and 5 one column
// per lane over an x run of 16 or 8 nodes, 6 one x node and 8 columns per
// lane of an 8-node run (spmm_zmarch_n).  Longer row runs per thread on the
// row-chunked tiles (R = 4, 8: R + 2 x windows per plane for 3 R products)
// spilled with the planes unrolled and ran 2.7-3.1 ms with them rolled,
// against 1.9 ms.
constexpr int kNumShapes = 7;
constexpr int kZShape = 4, kZShapeEnd = 7;
constexpr int kZNodeShape = 6;   // z-marching, one x node per lane
constexpr int kShapeR[kNumShapes] = {1, 1, 1, 2, 16, 8, kZN};
constexpr int kShapeC[kNumShapes] = {1, 2, 4, 2, kZK, kZK, kZK};

inline bool is_zshape(int shape) { return shape >= kZShape && shape < kZShapeEnd; }

inline bool shape_applies(int shape, int K, int dim)
{
    if (is_zshape(shape)) return K == kZK && dim == 3;
    return K % kShapeC[shape] == 0;
}

inline int pow2_at_least(int v)
{
    int p = 1;
    while (p < v) p <<= 1;
    return p;
}

// f(IC<KPC>, IC<R>, IC<CPT>) for shape `shape` and K columns
template <class F>
void dispatch_shape(int K, int shape, F&& f)
{
    const int R = kShapeR[shape], C = kShapeC[shape];
    RG_CHECK(K % C == 0, RG_VALUE_ERROR, "SpMM shape does not divide the column count");
    const int KPC = pow2_at_least(K / C);
    auto go = [&](auto RC, auto CC) {
        switch (KPC) {
            case 1: f(IC<1>{}, RC, CC); break;
            case 2: f(IC<2>{}, RC, CC); break;
            case 4: f(IC<4>{}, RC, CC); break;
            case 8: f(IC<8>{}, RC, CC); break;
            case 16: f(IC<16>{}, RC, CC); break;
            case 32: f(IC<32>{}, RC, CC); break;
            default: f(IC<64>{}, RC, CC); break;
        }
    };
    if (R == 1 && C == 1) go(IC<1>{}, IC<1>{});
    else if (R == 1 && C == 2) go(IC<1>{}, IC<2>{});
    else if (R == 1 && C == 4) go(IC<1>{}, IC<4>{});
    else if (R == 2 && C == 2) go(IC<2>{}, IC<2>{});
    else RG_CHECK(false, RG_VALUE_ERROR, "SpMM shape has no kernel");
}

struct MultiSolver {
    rg_matrix* m;
    int K, KP, shape;
    int rshape;     // shape of the row-chunked SpMM kernels (the FP64 residual replacement)
    bool fp32;
    bool zm;        // z-marching SpMM
    bool pf32;      // p and q = A p stored in FP32 (FP32 values on the z-marching SpMM)
    double* pd[2] = {nullptr, nullptr};   // the two direction blocks (FP64 or, with pf32, FP32)
    float* pfl[2] = {nullptr, nullptr};
    int cur = 0;    // direction block of the current iteration
    bool odd = false;   // parity of the current iteration (x is updated on odd ones)
    hipStream_t st;
    MArgs a;
    Launch L;       // row-chunked SpMM / vector kernels
    ZArgs z{};
    int Gs;         // workgroups (partials per column) of the iteration's SpMM

    MultiSolver(rg_matrix* mm, int k, int shp, bool f32) : m(mm), K(k), shape(shp), fp32(f32), st(mm->ctx->stream)
    {
        KP = pow2_at_least(K);
        zm = is_zshape(shape);
        pf32 = zm && fp32;
        RG_CHECK(shape_applies(shape, K, m->g.dim), RG_VALUE_ERROR,
                 "SpMM shape does not apply to this column count or dimension");
        rshape = zm ? 2 : shape;
        const int kpc = pow2_at_least(K / kShapeC[rshape]);
        const int64_t rpp = (int64_t)(kBlock / kpc) * kShapeR[rshape];
        L = chunked_launch((m->g.N + rpp - 1) / rpp * kBlock, 1, 2048);
        // chunked_launch counted threads; convert to rows per block
        L.chunk = L.chunk / kBlock * rpp;
        std::memset(&a, 0, sizeof(a));
        a.N = m->g.N; a.ld = m->g.ld; a.s1 = m->g.s1; a.s2 = m->g.s2; a.chunk = L.chunk; a.K = K;
        Gs = L.grid;
        if (zm) {
            z = zmarch_args(m->g, kShapeR[shape]);
            Gs = z.nbx * z.nby * z.nseg;
        }
    }

    // partials per column of the reducing kernels (SpMM and vector kernels)
    int max_grid() const { return std::max(L.grid, Gs); }

    template <class F>
    void by_tx(F&& f)
    {
        if (kShapeR[shape] == 16) f(IC<16>{});
        else f(IC<8>{});
    }

    // the z-marching kernels of this shape
    void z_spmm(hipStream_t s, const double* X, double* Y)
    {
        if (shape == kZNodeShape) {
            if (fp32) hipLaunchKernelGGL((k_mspmv_zn<float>), dim3(Gs), dim3(kBlock), 0, s, z, m->val32, X, Y);
            else hipLaunchKernelGGL((k_mspmv_zn<double>), dim3(Gs), dim3(kBlock), 0, s, z, m->val, X, Y);
            return;
        }
        by_tx([&](auto TX) {
            if (fp32) hipLaunchKernelGGL((k_mspmv_z<TX.value, float>), dim3(Gs), dim3(kBlock), 0, s, z, m->val32, X, Y);
            else hipLaunchKernelGGL((k_mspmv_z<TX.value, double>), dim3(Gs), dim3(kBlock), 0, s, z, m->val, X, Y);
        });
    }
    void z_spmm_dot(hipStream_t s)
    {
        if (shape == kZNodeShape) {
            if (fp32) hipLaunchKernelGGL((k_mspmv_dot_zn<float, float>), dim3(Gs), dim3(kBlock), 0, s, a, z, m->val32);
            else hipLaunchKernelGGL((k_mspmv_dot_zn<double, double>), dim3(Gs), dim3(kBlock), 0, s, a, z, m->val);
            return;
        }
        by_tx([&](auto TX) {
            if (fp32) hipLaunchKernelGGL((k_mspmv_dot_z<TX.value, float, float>), dim3(Gs), dim3(kBlock), 0, s, a, z, m->val32);
            else hipLaunchKernelGGL((k_mspmv_dot_z<TX.value, double, double>), dim3(Gs), dim3(kBlock), 0, s, a, z, m->val);
        });
    }
    void z_replace(hipStream_t s)
    {
        if (shape == kZNodeShape) {
            hipLaunchKernelGGL(k_mreplace_zn, dim3(Gs), dim3(kBlock), 0, s, a, z, m->val);
            return;
        }
        by_tx([&](auto TX) {
            hipLaunchKernelGGL((k_mreplace_z<TX.value>), dim3(Gs), dim3(kBlock), 0, s, a, z, m->val);
        });
    }

    void spmm(const double* X, double* Y)
    {
        const Geom& g = m->g;
        if (zm) {
            z_spmm(st, X, Y);
            RG_HIP(hipGetLastError());
            return;
        }
        dispatch_shape(K, shape, [&](auto KPC, auto RC, auto CC) {
            constexpr int P = KPC.value, Rv = RC.value, Cv = CC.value;
            if (g.dim == 3) {
                if (fp32) hipLaunchKernelGGL((k_mspmv<3, float, P, Rv, Cv>), dim3(L.grid), dim3(kBlock), 0, st, a, m->val32, X, Y);
                else hipLaunchKernelGGL((k_mspmv<3, double, P, Rv, Cv>), dim3(L.grid), dim3(kBlock), 0, st, a, m->val, X, Y);
            } else {
                if (fp32) hipLaunchKernelGGL((k_mspmv<2, float, P, Rv, Cv>), dim3(L.grid), dim3(kBlock), 0, st, a, m->val32, X, Y);
                else hipLaunchKernelGGL((k_mspmv<2, double, P, Rv, Cv>), dim3(L.grid), dim3(kBlock), 0, st, a, m->val, X, Y);
            }
        });
        RG_HIP(hipGetLastError());
    }

    // f(IC<KP>, PT{}) for the vector kernels: KP lanes per row, direction
    // vectors of type PT
    template <class F>
    void by_kp(F&& f)
    {
        if (pf32) f(IC<kZK>{}, float{});
        else dispatch_kp(K, [&](auto KPC) { f(KPC, double{}); });
    }

    template <int STAGE, int NV>
    void stage(hipStream_t s, int G = -1)
    {
        hipLaunchKernelGGL((k_mstage<STAGE, NV>), dim3(K), dim3(kBlock), 0, s, a.st, a.partials, G < 0 ? L.grid : G,
                           K);
    }

    void init(hipStream_t s)
    {
        by_kp([&](auto KPC, auto pt) {
            hipLaunchKernelGGL((k_minit<KPC.value, decltype(pt)>), dim3(L.grid), dim3(kBlock), 0, s, a);
        });
        stage<ST_BATCH_INIT, 2>(s);
        RG_HIP(hipGetLastError());
    }

    // One iteration.  The x update is deferred into the direction update of
    // every second (odd) iteration, which adds alpha_prev p_prev + alpha p
    // and so reads and writes x once per two iterations; the direction
    // blocks alternate (p in block cur, the next direction into cur ^ 1,
    // over the previous one).  A chunk starts on an even iteration and
    // ends on an odd one, whose x update goes into the r update (the
    // residual replacement reads x).  ST_CGX_SPMV clears the pending flag
    // of every column, ended ones included; ST_CGX2_SPMV moves it to the
    // previous one.
    void iteration(hipStream_t s, bool last)
    {
        RG_CHECK(!last || odd, RG_VALUE_ERROR, "solve_multi: a chunk must hold an even number of iterations");
        a.p = pd[cur];
        a.pf = pfl[cur];
        a.p2 = pd[cur ^ 1];
        a.pf2 = pfl[cur ^ 1];
        spmm_dot(s);
        if (odd) stage<ST_CGX2_SPMV, 1>(s, Gs);
        else stage<ST_CGX_SPMV, 1>(s, Gs);
        rest_of_iteration(s, last);
        cur ^= 1;
        odd = !odd;
    }

    // q = A p with the partials of p.q (a.p, a.q, a.partials)
    void spmm_dot(hipStream_t s)
    {
        const Geom& g = m->g;
        if (zm) {
            z_spmm_dot(s);
        } else {
            dispatch_shape(K, shape, [&](auto KPC, auto RC, auto CC) {
                constexpr int P = KPC.value, Rv = RC.value, Cv = CC.value;
                if (g.dim == 3) {
                    if (fp32) hipLaunchKernelGGL((k_mspmv_dot<3, float, P, Rv, Cv>), dim3(L.grid), dim3(kBlock), 0, s, a, m->val32);
                    else hipLaunchKernelGGL((k_mspmv_dot<3, double, P, Rv, Cv>), dim3(L.grid), dim3(kBlock), 0, s, a, m->val);
                } else {
                    if (fp32) hipLaunchKernelGGL((k_mspmv_dot<2, float, P, Rv, Cv>), dim3(L.grid), dim3(kBlock), 0, s, a, m->val32);
                    else hipLaunchKernelGGL((k_mspmv_dot<2, double, P, Rv, Cv>), dim3(L.grid), dim3(kBlock), 0, s, a, m->val);
                }
            });
        }
        RG_HIP(hipGetLastError());
    }

    void rest_of_iteration(hipStream_t s, bool last)
    {
        const Geom& g = m->g;
        by_kp([&](auto KPC, auto pt) {
            using PT = decltype(pt);
            if (last) hipLaunchKernelGGL((k_mupdate<KPC.value, true, PT>), dim3(L.grid), dim3(kBlock), 0, s, a);
            else hipLaunchKernelGGL((k_mupdate<KPC.value, false, PT>), dim3(L.grid), dim3(kBlock), 0, s, a);
        });
        if (last) stage<ST_CG_UPDATE, 2>(s);
        else stage<ST_CGX_UPDATE, 2>(s);
        if (last) {
            if (zm) {
                z_replace(s);
                stage<ST_CG_REPLACE, 2>(s, Gs);
            } else {
                dispatch_shape(K, rshape, [&](auto KPC, auto RC, auto CC) {
                    constexpr int P = KPC.value, Rv = RC.value, Cv = CC.value;
                    if (g.dim == 3) hipLaunchKernelGGL((k_mreplace<3, P, Rv, Cv>), dim3(L.grid), dim3(kBlock), 0, s, a, m->val);
                    else hipLaunchKernelGGL((k_mreplace<2, P, Rv, Cv>), dim3(L.grid), dim3(kBlock), 0, s, a, m->val);
                });
                stage<ST_CG_REPLACE, 2>(s);
            }
        }
        by_kp([&](auto KPC, auto pt) {
            using PT = decltype(pt);
            if (odd && !last) hipLaunchKernelGGL((k_mpupdate<KPC.value, PT, true>), dim3(L.grid), dim3(kBlock), 0, s, a);
            else hipLaunchKernelGGL((k_mpupdate<KPC.value, PT, false>), dim3(L.grid), dim3(kBlock), 0, s, a);
        });
        RG_HIP(hipGetLastError());
    }
};

// Times `reps` SpMMs of each applicable shape as the solver runs them (with
// the partial dot products p.q); returns the best.
int autotune_multi(rg_matrix* m, int K, bool fp32, const double* X, double* Y, int reps, double* times)
{
    hipStream_t st = m->ctx->stream;
    hipEvent_t e0, e1;
    RG_HIP(hipEventCreate(&e0));
    RG_HIP(hipEventCreate(&e1));
    int best = 0;
    float bt = 1e30f;
    DevArray<double> part;
    for (int c = 0; c < kNumShapes; ++c) {
        if (times) times[c] = -1.0;
        if (!shape_applies(c, K, m->g.dim)) continue;
        MultiSolver S(m, K, c, fp32);
        part.alloc((size_t)K * S.max_grid());
        S.a.p = const_cast<double*>(X);
        S.a.q = Y;
        // timing only: with FP32 directions the first half of X's and Y's
        // bytes serve as the FP32 blocks
        S.a.pf = reinterpret_cast<float*>(const_cast<double*>(X));
        S.a.qf = reinterpret_cast<float*>(Y);
        S.a.partials = part.p;
        S.spmm_dot(st);   // warm-up
        RG_HIP(hipEventRecord(e0, st));
        for (int k = 0; k < reps; ++k) S.spmm_dot(st);
        RG_HIP(hipEventRecord(e1, st));
        RG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        RG_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (times) times[c] = ms / 1e3 / reps;
        if (ms < bt) {
            bt = ms;
            best = c;
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return best;
}

int solve_multi(rg_matrix* m, double* X, const double* B, int K, const rg_options& opt, rg_diagnostics* dg)
{
    const double t0 = now_s();
    const Geom& g = m->g;
    RG_CHECK(g.bs == 1 && !m->symmetric, RG_VALUE_ERROR, "solve_multi: scalar matrices with full storage only");
    RG_CHECK(K >= 1 && K <= 64, RG_VALUE_ERROR, "solve_multi: 1 to 64 right-hand sides");
    RG_CHECK(opt.method == RG_METHOD_PCG_PLAIN || opt.method == RG_METHOD_PCG_MIXED, RG_VALUE_ERROR,
             "solve_multi: method must be RG_METHOD_PCG_PLAIN or RG_METHOD_PCG_MIXED");
    RG_CHECK(opt.preconditioner == RG_PRECONDITIONER_JACOBI && opt.sweeps <= 1, RG_VALUE_ERROR,
             "solve_multi: one Jacobi sweep is the only preconditioner");
    RG_CHECK(opt.tolerance > 0.0 && opt.tolerance <= 1.0, RG_VALUE_ERROR, "solve_multi: tolerance must be in (0, 1]");
    const bool fp32 = opt.method == RG_METHOD_PCG_MIXED;
    hipStream_t st = m->ctx->stream;
    const size_t nv = (size_t)g.N * K;
    DevArray<double> dinv, r, p, q, part;
    DevArray<float> pf, qf;
    DevArray<KrylovState> sts;
    dinv.alloc(g.ld);
    r.alloc(nv);
    sts.alloc(K);
    prepare_matrix(m, fp32, dinv.p);
    int shape = m->multi_shape - 1;
    if (shape < 0 || !shape_applies(shape, K, g.dim)) {
        shape = autotune_multi(m, K, fp32, B, r.p, 5, nullptr);
        m->multi_shape = shape + 1;
    }
    MultiSolver S(m, K, shape, fp32);
    DevArray<double> p2;
    DevArray<float> pf2;
    if (S.pf32) {
        pf.alloc(nv);
        pf2.alloc(nv);
        qf.alloc(nv);
    } else {
        p.alloc(nv);
        p2.alloc(nv);
        q.alloc(nv);
    }
    S.pd[0] = p.p; S.pd[1] = p2.p; S.pfl[0] = pf.p; S.pfl[1] = pf2.p;
    // the second direction block starts zero (an ended column's entries are
    // never read, but stay finite)
    if (S.pf32) RG_HIP(hipMemsetAsync(pf2.p, 0, sizeof(float) * nv, st));
    else RG_HIP(hipMemsetAsync(p2.p, 0, sizeof(double) * nv, st));
    part.alloc((size_t)2 * K * S.max_grid());
    S.a.dinv = dinv.p; S.a.b = B; S.a.x = X; S.a.r = r.p; S.a.p = p.p; S.a.q = q.p; S.a.pf = pf.p; S.a.qf = qf.p;
    S.a.st = sts.p; S.a.partials = part.p;
    init_states(st, sts.p, K, opt.tolerance, opt.iter_max);
    S.init(st);   // p into block 0 (S.cur)
    int check = opt.check_interval > 0 ? opt.check_interval : 32;
    check += check & 1;   // chunks of whole iteration pairs (the x update spans two)
    ChunkRunner run;
    run.stream = st;
    std::vector<KrylovState> fin;
    const std::vector<unsigned char> key;
    poll_until_done(st, sts.p, K, [&]() { run.run(key, check, [&](hipStream_t s, bool last) { S.iteration(s, last); }); },
                    fin);
    fill_diagnostics(fin, now_s() - t0, dg, K);
    return batch_status(dg, K, opt.accept_failed_convergence != 0);
}

void multi_vector(rg_matrix* m, const double* X, double* Y, int K)
{
    RG_CHECK(m->g.bs == 1 && !m->symmetric, RG_VALUE_ERROR, "matrix_vector_multi: scalar matrices with full storage only");
    RG_CHECK(K >= 1 && K <= 64, RG_VALUE_ERROR, "matrix_vector_multi: 1 to 64 columns");
    RG_CHECK(!m->is_balanced, RG_STATE_ERROR, "matrix_vector_multi: the matrix is balanced (set or reset values first)");
    int shape = m->multi_shape - 1;
    if (shape < 0 || !shape_applies(shape, K, m->g.dim)) shape = 0;
    MultiSolver S(m, K, shape, false);
    S.spmm(X, Y);
}

}  // namespace rg

// ---------------------------------------------------------- batch of systems
struct rg_batch {
    rg_context* ctx = nullptr;
    std::vector<rg_matrix*> mats;
    int dim = 2;
    std::vector<int64_t> offsets;   // count + 1 (aligned to 64 doubles)
    int tile_mult = 0;              // rows per tile = kBlock * R * tile_mult (0: autotune)
    rg::DevArray<rg::Inst> inst;
    rg::DevArray<rg::Tile> tiles;
    rg::DevArray<int> tile_begin;
    rg::DevArray<double> dinv;
    rg::DevArray<double> r, p, q, part;
    rg::DevArray<rg::KrylovState> sts;
    int G = 0;
    int built_mult = -1;
    bool built_fp32 = false;
    rg::ChunkRunner run;
};

namespace rg {

namespace {

constexpr int kTileCands[4] = {1, 2, 4, 8};

void build_tiles(rg_batch* B, int mult, bool fp32)
{
    const int R = fp32 ? spmv_rows_per_thread<float>() : spmv_rows_per_thread<double>();
    const int64_t tile = (int64_t)kBlock * R * mult;
    std::vector<Tile> t;
    std::vector<int> tb;
    for (size_t j = 0; j < B->mats.size(); ++j) {
        tb.push_back((int)t.size());
        const int64_t N = B->mats[j]->g.N;
        for (int64_t r0 = 0; r0 < N; r0 += tile) t.push_back(Tile{(int32_t)j, 0, r0, std::min(N, r0 + tile)});
    }
    tb.push_back((int)t.size());
    B->G = (int)t.size();
    hipStream_t st = B->ctx->stream;
    B->tiles.alloc(t.size());
    B->tile_begin.alloc(tb.size());
    RG_HIP(hipMemcpyAsync(B->tiles.p, t.data(), sizeof(Tile) * t.size(), hipMemcpyHostToDevice, st));
    RG_HIP(hipMemcpyAsync(B->tile_begin.p, tb.data(), sizeof(int) * tb.size(), hipMemcpyHostToDevice, st));
    RG_HIP(hipStreamSynchronize(st));
    B->built_mult = mult;
    B->built_fp32 = fp32;
}

void build_inst(rg_batch* B)
{
    std::vector<Inst> h;
    for (size_t j = 0; j < B->mats.size(); ++j) {
        const rg_matrix* m = B->mats[j];
        Inst in;
        in.val = m->val;
        in.val32 = m->val32;
        in.dinv = B->dinv.p + B->offsets[j];
        in.N = m->g.N; in.ld = m->g.ld; in.s1 = m->g.s1; in.s2 = m->g.s2;
        in.off = B->offsets[j];
        h.push_back(in);
    }
    B->inst.alloc(h.size());
    RG_HIP(hipMemcpyAsync(B->inst.p, h.data(), sizeof(Inst) * h.size(), hipMemcpyHostToDevice, B->ctx->stream));
    RG_HIP(hipStreamSynchronize(B->ctx->stream));
}

BArgs bargs(rg_batch* B, const double* b, double* x)
{
    BArgs a;
    a.inst = B->inst.p; a.tiles = B->tiles.p; a.b = b; a.x = x;
    a.r = B->r.p; a.p = B->p.p; a.q = B->q.p; a.st = B->sts.p; a.partials = B->part.p;
    return a;
}

void batch_spmv(rg_batch* B, bool fp32, const double* X, double* Y)
{
    BArgs a = bargs(B, nullptr, nullptr);
    hipStream_t st = B->ctx->stream;
    if (B->dim == 3) {
        if (fp32) hipLaunchKernelGGL((k_bspmv<3, float>), dim3(B->G), dim3(kBlock), 0, st, a, X, Y);
        else hipLaunchKernelGGL((k_bspmv<3, double>), dim3(B->G), dim3(kBlock), 0, st, a, X, Y);
    } else {
        if (fp32) hipLaunchKernelGGL((k_bspmv<2, float>), dim3(B->G), dim3(kBlock), 0, st, a, X, Y);
        else hipLaunchKernelGGL((k_bspmv<2, double>), dim3(B->G), dim3(kBlock), 0, st, a, X, Y);
    }
    RG_HIP(hipGetLastError());
}

template <int STAGE, int NV>
void bstage(rg_batch* B, hipStream_t s)
{
    hipLaunchKernelGGL((k_bstage<STAGE, NV>), dim3((int)B->mats.size()), dim3(kBlock), 0, s, B->sts.p, B->part.p, B->G,
                       B->tile_begin.p);
}

void batch_iteration(rg_batch* B, bool fp32, const BArgs& a, hipStream_t s, bool last)
{
    const int G = B->G;
    auto spmv_dot = [&](auto D) {
        if (fp32) hipLaunchKernelGGL((k_bspmv_dot<D.value, float>), dim3(G), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_bspmv_dot<D.value, double>), dim3(G), dim3(kBlock), 0, s, a);
    };
    if (B->dim == 3) spmv_dot(IC<3>{});
    else spmv_dot(IC<2>{});
    bstage<ST_CG_SPMV, 1>(B, s);
    hipLaunchKernelGGL(k_bupdate, dim3(G), dim3(kBlock), 0, s, a);
    bstage<ST_CG_UPDATE, 2>(B, s);
    if (last) {
        if (B->dim == 3) hipLaunchKernelGGL((k_breplace<3>), dim3(G), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_breplace<2>), dim3(G), dim3(kBlock), 0, s, a);
        bstage<ST_CG_REPLACE, 2>(B, s);
    }
    hipLaunchKernelGGL(k_bpupdate, dim3(G), dim3(kBlock), 0, s, a);
    RG_HIP(hipGetLastError());
}

void batch_prepare(rg_batch* B, bool fp32)
{
    const int64_t tot = B->offsets.back();
    B->dinv.alloc(tot);
    for (size_t j = 0; j < B->mats.size(); ++j) prepare_matrix(B->mats[j], fp32, B->dinv.p + B->offsets[j]);
    B->r.alloc(tot);
    B->p.alloc(tot);
    B->q.alloc(tot);
    B->sts.alloc(B->mats.size());
    build_inst(B);
}

}  // namespace

int batch_autotune(rg_batch* B, bool fp32, int reps, double* times)
{
    batch_prepare(B, fp32);
    hipStream_t st = B->ctx->stream;
    const int64_t tot = B->offsets.back();
    DevArray<double> x;
    x.alloc(tot);
    RG_HIP(hipMemsetAsync(x.p, 0, sizeof(double) * tot, st));
    hipEvent_t e0, e1;
    RG_HIP(hipEventCreate(&e0));
    RG_HIP(hipEventCreate(&e1));
    int best = 1;
    float bt = 1e30f;
    for (int c = 0; c < 4; ++c) {
        build_tiles(B, kTileCands[c], fp32);
        batch_spmv(B, fp32, x.p, B->q.p);
        RG_HIP(hipEventRecord(e0, st));
        for (int k = 0; k < reps; ++k) batch_spmv(B, fp32, x.p, B->q.p);
        RG_HIP(hipEventRecord(e1, st));
        RG_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        RG_HIP(hipEventElapsedTime(&ms, e0, e1));
        if (times) times[c] = ms / 1e3 / reps;
        if (ms < bt) {
            bt = ms;
            best = kTileCands[c];
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    B->tile_mult = best;
    return best;
}

int batch_solve(rg_batch* B, double* X, const double* b, const rg_options& opt, rg_diagnostics* dg)
{
    const double t0 = now_s();
    const int count = (int)B->mats.size();
    RG_CHECK(opt.method == RG_METHOD_PCG_PLAIN || opt.method == RG_METHOD_PCG_MIXED, RG_VALUE_ERROR,
             "batch_solve: method must be RG_METHOD_PCG_PLAIN or RG_METHOD_PCG_MIXED");
    RG_CHECK(opt.preconditioner == RG_PRECONDITIONER_JACOBI && opt.sweeps <= 1, RG_VALUE_ERROR,
             "batch_solve: one Jacobi sweep is the only preconditioner");
    RG_CHECK(opt.tolerance > 0.0 && opt.tolerance <= 1.0, RG_VALUE_ERROR, "batch_solve: tolerance must be in (0, 1]");
    const bool fp32 = opt.method == RG_METHOD_PCG_MIXED;
    if (B->tile_mult <= 0) batch_autotune(B, fp32, 5, nullptr);
    batch_prepare(B, fp32);
    if (B->built_mult != B->tile_mult || B->built_fp32 != fp32) build_tiles(B, B->tile_mult, fp32);
    B->part.alloc((size_t)2 * B->G);
    hipStream_t st = B->ctx->stream;
    init_states(st, B->sts.p, count, opt.tolerance, opt.iter_max);
    const BArgs a = bargs(B, b, X);
    hipLaunchKernelGGL(k_binit, dim3(B->G), dim3(kBlock), 0, st, a);
    bstage<ST_BATCH_INIT, 2>(B, st);
    RG_HIP(hipGetLastError());
    const int check = opt.check_interval > 0 ? opt.check_interval : 32;
    B->run.stream = st;
    // the captured launches depend on the argument block, the tile table and the precision
    std::vector<unsigned char> key(sizeof(BArgs) + 3 * sizeof(int));
    const int extra[3] = {fp32 ? 1 : 0, B->tile_mult, B->G};
    std::memcpy(key.data(), &a, sizeof(BArgs));
    std::memcpy(key.data() + sizeof(BArgs), extra, sizeof(extra));
    std::vector<KrylovState> fin;
    poll_until_done(st, B->sts.p, count,
                    [&]() { B->run.run(key, check, [&](hipStream_t s, bool last) { batch_iteration(B, fp32, a, s, last); }); },
                    fin);
    fill_diagnostics(fin, now_s() - t0, dg, count);
    return batch_status(dg, count, opt.accept_failed_convergence != 0);
}

void batch_vector(rg_batch* B, const double* X, double* Y)
{
    for (auto* m : B->mats)
        RG_CHECK(!m->is_balanced, RG_STATE_ERROR, "batch_vector: a matrix is balanced (set or reset values first)");
    if (B->inst.n == 0) {
        B->dinv.alloc(B->offsets.back());
        build_inst(B);
    }
    if (B->built_mult < 0) build_tiles(B, B->tile_mult > 0 ? B->tile_mult : 1, false);
    batch_spmv(B, false, X, Y);
}

}  // namespace rg

// ==================================================================== C ABI
namespace rg {
void set_last_error(const std::string& msg);
}

namespace {
template <class F>
int batch_guard(F&& f)
{
    try {
        rg::set_last_error("");
        return f();
    } catch (const rg::Error& e) {
        rg::set_last_error(e.what());
        (void)hipGetLastError();
        return e.code;
    } catch (const std::bad_alloc&) {
        rg::set_last_error("out of host memory");
        return RG_MEMORY_ERROR;
    } catch (const std::exception& e) {
        rg::set_last_error(e.what());
        (void)hipGetLastError();
        return RG_VALUE_ERROR;
    }
}
}  // namespace

extern "C" {

int rg_matrix_vector_multi(rg_matrix* m, const double* X, double* Y, int nrhs)
{
    return batch_guard([&] {
        RG_CHECK(m && X && Y, RG_VALUE_ERROR, "matrix_vector_multi: NULL argument");
        RG_HIP(hipSetDevice(m->ctx->device));
        rg::multi_vector(m, X, Y, nrhs);
        return RG_OK;
    });
}

int rg_solve_multi(rg_matrix* m, double* X, const double* B, int nrhs, const rg_options* opt, rg_diagnostics* diag)
{
    return batch_guard([&] {
        RG_CHECK(m && X && B && opt && diag, RG_VALUE_ERROR, "solve_multi: NULL argument");
        RG_HIP(hipSetDevice(m->ctx->device));
        return rg::solve_multi(m, X, B, nrhs, *opt, diag);
    });
}

int rg_multi_autotune(rg_matrix* m, int nrhs, int fp32, int reps, int* shape, double* times)
{
    return batch_guard([&] {
        RG_CHECK(m && shape, RG_VALUE_ERROR, "multi_autotune: NULL argument");
        RG_CHECK(m->g.bs == 1 && !m->symmetric, RG_VALUE_ERROR, "multi_autotune: scalar full-storage matrices only");
        RG_CHECK(nrhs >= 1 && nrhs <= 64 && reps >= 1, RG_VALUE_ERROR, "multi_autotune: 1..64 columns, reps >= 1");
        RG_CHECK(!m->is_balanced, RG_STATE_ERROR, "multi_autotune: the matrix is balanced (set or reset values first)");
        RG_HIP(hipSetDevice(m->ctx->device));
        rg::DevArray<double> x, y, d;
        x.alloc((size_t)m->g.N * nrhs);
        y.alloc((size_t)m->g.N * nrhs);
        RG_HIP(hipMemsetAsync(x.p, 0, sizeof(double) * m->g.N * nrhs, m->ctx->stream));
        d.alloc(m->g.ld);
        if (fp32) rg::prepare_matrix(m, true, d.p);
        *shape = rg::autotune_multi(m, nrhs, fp32 != 0, x.p, y.p, reps, times);
        m->multi_shape = *shape + 1;
        return RG_OK;
    });
}

int rg_multi_set_shape(rg_matrix* m, int shape)
{
    return batch_guard([&] {
        RG_CHECK(m, RG_VALUE_ERROR, "multi_set_shape: NULL matrix");
        RG_CHECK(shape >= -1 && shape < rg::kNumShapes, RG_VALUE_ERROR,
                 "multi_set_shape: shape must be -1 (autotune) or 0..6");
        m->multi_shape = shape + 1;
        return RG_OK;
    });
}

int rg_batch_create(rg_context* ctx, int count, rg_matrix* const* mats, rg_batch** out)
{
    return batch_guard([&] {
        RG_CHECK(ctx && mats && out && count >= 1, RG_VALUE_ERROR, "batch_create: NULL argument or empty batch");
        rg_batch* B = new rg_batch();
        B->ctx = ctx;
        int64_t off = 0;
        for (int j = 0; j < count; ++j) {
            rg_matrix* m = mats[j];
            if (!m || m->ctx != ctx || m->g.bs != 1 || m->symmetric || m->g.dim != mats[0]->g.dim) {
                delete B;
                throw rg::Error(RG_VALUE_ERROR,
                                "batch_create: scalar full-storage matrices of one dimension on this context only");
            }
            B->mats.push_back(m);
            B->offsets.push_back(off);
            off += (m->g.N + 63) / 64 * 64;
        }
        B->offsets.push_back(off);
        B->dim = mats[0]->g.dim;
        *out = B;
        return RG_OK;
    });
}

int rg_batch_destroy(rg_batch* b)
{
    if (b) {
        hipStreamSynchronize(b->ctx->stream);
        delete b;
    }
    return RG_OK;
}

int rg_batch_offsets(const rg_batch* b, int64_t* offsets)
{
    return batch_guard([&] {
        RG_CHECK(b && offsets, RG_VALUE_ERROR, "batch_offsets: NULL argument");
        for (size_t j = 0; j < b->offsets.size(); ++j) offsets[j] = b->offsets[j];
        return RG_OK;
    });
}

int rg_batch_vector(rg_batch* b, const double* x, double* y)
{
    return batch_guard([&] {
        RG_CHECK(b && x && y, RG_VALUE_ERROR, "batch_vector: NULL argument");
        RG_HIP(hipSetDevice(b->ctx->device));
        rg::batch_vector(b, x, y);
        return RG_OK;
    });
}

int rg_batch_autotune(rg_batch* b, int fp32, int reps, int* tile_mult, double* times)
{
    return batch_guard([&] {
        RG_CHECK(b && tile_mult && reps >= 1, RG_VALUE_ERROR, "batch_autotune: bad argument");
        RG_HIP(hipSetDevice(b->ctx->device));
        *tile_mult = rg::batch_autotune(b, fp32 != 0, reps, times);
        return RG_OK;
    });
}

int rg_batch_set_tile(rg_batch* b, int tile_mult)
{
    return batch_guard([&] {
        RG_CHECK(b && (tile_mult == 0 || tile_mult == 1 || tile_mult == 2 || tile_mult == 4 || tile_mult == 8),
                 RG_VALUE_ERROR, "batch_set_tile: tile multiplier must be 0 (autotune), 1, 2, 4 or 8");
        b->tile_mult = tile_mult;
        return RG_OK;
    });
}

int rg_batch_solve(rg_batch* b, double* x, const double* rhs, const rg_options* opt, rg_diagnostics* diag)
{
    return batch_guard([&] {
        RG_CHECK(b && x && rhs && opt && diag, RG_VALUE_ERROR, "batch_solve: NULL argument");
        RG_HIP(hipSetDevice(b->ctx->device));
        return rg::batch_solve(b, x, rhs, *opt, diag);
    });
}

}  // extern "C"
