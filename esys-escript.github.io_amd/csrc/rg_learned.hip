// Kernels of the learned preconditioner: a 3-level 3-D convolutional U-Net
// (16 -> 32 -> 64 channels, 3x3x3 kernels) that maps the Jacobi-scaled
// residual and log10(kappa) on the node grid to an error correction, used
// inside flexible CG.  Convolutions are implicit GEMMs on the BF16 matrix
// cores (v_mfma_f32_16x16x32_bf16, FP32 accumulate).
//
// Activation layout (HBM): channels-last BF16 on a zero-halo padded grid,
//   act[((z+1)*P1 + (y+1))*P0 + (x+1)][c],  P_d = D_d + 2,
// so every 3x3x3 tap of an interior voxel is in bounds and reads zeros
// outside the domain.  Weights: BF16 W[co][tap][ci] (the GEMM K index
// tap*Cin + ci is contiguous per output channel), tap = (dx+1) + 3(dy+1) +
// 9(dz+1).  All channel counts are multiples of 16.
//
// MFMA 16x16x32 operand mapping (lane l, q = l >> 4, n = l & 15):
//   A (16 x 32): row n, k = 8q .. 8q+7      B (32 x 16): col n, k = 8q .. 8q+7
//   D (16 x 16): col n, rows 4q .. 4q+3
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "rg_internal.h"

namespace rg {
namespace lpc {

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct Grid3 {
    int d0, d1, d2;      // interior voxels per axis
    int p0, p1;          // padded extents of axes 0 and 1
    int64_t nvox;        // interior voxel count
};

__device__ __forceinline__ int64_t padded_index(const Grid3& g, int64_t v)
{
    const int64_t x = v % g.d0, y = (v / g.d0) % g.d1, z = v / ((int64_t)g.d0 * g.d1);
    return ((z + 1) * g.p1 + (y + 1)) * g.p0 + (x + 1);
}

__device__ __forceinline__ int64_t tap_offset(const Grid3& g, int tap)
{
    const int dx = tap % 3 - 1, dy = (tap / 3) % 3 - 1, dz = tap / 9 - 1;
    return ((int64_t)dz * g.p1 + dy) * g.p0 + dx;
}

// ------------------------------------------------------------- conv (fwd)
// out[v][co] = act( bias[co] + sum_{tap,ci} in[v+off(tap)][ci] W[co][tap][ci] )
// A workgroup computes a 16 x 4 x 4 brick of output voxels: it stages the
// 18 x 6 x 6 input brick (all CIN channels) in LDS once, then each of its 4
// waves multiplies 4 rows of 16 voxels (M tiles) by all COUT/16 channel
// tiles, reading the A operand (16 voxels x 8 channels of one tap per lane
// group) from LDS and the weights (B) from global memory (shared by every
// workgroup, cache resident).
constexpr int kTX = 16, kTY = 4, kTZ = 4;                 // output brick
constexpr int kHX = kTX + 2, kHY = kTY + 2, kHZ = kTZ + 2;  // input brick with halo
constexpr int kHalo = kHX * kHY * kHZ;

struct ConvTiles {
    int nbx, nby, nbz;
};
inline ConvTiles conv_tiles(const Grid3& g)
{
    return {(g.d0 + kTX - 1) / kTX, (g.d1 + kTY - 1) / kTY, (g.d2 + kTZ - 1) / kTZ};
}

// Where a convolution's input comes from (SRC): the input tensor itself
// (kSrcPlain), or formed while the input brick is staged -- the decoder's
// concatenation of the nearest-neighbour x2 upsampled coarse tensor `in`
// (channels [0, CUP), grid gs) with the skip tensor `in2` (channels [CUP,
// CIN), the output grid) (kSrcCatUp), the encoder's 2x2x2 average pooling
// of the fine tensor `in` (grid gs) (kSrcPool), or the network input from
// the solver's vectors on the n0 x n1 x n2 node grid: channel 0 = scale *
// bal * dinv * r, channel 1 = logk, the others 0 (kSrcInput).  The staged
// values are the ones k_up_fwd / k_copy_channels / k_pool_fwd /
// k_make_input would have written.
enum { kSrcPlain = 0, kSrcCatUp = 1, kSrcPool = 2, kSrcInput = 3 };
struct ConvSrc {
    const bf16* in;
    const bf16* in2;
    Grid3 gs;
    const double* r = nullptr;      // kSrcInput
    const double* dinv = nullptr;
    const double* bal = nullptr;
    const float* logk = nullptr;
    const double* scale = nullptr;
    int n0 = 0, n1 = 0, n2 = 0;
};

// With ZOUT the convolution also hands its output channel 0 back to the
// solver, z[i] = y0 / (scale bal[i]) on the node grid (what k_take_output
// computes from the stored output).
struct ConvZOut {
    double* z = nullptr;
    const double* scale = nullptr;
    const double* bal = nullptr;
    int n0 = 0, n1 = 0, n2 = 0;
};

// On small grids (few output bricks) the COUT/16 channel tiles are split
// over blockIdx.y, NT tiles per workgroup, so that more workgroups share the
// work (a 9^3 grid has 9 bricks: with one 16-channel tile each, 4x the
// workgroups for COUT = 64).
template <int CIN, int COUT, bool RELU, int NT, int SRC = kSrcPlain, int CUP = 0, bool ZOUT = false>
__global__ __launch_bounds__(256) void k_conv_fwd(Grid3 g, ConvTiles T, const ConvSrc src,
                                                  const bf16* __restrict__ w, const float* __restrict__ bias,
                                                  bf16* __restrict__ out, const ConvZOut zo)
{
    const bf16* __restrict__ in = src.in;
    static_assert(COUT % (16 * NT) == 0, "channel tiles must divide COUT");
    const int jt0 = blockIdx.y * NT;        // first 16-channel tile of this workgroup
    constexpr int C8 = CIN / 8;             // 16-byte channel groups per voxel
    constexpr int G8 = 27 * C8;             // groups of 8 along K = 27 * CIN
    constexpr int STEPS = (G8 + 3) / 4;
    __shared__ bf16x8 tile[kHalo * C8];
    const int bx = blockIdx.x % T.nbx, by = (blockIdx.x / T.nbx) % T.nby, bz = blockIdx.x / (T.nbx * T.nby);
    const int x0 = bx * kTX, y0 = by * kTY, z0 = bz * kTZ;
    // stage the input brick: padded coordinates (x0 + hx, y0 + hy, z0 + hz);
    // the loads of SB passes are issued before their stores, so a
    // workgroup waits for one load latency per SB passes (up to 31 passes of
    // 256 x 16 bytes for CIN = 96; on the small grids of the solver, one
    // workgroup per CU, that chain was most of the kernel's time)
    const int p2 = g.d2 + 2;
    constexpr int NE = kHalo * C8;
    constexpr int NPASS = (NE + 255) / 256;
    constexpr int SB = SRC == kSrcPool ? 2 : 8;   // a pooled entry is 8 loads
    const Grid3& gs = src.gs;
    auto stage = [&](int pass0, auto UC) {
        constexpr int UU = decltype(UC)::value;
        bf16x8 v[UU];
#pragma unroll
        for (int u = 0; u < UU; ++u) {
            const int e = (int)threadIdx.x + 256 * (pass0 + u);
            const int v8 = e / C8, c8 = e % C8;
            const int hx = v8 % kHX, hy = (v8 / kHX) % kHY, hz = v8 / (kHX * kHY);
            const int px = x0 + hx, py = y0 + hy, pz = z0 + hz;
            v[u] = bf16x8{};
            if constexpr (SRC == kSrcPlain) {
                if (e < NE && px < g.p0 && py < g.p1 && pz < p2)
                    v[u] = *reinterpret_cast<const bf16x8*>(in + (((int64_t)pz * g.p1 + py) * g.p0 + px) * CIN + c8 * 8);
            } else {
                // interior voxels only (the padding is zero)
                if (e < NE && px >= 1 && px <= g.d0 && py >= 1 && py <= g.d1 && pz >= 1 && pz <= g.d2) {
                    const int x = px - 1, y = py - 1, z = pz - 1;
                    if constexpr (SRC == kSrcInput) {
                        if (c8 == 0 && x < src.n0 && y < src.n1 && z < src.n2) {
                            const int64_t i = x + (int64_t)src.n0 * (y + (int64_t)src.n1 * z);
                            const double sc = src.scale ? src.scale[0] : 1.0;
                            v[u][0] = (bf16)(float)(sc * (src.bal ? src.bal[i] : 1.0) *
                                                    (src.dinv ? src.dinv[i] : 1.0) * src.r[i]);
                            v[u][1] = (bf16)(src.logk ? src.logk[i] : 0.f);
                        }
                    } else if constexpr (SRC == kSrcCatUp) {
                        if (c8 * 8 < CUP) {
                            const int64_t pc = (((int64_t)(z / 2) + 1) * gs.p1 + (y / 2) + 1) * gs.p0 + (x / 2) + 1;
                            v[u] = *reinterpret_cast<const bf16x8*>(in + pc * CUP + c8 * 8);
                        } else {
                            const int64_t pf = ((int64_t)pz * g.p1 + py) * g.p0 + px;
                            v[u] = *reinterpret_cast<const bf16x8*>(src.in2 + pf * (CIN - CUP) + c8 * 8 - CUP);
                        }
                    } else {
                        float sum[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
                        for (int k = 0; k < 8; ++k) {
                            const int64_t pf = (((int64_t)(2 * z + (k >> 2)) + 1) * gs.p1 + (2 * y + ((k >> 1) & 1)) + 1) *
                                                   gs.p0 +
                                               (2 * x + (k & 1)) + 1;
                            const bf16x8 t = *reinterpret_cast<const bf16x8*>(in + pf * CIN + c8 * 8);
#pragma unroll
                            for (int j = 0; j < 8; ++j) sum[j] += (float)t[j];
                        }
#pragma unroll
                        for (int j = 0; j < 8; ++j) v[u][j] = (bf16)(0.125f * sum[j]);
                    }
                }
            }
        }
#pragma unroll
        for (int u = 0; u < UU; ++u) {
            const int e = (int)threadIdx.x + 256 * (pass0 + u);
            if (e < NE) tile[e] = v[u];
        }
    };
    {
        int pass = 0;
        for (; pass + SB <= NPASS; pass += SB) stage(pass, std::integral_constant<int, SB>{});
        if constexpr (NPASS % SB != 0) stage(pass, std::integral_constant<int, NPASS % SB>{});
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, n = lane & 15;
    f32x4 acc[4][NT];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int j = 0; j < NT; ++j) acc[m][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    // K steps in groups of U: the U weight loads of a group (global memory,
    // L2 resident) are issued together before its MFMAs, so one load latency
    // is paid per group rather than per step, and the A operands (LDS) are
    // read one step ahead of their MFMAs -- on the small grids of the coarse
    // levels a workgroup's K loop is the whole kernel, and these latency
    // chains set its length.  (Loading each group's weights one group ahead
    // as well measured no faster.)
    constexpr int U = NT <= 2 ? 8 : 4;
    auto steps = [&](int s0, auto UC) {
        constexpr int UU = decltype(UC)::value;
        bf16x8 b[UU][NT];
        int tapv[UU], cigv[UU];
        bool gokv[UU];
#pragma unroll
        for (int u = 0; u < UU; ++u) {
            const int gi = 4 * (s0 + u) + q;
            gokv[u] = gi < G8;
            const int gg = gokv[u] ? gi : 0;
            tapv[u] = gg / C8;
            cigv[u] = gg % C8;
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                b[u][j] = *reinterpret_cast<const bf16x8*>(w + ((int64_t)(16 * (jt0 + j) + n) * 27 + tapv[u]) * CIN +
                                                           cigv[u] * 8);
                if (!gokv[u]) b[u][j] = bf16x8{};
            }
        }
        auto load_a = [&](int u, bf16x8 (&a)[4]) {
            const int tap = tapv[u];
            const int dx = tap % 3, dy = (tap / 3) % 3, dz = tap / 9;   // offsets + 1
#pragma unroll
            for (int m = 0; m < 4; ++m) {
                const int row = 4 * wave + m;                 // (ty, tz) row of the brick
                const int ty = row % kTY, tz = row / kTY;
                a[m] = tile[(((tz + dz) * kHY + (ty + dy)) * kHX + (n + dx)) * C8 + cigv[u]];
                if (!gokv[u]) a[m] = bf16x8{};
            }
        };
        bf16x8 a0[4], a1[4];
        load_a(0, a0);
        // scheduling order of the group (the compiler otherwise kept one LDS
        // read in flight): its UU * NT weight loads (VMEM reads, mask
        // 0x020), the first step's 4 LDS reads (0x100), then per step the
        // next step's LDS reads and this step's MFMAs (0x008)
        __builtin_amdgcn_sched_group_barrier(0x020, UU * NT, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
#pragma unroll
        for (int u = 0; u < UU; ++u) {
            bf16x8 (&cur)[4] = (u & 1) ? a1 : a0;
            bf16x8 (&nxt)[4] = (u & 1) ? a0 : a1;
            if (u + 1 < UU) load_a(u + 1, nxt);
#pragma unroll
            for (int m = 0; m < 4; ++m)
#pragma unroll
                for (int j = 0; j < NT; ++j)
                    acc[m][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[m], b[u][j], acc[m][j], 0, 0, 0);
            if (u + 1 < UU) __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);
            __builtin_amdgcn_sched_group_barrier(0x008, 4 * NT, 0);
        }
    };
    int s = 0;
    for (; s + U <= STEPS; s += U) steps(s, std::integral_constant<int, U>{});
    if constexpr (STEPS % U != 0) steps(s, std::integral_constant<int, STEPS % U>{});
    // D rows 4q..4q+3 are the voxels x0 + 4q + r of the row, column n the channel
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        const int row = 4 * wave + m;
        const int y = y0 + row % kTY, z = z0 + row / kTY;
        if (y >= g.d1 || z >= g.d2) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int x = x0 + 4 * q + r;
            if (x >= g.d0) continue;
            const int64_t p = (((int64_t)z + 1) * g.p1 + (y + 1)) * g.p0 + (x + 1);
#pragma unroll
            for (int j = 0; j < NT; ++j) {
                const int co = 16 * (jt0 + j) + n;
                float o = acc[m][j][r] + (bias ? bias[co] : 0.f);
                if (RELU) o = fmaxf(o, 0.f);
                out[p * COUT + co] = (bf16)o;
                if constexpr (ZOUT) {
                    if (co == 0 && x < zo.n0 && y < zo.n1 && z < zo.n2) {
                        const int64_t i = x + (int64_t)zo.n0 * (y + (int64_t)zo.n1 * z);
                        const double inv = zo.scale ? 1.0 / zo.scale[0] : 1.0;
                        zo.z[i] = (double)(float)(bf16)o * inv / (zo.bal ? zo.bal[i] : 1.0);
                    }
                }
            }
        }
    }
}

// ------------------------------------------------- conv (weight gradient)
// dW[co][tap][ci] = sum_v in[v+off(tap)][ci] dy[v][co], accumulated over a
// voxel chunk per block (blockIdx.y = tap) and written as FP32 partials
// part[blockIdx.x][tap][co][ci] (summed in fixed order by k_sum_partials).
// Each step stages 128 voxels of in (shifted by the tap) and dy, transposed
// to channel-major, in LDS; each wave then multiplies its 32-voxel slice.
template <int CIN, int COUT>
__global__ __launch_bounds__(256) void k_conv_wgrad(Grid3 g, const bf16* __restrict__ in, const bf16* __restrict__ dy,
                                                    float* __restrict__ part, int64_t vox_per_block)
{
    constexpr int MT = COUT / 16, NTI = CIN / 16;
    __shared__ bf16 xs[CIN][128 + 8];
    __shared__ bf16 ys[COUT][128 + 8];
    const int tap = blockIdx.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, n = lane & 15;
    const Grid3 gg = g;
    const int64_t off = tap_offset(gg, tap);
    const int64_t v0 = (int64_t)blockIdx.x * vox_per_block;
    int64_t v1 = v0 + vox_per_block;
    if (v1 > g.nvox) v1 = g.nvox;
    f32x4 acc[MT][NTI];
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int j = 0; j < NTI; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (int64_t vs = v0; vs < v1; vs += 128) {
        __syncthreads();
        // stage: element e of (128 voxels x C) handled by thread t
        for (int e = threadIdx.x; e < 128 * CIN; e += 256) {
            const int vv = e / CIN, c = e % CIN;
            const int64_t v = vs + vv;
            bf16 val = (bf16)0.f;
            if (v < v1) val = in[(padded_index(gg, v) + off) * CIN + c];
            xs[c][vv] = val;
        }
        for (int e = threadIdx.x; e < 128 * COUT; e += 256) {
            const int vv = e / COUT, c = e % COUT;
            const int64_t v = vs + vv;
            bf16 val = (bf16)0.f;
            if (v < v1) val = dy[padded_index(gg, v) * COUT + c];
            ys[c][vv] = val;
        }
        __syncthreads();
        const int kv = 32 * wave + 8 * q;
#pragma unroll
        for (int i = 0; i < MT; ++i) {
            bf16x8 a;
#pragma unroll
            for (int j = 0; j < 8; ++j) a[j] = ys[16 * i + n][kv + j];
#pragma unroll
            for (int jn = 0; jn < NTI; ++jn) {
                bf16x8 b;
#pragma unroll
                for (int j = 0; j < 8; ++j) b[j] = xs[16 * jn + n][kv + j];
                acc[i][jn] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc[i][jn], 0, 0, 0);
            }
        }
    }
    // D tile (i, jn): rows co = 16i + 4q + r, column ci = 16 jn + n; the four
    // waves' partials are written separately
    float* dst = part + (((int64_t)blockIdx.x * 4 + wave) * 27 + tap) * COUT * CIN;
#pragma unroll
    for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int jn = 0; jn < NTI; ++jn)
#pragma unroll
            for (int r = 0; r < 4; ++r) dst[(16 * i + 4 * q + r) * CIN + 16 * jn + n] = acc[i][jn][r];
}

// dW[co][tap][ci] = sum over np partials of part[p][tap][co][ci] (fixed order),
// and db[co] from the bias partials when given.
__global__ void k_sum_wgrad(const float* __restrict__ part, int np, int cout, int cin, float* __restrict__ dw)
{
    const int64_t total = (int64_t)27 * cout * cin;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int ci = e % cin, co = (e / cin) % cout, tap = e / ((int64_t)cin * cout);
        float s = 0.f;
        for (int p = 0; p < np; ++p) s += part[(((int64_t)p * 27 + tap) * cout + co) * cin + ci];
        dw[((int64_t)co * 27 + tap) * cin + ci] = s;
    }
}

// db[co] = sum_v dy[v][co]: per-block partials (fixed order), then summed.
__global__ void k_bias_grad_part(Grid3 g, const bf16* __restrict__ dy, int cout, float* __restrict__ part,
                                 int64_t vox_per_block)
{
    __shared__ float red[256];
    const int64_t v0 = (int64_t)blockIdx.x * vox_per_block;
    int64_t v1 = v0 + vox_per_block;
    if (v1 > g.nvox) v1 = g.nvox;
    for (int co = 0; co < cout; ++co) {
        float s = 0.f;
        for (int64_t v = v0 + threadIdx.x; v < v1; v += 256) s += (float)dy[padded_index(g, v) * cout + co];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int o = 128; o > 0; o >>= 1) {
            if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
            __syncthreads();
        }
        if (threadIdx.x == 0) part[(int64_t)blockIdx.x * cout + co] = red[0];
        __syncthreads();
    }
}

__global__ void k_sum_rows(const float* __restrict__ part, int np, int n, float* __restrict__ out)
{
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
        float s = 0.f;
        for (int p = 0; p < np; ++p) s += part[(int64_t)p * n + e];
        out[e] = s;
    }
}

// W^T with flipped taps for the data gradient: wt[ci][tap][co] = w[co][26-tap][ci].
__global__ void k_flip_weights(const bf16* __restrict__ w, int cout, int cin, bf16* __restrict__ wt)
{
    const int64_t total = (int64_t)27 * cout * cin;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int co = e % cout, tap = (e / cout) % 27, ci = e / ((int64_t)cout * 27);
        wt[((int64_t)ci * 27 + tap) * cout + co] = w[((int64_t)co * 27 + (26 - tap)) * cin + ci];
    }
}

// ------------------------------------------------------- elementwise ops
// All channel counts and offsets are multiples of 8: each thread moves one
// 16-byte group of 8 channels of one voxel (element e = voxel * c/8 + group).
__device__ __forceinline__ bf16x8 ld8(const bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ __forceinline__ void st8(bf16* p, const bf16x8& v) { *reinterpret_cast<bf16x8*>(p) = v; }

__device__ __forceinline__ int64_t fine_padded(const Grid3& gf, int64_t x, int64_t y, int64_t z)
{
    return ((z + 1) * gf.p1 + (y + 1)) * gf.p0 + (x + 1);
}

// dpre = dout * (act > 0)  (ReLU backward), interior voxels only
__global__ void k_relu_bwd(Grid3 g, int c, const bf16* __restrict__ dout, const bf16* __restrict__ act,
                           bf16* __restrict__ dpre)
{
    const int c8 = c / 8;
    const int64_t total = g.nvox * c8;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = padded_index(g, e / c8) * c + (e % c8) * 8;
        const bf16x8 a = ld8(act + p), d = ld8(dout + p);
        bf16x8 o;
#pragma unroll
        for (int k = 0; k < 8; ++k) o[k] = ((float)a[k] > 0.f) ? d[k] : (bf16)0.f;
        st8(dpre + p, o);
    }
}

// 2x2x2 average pooling (fine grid gf -> coarse grid gc, gc.d = gf.d / 2)
__global__ void k_pool_fwd(Grid3 gf, Grid3 gc, int c, const bf16* __restrict__ in, bf16* __restrict__ out)
{
    const int c8 = c / 8;
    const int64_t total = gc.nvox * c8;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t vc = e / c8;
        const int ch = (int)(e % c8) * 8;
        const int64_t X = vc % gc.d0, Y = (vc / gc.d0) % gc.d1, Z = vc / ((int64_t)gc.d0 * gc.d1);
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const bf16x8 v = ld8(in + fine_padded(gf, 2 * X + (k & 1), 2 * Y + ((k >> 1) & 1), 2 * Z + (k >> 2)) * c + ch);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += (float)v[j];
        }
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = (bf16)(0.125f * s[j]);
        st8(out + padded_index(gc, vc) * c + ch, o);
    }
}

__global__ void k_pool_bwd(Grid3 gf, Grid3 gc, int c, const bf16* __restrict__ dout, bf16* __restrict__ din)
{
    const int c8 = c / 8;
    const int64_t total = gf.nvox * c8;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = e / c8;
        const int ch = (int)(e % c8) * 8;
        const int64_t x = v % gf.d0, y = (v / gf.d0) % gf.d1, z = v / ((int64_t)gf.d0 * gf.d1);
        const bf16x8 d = ld8(dout + ((((z / 2) + 1) * gc.p1 + (y / 2) + 1) * gc.p0 + (x / 2) + 1) * c + ch);
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = (bf16)(0.125f * (float)d[j]);
        st8(din + padded_index(gf, v) * c + ch, o);
    }
}

// nearest-neighbour x2 upsampling of channels [0, c) of `in` into channels
// [coff, coff + c) of `out` (which has cout channels): the decoder's
// upsample + concat in one pass
__global__ void k_up_fwd(Grid3 gf, Grid3 gc, int c, const bf16* __restrict__ in, int cout, int coff,
                         bf16* __restrict__ out)
{
    const int c8 = c / 8;
    const int64_t total = gf.nvox * c8;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t v = e / c8;
        const int ch = (int)(e % c8) * 8;
        const int64_t x = v % gf.d0, y = (v / gf.d0) % gf.d1, z = v / ((int64_t)gf.d0 * gf.d1);
        const int64_t pc = (((z / 2) + 1) * gc.p1 + (y / 2) + 1) * gc.p0 + (x / 2) + 1;
        st8(out + padded_index(gf, v) * cout + coff + ch, ld8(in + pc * c + ch));
    }
}

__global__ void k_up_bwd(Grid3 gf, Grid3 gc, int c, const bf16* __restrict__ dout, int cout, int coff,
                         bf16* __restrict__ din)
{
    const int c8 = c / 8;
    const int64_t total = gc.nvox * c8;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t vc = e / c8;
        const int ch = (int)(e % c8) * 8;
        const int64_t X = vc % gc.d0, Y = (vc / gc.d0) % gc.d1, Z = vc / ((int64_t)gc.d0 * gc.d1);
        float s[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const bf16x8 v =
                ld8(dout + fine_padded(gf, 2 * X + (k & 1), 2 * Y + ((k >> 1) & 1), 2 * Z + (k >> 2)) * cout + coff + ch);
#pragma unroll
            for (int j = 0; j < 8; ++j) s[j] += (float)v[j];
        }
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = (bf16)s[j];
        st8(din + padded_index(gc, vc) * c + ch, o);
    }
}

// channel slice copy: out[v][ooff + k] = in[v][ioff + k], k < c (interior voxels)
__global__ void k_copy_channels(Grid3 g, int c, const bf16* __restrict__ in, int cin, int ioff,
                                bf16* __restrict__ out, int cout, int ooff)
{
    const int c8 = c / 8;
    const int64_t total = g.nvox * c8;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = padded_index(g, e / c8);
        const int ch = (int)(e % c8) * 8;
        st8(out + p * cout + ooff + ch, ld8(in + p * cin + ioff + ch));
    }
}

// dst[v][ooff + k] += src[v][ioff + k]  (sum of two gradient paths)
__global__ void k_add_channels(Grid3 g, int c, const bf16* __restrict__ src, int cs, int ioff,
                               bf16* __restrict__ dst, int cd, int ooff)
{
    const int c8 = c / 8;
    const int64_t total = g.nvox * c8;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t p = padded_index(g, e / c8);
        const int ch = (int)(e % c8) * 8;
        const bf16x8 a = ld8(dst + p * cd + ooff + ch), b = ld8(src + p * cs + ioff + ch);
        bf16x8 o;
#pragma unroll
        for (int j = 0; j < 8; ++j) o[j] = (bf16)((float)a[j] + (float)b[j]);
        st8(dst + p * cd + ooff + ch, o);
    }
}

// Network input from the solver's vectors on the node grid (nodes = voxels
// of the interior, node grid n0 x n1 x n2 inside a grid padded to multiples
// of 4): channel 0 = scale * dinv[i] * r[i], channel 1 = logk[i]; others 0.
__global__ void k_make_input(Grid3 g, int n0, int n1, int n2, const double* __restrict__ r,
                             const double* __restrict__ dinv, const double* __restrict__ bal,
                             const float* __restrict__ logk, const double* __restrict__ scale, int cin,
                             bf16* __restrict__ out)
{
    const double sc = scale ? scale[0] : 1.0;
    for (int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < g.nvox; v += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = v % g.d0, y = (v / g.d0) % g.d1, z = v / ((int64_t)g.d0 * g.d1);
        float u = 0.f, lk = 0.f;
        if (x < n0 && y < n1 && z < n2) {
            const int64_t i = x + (int64_t)n0 * (y + (int64_t)n1 * z);
            u = (float)(sc * (bal ? bal[i] : 1.0) * (dinv ? dinv[i] : 1.0) * r[i]);
            lk = logk ? logk[i] : 0.f;
        }
        const int64_t p = padded_index(g, v) * cin;
        bf16x8 first = bf16x8{};
        first[0] = (bf16)u;
        first[1] = (bf16)lk;
        st8(out + p, first);
        for (int ch = 8; ch < cin; ch += 8) st8(out + p + ch, bf16x8{});
    }
}

// z[i] = out[v][0] / (scale * bal[i]) for the nodes of the grid
__global__ void k_take_output(Grid3 g, int n0, int n1, int n2, const bf16* __restrict__ out, int cout,
                              const double* __restrict__ scale, const double* __restrict__ bal,
                              double* __restrict__ z)
{
    const double inv = scale ? 1.0 / scale[0] : 1.0;
    const int64_t nn = (int64_t)n0 * n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = i % n0, y = (i / n0) % n1, z3 = i / ((int64_t)n0 * n1);
        const int64_t p = (((z3 + 1) * g.p1 + (y + 1)) * g.p0 + (x + 1)) * cout;
        z[i] = (double)(float)out[p] * inv / (bal ? bal[i] : 1.0);
    }
}

}  // namespace lpc
}  // namespace rg

// ===================================================================== host
#include <cmath>
#include <cstring>
#include <vector>

namespace rg {
namespace lpc {

// Layers of the network: (cin, cout, relu).  Layer 0 reads 2 real input
// channels padded to 16; layer 8 produces 16 channels of which channel 0 is
// the output.
struct LayerSpec {
    int cin, cout;
    bool relu;
};
static const LayerSpec kLayers[9] = {
    {16, 16, true},  {16, 16, true},            // level 1
    {16, 32, true},  {32, 32, true},            // level 2
    {32, 64, true},  {64, 64, true},            // level 3
    {96, 32, true},                             // level 2 decoder (up(64) + skip 32)
    {48, 16, true},                             // level 1 decoder (up(32) + skip 16)
    {16, 16, false},                            // output
};
constexpr int kNumLayers = 9;

inline int64_t layer_wsize(int l) { return (int64_t)27 * kLayers[l].cin * kLayers[l].cout; }
inline int64_t total_params()
{
    int64_t s = 0;
    for (int l = 0; l < kNumLayers; ++l) s += layer_wsize(l) + kLayers[l].cout;
    return s;
}

inline Grid3 make_grid(int d0, int d1, int d2)
{
    Grid3 g;
    g.d0 = d0; g.d1 = d1; g.d2 = d2;
    g.p0 = d0 + 2; g.p1 = d1 + 2;
    g.nvox = (int64_t)d0 * d1 * d2;
    return g;
}
inline int64_t padded_size(const Grid3& g) { return (int64_t)(g.d0 + 2) * (g.d1 + 2) * (g.d2 + 2); }

template <int CIN, int COUT, int SRC = kSrcPlain, int CUP = 0, bool ZOUT = false>
void conv_fwd_t(hipStream_t st, const Grid3& g, const ConvSrc& in, const bf16* w, const float* b, bf16* out, bool relu,
                const ConvZOut& zo = {})
{
    const ConvTiles T = conv_tiles(g);
    const int grid = T.nbx * T.nby * T.nbz;
    constexpr int NT = COUT / 16;
    // fewer bricks than half the CUs: one channel tile per workgroup
    if (grid < 128) {
        if (relu)
            hipLaunchKernelGGL((k_conv_fwd<CIN, COUT, true, 1, SRC, CUP, ZOUT>), dim3(grid, NT), dim3(256), 0, st, g, T, in, w, b, out, zo);
        else
            hipLaunchKernelGGL((k_conv_fwd<CIN, COUT, false, 1, SRC, CUP, ZOUT>), dim3(grid, NT), dim3(256), 0, st, g, T, in, w, b, out, zo);
    } else {
        if (relu)
            hipLaunchKernelGGL((k_conv_fwd<CIN, COUT, true, NT, SRC, CUP, ZOUT>), dim3(grid), dim3(256), 0, st, g, T, in, w, b, out, zo);
        else
            hipLaunchKernelGGL((k_conv_fwd<CIN, COUT, false, NT, SRC, CUP, ZOUT>), dim3(grid), dim3(256), 0, st, g, T, in, w, b, out, zo);
    }
}

void conv_fwd(hipStream_t st, const Grid3& g, int cin, int cout, const bf16* in, const bf16* w, const float* b,
              bf16* out, bool relu)
{
#define RG_CONV(CI, CO) \
    if (cin == CI && cout == CO) return conv_fwd_t<CI, CO>(st, g, ConvSrc{in, nullptr, g}, w, b, out, relu);
    RG_CONV(16, 16) RG_CONV(16, 32) RG_CONV(32, 32) RG_CONV(32, 64) RG_CONV(64, 64) RG_CONV(96, 32)
    RG_CONV(48, 16) RG_CONV(32, 16) RG_CONV(64, 32) RG_CONV(32, 96) RG_CONV(16, 48)
#undef RG_CONV
    throw Error(RG_VALUE_ERROR, "conv: unsupported channel pair");
}

template <int CIN, int COUT>
void conv_wgrad_t(hipStream_t st, const Grid3& g, const bf16* x, const bf16* dy, float* part, int nchunk)
{
    const int64_t vpb = ((g.nvox + nchunk - 1) / nchunk + 127) / 128 * 128;
    hipLaunchKernelGGL((k_conv_wgrad<CIN, COUT>), dim3(nchunk, 27), dim3(256), 0, st, g, x, dy, part, vpb);
}

void conv_wgrad(hipStream_t st, const Grid3& g, int cin, int cout, const bf16* x, const bf16* dy, float* part,
                int nchunk)
{
#define RG_WG(CI, CO) \
    if (cin == CI && cout == CO) return conv_wgrad_t<CI, CO>(st, g, x, dy, part, nchunk);
    RG_WG(16, 16) RG_WG(16, 32) RG_WG(32, 32) RG_WG(32, 64) RG_WG(64, 64) RG_WG(96, 32) RG_WG(48, 16)
#undef RG_WG
    throw Error(RG_VALUE_ERROR, "conv wgrad: unsupported channel pair");
}

inline int ew_grid(int64_t n) { return (int)std::min<int64_t>((n + 255) / 256, 8192); }

struct Tensor {
    bf16* p = nullptr;
    int c = 0;
};

struct Net {
    rg_context* ctx = nullptr;
    int n[3] = {0, 0, 0};          // node grid of the problem
    Grid3 g1, g2, g3;              // levels (interior padded to multiples of 4)
    std::vector<float*> W32, B32;  // FP32 master parameters (device)
    std::vector<bf16*> W16;        // BF16 weights used by the convolutions
    std::vector<bf16*> WT16;       // flipped/transposed weights for data gradients
    std::vector<float*> gW, gB;    // gradients
    std::vector<float*> mW, vW, mB, vB;   // Adam moments
    int adam_t = 0;
    // activations (forward) and their gradients (backward)
    Tensor x0, a1, s1, p1, a2, s2, p2, a3, a3b, cat2, a4, cat1, a5, y;
    Tensor gy, ga5, gcat1, gs1, gs1b, ga4, gcat2, gs2, gs2b, ga3b, ga3, gp2, ga2, gp1, ga1, gtmp1;
    float* part = nullptr;         // weight-gradient partials
    int64_t part_elems = 0;
    int nchunk = 64;
    double* scale = nullptr;       // device scalar: input normalisation
    double* red = nullptr;         // reduction scratch (partials + results)
    rg_matrix* loss_op = nullptr;  // energy-norm loss operator (borrowed; null: cosine loss)
    double* zbuf = nullptr;        // energy loss: the output z on the nodes and A z
    double* azbuf = nullptr;
    std::vector<void*> allocs;

    void* alloc(size_t bytes)
    {
        void* p = nullptr;
        RG_HIP(hipMalloc(&p, bytes));
        RG_HIP(hipMemsetAsync(p, 0, bytes, ctx->stream));
        allocs.push_back(p);
        return p;
    }
    Tensor tensor(const Grid3& g, int c)
    {
        Tensor t;
        t.c = c;
        t.p = (bf16*)alloc(sizeof(bf16) * padded_size(g) * c);
        return t;
    }
    ~Net()
    {
        for (void* p : allocs) hipFree(p);
    }
};

void net_init(Net& N, rg_context* ctx, const int32_t* nodes, bool training)
{
    N.ctx = ctx;
    for (int d = 0; d < 3; ++d) N.n[d] = nodes[d];
    auto up4 = [](int v) { return ((v + 3) / 4) * 4; };
    N.g1 = make_grid(up4(nodes[0]), up4(nodes[1]), up4(nodes[2]));
    N.g2 = make_grid(N.g1.d0 / 2, N.g1.d1 / 2, N.g1.d2 / 2);
    N.g3 = make_grid(N.g2.d0 / 2, N.g2.d1 / 2, N.g2.d2 / 2);
    for (int l = 0; l < kNumLayers; ++l) {
        const int64_t ws = layer_wsize(l);
        N.W32.push_back((float*)N.alloc(sizeof(float) * ws));
        N.B32.push_back((float*)N.alloc(sizeof(float) * kLayers[l].cout));
        N.W16.push_back((bf16*)N.alloc(sizeof(bf16) * ws));
        if (training) {
            N.WT16.push_back((bf16*)N.alloc(sizeof(bf16) * ws));
            N.gW.push_back((float*)N.alloc(sizeof(float) * ws));
            N.gB.push_back((float*)N.alloc(sizeof(float) * kLayers[l].cout));
            N.mW.push_back((float*)N.alloc(sizeof(float) * ws));
            N.vW.push_back((float*)N.alloc(sizeof(float) * ws));
            N.mB.push_back((float*)N.alloc(sizeof(float) * kLayers[l].cout));
            N.vB.push_back((float*)N.alloc(sizeof(float) * kLayers[l].cout));
        }
    }
    const Grid3 &g1 = N.g1, &g2 = N.g2, &g3 = N.g3;
    N.x0 = N.tensor(g1, 16); N.a1 = N.tensor(g1, 16); N.s1 = N.tensor(g1, 16);
    N.p1 = N.tensor(g2, 16); N.a2 = N.tensor(g2, 32); N.s2 = N.tensor(g2, 32);
    N.p2 = N.tensor(g3, 32); N.a3 = N.tensor(g3, 64); N.a3b = N.tensor(g3, 64);
    N.cat2 = N.tensor(g2, 96); N.a4 = N.tensor(g2, 32);
    N.cat1 = N.tensor(g1, 48); N.a5 = N.tensor(g1, 16); N.y = N.tensor(g1, 16);
    if (training) {
        N.gy = N.tensor(g1, 16); N.ga5 = N.tensor(g1, 16); N.gcat1 = N.tensor(g1, 48);
        N.gs1 = N.tensor(g1, 16); N.gs1b = N.tensor(g1, 16); N.ga1 = N.tensor(g1, 16); N.gtmp1 = N.tensor(g1, 16);
        N.ga4 = N.tensor(g2, 32); N.gcat2 = N.tensor(g2, 96); N.gs2 = N.tensor(g2, 32); N.gs2b = N.tensor(g2, 32);
        N.ga2 = N.tensor(g2, 32); N.gp1 = N.tensor(g2, 16);
        N.ga3b = N.tensor(g3, 64); N.ga3 = N.tensor(g3, 64); N.gp2 = N.tensor(g3, 32);
        int64_t mx = 0;
        for (int l = 0; l < kNumLayers; ++l) mx = std::max<int64_t>(mx, layer_wsize(l));
        N.part_elems = (int64_t)N.nchunk * 4 * mx;
        N.part = (float*)N.alloc(sizeof(float) * N.part_elems);
    }
    N.scale = (double*)N.alloc(sizeof(double) * 4);
    N.red = (double*)N.alloc(sizeof(double) * (3 * 1024 + 8));   // partials, then 3 loss dots at 3*1024
    if (training) {
        const int64_t nn = (int64_t)nodes[0] * nodes[1] * nodes[2];
        N.zbuf = (double*)N.alloc(sizeof(double) * nn);
        N.azbuf = (double*)N.alloc(sizeof(double) * nn);
    }
}

__global__ void k_to_bf16(const float* __restrict__ a, bf16* __restrict__ b, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (bf16)a[i];
}

void refresh_bf16(Net& N)
{
    for (int l = 0; l < kNumLayers; ++l)
        hipLaunchKernelGGL(k_to_bf16, dim3(ew_grid(layer_wsize(l))), dim3(256), 0, N.ctx->stream, N.W32[l], N.W16[l],
                           layer_wsize(l));
    RG_HIP(hipGetLastError());
}

// RMS of u = dinv * r over the nodes -> scale = 1 / rms (device scalar).
__global__ void k_rms_part(const double* __restrict__ r, const double* __restrict__ dinv,
                           const double* __restrict__ bal, int64_t n, double* __restrict__ part)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double u = (bal ? bal[i] : 1.0) * (dinv ? dinv[i] : 1.0) * r[i];
        s += u * u;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// The same scale from one workgroup, for the small levels the solver applies
// the network on (one launch instead of two)
__global__ __launch_bounds__(1024) void k_rms_small(const double* __restrict__ r, const double* __restrict__ dinv,
                                                    const double* __restrict__ bal, int64_t n, double* __restrict__ scale)
{
    __shared__ double red[1024];
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += 1024) {
        const double u = (bal ? bal[i] : 1.0) * (dinv ? dinv[i] : 1.0) * r[i];
        s += u * u;
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 512; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double rms = sqrt(red[0] / (double)n);
        scale[0] = rms > 0.0 ? 1.0 / rms : 1.0;
    }
}

__global__ void k_rms_final(const double* __restrict__ part, int np, int64_t n, double* __restrict__ scale)
{
    __shared__ double red[256];
    double s = 0.0;
    for (int b = threadIdx.x; b < np; b += 256) s += part[b];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double rms = sqrt(red[0] / (double)n);
        scale[0] = rms > 0.0 ? 1.0 / rms : 1.0;
    }
}

// Forward pass: y = Net(u / rms(u), logk) with u = bal * dinv * r; z = y[...,0] / (scale bal).
// With bal (the balance vector of a balanced system, A_b = D A D) the network
// sees the unbalanced residual's Jacobi-scaled form and returns the balanced
// correction D^-1 e.  With materialize (training: the backward pass reads
// them) the pooled and concatenated layer inputs are written out; otherwise
// the convolutions that read them form them while staging their input.
void net_forward(Net& N, const double* r, const double* dinv, const double* bal, const float* logk, double* z,
                 hipStream_t st, bool materialize = false)
{
    const int64_t nn = (int64_t)N.n[0] * N.n[1] * N.n[2];
    const Grid3 &g1 = N.g1, &g2 = N.g2, &g3 = N.g3;
    if (!materialize) {
        // the input scale in one launch on small grids, the input formed by
        // the first convolution, z handed back by the last
        if (nn <= (int64_t)1 << 17) {
            hipLaunchKernelGGL(k_rms_small, dim3(1), dim3(1024), 0, st, r, dinv, bal, nn, N.scale);
        } else {
            hipLaunchKernelGGL(k_rms_part, dim3(1024), dim3(256), 0, st, r, dinv, bal, nn, N.red);
            hipLaunchKernelGGL(k_rms_final, dim3(1), dim3(256), 0, st, N.red, 1024, nn, N.scale);
        }
        ConvSrc in0{nullptr, nullptr, g1};
        in0.r = r; in0.dinv = dinv; in0.bal = bal; in0.logk = logk; in0.scale = N.scale;
        in0.n0 = N.n[0]; in0.n1 = N.n[1]; in0.n2 = N.n[2];
        conv_fwd_t<16, 16, kSrcInput>(st, g1, in0, N.W16[0], N.B32[0], N.a1.p, true);
        conv_fwd(st, g1, 16, 16, N.a1.p, N.W16[1], N.B32[1], N.s1.p, true);
        conv_fwd_t<16, 32, kSrcPool>(st, g2, ConvSrc{N.s1.p, nullptr, g1}, N.W16[2], N.B32[2], N.a2.p, true);
        conv_fwd(st, g2, 32, 32, N.a2.p, N.W16[3], N.B32[3], N.s2.p, true);
        conv_fwd_t<32, 64, kSrcPool>(st, g3, ConvSrc{N.s2.p, nullptr, g2}, N.W16[4], N.B32[4], N.a3.p, true);
        conv_fwd(st, g3, 64, 64, N.a3.p, N.W16[5], N.B32[5], N.a3b.p, true);
        conv_fwd_t<96, 32, kSrcCatUp, 64>(st, g2, ConvSrc{N.a3b.p, N.s2.p, g3}, N.W16[6], N.B32[6], N.a4.p, true);
        conv_fwd_t<48, 16, kSrcCatUp, 32>(st, g1, ConvSrc{N.a4.p, N.s1.p, g2}, N.W16[7], N.B32[7], N.a5.p, true);
        if (z) {
            ConvZOut zo;
            zo.z = z; zo.scale = N.scale; zo.bal = bal; zo.n0 = N.n[0]; zo.n1 = N.n[1]; zo.n2 = N.n[2];
            conv_fwd_t<16, 16, kSrcPlain, 0, true>(st, g1, ConvSrc{N.a5.p, nullptr, g1}, N.W16[8], N.B32[8], N.y.p,
                                                   false, zo);
        } else {
            conv_fwd(st, g1, 16, 16, N.a5.p, N.W16[8], N.B32[8], N.y.p, false);
        }
        RG_HIP(hipGetLastError());
        return;
    }
    hipLaunchKernelGGL(k_rms_part, dim3(1024), dim3(256), 0, st, r, dinv, bal, nn, N.red);
    hipLaunchKernelGGL(k_rms_final, dim3(1), dim3(256), 0, st, N.red, 1024, nn, N.scale);
    hipLaunchKernelGGL(k_make_input, dim3(ew_grid(g1.nvox)), dim3(256), 0, st, g1, N.n[0], N.n[1], N.n[2], r, dinv,
                       bal, logk, N.scale, 16, N.x0.p);
    conv_fwd(st, g1, 16, 16, N.x0.p, N.W16[0], N.B32[0], N.a1.p, true);
    conv_fwd(st, g1, 16, 16, N.a1.p, N.W16[1], N.B32[1], N.s1.p, true);
    hipLaunchKernelGGL(k_pool_fwd, dim3(ew_grid(g2.nvox * 2)), dim3(256), 0, st, g1, g2, 16, N.s1.p, N.p1.p);
    conv_fwd(st, g2, 16, 32, N.p1.p, N.W16[2], N.B32[2], N.a2.p, true);
    conv_fwd(st, g2, 32, 32, N.a2.p, N.W16[3], N.B32[3], N.s2.p, true);
    hipLaunchKernelGGL(k_pool_fwd, dim3(ew_grid(g3.nvox * 4)), dim3(256), 0, st, g2, g3, 32, N.s2.p, N.p2.p);
    conv_fwd(st, g3, 32, 64, N.p2.p, N.W16[4], N.B32[4], N.a3.p, true);
    conv_fwd(st, g3, 64, 64, N.a3.p, N.W16[5], N.B32[5], N.a3b.p, true);
    hipLaunchKernelGGL(k_up_fwd, dim3(ew_grid(g2.nvox * 8)), dim3(256), 0, st, g2, g3, 64, N.a3b.p, 96, 0, N.cat2.p);
    hipLaunchKernelGGL(k_copy_channels, dim3(ew_grid(g2.nvox * 4)), dim3(256), 0, st, g2, 32, N.s2.p, 32, 0,
                       N.cat2.p, 96, 64);
    conv_fwd(st, g2, 96, 32, N.cat2.p, N.W16[6], N.B32[6], N.a4.p, true);
    hipLaunchKernelGGL(k_up_fwd, dim3(ew_grid(g1.nvox * 4)), dim3(256), 0, st, g1, g2, 32, N.a4.p, 48, 0, N.cat1.p);
    hipLaunchKernelGGL(k_copy_channels, dim3(ew_grid(g1.nvox * 2)), dim3(256), 0, st, g1, 16, N.s1.p, 16, 0,
                       N.cat1.p, 48, 32);
    conv_fwd(st, g1, 48, 16, N.cat1.p, N.W16[7], N.B32[7], N.a5.p, true);
    conv_fwd(st, g1, 16, 16, N.a5.p, N.W16[8], N.B32[8], N.y.p, false);
    if (z)
        hipLaunchKernelGGL(k_take_output, dim3(ew_grid(nn)), dim3(256), 0, st, g1, N.n[0], N.n[1], N.n[2], N.y.p, 16,
                           N.scale, bal, z);
    RG_HIP(hipGetLastError());
}

// Cosine loss L = 1 - (y.e)/(|y||e|) over the nodes (y = output channel 0 in
// normalised units, e = target); dL/dy written into gy channel 0.
__global__ void k_loss_part(Grid3 g, int n0, int n1, int n2, const bf16* __restrict__ y, const double* __restrict__ e,
                            double* __restrict__ part)
{
    __shared__ double red[3][256];
    double s[3] = {0.0, 0.0, 0.0};
    const int64_t nn = (int64_t)n0 * n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = i % n0, yy = (i / n0) % n1, z = i / ((int64_t)n0 * n1);
        const double yv = (double)(float)y[(((z + 1) * g.p1 + (yy + 1)) * g.p0 + (x + 1)) * 16];
        const double ev = e[i];
        s[0] += yv * ev;
        s[1] += yv * yv;
        s[2] += ev * ev;
    }
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[k * gridDim.x + blockIdx.x] = red[k][0];
}

__global__ void k_loss_final(const double* __restrict__ part, int np, double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3; ++k)
        for (int b = 0; b < np; ++b) s[k] += part[k * np + b];
    out[0] = s[0];
    out[1] = s[1];
    out[2] = s[2];
}

__global__ void k_loss_grad(Grid3 g, int n0, int n1, int n2, const bf16* __restrict__ y, const double* __restrict__ e,
                            const double* __restrict__ dots, bf16* __restrict__ gy)
{
    const double ye = dots[0], yy = dots[1], ee = dots[2];
    const double ny = sqrt(fmax(yy, 1e-300)), ne = sqrt(fmax(ee, 1e-300));
    const int64_t nn = (int64_t)n0 * n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = i % n0, yv = (i / n0) % n1, z = i / ((int64_t)n0 * n1);
        const int64_t p = (((z + 1) * g.p1 + (yv + 1)) * g.p0 + (x + 1)) * 16;
        const double yi = (double)(float)y[p];
        // d/dy [ -(y.e)/(|y||e|) ] = -e/(|y||e|) + (y.e) y /(|y|^3 |e|)
        const double gval = -e[i] / (ny * ne) + ye * yi / (ny * ny * ny * ne);
        gy[p] = (bf16)(float)gval;
    }
}

// Energy-norm loss of the output z (physical units) of a level with operator
// A, for a pair with r = A e:  L = 1 - (r.z)^2 / ((r.e)(z.A z)), the fraction
// of the error energy e.A e that the energy-optimal step along z leaves.
// Partials of (r.z, z.Az, r.e).
__global__ void k_energy_part(int64_t n, const double* __restrict__ r, const double* __restrict__ z,
                              const double* __restrict__ az, const double* __restrict__ e, double* __restrict__ part)
{
    __shared__ double red[3][256];
    double s[3] = {0.0, 0.0, 0.0};
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        s[0] += r[i] * z[i];
        s[1] += z[i] * az[i];
        s[2] += r[i] * e[i];
    }
    for (int k = 0; k < 3; ++k) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) part[k * gridDim.x + blockIdx.x] = red[k][0];
}

// dL/dy into gy channel 0: with c = r.z, q = z.Az, E = r.e,
// dL/dz = -(2c / (E q)) (r - (c/q) A z), and z = y / scale.
__global__ void k_energy_grad(Grid3 g, int n0, int n1, int n2, const double* __restrict__ r,
                              const double* __restrict__ az, const double* __restrict__ dots,
                              const double* __restrict__ scale, bf16* __restrict__ gy)
{
    const double c = dots[0], q = fmax(dots[1], 1e-300), E = fmax(dots[2], 1e-300);
    const double f = -2.0 * c / (E * q) / scale[0], t = c / q;
    const int64_t nn = (int64_t)n0 * n1 * n2;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t x = i % n0, yv = (i / n0) % n1, z = i / ((int64_t)n0 * n1);
        const int64_t p = (((z + 1) * g.p1 + (yv + 1)) * g.p0 + (x + 1)) * 16;
        gy[p] = (bf16)(float)(f * (r[i] - t * az[i]));
    }
}

__global__ void k_adam(float* __restrict__ w, const float* __restrict__ gw, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gg = gw[i];
        const float mi = b1 * m[i] + (1.f - b1) * gg;
        const float vi = b2 * v[i] + (1.f - b2) * gg * gg;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}

// dW, db of layer l from its input x (cin channels) and pre-activation
// gradient dpre (cout channels), both on grid g.
void layer_wgrad(Net& N, int l, const Grid3& g, const bf16* x, const bf16* dpre)
{
    hipStream_t st = N.ctx->stream;
    const int cin = kLayers[l].cin, cout = kLayers[l].cout;
    conv_wgrad(st, g, cin, cout, x, dpre, N.part, N.nchunk);
    hipLaunchKernelGGL(k_sum_wgrad, dim3(ew_grid(layer_wsize(l))), dim3(256), 0, st, N.part, N.nchunk * 4, cout, cin,
                       N.gW[l]);
    const int64_t vpb = (g.nvox + N.nchunk - 1) / N.nchunk;
    hipLaunchKernelGGL(k_bias_grad_part, dim3(N.nchunk), dim3(256), 0, st, g, dpre, cout, N.part, vpb);
    hipLaunchKernelGGL(k_sum_rows, dim3(1), dim3(256), 0, st, N.part, N.nchunk, cout, N.gB[l]);
}

// dx = conv^T(dpre) with the flipped weights of layer l (cout -> cin channels)
void layer_dgrad(Net& N, int l, const Grid3& g, const bf16* dpre, bf16* dx)
{
    hipStream_t st = N.ctx->stream;
    const int cin = kLayers[l].cin, cout = kLayers[l].cout;
    hipLaunchKernelGGL(k_flip_weights, dim3(ew_grid(layer_wsize(l))), dim3(256), 0, st, N.W16[l], cout, cin,
                       N.WT16[l]);
    conv_fwd(st, g, cout, cin, dpre, N.WT16[l], nullptr, dx, false);
}

void relu_bwd(Net& N, const Grid3& g, int c, const bf16* dout, const bf16* act, bf16* dpre)
{
    hipLaunchKernelGGL(k_relu_bwd, dim3(ew_grid(g.nvox * c / 8)), dim3(256), 0, N.ctx->stream, g, c, dout, act, dpre);
}

// Forward + loss + backward; gradients in N.gW / N.gB; returns the loss.
double net_gradients(Net& N, const double* r, const double* dinv, const float* logk, const double* e)
{
    hipStream_t st = N.ctx->stream;
    const Grid3 &g1 = N.g1, &g2 = N.g2, &g3 = N.g3;
    const int64_t nn = (int64_t)N.n[0] * N.n[1] * N.n[2];
    const bool energy = N.loss_op != nullptr;
    net_forward(N, r, dinv, nullptr, logk, energy ? N.zbuf : nullptr, st, true);
    const int np = 256;
    RG_HIP(hipMemsetAsync(N.gy.p, 0, sizeof(bf16) * padded_size(g1) * 16, st));
    if (energy) {
        launch_spmv_on(N.loss_op, 1.0, N.zbuf, 0.0, N.azbuf, st);
        hipLaunchKernelGGL(k_energy_part, dim3(np), dim3(256), 0, st, nn, r, N.zbuf, N.azbuf, e, N.red);
        hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, st, N.red, np, N.red + 3 * 1024);
        hipLaunchKernelGGL(k_energy_grad, dim3(ew_grid(nn)), dim3(256), 0, st, g1, N.n[0], N.n[1], N.n[2], r,
                           N.azbuf, N.red + 3 * 1024, N.scale, N.gy.p);
    } else {
        // cosine loss on the output in normalised units (the output is y / scale)
        hipLaunchKernelGGL(k_loss_part, dim3(np), dim3(256), 0, st, g1, N.n[0], N.n[1], N.n[2], N.y.p, e, N.red);
        hipLaunchKernelGGL(k_loss_final, dim3(1), dim3(64), 0, st, N.red, np, N.red + 3 * 1024);
        hipLaunchKernelGGL(k_loss_grad, dim3(ew_grid(nn)), dim3(256), 0, st, g1, N.n[0], N.n[1], N.n[2], N.y.p, e,
                           N.red + 3 * 1024, N.gy.p);
    }
    // output layer (linear)
    layer_wgrad(N, 8, g1, N.a5.p, N.gy.p);
    layer_dgrad(N, 8, g1, N.gy.p, N.ga5.p);
    // level-1 decoder: a5 = relu(c5(cat1))
    relu_bwd(N, g1, 16, N.ga5.p, N.a5.p, N.gtmp1.p);
    layer_wgrad(N, 7, g1, N.cat1.p, N.gtmp1.p);
    layer_dgrad(N, 7, g1, N.gtmp1.p, N.gcat1.p);
    // cat1 = [up(a4) (32), s1 (16)]
    hipLaunchKernelGGL(k_up_bwd, dim3(ew_grid(g2.nvox * 4)), dim3(256), 0, st, g1, g2, 32, N.gcat1.p, 48, 0, N.ga4.p);
    hipLaunchKernelGGL(k_copy_channels, dim3(ew_grid(g1.nvox * 2)), dim3(256), 0, st, g1, 16, N.gcat1.p, 48, 32,
                       N.gs1.p, 16, 0);
    // level-2 decoder: a4 = relu(c4(cat2))
    relu_bwd(N, g2, 32, N.ga4.p, N.a4.p, N.ga2.p);   // ga2 used as scratch here
    layer_wgrad(N, 6, g2, N.cat2.p, N.ga2.p);
    layer_dgrad(N, 6, g2, N.ga2.p, N.gcat2.p);
    // cat2 = [up(a3b) (64), s2 (32)]
    hipLaunchKernelGGL(k_up_bwd, dim3(ew_grid(g3.nvox * 8)), dim3(256), 0, st, g2, g3, 64, N.gcat2.p, 96, 0, N.ga3b.p);
    hipLaunchKernelGGL(k_copy_channels, dim3(ew_grid(g2.nvox * 4)), dim3(256), 0, st, g2, 32, N.gcat2.p, 96, 64,
                       N.gs2.p, 32, 0);
    // level 3: a3b = relu(c5(a3)), a3 = relu(c4(p2))
    relu_bwd(N, g3, 64, N.ga3b.p, N.a3b.p, N.ga3b.p);
    layer_wgrad(N, 5, g3, N.a3.p, N.ga3b.p);
    layer_dgrad(N, 5, g3, N.ga3b.p, N.ga3.p);
    relu_bwd(N, g3, 64, N.ga3.p, N.a3.p, N.ga3.p);
    layer_wgrad(N, 4, g3, N.p2.p, N.ga3.p);
    layer_dgrad(N, 4, g3, N.ga3.p, N.gp2.p);
    // p2 = pool(s2): add the pooled path into gs2
    hipLaunchKernelGGL(k_pool_bwd, dim3(ew_grid(g2.nvox * 4)), dim3(256), 0, st, g2, g3, 32, N.gp2.p, N.gs2b.p);
    hipLaunchKernelGGL(k_add_channels, dim3(ew_grid(g2.nvox * 4)), dim3(256), 0, st, g2, 32, N.gs2b.p, 32, 0,
                       N.gs2.p, 32, 0);
    // level 2: s2 = relu(c3(a2)), a2 = relu(c2(p1))
    relu_bwd(N, g2, 32, N.gs2.p, N.s2.p, N.gs2.p);
    layer_wgrad(N, 3, g2, N.a2.p, N.gs2.p);
    layer_dgrad(N, 3, g2, N.gs2.p, N.ga2.p);
    relu_bwd(N, g2, 32, N.ga2.p, N.a2.p, N.ga2.p);
    layer_wgrad(N, 2, g2, N.p1.p, N.ga2.p);
    layer_dgrad(N, 2, g2, N.ga2.p, N.gp1.p);
    hipLaunchKernelGGL(k_pool_bwd, dim3(ew_grid(g1.nvox * 2)), dim3(256), 0, st, g1, g2, 16, N.gp1.p, N.gs1b.p);
    hipLaunchKernelGGL(k_add_channels, dim3(ew_grid(g1.nvox * 2)), dim3(256), 0, st, g1, 16, N.gs1b.p, 16, 0,
                       N.gs1.p, 16, 0);
    // level 1: s1 = relu(c1(a1)), a1 = relu(c0(x0))
    relu_bwd(N, g1, 16, N.gs1.p, N.s1.p, N.gs1.p);
    layer_wgrad(N, 1, g1, N.a1.p, N.gs1.p);
    layer_dgrad(N, 1, g1, N.gs1.p, N.ga1.p);
    relu_bwd(N, g1, 16, N.ga1.p, N.a1.p, N.ga1.p);
    layer_wgrad(N, 0, g1, N.x0.p, N.ga1.p);
    RG_HIP(hipGetLastError());
    double dots[3];
    RG_HIP(hipMemcpyAsync(dots, N.red + 3 * 1024, 3 * sizeof(double), hipMemcpyDeviceToHost, st));
    RG_HIP(hipStreamSynchronize(st));
    if (energy) return 1.0 - dots[0] * dots[0] / (std::max(dots[1], 1e-300) * std::max(dots[2], 1e-300));
    const double ny = std::sqrt(std::max(dots[1], 1e-300)), ne = std::sqrt(std::max(dots[2], 1e-300));
    return 1.0 - dots[0] / (ny * ne);
}

void adam_step(Net& N, float lr)
{
    hipStream_t st = N.ctx->stream;
    N.adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = 1.f - std::pow(b1, (float)N.adam_t), c2 = 1.f - std::pow(b2, (float)N.adam_t);
    for (int l = 0; l < kNumLayers; ++l) {
        hipLaunchKernelGGL(k_adam, dim3(ew_grid(layer_wsize(l))), dim3(256), 0, st, N.W32[l], N.gW[l], N.mW[l],
                           N.vW[l], layer_wsize(l), lr, b1, b2, eps, c1, c2);
        hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, st, N.B32[l], N.gB[l], N.mB[l], N.vB[l],
                           (int64_t)kLayers[l].cout, lr, b1, b2, eps, c1, c2);
    }
    refresh_bf16(N);
    RG_HIP(hipGetLastError());
}

}  // namespace lpc
}  // namespace rg

struct rg_lpc {
    rg::lpc::Net net;
    bool training = false;
};

namespace rg {
void lpc_nodes(const rg_lpc* net, int32_t nodes[3])
{
    for (int d = 0; d < 3; ++d) nodes[d] = net->net.n[d];
}

void lpc_forward_balanced(rg_lpc* net, const double* r, const double* dinv, const double* bal, const float* logk,
                          double* z, hipStream_t st)
{
    lpc::net_forward(net->net, r, dinv, bal, logk, z, st);
}
}  // namespace rg

// ====================================================================== C ABI
namespace rg {
// shared with rg_capi.hip: message slot of the last failure
void set_last_error(const std::string& msg);
}

template <class F>
static int lpc_guard(F&& f)
{
    try {
        return f();
    } catch (const rg::Error& e) {
        rg::set_last_error(e.what());
        (void)hipGetLastError();   // do not leave a failed call's error for the next caller
        return e.code;
    } catch (const std::exception& e) {
        rg::set_last_error(e.what());
        (void)hipGetLastError();
        return RG_VALUE_ERROR;
    }
}

extern "C" {

int64_t rg_lpc_num_params(void) { return rg::lpc::total_params(); }

int rg_lpc_create(rg_context* ctx, const int32_t* nodes, int training, rg_lpc** out)
{
    return lpc_guard([&] {
        RG_CHECK(ctx && nodes && out, RG_VALUE_ERROR, "rg_lpc_create: NULL argument");
        for (int d = 0; d < 3; ++d) RG_CHECK(nodes[d] >= 4, RG_VALUE_ERROR, "rg_lpc_create: grid too small");
        RG_HIP(hipSetDevice(ctx->device));
        rg_lpc* h = new rg_lpc();
        try {
            h->training = training != 0;
            rg::lpc::net_init(h->net, ctx, nodes, h->training);
            RG_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
        return RG_OK;
    });
}

int rg_lpc_destroy(rg_lpc* h)
{
    if (h) {
        hipStreamSynchronize(h->net.ctx->stream);
        delete h;
    }
    return RG_OK;
}

// params: per layer W[co][27][ci] then b[co], FP32, layers in network order
int rg_lpc_set_params(rg_lpc* h, const float* params)
{
    return lpc_guard([&] {
        RG_CHECK(h && params, RG_VALUE_ERROR, "rg_lpc_set_params: NULL argument");
        rg::lpc::Net& N = h->net;
        int64_t o = 0;
        for (int l = 0; l < rg::lpc::kNumLayers; ++l) {
            const int64_t ws = rg::lpc::layer_wsize(l);
            RG_HIP(hipMemcpyAsync(N.W32[l], params + o, ws * sizeof(float), hipMemcpyHostToDevice, N.ctx->stream));
            o += ws;
            const int co = rg::lpc::kLayers[l].cout;
            RG_HIP(hipMemcpyAsync(N.B32[l], params + o, co * sizeof(float), hipMemcpyHostToDevice, N.ctx->stream));
            o += co;
        }
        rg::lpc::refresh_bf16(N);
        RG_HIP(hipStreamSynchronize(N.ctx->stream));
        return RG_OK;
    });
}

int rg_lpc_get_params(rg_lpc* h, float* params)
{
    return lpc_guard([&] {
        RG_CHECK(h && params, RG_VALUE_ERROR, "rg_lpc_get_params: NULL argument");
        rg::lpc::Net& N = h->net;
        int64_t o = 0;
        for (int l = 0; l < rg::lpc::kNumLayers; ++l) {
            const int64_t ws = rg::lpc::layer_wsize(l);
            RG_HIP(hipMemcpyAsync(params + o, N.W32[l], ws * sizeof(float), hipMemcpyDeviceToHost, N.ctx->stream));
            o += ws;
            const int co = rg::lpc::kLayers[l].cout;
            RG_HIP(hipMemcpyAsync(params + o, N.B32[l], co * sizeof(float), hipMemcpyDeviceToHost, N.ctx->stream));
            o += co;
        }
        RG_HIP(hipStreamSynchronize(N.ctx->stream));
        return RG_OK;
    });
}

int rg_lpc_apply(rg_lpc* h, const double* r, const double* dinv, const float* logk, double* z)
{
    return lpc_guard([&] {
        RG_CHECK(h && r && z, RG_VALUE_ERROR, "rg_lpc_apply: NULL argument");
        RG_HIP(hipSetDevice(h->net.ctx->device));
        rg::lpc::net_forward(h->net, r, dinv, nullptr, logk, z, h->net.ctx->stream);
        return RG_OK;
    });
}

// grads (host, may be NULL): same layout as the parameters.
int rg_lpc_gradients(rg_lpc* h, const double* r, const double* dinv, const float* logk, const double* e,
                     double* loss, float* grads)
{
    return lpc_guard([&] {
        RG_CHECK(h && h->training && r && e, RG_VALUE_ERROR,
                 "rg_lpc_gradients: needs a network created for training and r, e");
        RG_HIP(hipSetDevice(h->net.ctx->device));
        rg::lpc::Net& N = h->net;
        const double L = rg::lpc::net_gradients(N, r, dinv, logk, e);
        if (loss) *loss = L;
        if (grads) {
            int64_t o = 0;
            for (int l = 0; l < rg::lpc::kNumLayers; ++l) {
                const int64_t ws = rg::lpc::layer_wsize(l);
                RG_HIP(hipMemcpyAsync(grads + o, N.gW[l], ws * sizeof(float), hipMemcpyDeviceToHost, N.ctx->stream));
                o += ws;
                const int co = rg::lpc::kLayers[l].cout;
                RG_HIP(hipMemcpyAsync(grads + o, N.gB[l], co * sizeof(float), hipMemcpyDeviceToHost, N.ctx->stream));
                o += co;
            }
            RG_HIP(hipStreamSynchronize(N.ctx->stream));
        }
        return RG_OK;
    });
}

int rg_lpc_set_loss_operator(rg_lpc* h, rg_matrix* A)
{
    return lpc_guard([&] {
        RG_CHECK(h && h->training, RG_VALUE_ERROR, "rg_lpc_set_loss_operator: needs a network created for training");
        if (A) {
            RG_CHECK(A->g.dim == 3 && A->g.bs == 1 && A->g.nn[0] == h->net.n[0] && A->g.nn[1] == h->net.n[1] &&
                         A->g.nn[2] == h->net.n[2],
                     RG_VALUE_ERROR, "rg_lpc_set_loss_operator: the operator's grid must be the network's");
            RG_CHECK(!A->is_balanced, RG_STATE_ERROR, "rg_lpc_set_loss_operator: balanced matrix is not supported.");
        }
        h->net.loss_op = A;
        return RG_OK;
    });
}

int rg_lpc_train_step(rg_lpc* h, const double* r, const double* dinv, const float* logk, const double* e, float lr,
                      double* loss)
{
    return lpc_guard([&] {
        RG_CHECK(h && h->training && r && e, RG_VALUE_ERROR,
                 "rg_lpc_train_step: needs a network created for training and r, e");
        RG_HIP(hipSetDevice(h->net.ctx->device));
        const double L = rg::lpc::net_gradients(h->net, r, dinv, logk, e);
        rg::lpc::adam_step(h->net, lr);
        if (loss) *loss = L;
        return RG_OK;
    });
}

}  // extern "C"
