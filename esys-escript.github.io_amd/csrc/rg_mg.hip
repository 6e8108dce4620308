// Multilevel preconditioner and flexible CG of RG_METHOD_FCG_LEARNED.
//
// No counterpart in the reference: PASO's preconditioners are Jacobi,
// Gauss-Seidel, ILU0, RILU and an algebraic multigrid
// (paso/src/Preconditioner.cpp:46-137).  This is geometric multigrid on
// ripley's vertex-centred Brick grid, with the learned network correcting
// one of its levels:
//
//   * level l+1 has every other node of level l (ne/2 elements per axis);
//     prolongation P is trilinear interpolation (per axis P[2I, I] = 1,
//     P[2I+1, I] = P[2I+1, I+1] = 1/2), restriction is P^T;
//   * A_{l+1} = P^T A_l P (Galerkin), each coarse operator an ordinary
//     27-point stencil matrix of the coarse Brick; A_0 is the unbalanced
//     operator, D^-1 A_bal D^-1 of the balanced matrix the solver holds
//     (D = diag(balance), paso/src/SystemMatrix.cpp:262-312);
//   * smoothing: `degree` Chebyshev steps on D_l^-1 A_l (D_l = diag A_l)
//     over [lmax / 10, lmax], lmax the Gershgorin bound
//     max_i sum_j |a_ij| / a_ii, before (from zero) and after the coarse
//     correction, values in FP32, vectors in FP64;
//   * coarsest level: `coarse_steps` Chebyshev steps from zero;
//   * the network's level (any level below the finest, the coarsest
//     included), after its post-smoothing: y = Net(r') for the residual
//     r' = b_l - A_l x_l left there (inputs the Jacobi-scaled residual and
//     log10 of the level's diagonal relative to that of a unit coefficient),
//     low-pass filtered by `coarse_steps` Chebyshev steps on A_l y = 0 from
//     y, and x_l += alpha y with alpha = (r'.y)/(y.A_l y), the step that
//     minimises the A_l-norm error along y -- the network corrects what the
//     sub-cycle below and the smoother leave, and never increases the
//     level's error energy.
//
// The V-cycle runs in the physical (unbalanced) variables below level 0:
// r_1 = P^T D^-1 r_bal going down, y_bal += D^-1 P x_1 coming up, so that
// the interpolated corrections are smooth in the coefficient-weighted sense
// and not scaled by sqrt(kappa).
//
// The outer solver is Notay's flexible CG truncated to the last kDirs
// directions (FCG(kDirs)): each new direction A-orthogonalised against the
// kept ones, FP64 operator, FP64 vectors and scalars:
//   beta_j = (z.q_j) / (p_j.q_j)                    k_fcgm_dots    -> k_fcgm_beta
//   p = z - sum_j beta_j p_j                        k_fcgm_dir
//   q = A p, alpha = (p.r) / (p.q)                  k_fcgm_spmv    -> k_fcgm_alpha
//   x += alpha p, r -= alpha q, ||r||               k_fcgm_update  -> ST_FCG_UPDATE
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "rg_device.h"
#include "rg_dispatch.h"
#include "rg_krylov.h"
#include "rg_mg.h"
#include "rg_stencil.h"

namespace rg {
namespace {

constexpr int C3 = 13;   // main slot of the 27-point stencil

struct MgGeom {
    int64_t n0, n1, n2, N, ld;
};

MgGeom geom_of(const rg_matrix* a)
{
    return {a->g.nn[0], a->g.nn[1], a->g.nn[2], a->g.N, a->g.ld};
}

__device__ __forceinline__ bool guard_done(const KrylovState* g) { return g && g->done; }

// Block maximum of v into part[blockIdx.x] (plain store; a one-workgroup
// kernel launched next reads it).
__device__ __forceinline__ void publish_block_max(double v, double* part)
{
    __shared__ double red[kBlock / 64];
    const double w = wave_max(v);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
#pragma unroll
        for (int k = 1; k < kBlock / 64; ++k) m = fmax(m, red[k]);
        part[blockIdx.x] = m;
    }
}

// Entry a(i, j) of a 27-point operator, j = i + offset of full slot s (the
// caller passes ok = both nodes in the grid; otherwise 0).  With symmetric
// storage a slot below the main one is the transposed upper slot of row j.
template <bool SYM>
__device__ __forceinline__ double entry27(const double* __restrict__ val, int64_t ld, int64_t i, int64_t j, int s,
                                          bool ok)
{
    int64_t k;
    if (!SYM) k = (int64_t)s * ld + i;
    else if (s >= C3) k = (int64_t)(s - C3) * ld + i;
    else k = (int64_t)(C3 - s) * ld + j;
    return ok ? val[ok ? k : 0] : 0.0;
}

// ------------------------------------------------------------ set-up kernels
// One axis of the Galerkin product: out = P_AX^T in P_AX with P_AX the
// linear interpolation along axis AX alone (P = P_z P_y P_x, so three such
// products, x then y then z, give P^T A P).  Both operators are 27-point
// stencils (the output on the grid with axis AX halved); one output row per
// thread: the fine rows i = 2I + r along AX (r = -1, 0, 1, weight 1 - |r|/2)
// and their stencil neighbours j = i + s; along AX, j sits at offset
// e = r + s_AX from 2I, which the interpolation maps to the coarse offset e/2
// (e even, weight 1) or to (e-1)/2 and (e+1)/2 (e odd, weight 1/2 each);
// across the other axes the offset stays s.  Each output row reads 3 x 27
// input entries (the direct product P^T A P per coarse node reads 27 x 27).
// With BAL the input is the balanced level-0 matrix and its entries are
// multiplied by ibal_i ibal_j (ibal = 1 / balance).  With SYMO only the main
// and upper output slots are written.
template <int AX, bool SYMI, bool SYMO, bool BAL>
__global__ __launch_bounds__(kBlock) void k_galerkin_axis(const double* __restrict__ in, MgGeom fi,
                                                          const double* __restrict__ ibal, double* __restrict__ out,
                                                          MgGeom fo)
{
    const int64_t I = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (I >= fo.N) return;
    const int64_t o[3] = {I % fo.n0, (I / fo.n0) % fo.n1, I / (fo.n0 * fo.n1)};
    const int64_t fn[3] = {fi.n0, fi.n1, fi.n2};
    double acc[27];
#pragma unroll
    for (int t = 0; t < 27; ++t) acc[t] = 0.0;
#pragma unroll
    for (int r = -1; r <= 1; ++r) {
        int64_t f[3] = {o[0], o[1], o[2]};
        f[AX] = 2 * o[AX] + r;
        if (f[AX] < 0 || f[AX] >= fn[AX]) continue;
        const int64_t i = f[0] + fi.n0 * (f[1] + fi.n1 * f[2]);
        const double wr = r ? 0.5 : 1.0;
        const double ibi = BAL ? ibal[i] : 1.0;
#pragma unroll
        for (int s = 0; s < 27; ++s) {
            const int sd[3] = {s % 3 - 1, (s / 3) % 3 - 1, s / 9 - 1};
            const int64_t jx = f[0] + sd[0], jy = f[1] + sd[1], jz = f[2] + sd[2];
            const bool ok = jx >= 0 && jx < fi.n0 && jy >= 0 && jy < fi.n1 && jz >= 0 && jz < fi.n2;
            const int64_t j = ok ? jx + fi.n0 * (jy + fi.n1 * jz) : 0;
            double a = entry27<SYMI>(in, fi.ld, i, j, s, ok);
            if (BAL) a *= ibi * ibal[j];
            const int e = r + sd[AX];
            const int lo = (e & 1) ? (e - 1) / 2 : e / 2;
            const int hi = (e & 1) ? (e + 1) / 2 : e / 2;
            const double w = ((e & 1) ? 0.5 : 1.0) * wr;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                if (h == 1 && !(e & 1)) continue;
                int d[3] = {sd[0], sd[1], sd[2]};
                d[AX] = h ? hi : lo;
                const int t = (d[0] + 1) + 3 * (d[1] + 1) + 9 * (d[2] + 1);
                acc[t] = fma(w, a, acc[t]);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < 27; ++t) {
        if (SYMO && t < C3) continue;
        out[(int64_t)(SYMO ? t - C3 : t) * fo.ld + I] = acc[t];
    }
}

// Gershgorin ratio max_i sum_j |a_ij| / a_ii of a 27-point operator (block
// maxima).
template <bool SYM>
__global__ __launch_bounds__(kBlock) void k_gershgorin(const double* __restrict__ val, MgGeom a,
                                                       double* __restrict__ gpart)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double g = 0.0;
    if (i < a.N) {
        const int64_t x = i % a.n0, y = (i / a.n0) % a.n1, z = i / (a.n0 * a.n1);
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const int64_t jx = x + t % 3 - 1, jy = y + (t / 3) % 3 - 1, jz = z + t / 9 - 1;
            const bool ok = jx >= 0 && jx < a.n0 && jy >= 0 && jy < a.n1 && jz >= 0 && jz < a.n2;
            const int64_t j = ok ? jx + a.n0 * (jy + a.n1 * jz) : 0;
            s += fabs(entry27<SYM>(val, a.ld, i, j, t, ok));
        }
        g = s / entry27<SYM>(val, a.ld, i, i, C3, true);
    }
    publish_block_max(g, gpart);
}

// Chebyshev coefficients of one level from its Gershgorin bound (one
// workgroup): coef[2k] multiplies the previous update d, coef[2k+1] the
// Jacobi-scaled residual, for steps k = 0 .. kChebSteps-1; coef[2 kChebSteps]
// keeps lmax.
__global__ void k_cheb_table(const double* __restrict__ gpart, int G, double ratio, double* __restrict__ coef)
{
    __shared__ double red[kBlock / 64];
    double m = 0.0;
    for (int b = threadIdx.x; b < G; b += blockDim.x) m = fmax(m, gpart[b]);
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double lmax = red[0];
    for (int k = 1; k < (int)(blockDim.x / 64); ++k) lmax = fmax(lmax, red[k]);
    const double lmin = lmax / ratio;
    const double theta = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = theta / delta;
    double rho_old = 1.0 / sigma;
    coef[0] = 0.0;
    coef[1] = 1.0 / theta;
    for (int k = 1; k < kChebSteps; ++k) {
        const double rho = 1.0 / (2.0 * sigma - rho_old);
        coef[2 * k] = rho * rho_old;
        coef[2 * k + 1] = 2.0 * rho / delta;
        rho_old = rho;
    }
    coef[2 * kChebSteps] = lmax;
}

// Network input of the coarsest level: log10 of a_ii relative to the main
// entry aref of a unit coefficient on that grid (float per node).
__global__ void k_level_logk(const double* __restrict__ main_slot, int64_t N, double aref, float* __restrict__ logk)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < N) logk[i] = (float)log10(fmax(main_slot[i], 1e-300) / aref);
}

// out = 1 / v
__global__ void k_reciprocal(const double* __restrict__ v, int64_t N, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < N) out[i] = 1.0 / v[i];
}

// dinv = 1 / a_ii from the main slot
__global__ void k_inv_main_slot(const double* __restrict__ main_slot, int64_t N, double* __restrict__ dinv)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < N) dinv[i] = 1.0 / main_slot[i];
}

// ------------------------------------------------------------ V-cycle kernels
// First Chebyshev step from x = 0: d = c D^-1 b, x = d.
__global__ __launch_bounds__(kBlock) void k_cheb_first(int64_t N, const double* __restrict__ b,
                                                       const double* __restrict__ dinv, const double* __restrict__ coef,
                                                       double* __restrict__ x, double* __restrict__ d,
                                                       const KrylovState* guard)
{
    if (guard_done(guard)) return;
    const double c = coef[1];
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (int64_t)gridDim.x * kBlock) {
        const double v = c * dinv[i] * b[i];
        d[i] = v;
        x[i] = v;
    }
}

// One Chebyshev step k from x_in: r = b - A x_in, d = c1 d + c2 D^-1 r,
// x_out = x_in + d (coef points at the step's pair (c1, c2); c1 = 0 for the
// first step from a non-zero x).
template <bool SYM, typename VT>
__global__ __launch_bounds__(kBlock) void k_cheb_step(const VT* __restrict__ val, int64_t ld, int64_t N, int64_t s1,
                                                      int64_t s2, int64_t chunk, const double* __restrict__ dinv,
                                                      const double* __restrict__ b, const double* __restrict__ xin,
                                                      double* __restrict__ xout, double* __restrict__ d,
                                                      const double* __restrict__ coef, const KrylovState* guard)
{
    if (guard_done(guard)) return;
    const double c1 = coef[0], c2 = coef[1];
    const RowRange rr = block_rows(N, chunk);
    auto row = [&](int64_t i, const double (&acc)[1]) {
        const double r = b[i] - acc[0];
        // c1 = 0 (a first step): the previous update is not read
        const double dn = c2 * dinv[i] * r + (c1 != 0.0 ? c1 * d[i] : 0.0);
        d[i] = dn;
        xout[i] = xin[i] + dn;
    };
    constexpr int R = spmv_rows_per_thread<VT>();
    // the vector streams R rows per access where every one is 16-byte
    // aligned at the run starts (multiples of R rows)
    const bool runs = ((reinterpret_cast<uintptr_t>(b) | reinterpret_cast<uintptr_t>(dinv) |
                        reinterpret_cast<uintptr_t>(xin) | reinterpret_cast<uintptr_t>(xout) |
                        reinterpret_cast<uintptr_t>(d)) & 15) == 0;
    if (runs) {
        stencil_block_runs<3, SYM, VT>(val, ld, N, s1, s2, rr.r0, rr.r1, xin, [&](int64_t i0, const double (&acc)[R]) {
            double bv[R], dv[R], xv[R], dn[R], xo[R];
            load_run<R>(b + i0, bv);
            load_run<R>(dinv + i0, dv);
            load_run<R>(xin + i0, xv);
            double dp[R];
            if (c1 != 0.0) load_run<R>(d + i0, dp);
#pragma unroll
            for (int k = 0; k < R; ++k) {
                dn[k] = c2 * dv[k] * (bv[k] - acc[k]) + (c1 != 0.0 ? c1 * dp[k] : 0.0);
                xo[k] = xv[k] + dn[k];
            }
            store_run<R>(d + i0, dn);
            store_run<R>(xout + i0, xo);
        }, row);
    } else {
        stencil_block<3, 1, SYM, VT>(val, ld, N, s1, s2, rr.r0, rr.r1, xin, row);
    }
}

// All `steps` Chebyshev steps of a small level (N <= kSmallNodes, one node
// per thread of one workgroup): the thread keeps its row's 27 values and
// neighbour indices in registers, x (double-buffered) and the previous
// update d live in LDS, and a barrier separates the steps (each reads its
// neighbours' x of the previous step).  The same recurrence as
// k_cheb_first + k_cheb_step (FROM_ZERO: the first step from x = 0 with
// coef[1]; otherwise from xin with step 0's c1 = 0), one launch instead of
// `steps` on the coarsest levels (9^3, 7^3 nodes), where each launch costs
// more than its work.  (With several nodes per thread and the values
// re-read from L2 each step, a 17^3 level ran slower this way than as
// separate launches: each step then waits on those loads.)
constexpr int kSmallNodes = 1024;

template <bool SYM, bool FROM_ZERO>
__global__ __launch_bounds__(kSmallNodes) void k_cheb_small(const float* __restrict__ val, MgGeom a,
                                                            const double* __restrict__ dinv,
                                                            const double* __restrict__ b,
                                                            const double* __restrict__ xin, double* __restrict__ xout,
                                                            double* __restrict__ dout, const double* __restrict__ coef,
                                                            int steps, const KrylovState* guard)
{
    if (guard_done(guard)) return;
    __shared__ double xs[2][kSmallNodes];
    const int N = (int)a.N;
    const int i = threadIdx.x;
    const bool own = i < N;
    // this row's values and neighbours (value 0 and neighbour i outside the grid)
    float v[27];
    int nb[27];
    {
        const int n0 = (int)a.n0, n1 = (int)a.n1, n2 = (int)a.n2, ld = (int)a.ld;
        const int x = i % n0, y = (i / n0) % n1, z = i / (n0 * n1);
#pragma unroll
        for (int t = 0; t < 27; ++t) {
            const int jx = x + t % 3 - 1, jy = y + (t / 3) % 3 - 1, jz = z + t / 9 - 1;
            const bool ok = own && jx >= 0 && jx < n0 && jy >= 0 && jy < n1 && jz >= 0 && jz < n2;
            const int j = ok ? jx + n0 * (jy + n1 * jz) : 0;
            int k;
            if (!SYM) k = t * ld + i;
            else if (t >= C3) k = (t - C3) * ld + i;
            else k = (C3 - t) * ld + j;
            v[t] = ok ? val[k] : 0.f;
            nb[t] = j;
        }
    }
    const double bi = own ? b[i] : 0.0, di = own ? dinv[i] : 0.0;
    double d = 0.0;
    int k0 = 0;
    if (FROM_ZERO) {
        d = coef[1] * di * bi;
        k0 = 1;
    }
    if (own) xs[0][i] = FROM_ZERO ? d : xin[i];
    __syncthreads();
    int cur = 0;
    for (int k = k0; k < steps; ++k) {
        const double c1 = coef[2 * k], c2 = coef[2 * k + 1];
        if (own) {
            double acc = 0.0;
#pragma unroll
            for (int t = 0; t < 27; ++t) acc = fma((double)v[t], xs[cur][nb[t]], acc);
            // c1 = 0 (a first step): the previous update is not read
            d = c2 * di * (bi - acc) + (c1 != 0.0 ? c1 * d : 0.0);
            xs[cur ^ 1][i] = xs[cur][i] + d;
        }
        cur ^= 1;
        __syncthreads();   // the step's x is complete before the next reads it
    }
    if (own) {
        xout[i] = xs[cur][i];
        dout[i] = d;
    }
}

// r = b - A x
template <bool SYM, typename VT>
__global__ __launch_bounds__(kBlock) void k_residual(const VT* __restrict__ val, int64_t ld, int64_t N, int64_t s1,
                                                     int64_t s2, int64_t chunk, const double* __restrict__ b,
                                                     const double* __restrict__ x, double* __restrict__ r,
                                                     const KrylovState* guard)
{
    if (guard_done(guard)) return;
    const RowRange rr = block_rows(N, chunk);
    auto row = [&](int64_t i, const double (&acc)[1]) { r[i] = b[i] - acc[0]; };
    constexpr int R = spmv_rows_per_thread<VT>();
    if (((reinterpret_cast<uintptr_t>(b) | reinterpret_cast<uintptr_t>(r)) & 15) == 0) {
        stencil_block_runs<3, SYM, VT>(val, ld, N, s1, s2, rr.r0, rr.r1, x, [&](int64_t i0, const double (&acc)[R]) {
            double bv[R], rv[R];
            load_run<R>(b + i0, bv);
#pragma unroll
            for (int k = 0; k < R; ++k) rv[k] = bv[k] - acc[k];
            store_run<R>(r + i0, rv);
        }, row);
    } else {
        stencil_block<3, 1, SYM, VT>(val, ld, N, s1, s2, rr.r0, rr.r1, x, row);
    }
}

// b_c = P^T (r / bal): one coarse node per thread over its 27 fine nodes.
template <bool BAL>
__global__ __launch_bounds__(kBlock) void k_restrict(MgGeom f, MgGeom c, const double* __restrict__ r,
                                                     const double* __restrict__ bal, double* __restrict__ bc,
                                                     const KrylovState* guard)
{
    if (guard_done(guard)) return;
    const int64_t I = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (I >= c.N) return;
    const int64_t X = I % c.n0, Y = (I / c.n0) % c.n1, Z = I / (c.n0 * c.n1);
    double s = 0.0;
#pragma unroll
    for (int di = 0; di < 27; ++di) {
        const int dx = di % 3 - 1, dy = (di / 3) % 3 - 1, dz = di / 9 - 1;
        const int64_t fx = 2 * X + dx, fy = 2 * Y + dy, fz = 2 * Z + dz;
        const bool ok = fx >= 0 && fx < f.n0 && fy >= 0 && fy < f.n1 && fz >= 0 && fz < f.n2;
        const int64_t i = ok ? fx + f.n0 * (fy + f.n1 * fz) : 0;
        const double w = (dx ? 0.5 : 1.0) * (dy ? 0.5 : 1.0) * (dz ? 0.5 : 1.0);
        const double v = BAL ? r[i] / bal[i] : r[i];
        s += ok ? w * v : 0.0;
    }
    bc[I] = s;
}

// x_f += (P x_c) / bal: one fine node per thread over the (up to 8) coarse
// nodes that interpolate it.
template <bool BAL>
__global__ __launch_bounds__(kBlock) void k_prolong(MgGeom f, MgGeom c, const double* __restrict__ xc,
                                                    const double* __restrict__ bal, double* __restrict__ xf,
                                                    const KrylovState* guard)
{
    if (guard_done(guard)) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= f.N) return;
    const int64_t x = i % f.n0, y = (i / f.n0) % f.n1, z = i / (f.n0 * f.n1);
    const int64_t X0 = x >> 1, Y0 = y >> 1, Z0 = z >> 1;
    const int ox = (int)(x & 1), oy = (int)(y & 1), oz = (int)(z & 1);
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const int kx = k & 1, ky = (k >> 1) & 1, kz = (k >> 2) & 1;
        const bool use = (kx <= ox) && (ky <= oy) && (kz <= oz);
        const int64_t J = use ? (X0 + kx) + c.n0 * ((Y0 + ky) + c.n1 * (Z0 + kz)) : 0;
        const double w = (ox ? 0.5 : 1.0) * (oy ? 0.5 : 1.0) * (oz ? 0.5 : 1.0);
        s += use ? w * xc[J] : 0.0;
    }
    if (BAL) s /= bal[i];
    xf[i] += s;
}

// Ay for the coarse step length: partials (r.y, y.Ay) of A y.
template <bool SYM>
__global__ __launch_bounds__(kBlock) void k_coarse_energy(const double* __restrict__ val, int64_t ld, int64_t N,
                                                          int64_t s1, int64_t s2, int64_t chunk,
                                                          const double* __restrict__ y, const double* __restrict__ r,
                                                          double* __restrict__ partials, const KrylovState* guard)
{
    if (guard_done(guard)) return;
    const RowRange rr = block_rows(N, chunk);
    double s[2] = {0.0, 0.0};
    stencil_block<3, 1, SYM, double>(val, ld, N, s1, s2, rr.r0, rr.r1, y, [&](int64_t i, const double (&acc)[1]) {
        s[0] = fma(r[i], y[i], s[0]);
        s[1] = fma(y[i], acc[0], s[1]);
    });
    publish_block_sums<2>(s, partials);
}

// alpha = (r.y) / (y.A y), 0 when y.A y is not positive.
// x += alpha y with the energy-optimal alpha = (r'.y)/(y.A y) (0 when
// y.A y is not positive), reduced from k_coarse_energy's partials by every
// workgroup itself, in the same fixed order everywhere (one launch instead
// of a one-workgroup stage and the update)
__global__ __launch_bounds__(kBlock) void k_coarse_step(int64_t N, const double* __restrict__ partials, int G,
                                                        const double* __restrict__ y, double* __restrict__ x,
                                                        const KrylovState* guard)
{
    if (guard_done(guard)) return;
    double tot[2];
    reduce_partials<2>(partials, G, tot);
    const double a = tot[1] > 0.0 ? tot[0] / tot[1] : 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < N; i += (int64_t)gridDim.x * kBlock)
        x[i] = fma(a, y[i], x[i]);
}

// ------------------------------------------------------------ FCG(kDirs)
// Direction t of the ring lives in slot t % kDirs; the iteration with
// st->num_iter = k forms direction k from the kept directions k-1 .. k-kDirs.
struct FcgArgs {
    double* x;
    double* r;
    const double* z;
    double* P;          // kDirs vectors of n, ld apart
    double* Q;
    double* coef;       // [0, kDirs): p_j.q_j per slot; [kDirs, 2 kDirs): beta_j
    KrylovState* st;
    double* partials;
    int64_t n, vld, chunk;
    const double* wn = nullptr;   // k_fcgm_update: the residual norm is taken of r * wn (wn = 1 / balance)
};

__global__ __launch_bounds__(kBlock) void k_fcgm_dots(FcgArgs a)
{
    if (a.st->done) return;
    const int k = a.st->num_iter;
    const double* q[kDirs];
    bool use[kDirs];
#pragma unroll
    for (int j = 0; j < kDirs; ++j) {
        const int t = k - 1 - j;
        use[j] = t >= 0;
        q[j] = a.Q + (int64_t)((t >= 0 ? t : 0) % kDirs) * a.vld;
    }
    double s[kDirs];
#pragma unroll
    for (int j = 0; j < kDirs; ++j) s[j] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlock) {
        const double zi = a.z[i];
#pragma unroll
        for (int j = 0; j < kDirs; ++j)
            if (use[j]) s[j] = fma(zi, q[j][i], s[j]);
    }
    publish_block_sums<kDirs>(s, a.partials);
}

__global__ __launch_bounds__(kBlock) void k_fcgm_beta(FcgArgs a, int G)
{
    if (a.st->done) return;
    double tot[kDirs];
    reduce_partials<kDirs>(a.partials, G, tot);
    if (threadIdx.x == 0) {
        const int k = a.st->num_iter;
#pragma unroll
        for (int j = 0; j < kDirs; ++j) {
            const int t = k - 1 - j;
            a.coef[kDirs + j] = (t >= 0) ? tot[j] / a.coef[t % kDirs] : 0.0;
        }
    }
}

// p_k = z - sum_j beta_j p_{k-1-j}; the slot of p_k is that of p_{k-kDirs},
// read by the same thread before it is overwritten.
__global__ __launch_bounds__(kBlock) void k_fcgm_dir(FcgArgs a)
{
    if (a.st->done) return;
    const int k = a.st->num_iter;
    const double* p[kDirs];
    double beta[kDirs];
#pragma unroll
    for (int j = 0; j < kDirs; ++j) {
        const int t = k - 1 - j;
        beta[j] = a.coef[kDirs + j];
        p[j] = a.P + (int64_t)((t >= 0 ? t : 0) % kDirs) * a.vld;
    }
    double* pk = a.P + (int64_t)(k % kDirs) * a.vld;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlock) {
        double v = a.z[i];
#pragma unroll
        for (int j = 0; j < kDirs; ++j) v = fma(-beta[j], beta[j] != 0.0 ? p[j][i] : 0.0, v);
        pk[i] = v;
    }
}

// q_k = A p_k (FP64 values) with the partials (p.q, p.r).
template <bool SYM>
__global__ __launch_bounds__(kBlock) void k_fcgm_spmv(FcgArgs a, const double* __restrict__ val, int64_t ld,
                                                      int64_t s1, int64_t s2)
{
    if (a.st->done) return;
    const int k = a.st->num_iter;
    const double* __restrict__ p = a.P + (int64_t)(k % kDirs) * a.vld;
    double* __restrict__ q = a.Q + (int64_t)(k % kDirs) * a.vld;
    const RowRange rr = block_rows(a.n, a.chunk);
    double s[2] = {0.0, 0.0};
    stencil_block<3, 1, SYM, double>(val, ld, a.n, s1, s2, rr.r0, rr.r1, p, [&](int64_t i, const double (&acc)[1]) {
        q[i] = acc[0];
        s[0] = fma(acc[0], p[i], s[0]);
        s[1] = fma(a.r[i], p[i], s[1]);
    });
    publish_block_sums<2>(s, a.partials);
}

__global__ __launch_bounds__(kBlock) void k_fcgm_alpha(FcgArgs a, int G)
{
    if (a.st->done) return;
    double tot[2];
    reduce_partials<2>(a.partials, G, tot);
    if (threadIdx.x == 0) {
        KrylovState* st = a.st;
        if (!(tot[0] > 0.0)) {
            st->done = 1;
            st->status = KS_BREAKDOWN;
        } else {
            st->alpha = tot[1] / tot[0];
            a.coef[st->num_iter % kDirs] = tot[0];
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_fcgm_update(FcgArgs a)
{
    if (a.st->done) return;
    const double alpha = a.st->alpha;
    const int k = a.st->num_iter;
    const double* __restrict__ p = a.P + (int64_t)(k % kDirs) * a.vld;
    const double* __restrict__ q = a.Q + (int64_t)(k % kDirs) * a.vld;
    double s[1] = {0.0};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * kBlock) {
        a.x[i] = fma(alpha, p[i], a.x[i]);
        const double rn = fma(-alpha, q[i], a.r[i]);
        a.r[i] = rn;
        const double rw = a.wn ? rn * a.wn[i] : rn;
        s[0] = fma(rw, rw, s[0]);
    }
    publish_block_sums<1>(s, a.partials);
}

__global__ __launch_bounds__(kBlock) void k_fcgm_end(KrylovState* st, const double* __restrict__ partials, int G)
{
    if (st->done) return;
    double tot[1];
    reduce_partials<1>(partials, G, tot);
    if (threadIdx.x == 0) stage_apply<ST_FCG_UPDATE, 1>(st, tot);
}

int grid_for(int64_t n) { return (int)std::min<int64_t>((n + kBlock - 1) / kBlock, 2048); }

// the grid of a one-node-per-thread kernel
unsigned grid1(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

MgGeom make_mg_geom(int64_t n0, int64_t n1, int64_t n2)
{
    const int64_t N = n0 * n1 * n2;
    return {n0, n1, n2, N, ((N + 63) / 64) * 64};
}

// The two partly coarsened grids of the one-axis products from f to c.
void galerkin_grids(const MgGeom& f, const MgGeom& c, MgGeom* g1, MgGeom* g2)
{
    *g1 = make_mg_geom(c.n0, f.n1, f.n2);
    *g2 = make_mg_geom(c.n0, c.n1, f.n2);
}

// Doubles of the intermediate operators of galerkin_product (nslot values per
// row, the coarse storage's slot count).
int64_t galerkin_tmp_doubles(const MgGeom& f, const MgGeom& c, int nslot)
{
    MgGeom g1, g2;
    galerkin_grids(f, c, &g1, &g2);
    return (int64_t)nslot * (g1.ld + g2.ld);
}

// cv = P^T A_f P (fv on grid f, symmetric storage sf; cv on grid c,
// storage sc) as the x, y and z one-axis products, the two intermediate
// operators in tmp (galerkin_tmp_doubles) with the coarse storage; ibal
// (level 0 of a balanced hierarchy, else null) unbalances the fine entries.
void galerkin_product(const double* fv, const MgGeom& f, bool sf, const double* ibal, double* cv, const MgGeom& c,
                      bool sc, double* tmp, hipStream_t st)
{
    MgGeom g1, g2;
    galerkin_grids(f, c, &g1, &g2);
    double* t1 = tmp;
    double* t2 = tmp + (int64_t)(sc ? 14 : 27) * g1.ld;
    auto first = [&](auto SI, auto SO) {
        if (ibal)
            hipLaunchKernelGGL((k_galerkin_axis<0, SI.value, SO.value, true>), dim3(grid1(g1.N)), dim3(kBlock), 0, st,
                               fv, f, ibal, t1, g1);
        else
            hipLaunchKernelGGL((k_galerkin_axis<0, SI.value, SO.value, false>), dim3(grid1(g1.N)), dim3(kBlock), 0, st,
                               fv, f, nullptr, t1, g1);
    };
    auto rest = [&](auto S) {
        hipLaunchKernelGGL((k_galerkin_axis<1, S.value, S.value, false>), dim3(grid1(g2.N)), dim3(kBlock), 0, st, t1,
                           g1, nullptr, t2, g2);
        hipLaunchKernelGGL((k_galerkin_axis<2, S.value, S.value, false>), dim3(grid1(c.N)), dim3(kBlock), 0, st, t2,
                           g2, nullptr, cv, c);
    };
    if (sf && sc) first(BC<true>{}, BC<true>{});
    else if (sf) first(BC<true>{}, BC<false>{});
    else if (sc) first(BC<false>{}, BC<true>{});
    else first(BC<false>{}, BC<false>{});
    if (sc) rest(BC<true>{});
    else rest(BC<false>{});
    RG_HIP(hipGetLastError());
}

}  // namespace

// ================================================================== host side
struct MgLevel {
    rg_matrix* A = nullptr;        // level >= 1: owned coarse operator (FP64 values + FP32 copy)
    double* dinv = nullptr;        // level >= 1: owned 1/a_ii; level 0: the solver's Jacobi inverse
    double *xa = nullptr, *xb = nullptr, *b = nullptr, *r = nullptr, *d = nullptr;
    double* coef = nullptr;        // Chebyshev table (device), 2 kChebSteps + 1 doubles
    MgGeom g{};
};

struct rg_mg_state {
    int nlev = 0;
    MgLevel lev[kMaxLevels];
    double* coefs = nullptr;       // kMaxLevels Chebyshev tables
    double* part = nullptr;        // block partials of the coarse-step reduction
    double* gpart = nullptr;       // block maxima of the set-up's Gershgorin bounds (one per 256 fine rows)
    double* ibal = nullptr;        // 1 / balance of the level-0 matrix
    double* gtmp = nullptr;        // intermediate operators of the Galerkin products
    float* logk = nullptr;         // network input of the coarsest level
    double* y = nullptr;           // network output and its filtered copies
    double* y2 = nullptr;
    double* zero = nullptr;        // zero right-hand side of the output filter
    double* P = nullptr;           // FCG ring of directions and their products
    double* Q = nullptr;
    double* fcg = nullptr;
    int64_t vld = 0;
    bool with_net = false;
    int net_level = -1;            // level the network corrects (-1: none)
    bool balanced = true;          // level 0 is a balanced matrix (D A D, D = m->balance)
};

namespace {

void free_level_vectors(MgLevel& L)
{
    for (double* p : {L.xa, L.xb, L.b, L.r, L.d})
        if (p) hipFree(p);
    L.xa = L.xb = L.b = L.r = L.d = nullptr;
}

double* dalloc(int64_t n)
{
    double* p = nullptr;
    RG_HIP(hipMalloc(&p, sizeof(double) * std::max<int64_t>(n, 1)));
    return p;
}

// Levels of the hierarchy of m: m->mg_levels, or (0) until a level has at
// most 9^3 nodes or cannot be halved.  With a network attached, *net_level
// is the level whose node grid is the network's (at least one halving; the
// hierarchy reaches it), else -1.
int plan_levels(const rg_matrix* m, int* net_level)
{
    const Geom& g = m->g;
    int ne[3] = {g.ne[0], g.ne[1], g.ne[2]};
    int32_t target[3] = {0, 0, 0};
    const bool with_net = m->lpc != nullptr;
    if (with_net) lpc_nodes(m->lpc, target);
    *net_level = -1;
    int levels = 1;
    for (;;) {
        const int64_t nodes = (int64_t)(ne[0] + 1) * (ne[1] + 1) * (ne[2] + 1);
        if (with_net && ne[0] + 1 == target[0] && ne[1] + 1 == target[1] && ne[2] + 1 == target[2]) {
            RG_CHECK(levels >= 2, RG_VALUE_ERROR,
                     "FCG_LEARNED: the network's node grid must be a coarse level (at least one halving) of the "
                     "matrix's grid");
            *net_level = levels - 1;
        }
        const bool deep_enough = m->mg_levels > 0 ? levels >= m->mg_levels : nodes <= 9 * 9 * 9;
        if (deep_enough && (!with_net || *net_level >= 0)) return levels;
        const bool halvable = ne[0] % 2 == 0 && ne[1] % 2 == 0 && ne[2] % 2 == 0 && ne[0] >= 2 && ne[1] >= 2 &&
                              ne[2] >= 2;
        if (!halvable || levels == kMaxLevels) {
            RG_CHECK(!with_net || *net_level >= 0, RG_VALUE_ERROR,
                     "FCG_LEARNED: the network's node grid is not a level of this matrix's grid hierarchy "
                     "(every other node per level)");
            RG_CHECK(m->mg_levels <= 0, RG_VALUE_ERROR, "FCG_LEARNED: the grid cannot be halved that many times");
            return levels;
        }
        for (int d = 0; d < 3; ++d) ne[d] /= 2;
        ++levels;
    }
}

}  // namespace

void mg_free(rg_matrix* m)
{
    rg_mg_state* s = m->mg;
    if (!s) return;
    hipStreamSynchronize(m->ctx->stream);
    for (int l = 0; l < s->nlev; ++l) {
        MgLevel& L = s->lev[l];
        free_level_vectors(L);
        if (l > 0) {
            if (L.dinv) hipFree(L.dinv);
            if (L.A) rg_matrix_destroy(L.A);
        }
    }
    for (void* p : {(void*)s->coefs, (void*)s->part, (void*)s->gpart, (void*)s->ibal, (void*)s->gtmp,
                    (void*)s->logk, (void*)s->y, (void*)s->y2, (void*)s->zero, (void*)s->P,
                    (void*)s->Q, (void*)s->fcg})
        if (p) hipFree(p);
    delete s;
    m->mg = nullptr;
}

// Builds (on first use or when the plan changes) and refreshes the
// hierarchy of the matrix m -- balanced (the solver's) or not: Galerkin
// operators, their FP32 copies, Jacobi inverses, Chebyshev tables and the
// network's coefficient input.
void mg_setup(rg_matrix* m, bool balanced)
{
    const Geom& g = m->g;
    RG_CHECK(g.dim == 3 && g.bs == 1, RG_VALUE_ERROR, "FCG_LEARNED: scalar Brick matrices only");
    RG_CHECK(m->is_balanced == balanced && m->val32_valid, RG_STATE_ERROR, "FCG_LEARNED: matrix not prepared");
    int net_level = -1;
    const int nlev = plan_levels(m, &net_level);
    const bool with_net = net_level >= 0;
    if (m->mg && (m->mg->nlev != nlev || m->mg->net_level != net_level)) mg_free(m);
    hipStream_t st = m->ctx->stream;
    if (!m->mg) {
        if (m->work) m->work->graph_method = -1;   // captured iterations hold the old buffers
        rg_mg_state* s = new rg_mg_state();
        m->mg = s;
        s->nlev = nlev;
        s->with_net = with_net;
        s->net_level = net_level;
        s->coefs = dalloc((int64_t)kMaxLevels * (2 * kChebSteps + 1));
        s->part = dalloc(2 * kMaxGrid);
        s->gpart = dalloc(grid1(g.N));
        s->ibal = dalloc(g.N);
        s->vld = ((g.N + 63) / 64) * 64;
        s->P = dalloc(kDirs * s->vld);
        s->Q = dalloc(kDirs * s->vld);
        s->fcg = dalloc(2 * kDirs);
        for (int l = 0; l < nlev; ++l) {
            MgLevel& L = s->lev[l];
            L.coef = s->coefs + (int64_t)l * (2 * kChebSteps + 1);
            if (l == 0) {
                L.g = geom_of(m);
            } else {
                const rg_matrix* f = s->lev[l - 1].A ? s->lev[l - 1].A : m;
                int32_t ne[3];
                double len[3];
                for (int d = 0; d < 3; ++d) {
                    ne[d] = f->g.ne[d] / 2;
                    len[d] = f->g.len[d];
                }
                const int rc = rg_matrix_create(m->ctx, 3, ne, len, 1, m->symmetric ? RG_MATRIX_SYMMETRIC : 0, &L.A);
                RG_CHECK(rc == RG_OK, rc, "FCG_LEARNED: cannot create a coarse level");
                RG_HIP(hipMalloc(&L.A->val32, value_count(L.A->g) * sizeof(float)));
                L.dinv = dalloc(L.A->g.N);
                L.g = geom_of(L.A);
                L.b = dalloc(L.g.N);
            }
            L.xa = dalloc(L.g.N);
            L.xb = dalloc(L.g.N);
            L.r = dalloc(L.g.N);
            L.d = dalloc(L.g.N);
        }
        if (nlev > 1)   // the first product is the largest
            s->gtmp = dalloc(galerkin_tmp_doubles(s->lev[0].g, s->lev[1].g, m->symmetric ? 14 : 27));
        const MgGeom& gc = s->lev[std::max(net_level, 0)].g;
        if (with_net) {
            RG_HIP(hipMalloc(&s->logk, sizeof(float) * gc.N));
            s->y = dalloc(gc.N);
            s->y2 = dalloc(gc.N);
            s->zero = dalloc(gc.N);
            RG_HIP(hipMemsetAsync(s->zero, 0, sizeof(double) * gc.N, st));
        }
    }
    rg_mg_state* s = m->mg;
    if (s->balanced != balanced && m->work) m->work->graph_method = -1;
    s->balanced = balanced;
    s->lev[0].A = nullptr;
    s->lev[0].dinv = m->dinv;
    if (!balanced) {
        // 1 / a_ii of the unbalanced matrix in the Jacobi slot, which a solve
        // then rebuilds for its balanced matrix
        const double* main_slot = m->val + (int64_t)(m->symmetric ? 0 : C3) * g.ld;
        hipLaunchKernelGGL(k_inv_main_slot, dim3(grid1(g.N)), dim3(kBlock), 0, st, main_slot, g.N, m->dinv);
        RG_HIP(hipGetLastError());
        m->have_prec = false;
    }
    // level 0: Gershgorin bound of the balanced operator (the spectrum of
    // D^-1 A does not change under the balancing)
    {
        const MgGeom a = s->lev[0].g;
        const unsigned G = grid1(a.N);
        if (m->symmetric) hipLaunchKernelGGL(k_gershgorin<true>, dim3(G), dim3(kBlock), 0, st, m->val, a, s->gpart);
        else hipLaunchKernelGGL(k_gershgorin<false>, dim3(G), dim3(kBlock), 0, st, m->val, a, s->gpart);
        hipLaunchKernelGGL(k_cheb_table, dim3(1), dim3(kBlock), 0, st, s->gpart, (int)G, kChebRatio, s->lev[0].coef);
        RG_HIP(hipGetLastError());
    }
    // 1 / balance: the level-0 Galerkin product's scaling and the flexible
    // CG's unbalanced residual norm
    if (balanced) {
        hipLaunchKernelGGL(k_reciprocal, dim3(grid1(g.N)), dim3(kBlock), 0, st, m->balance, g.N, s->ibal);
        RG_HIP(hipGetLastError());
    }
    for (int l = 1; l < s->nlev; ++l) {
        MgLevel& L = s->lev[l];
        const rg_matrix* F = (l == 1) ? m : s->lev[l - 1].A;
        const MgGeom f = s->lev[l - 1].g, c = L.g;
        const unsigned G = grid1(c.N);
        const double* bal = (l == 1 && balanced) ? s->ibal : nullptr;
        galerkin_product(F->val, f, F->symmetric, bal, L.A->val, c, L.A->symmetric, s->gtmp, st);
        // its Jacobi inverse and Gershgorin bound
        const double* main_slot = L.A->val + (int64_t)(L.A->symmetric ? 0 : C3) * c.ld;
        hipLaunchKernelGGL(k_inv_main_slot, dim3(G), dim3(kBlock), 0, st, main_slot, c.N, L.dinv);
        if (L.A->symmetric)
            hipLaunchKernelGGL(k_gershgorin<true>, dim3(G), dim3(kBlock), 0, st, L.A->val, c, s->gpart);
        else
            hipLaunchKernelGGL(k_gershgorin<false>, dim3(G), dim3(kBlock), 0, st, L.A->val, c, s->gpart);
        hipLaunchKernelGGL(k_cheb_table, dim3(1), dim3(kBlock), 0, st, s->gpart, (int)G, kChebRatio, L.coef);
        RG_HIP(hipGetLastError());
        L.A->val_zero = false;
        L.A->is_balanced = false;
        launch_to_fp32(L.A);
        L.A->val32_valid = true;
    }
    if (s->with_net) {
        const MgLevel& L = s->lev[s->net_level];
        const rg_matrix* A = L.A ? L.A : m;
        // main entry of a unit coefficient: 8 elements x sum_d vol / (9 h_d^2)
        double aref = 0.0;
        const double vol = A->g.h[0] * A->g.h[1] * A->g.h[2];
        for (int d = 0; d < 3; ++d) aref += 8.0 * vol / (9.0 * A->g.h[d] * A->g.h[d]);
        const double* main_slot = A->val + (int64_t)(A->symmetric ? 0 : C3) * A->g.ld;
        hipLaunchKernelGGL(k_level_logk, dim3(grid1(L.g.N)), dim3(kBlock), 0, st, main_slot, L.g.N, aref, s->logk);
        RG_HIP(hipGetLastError());
    }
}

namespace {

template <class F>
void by_sym_vt(bool sym, F&& f)
{
    if (sym) f(BC<true>{});
    else f(BC<false>{});
}

// Chebyshev smoothing of level l: `steps` steps from x (x_from_zero: from 0,
// b only).  Returns the buffer holding the result (xa or xb of the level, or
// `out` for the last step when out is not null).
// steps Chebyshev steps of a level with at most kSmallNodes nodes in one
// launch (k_cheb_small), from zero (x null) or from x, into dst.
void cheb_small(rg_mg_state* s, const rg_matrix* A, int l, const double* b, const double* x, double* dst, int steps,
                hipStream_t st, const KrylovState* guard)
{
    MgLevel& L = s->lev[l];
    by_sym_vt(A->symmetric, [&](auto S) {
        if (x)
            hipLaunchKernelGGL((k_cheb_small<S.value, false>), dim3(1), dim3(kSmallNodes), 0, st, A->val32, L.g,
                               L.dinv, b, x, dst, L.d, L.coef, steps, guard);
        else
            hipLaunchKernelGGL((k_cheb_small<S.value, true>), dim3(1), dim3(kSmallNodes), 0, st, A->val32, L.g,
                               L.dinv, b, nullptr, dst, L.d, L.coef, steps, guard);
    });
    RG_HIP(hipGetLastError());
}

double* smooth(rg_mg_state* s, const rg_matrix* Am, int l, const double* b, double* x, bool from_zero, int steps,
               double* out, hipStream_t st, const KrylovState* guard)
{
    MgLevel& L = s->lev[l];
    const rg_matrix* A = L.A ? L.A : Am;
    const Geom& g = A->g;
    if (g.N <= kSmallNodes && steps >= 2) {
        double* dst = out ? out : (x == L.xa ? L.xb : L.xa);
        cheb_small(s, A, l, b, from_zero ? nullptr : x, dst, steps, st, guard);
        return dst;
    }
    const Launch Ls = chunked_launch(g.N, spmv_rows_per_thread<float>());
    double* cur = x;
    int k = 0;
    if (from_zero) {
        double* dst = (steps == 1 && out) ? out : L.xa;
        hipLaunchKernelGGL(k_cheb_first, dim3(grid_for(g.N)), dim3(kBlock), 0, st, g.N, b, L.dinv, L.coef, dst, L.d,
                           guard);
        cur = dst;
        k = 1;
    }
    for (; k < steps; ++k) {
        double* dst = (k == steps - 1 && out) ? out : (cur == L.xa ? L.xb : L.xa);
        const double* cf = L.coef + 2 * k;
        by_sym_vt(A->symmetric, [&](auto S) {
            hipLaunchKernelGGL((k_cheb_step<S.value, float>), dim3(Ls.grid), dim3(kBlock), 0, st, A->val32, g.ld, g.N,
                               g.s1, g.s2, Ls.chunk, L.dinv, b, cur, dst, L.d, cf, guard);
        });
        cur = dst;
    }
    RG_HIP(hipGetLastError());
    return cur;
}

}  // namespace

namespace {

// The network's correction of level l's iterate x for A_l x = b: the residual
// r' = b - A_l x (FP32 values), y = Net(r'), low-pass filtered by the level's
// Chebyshev steps on A_l y = 0 from y (they damp the rough components of y,
// whose energy would otherwise dominate the step; the error the smoothing
// left is smooth and passes), and x += alpha y with the energy-optimal
// alpha = (r'.y)/(y.A_l y) (FP64 values), which never raises the level's
// error energy.
void net_correct(rg_mg_state* s, rg_matrix* m, int l, const double* b, double* x, hipStream_t st,
                 const KrylovState* guard)
{
    MgLevel& L = s->lev[l];
    const rg_matrix* A = L.A;
    const Launch Ls = chunked_launch(A->g.N, spmv_rows_per_thread<float>());
    by_sym_vt(A->symmetric, [&](auto S) {
        hipLaunchKernelGGL((k_residual<S.value, float>), dim3(Ls.grid), dim3(kBlock), 0, st, A->val32, A->g.ld,
                           A->g.N, A->g.s1, A->g.s2, Ls.chunk, b, x, L.r, guard);
    });
    lpc_forward_balanced(m->lpc, L.r, L.dinv, nullptr, s->logk, s->y, st);
    double* yf = s->y;
    const int fsteps = std::min(m->mg_coarse_steps, kChebSteps);
    if (A->g.N <= kSmallNodes && fsteps >= 2) {
        cheb_small(s, A, l, s->zero, s->y, s->y2, fsteps, st, guard);
        yf = s->y2;
    } else for (int k = 0; k < fsteps; ++k) {
        double* dst = (yf == s->y) ? s->y2 : s->y;
        by_sym_vt(A->symmetric, [&](auto S) {
            hipLaunchKernelGGL((k_cheb_step<S.value, float>), dim3(Ls.grid), dim3(kBlock), 0, st, A->val32, A->g.ld,
                               A->g.N, A->g.s1, A->g.s2, Ls.chunk, L.dinv, s->zero, yf, dst, L.d, L.coef + 2 * k,
                               guard);
        });
        yf = dst;
    }
    const Launch Le = chunked_launch(A->g.N, spmv_rows_per_thread<double>(), 1024);
    by_sym_vt(A->symmetric, [&](auto S) {
        hipLaunchKernelGGL((k_coarse_energy<S.value>), dim3(Le.grid), dim3(kBlock), 0, st, A->val, A->g.ld, A->g.N,
                           A->g.s1, A->g.s2, Le.chunk, yf, L.r, s->part, guard);
    });
    hipLaunchKernelGGL(k_coarse_step, dim3(grid_for(A->g.N)), dim3(kBlock), 0, st, A->g.N, s->part, Le.grid, yf, x,
                       guard);
    RG_HIP(hipGetLastError());
}

}  // namespace

// Launches one V-cycle: z = M(r) on the level-0 vectors (balanced unless the
// hierarchy was set up for an unbalanced matrix; r, z distinct, z written in
// full).  Down: `degree` Chebyshev steps from zero, residual, restriction;
// the coarsest level: coarse_steps Chebyshev steps from zero; up:
// prolongation, `degree` Chebyshev steps, and at the network's level its
// correction.
void mg_vcycle(rg_matrix* m, const double* r, double* z, hipStream_t st, const KrylovState* guard)
{
    rg_mg_state* s = m->mg;
    RG_CHECK(s != nullptr, RG_STATE_ERROR, "FCG_LEARNED: no multilevel hierarchy");
    const int deg = std::max(1, std::min(m->mg_degree, kChebSteps));
    const int nl = s->nlev;
    const double* bl[kMaxLevels];
    double* xl[kMaxLevels];
    bl[0] = r;
    // down
    for (int l = 0; l + 1 < nl; ++l) {
        MgLevel& L = s->lev[l];
        const rg_matrix* A = L.A ? L.A : m;
        xl[l] = smooth(s, m, l, bl[l], nullptr, true, deg, nullptr, st, guard);
        const Launch Ls = chunked_launch(A->g.N, spmv_rows_per_thread<float>());
        by_sym_vt(A->symmetric, [&](auto S) {
            hipLaunchKernelGGL((k_residual<S.value, float>), dim3(Ls.grid), dim3(kBlock), 0, st, A->val32, A->g.ld,
                               A->g.N, A->g.s1, A->g.s2, Ls.chunk, bl[l], xl[l], L.r, guard);
        });
        MgLevel& C = s->lev[l + 1];
        if (l == 0 && s->balanced)
            hipLaunchKernelGGL(k_restrict<true>, dim3(grid1(C.g.N)), dim3(kBlock), 0, st, L.g, C.g, L.r, m->balance,
                               C.b, guard);
        else
            hipLaunchKernelGGL(k_restrict<false>, dim3(grid1(C.g.N)), dim3(kBlock), 0, st, L.g, C.g, L.r, nullptr,
                               C.b, guard);
        bl[l + 1] = C.b;
        RG_HIP(hipGetLastError());
    }
    // coarsest level
    if (nl == 1) {
        smooth(s, m, 0, r, nullptr, true, std::max(deg, m->mg_coarse_steps), z, st, guard);
        return;
    }
    xl[nl - 1] = smooth(s, m, nl - 1, bl[nl - 1], nullptr, true, std::min(m->mg_coarse_steps, kChebSteps), nullptr,
                        st, guard);
    if (s->net_level == nl - 1) net_correct(s, m, nl - 1, bl[nl - 1], xl[nl - 1], st, guard);
    // up
    for (int l = nl - 2; l >= 0; --l) {
        MgLevel& L = s->lev[l];
        const MgLevel& C = s->lev[l + 1];
        if (l == 0 && s->balanced)
            hipLaunchKernelGGL(k_prolong<true>, dim3(grid1(L.g.N)), dim3(kBlock), 0, st, L.g, C.g, xl[l + 1],
                               m->balance, xl[l], guard);
        else
            hipLaunchKernelGGL(k_prolong<false>, dim3(grid1(L.g.N)), dim3(kBlock), 0, st, L.g, C.g, xl[l + 1],
                               nullptr, xl[l], guard);
        RG_HIP(hipGetLastError());
        // post-smoothing from the corrected iterate: the Chebyshev recursion
        // restarts (its first step has no previous update)
        const rg_matrix* A = L.A ? L.A : m;
        const Geom& g = A->g;
        const Launch Ls = chunked_launch(g.N, spmv_rows_per_thread<float>());
        double* cur = xl[l];
        if (g.N <= kSmallNodes && deg >= 2) {
            double* dst = (l == 0) ? z : (cur == L.xa ? L.xb : L.xa);
            cheb_small(s, A, l, bl[l], cur, dst, deg, st, guard);
            cur = dst;
        } else for (int k = 0; k < deg; ++k) {
            double* dst = (l == 0 && k == deg - 1) ? z : (cur == L.xa ? L.xb : L.xa);
            by_sym_vt(A->symmetric, [&](auto S) {
                hipLaunchKernelGGL((k_cheb_step<S.value, float>), dim3(Ls.grid), dim3(kBlock), 0, st, A->val32, g.ld,
                                   g.N, g.s1, g.s2, Ls.chunk, L.dinv, bl[l], cur, dst, L.d, L.coef + 2 * k, guard);
            });
            cur = dst;
        }
        xl[l] = cur;
        RG_HIP(hipGetLastError());
        if (l == s->net_level) net_correct(s, m, l, bl[l], xl[l], st, guard);
    }
}

// The Galerkin operator `levels` halvings below the (unbalanced) matrix m, as
// a new matrix of the coarse Brick with m's storage.
rg_matrix* coarse_operator(rg_matrix* m, int levels)
{
    RG_CHECK(m->g.dim == 3 && m->g.bs == 1, RG_VALUE_ERROR, "coarse_operator: scalar Brick matrices only");
    RG_CHECK(!m->is_balanced, RG_STATE_ERROR, "coarse_operator: balanced matrix is not supported.");
    RG_CHECK(levels >= 1, RG_VALUE_ERROR, "coarse_operator: levels must be at least 1");
    hipStream_t st = m->ctx->stream;
    rg_matrix* F = m;
    for (int l = 0; l < levels; ++l) {
        int32_t ne[3];
        for (int d = 0; d < 3; ++d) {
            RG_CHECK(F->g.ne[d] % 2 == 0 && F->g.ne[d] >= 2, RG_VALUE_ERROR,
                     "coarse_operator: the grid cannot be halved that many times");
            ne[d] = F->g.ne[d] / 2;
        }
        rg_matrix* C = nullptr;
        const int rc = rg_matrix_create(m->ctx, 3, ne, F->g.len, 1, m->symmetric ? RG_MATRIX_SYMMETRIC : 0, &C);
        if (rc != RG_OK) {
            if (F != m) rg_matrix_destroy(F);
            throw Error(rc, "coarse_operator: cannot create the coarse matrix");
        }
        const MgGeom f = geom_of(F), c = geom_of(C);
        double* tmp = nullptr;
        try {
            RG_HIP(hipMalloc(&tmp, sizeof(double) * galerkin_tmp_doubles(f, c, C->symmetric ? 14 : 27)));
            galerkin_product(F->val, f, F->symmetric, nullptr, C->val, c, C->symmetric, tmp, st);
            RG_HIP(hipStreamSynchronize(st));
        } catch (...) {
            if (tmp) hipFree(tmp);
            rg_matrix_destroy(C);
            if (F != m) rg_matrix_destroy(F);
            throw;
        }
        hipFree(tmp);
        C->val_zero = false;
        if (F != m) rg_matrix_destroy(F);
        F = C;
    }
    return F;
}

// x = one V-cycle for A x = b on the unbalanced matrix m (its hierarchy set
// up in physical variables).
void vcycle_unbalanced(rg_matrix* m, const double* b, double* x)
{
    RG_CHECK(m->g.dim == 3 && m->g.bs == 1, RG_VALUE_ERROR, "vcycle: scalar Brick matrices only");
    RG_CHECK(!m->is_balanced, RG_STATE_ERROR, "vcycle: balanced matrix is not supported.");
    if (!m->val32_valid) {
        if (!m->val32) RG_HIP(hipMalloc(&m->val32, value_count(m->g) * sizeof(float)));
        launch_to_fp32(m);
        m->val32_valid = true;
    }
    mg_setup(m, false);
    mg_vcycle(m, b, x, m->ctx->stream, nullptr);
    RG_HIP(hipStreamSynchronize(m->ctx->stream));
}

// x = `steps` Chebyshev steps from zero on the unbalanced matrix m (FP32
// values, FP64 vectors, the bound and interval of the multilevel levels).
void smooth_unbalanced(rg_matrix* m, const double* b, double* x, int steps)
{
    RG_CHECK(m->g.dim == 3 && m->g.bs == 1, RG_VALUE_ERROR, "smooth: scalar Brick matrices only");
    RG_CHECK(!m->is_balanced, RG_STATE_ERROR, "smooth: balanced matrix is not supported.");
    RG_CHECK(steps >= 1 && steps <= kChebSteps, RG_VALUE_ERROR, "smooth: steps must be 1..16");
    hipStream_t st = m->ctx->stream;
    const Geom& g = m->g;
    if (!m->val32_valid) {
        if (!m->val32) RG_HIP(hipMalloc(&m->val32, value_count(g) * sizeof(float)));
        launch_to_fp32(m);
        m->val32_valid = true;
    }
    const MgGeom a = geom_of(m);
    const unsigned G = grid1(a.N);
    double* buf = nullptr;
    const int64_t nc = 2 * kChebSteps + 1;
    RG_HIP(hipMalloc(&buf, sizeof(double) * (G + nc + 4 * a.N)));
    double* gp = buf;
    double* coef = gp + G;
    double* dinv = coef + nc;
    double* d = dinv + a.N;
    double* xa = d + a.N;
    double* xb = xa + a.N;
    try {
        if (m->symmetric) hipLaunchKernelGGL(k_gershgorin<true>, dim3(G), dim3(kBlock), 0, st, m->val, a, gp);
        else hipLaunchKernelGGL(k_gershgorin<false>, dim3(G), dim3(kBlock), 0, st, m->val, a, gp);
        hipLaunchKernelGGL(k_cheb_table, dim3(1), dim3(kBlock), 0, st, gp, (int)G, kChebRatio, coef);
        const double* main_slot = m->val + (int64_t)(m->symmetric ? 0 : C3) * g.ld;
        hipLaunchKernelGGL(k_inv_main_slot, dim3(G), dim3(kBlock), 0, st, main_slot, a.N, dinv);
        hipLaunchKernelGGL(k_cheb_first, dim3(grid_for(g.N)), dim3(kBlock), 0, st, g.N, b, dinv, coef,
                           steps == 1 ? x : xa, d, nullptr);
        const double* cur = xa;
        const Launch Ls = chunked_launch(g.N, spmv_rows_per_thread<float>());
        for (int k = 1; k < steps; ++k) {
            double* dst = (k == steps - 1) ? x : (cur == xa ? xb : xa);
            by_sym_vt(m->symmetric, [&](auto S) {
                hipLaunchKernelGGL((k_cheb_step<S.value, float>), dim3(Ls.grid), dim3(kBlock), 0, st, m->val32, g.ld,
                                   g.N, g.s1, g.s2, Ls.chunk, dinv, b, cur, dst, d, coef + 2 * k, nullptr);
            });
            cur = dst;
        }
        RG_HIP(hipGetLastError());
        RG_HIP(hipStreamSynchronize(st));
    } catch (...) {
        hipFree(buf);
        throw;
    }
    hipFree(buf);
}

// One FCG(kDirs) iteration: z = M(r), the direction, q = A p, the updates.
void mg_fcg_iteration(rg_matrix* m, double* x, double* r, double* z, KrylovState* st_dev, hipStream_t st,
                      const double* norm_w)
{
    rg_mg_state* s = m->mg;
    mg_vcycle(m, r, z, st, st_dev);
    const Geom& g = m->g;
    FcgArgs a;
    a.x = x;
    a.r = r;
    a.z = z;
    a.P = s->P;
    a.Q = s->Q;
    a.coef = s->fcg;
    a.st = st_dev;
    a.partials = m->ctx->partials;
    a.n = g.N;
    a.vld = s->vld;
    RG_CHECK(!norm_w || (m->is_balanced && norm_w == m->balance), RG_STATE_ERROR,
             "FCG_LEARNED: the unbalanced residual norm needs the balanced matrix's balance vector");
    a.wn = norm_w ? s->ibal : nullptr;   // norm_w is the balance; ibal its reciprocal
    const int vg = 1024;
    a.chunk = 0;
    hipLaunchKernelGGL(k_fcgm_dots, dim3(vg), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(k_fcgm_beta, dim3(1), dim3(kBlock), 0, st, a, vg);
    hipLaunchKernelGGL(k_fcgm_dir, dim3(vg), dim3(kBlock), 0, st, a);
    const Launch Lsp = chunked_launch(g.N, spmv_rows_per_thread<double>());
    FcgArgs asp = a;
    asp.chunk = Lsp.chunk;
    by_sym_vt(m->symmetric, [&](auto S) {
        hipLaunchKernelGGL((k_fcgm_spmv<S.value>), dim3(Lsp.grid), dim3(kBlock), 0, st, asp, m->val, g.ld, g.s1,
                           g.s2);
    });
    hipLaunchKernelGGL(k_fcgm_alpha, dim3(1), dim3(kBlock), 0, st, a, Lsp.grid);
    hipLaunchKernelGGL(k_fcgm_update, dim3(vg), dim3(kBlock), 0, st, a);
    hipLaunchKernelGGL(k_fcgm_end, dim3(1), dim3(kBlock), 0, st, st_dev, m->ctx->partials, vg);
    RG_HIP(hipGetLastError());
}

}  // namespace rg
