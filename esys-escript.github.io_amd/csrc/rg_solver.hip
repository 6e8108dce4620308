// Solver kernels and host driver of libripley_hip: PASO's balance, Jacobi
// preconditioner, residual-smoothed PCG (and the textbook / mixed-precision
// variants) and the outer refinement loop of paso::Solver.
//
// Every Krylov scalar lives in device memory (KrylovState); each reducing
// kernel publishes per-workgroup partial sums and its last workgroup reduces
// them in a fixed order and updates the scalars, so a chunk of iterations is
// a fixed sequence of launches with no host round trip.  The chunk is
// captured once into a hipGraph and replayed; kernels return at entry once
// the state says the iteration has stopped, and the host polls the state
// between chunks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>

#include "rg_device.h"
#include "rg_dispatch.h"
#include "rg_krylov.h"
#include "rg_mg.h"
#include "rg_stencil.h"

namespace rg {

namespace {

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ------------------------------------------------------------ Jacobi apply
// BlockOps_solveAll (paso/src/BlockOps.h:219-245): z_i = inv(A_ii) r_i with
// the inverse stored structure-of-arrays dinv[(ir + bs*ic)*ld + i].
template <int BS, typename DT>
__device__ __forceinline__ void jacobi_node(const DT* __restrict__ dinv, int64_t ld, int64_t i,
                                            const double* __restrict__ r, double (&z)[BS])
{
    if (BS == 1) {
        z[0] = (double)dinv[i] * r[i];
    } else {
        double rv[BS];
#pragma unroll
        for (int c = 0; c < BS; ++c) rv[c] = r[i * BS + c];
#pragma unroll
        for (int a = 0; a < BS; ++a) {
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < BS; ++c) s = fma((double)dinv[(int64_t)(a + BS * c) * ld + i], rv[c], s);
            z[a] = s;
        }
    }
}

// z = inv(D_i) r for a node whose r values are in registers.
template <int BS, typename DT>
__device__ __forceinline__ void jacobi_regs(const DT* __restrict__ dinv, int64_t ld, int64_t i,
                                            const double (&r)[BS], double (&z)[BS])
{
#pragma unroll
    for (int a = 0; a < BS; ++a) {
        double s = 0.0;
#pragma unroll
        for (int c = 0; c < BS; ++c) s = fma((double)dinv[(int64_t)(a + BS * c) * ld + i], r[c], s);
        z[a] = s;
    }
}

// z = M^-1 r for node i: the stored sweep result if zbuf is given, else one
// Jacobi sweep computed on the fly.
template <int BS, typename DT>
__device__ __forceinline__ void prec_node(const DT* dinv, int64_t ld, int64_t i, const double* r,
                                          const double* zbuf, double (&z)[BS])
{
    if (zbuf) {
#pragma unroll
        for (int c = 0; c < BS; ++c) z[c] = zbuf[i * BS + c];
    } else {
        jacobi_node<BS>(dinv, ld, i, r, z);
    }
}

// ----------------------------------------------------------------- balance
// SystemMatrix::balance (paso/src/SystemMatrix.cpp:262-312) with
// SparseMatrix::maxAbsRow_CSR_OFFSET0 (SparseMatrix.cpp:266-283):
// d[i r] = sqrt(1/max_(j,c) |a(i r, j c)|), or 1 for a zero row.
template <int DIM, int BS, bool SYM>
__global__ __launch_bounds__(kBlock) void k_balance_vector(const double* __restrict__ val, int64_t ld,
                                                           int64_t N, int64_t s1, int64_t s2,
                                                           double* __restrict__ d)
{
    constexpr int NS = StencilShape<DIM>::NS;
    constexpr int C = StencilShape<DIM>::CENTER;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    double fac[BS];
#pragma unroll
    for (int r = 0; r < BS; ++r) fac[r] = 0.0;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        int dx, dy, dz;
        slot_delta<DIM>(s, dx, dy, dz);
        const int64_t j = i + dx + dy * s1 + dz * s2;
        if ((uint64_t)j >= (uint64_t)N) continue;
#pragma unroll
        for (int c = 0; c < BS; ++c)
#pragma unroll
            for (int r = 0; r < BS; ++r) {
                double a;
                if (!SYM || s >= C) a = val[(int64_t)((SYM ? s - C : s) * BS * BS + r + BS * c) * ld + i];
                else a = val[(int64_t)((C - s) * BS * BS + c + BS * r) * ld + j];
                fac[r] = fmax(fac[r], fabs(a));
            }
    }
#pragma unroll
    for (int r = 0; r < BS; ++r) d[i * BS + r] = (fac[r] > 0.0) ? sqrt(1.0 / fac[r]) : 1.0;
}

// SparseMatrix::applyDiagonal_CSR_OFFSET0 (SparseMatrix.cpp:361-384):
// a(i r, j c) *= d[i r] * d[j c].
template <int DIM, int BS, bool SYM>
__global__ __launch_bounds__(kBlock) void k_apply_balance(double* __restrict__ val, int64_t ld, int64_t N,
                                                          int64_t s1, int64_t s2,
                                                          const double* __restrict__ d)
{
    constexpr int NST = stored_slots<DIM, SYM>();
    constexpr int FIRST = SYM ? StencilShape<DIM>::CENTER : 0;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    double di[BS];
#pragma unroll
    for (int r = 0; r < BS; ++r) di[r] = d[i * BS + r];
#pragma unroll
    for (int t = 0; t < NST; ++t) {
        int dx, dy, dz;
        slot_delta<DIM>(FIRST + t, dx, dy, dz);
        const int64_t j = i + dx + dy * s1 + dz * s2;
        if ((uint64_t)j >= (uint64_t)N) continue;
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const double dj = d[j * BS + c];
#pragma unroll
            for (int r = 0; r < BS; ++r) {
                const int64_t idx = (int64_t)(t * BS * BS + r + BS * c) * ld + i;
                val[idx] *= di[r] * dj;
            }
        }
    }
}

// SparseMatrix::invMain (SparseMatrix.cpp:387-435) with the closed-form
// inverses BlockOps_invM_2/3 (BlockOps.h:100-145); a singular block raises
// *failed.
template <int BS, bool SYM, int DIM>
__global__ __launch_bounds__(kBlock) void k_inv_main(const double* __restrict__ val, int64_t ld, int64_t N,
                                                     double* __restrict__ dinv, float* __restrict__ dinv32,
                                                     int* failed)
{
    constexpr int T = SYM ? 0 : StencilShape<DIM>::CENTER;   // stored slot of the main diagonal
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    double A[BS * BS], I[BS * BS];
#pragma unroll
    for (int e = 0; e < BS * BS; ++e) A[e] = val[(int64_t)(T * BS * BS + e) * ld + i];
    bool ok = true;
    if (BS == 1) {
        ok = fabs(A[0]) > 0.0;
        I[0] = 1.0 / A[0];
    } else if (BS == 2) {
        const double A11 = A[0], A12 = A[2], A21 = A[1], A22 = A[3];
        double D = A11 * A22 - A12 * A21;
        ok = fabs(D) > 0.0;
        D = 1.0 / D;
        I[0] = A22 * D; I[1] = -A21 * D; I[2] = -A12 * D; I[3] = A11 * D;
    } else if (BS == 3) {
        const double A11 = A[0], A21 = A[1], A31 = A[2], A12 = A[3], A22 = A[4], A32 = A[5], A13 = A[6],
                     A23 = A[7], A33 = A[8];
        double D = A11 * (A22 * A33 - A23 * A32) + A12 * (A31 * A23 - A21 * A33) + A13 * (A21 * A32 - A31 * A22);
        ok = fabs(D) > 0.0;
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
    }
    if (!ok) {
        *failed = 1;
        return;
    }
#pragma unroll
    for (int e = 0; e < BS * BS; ++e) {
        dinv[(int64_t)e * ld + i] = I[e];
        dinv32[(int64_t)e * ld + i] = (float)I[e];
    }
}

// ------------------------------------------------------- vector utilities
__global__ void k_mul(double* __restrict__ out, const double* __restrict__ a, const double* __restrict__ b,
                      int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        out[i] = a[i] * b[i];
}

__global__ void k_axpy(double* __restrict__ y, double a, const double* __restrict__ x, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        y[i] += a * x[i];
}

__global__ void k_fill(double* __restrict__ y, double v, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        y[i] = v;
}

// ||v||_2^2 and ||v||_inf in one pass (Solver.cpp:106-130, :201-225); with
// d given, of v / d (the unbalanced residual of a balanced one).  Block sums
// go to partial row 0, block maxima to row 1; k_norms_stage combines them.
__global__ __launch_bounds__(kBlock) void k_norms(const double* __restrict__ v, const double* __restrict__ d,
                                                  int64_t n, double* partials)
{
    double s[1] = {0.0};
    double mx = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double a = d ? v[i] / d[i] : v[i];
        s[0] += a * a;
        mx = fmax(mx, fabs(a));
    }
    mx = wave_max(mx);
    __shared__ double smax[kBlock / 64];
    if ((threadIdx.x & 63) == 0) smax[threadIdx.x >> 6] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) m = fmax(m, smax[w]);
        partials[gridDim.x + blockIdx.x] = m;
    }
    publish_block_sums<1>(s, partials);
}

__global__ __launch_bounds__(kBlock) void k_norms_stage(const double* __restrict__ partials, int G, double* out)
{
    double t[1];
    reduce_partials<1>(partials, G, t);
    double m = 0.0;
    for (int b = threadIdx.x; b < G; b += kBlock) m = fmax(m, partials[G + b]);
    m = wave_max(m);
    __shared__ double sm[kBlock / 64];
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double mm = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) mm = fmax(mm, sm[w]);
        out[0] = t[0];
        out[1] = mm;
    }
}

// ------------------------------------------------ PASO residual-smoothed PCG
// Solver_PCG (paso/src/PCG.cpp:62-368).  Iteration k:
//   (1)-(3) v = M^-1 r, tau = v.r, p = v + (tau/tau_old) p      k_paso_p
//   (4)-(5) v = A p, delta = p.v, alpha = tau/delta             k_paso_spmv
//   (6)     r -= alpha v, d = r - rs, s3 = d.d, s4 = d.rs     k_paso_r
//   (6)-(7) x2 += alpha p, gamma = -s4/s3,
//           rs = (1-gamma) rs + gamma r,
//           x = (1-gamma) x + gamma x2, s5 = rs.rs              k_paso_x
//   (8)     stop if sqrt(s5) <= tol, iter > maxit or tau == 0
// tau of iteration k+1 is reduced in k_paso_x of iteration k (r is final
// there), so each iteration is four launches.  Streaming kernels of block
// size 1 move two entries per thread with 16-byte accesses.
struct PcgArgs {
    double *r, *x, *rs, *p, *v, *x2;
    double *zp, *z2p, *s;   // BiCGStab: phat, shat, s
    const double* z;        // stored M^-1 r (Jacobi sweeps > 1) or nullptr
    const double* b;        // right-hand side of a residual replacement
    KrylovState* st;
    double* partials;
    int64_t N, ld, s1, s2, chunk;
    const double* fold_partials;   // k_p_update<.., FOLD>: the update kernel's partials (2 x fold_G)
    int fold_G;
    const double* wn;       // k_cg_init: the residual norm is taken of r / wn (nullptr: of r)
};

// Loop over the node range of this block, two nodes per thread for BS == 1.
// f(i, two) is called with the first node of a pair (two == true) or a single
// node (two == false).
template <int BS, class F>
__device__ __forceinline__ void node_loop(const RowRange& rr, F&& f)
{
    if (BS == 1) {
        for (int64_t i = rr.r0 + 2 * (int64_t)threadIdx.x; i < rr.r1; i += 2 * (int64_t)kBlock)
            f(i, i + 1 < rr.r1);
    } else {
        for (int64_t i = rr.r0 + threadIdx.x; i < rr.r1; i += kBlock) f(i, false);
    }
}

template <typename T>
struct Two {
    T a, b;
};
// Vector streams of the Krylov update kernels, cached or non-temporal (NT).
// NT keeps a matrix resident in the 256 MB memory-side cache (MALL) when its
// values fit there but values plus the vectors streamed per iteration do
// not (stream_policy): measured on the 128^3 Brick, FP64 symmetric storage
// (240 MB of values) PCG 0.171 -> 0.154 s and PASO PCG 0.221 -> 0.200 s;
// FP32 symmetric values (120 MB, everything fits) 0.158 -> 0.166 s, full
// storage within 1 %.
template <bool NT>
__device__ __forceinline__ Two<double> ld2(const double* p, bool two)
{
    if (two) {
        if constexpr (NT) {
            const rg_dbl2 w = __builtin_nontemporal_load(reinterpret_cast<const rg_dbl2*>(p));
            return {w[0], w[1]};
        } else {
            const double2 v = *reinterpret_cast<const double2*>(p);
            return {v.x, v.y};
        }
    }
    return {p[0], 0.0};
}
template <bool NT>
__device__ __forceinline__ Two<double> ld2(const float* p, bool two)
{
    if (two) {
        const float2 v = *reinterpret_cast<const float2*>(p);
        return {(double)v.x, (double)v.y};
    }
    return {(double)p[0], 0.0};
}
template <bool NT>
__device__ __forceinline__ void st2(double* p, Two<double> v, bool two)
{
    if (two) {
        if constexpr (NT) __builtin_nontemporal_store(rg_dbl2{v.a, v.b}, reinterpret_cast<rg_dbl2*>(p));
        else *reinterpret_cast<double2*>(p) = make_double2(v.a, v.b);
    } else {
        p[0] = v.a;
    }
}

template <bool NT>
__device__ __forceinline__ double ld1(const double* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ __forceinline__ void st1(double* p, double v)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <int BS, typename DT>
__global__ __launch_bounds__(kBlock) void k_paso_init(PcgArgs a, const DT* __restrict__ dinv)
{
    const RowRange rr = block_rows(a.N, a.chunk);
    double* __restrict__ r = a.r;
    double* __restrict__ rs = a.rs;
    double* __restrict__ x2 = a.x2;
    double* __restrict__ p = a.p;
    const double* __restrict__ x = a.x;
    double t[1] = {0.0};
    for (int64_t i = rr.r0 + threadIdx.x; i < rr.r1; i += kBlock) {
        double z[BS];
        prec_node<BS>(dinv, a.ld, i, r, a.z, z);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const int64_t k = i * BS + c;
            const double rk = r[k];
            rs[k] = rk;
            x2[k] = x[k];
            p[k] = 0.0;
            t[0] += z[c] * rk;
        }
    }
    publish_block_sums<1>(t, a.partials);
}

// p = M^-1 r + beta p (paso k_paso_p and textbook k_cg_p).  With XD the
// textbook PCG's x += alpha p of the same iteration happens here instead of
// in k_cg_update (the p read is shared: one vector pass less per
// iteration); it is pending (st->early) after every update stage, also the
// one that ends the solve, when p itself is left alone.
//
// With FOLD (and XD) there is no update stage before this kernel: every
// block reduces the update kernel's partials itself (fixed order, the same
// totals in every block), takes the stopping tests of cg_end for p and
// beta = r.z / rho, and block 0 leaves the totals for cg_commit at the next
// stage; the x update is then always this iteration's.
template <int BS, typename DT, bool XD = false, bool NT = false, bool FOLD = false>
__global__ __launch_bounds__(kBlock) void k_p_update(PcgArgs a, const DT* __restrict__ dinv)
{
    static_assert(!FOLD || XD, "the folded update stage defers x");
    if (FOLD ? a.st->done : (XD ? !a.st->early : a.st->done)) return;
    bool upd = !XD || !a.st->done;
    double beta = a.st->beta;
    const double alpha = a.st->alpha;
    if constexpr (FOLD) {
        double tot[2];
        reduce_partials_wide<2, kBlock>(a.fold_partials, a.fold_G, tot);
        const KrylovState* st = a.st;
        const double rz = tot[0], rnorm2 = tot[1];
        const bool fin = (sqrt(rnorm2) <= st->tol && !st->replace_mode) || st->num_iter + 1 >= st->maxit ||
                         fabs(rz) <= 0.0;
        upd = !fin;
        beta = rz / st->rho;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            a.st->pend_rz = rz;
            a.st->pend_rr = rnorm2;
            a.st->pend = 1;
        }
    }
    const RowRange rr = block_rows(a.N, a.chunk);
    const double* __restrict__ r = a.r;
    const double* __restrict__ zb = a.z;
    double* __restrict__ p = a.p;
    double* __restrict__ x = a.x;
    node_loop<BS>(rr, [&](int64_t i, bool two) {
        if (BS == 1) {
            const Two<double> pv = ld2<NT>(p + i, two);
            if (XD) {
                const Two<double> xv = ld2<NT>(x + i, two);
                st2<NT>(x + i, {fma(alpha, pv.a, xv.a), fma(alpha, pv.b, xv.b)}, two);
            }
            if (upd) {
                Two<double> z;
                if (zb) z = ld2<NT>(zb + i, two);
                else {
                    const Two<double> rv = ld2<NT>(r + i, two), dv = ld2<NT>(dinv + i, two);
                    z = {dv.a * rv.a, dv.b * rv.b};
                }
                st2<NT>(p + i, {fma(beta, pv.a, z.a), fma(beta, pv.b, z.b)}, two);
            }
        } else {
            if (XD) {
#pragma unroll
                for (int c = 0; c < BS; ++c) x[i * BS + c] = fma(alpha, p[i * BS + c], x[i * BS + c]);
            }
            if (upd) {
                double z[BS];
                prec_node<BS>(dinv, a.ld, i, r, zb, z);
#pragma unroll
                for (int c = 0; c < BS; ++c) p[i * BS + c] = fma(beta, p[i * BS + c], z[c]);
            }
        }
    });
}

// v = A p and the block partials of p.v (alpha = (tau or rho) / p.v follows
// in the stage).  A non-temporal store of v under the stream policy measured
// the same as this cached one.
template <int DIM, int BS, bool SYM, typename VT, bool PASO>
__global__ __launch_bounds__(kBlock) void k_spmv_dot(PcgArgs a, const VT* __restrict__ val)
{
    if (a.st->done) return;
    const RowRange rr = block_rows(a.N, a.chunk);
    const double* __restrict__ p = a.p;
    double* __restrict__ v = a.v;
    double dl[1] = {0.0};
    stencil_block<DIM, BS, SYM, VT>(val, a.ld, a.N, a.s1, a.s2, rr.r0, rr.r1, p,
                                    [&](int64_t i, const double (&acc)[BS]) {
#pragma unroll
                                        for (int r = 0; r < BS; ++r) {
                                            v[i * BS + r] = acc[r];
                                            dl[0] = fma(acc[r], p[i * BS + r], dl[0]);
                                        }
                                    });
    publish_block_sums<1>(dl, a.partials);
}

// a.N counts scalar entries here (nodes * block size).
template <bool NT = false>
__global__ __launch_bounds__(kBlock) void k_paso_r(PcgArgs a)
{
    if (a.st->done) return;
    const double alpha = a.st->alpha;
    const RowRange rr = block_rows(a.N, a.chunk);
    double* __restrict__ r = a.r;
    const double* __restrict__ rs = a.rs;
    const double* __restrict__ v = a.v;
    double s[2] = {0.0, 0.0};
    node_loop<1>(rr, [&](int64_t i, bool two) {
        const Two<double> rv = ld2<NT>(r + i, two), vv = ld2<NT>(v + i, two), rsv = ld2<NT>(rs + i, two);
        const Two<double> rn = {fma(-alpha, vv.a, rv.a), fma(-alpha, vv.b, rv.b)};
        st2<NT>(r + i, rn, two);
        const double d0 = rn.a - rsv.a, d1 = two ? rn.b - rsv.b : 0.0;
        s[0] += d0 * d0 + d1 * d1;
        s[1] += d0 * rsv.a + (two ? d1 * rsv.b : 0.0);
    });
    publish_block_sums<2>(s, a.partials);
}

// PCG.cpp's x2 += alpha p of the same iteration happens here (not in
// k_paso_r, which then streams three vectors instead of five; p and x2 are
// read here once).  With FOLD there is no stage between k_paso_r and this
// kernel: every block reduces k_paso_r's partials itself (fixed order, the
// same totals in every block) and forms gamma as ST_PASO_R would.
template <int BS, typename DT, bool NT = false, bool FOLD = false>
__global__ __launch_bounds__(kBlock) void k_paso_x(PcgArgs a, const DT* __restrict__ dinv)
{
    if (a.st->done) return;
    const double alpha = a.st->alpha;
    double g1 = a.st->gamma1;
    if constexpr (FOLD) {
        double tot[2];
        reduce_partials_wide<2, kBlock>(a.fold_partials, a.fold_G, tot);
        g1 = (fabs(tot[0]) <= 0.0) ? 0.0 : -tot[1] / tot[0];
    }
    const double g2 = 1.0 - g1;
    const RowRange rr = block_rows(a.N, a.chunk);
    const double* __restrict__ r = a.r;
    const double* __restrict__ p = a.p;
    double* __restrict__ x2 = a.x2;
    const double* __restrict__ zb = a.z;
    double* __restrict__ rs = a.rs;
    double* __restrict__ x = a.x;
    double s[2] = {0.0, 0.0};
    node_loop<BS>(rr, [&](int64_t i, bool two) {
        if (BS == 1) {
            const Two<double> rv = ld2<NT>(r + i, two), rsv = ld2<NT>(rs + i, two), xv = ld2<NT>(x + i, two);
            const Two<double> pv = ld2<NT>(p + i, two), x2o = ld2<NT>(x2 + i, two);
            const Two<double> x2v = {fma(alpha, pv.a, x2o.a), fma(alpha, pv.b, x2o.b)};
            st2<NT>(x2 + i, x2v, two);
            Two<double> z;
            if (zb) z = ld2<NT>(zb + i, two);
            else {
                const Two<double> dv = ld2<NT>(dinv + i, two);
                z = {dv.a * rv.a, dv.b * rv.b};
            }
            const Two<double> rsn = {g2 * rsv.a + g1 * rv.a, g2 * rsv.b + g1 * rv.b};
            st2<NT>(rs + i, rsn, two);
            st2<NT>(x + i, {g2 * xv.a + g1 * x2v.a, g2 * xv.b + g1 * x2v.b}, two);
            s[0] += rsn.a * rsn.a + (two ? rsn.b * rsn.b : 0.0);
            s[1] += z.a * rv.a + (two ? z.b * rv.b : 0.0);
        } else {
            double z[BS];
            prec_node<BS>(dinv, a.ld, i, r, zb, z);
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                const int64_t k = i * BS + c;
                const double rk = r[k];
                const double rsk = g2 * rs[k] + g1 * rk;
                const double x2k = fma(alpha, p[k], x2[k]);
                x2[k] = x2k;
                rs[k] = rsk;
                x[k] = g2 * x[k] + g1 * x2k;
                s[0] += rsk * rsk;
                s[1] += z[c] * rk;
            }
        }
    });
    publish_block_sums<2>(s, a.partials);
}

// ------------------------------------------------------ textbook Jacobi-PCG
//   q = A p, delta = p.q, alpha = rho/delta                  k_spmv_dot
//   x += alpha p, r -= alpha q, rho' = r.M^-1 r, ||r||       k_cg_update
//   p = M^-1 r + (rho'/rho) p                                k_p_update
// Stops when ||r||_2 <= tol (or iter >= maxit, or a zero scalar).
template <int BS, typename DT>
__global__ __launch_bounds__(kBlock) void k_cg_init(PcgArgs a, const DT* __restrict__ dinv)
{
    const RowRange rr = block_rows(a.N, a.chunk);
    const double* __restrict__ r = a.r;
    double* __restrict__ p = a.p;
    double s[2] = {0.0, 0.0};
    for (int64_t i = rr.r0 + threadIdx.x; i < rr.r1; i += kBlock) {
        double z[BS];
        prec_node<BS>(dinv, a.ld, i, r, a.z, z);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const int64_t k = i * BS + c;
            const double rk = r[k];
            p[k] = z[c];
            s[0] += rk * z[c];
            const double rw = a.wn ? rk / a.wn[k] : rk;
            s[1] += rw * rw;
        }
    }
    publish_block_sums<2>(s, a.partials);
}

// With XD the x update is left to k_p_update<.., true>.
template <int BS, typename DT, bool XD = false, bool NT = false>
__global__ __launch_bounds__(kBlock) void k_cg_update(PcgArgs a, const DT* __restrict__ dinv)
{
    if (a.st->done) return;
    const double alpha = a.st->alpha;
    const RowRange rr = block_rows(a.N, a.chunk);
    double* __restrict__ x = a.x;
    double* __restrict__ r = a.r;
    const double* __restrict__ p = a.p;
    const double* __restrict__ v = a.v;
    const bool have_z = a.z != nullptr;
    double s[2] = {0.0, 0.0};
    node_loop<BS>(rr, [&](int64_t i, bool two) {
        if (BS == 1) {
            const Two<double> rv = ld2<NT>(r + i, two), vv = ld2<NT>(v + i, two), dv = ld2<NT>(dinv + i, two);
            if (!XD) {
                const Two<double> xv = ld2<NT>(x + i, two), pv = ld2<NT>(p + i, two);
                st2<NT>(x + i, {fma(alpha, pv.a, xv.a), fma(alpha, pv.b, xv.b)}, two);
            }
            const Two<double> rn = {fma(-alpha, vv.a, rv.a), fma(-alpha, vv.b, rv.b)};
            st2<NT>(r + i, rn, two);
            s[0] += dv.a * rn.a * rn.a + (two ? dv.b * rn.b * rn.b : 0.0);
            s[1] += rn.a * rn.a + (two ? rn.b * rn.b : 0.0);
        } else {
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                const int64_t k = i * BS + c;
                if (!XD) x[k] = fma(alpha, p[k], x[k]);
                r[k] = fma(-alpha, v[k], r[k]);
            }
            double z[BS];
            jacobi_node<BS>(dinv, a.ld, i, r, z);
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                const double rk = r[i * BS + c];
                s[0] += rk * z[c];
                s[1] += rk * rk;
            }
        }
    });
    publish_block_sums<2>(s, a.partials);
}

// r.z for the sweeps > 1 path of the textbook PCG (z stored).
__global__ __launch_bounds__(kBlock) void k_cg_rz(PcgArgs a, int64_t n)
{
    if (a.st->done) return;
    double s[1] = {0.0};
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
        s[0] += a.r[k] * a.z[k];
    publish_block_sums<1>(s, a.partials);
}

// Residual replacement of the mixed-precision PCG: r = b - A x with the FP64
// matrix, then r.M^-1 r and ||r|| as at the end of an iteration.
template <int DIM, int BS, bool SYM, typename DT>
__global__ __launch_bounds__(kBlock) void k_cg_replace(PcgArgs a, const double* __restrict__ val,
                                                       const DT* __restrict__ dinv)
{
    if (a.st->done) return;
    const RowRange rr = block_rows(a.N, a.chunk);
    const double* __restrict__ b = a.b;
    double* __restrict__ r = a.r;
    double s[2] = {0.0, 0.0};
    stencil_block<DIM, BS, SYM, double>(val, a.ld, a.N, a.s1, a.s2, rr.r0, rr.r1, a.x,
                                        [&](int64_t i, const double (&acc)[BS]) {
                                            double rv[BS];
#pragma unroll
                                            for (int c = 0; c < BS; ++c) {
                                                rv[c] = b[i * BS + c] - acc[c];
                                                r[i * BS + c] = rv[c];
                                            }
                                            double z[BS];
                                            jacobi_node<BS>(dinv, a.ld, i, r, z);
#pragma unroll
                                            for (int c = 0; c < BS; ++c) {
                                                s[0] += rv[c] * z[c];
                                                s[1] += rv[c] * rv[c];
                                            }
                                        });
    publish_block_sums<2>(s, a.partials);
}

// ------------------------------------------------------------ BiCGStab
// Solver_BiCGStab (paso/src/BiCGStab.cpp:77-257), Jacobi preconditioner,
// vectors: rtld = rs, p, v, t = x2, phat = zp, shat = z2p, s.
//   rho = rtld.r                                   k_bicg_rho     -> ST_BICG_RHO
//   p = r + beta (p - omega v), phat = M^-1 p        k_bicg_p
//   v = A phat, rtld.v                              k_bicg_spmv_v  -> ST_BICG_ALPHA
//   r -= alpha v, s = r, shat = M^-1 s, ||s||        k_bicg_s       -> ST_BICG_S
//   (converged on ||s||: x += alpha phat)           k_bicg_xa      -> ST_BICG_XA
//   t = A shat, t.s, t.t                            k_bicg_spmv_t  -> ST_BICG_OMEGA
//   x += alpha phat + omega shat, r = s - omega t    k_bicg_update  -> ST_BICG_END
template <int BS>
__global__ __launch_bounds__(kBlock) void k_bicg_init(PcgArgs a)
{
    const RowRange rr = block_rows(a.N * BS, a.chunk);
    for (int64_t k = rr.r0 + threadIdx.x; k < rr.r1; k += kBlock) {
        a.rs[k] = a.r[k];
        a.p[k] = 0.0;
        a.v[k] = 0.0;
    }
}

template <bool NT = false>
__global__ __launch_bounds__(kBlock) void k_bicg_rho(PcgArgs a, int64_t n)
{
    if (a.st->done) return;
    double s[1] = {0.0};
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
        s[0] += ld1<NT>(a.rs + k) * ld1<NT>(a.r + k);
    publish_block_sums<1>(s, a.partials);
}

template <int BS, bool NT = false>
__global__ __launch_bounds__(kBlock) void k_bicg_p(PcgArgs a, const double* __restrict__ dinv)
{
    if (a.st->done) return;
    const double beta = a.st->beta, omega = a.st->gamma1;
    const bool first = a.st->num_iter <= 1;
    const RowRange rr = block_rows(a.N, a.chunk);
    for (int64_t i = rr.r0 + threadIdx.x; i < rr.r1; i += kBlock) {
        double pv[BS];
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const int64_t k = i * BS + c;
            const double rk = ld1<NT>(a.r + k);
            pv[c] = first ? rk : rk + beta * (ld1<NT>(a.p + k) - omega * ld1<NT>(a.v + k));
            st1<NT>(a.p + k, pv[c]);
        }
        if (dinv) {   // one fused Jacobi sweep; otherwise phat is computed afterwards
            double z[BS];
            jacobi_regs<BS>(dinv, a.ld, i, pv, z);
#pragma unroll
            for (int c = 0; c < BS; ++c) st1<NT>(a.zp + i * BS + c, z[c]);
        }
    }
}

// out = A in (stencil), partials: MODE 0: rtld.out; MODE 1: out.s, out.out
template <int DIM, int BS, bool SYM, int MODE>
__global__ __launch_bounds__(kBlock) void k_bicg_spmv(PcgArgs a, const double* __restrict__ val,
                                                      const double* __restrict__ in, double* __restrict__ out)
{
    if (a.st->done) return;
    const RowRange rr = block_rows(a.N, a.chunk);
    double s[2] = {0.0, 0.0};
    stencil_block<DIM, BS, SYM, double>(val, a.ld, a.N, a.s1, a.s2, rr.r0, rr.r1, in,
                                        [&](int64_t i, const double (&acc)[BS]) {
#pragma unroll
                                            for (int c = 0; c < BS; ++c) {
                                                const int64_t k = i * BS + c;
                                                out[k] = acc[c];
                                                if (MODE == 0) {
                                                    s[0] += a.rs[k] * acc[c];
                                                } else {
                                                    s[0] += acc[c] * a.s[k];
                                                    s[1] += acc[c] * acc[c];
                                                }
                                            }
                                        });
    publish_block_sums<2>(s, a.partials);
}

template <int BS, bool NT = false>
__global__ __launch_bounds__(kBlock) void k_bicg_s(PcgArgs a, const double* __restrict__ dinv)
{
    if (a.st->done) return;
    const double alpha = a.st->alpha;
    const RowRange rr = block_rows(a.N, a.chunk);
    double s[1] = {0.0};
    for (int64_t i = rr.r0 + threadIdx.x; i < rr.r1; i += kBlock) {
        double sv[BS];
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const int64_t k = i * BS + c;
            const double rv = ld1<NT>(a.r + k) - alpha * ld1<NT>(a.v + k);
            st1<NT>(a.r + k, rv);
            st1<NT>(a.s + k, rv);
            sv[c] = rv;
            s[0] += rv * rv;
        }
        if (dinv) {
            double z[BS];
            jacobi_regs<BS>(dinv, a.ld, i, sv, z);
#pragma unroll
            for (int c = 0; c < BS; ++c) st1<NT>(a.z2p + i * BS + c, z[c]);
        }
    }
    publish_block_sums<1>(s, a.partials);
}

// x += alpha phat after convergence on ||s|| (BiCGStab.cpp:181-186)
template <bool NT = false>
__global__ __launch_bounds__(kBlock) void k_bicg_xa(PcgArgs a, int64_t n)
{
    if (!a.st->early) return;
    const double alpha = a.st->alpha;
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock)
        st1<NT>(a.x + k, ld1<NT>(a.x + k) + alpha * ld1<NT>(a.zp + k));
}

template <bool NT = false>
__global__ __launch_bounds__(kBlock) void k_bicg_update(PcgArgs a, int64_t n)
{
    if (a.st->done) return;
    const double alpha = a.st->alpha, omega = a.st->gamma1;
    double s[1] = {0.0};
    for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBlock) {
        st1<NT>(a.x + k, ld1<NT>(a.x + k) + (alpha * ld1<NT>(a.zp + k) + omega * ld1<NT>(a.z2p + k)));
        const double rv = ld1<NT>(a.s + k) - omega * ld1<NT>(a.x2 + k);
        st1<NT>(a.r + k, rv);
        s[0] += rv * rv;
    }
    publish_block_sums<1>(s, a.partials);
}

// ------------------------------------------------------------ MINRES
// Solver_MINRES (paso/src/MINRES.cpp:59-191); vectors: r, r_old = rs,
// Az = v, w = p, w_old = x2, z = M^-1 r (unscaled) and its scaled copy
// z / gamma = s.
//   z = M^-1 r, r.z, r.r                                  k_mr_init  -> ST_MR_INIT
//   s = z / gamma, Az = A s, Az.s                          k_mr_spmv  -> ST_MR_DELTA
//   r_old, r = r, Az - delta/gamma r - gamma/gamma_old r_old,
//   z = M^-1 r (fused Jacobi), r.z                         k_mr_r     -> ST_MR_GAMMA
//   w_old, w = w, (s - a3 w_old - a2 w) / a1, x += c eta w  k_mr_w
// Terms of the first two iterations that the reference skips are skipped
// (not multiplied by zero): the buffers hold stale values then.
template <int BS>
__global__ __launch_bounds__(kBlock) void k_mr_init(PcgArgs a, const double* __restrict__ dinv)
{
    const RowRange rr = block_rows(a.N, a.chunk);
    const double* __restrict__ r = a.r;
    double s[2] = {0.0, 0.0};
    for (int64_t i = rr.r0 + threadIdx.x; i < rr.r1; i += kBlock) {
        double z[BS];
        prec_node<BS>(dinv, a.ld, i, r, a.z, z);
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const int64_t k = i * BS + c;
            if (!a.z) a.zp[k] = z[c];
            s[0] += r[k] * z[c];
            s[1] += r[k] * r[k];
        }
    }
    publish_block_sums<2>(s, a.partials);
}

template <int DIM, int BS, bool SYM>
__global__ __launch_bounds__(kBlock) void k_mr_spmv(PcgArgs a, const double* __restrict__ val)
{
    if (a.st->done) return;
    const double ig = 1.0 / a.st->rho;
    const RowRange rr = block_rows(a.N, a.chunk);
    const double* __restrict__ z = a.zp;
    double* __restrict__ az = a.v;
    double* __restrict__ zs = a.s;
    double d[1] = {0.0};
    stencil_block<DIM, BS, SYM, double>(val, a.ld, a.N, a.s1, a.s2, rr.r0, rr.r1, z,
                                        [&](int64_t i, const double (&acc)[BS]) {
#pragma unroll
                                            for (int c = 0; c < BS; ++c) {
                                                const int64_t k = i * BS + c;
                                                const double zk = z[k] * ig, azk = acc[c] * ig;
                                                zs[k] = zk;
                                                az[k] = azk;
                                                d[0] = fma(azk, zk, d[0]);
                                            }
                                        });
    publish_block_sums<1>(d, a.partials);
}

template <int BS, bool NT = false>
__global__ __launch_bounds__(kBlock) void k_mr_r(PcgArgs a, const double* __restrict__ dinv)
{
    if (a.st->done) return;
    const double g = a.st->rho, f1 = a.st->delta / g;
    const bool second = a.st->num_iter > 0;
    const double f2 = second ? g / a.st->rho_prev : 0.0;
    const RowRange rr = block_rows(a.N, a.chunk);
    double s[1] = {0.0};
    for (int64_t i = rr.r0 + threadIdx.x; i < rr.r1; i += kBlock) {
        double rn[BS];
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            const int64_t k = i * BS + c;
            const double rk = ld1<NT>(a.r + k);
            double v = fma(-f1, rk, ld1<NT>(a.v + k));
            if (second) v = fma(-f2, ld1<NT>(a.rs + k), v);
            st1<NT>(a.rs + k, rk);
            st1<NT>(a.r + k, v);
            rn[c] = v;
        }
        if (dinv) {   // one fused Jacobi sweep; otherwise z is computed afterwards
            double z[BS];
            jacobi_regs<BS>(dinv, a.ld, i, rn, z);
#pragma unroll
            for (int c = 0; c < BS; ++c) {
                a.zp[i * BS + c] = z[c];
                s[0] += rn[c] * z[c];
            }
        }
    }
    publish_block_sums<1>(s, a.partials);
}

template <bool NT = false>
__global__ __launch_bounds__(kBlock) void k_mr_w(PcgArgs a, int64_t n)
{
    if (!a.st->early) return;
    const double ia1 = 1.0 / a.st->a1, a2 = a.st->a2, a3 = a.st->a3, ceta = a.st->ceta;
    const int it = a.st->num_iter - 1;   // the iteration whose update this is
    const bool k1 = it > 0, k2 = it > 1;
    for (int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x; q < n; q += (int64_t)gridDim.x * kBlock) {
        double wv = ld1<NT>(a.s + q);
        if (k2) wv = fma(-a3, ld1<NT>(a.x2 + q), wv);   // w_ancient = w_old
        if (k1) {
            const double w = ld1<NT>(a.p + q);          // w_old = w
            wv = fma(-a2, w, wv);
            st1<NT>(a.x2 + q, w);
        }
        wv *= ia1;
        st1<NT>(a.p + q, wv);
        st1<NT>(a.x + q, fma(ceta, wv, ld1<NT>(a.x + q)));
    }
}

// ------------------------------------------------------------ stage kernels
// One workgroup after each reducing kernel (recurrences in rg_krylov.h).
constexpr int kStageThreads = 1024;

// The done flag is loaded while the partials are in flight (one memory
// round trip instead of two dependent ones; the partials of a finished solve
// are reduced and ignored), which also brings the state's first cache line,
// holding the common scalars, into L2 for thread 0's recurrence.
template <int STAGE, int NV>
__global__ __launch_bounds__(kStageThreads) void k_stage(KrylovState* st, const double* __restrict__ partials, int G)
{
    int done = 0;
    double tot[NV];
    reduce_partials_wide<NV, kStageThreads>(partials, G, tot, [&] { done = st->done; });
    if (threadIdx.x != 0) return;
    if (!stage_starts<STAGE>() && done) return;
    stage_apply<STAGE, NV>(st, tot);
}

// Jacobi sweeps 2..n (Smoother.cpp:100-127): z += inv(D) (r - A z), into
// zout from zin (ping-pong buffers).
template <int DIM, int BS, bool SYM>
__global__ __launch_bounds__(kBlock) void k_jacobi_sweep(const double* __restrict__ val, const double* __restrict__ dinv,
                                                         int64_t ld, int64_t N, int64_t s1, int64_t s2,
                                                         int64_t chunk, const double* __restrict__ r,
                                                         const double* __restrict__ zin, double* __restrict__ zout,
                                                         const KrylovState* st)
{
    if (st && st->done) return;
    const RowRange rr = block_rows(N, chunk);
    stencil_block<DIM, BS, SYM, double>(val, ld, N, s1, s2, rr.r0, rr.r1, zin, [&](int64_t i, const double (&acc)[BS]) {
        double res[BS];
#pragma unroll
        for (int q = 0; q < BS; ++q) res[q] = r[i * BS + q] - acc[q];
#pragma unroll
        for (int q = 0; q < BS; ++q) {
            double sacc = 0.0;
#pragma unroll
            for (int c = 0; c < BS; ++c) sacc += dinv[(int64_t)(q + BS * c) * ld + i] * res[c];
            zout[i * BS + q] = zin[i * BS + q] + sacc;
        }
    });
}

template <int BS>
__global__ __launch_bounds__(kBlock) void k_jacobi_first(const double* __restrict__ dinv, int64_t ld, int64_t N,
                                                         int64_t chunk, const double* __restrict__ r,
                                                         double* __restrict__ z, const KrylovState* st)
{
    if (st && st->done) return;
    const RowRange rr = block_rows(N, chunk);
    for (int64_t i = rr.r0 + threadIdx.x; i < rr.r1; i += kBlock) {
        double zz[BS];
        jacobi_node<BS>(dinv, ld, i, r, zz);
#pragma unroll
        for (int q = 0; q < BS; ++q) z[i * BS + q] = zz[q];
    }
}

}  // namespace

// ====================================================================== host

struct SolverImpl {
    rg_matrix* m;
    rg_solver_work* w;
    hipStream_t st;
    Launch L;     // node kernels without pairing (init, Jacobi sweeps)
    int prec = RG_PRECONDITIONER_JACOBI;
    bool nt = false;   // non-temporal vector streams in the update kernels (stream_policy)
    const double* norm_w = nullptr;   // the flexible CG's residual norm is taken of r / norm_w

    explicit SolverImpl(rg_matrix* mm) : m(mm), w(mm->work), st(mm->ctx->stream), L(chunked_launch(mm->g.N)) {}

    PcgArgs args(const double* zbuf, const Launch& l)
    {
        PcgArgs a;
        a.r = w->r; a.x = w->xw; a.rs = w->rs; a.p = w->p; a.v = w->v; a.x2 = w->x2;
        a.zp = w->z; a.z2p = w->z2; a.s = w->s;
        a.z = zbuf;
        a.b = w->bb;
        a.st = w->st;
        a.partials = m->ctx->partials;
        a.fold_partials = nullptr;
        a.fold_G = 0;
        a.wn = norm_w;
        a.N = m->g.N; a.ld = m->g.ld; a.s1 = m->g.s1; a.s2 = m->g.s2; a.chunk = l.chunk;
        return a;
    }
    template <int STAGE, int NV>
    void stage(int G)
    {
        hipLaunchKernelGGL((k_stage<STAGE, NV>), dim3(1), dim3(kStageThreads), 0, st, w->st, m->ctx->partials, G);
    }

    // launch shapes: stencil kernels (R rows per thread), paired node kernels
    template <typename VT>
    Launch spmv_launch() const { return chunked_launch(m->g.N, spmv_rows_per_thread<VT>()); }
    Launch node_launch() const { return chunked_launch(m->g.N, m->g.bs == 1 ? 2 : 1); }

    // z = M^-1 r with `sweeps` Jacobi sweeps, into w->z (returns the buffer).
    double* jacobi(const double* r, int sweeps, const KrylovState* guard)
    {
        const Geom& g = m->g;
        double* za = w->z;
        double* zb = w->z2;
        dispatch_bs(g.bs, [&](auto B) {
            hipLaunchKernelGGL((k_jacobi_first<B.value>), dim3(L.grid), dim3(kBlock), 0, st, m->dinv, g.ld, g.N,
                               L.chunk, r, za, guard);
        });
        for (int s = 1; s < sweeps; ++s) {
            dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
                hipLaunchKernelGGL((k_jacobi_sweep<D.value, B.value, S.value>), dim3(L.grid), dim3(kBlock), 0, st,
                                   m->val, m->dinv, g.ld, g.N, g.s1, g.s2, L.chunk, r, za, zb, guard);
            });
            std::swap(za, zb);
        }
        RG_HIP(hipGetLastError());
        return za;
    }

    // Whether z = M^-1 r is computed by its own launches into w->z (any
    // preconditioner but a single Jacobi sweep, which the update kernels fuse).
    bool buffered(int sweeps) const { return prec != RG_PRECONDITIONER_JACOBI || sweeps > 1; }

    // out = M^-1 in (Preconditioner_solve); in and out distinct
    void precondition(const double* in, double* out, int sweeps, const KrylovState* guard)
    {
        if (prec == RG_PRECONDITIONER_JACOBI) {
            double* z = jacobi(in, sweeps, guard);
            if (z != out) RG_HIP(hipMemcpyAsync(out, z, sizeof(double) * w->n, hipMemcpyDeviceToDevice, st));
        } else {
            prec_apply(m, prec, sweeps, in, out, st, guard);
        }
    }

    void sweeps_into_z(int sweeps, const KrylovState* guard) { precondition(w->r, w->z, sweeps, guard); }

    // One iteration's launches of the PASO residual-smoothed PCG.
    void paso_iteration(int sweeps)
    {
        if (nt) paso_iteration_nt<true>(sweeps);
        else paso_iteration_nt<false>(sweeps);
    }

    template <bool NT>
    void paso_iteration_nt(int sweeps)
    {
        const Geom& g = m->g;
        const bool buf = buffered(sweeps);
        const double* zb = buf ? w->z : nullptr;
        const Launch Lsp = spmv_launch<double>(), Ln = node_launch();
        const PcgArgs an = args(zb, Ln), asp = args(zb, Lsp);
        PcgArgs ar = args(zb, chunked_launch(g.N * g.bs, 2));
        ar.N = g.N * g.bs;
        const Launch Lr = chunked_launch(ar.N, 2);
        dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
            hipLaunchKernelGGL((k_p_update<B.value, double, false, NT>), dim3(Ln.grid), dim3(kBlock), 0, st, an,
                               m->dinv);
            hipLaunchKernelGGL((k_spmv_dot<D.value, B.value, S.value, double, true>), dim3(Lsp.grid), dim3(kBlock), 0,
                               st, asp, m->val);
        });
        stage<ST_PASO_SPMV, 1>(Lsp.grid);
        // gamma's reduction folded into k_paso_x (both on the fold shape):
        // 128^3 Brick, full storage 0.2676 -> 0.2632 s per PASO solve; with
        // the non-temporal streams of symmetric storage 0.2004 -> 0.2012 s,
        // so only with cached streams
        if (!buf && !NT) {
            const Launch Lrf = chunked_launch(g.N * g.bs, 2, kFoldGrid), Lf = fold_launch();
            PcgArgs arf = args(zb, Lrf);
            arf.N = g.N * g.bs;
            arf.partials = m->ctx->partials + 8 * kMaxGrid;   // rows 8.. of the partials: untouched by the SpMV
            PcgArgs axf = args(zb, Lf);
            axf.fold_partials = arf.partials;
            axf.fold_G = Lrf.grid;
            hipLaunchKernelGGL(k_paso_r<NT>, dim3(Lrf.grid), dim3(kBlock), 0, st, arf);
            dispatch_bs(g.bs, [&](auto B) {
                hipLaunchKernelGGL((k_paso_x<B.value, double, NT, true>), dim3(Lf.grid), dim3(kBlock), 0, st, axf,
                                   m->dinv);
            });
            stage<ST_PASO_X, 2>(Lf.grid);
        } else {
            hipLaunchKernelGGL(k_paso_r<NT>, dim3(Lr.grid), dim3(kBlock), 0, st, ar);
            stage<ST_PASO_R, 2>(Lr.grid);
            if (buf) sweeps_into_z(sweeps, w->st);
            dispatch_bs(g.bs, [&](auto B) {
                hipLaunchKernelGGL((k_paso_x<B.value, double, NT>), dim3(Ln.grid), dim3(kBlock), 0, st, an, m->dinv);
            });
            stage<ST_PASO_X, 2>(Ln.grid);
        }
        RG_HIP(hipGetLastError());
    }

    // One iteration's launches of the textbook PCG (VT: value precision, DT:
    // Jacobi precision); with `replace` the iteration ends with a residual
    // replacement r = D b - A x in FP64.
    template <typename VT, typename DT>
    void cg_iteration(const VT* val, const DT* dinv, int sweeps, bool replace, bool fold = false)
    {
        if (nt) cg_iteration_nt<VT, DT, true>(val, dinv, sweeps, replace, fold);
        else cg_iteration_nt<VT, DT, false>(val, dinv, sweeps, replace, fold);
    }

    // Launch shape of the folded update: at most kFoldGrid workgroups, so
    // that every workgroup of k_p_update<.., FOLD> reduces its partials in
    // one round trip (8 per thread and row)
    // (128^3 Brick, symmetric storage, per solve: 0.1528 s unfolded; folded
    // with 840 / 1400 workgroups 0.1510 / 0.1497 s; the direction update on
    // the full 4200-workgroup shape 0.1534 s)
    static constexpr int kFoldGrid = 2048;
    Launch fold_launch() const { return chunked_launch(m->g.N, m->g.bs == 1 ? 2 : 1, kFoldGrid); }
    // whether an iteration of `method` folds its update stage (chunk() then
    // ends with a commit stage)
    bool folds(int method, int sweeps) const { return method == RG_METHOD_PCG_PLAIN && !buffered(sweeps); }

    template <typename VT, typename DT, bool NT>
    void cg_iteration_nt(const VT* val, const DT* dinv, int sweeps, bool replace, bool fold)
    {
        const Geom& g = m->g;
        const bool buf = buffered(sweeps);
        const double* zb = buf ? w->z : nullptr;
        const Launch Lsp = spmv_launch<VT>(), Ln = node_launch(), L64 = spmv_launch<double>();
        const PcgArgs an = args(zb, Ln), asp = args(zb, Lsp), a64 = args(nullptr, L64);
        // x += alpha p moves into the direction update unless a residual
        // replacement needs x first or the preconditioner has its own
        // launches between the two
        const bool xd = !replace && !buf;
        const bool fd = xd && fold;
        const Launch Lf = fold_launch();
        PcgArgs af = args(zb, Lf);
        af.partials = m->ctx->partials + 8 * kMaxGrid;   // rows 8.. of the partials: untouched by the SpMV
        af.fold_partials = af.partials;
        af.fold_G = Lf.grid;
        dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
            hipLaunchKernelGGL((k_spmv_dot<D.value, B.value, S.value, VT, false>), dim3(Lsp.grid), dim3(kBlock), 0,
                               st, asp, val);
            if (fd) {
                stage<ST_CGX_SPMV, 1>(Lsp.grid);
                hipLaunchKernelGGL((k_cg_update<B.value, DT, true, NT>), dim3(Lf.grid), dim3(kBlock), 0, st, af, dinv);
            } else if (xd) {
                stage<ST_CGX_SPMV, 1>(Lsp.grid);
                hipLaunchKernelGGL((k_cg_update<B.value, DT, true, NT>), dim3(Ln.grid), dim3(kBlock), 0, st, an, dinv);
                stage<ST_CGX_UPDATE, 2>(Ln.grid);
            } else {
                stage<ST_CG_SPMV, 1>(Lsp.grid);
                hipLaunchKernelGGL((k_cg_update<B.value, DT, false, NT>), dim3(Ln.grid), dim3(kBlock), 0, st, an, dinv);
                if (buf) stage<ST_CG_UPDATE_NORZ, 2>(Ln.grid);
                else stage<ST_CG_UPDATE, 2>(Ln.grid);
            }
            if (replace) {
                hipLaunchKernelGGL((k_cg_replace<D.value, B.value, S.value, DT>), dim3(L64.grid), dim3(kBlock), 0,
                                   st, a64, m->val, dinv);
                stage<ST_CG_REPLACE, 2>(L64.grid);
            }
        });
        if (buf) {
            sweeps_into_z(sweeps, w->st);
            hipLaunchKernelGGL(k_cg_rz, dim3(L.grid), dim3(kBlock), 0, st, args(zb, L), w->n);
            stage<ST_CG_RZ, 1>(L.grid);
        }
        dispatch_bs(g.bs, [&](auto B) {
            if (fd)
                hipLaunchKernelGGL((k_p_update<B.value, DT, true, NT, true>), dim3(Lf.grid), dim3(kBlock), 0, st, af,
                                   dinv);
            else if (xd)
                hipLaunchKernelGGL((k_p_update<B.value, DT, true, NT>), dim3(Ln.grid), dim3(kBlock), 0, st, an, dinv);
            else
                hipLaunchKernelGGL((k_p_update<B.value, DT, false, NT>), dim3(Ln.grid), dim3(kBlock), 0, st, an, dinv);
        });
        RG_HIP(hipGetLastError());
    }

    // One iteration of PASO's BiCGStab (FP64, one Jacobi sweep).
    void bicg_iteration(int sweeps)
    {
        if (nt) bicg_iteration_nt<true>(sweeps);
        else bicg_iteration_nt<false>(sweeps);
    }

    template <bool NT>
    void bicg_iteration_nt(int sweeps)
    {
        const Geom& g = m->g;
        const Launch Lsp = spmv_launch<double>(), Ln = chunked_launch(g.N);
        const PcgArgs an = args(nullptr, Ln), asp = args(nullptr, Lsp);
        const int vg = 1024;
        const bool buf = buffered(sweeps);
        const double* dinv = buf ? nullptr : m->dinv;
        hipLaunchKernelGGL(k_bicg_rho<NT>, dim3(vg), dim3(kBlock), 0, st, an, w->n);
        stage<ST_BICG_RHO, 1>(vg);
        dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
            hipLaunchKernelGGL((k_bicg_p<B.value, NT>), dim3(Ln.grid), dim3(kBlock), 0, st, an, dinv);
            if (buf) precondition(w->p, w->z, sweeps, w->st);   // phat = M^-1 p
            hipLaunchKernelGGL((k_bicg_spmv<D.value, B.value, S.value, 0>), dim3(Lsp.grid), dim3(kBlock), 0, st, asp,
                               m->val, w->z, w->v);
            stage<ST_BICG_ALPHA, 2>(Lsp.grid);
            hipLaunchKernelGGL((k_bicg_s<B.value, NT>), dim3(Ln.grid), dim3(kBlock), 0, st, an, dinv);
            stage<ST_BICG_S, 1>(Ln.grid);
            if (buf) precondition(w->s, w->z2, sweeps, w->st);  // shat = M^-1 s
            hipLaunchKernelGGL(k_bicg_xa<NT>, dim3(vg), dim3(kBlock), 0, st, an, w->n);
            stage<ST_BICG_XA, 1>(vg);
            hipLaunchKernelGGL((k_bicg_spmv<D.value, B.value, S.value, 1>), dim3(Lsp.grid), dim3(kBlock), 0, st, asp,
                               m->val, w->z2, w->x2);
            stage<ST_BICG_OMEGA, 2>(Lsp.grid);
        });
        hipLaunchKernelGGL(k_bicg_update<NT>, dim3(vg), dim3(kBlock), 0, st, an, w->n);
        stage<ST_BICG_END, 1>(vg);
        RG_HIP(hipGetLastError());
    }

    // One iteration of PASO's MINRES (FP64, any preconditioner).
    void mr_iteration(int sweeps)
    {
        if (nt) mr_iteration_nt<true>(sweeps);
        else mr_iteration_nt<false>(sweeps);
    }

    template <bool NT>
    void mr_iteration_nt(int sweeps)
    {
        const Geom& g = m->g;
        const bool buf = buffered(sweeps);
        const Launch Lsp = spmv_launch<double>(), Ln = chunked_launch(g.N);
        const PcgArgs an = args(nullptr, Ln), asp = args(nullptr, Lsp);
        const int vg = 1024;
        dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
            hipLaunchKernelGGL((k_mr_spmv<D.value, B.value, S.value>), dim3(Lsp.grid), dim3(kBlock), 0, st, asp,
                               m->val);
            stage<ST_MR_DELTA, 1>(Lsp.grid);
            hipLaunchKernelGGL((k_mr_r<B.value, NT>), dim3(Ln.grid), dim3(kBlock), 0, st, an,
                               buf ? nullptr : m->dinv);
        });
        if (buf) {
            sweeps_into_z(sweeps, w->st);
            hipLaunchKernelGGL(k_cg_rz, dim3(L.grid), dim3(kBlock), 0, st, args(w->z, L), w->n);
            stage<ST_MR_GAMMA, 1>(L.grid);
        } else {
            stage<ST_MR_GAMMA, 1>(Ln.grid);
        }
        hipLaunchKernelGGL(k_mr_w<NT>, dim3(vg), dim3(kBlock), 0, st, an, w->n);
        RG_HIP(hipGetLastError());
    }

    void iteration(int method, int sweeps, bool last)
    {
        if (method == RG_METHOD_BICGSTAB) bicg_iteration(sweeps);
        else if (method == RG_METHOD_MINRES) mr_iteration(sweeps);
        else if (method == RG_METHOD_FCG_LEARNED) mg_fcg_iteration(m, w->xw, w->r, w->z, w->st, st, norm_w);
        else if (method == RG_METHOD_PCG) paso_iteration(sweeps);
        else if (method == RG_METHOD_PCG_PLAIN)
            cg_iteration<double, double>(m->val, m->dinv, sweeps, false, folds(method, sweeps));
        else cg_iteration<float, float>(m->val32, m->dinv32, 1, last);
    }

    void init(int method, int sweeps)
    {
        const Geom& g = m->g;
        double* zb = nullptr;
        if (buffered(sweeps) && method != RG_METHOD_BICGSTAB) {
            sweeps_into_z(sweeps, nullptr);
            zb = w->z;
        }
        PcgArgs a = args(zb, L);
        dispatch_bs(g.bs, [&](auto B) {
            if (method == RG_METHOD_PCG) {
                hipLaunchKernelGGL((k_paso_init<B.value, double>), dim3(L.grid), dim3(kBlock), 0, st, a, m->dinv);
                stage<ST_PASO_INIT, 1>(L.grid);
            } else if (method == RG_METHOD_PCG_PLAIN) {
                hipLaunchKernelGGL((k_cg_init<B.value, double>), dim3(L.grid), dim3(kBlock), 0, st, a, m->dinv);
                stage<ST_CG_INIT, 2>(L.grid);
            } else if (method == RG_METHOD_BICGSTAB) {
                PcgArgs ai = a;
                const Launch Lv = chunked_launch(g.N * B.value);
                ai.chunk = Lv.chunk;
                hipLaunchKernelGGL((k_bicg_init<B.value>), dim3(Lv.grid), dim3(kBlock), 0, st, ai);
            } else if (method == RG_METHOD_MINRES) {
                hipLaunchKernelGGL((k_mr_init<B.value>), dim3(L.grid), dim3(kBlock), 0, st, a, m->dinv);
                stage<ST_MR_INIT, 2>(L.grid);
            } else if (method == RG_METHOD_FCG_LEARNED) {
                hipLaunchKernelGGL((k_cg_init<B.value, double>), dim3(L.grid), dim3(kBlock), 0, st, a, m->dinv);
                stage<ST_FCG_INIT, 2>(L.grid);
            } else {
                hipLaunchKernelGGL((k_cg_init<B.value, float>), dim3(L.grid), dim3(kBlock), 0, st, a, m->dinv32);
                stage<ST_CG_INIT, 2>(L.grid);
            }
        });
        RG_HIP(hipGetLastError());
    }

    // A chunk of `iters` iterations, replayed from a cached hipGraph.
    void chunk(int method, int sweeps, int iters, bool use_graph)
    {
        if (!use_graph) {
            for (int k = 0; k < iters; ++k) iteration(method, sweeps, k == iters - 1);
            if (folds(method, sweeps)) stage<ST_CG_COMMIT, 1>(1);
            return;
        }
        const int key = method + 8 * prec + (norm_w ? 64 : 0);
        if (!w->graph || w->graph_method != key || w->graph_sweeps != sweeps || w->graph_iters != iters) {
            if (w->graph) {
                RG_HIP(hipGraphExecDestroy(w->graph));
                w->graph = nullptr;
            }
            if (!w->capture) RG_HIP(hipStreamCreateWithFlags(&w->capture, hipStreamNonBlocking));
            hipStream_t keep = st;
            st = w->capture;
            hipGraph_t graph;
            RG_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            try {
                for (int k = 0; k < iters; ++k) iteration(method, sweeps, k == iters - 1);
                // the chunk's state is complete when the host polls it
                if (folds(method, sweeps)) stage<ST_CG_COMMIT, 1>(1);
            } catch (...) {
                hipStreamEndCapture(st, &graph);
                st = keep;
                throw;
            }
            RG_HIP(hipStreamEndCapture(st, &graph));
            st = keep;
            RG_HIP(hipGraphInstantiate(&w->graph, graph, nullptr, nullptr, 0));
            RG_HIP(hipGraphDestroy(graph));
            w->graph_method = key;
            w->graph_sweeps = sweeps;
            w->graph_iters = iters;
        }
        RG_HIP(hipGraphLaunch(w->graph, st));
    }

    // Runs the Krylov loop from w->r, w->xw until the state reports done.
    // Returns (status, iterations, final norm).
    void krylov(int method, int sweeps, int maxit, double tol, int check, bool use_graph, int* status,
                int* iters, double* norm)
    {
        // the host copy and two polling slots sit below the flag words at pinned + 128
        static_assert(3 * sizeof(KrylovState) <= 128 * sizeof(double), "KrylovState outgrew the pinned buffer");
        KrylovState h;
        std::memset(&h, 0, sizeof(h));
        h.maxit = maxit;
        h.tol = tol;
        h.replace_mode = (method == RG_METHOD_PCG_MIXED) ? 1 : 0;
        KrylovState* pin = reinterpret_cast<KrylovState*>(m->ctx->pinned);
        *pin = h;
        RG_HIP(hipMemcpyAsync(w->st, pin, sizeof(KrylovState), hipMemcpyHostToDevice, st));
        // the copy above reads pinned memory asynchronously: wait before reusing it
        RG_HIP(hipStreamSynchronize(st));
        init(method, sweeps);
        KrylovState* slots = pin + 1;   // two polling slots
        hipEvent_t evs[2];
        RG_HIP(hipEventCreateWithFlags(&evs[0], hipEventDisableTiming));
        RG_HIP(hipEventCreateWithFlags(&evs[1], hipEventDisableTiming));
        int launched = 0;
        bool done = false;
        std::vector<double> hist;
        bool stagnated = false;
        try {
            // keep two chunks in flight; poll the older one
            while (!done) {
                chunk(method, sweeps, check, use_graph);
                RG_HIP(hipMemcpyAsync(&slots[launched % 2], w->st, sizeof(KrylovState), hipMemcpyDeviceToHost, st));
                RG_HIP(hipEventRecord(evs[launched % 2], st));
                ++launched;
                if (launched >= 2) {
                    RG_HIP(hipEventSynchronize(evs[(launched - 2) % 2]));
                    const KrylovState& sl = slots[(launched - 2) % 2];
                    done = sl.done != 0;
                    // flexible CG with a learned preconditioner: stop on
                    // stagnation -- no new best residual (1% below the best
                    // so far) for max(4 polls, 64 iterations)
                    if (!done && method == RG_METHOD_FCG_LEARNED && !stagnated) {
                        hist.push_back(sl.norm);
                        const size_t k = hist.size();
                        const size_t window = std::max<size_t>(4, (size_t)((64 + check - 1) / check));
                        bool improving = k <= window;
                        if (!improving) {
                            double best_before = hist[0];
                            for (size_t j = 1; j + window < k; ++j) best_before = std::min(best_before, hist[j]);
                            for (size_t j = k - window; j < k; ++j) improving = improving || hist[j] < 0.99 * best_before;
                        }
                        if (!improving) {
                            stagnated = true;
                            int32_t flag[2] = {1, KS_STAGNATED};
                            int32_t* pin32 = reinterpret_cast<int32_t*>(m->ctx->pinned + 128);
                            pin32[0] = flag[0];
                            pin32[1] = flag[1];
                            RG_HIP(hipMemcpyAsync(&w->st->done, pin32, 2 * sizeof(int32_t), hipMemcpyHostToDevice, st));
                        }
                    }
                }
            }
        } catch (...) {
            hipEventDestroy(evs[0]);
            hipEventDestroy(evs[1]);
            throw;
        }
        RG_HIP(hipStreamSynchronize(st));
        RG_HIP(hipEventDestroy(evs[0]));
        RG_HIP(hipEventDestroy(evs[1]));
        RG_HIP(hipMemcpy(pin, w->st, sizeof(KrylovState), hipMemcpyDeviceToHost));
        *status = pin->status;
        *iters = pin->num_iter;
        *norm = pin->norm;
    }

    void norms(const double* v, double* n2, double* nmax, const double* d = nullptr)
    {
        const int grid = 1024;
        hipLaunchKernelGGL(k_norms, dim3(grid), dim3(kBlock), 0, st, v, d, w->n, m->ctx->partials);
        hipLaunchKernelGGL(k_norms_stage, dim3(1), dim3(kBlock), 0, st, m->ctx->partials, grid, w->norms);
        RG_HIP(hipGetLastError());
        RG_HIP(hipMemcpyAsync(m->ctx->pinned, w->norms, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
        RG_HIP(hipStreamSynchronize(st));
        *n2 = std::sqrt(m->ctx->pinned[0]);
        *nmax = m->ctx->pinned[1];
    }
};

// MI355X's memory-side cache (Infinity Cache, MALL).
constexpr int64_t kMallBytes = 256ll << 20;

// Whether the update kernels stream their vectors non-temporally: when the
// SpMV's values fit in the MALL but values plus the vectors an iteration
// touches (textbook CG: p, q, r, x, M^-1; PASO adds rs, x2; MINRES: r,
// r_old, Az, z, z / gamma, w, w_old, x, M^-1; BiCGStab: r, r~, p, v, phat,
// s, shat, t, x, M^-1) do not, the vectors would otherwise evict the values
// between SpMVs.
bool stream_policy(const rg_matrix* m, int method)
{
    if (method != RG_METHOD_PCG && method != RG_METHOD_PCG_PLAIN && method != RG_METHOD_PCG_MIXED &&
        method != RG_METHOD_MINRES && method != RG_METHOD_BICGSTAB)
        return false;
    const int64_t vsize = method == RG_METHOD_PCG_MIXED ? (int64_t)sizeof(float) : (int64_t)sizeof(double);
    const int64_t val_bytes = (int64_t)m->g.nstore * m->g.bs * m->g.bs * m->g.ld * vsize;
    const int64_t nvec = method == RG_METHOD_PCG        ? 7
                         : method == RG_METHOD_MINRES   ? 9
                         : method == RG_METHOD_BICGSTAB ? 10
                                                        : 5;
    const int64_t vec_bytes = nvec * (int64_t)m->g.N * m->g.bs * (int64_t)sizeof(double);
    return val_bytes <= kMallBytes && val_bytes + vec_bytes > kMallBytes;
}

void ensure_work(rg_matrix* m)
{
    if (m->work) return;
    rg_solver_work* w = new rg_solver_work();
    w->n = m->g.N * m->g.bs;
    const size_t bytes = sizeof(double) * (size_t)w->n;
    try {
        RG_HIP(hipMalloc(&w->r, bytes));
        RG_HIP(hipMalloc(&w->xw, bytes));
        RG_HIP(hipMalloc(&w->rs, bytes));
        RG_HIP(hipMalloc(&w->p, bytes));
        RG_HIP(hipMalloc(&w->v, bytes));
        RG_HIP(hipMalloc(&w->x2, bytes));
        RG_HIP(hipMalloc(&w->z, bytes));
        RG_HIP(hipMalloc(&w->z2, bytes));
        RG_HIP(hipMalloc(&w->s, bytes));
        RG_HIP(hipMalloc(&w->bb, bytes));
        RG_HIP(hipMalloc(&w->norms, 4 * sizeof(double)));
        RG_HIP(hipMalloc(&w->st, sizeof(KrylovState)));
        RG_HIP(hipMalloc(&w->flag, sizeof(int)));
    } catch (...) {
        m->work = w;
        throw;
    }
    m->work = w;
}

void free_work(rg_solver_work* w)
{
    if (!w) return;
    if (w->graph) hipGraphExecDestroy(w->graph);
    if (w->capture) hipStreamDestroy(w->capture);
    double* bufs[] = {w->r, w->xw, w->rs, w->p, w->v, w->x2, w->z, w->z2, w->s, w->bb, w->norms, w->pb1, w->pb2};
    for (double* b : bufs)
        if (b) hipFree(b);
    if (w->st) hipFree(w->st);
    if (w->flag) hipFree(w->flag);
    delete w;
}

// SystemMatrix::balance (once per set of values) and the Jacobi inverse of
// the balanced matrix (Preconditioner_alloc -> invMain; kept until the values
// are reset without preserving solver data).
void prepare_solve(rg_matrix* m, int prec)
{
    const Geom& g = m->g;
    hipStream_t st = m->ctx->stream;
    const int grid = (int)((g.N + kBlock - 1) / kBlock);
    if (!m->is_balanced) {
        if (m->pval_state == 1) m->pval_state = 0;   // the Gauss-Seidel copy holds the old values
        dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
            hipLaunchKernelGGL((k_balance_vector<D.value, B.value, S.value>), dim3(grid), dim3(kBlock), 0, st,
                               m->val, g.ld, g.N, g.s1, g.s2, m->balance);
            hipLaunchKernelGGL((k_apply_balance<D.value, B.value, S.value>), dim3(grid), dim3(kBlock), 0, st,
                               m->val, g.ld, g.N, g.s1, g.s2, m->balance);
        });
        RG_HIP(hipGetLastError());
        m->is_balanced = true;
        m->val32_valid = false;
        m->val_zero = false;
    }
    const bool needs_inverse = prec == RG_PRECONDITIONER_JACOBI || prec == RG_PRECONDITIONER_GAUSS_SEIDEL;
    if (needs_inverse && !m->have_prec) {
        RG_CHECK(g.bs <= 3, RG_VALUE_ERROR,
                 "SparseMatrix::invMain: block size > 3 needs LAPACK (not supported)");
        RG_HIP(hipMemsetAsync(m->work->flag, 0, sizeof(int), st));
        dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
            hipLaunchKernelGGL((k_inv_main<B.value, S.value, D.value>), dim3(grid), dim3(kBlock), 0, st, m->val,
                               g.ld, g.N, m->dinv, m->dinv32, m->work->flag);
        });
        RG_HIP(hipGetLastError());
        int failed = 0;
        RG_HIP(hipMemcpyAsync(&failed, m->work->flag, sizeof(int), hipMemcpyDeviceToHost, st));
        RG_HIP(hipStreamSynchronize(st));
        RG_CHECK(!failed, RG_VALUE_ERROR, "SparseMatrix::invMain: non-regular main diagonal block.");
        m->have_prec = true;
    }
    prec_setup(m, prec);
}

// paso::Solver (paso/src/Solver.cpp:46-331) on device vectors.
// With RG_CRITERION_TRUE_RESIDUAL the outer loop stops on the unbalanced
// FP64 residual, ||b - A x||_2 <= rtol ||b||_2, instead of PASO's balanced
// two-norm/max-norm pair; the inner tolerance is scaled by the observed
// ratio of balanced to unbalanced residual norms.
int solve(rg_matrix* m, double* x, const double* b, const rg_options& opt, rg_diagnostics& dg)
{
    const double EPS = 2.220446049250313e-16;
    const double t0 = now_s();
    dg.num_iter = 0; dg.converged = 0; dg.status = 0; dg.outer_steps = 0; dg.fallbacks = 0; dg.reserved = 0;
    dg.residual_norm = -1; dg.set_up_time = -1; dg.net_time = -1; dg.time = -1;
    const double tolerance = opt.tolerance;
    RG_CHECK(tolerance >= 100. * EPS, RG_VALUE_ERROR, "Solver: Tolerance is too small.");
    RG_CHECK(tolerance <= 1., RG_VALUE_ERROR, "Solver: Tolerance must be less than one.");
    RG_CHECK(opt.method >= RG_METHOD_PCG && opt.method <= RG_METHOD_MINRES, RG_VALUE_ERROR,
             "Solver: unknown solver method.");
    const int prec = opt.preconditioner;
    RG_CHECK(prec >= RG_PRECONDITIONER_JACOBI && prec <= RG_PRECONDITIONER_NONE, RG_VALUE_ERROR,
             "Solver: unknown preconditioner.");
    RG_CHECK(prec == RG_PRECONDITIONER_JACOBI ||
                 (opt.method != RG_METHOD_PCG_MIXED && opt.method != RG_METHOD_FCG_LEARNED),
             RG_VALUE_ERROR, "Solver: the FP32 and learned methods use the Jacobi preconditioner.");
    RG_CHECK(opt.method != RG_METHOD_BICGSTAB || prec != RG_PRECONDITIONER_JACOBI || opt.sweeps <= 1,
             RG_VALUE_ERROR, "Solver: BiCGStab runs with one Jacobi sweep.");
    RG_CHECK(opt.method != RG_METHOD_FCG_LEARNED || (m->g.bs == 1 && m->g.dim == 3), RG_VALUE_ERROR,
             "Solver: the flexible CG with the multilevel preconditioner needs a scalar Brick matrix.");
    RG_CHECK(opt.criterion == RG_CRITERION_PASO || opt.criterion == RG_CRITERION_TRUE_RESIDUAL, RG_VALUE_ERROR,
             "Solver: unknown stopping criterion.");
    const bool true_res = opt.criterion == RG_CRITERION_TRUE_RESIDUAL;
    ensure_work(m);
    rg_solver_work* w = m->work;
    hipStream_t st = m->ctx->stream;
    SolverImpl S(m);
    S.prec = prec;
    S.nt = stream_policy(m, opt.method);
    const int64_t n = w->n;
    const int vgrid = 2048;
    prepare_solve(m, prec);
    if (prec == RG_PRECONDITIONER_GAUSS_SEIDEL && opt.sweeps > 1) {   // before any graph capture
        if (!w->pb1) RG_HIP(hipMalloc(&w->pb1, sizeof(double) * n));
        if (!w->pb2) RG_HIP(hipMalloc(&w->pb2, sizeof(double) * n));
    }
    int method = opt.method;
    if ((method == RG_METHOD_PCG_MIXED || method == RG_METHOD_FCG_LEARNED) && !m->val32_valid) {
        if (!m->val32) {
            RG_HIP(hipMalloc(&m->val32, value_count(m->g) * sizeof(float)));
        }
        launch_to_fp32(m);
        m->val32_valid = true;
    }
    if (method == RG_METHOD_FCG_LEARNED) mg_setup(m);
    // r = D b, norms of D b (Solver.cpp:104-130)
    hipLaunchKernelGGL(k_mul, dim3(vgrid), dim3(kBlock), 0, st, w->bb, b, m->balance, n);
    RG_HIP(hipMemcpyAsync(w->r, w->bb, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
    double norm2_of_b, norm_max_of_b;
    S.norms(w->r, &norm2_of_b, &norm_max_of_b);
    RG_CHECK(!std::isnan(norm2_of_b) && !std::isnan(norm_max_of_b), RG_VALUE_ERROR,
             "Solver: Matrix or right hand side contains undefined values.");
    double true2_of_b = 0.0, truemax_of_b = 0.0;
    if (true_res) S.norms(b, &true2_of_b, &truemax_of_b);
    int status = RG_OK;
    const bool use_graph = true;
    if (norm2_of_b <= 0.) {
        hipLaunchKernelGGL(k_fill, dim3(vgrid), dim3(kBlock), 0, st, x, 0.0, n);
        RG_HIP(hipGetLastError());
        if (opt.verbose) printf("right hand side is identical to zero.\n");
    } else {
        dg.set_up_time = now_s() - t0;
        // initial guess x = M^-1 D b (Solver.cpp:181)
        S.precondition(w->r, w->xw, std::max(1, opt.sweeps), nullptr);
        int totIter = 1;
        bool finalize = false;
        double last2 = norm2_of_b, lastmax = norm_max_of_b;
        const double tn = now_s();
        while (!finalize) {
            const int cntIter = opt.iter_max - totIter;
            finalize = true;
            if (totIter > 1) RG_HIP(hipMemcpyAsync(w->r, w->bb, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
            launch_spmv(m, -1.0, w->xw, 1.0, w->r);
            double norm2, normmax;
            S.norms(w->r, &norm2, &normmax);
            double true2 = 0.0, truemax = 0.0;
            if (true_res) S.norms(w->r, &true2, &truemax, m->balance);
            dg.residual_norm = norm2;
            dg.outer_steps++;
            if (opt.verbose)
                printf("Solver: Step %d: l2/lmax-norm of residual is %g/%g\n", totIter, norm2, normmax);
            if (totIter > 1 && norm2 >= last2 && normmax >= lastmax) {
                throw Error(RG_DIVERGENCE, "Solver: No improvement during iteration. Iterative solver gives up.");
            }
            const bool more = true_res ? (true2 > tolerance * true2_of_b)
                                       : (norm2 > tolerance * norm2_of_b || normmax > tolerance * norm_max_of_b);
            if (more) {
                // the flexible CG tracks the unbalanced residual norm itself
                // with the true-residual criterion (its r / balance), so its
                // tolerance is the target with a margin for the recursive
                // residual's drift; the other methods scale the target by
                // the observed ratio of balanced to unbalanced norms
                const bool weighted = true_res && method == RG_METHOD_FCG_LEARNED;
                S.norm_w = weighted ? m->balance : nullptr;
                double tol = tolerance * std::min(norm2_of_b, 0.1 * norm2 / normmax * norm_max_of_b);
                if (weighted) tol = 0.8 * tolerance * true2_of_b;
                else if (true_res) tol = 0.5 * tolerance * true2_of_b * (norm2 / std::max(true2, 1e-300));
                // iterations per replayed graph and host poll: the flexible
                // CG's iteration is some 50 launches, and the chunk launched
                // behind the converged one runs as no-ops
                const int check = opt.check_interval > 0 ? opt.check_interval
                                                         : (method == RG_METHOD_FCG_LEARNED ? 1 : 32);
                if (opt.verbose) printf("Solver: new tolerance = %g\n", tol);
                last2 = norm2;
                lastmax = normmax;
                int ks = 0, it = 0;
                double kn = 0;
                S.krylov(method, std::max(1, opt.sweeps), cntIter, tol, check, use_graph, &ks, &it, &kn);
                RG_HIP(hipGetLastError());
                totIter += it;
                if (ks == KS_CONVERGED) {
                    finalize = false;
                } else if (ks == KS_STAGNATED || (ks == KS_BREAKDOWN && method == RG_METHOD_FCG_LEARNED)) {
                    // the learned preconditioner stopped helping (or gave a
                    // zero direction): continue from the current iterate
                    // with the Jacobi-PCG
                    if (opt.verbose) printf("Solver: learned preconditioner stagnated, falling back to PCG\n");
                    method = RG_METHOD_PCG_MIXED;
                    dg.fallbacks++;
                    finalize = false;
                    // the switch is the remedy for a lack of progress: the
                    // next outer step is not judged against this one
                    last2 = lastmax = std::numeric_limits<double>::infinity();
                } else if (ks == KS_MAXITER) {
                    status = RG_MAX_ITER_REACHED;
                    break;
                } else if (ks == KS_NEGATIVE) {
                    status = RG_NEGATIVE_NORM;
                    break;
                } else if (ks == KS_BREAKDOWN) {
                    if (it > 1) {
                        finalize = false;
                    } else {
                        status = RG_BREAKDOWN;
                        break;
                    }
                }
            } else {
                dg.converged = 1;
            }
        }
        dg.net_time = now_s() - tn;
        dg.num_iter = totIter;
        hipLaunchKernelGGL(k_mul, dim3(vgrid), dim3(kBlock), 0, st, x, w->xw, m->balance, n);
        RG_HIP(hipGetLastError());
    }
    RG_HIP(hipStreamSynchronize(st));
    dg.status = status;
    dg.time = now_s() - t0;
    return status;
}

// SystemMatrix::setPreconditioner + solvePreconditioner
// (paso/src/SystemMatrix.cpp:42-56) on the balanced matrix.
void solve_preconditioner(rg_matrix* m, const rg_options& opt, double* x, const double* b)
{
    const int prec = opt.preconditioner;
    RG_CHECK(prec >= RG_PRECONDITIONER_JACOBI && prec <= RG_PRECONDITIONER_NONE, RG_VALUE_ERROR,
             "solvePreconditioner: unknown preconditioner.");
    RG_CHECK(x != b, RG_VALUE_ERROR, "solvePreconditioner: x and b must be distinct");
    ensure_work(m);
    prepare_solve(m, prec);
    if (opt.method == RG_METHOD_FCG_LEARNED) {
        // one V-cycle of the multilevel preconditioner
        RG_CHECK(prec == RG_PRECONDITIONER_JACOBI && m->g.bs == 1 && m->g.dim == 3, RG_VALUE_ERROR,
                 "solvePreconditioner: the multilevel preconditioner needs a scalar Brick matrix and Jacobi");
        if (!m->val32_valid) {
            if (!m->val32) RG_HIP(hipMalloc(&m->val32, value_count(m->g) * sizeof(float)));
            launch_to_fp32(m);
            m->val32_valid = true;
        }
        mg_setup(m);
        mg_vcycle(m, b, x, m->ctx->stream, nullptr);
        RG_HIP(hipStreamSynchronize(m->ctx->stream));
        return;
    }
    SolverImpl S(m);
    S.prec = prec;
    S.precondition(b, x, std::max(1, opt.sweeps), nullptr);
    RG_HIP(hipStreamSynchronize(m->ctx->stream));
}

}  // namespace rg
