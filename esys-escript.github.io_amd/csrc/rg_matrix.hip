// Matrix kernels of libripley_hip: value initialisation, element-block
// insertion, colour-free gather assembly, stencil SpMV, Dirichlet
// nullification and the PASO CSR export.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>

#include "rg_device.h"
#include "rg_dispatch.h"
#include "rg_stencil.h"

namespace rg {

// ------------------------------------------------------------- set values
// paso::SystemMatrix::setValues (paso/src/SystemMatrix.h:273-279): every
// entry of the PASO pattern := value.  Stencil slots whose neighbour lies
// outside the grid are not in the pattern and are set to zero.
template <int DIM, int BS, bool SYM>
__global__ __launch_bounds__(kBlock) void k_set_values(double* __restrict__ val, int64_t ld, int64_t N,
                                                       int64_t n0, int64_t n1, int64_t n2,
                                                       double value)
{
    constexpr int NST = stored_slots<DIM, SYM>();
    constexpr int FIRST = SYM ? StencilShape<DIM>::CENTER : 0;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const int64_t x = i % n0, y = (i / n0) % n1, z = i / (n0 * n1);
#pragma unroll
    for (int t = 0; t < NST; ++t) {
        const double v = slot_in_grid<DIM>(FIRST + t, x, y, z, n0, n1, n2) ? value : 0.0;
#pragma unroll
        for (int e = 0; e < BS * BS; ++e) val[(int64_t)(t * BS * BS + e) * ld + i] = v;
    }
}

void launch_set_values(rg_matrix* m, double value)
{
    const Geom& g = m->g;
    const int grid = (int)((g.N + kBlock - 1) / kBlock);
    dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
        hipLaunchKernelGGL((k_set_values<D.value, B.value, S.value>), dim3(grid), dim3(kBlock), 0,
                           m->ctx->stream, m->val, g.ld, g.N, g.nn[0], g.nn[1], g.nn[2], value);
    });
    RG_HIP(hipGetLastError());
    m->val_zero = (value == 0.0);
}

// -------------------------------------------------- element block insertion
// ripley::SystemMatrix::add (ripley/src/RipleySystemMatrix.cu:111-155), the
// stencil-matrix counterpart of addToPasoMatrix
// (ripley/src/RipleyDomain.cpp:1488-1517): the pair (a, b) of local element
// nodes lands in the slot given by their local positions; with symmetric
// storage only pairs with b >= a are added.
//
// One launch per element colour: the colour of an element is the parity of
// the grid coordinates of its local node 0 (2^DIM colours), so two distinct
// grid cells of one colour share no node -- the slab colouring that makes
// the reference's assembler loops race-free (DefaultAssembler3D.cpp:243-245),
// taken in all axes.  Inside a launch every thread owns one (element, a, b,
// k, m) target.  For distinct grid cells each matrix entry then receives at
// most one contribution per launch and the sums are formed in colour order,
// the same on every run.  The adds are atomic, so inputs whose "elements"
// share nodes within a colour (repeated elements, node lists that are not
// grid cells) stay race-free; only their summation order is then open.
template <int DIM, int BS, bool SYM>
__global__ __launch_bounds__(kBlock) void k_add_elements(double* __restrict__ val, int64_t ld, int64_t n0,
                                                         int64_t n1, int64_t num_el,
                                                         const int32_t* __restrict__ nodes,
                                                         const double* __restrict__ em, int colour)
{
    constexpr int NN = 1 << DIM;
    constexpr int C = StencilShape<DIM>::CENTER;
    constexpr int NT = NN * NN * BS * BS;
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= num_el * NT) return;
    const int64_t e = t / NT;
    const int j = (int)(t - e * NT);
    const int64_t base = nodes[e * NN];
    const int64_t bx = base % n0, by = (base / n0) % n1, bz = base / (n0 * n1);
    const int c = (int)((bx & 1) | ((by & 1) << 1) | (DIM == 3 ? (bz & 1) << 2 : 0));
    if (c != colour) return;
    const int k = j % BS, mm = (j / BS) % BS, a = (j / (BS * BS)) % NN, b = j / (BS * BS * NN);
    if (SYM && b < a) return;
    const int s = ((b & 1) - (a & 1) + 1) + 3 * (((b >> 1) & 1) - ((a >> 1) & 1) + 1) +
                  (DIM == 3 ? 9 * (((b >> 2) & 1) - ((a >> 2) & 1) + 1) : 0);
    const int st = SYM ? s - C : s;
    const int64_t row = nodes[e * NN + a];
    atomicAdd(&val[(int64_t)(st * BS * BS + k + BS * mm) * ld + row], em[t]);
}

void launch_add_element_blocks(rg_matrix* m, int64_t ne, int npe, const int32_t* d_nodes,
                               const double* d_em)
{
    const Geom& g = m->g;
    RG_CHECK(npe == (1 << g.dim), RG_VALUE_ERROR,
             "add_element_blocks: nodes_per_element must be 4 (2-D) or 8 (3-D)");
    if (ne == 0) return;
    const int64_t targets = ne * npe * npe * g.bs * g.bs;
    const int64_t grid = (targets + kBlock - 1) / kBlock;
    RG_CHECK(grid < ((int64_t)1 << 31), RG_VALUE_ERROR, "add_element_blocks: too many elements in one call");
    dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
        for (int c = 0; c < npe; ++c)
            hipLaunchKernelGGL((k_add_elements<D.value, B.value, S.value>), dim3((unsigned)grid), dim3(kBlock), 0,
                               m->ctx->stream, m->val, g.ld, g.nn[0], g.nn[1], ne, d_nodes, d_em, c);
    });
    RG_HIP(hipGetLastError());
    m->val_zero = false;
}

// ---------------------------------------------------- allocation-free scratch
// Device memory of the entry points that allocate nothing
// (rg_matrix_add_element_blocks_noalloc): one array per device, held until
// the process ends.  Host code must not pass the array's name to a kernel
// (it is a host-side stand-in); it uses the device address that
// scratch_address() looks up once per device with that device current.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

void* scratch_address()
{
    constexpr int kMaxDevices = 64;
    static std::mutex mu;
    static void* addr[kMaxDevices] = {};
    int dev = 0;
    RG_HIP(hipGetDevice(&dev));
    RG_CHECK(dev >= 0 && dev < kMaxDevices, RG_DEVICE_ERROR, "scratch: device id out of range");
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        RG_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        RG_CHECK(p != nullptr, RG_DEVICE_ERROR, "hipGetSymbolAddress(scratch) returned NULL");
        addr[dev] = p;
    }
    return addr[dev];
}

// ---------------------------------------------------------------- assembly
struct CoefDev {
    const double* p;
    int mode;
    int stride;
    int iso;
};
struct AsmArgs {
    CoefDev A, B, C, D, X, Y;
    int64_t n0, n1, n2;     // nodes per axis
    int ne0, ne1, ne2;      // elements per axis
    double h0, h1, h2;
    int reduced;
};

__device__ __forceinline__ const double* coef_pt(const CoefDev& c, int64_t e, int q, int dps)
{
    const double* p = c.p + (int64_t)c.stride * e;
    return (c.mode == RG_COEF_EXPANDED) ? p + q * dps : p;
}

// Colour-free gather assembly.  One thread per (node i, equation k,
// component m): it visits the 2^DIM elements around node i, computes row a
// (the node's local index) of each element matrix by Gauss quadrature and
// adds entry (a, b) to the stencil slot of local node b:
//   EM_S[k,m,a,b] = sum_q w ( A[k,i,m,j] dN_a/dx_i dN_b/dx_j + B[k,i,m] N_b dN_a/dx_i
//                           + C[k,m,j] N_a dN_b/dx_j + D[k,m] N_a N_b )
//   EM_F[k,a]     = sum_q w ( X[k,i] dN_a/dx_i + Y[k] N_a )
// -- the weak form that DefaultAssembler3D::assemblePDESingle/System
// integrate with the same 2-point Gauss rule (points (1-+1/sqrt3)/2, x
// fastest; ripley/src/DefaultAssembler3D.cpp:153-2229, :3631-5745) and, for
// reduced, the one-point rule of assemblePDE*Reduced (:2891-3270).  Each
// matrix entry and rhs entry is written by exactly one thread: no atomics.
template <int DIM, int NB, bool SYM>
__global__ __launch_bounds__(kBlock) void k_assemble(AsmArgs a, double* __restrict__ val, int64_t ld,
                                                     int64_t N, double* __restrict__ rhs)
{
    constexpr int NS = StencilShape<DIM>::NS;
    constexpr int CEN = StencilShape<DIM>::CENTER;
    constexpr int NST = stored_slots<DIM, SYM>();
    constexpr int FIRST = SYM ? CEN : 0;
    constexpr int NNODE = 1 << DIM;
    const int km = blockIdx.y;
    const int k = km % NB, m = km / NB;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const int64_t x = i % a.n0, y = (i / a.n0) % a.n1, z = (DIM == 3) ? i / (a.n0 * a.n1) : 0;
    const double h[3] = {a.h0, a.h1, a.h2};
    double vol = a.h0 * a.h1;
    if (DIM == 3) vol *= a.h2;
    const int nq = a.reduced ? 1 : NNODE;
    const double w = vol / nq;
    const double g0 = 0.5 - 0.5 / 1.73205080756887719318, g1 = 0.5 + 0.5 / 1.73205080756887719318;
    const bool doS = (a.A.mode | a.B.mode | a.C.mode | a.D.mode) != 0;
    const bool doF = rhs != nullptr && m == 0 && (a.X.mode | a.Y.mode) != 0;
    const int dpsA = NB * DIM * NB * DIM, dpsB = NB * DIM * NB, dpsC = NB * NB * DIM, dpsD = NB * NB;
    const int dpsX = NB * DIM, dpsY = NB;

    double acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = 0.0;
    double f = 0.0;

#pragma unroll
    for (int ez = 0; ez < (DIM == 3 ? 2 : 1); ++ez)
#pragma unroll
        for (int ey = 0; ey < 2; ++ey)
#pragma unroll
            for (int ex = 0; ex < 2; ++ex) {
                const int64_t k0 = x - 1 + ex, k1 = y - 1 + ey, k2 = (DIM == 3) ? z - 1 + ez : 0;
                if (k0 < 0 || k0 >= a.ne0 || k1 < 0 || k1 >= a.ne1) continue;
                if (DIM == 3 && (k2 < 0 || k2 >= a.ne2)) continue;
                const int64_t e = k0 + (int64_t)a.ne0 * (k1 + (int64_t)a.ne1 * k2);
                // local index bits of node i in this element
                const int abit[3] = {1 - ex, 1 - ey, 1 - ez};
                for (int q = 0; q < nq; ++q) {
                    double xi[3];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) xi[d] = a.reduced ? 0.5 : (((q >> d) & 1) ? g1 : g0);
                    double fa[3], Ga[3];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) fa[d] = abit[d] ? xi[d] : 1.0 - xi[d];
                    double Na = fa[0] * fa[1];
                    if (DIM == 3) Na *= fa[2];
#pragma unroll
                    for (int d = 0; d < DIM; ++d) {
                        double gq = (abit[d] ? 1.0 : -1.0) / h[d];
#pragma unroll
                        for (int o = 0; o < DIM; ++o)
                            if (o != d) gq *= fa[o];
                        Ga[d] = gq;
                    }
                    if (doS) {
                        double u[3] = {0.0, 0.0, 0.0};
                        double c0 = 0.0;
                        if (a.A.mode) {
                            const double* Ap = coef_pt(a.A, e, q, a.A.iso ? 1 : dpsA);
                            if (a.A.iso) {
#pragma unroll
                                for (int j = 0; j < DIM; ++j) u[j] = Ap[0] * Ga[j];
                            } else {
#pragma unroll
                                for (int j = 0; j < DIM; ++j)
#pragma unroll
                                    for (int ii = 0; ii < DIM; ++ii)
                                        u[j] += Ap[k + NB * (ii + DIM * (m + NB * j))] * Ga[ii];
                            }
                        }
                        if (a.C.mode) {
                            const double* Cp = coef_pt(a.C, e, q, dpsC);
#pragma unroll
                            for (int j = 0; j < DIM; ++j) u[j] += Na * Cp[k + NB * (m + NB * j)];
                        }
                        if (a.B.mode) {
                            const double* Bp = coef_pt(a.B, e, q, dpsB);
#pragma unroll
                            for (int ii = 0; ii < DIM; ++ii) c0 += Bp[k + NB * (ii + DIM * m)] * Ga[ii];
                        }
                        if (a.D.mode) {
                            const double* Dp = coef_pt(a.D, e, q, dpsD);
                            c0 += Dp[k + NB * m] * Na;
                        }
#pragma unroll
                        for (int b = 0; b < NNODE; ++b) {
                            const int bb[3] = {b & 1, (b >> 1) & 1, (b >> 2) & 1};
                            double fb[3];
#pragma unroll
                            for (int d = 0; d < DIM; ++d) fb[d] = bb[d] ? xi[d] : 1.0 - xi[d];
                            double Nb = fb[0] * fb[1];
                            if (DIM == 3) Nb *= fb[2];
                            double sdot = c0 * Nb;
#pragma unroll
                            for (int d = 0; d < DIM; ++d) {
                                double gq = (bb[d] ? 1.0 : -1.0) / h[d];
#pragma unroll
                                for (int o = 0; o < DIM; ++o)
                                    if (o != d) gq *= fb[o];
                                sdot += u[d] * gq;
                            }
                            const int s = (ex + bb[0]) + 3 * (ey + bb[1]) + (DIM == 3 ? 9 * (ez + bb[2]) : 0);
                            acc[s] += w * sdot;
                        }
                    }
                    if (doF) {
                        double sf = 0.0;
                        if (a.X.mode) {
                            const double* Xp = coef_pt(a.X, e, q, dpsX);
#pragma unroll
                            for (int ii = 0; ii < DIM; ++ii) sf += Xp[k + NB * ii] * Ga[ii];
                        }
                        if (a.Y.mode) {
                            const double* Yp = coef_pt(a.Y, e, q, dpsY);
                            sf += Yp[k] * Na;
                        }
                        f += w * sf;
                    }
                }
            }
    if (doS) {
#pragma unroll
        for (int t = 0; t < NST; ++t)
            val[(int64_t)(t * NB * NB + k + NB * m) * ld + i] += acc[FIRST + t];
    }
    if (doF) rhs[i * NB + k] += f;
}

// Element-constant isotropic A = kappa_e I (single PDE, full quadrature),
// optionally with element-constant Y and no other coefficient: the
// reference's constant-coefficient branch (DefaultAssembler3D.cpp:906-993,
// load :2205-2214; DefaultAssembler2D.cpp:301-326, :578-583).  The element
// matrix is then kappa_e K with K(a, b) = sum_d s_d(a, b) / h_d
// prod_{o != d} h_o m_o(a, b), s = +1 / -1 and m = 1/3 / 1/6 for equal /
// different local coordinates along an axis (the 2-point Gauss rule
// integrates these products exactly), so K depends only on which axes a and
// b agree in: kt[bit d set = agree along d] (2^DIM values, from the host).
// One thread per node gathers the 2^DIM elements around it; each value slot
// is written once (plainly when the matrix is known to hold zeros, ZERO),
// and Y contributes Y_e vol / 2^DIM to the node's rhs.
struct IsoArgs {
    const double* kappa;
    int64_t kstride;       // doubles between elements' kappa (0: one value for all)
    const double* Y;
    int64_t ystride;
    int64_t n0, n1, n2;
    int ne0, ne1, ne2;
    double kt[8];
    double fy;             // vol / 2^DIM
};

template <int DIM, bool SYM, bool ZERO>
__global__ __launch_bounds__(kBlock) void k_assemble_iso(IsoArgs a, double* __restrict__ val, int64_t ld, int64_t N,
                                                         double* __restrict__ rhs)
{
    constexpr int NS = StencilShape<DIM>::NS;
    constexpr int CEN = StencilShape<DIM>::CENTER;
    constexpr int NST = stored_slots<DIM, SYM>();
    constexpr int FIRST = SYM ? CEN : 0;
    constexpr int NNODE = 1 << DIM;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const int64_t x = i % a.n0, y = (i / a.n0) % a.n1, z = (DIM == 3) ? i / (a.n0 * a.n1) : 0;
    double kap[NNODE];
#pragma unroll
    for (int e = 0; e < NNODE; ++e) {
        const int ex = e & 1, ey = (e >> 1) & 1, ez = (e >> 2) & 1;
        const int64_t k0 = x - 1 + ex, k1 = y - 1 + ey, k2 = (DIM == 3) ? z - 1 + ez : 0;
        const bool in = k0 >= 0 && k0 < a.ne0 && k1 >= 0 && k1 < a.ne1 && (DIM == 2 || (k2 >= 0 && k2 < a.ne2));
        const int64_t el = in ? k0 + (int64_t)a.ne0 * (k1 + (int64_t)a.ne1 * k2) : 0;
        kap[e] = in ? a.kappa[a.kstride * el] : 0.0;
    }
    double acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = 0.0;
    // element e = (ex, ey, ez) holds node i at local bits (1-ex, 1-ey, 1-ez);
    // local node b of e is the neighbour at offset (ex+b0-1, ey+b1-1, ez+b2-1)
#pragma unroll
    for (int e = 0; e < NNODE; ++e) {
#pragma unroll
        for (int b = 0; b < NNODE; ++b) {
            const int agree = ((b & 1) == 1 - (e & 1) ? 1 : 0) | (((b >> 1) & 1) == 1 - ((e >> 1) & 1) ? 2 : 0) |
                              (DIM == 3 && ((b >> 2) & 1) == 1 - ((e >> 2) & 1) ? 4 : 0);
            const int s = ((e & 1) + (b & 1)) + 3 * (((e >> 1) & 1) + ((b >> 1) & 1)) +
                          (DIM == 3 ? 9 * (((e >> 2) & 1) + ((b >> 2) & 1)) : 0);
            acc[s] = fma(kap[e], a.kt[agree], acc[s]);
        }
    }
#pragma unroll
    for (int t = 0; t < NST; ++t) {
        double* p = &val[(int64_t)t * ld + i];
        *p = ZERO ? acc[FIRST + t] : *p + acc[FIRST + t];
    }
    if (rhs && a.Y) {
        double f = 0.0;
#pragma unroll
        for (int e = 0; e < NNODE; ++e) {
            const int ex = e & 1, ey = (e >> 1) & 1, ez = (e >> 2) & 1;
            const int64_t k0 = x - 1 + ex, k1 = y - 1 + ey, k2 = (DIM == 3) ? z - 1 + ez : 0;
            const bool in = k0 >= 0 && k0 < a.ne0 && k1 >= 0 && k1 < a.ne1 && (DIM == 2 || (k2 >= 0 && k2 < a.ne2));
            const int64_t el = in ? k0 + (int64_t)a.ne0 * (k1 + (int64_t)a.ne1 * k2) : 0;
            f += in ? a.Y[a.ystride * el] * a.fy : 0.0;
        }
        rhs[i] += f;
    }
}

// Boundary terms, colour free: one thread per (node i, equation k,
// component m) visits the face elements around node i on every face of the
// domain the node lies on and adds row a (node i's local face index) of
//   EM_S[k,m,a,b] = sum_q w d[k + n m] N_a N_b,   EM_F[k,a] = sum_q w y[k] N_a
// -- the face integrals of DefaultAssembler2D/3D::assemblePDEBoundarySingle/
// System (2-point Gauss rule along each tangential axis, weight face measure
// / points; DefaultAssembler2D.cpp:601-840, DefaultAssembler3D.cpp:2233-2890)
// or, reduced, the one-point rule at the face centre (:1014-1200,
// :3274-3630).  Face elements are indexed as the reference's
// FunctionOnBoundary samples: faces x0=0, x0=L, x1=0, x1=L[, x2=0, x2=L],
// within a face the lower tangential element index fastest (m_faceOffset).
struct BndArgs {
    CoefDev d, y;
    int64_t n[3];        // nodes per axis
    int ne[3];           // elements per axis
    double h[3];
    int64_t face_off[6];
    int reduced;
};

template <int DIM, int NB, bool SYM>
__global__ __launch_bounds__(kBlock) void k_assemble_boundary(BndArgs a, double* __restrict__ val, int64_t ld,
                                                              int64_t N, double* __restrict__ rhs)
{
    constexpr int NS = StencilShape<DIM>::NS;
    constexpr int CEN = StencilShape<DIM>::CENTER;
    constexpr int NST = stored_slots<DIM, SYM>();
    constexpr int FIRST = SYM ? CEN : 0;
    constexpr int NFN = 1 << (DIM - 1);                // nodes per face element
    const int km = blockIdx.y;
    const int k = km % NB, m = km / NB;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const int64_t c[3] = {i % a.n[0], (i / a.n[0]) % a.n[1], DIM == 3 ? i / (a.n[0] * a.n[1]) : 0};
    const bool doS = a.d.mode != 0;
    const bool doF = rhs != nullptr && m == 0 && a.y.mode != 0;
    const int nq = a.reduced ? 1 : NFN;
    const double g0 = 0.5 - 0.5 / 1.73205080756887719318, g1 = 0.5 + 0.5 / 1.73205080756887719318;
    double acc[NS];
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s] = 0.0;
    double f = 0.0;
    bool any = false;
#pragma unroll
    for (int face = 0; face < 2 * DIM; ++face) {
        const int ax = face / 2, side = face % 2;
        if (c[ax] != (side ? a.n[ax] - 1 : 0)) continue;
        any = true;
        const int t0 = (ax == 0) ? 1 : 0;
        const int t1 = (DIM == 3) ? ((ax == 2) ? 1 : 2) : -1;
        const double meas = a.h[t0] * (DIM == 3 ? a.h[t1] : 1.0);
        const double w = meas / nq;
#pragma unroll
        for (int e1 = 0; e1 < (DIM == 3 ? 2 : 1); ++e1)
#pragma unroll
            for (int e0 = 0; e0 < 2; ++e0) {
                const int64_t j0 = c[t0] - 1 + e0;
                const int64_t j1 = (DIM == 3) ? c[t1] - 1 + e1 : 0;
                if (j0 < 0 || j0 >= a.ne[t0]) continue;
                if (DIM == 3 && (j1 < 0 || j1 >= a.ne[t1])) continue;
                const int64_t fe = a.face_off[face] + j0 + (int64_t)a.ne[t0] * j1;
                const int ab0 = 1 - e0, ab1 = 1 - e1;      // node i's local face bits
                for (int q = 0; q < nq; ++q) {
                    const double xi0 = a.reduced ? 0.5 : ((q & 1) ? g1 : g0);
                    const double xi1 = a.reduced ? 0.5 : (((q >> 1) & 1) ? g1 : g0);
                    const double Na = (ab0 ? xi0 : 1.0 - xi0) * (DIM == 3 ? (ab1 ? xi1 : 1.0 - xi1) : 1.0);
                    if (doS) {
                        const double dv = coef_pt(a.d, fe, q, NB * NB)[k + NB * m];
#pragma unroll
                        for (int b = 0; b < NFN; ++b) {
                            const int bb0 = b & 1, bb1 = (b >> 1) & 1;
                            const double Nb = (bb0 ? xi0 : 1.0 - xi0) * (DIM == 3 ? (bb1 ? xi1 : 1.0 - xi1) : 1.0);
                            int off[3] = {1, 1, 1};
                            off[t0] = e0 + bb0;
                            if (DIM == 3) off[t1] = e1 + bb1;
                            const int s = off[0] + 3 * off[1] + (DIM == 3 ? 9 * off[2] : 0);
                            acc[s] += w * dv * Na * Nb;
                        }
                    }
                    if (doF) f += w * coef_pt(a.y, fe, q, NB)[k] * Na;
                }
            }
    }
    if (!any) return;
    if (doS) {
#pragma unroll
        for (int t = 0; t < NST; ++t)
            val[(int64_t)(t * NB * NB + k + NB * m) * ld + i] += acc[FIRST + t];
    }
    if (doF) rhs[i * NB + k] += f;
}

// Dirac points in order (RipleyDomain::assemblePDEDirac, RipleyDomain.cpp:
// 1756-1800): rhs[k + n*node] += y[p][k]; diagonal block += d[p] (k + n*m).
template <int NB>
__global__ void k_add_dirac(int64_t count, const int64_t* __restrict__ nodes, const double* __restrict__ d,
                            const double* __restrict__ y, double* __restrict__ val, int64_t ld, int diag_slot,
                            double* __restrict__ rhs)
{
    if (threadIdx.x >= NB * NB) return;
    const int k = threadIdx.x % NB, m = threadIdx.x / NB;
    for (int64_t p = 0; p < count; ++p) {
        const int64_t i = nodes[p];
        if (d) val[(int64_t)(diag_slot * NB * NB + k + NB * m) * ld + i] += d[p * NB * NB + k + NB * m];
        if (y && rhs && m == 0) rhs[i * NB + k] += y[p * NB + k];
    }
}

void launch_assemble_boundary(rg_matrix* m, const rg_coef* d, const rg_coef* y, int reduced, double* rhs)
{
    const Geom& g = m->g;
    auto cv = [](const rg_coef* x) {
        return x ? CoefDev{x->data, x->data ? x->mode : 0, x->elem_stride, 0} : CoefDev{nullptr, 0, 0, 0};
    };
    BndArgs a;
    a.d = cv(d);
    a.y = cv(y);
    for (const rg_coef* x : {d, y}) {
        if (!x) continue;
        RG_CHECK(x->mode >= RG_COEF_EMPTY && x->mode <= RG_COEF_EXPANDED, RG_VALUE_ERROR,
                 "assemble_boundary: unknown coefficient mode");
        RG_CHECK(x->mode == RG_COEF_EMPTY || x->data != nullptr, RG_VALUE_ERROR,
                 "assemble_boundary: coefficient data is NULL");
        RG_CHECK(!(reduced && x->mode == RG_COEF_EXPANDED), RG_VALUE_ERROR,
                 "assemble_boundary: reduced-order coefficients cannot be expanded");
        RG_CHECK(!x->isotropic, RG_VALUE_ERROR, "assemble_boundary: isotropic is defined for A only");
    }
    for (int t = 0; t < 3; ++t) {
        a.n[t] = g.nn[t];
        a.ne[t] = g.ne[t];
        a.h[t] = g.h[t];
    }
    int64_t off = 0;
    for (int f = 0; f < 6; ++f) {
        a.face_off[f] = off;
        if (f >= 2 * g.dim) continue;
        const int ax = f / 2;
        int64_t cnt = 1;
        for (int t = 0; t < g.dim; ++t)
            if (t != ax) cnt *= g.ne[t];
        off += cnt;
    }
    a.reduced = reduced;
    if (!a.d.mode && !(a.y.mode && rhs)) return;
    const dim3 grid((unsigned)((g.N + kBlock - 1) / kBlock), (unsigned)(g.bs * g.bs));
    dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
        hipLaunchKernelGGL((k_assemble_boundary<D.value, B.value, S.value>), grid, dim3(kBlock), 0, m->ctx->stream, a,
                           m->val, g.ld, g.N, rhs);
    });
    RG_HIP(hipGetLastError());
    m->val32_valid = false;
    if (a.d.mode) m->val_zero = false;
}

void launch_add_dirac(rg_matrix* m, int64_t count, const int64_t* d_nodes, const double* d_d, const double* d_y,
                      double* rhs)
{
    const Geom& g = m->g;
    const int diag_slot = m->symmetric ? 0 : (g.dim == 3 ? 13 : 4);
    dispatch_bs(g.bs, [&](auto B) {
        hipLaunchKernelGGL((k_add_dirac<B.value>), dim3(1), dim3(64), 0, m->ctx->stream, count, d_nodes, d_d, d_y,
                           m->val, g.ld, diag_slot, rhs);
    });
    RG_HIP(hipGetLastError());
    m->val32_valid = false;
    if (d_d) m->val_zero = false;
}

void launch_assemble(rg_matrix* m, const rg_pde_coefs& c, int reduced, double* rhs)
{
    const Geom& g = m->g;
    AsmArgs a;
    auto cv = [](const rg_coef& x) { return CoefDev{x.data, x.data ? x.mode : 0, x.elem_stride, x.isotropic}; };
    a.A = cv(c.A); a.B = cv(c.B); a.C = cv(c.C); a.D = cv(c.D); a.X = cv(c.X); a.Y = cv(c.Y);
    const rg_coef* all[6] = {&c.A, &c.B, &c.C, &c.D, &c.X, &c.Y};
    for (int t = 0; t < 6; ++t) {
        const rg_coef& x = *all[t];
        RG_CHECK(x.mode >= RG_COEF_EMPTY && x.mode <= RG_COEF_EXPANDED, RG_VALUE_ERROR,
                 "assemble: unknown coefficient mode");
        RG_CHECK(x.mode == RG_COEF_EMPTY || x.data != nullptr, RG_VALUE_ERROR,
                 "assemble: coefficient data is NULL");
        RG_CHECK(!(reduced && x.mode == RG_COEF_EXPANDED), RG_VALUE_ERROR,
                 "assemble: reduced-order coefficients cannot be expanded");
    }
    RG_CHECK(!c.A.isotropic || g.bs == 1, RG_VALUE_ERROR,
             "assemble: isotropic A is defined for a single PDE only");
    a.n0 = g.nn[0]; a.n1 = g.nn[1]; a.n2 = g.nn[2];
    a.ne0 = g.ne[0]; a.ne1 = g.ne[1]; a.ne2 = g.ne[2];
    a.h0 = g.h[0]; a.h1 = g.h[1]; a.h2 = g.h[2];
    a.reduced = reduced;
    const bool iso_const = g.bs == 1 && !reduced && c.A.mode == RG_COEF_ELEMENT && c.A.isotropic && c.A.data &&
                           !c.B.mode && !c.C.mode && !c.D.mode && !c.X.mode &&
                           (c.Y.mode == RG_COEF_EMPTY || c.Y.mode == RG_COEF_ELEMENT);
    if (iso_const) {
        IsoArgs ia;
        ia.kappa = c.A.data;
        ia.kstride = c.A.elem_stride;
        ia.Y = (c.Y.mode == RG_COEF_ELEMENT) ? c.Y.data : nullptr;
        ia.ystride = c.Y.elem_stride;
        ia.n0 = g.nn[0]; ia.n1 = g.nn[1]; ia.n2 = g.nn[2];
        ia.ne0 = g.ne[0]; ia.ne1 = g.ne[1]; ia.ne2 = g.ne[2];
        double vol = 1.0;
        for (int d = 0; d < g.dim; ++d) vol *= g.h[d];
        ia.fy = vol / (1 << g.dim);
        // K(agree) = sum_d s_d / h_d prod_{o != d} h_o m_o
        for (int ag = 0; ag < (1 << g.dim); ++ag) {
            double kt = 0.0;
            for (int d = 0; d < g.dim; ++d) {
                double t = (((ag >> d) & 1) ? 1.0 : -1.0) / g.h[d];
                for (int o = 0; o < g.dim; ++o)
                    if (o != d) t *= g.h[o] * (((ag >> o) & 1) ? 1.0 / 3.0 : 1.0 / 6.0);
                kt += t;
            }
            ia.kt[ag] = kt;
        }
        for (int ag = 1 << g.dim; ag < 8; ++ag) ia.kt[ag] = 0.0;
        const dim3 grid((unsigned)((g.N + kBlock - 1) / kBlock));
        auto go = [&](auto D, auto S, auto Z) {
            hipLaunchKernelGGL((k_assemble_iso<D.value, S.value, Z.value>), grid, dim3(kBlock), 0, m->ctx->stream, ia,
                               m->val, g.ld, g.N, rhs);
        };
        auto byz = [&](auto D, auto S) {
            if (m->val_zero) go(D, S, BC<true>{});
            else go(D, S, BC<false>{});
        };
        if (g.dim == 3) {
            if (m->symmetric) byz(IC<3>{}, BC<true>{});
            else byz(IC<3>{}, BC<false>{});
        } else {
            if (m->symmetric) byz(IC<2>{}, BC<true>{});
            else byz(IC<2>{}, BC<false>{});
        }
    } else {
        const dim3 grid((unsigned)((g.N + kBlock - 1) / kBlock), (unsigned)(g.bs * g.bs));
        dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
            hipLaunchKernelGGL((k_assemble<D.value, B.value, S.value>), grid, dim3(kBlock), 0,
                               m->ctx->stream, a, m->val, g.ld, g.N, rhs);
        });
    }
    RG_HIP(hipGetLastError());
    m->val32_valid = false;
    if (c.A.mode || c.B.mode || c.C.mode || c.D.mode) m->val_zero = false;
}

// -------------------------------------------------------------------- SpMV
// paso SystemMatrix::MatrixVector -> SparseMatrix_MatrixVector_CSR_OFFSET0
// (paso/src/SparseMatrix_MatrixVector.cpp:191-336): out = alpha*A*in +
// beta*out, out first scaled by beta (zeroed for beta == 0).  BETA selects
// the epilogue at compile time (0: beta == 0, 1: beta == 1, 2: general): a
// run-time branch on beta inside the row loop cost 150 registers.
template <int DIM, int BS, bool SYM, typename VT, int BETA>
__global__ __launch_bounds__(kBlock) void k_spmv(const VT* __restrict__ val, int64_t ld, int64_t N,
                                                 int64_t s1, int64_t s2, int64_t chunk,
                                                 const double* __restrict__ x, double* __restrict__ y,
                                                 double alpha, double beta)
{
    const RowRange rr = block_rows(N, chunk);
    stencil_block<DIM, BS, SYM, VT>(val, ld, N, s1, s2, rr.r0, rr.r1, x, [&](int64_t i, const double (&acc)[BS]) {
#pragma unroll
        for (int r = 0; r < BS; ++r) {
            if (BETA == 0) y[i * BS + r] = alpha * acc[r];
            else if (BETA == 1) y[i * BS + r] = fma(alpha, acc[r], y[i * BS + r]);
            else y[i * BS + r] = fma(alpha, acc[r], beta * y[i * BS + r]);
        }
    });
}

void launch_spmv(rg_matrix* m, double alpha, const double* x, double beta, double* y)
{
    launch_spmv_on(m, alpha, x, beta, y, m->ctx->stream);
}

void launch_spmv_on(rg_matrix* m, double alpha, const double* x, double beta, double* y, hipStream_t st)
{
    const Geom& g = m->g;
    const Launch L = chunked_launch(g.N, spmv_rows_per_thread<double>());
    dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
        if (beta == 0.0)
            hipLaunchKernelGGL((k_spmv<D.value, B.value, S.value, double, 0>), dim3(L.grid), dim3(kBlock), 0,
                               st, m->val, g.ld, g.N, g.s1, g.s2, L.chunk, x, y, alpha, beta);
        else if (beta == 1.0)
            hipLaunchKernelGGL((k_spmv<D.value, B.value, S.value, double, 1>), dim3(L.grid), dim3(kBlock), 0,
                               st, m->val, g.ld, g.N, g.s1, g.s2, L.chunk, x, y, alpha, beta);
        else
            hipLaunchKernelGGL((k_spmv<D.value, B.value, S.value, double, 2>), dim3(L.grid), dim3(kBlock), 0,
                               st, m->val, g.ld, g.N, g.s1, g.s2, L.chunk, x, y, alpha, beta);
    });
    RG_HIP(hipGetLastError());
}

void launch_spmv32(rg_matrix* m, const double* x, double* y)
{
    const Geom& g = m->g;
    const Launch L = chunked_launch(g.N, spmv_rows_per_thread<float>());
    dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
        hipLaunchKernelGGL((k_spmv<D.value, B.value, S.value, float, 0>), dim3(L.grid), dim3(kBlock), 0,
                           m->ctx->stream, m->val32, g.ld, g.N, g.s1, g.s2, L.chunk, x, y, 1.0, 0.0);
    });
    RG_HIP(hipGetLastError());
}

// ---------------------------------------------------------------- nullify
// SparseMatrix::nullifyRowsAndCols_CSR(_BLK1) (paso/src/SparseMatrix.h:
// 494-535): a(i r, j c) := (i r == j c ? mdv : 0) where row_mask[i r] > 0 or
// col_mask[j c] > 0, over the PASO pattern.  With symmetric storage the two
// masks must be equal (checked by k_mask_differs).
template <int DIM, int BS, bool SYM>
__global__ __launch_bounds__(kBlock) void k_nullify(double* __restrict__ val, int64_t ld, int64_t N,
                                                    int64_t n0, int64_t n1, int64_t n2,
                                                    const double* __restrict__ mrow,
                                                    const double* __restrict__ mcol, double mdv)
{
    constexpr int NST = stored_slots<DIM, SYM>();
    constexpr int FIRST = SYM ? StencilShape<DIM>::CENTER : 0;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const int64_t x = i % n0, y = (i / n0) % n1, z = i / (n0 * n1);
    bool anyrow = false;
    double mr[BS];
#pragma unroll
    for (int r = 0; r < BS; ++r) {
        mr[r] = mrow[i * BS + r];
        anyrow |= mr[r] > 0.0;
    }
#pragma unroll
    for (int t = 0; t < NST; ++t) {
        const int s = FIRST + t;
        if (!slot_in_grid<DIM>(s, x, y, z, n0, n1, n2)) continue;
        int dx, dy, dz;
        slot_delta<DIM>(s, dx, dy, dz);
        const int64_t j = i + dx + dy * n0 + dz * n0 * n1;
        bool anycol = false;
        double mc[BS];
#pragma unroll
        for (int c = 0; c < BS; ++c) {
            mc[c] = mcol[j * BS + c];
            anycol |= mc[c] > 0.0;
        }
        if (!anyrow && !anycol) continue;
#pragma unroll
        for (int c = 0; c < BS; ++c)
#pragma unroll
            for (int r = 0; r < BS; ++r)
                if (mr[r] > 0.0 || mc[c] > 0.0)
                    val[(int64_t)(t * BS * BS + r + BS * c) * ld + i] = (j == i && r == c) ? mdv : 0.0;
    }
}

__global__ void k_mask_differs(const double* a, const double* b, int64_t n, int* flag)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        if ((a[i] > 0.0) != (b[i] > 0.0)) *flag = 1;
}

void launch_nullify(rg_matrix* m, const double* row_mask, const double* col_mask, double mdv, int* d_flag)
{
    const Geom& g = m->g;
    const int grid = (int)((g.N + kBlock - 1) / kBlock);
    if (m->symmetric && row_mask != col_mask) {
        hipLaunchKernelGGL(k_mask_differs, dim3(std::min(grid, 1024)), dim3(kBlock), 0, m->ctx->stream,
                           row_mask, col_mask, g.N * g.bs, d_flag);
        RG_HIP(hipGetLastError());
        int h = 0;
        RG_HIP(hipMemcpyAsync(&h, d_flag, sizeof(int), hipMemcpyDeviceToHost, m->ctx->stream));
        RG_HIP(hipStreamSynchronize(m->ctx->stream));
        RG_CHECK(h == 0, RG_VALUE_ERROR,
                 "nullifyRowsAndCols: a matrix with symmetric storage needs equal row and column masks");
    }
    dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
        hipLaunchKernelGGL((k_nullify<D.value, B.value, S.value>), dim3(grid), dim3(kBlock), 0,
                           m->ctx->stream, m->val, g.ld, g.N, g.nn[0], g.nn[1], g.nn[2], row_mask,
                           col_mask, mdv);
    });
    RG_HIP(hipGetLastError());
    m->val32_valid = false;
    m->val_zero = false;
}

// ------------------------------------------------------------- CSR export
// The PASO main block in CSR form: row i lists its in-grid neighbours in
// ascending order (ripley/src/Brick.cpp:3000-3061) with the block values
// val[k*bs*bs + ir + bs*ic] (paso/src/SystemMatrix.h:404-492).
template <int DIM, int BS, bool SYM>
__global__ __launch_bounds__(kBlock) void k_export_csr(const double* __restrict__ val, int64_t ld,
                                                       int64_t N, int64_t n0, int64_t n1, int64_t n2,
                                                       const int32_t* __restrict__ ptr,
                                                       int32_t* __restrict__ index,
                                                       double* __restrict__ out)
{
    constexpr int NS = StencilShape<DIM>::NS;
    constexpr int C = StencilShape<DIM>::CENTER;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= N) return;
    const int64_t x = i % n0, y = (i / n0) % n1, z = i / (n0 * n1);
    int64_t k = ptr[i];
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        if (!slot_in_grid<DIM>(s, x, y, z, n0, n1, n2)) continue;
        int dx, dy, dz;
        slot_delta<DIM>(s, dx, dy, dz);
        const int64_t j = i + dx + dy * n0 + dz * n0 * n1;
        if (index) index[k] = (int32_t)j;
        if (out) {
#pragma unroll
            for (int c = 0; c < BS; ++c)
#pragma unroll
                for (int r = 0; r < BS; ++r) {
                    double v;
                    if (!SYM || s >= C)
                        v = val[(int64_t)((SYM ? s - C : s) * BS * BS + r + BS * c) * ld + i];
                    else
                        v = val[(int64_t)((C - s) * BS * BS + c + BS * r) * ld + j];
                    out[k * BS * BS + r + BS * c] = v;
                }
        }
        ++k;
    }
}

void launch_export_csr_values(rg_matrix* m, const int32_t* d_ptr, int32_t* d_index, double* d_val)
{
    const Geom& g = m->g;
    const int grid = (int)((g.N + kBlock - 1) / kBlock);
    dispatch(g, m->symmetric, [&](auto D, auto B, auto S) {
        hipLaunchKernelGGL((k_export_csr<D.value, B.value, S.value>), dim3(grid), dim3(kBlock), 0,
                           m->ctx->stream, m->val, g.ld, g.N, g.nn[0], g.nn[1], g.nn[2], d_ptr,
                           d_index, d_val);
    });
    RG_HIP(hipGetLastError());
}

// ---------------------------------------------------------- FP32 values
__global__ void k_to_fp32(const double* __restrict__ a, float* __restrict__ b, int64_t n)
{
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        b[i] = (float)a[i];
}

void launch_to_fp32(rg_matrix* m)
{
    const Geom& g = m->g;
    const int64_t n = (int64_t)g.nstore * g.bs * g.bs * g.ld;
    hipLaunchKernelGGL(k_to_fp32, dim3(4096), dim3(kBlock), 0, m->ctx->stream, m->val, m->val32, n);
    RG_HIP(hipGetLastError());
}

// ----------------------------------------------------------- host pattern
void csr_row_ptr(const Geom& g, int32_t* ptr)
{
    const int64_t n0 = g.nn[0], n1 = g.nn[1], n2 = g.nn[2];
    ptr[0] = 0;
    int64_t acc = 0;
    for (int64_t i = 0; i < g.N; ++i) {
        const int64_t x = i % n0, y = (i / n0) % n1, z = i / (n0 * n1);
        const int cx = 1 + (x > 0) + (x < n0 - 1);
        const int cy = 1 + (y > 0) + (y < n1 - 1);
        const int cz = (g.dim == 3) ? 1 + (z > 0) + (z < n2 - 1) : 1;
        acc += cx * cy * cz;
        ptr[i + 1] = (int32_t)acc;
    }
}

int64_t csr_nnz(const Geom& g)
{
    int64_t n = 1;
    for (int d = 0; d < g.dim; ++d) n *= 3 * g.nn[d] - 2;
    return n;
}

}  // namespace rg
