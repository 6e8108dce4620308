// ORACLE -- test infrastructure only.
//
// A CPU restatement (C++17 + OpenMP) of the esys-escript hot path that the
// HIP library replaces: ripley Rectangle/Brick Q1 assembly into a PASO CSR
// system matrix, Dirichlet constraints, and PASO's Jacobi-preconditioned CG
// driver, all at one MPI rank.  It is used only as the checker by tests/,
// by __graft_entry__.smoke(), and as the timed CPU baseline ("port") leg of
// bench.py.  Nothing in the product links or calls it.
//
// Every function cites the reference file:line it restates.  How each part is
// pinned (golden vectors, closed forms, invariants) is listed next to the
// tests that do it (tests/test_oracle_*.py).  Functions with no such check
// are marked "parity unpinned" in their comment; at present there are none.
//
// Conventions (all taken from the reference):
//   * node / DOF numbering: lexicographic, x fastest (Brick.cpp:3000-3061)
//   * element e = k0 + NE0*(k1 + NE1*k2); firstNode = k0 + NN0*(k1 + NN1*k2)
//     and local node a has offset bits (a&1, (a>>1)&1, (a>>2)&1)
//     (Brick.cpp:3244-3251, Rectangle.cpp:2233-2237)
//   * element matrix EM_S[k + n*(m + n*(a + nN*b))], EM_F[k + n*a]
//     (RipleyDomain.cpp:1508-1517, Brick.cpp:3253-3260)
//   * PASO block value val[iptr*bs*bs + ir + bs*ic] (column-major block)
//   * coefficient data points are escript's column-major layout; expanded
//     data stores the quadrature point index after the data-point components
//     (DefaultAssembler3D.cpp:259-330 A_p[INDEX3(i,j,q,3,3)]).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>
#include <omp.h>

typedef int32_t index_t;
#define ORC_API extern "C"

// ---------------------------------------------------------------- geometry

// Number of nodes of a Rectangle (dim 2) or Brick (dim 3) with ne elements
// per axis at one rank (Rectangle.h:352-424, Brick.h:352-416).
ORC_API int64_t orc_num_nodes(int dim, const int* ne)
{
    int64_t n = 1;
    for (int d = 0; d < dim; ++d) n *= (int64_t)(ne[d] + 1);
    return n;
}

// Stencil offsets in getDiagonalIndices order: i2 outer, i0 inner, each in
// {-1,0,1}; offset = i2*nDOF0*nDOF1 + i1*nDOF0 + i0.  upper_only keeps the
// non-negative ones (Brick.cpp:2484-2507, Rectangle.cpp:1766-1785).
ORC_API int orc_diagonal_indices(int dim, const int* ndof, int upper_only, int32_t* out)
{
    int idx = 0;
    if (dim == 2) {
        for (int i1 = -1; i1 < 2; ++i1)
            for (int i0 = -1; i0 < 2; ++i0) {
                const int o = i1 * ndof[0] + i0;
                if (!upper_only || o >= 0) out[idx++] = o;
            }
    } else {
        for (int i2 = -1; i2 < 2; ++i2)
            for (int i1 = -1; i1 < 2; ++i1)
                for (int i0 = -1; i0 < 2; ++i0) {
                    const int o = i2 * ndof[0] * ndof[1] + i1 * ndof[0] + i0;
                    if (!upper_only || o >= 0) out[idx++] = o;
                }
    }
    return idx;
}

// PASO main-block pattern of a one-rank Rectangle/Brick: row i holds the
// in-bounds neighbours of node i, generated in (i2,i1,i0) loop order, which is
// ascending (Brick.cpp:3000-3061 getConnections, RipleyDomain.cpp:1364-1383
// createPasoPattern, Rectangle.cpp:2035-2085).  ptr has N+1 entries.  If
// index is NULL only ptr is filled.  Returns nnz (number of block entries).
ORC_API int64_t orc_pattern(int dim, const int* nn, index_t* ptr, index_t* index)
{
    const int64_t n0 = nn[0], n1 = nn[1], n2 = (dim == 3 ? nn[2] : 1);
    const int64_t N = n0 * n1 * n2;
    ptr[0] = 0;
    for (int64_t i = 0; i < N; ++i) {
        const int64_t x = i % n0, y = (i / n0) % n1, z = i / (n0 * n1);
        const int cx = 1 + (x > 0) + (x < n0 - 1);
        const int cy = 1 + (y > 0) + (y < n1 - 1);
        const int cz = (dim == 3) ? 1 + (z > 0) + (z < n2 - 1) : 1;
        ptr[i + 1] = ptr[i] + cx * cy * cz;
    }
    if (index) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < N; ++i) {
            const int64_t x = i % n0, y = (i / n0) % n1, z = i / (n0 * n1);
            index_t k = ptr[i];
            const int zlo = (dim == 3) ? -1 : 0, zhi = (dim == 3) ? 1 : 0;
            for (int dz = zlo; dz <= zhi; ++dz) {
                const int64_t zz = z + dz;
                if (zz < 0 || zz >= n2) continue;
                for (int dy = -1; dy <= 1; ++dy) {
                    const int64_t yy = y + dy;
                    if (yy < 0 || yy >= n1) continue;
                    for (int dx = -1; dx <= 1; ++dx) {
                        const int64_t xx = x + dx;
                        if (xx < 0 || xx >= n0) continue;
                        index[k++] = (index_t)(xx + n0 * (yy + n1 * zz));
                    }
                }
            }
        }
    }
    return ptr[N];
}

// ------------------------------------------------------- element matrices

// Coefficient descriptor.  mode 0: empty; 1: one data point per element
// (the reference's "constant data" branch, which reads getSampleDataRO(e) and
// so covers DataConstant and per-element DataTagged values); 2: expanded, one
// data point per quadrature point ("actsExpanded" branch).  For mode 1 a
// stride of 0 means the same data point for every element.
typedef struct {
    const double* data;
    int mode;
    int elem_stride;   // doubles between elements (0 = shared constant)
} orc_coef;

static const double SQRT3 = 1.73205080756887719318;

// 2-point Gauss points on [0,1]: index 0 is (1-1/sqrt3)/2, next to the
// element's local node 0; quadrature points are numbered x fastest
// (DefaultAssembler2D.cpp:563-586, DefaultAssembler3D.cpp:2168-2215).
static inline double gauss_pt(int bit) { return bit ? 0.5 + 0.5 / SQRT3 : 0.5 - 0.5 / SQRT3; }

struct ShapeTab {
    int dim, nN, nQ;
    double w;                 // quadrature weight (element volume / nQ)
    double N[8][8];           // N[q][a]
    double G[8][8][3];        // G[q][a][i] = dN_a/dx_i at q
};

// reduced != 0 selects the one-point rule at the element centre used by
// assemblePDESingleReduced/assemblePDESystemReduced
// (DefaultAssembler3D.cpp:2891-3270: D weight h0h1h2/64 = vol*(1/8)^2).
static void make_shape(int dim, const double* h, ShapeTab* t, int reduced = 0)
{
    t->dim = dim;
    t->nN = 1 << dim;
    t->nQ = reduced ? 1 : 1 << dim;
    double vol = 1.0;
    for (int d = 0; d < dim; ++d) vol *= h[d];
    t->w = vol / t->nQ;
    for (int q = 0; q < t->nQ; ++q) {
        double xi[3];
        for (int d = 0; d < dim; ++d) xi[d] = reduced ? 0.5 : gauss_pt((q >> d) & 1);
        for (int a = 0; a < t->nN; ++a) {
            double f[3], df[3];
            for (int d = 0; d < dim; ++d) {
                const int b = (a >> d) & 1;
                f[d] = b ? xi[d] : 1.0 - xi[d];
                df[d] = (b ? 1.0 : -1.0) / h[d];
            }
            double prod = 1.0;
            for (int d = 0; d < dim; ++d) prod *= f[d];
            t->N[q][a] = prod;
            for (int i = 0; i < dim; ++i) {
                double g = df[i];
                for (int d = 0; d < dim; ++d)
                    if (d != i) g *= f[d];
                t->G[q][a][i] = g;
            }
        }
    }
}

static inline const double* coef_point(const orc_coef& c, int64_t e, int q, int dps)
{
    if (c.mode == 1) return c.data + (int64_t)c.elem_stride * e;
    return c.data + (int64_t)c.elem_stride * e + (int64_t)q * dps;
}

// Element matrix and load vector of one element by 2x2(x2) Gauss quadrature:
//   EM_S[k,m,a,b] += sum_q w ( A[k,i,m,j] dN_a/dx_i dN_b/dx_j
//                             + B[k,i,m] N_b dN_a/dx_i
//                             + C[k,m,j] dN_b/dx_j N_a + D[k,m] N_a N_b )
//   EM_F[k,a]     += sum_q w ( X[k,i] dN_a/dx_i + Y[k] N_a )
// which is the weak form integrated by DefaultAssembler{2D,3D}::
// assemblePDESingle/assemblePDESystem (e.g. DefaultAssembler3D.cpp:906-993
// for constant A, :1340-1360 constant B, :1784 constant C, :2160-2215 Y).
// Coefficient layouts per data point (numEq = numComp = n):
//   A[k + n*(i + dim*(m + n*j))], B[k + n*(i + dim*m)], C[k + n*(m + n*j)],
//   D[k + n*m], X[k + n*i], Y[k]   (for n == 1 these reduce to the single
//   PDE layouts A[i + dim*j], B[i], C[j], D, X[i], Y).
static void element_matrices(const ShapeTab& t, int n, int64_t e,
                             const orc_coef* A, const orc_coef* B, const orc_coef* C,
                             const orc_coef* D, const orc_coef* X, const orc_coef* Y,
                             double* EM_S, double* EM_F)
{
    const int dim = t.dim, nN = t.nN;
    const int dpsA = n * dim * n * dim, dpsB = n * dim * n, dpsC = n * n * dim;
    const int dpsD = n * n, dpsX = n * dim, dpsY = n;
    for (int q = 0; q < t.nQ; ++q) {
        const double* Ap = A->mode ? coef_point(*A, e, q, dpsA) : nullptr;
        const double* Bp = B->mode ? coef_point(*B, e, q, dpsB) : nullptr;
        const double* Cp = C->mode ? coef_point(*C, e, q, dpsC) : nullptr;
        const double* Dp = D->mode ? coef_point(*D, e, q, dpsD) : nullptr;
        const double* Xp = X->mode ? coef_point(*X, e, q, dpsX) : nullptr;
        const double* Yp = Y->mode ? coef_point(*Y, e, q, dpsY) : nullptr;
        if (EM_S && (Ap || Bp || Cp || Dp)) {
            for (int b = 0; b < nN; ++b)
                for (int a = 0; a < nN; ++a)
                    for (int m = 0; m < n; ++m)
                        for (int k = 0; k < n; ++k) {
                            double s = 0.0;
                            if (Ap)
                                for (int j = 0; j < dim; ++j)
                                    for (int i = 0; i < dim; ++i)
                                        s += Ap[k + n * (i + dim * (m + n * j))] * t.G[q][a][i] * t.G[q][b][j];
                            if (Bp)
                                for (int i = 0; i < dim; ++i)
                                    s += Bp[k + n * (i + dim * m)] * t.N[q][b] * t.G[q][a][i];
                            if (Cp)
                                for (int j = 0; j < dim; ++j)
                                    s += Cp[k + n * (m + n * j)] * t.G[q][b][j] * t.N[q][a];
                            if (Dp) s += Dp[k + n * m] * t.N[q][a] * t.N[q][b];
                            EM_S[k + n * (m + n * (a + nN * b))] += t.w * s;
                        }
        }
        if (EM_F && (Xp || Yp)) {
            for (int a = 0; a < nN; ++a)
                for (int k = 0; k < n; ++k) {
                    double s = 0.0;
                    if (Xp)
                        for (int i = 0; i < dim; ++i) s += Xp[k + n * i] * t.G[q][a][i];
                    if (Yp) s += Yp[k] * t.N[q][a];
                    EM_F[k + n * a] += t.w * s;
                }
        }
    }
}

// One element's EM_S / EM_F (exported for the closed-form tests).
ORC_API void orc_element_matrices(int dim, const double* h, int n, int reduced,
                          const orc_coef* A, const orc_coef* B, const orc_coef* C,
                          const orc_coef* D, const orc_coef* X, const orc_coef* Y,
                          double* EM_S, double* EM_F)
{
    ShapeTab t;
    make_shape(dim, h, &t, reduced);
    const int nN = t.nN;
    if (EM_S) std::fill(EM_S, EM_S + n * n * nN * nN, 0.0);
    if (EM_F) std::fill(EM_F, EM_F + n * nN, 0.0);
    element_matrices(t, n, 0, A, B, C, D, X, Y, EM_S, EM_F);
}

// ------------------------------------------------------------- insertion

// addToPasoMatrix, CSR branch at one rank: for each (row node, column node)
// pair, linear search of the row for the column, then UPDATE_BLOCK
// val[k*bs*bs + ir + bs*ic] += EM[ir + n*(ic + n*(kEq + nN*kSol))]
// (RipleyDomain.cpp:1488-1517, :1566-1613).  Returns the number of
// (row, column) pairs that were not found in the pattern.
ORC_API int orc_add_to_csr(int bs, int64_t nrows, const index_t* ptr, const index_t* index, double* val,
                   int nN, const index_t* nodes, const double* EM)
{
    int missing = 0;
    const int n = bs;
    for (int kEq = 0; kEq < nN; ++kEq) {
        const index_t row = nodes[kEq];
        if (row < 0 || row >= nrows) continue;
        for (int kSol = 0; kSol < nN; ++kSol) {
            const index_t col = nodes[kSol];
            bool found = false;
            for (index_t k = ptr[row]; k < ptr[row + 1]; ++k) {
                if (index[k] == col) {
                    for (int ic = 0; ic < bs; ++ic)
                        for (int ir = 0; ir < bs; ++ir)
                            val[(int64_t)k * bs * bs + ir + bs * ic] +=
                                EM[ir + n * (ic + n * (kEq + nN * kSol))];
                    found = true;
                    break;
                }
            }
            if (!found) ++missing;
        }
    }
    return missing;
}

// assemblePDESingle / assemblePDESystem over all elements followed by
// addToMatrixAndRHS (DefaultAssembler2D.cpp:152-597,
// DefaultAssembler3D.cpp:153-2229 and :3631-5745; Brick.cpp:3239-3265;
// Rectangle.cpp:2228-2250).  Elements are visited in the reference's order:
// colour (parity of the outermost element index), slab, then inner indices;
// slabs of one colour run in parallel, which is race free as in the reference.
// n = numEq = numComp = block size.  val/rhs may be NULL to skip them.
// reduced != 0 assembles the *_reduced coefficients with the one-point rule
// (assemblePDESingleReduced DefaultAssembler3D.cpp:2891-3270,
// DefaultAssembler2D.cpp:842-1011).
ORC_API void orc_assemble(int dim, const int* ne, const double* len, int n, int reduced,
                  const orc_coef* A, const orc_coef* B, const orc_coef* C,
                  const orc_coef* D, const orc_coef* X, const orc_coef* Y,
                  const index_t* ptr, const index_t* index, double* val, double* rhs)
{
    double h[3] = {1.0, 1.0, 1.0};
    for (int d = 0; d < dim; ++d) h[d] = len[d] / ne[d];
    ShapeTab t;
    make_shape(dim, h, &t, reduced);
    const int nN = t.nN;
    int nn[3] = {ne[0] + 1, ne[1] + 1, dim == 3 ? ne[2] + 1 : 1};
    const int64_t N = (int64_t)nn[0] * nn[1] * nn[2];
    const bool addS = (A->mode || B->mode || C->mode || D->mode) && val;
    const bool addF = (X->mode || Y->mode) && rhs;
    const int outer = (dim == 3) ? ne[2] : ne[1];
    const int mid = (dim == 3) ? ne[1] : 1;
#pragma omp parallel
    {
        std::vector<double> EM_S(n * n * nN * nN), EM_F(n * nN);
        index_t nodes[8];
        for (int c0 = 0; c0 < 2; ++c0) {
#pragma omp for schedule(static)
            for (int ko = c0; ko < outer; ko += 2) {
                for (int km = 0; km < mid; ++km)
                    for (int k0 = 0; k0 < ne[0]; ++k0) {
                        const int k1 = (dim == 3) ? km : ko;
                        const int k2 = (dim == 3) ? ko : 0;
                        const int64_t e = k0 + (int64_t)ne[0] * (k1 + (int64_t)ne[1] * k2);
                        std::fill(EM_S.begin(), EM_S.end(), 0.0);
                        std::fill(EM_F.begin(), EM_F.end(), 0.0);
                        element_matrices(t, n, e, A, B, C, D, X, Y,
                                         addS ? EM_S.data() : nullptr,
                                         addF ? EM_F.data() : nullptr);
                        const int64_t first = k0 + (int64_t)nn[0] * (k1 + (int64_t)nn[1] * k2);
                        for (int a = 0; a < nN; ++a)
                            nodes[a] = (index_t)(first + (a & 1) + ((a >> 1) & 1) * nn[0] +
                                                 ((a >> 2) & 1) * (int64_t)nn[0] * nn[1]);
                        if (addF)
                            for (int a = 0; a < nN; ++a)
                                for (int k = 0; k < n; ++k)
                                    rhs[(int64_t)k + (int64_t)n * nodes[a]] += EM_F[k + n * a];
                        if (addS) orc_add_to_csr(n, N, ptr, index, val, nN, nodes, EM_S.data());
                    }
            }
        }
    }
}

// ------------------------------------------------------ boundary, Dirac

// Face-element matrices of assemblePDEBoundarySingle/System and their
// Reduced variants (DefaultAssembler2D.cpp:601-840, :1014-1200;
// DefaultAssembler3D.cpp:2233-2890, :3274-3630), and the element loop with
// addToMatrixAndRHS: faces in the order x0=0, x0=L, x1=0, x1=L[, x2=0,
// x2=L]; face elements of a face numbered by their two tangential element
// indices, lower axis fastest, offset by the element counts of the faces
// before (m_faceOffset); quadrature points on a face element 2 (2-D) or 2x2
// (3-D), lower tangential axis fastest, at the Gauss points of the element's
// edge/face, weight = face measure / points; the one-point rule at the face
// centre for reduced.  Per face element and point:
//   EM_S[k,m,a,b] += w d[k + n m] N_a N_b,   EM_F[k,a] += w y[k] N_a
// for the element nodes a, b on the face (the element's other nodes get 0),
// then added with addToMatrixAndRHS at the element's first node
// (e.g. DefaultAssembler3D.cpp:2360-2364).
ORC_API void orc_assemble_boundary(int dim, const int* ne, const double* len, int n, int reduced,
                                   const orc_coef* d, const orc_coef* y, const index_t* ptr, const index_t* index,
                                   double* val, double* rhs)
{
    double h[3] = {1.0, 1.0, 1.0};
    for (int a = 0; a < dim; ++a) h[a] = len[a] / ne[a];
    const int nN = 1 << dim;
    const int nn[3] = {ne[0] + 1, ne[1] + 1, dim == 3 ? ne[2] + 1 : 1};
    const int64_t N = (int64_t)nn[0] * nn[1] * nn[2];
    const bool addS = d->mode && val;
    const bool addF = y->mode && rhs;
    const int nq = reduced ? 1 : (1 << (dim - 1));
    std::vector<double> EM_S(n * n * nN * nN), EM_F(n * nN);
    int64_t offset = 0;
    for (int f = 0; f < 2 * dim; ++f) {
        const int ax = f / 2, side = f % 2;
        int t[2] = {-1, -1}, nt = 0;
        for (int a = 0; a < dim; ++a)
            if (a != ax) t[nt++] = a;
        const int ne0 = ne[t[0]], ne1 = (dim == 3) ? ne[t[1]] : 1;
        const double meas = h[t[0]] * (dim == 3 ? h[t[1]] : 1.0);
        const double w = meas / nq;
        for (int j1 = 0; j1 < ne1; ++j1)
            for (int j0 = 0; j0 < ne0; ++j0) {
                const int64_t e = offset + j0 + (int64_t)ne0 * j1;
                std::fill(EM_S.begin(), EM_S.end(), 0.0);
                std::fill(EM_F.begin(), EM_F.end(), 0.0);
                for (int q = 0; q < nq; ++q) {
                    const double xi0 = reduced ? 0.5 : gauss_pt(q & 1);
                    const double xi1 = reduced ? 0.5 : gauss_pt((q >> 1) & 1);
                    const double* dp = addS ? coef_point(*d, e, q, n * n) : nullptr;
                    const double* yp = addF ? coef_point(*y, e, q, n) : nullptr;
                    for (int a = 0; a < nN; ++a) {
                        const int ab[3] = {a & 1, (a >> 1) & 1, (a >> 2) & 1};
                        if (ab[ax] != side) continue;
                        const double Na = (ab[t[0]] ? xi0 : 1.0 - xi0) *
                                          (dim == 3 ? (ab[t[1]] ? xi1 : 1.0 - xi1) : 1.0);
                        if (addF)
                            for (int k = 0; k < n; ++k) EM_F[k + n * a] += w * yp[k] * Na;
                        if (addS)
                            for (int b = 0; b < nN; ++b) {
                                const int bb[3] = {b & 1, (b >> 1) & 1, (b >> 2) & 1};
                                if (bb[ax] != side) continue;
                                const double Nb = (bb[t[0]] ? xi0 : 1.0 - xi0) *
                                                  (dim == 3 ? (bb[t[1]] ? xi1 : 1.0 - xi1) : 1.0);
                                for (int m = 0; m < n; ++m)
                                    for (int k = 0; k < n; ++k)
                                        EM_S[k + n * (m + n * (a + nN * b))] += w * dp[k + n * m] * Na * Nb;
                            }
                    }
                }
                // the volume element carrying the face element
                int kk[3] = {0, 0, 0};
                kk[ax] = side ? ne[ax] - 1 : 0;
                kk[t[0]] = j0;
                if (dim == 3) kk[t[1]] = j1;
                const int64_t first = kk[0] + (int64_t)nn[0] * (kk[1] + (int64_t)nn[1] * kk[2]);
                index_t nodes[8];
                for (int a = 0; a < nN; ++a)
                    nodes[a] = (index_t)(first + (a & 1) + ((a >> 1) & 1) * nn[0] +
                                         ((a >> 2) & 1) * (int64_t)nn[0] * nn[1]);
                if (addF)
                    for (int a = 0; a < nN; ++a)
                        for (int k = 0; k < n; ++k) rhs[(int64_t)k + (int64_t)n * nodes[a]] += EM_F[k + n * a];
                if (addS) orc_add_to_csr(n, N, ptr, index, val, nN, nodes, EM_S.data());
            }
        offset += (int64_t)ne0 * ne1;
    }
}

// RipleyDomain::assemblePDEDirac (RipleyDomain.cpp:1756-1800): for each
// Dirac point in order, rhs[eq + n*dof] += y_i[eq] and the n x n block d_i
// added to the diagonal block of row dof (addToSystemMatrix with one node).
ORC_API void orc_add_dirac(int n, int64_t nrows, const index_t* ptr, const index_t* index, double* val,
                           int64_t count, const int64_t* nodes, const double* d, const double* y, double* rhs)
{
    for (int64_t p = 0; p < count; ++p) {
        const index_t dof = (index_t)nodes[p];
        if (y && rhs)
            for (int k = 0; k < n; ++k) rhs[k + (int64_t)n * dof] += y[p * n + k];
        if (d && val) orc_add_to_csr(n, nrows, ptr, index, val, 1, &dof, d + p * n * n);
    }
}

// ------------------------------------------------------- matrix operations

// SparseMatrix_MatrixVector_CSR_OFFSET0: out = alpha*A*in + beta*out with the
// rows split into static stripes, one per thread (SparseMatrix_MatrixVector.cpp:
// 191-336; bs=1 :262-270, bs=3 column-major block :288-320).
ORC_API void orc_spmv(int bs, int64_t nrows, const index_t* ptr, const index_t* index, const double* val,
              double alpha, const double* in, double beta, double* out)
{
#pragma omp parallel for schedule(static)
    for (int64_t ir = 0; ir < nrows; ++ir) {
        for (int r = 0; r < bs; ++r) {
            const int64_t irow = r + (int64_t)bs * ir;
            if (beta != 1.0) out[irow] = (beta == 0.0) ? 0.0 : beta * out[irow];
        }
        if (alpha == 0.0) continue;
        if (bs == 1) {
            double reg = 0.0;
            for (index_t k = ptr[ir]; k < ptr[ir + 1]; ++k) reg += val[k] * in[index[k]];
            out[ir] += alpha * reg;
        } else {
            double reg[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (index_t k = ptr[ir]; k < ptr[ir + 1]; ++k) {
                const int64_t ic0 = (int64_t)bs * index[k];
                for (int c = 0; c < bs; ++c)
                    for (int r = 0; r < bs; ++r)
                        reg[r] += val[(int64_t)k * bs * bs + r + bs * c] * in[ic0 + c];
            }
            for (int r = 0; r < bs; ++r) out[r + (int64_t)bs * ir] += alpha * reg[r];
        }
    }
}

// nullifyRowsAndCols_CSR_BLK1 / _CSR: a_ij = (i==j ? mdv : 0) wherever
// mask_row[i] > 0 or mask_col[j] > 0 (SparseMatrix.h:494-535; driver
// SystemMatrix.h:544-605).  Masks have nrows*bs entries.
ORC_API void orc_nullify(int bs, int64_t nrows, const index_t* ptr, const index_t* index, double* val,
                 const double* mask_row, const double* mask_col, double mdv)
{
#pragma omp parallel for schedule(static)
    for (int64_t ir = 0; ir < nrows; ++ir)
        for (index_t k = ptr[ir]; k < ptr[ir + 1]; ++k)
            for (int r = 0; r < bs; ++r)
                for (int c = 0; c < bs; ++c) {
                    const int64_t irow = r + (int64_t)bs * ir;
                    const int64_t icol = c + (int64_t)bs * index[k];
                    if (mask_col[icol] > 0. || mask_row[irow] > 0.)
                        val[(int64_t)k * bs * bs + r + bs * c] = (irow == icol ? mdv : 0.);
                }
}

// SystemMatrix::balance: d_i = sqrt(1/max_j|a_ij|) (1 for an empty row), then
// a_ij <- d_i a_ij d_j (SystemMatrix.cpp:262-312, SparseMatrix.cpp:266-283
// maxAbsRow, :361-384 applyDiagonal).  d has nrows*bs entries.
ORC_API void orc_balance(int bs, int64_t nrows, const index_t* ptr, const index_t* index, double* val, double* d)
{
#pragma omp parallel for schedule(static)
    for (int64_t ir = 0; ir < nrows; ++ir)
        for (int r = 0; r < bs; ++r) {
            double fac = 0.;
            for (index_t k = ptr[ir]; k < ptr[ir + 1]; ++k)
                for (int c = 0; c < bs; ++c)
                    fac = std::max(fac, std::fabs(val[(int64_t)k * bs * bs + r + bs * c]));
            d[r + (int64_t)bs * ir] = (fac > 0) ? std::sqrt(1. / fac) : 1.;
        }
#pragma omp parallel for schedule(static)
    for (int64_t ir = 0; ir < nrows; ++ir)
        for (int r = 0; r < bs; ++r) {
            const double rt = d[r + (int64_t)bs * ir];
            for (index_t k = ptr[ir]; k < ptr[ir + 1]; ++k)
                for (int c = 0; c < bs; ++c)
                    val[(int64_t)k * bs * bs + r + bs * c] *= rt * d[c + (int64_t)bs * index[k]];
        }
}

// SparseMatrix::invMain for bs 1, 2, 3: inverse of the main diagonal block by
// the explicit formulas of BlockOps_invM_2/3 (SparseMatrix.cpp:387-435,
// BlockOps.h:100-145).  Returns 1 if a block is singular ("non-regular main
// diagonal block"), 2 for an unsupported block size, else 0.
ORC_API int orc_inv_main(int bs, int64_t nrows, const index_t* ptr, const index_t* index, const double* val,
                 double* inv)
{
    int failed = 0;
    if (bs > 3) return 2;
#pragma omp parallel for schedule(static) reduction(| : failed)
    for (int64_t i = 0; i < nrows; ++i) {
        index_t kd = -1;
        for (index_t k = ptr[i]; k < ptr[i + 1]; ++k)
            if (index[k] == i) { kd = k; break; }
        if (kd < 0) { failed |= 1; continue; }
        const double* A = val + (int64_t)kd * bs * bs;
        double* I = inv + (int64_t)i * bs * bs;
        if (bs == 1) {
            if (std::fabs(A[0]) > 0.) I[0] = 1. / A[0]; else failed |= 1;
        } else if (bs == 2) {
            const double A11 = A[0], A12 = A[2], A21 = A[1], A22 = A[3];
            double D = A11 * A22 - A12 * A21;
            if (std::fabs(D) > 0) {
                D = 1. / D;
                I[0] = A22 * D; I[1] = -A21 * D; I[2] = -A12 * D; I[3] = A11 * D;
            } else failed |= 1;
        } else {
            const double A11 = A[0], A21 = A[1], A31 = A[2], A12 = A[3], A22 = A[4], A32 = A[5],
                         A13 = A[6], A23 = A[7], A33 = A[8];
            double D = A11 * (A22 * A33 - A23 * A32) + A12 * (A31 * A23 - A21 * A33) +
                       A13 * (A21 * A32 - A31 * A22);
            if (std::fabs(D) > 0) {
                D = 1. / D;
                I[0] = (A22 * A33 - A23 * A32) * D;
                I[1] = (A31 * A23 - A21 * A33) * D;
                I[2] = (A21 * A32 - A31 * A22) * D;
                I[3] = (A13 * A32 - A12 * A33) * D;
                I[4] = (A11 * A33 - A31 * A13) * D;
                I[5] = (A12 * A31 - A11 * A32) * D;
                I[6] = (A12 * A23 - A13 * A22) * D;
                I[7] = (A13 * A21 - A11 * A23) * D;
                I[8] = (A11 * A22 - A12 * A21) * D;
            } else failed |= 1;
        }
    }
    return failed;
}

// BlockOps_solveAll for the Jacobi smoother: x_i <- inv_i x_i
// (BlockOps.h:219-245).
static void block_solve_all(int bs, int64_t nrows, const double* inv, double* x)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nrows; ++i) {
        if (bs == 1) {
            x[i] *= inv[i];
        } else {
            double v[3], o[3];
            for (int r = 0; r < bs; ++r) v[r] = x[(int64_t)bs * i + r];
            for (int r = 0; r < bs; ++r) {
                double s = 0;
                for (int c = 0; c < bs; ++c) s += inv[(int64_t)i * bs * bs + r + bs * c] * v[c];
                o[r] = s;
            }
            for (int r = 0; r < bs; ++r) x[(int64_t)bs * i + r] = o[r];
        }
    }
}

// Preconditioner_Smoother_solve with Jacobi (x not initial): x = D^-1 b, then
// (sweeps-1) times x += D^-1 (b - A x) (Smoother.cpp:100-127, :201-216).
static void jacobi_apply(int bs, int64_t nrows, const index_t* ptr, const index_t* index,
                         const double* val, const double* inv, int sweeps, double* x,
                         const double* b, double* buf)
{
    const int64_t n = nrows * bs;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] = b[i];
    block_solve_all(bs, nrows, inv, x);
    for (int s = 1; s < sweeps; ++s) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) buf[i] = b[i];
        orc_spmv(bs, nrows, ptr, index, val, -1., x, 1., buf);
        block_solve_all(bs, nrows, inv, buf);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) x[i] += buf[i];
    }
}

// ------------------------------------------- colouring, Gauss-Seidel, ILU0

// Pattern::mis (paso/src/Pattern_mis.cpp:44-110): Luby-style maximal
// independent set with fixed priorities value_i = fmod(0.4142135623730951 *
// (i+1), 1) for available vertices (2 for the others).  Each round puts the
// available vertices whose value is below that of every neighbour into the
// set and marks their neighbours connected.  On entry mis_marker is -1 for
// available vertices; on exit it is 1 for the set, 0 otherwise.
static void pattern_mis(int64_t n, const index_t* ptr, const index_t* index, index_t* mis_marker)
{
    const index_t AVAILABLE = -1, IN_MIS_NOW = -2, IN_MIS = -3, CONNECTED = -4;
    const double seed = .4142135623730951;
    std::vector<double> value(n);
    for (;;) {
        bool any = false;
        for (int64_t i = 0; i < n; ++i) any = any || mis_marker[i] == AVAILABLE;
        if (!any) break;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i)
            value[i] = (mis_marker[i] == AVAILABLE) ? std::fmod(seed * (double)(i + 1), 1.) : 2.;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            if (mis_marker[i] != AVAILABLE) continue;
            index_t flag = IN_MIS_NOW;
            for (index_t k = ptr[i]; k < ptr[i + 1]; ++k) {
                const index_t j = index[k];
                if (j != i && value[j] <= value[i]) { flag = AVAILABLE; break; }
            }
            mis_marker[i] = flag;
        }
        // the set is independent, so the writes below never race with a
        // vertex of the set (as in the reference's parallel loop)
        for (int64_t i = 0; i < n; ++i) {
            if (mis_marker[i] != IN_MIS_NOW) continue;
            for (index_t k = ptr[i]; k < ptr[i + 1]; ++k)
                if (index[k] != i) mis_marker[index[k]] = CONNECTED;
            mis_marker[i] = IN_MIS;
        }
    }
    for (int64_t i = 0; i < n; ++i) mis_marker[i] = (mis_marker[i] == IN_MIS);
}

// Pattern::borrowColoringPointer (paso/src/Pattern.cpp:166-195): colour c is
// the MIS (above) of the vertices left uncoloured by colours 0..c-1.
// Returns the number of colours.
ORC_API int orc_coloring(int64_t n, const index_t* ptr, const index_t* index, index_t* coloring)
{
    std::vector<index_t> marker(n, -1);
    for (int64_t i = 0; i < n; ++i) coloring[i] = -1;
    int out = 0;
    for (;;) {
        bool any = false;
        for (int64_t i = 0; i < n; ++i) any = any || coloring[i] == -1;
        if (!any) break;
        pattern_mis(n, ptr, index, marker.data());
        for (int64_t i = 0; i < n; ++i) {
            if (marker[i]) coloring[i] = out;
            marker[i] = coloring[i];
        }
        ++out;
    }
    return out;
}

// position of the main diagonal entry of row i (Pattern::borrowMainDiagonalPointer)
static inline index_t main_ptr(const index_t* ptr, const index_t* index, int64_t i)
{
    for (index_t k = ptr[i]; k < ptr[i + 1]; ++k)
        if (index[k] == i) return k;
    return -1;
}

// R -= M V for a column-major bs x bs block (BlockOps_SMV_2/3, BlockOps.h:46-85)
static inline void block_smv(int bs, double* R, const double* M, const double* V)
{
    double t[3];
    for (int r = 0; r < bs; ++r) {
        double s = 0.0;
        for (int c = 0; c < bs; ++c) s += M[r + bs * c] * V[c];
        t[r] = s;
    }
    for (int r = 0; r < bs; ++r) R[r] -= t[r];
}

// V <- M V in place (BlockOps_MViP_2/3, BlockOps.h:187-217)
static inline void block_mvip(int bs, const double* M, double* V)
{
    double t[3];
    for (int r = 0; r < bs; ++r) {
        double s = 0.0;
        for (int c = 0; c < bs; ++c) s += M[r + bs * c] * V[c];
        t[r] = s;
    }
    for (int r = 0; r < bs; ++r) V[r] = t[r];
}

// Preconditioner_LocalSmoother_Sweep_colored (paso/src/Smoother.cpp:
// 331-507), the sweep PASO runs whenever OpenMP has two or more threads:
// forward over the colours, x_i <- inv(A_ii) (x_i - sum_{c(k)<c(i)} A_ik x_k);
// backward over colours nc-2..0, x_i <- inv(A_ii) (A_ii x_i -
// sum_{c(k)>c(i)} A_ik x_k).  x holds the right-hand side on entry.
ORC_API void orc_gs_sweep(int bs, int64_t nrows, const index_t* ptr, const index_t* index, const double* val,
                          const double* inv, const index_t* coloring, int ncolors, double* x)
{
    const int b2 = bs * bs;
    for (int color = 0; color < ncolors; ++color) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nrows; ++i) {
            if (coloring[i] != color) continue;
            double* xi = x + (int64_t)bs * i;
            if (color > 0)
                for (index_t k = ptr[i]; k < ptr[i + 1]; ++k)
                    if (coloring[index[k]] < color) block_smv(bs, xi, val + (int64_t)k * b2, x + (int64_t)bs * index[k]);
            block_mvip(bs, inv + (int64_t)i * b2, xi);
        }
    }
    for (int color = ncolors - 2; color > -1; --color) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nrows; ++i) {
            if (coloring[i] != color) continue;
            double* xi = x + (int64_t)bs * i;
            block_mvip(bs, val + (int64_t)main_ptr(ptr, index, i) * b2, xi);
            for (index_t k = ptr[i]; k < ptr[i + 1]; ++k)
                if (coloring[index[k]] > color) block_smv(bs, xi, val + (int64_t)k * b2, x + (int64_t)bs * index[k]);
            block_mvip(bs, inv + (int64_t)i * b2, xi);
        }
    }
}

// inverse of a column-major bs x bs block by the closed forms of
// BlockOps_invM_2/3 (BlockOps.h:100-145, also used by Solver_getILU);
// returns false for a zero determinant
static bool block_inverse(int bs, const double* A, double* I)
{
    if (bs == 1) {
        if (!(std::fabs(A[0]) > 0.)) return false;
        I[0] = 1. / A[0];
        return true;
    }
    if (bs == 2) {
        const double A11 = A[0], A12 = A[2], A21 = A[1], A22 = A[3];
        double D = A11 * A22 - A12 * A21;
        if (!(std::fabs(D) > 0.)) return false;
        D = 1. / D;
        I[0] = A22 * D; I[1] = -A21 * D; I[2] = -A12 * D; I[3] = A11 * D;
        return true;
    }
    const double A11 = A[0], A21 = A[1], A31 = A[2], A12 = A[3], A22 = A[4], A32 = A[5], A13 = A[6],
                 A23 = A[7], A33 = A[8];
    double D = A11 * (A22 * A33 - A23 * A32) + A12 * (A31 * A23 - A21 * A33) + A13 * (A21 * A32 - A31 * A22);
    if (!(std::fabs(D) > 0.)) return false;
    D = 1. / D;
    I[0] = (A22 * A33 - A23 * A32) * D;
    I[1] = (A31 * A23 - A21 * A33) * D;
    I[2] = (A21 * A32 - A31 * A22) * D;
    I[3] = (A13 * A32 - A12 * A33) * D;
    I[4] = (A11 * A33 - A31 * A13) * D;
    I[5] = (A12 * A31 - A11 * A32) * D;
    I[6] = (A12 * A23 - A13 * A22) * D;
    I[7] = (A13 * A21 - A11 * A23) * D;
    I[8] = (A11 * A22 - A12 * A21) * D;
    return true;
}

// Solver_getILU (paso/src/ILU.cpp:43-300): incomplete block LU in colour
// order.  Row i of colour c: for colours c2 < c, for neighbours k of colour
// c2 (row order), a_ij -= a_ik a_kj for every neighbour j of k with colour >
// c2 that is in row i; then a_ii <- inv(a_ii) and a_ik <- inv(a_ii) a_ik for
// neighbours of higher colour.  Returns 1 for a singular pivot block, 2 for a
// block size above 3.
ORC_API int orc_ilu0_factor(int bs, int64_t nrows, const index_t* ptr, const index_t* index, const double* val,
                            const index_t* coloring, int ncolors, double* F)
{
    if (bs > 3) return 2;
    const int b2 = bs * bs;
    const int64_t len = (int64_t)ptr[nrows] * b2;
    for (int64_t e = 0; e < len; ++e) F[e] = val[e];
    int failed = 0;
    for (int color = 0; color < ncolors; ++color) {
#pragma omp parallel for schedule(static) reduction(| : failed)
        for (int64_t i = 0; i < nrows; ++i) {
            if (coloring[i] != color) continue;
            for (int color2 = 0; color2 < color; ++color2)
                for (index_t ik = ptr[i]; ik < ptr[i + 1]; ++ik) {
                    const index_t k = index[ik];
                    if (coloring[k] != color2) continue;
                    double Aik[9];
                    for (int e = 0; e < b2; ++e) Aik[e] = F[(int64_t)ik * b2 + e];
                    for (index_t kj = ptr[k]; kj < ptr[k + 1]; ++kj) {
                        const index_t j = index[kj];
                        if (coloring[j] <= color2) continue;
                        for (index_t ij = ptr[i]; ij < ptr[i + 1]; ++ij) {
                            if (index[ij] != j) continue;
                            const double* S = F + (int64_t)kj * b2;
                            double* T = F + (int64_t)ij * b2;
                            for (int c = 0; c < bs; ++c)
                                for (int r = 0; r < bs; ++r) {
                                    double s = 0.0;
                                    for (int m = 0; m < bs; ++m) s += Aik[r + bs * m] * S[m + bs * c];
                                    T[r + bs * c] -= s;
                                }
                            break;
                        }
                    }
                }
            const index_t mm = main_ptr(ptr, index, i);
            double Di[9];
            if (mm < 0 || !block_inverse(bs, F + (int64_t)mm * b2, Di)) {
                failed |= 1;
                continue;
            }
            for (int e = 0; e < b2; ++e) F[(int64_t)mm * b2 + e] = Di[e];
            for (index_t ik = ptr[i]; ik < ptr[i + 1]; ++ik) {
                if (coloring[index[ik]] <= color) continue;
                double* T = F + (int64_t)ik * b2;
                double t[9];
                for (int c = 0; c < bs; ++c)
                    for (int r = 0; r < bs; ++r) {
                        double s = 0.0;
                        for (int m = 0; m < bs; ++m) s += Di[r + bs * m] * T[m + bs * c];
                        t[r + bs * c] = s;
                    }
                for (int e = 0; e < b2; ++e) T[e] = t[e];
            }
        }
    }
    return failed;
}

// Solver_solveILU (paso/src/ILU.cpp:316-462): x = b; forward over colours
// x_i <- inv(a_ii) (x_i - sum_{c(k)<c(i)} a_ik x_k); backward over colours
// nc-1..0, x_i <- x_i - sum_{c(k)>c(i)} a_ik x_k.
ORC_API void orc_ilu0_solve(int bs, int64_t nrows, const index_t* ptr, const index_t* index, const double* F,
                            const index_t* coloring, int ncolors, double* x, const double* b)
{
    const int b2 = bs * bs;
    for (int64_t e = 0; e < nrows * bs; ++e) x[e] = b[e];
    for (int color = 0; color < ncolors; ++color) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nrows; ++i) {
            if (coloring[i] != color) continue;
            double* xi = x + (int64_t)bs * i;
            for (index_t k = ptr[i]; k < ptr[i + 1]; ++k)
                if (coloring[index[k]] < color) block_smv(bs, xi, F + (int64_t)k * b2, x + (int64_t)bs * index[k]);
            block_mvip(bs, F + (int64_t)main_ptr(ptr, index, i) * b2, xi);
        }
    }
    for (int color = ncolors - 1; color > -1; --color) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < nrows; ++i) {
            if (coloring[i] != color) continue;
            double* xi = x + (int64_t)bs * i;
            for (index_t k = ptr[i]; k < ptr[i + 1]; ++k)
                if (coloring[index[k]] > color) block_smv(bs, xi, F + (int64_t)k * b2, x + (int64_t)bs * index[k]);
        }
    }
}

// Preconditioner_Smoother_solve with the Gauss-Seidel sweep (x not
// initial): x = S^-1 b, then (sweeps-1) times x += S^-1 (b - A x)
// (Smoother.cpp:100-127, :201-216).
static void gs_apply(int bs, int64_t nrows, const index_t* ptr, const index_t* index, const double* val,
                     const double* inv, const index_t* coloring, int ncolors, int sweeps, double* x,
                     const double* b, double* buf)
{
    const int64_t n = nrows * bs;
    for (int64_t i = 0; i < n; ++i) x[i] = b[i];
    orc_gs_sweep(bs, nrows, ptr, index, val, inv, coloring, ncolors, x);
    for (int s = 1; s < sweeps; ++s) {
        for (int64_t i = 0; i < n; ++i) buf[i] = b[i];
        orc_spmv(bs, nrows, ptr, index, val, -1., x, 1., buf);
        orc_gs_sweep(bs, nrows, ptr, index, val, inv, coloring, ncolors, buf);
        for (int64_t i = 0; i < n; ++i) x[i] += buf[i];
    }
}

// Preconditioner_Smoother_solve with the coloured Gauss-Seidel sweep, exported
// for the parity tests of a single preconditioner application.
ORC_API void orc_gs_apply(int bs, int64_t nrows, const index_t* ptr, const index_t* index, const double* val,
                          const double* inv, const index_t* coloring, int ncolors, int sweeps, double* x,
                          const double* b)
{
    std::vector<double> buf((size_t)nrows * bs);
    gs_apply(bs, nrows, ptr, index, val, inv, coloring, ncolors, sweeps, x, b, buf.data());
}

// ---------------------------------------------------------- reductions
// PASO accumulates per-thread partial sums under "omp critical", so its sums
// depend on thread timing.  Here the partials are combined in thread order,
// which is deterministic for a fixed thread count.
template <class F>
static double par_sum(int64_t n, F f)
{
    const int nt = omp_get_max_threads();
    std::vector<double> part(nt, 0.0);
#pragma omp parallel
    {
        const int t = omp_get_thread_num();
        double s = 0.0;
#pragma omp for schedule(static)
        for (int64_t i = 0; i < n; ++i) s += f(i);
        part[t] = s;
    }
    double s = 0.0;
    for (double p : part) s += p;
    return s;
}

static double par_max_abs(int64_t n, const double* v)
{
    double m = 0.0;
#pragma omp parallel for schedule(static) reduction(max : m)
    for (int64_t i = 0; i < n; ++i) m = std::max(m, std::fabs(v[i]));
    return m;
}

// ---------------------------------------------- test hooks: trace, plant
// orc_trace_begin(buf, cap, n) makes the next Krylov solve (the first inner
// call of the next orc_solve) record, after each of its iterations, four
// vectors of n doubles and four scalars into buf (stride 4n + 4, at most cap
// iterations); recording stops when that solve returns.  The tests check the
// recorded iterates against invariants of the recurrences and against
// library implementations of the same methods.  orc_set_plant(p) switches on
// a known mistake in one recurrence so that those tests can show they catch
// it (0: none; 1: PCG without residual smoothing, gamma = 1; 2: PCG with the
// smoothing coefficient's sign flipped; 3: BiCGStab's beta without the
// alpha/omega factor; 4: BiCGStab's omega with its sign flipped; 5: MINRES's
// w recurrence without the a3 term).
static double* g_trace = nullptr;
static int g_trace_cap = 0, g_trace_count = 0, g_trace_on = 0;
static int64_t g_trace_n = 0;
static int g_plant = 0;

ORC_API void orc_trace_begin(double* buf, int cap, int64_t n)
{
    g_trace = buf;
    g_trace_cap = cap;
    g_trace_n = n;
    g_trace_count = 0;
    g_trace_on = buf != nullptr && cap > 0;
}

ORC_API int orc_trace_count(void) { return g_trace_count; }

ORC_API void orc_set_plant(int p) { g_plant = p; }

static void trace_record(const double* v0, const double* v1, const double* v2, const double* v3, double s0,
                         double s1, double s2, double s3)
{
    if (!g_trace_on || g_trace_count >= g_trace_cap) return;
    const int64_t n = g_trace_n;
    double* rec = g_trace + (size_t)g_trace_count * (4 * n + 4);
    const double* v[4] = {v0, v1, v2, v3};
    for (int k = 0; k < 4; ++k)
        for (int64_t i = 0; i < n; ++i) rec[k * n + i] = v[k] ? v[k][i] : 0.0;
    rec[4 * n] = s0;
    rec[4 * n + 1] = s1;
    rec[4 * n + 2] = s2;
    rec[4 * n + 3] = s3;
    ++g_trace_count;
}

static void trace_end() { g_trace_on = 0; }

// ------------------------------------------------------------- solvers

typedef struct {
    double tolerance;      // SolverBuddy default 1e-8 (SolverOptions.cpp:40-72)
    int iter_max;          // default 100000
    int sweeps;            // Jacobi sweeps, SolverBuddy default 1
    int plain;             // 0: PASO residual-smoothed PCG; 1: textbook PCG; 2: BiCGStab; 3: MINRES
    int verbose;
    int preconditioner;    // 0: Jacobi, 1: Gauss-Seidel, 2: ILU0, 3: none (Preconditioner.cpp:46-107)
} orc_options;

typedef struct {
    int num_iter;          // Solver.cpp:324-325
    int converged;
    int status;            // SolverResult (Paso.h:44-52)
    int outer_steps;
    double residual_norm;  // last ||r||_2 in the balanced space
    double set_up_time, net_time, time;
} orc_diagnostics;

enum { NoError = 0, MaxIterReached, InputError, MemoryError, Breakdown, NegativeNormError, Divergence };

struct Sys {
    int bs;
    int64_t nrows, n;
    const index_t *ptr, *index;
    const double* val;
    const double* inv;
    int sweeps;
    std::vector<double> buf;
    int precond = 0;
    const index_t* coloring = nullptr;
    int ncolors = 0;
    const double* F = nullptr;
    // Preconditioner_solve (paso/src/Preconditioner.cpp:111-137)
    void prec(double* x, const double* b)
    {
        if (precond == 1) gs_apply(bs, nrows, ptr, index, val, inv, coloring, ncolors, sweeps, x, b, buf.data());
        else if (precond == 2) orc_ilu0_solve(bs, nrows, ptr, index, F, coloring, ncolors, x, b);
        else if (precond == 3) std::copy(b, b + n, x);
        else jacobi_apply(bs, nrows, ptr, index, val, inv, sweeps, x, b, buf.data());
    }
    void mv(double alpha, const double* in, double beta, double* out) { orc_spmv(bs, nrows, ptr, index, val, alpha, in, beta, out); }
};

// Solver_PCG: residual-smoothed PCG (PCG.cpp:62-368).  On entry r = b - A x
// for the current x; x is updated in place.  *iter is the limit on entry and
// the number of iterations on exit; *tolerance likewise becomes ||rs||_2.
static int solver_pcg_paso(Sys& S, double* r, double* x, int* iter, double* tolerance)
{
    const int64_t n = S.n;
    std::vector<double> rs(r, r + n), p(n, 0.0), v(n, 0.0), x2(x, x + n);
    const int maxit = *iter;
    const double tol = *tolerance;
    int num_iter = 0;
    double tau = 0, tau_old, norm_of_residual = 0;
    bool convergeFlag = false, maxIterFlag = false, breakFlag = false;
    while (!(convergeFlag || maxIterFlag || breakFlag)) {
        ++num_iter;
        S.prec(v.data(), r);                                                     // :152
        const double sum_1 = par_sum(n, [&](int64_t i) { return v[i] * r[i]; }); // :156-191
        tau_old = tau;
        tau = sum_1;
        if (num_iter == 1) {                                                     // :193-219
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) p[i] = v[i];
        } else {
            const double beta = tau / tau_old;
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) p[i] = v[i] + beta * p[i];
        }
        S.mv(1.0, p.data(), 0.0, v.data());                                      // :223
        const double delta = par_sum(n, [&](int64_t i) { return v[i] * p[i]; }); // :228-260
        const double alpha = tau / delta;
        breakFlag = (std::fabs(delta) <= 0.0);
        if (!breakFlag) {
            double sum_3, sum_4;                                                 // :262-303
            {
                const int nt = omp_get_max_threads();
                std::vector<double> p3(nt, 0.0), p4(nt, 0.0);
#pragma omp parallel
                {
                    double ss = 0, ss1 = 0;
#pragma omp for schedule(static)
                    for (int64_t i = 0; i < n; ++i) {
                        r[i] -= alpha * v[i];
                        const double d = r[i] - rs[i];
                        ss += d * d;
                        ss1 += d * rs[i];
                    }
                    p3[omp_get_thread_num()] = ss;
                    p4[omp_get_thread_num()] = ss1;
                }
                sum_3 = 0; sum_4 = 0;
                for (int t = 0; t < nt; ++t) { sum_3 += p3[t]; sum_4 += p4[t]; }
            }
            double gamma_1 = (std::fabs(sum_3) <= 0.0) ? 0 : -sum_4 / sum_3;       // :306-343
            if (g_plant == 1) gamma_1 = 1.0;
            if (g_plant == 2) gamma_1 = -gamma_1;
            const double gamma_2 = 1.0 - gamma_1;
            const double sum_5 = par_sum(n, [&](int64_t i) {
                rs[i] = gamma_2 * rs[i] + gamma_1 * r[i];
                x2[i] += alpha * p[i];
                x[i] = gamma_2 * x[i] + gamma_1 * x2[i];
                return rs[i] * rs[i];
            });
            norm_of_residual = std::sqrt(sum_5);                                 // :344-347
            convergeFlag = norm_of_residual <= tol;
            maxIterFlag = num_iter > maxit;
            breakFlag = (std::fabs(tau) <= 0.0);
            trace_record(x, rs.data(), x2.data(), r, alpha, tau, gamma_1, norm_of_residual);
        }
    }
    trace_end();
    *iter = num_iter;
    *tolerance = norm_of_residual;
    if (maxIterFlag) return MaxIterReached;
    if (breakFlag) return Breakdown;
    return NoError;
}

// Solver_BiCGStab (paso/src/BiCGStab.cpp:77-257): right-preconditioned
// BiCGStab with the Jacobi smoother; on entry r = b - A x; early exit when
// ||s||_2 <= tol (x += alpha phat), else the stabilising step; stops when
// ||r||_2 <= tol, num_iter > maxit, or a zero rho, rtld.v, t.t or omega.
static int solver_bicgstab(Sys& S, double* r, double* x, int* iter, double* tolerance)
{
    const int64_t n = S.n;
    std::vector<double> rtld(r, r + n), p(n, 0.0), v(n, 0.0), t(n, 0.0), phat(n, 0.0), shat(n, 0.0), s(n, 0.0);
    const int maxit = *iter;
    const double tol = *tolerance;
    int num_iter = 0;
    double alpha = 0, omega = 0, rho = 0, rho1 = 0, norm_of_residual = 0;
    bool breakFlag = false, maxIterFlag = false, convergeFlag = false;
    while (true) {
        ++num_iter;                                                                  // :125
        rho = par_sum(n, [&](int64_t i) { return rtld[i] * r[i]; });                 // :133
        if ((breakFlag = (std::fabs(rho) <= 0.0))) break;                            // :139
        if (num_iter > 1) {                                                          // :142-150
            const double beta = g_plant == 3 ? rho / rho1 : rho / rho1 * (alpha / omega);
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        } else {
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) p[i] = r[i];
        }
        S.prec(phat.data(), p.data());                                               // :154
        S.mv(1.0, phat.data(), 0.0, v.data());                                       // :155
        const double sum_2 = par_sum(n, [&](int64_t i) { return rtld[i] * v[i]; });  // :158
        if ((breakFlag = (std::fabs(sum_2) <= 0.0))) break;                          // :164
        alpha = rho / sum_2;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {                                            // :168-173
            r[i] -= alpha * v[i];
            s[i] = r[i];
        }
        norm_of_residual = std::sqrt(par_sum(n, [&](int64_t i) { return s[i] * s[i]; }));
        if ((convergeFlag = (norm_of_residual <= tol))) {                            // :181-186
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < n; ++i) x[i] += alpha * phat[i];
            maxIterFlag = false;
            breakFlag = false;
            trace_record(x, r, s.data(), nullptr, rho, alpha, 0.0, norm_of_residual);
            break;
        }
        S.prec(shat.data(), s.data());                                               // :189
        S.mv(1.0, shat.data(), 0.0, t.data());                                       // :190
        const double omegaNumtr = par_sum(n, [&](int64_t i) { return t[i] * s[i]; }); // :193-196
        const double omegaDenumtr = par_sum(n, [&](int64_t i) { return t[i] * t[i]; });
        if ((breakFlag = (std::fabs(omegaDenumtr) <= 0.0))) break;                   // :206
        omega = omegaNumtr / omegaDenumtr;
        if (g_plant == 4) omega = -omega;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) {                                            // :209-214
            x[i] += alpha * phat[i] + omega * shat[i];
            r[i] = s[i] - omega * t[i];
        }
        norm_of_residual = std::sqrt(par_sum(n, [&](int64_t i) { return r[i] * r[i]; }));
        convergeFlag = norm_of_residual <= tol;                                      // :221-223
        maxIterFlag = num_iter > maxit;
        breakFlag = std::fabs(omega) <= 0.0;
        trace_record(x, r, s.data(), t.data(), rho, alpha, omega, norm_of_residual);
        if (convergeFlag || maxIterFlag || breakFlag) break;
        rho1 = rho;                                                                  // :228
    }
    trace_end();
    *iter = num_iter;
    *tolerance = norm_of_residual;
    if (maxIterFlag) return MaxIterReached;
    if (breakFlag) return Breakdown;
    return NoError;
}

// Solver_MINRES (paso/src/MINRES.cpp:59-191): preconditioned MINRES by the
// Lanczos recurrence r_new = A z - (delta/gamma) r - (gamma/gamma_old) r_old
// with z = M^-1 r / gamma, Givens rotations of the tridiagonal matrix, and
// x += c eta w.  Stops when the preconditioned residual estimate
// rnorm_prec <= tol * ||r||_M / ||r||_2, returns NegativeNormError for
// r.M^-1 r < 0, converges on r.M^-1 r == 0 (the reference's test
// |dp| > 0.e-13 |dp0| is |dp| > 0).  On exit *tolerance holds
// rnorm_prec ||r||_2 / ||r||_M.
static int solver_minres(Sys& S, double* r, double* x, int* iter, double* tolerance)
{
    const int64_t n = S.n;
    const int maxit = *iter;
    if (maxit <= 0) {
        trace_end();
        return InputError;
    }
    std::vector<double> znew(n), z(n), az(n), w(n, 0.0), r_old(n, 0.0), w_old(n, 0.0), r_anc(n, 0.0),
        w_anc(n, 0.0);
    double delta = 0., gamma = 0., gamma_old = 0., eta = 0., dp0 = 0., c = 1.0, c_old = 1.0;
    double c_anc = 1., s = 0.0, s_old = 0.0, s_anc, rnorm_prec = 1;
    double tol = 1., norm_scal = 1.;
    int num_iter = 0;
    bool converge = false;
    int status = NoError;
    S.prec(z.data(), r);                                                       // :77-79
    double dp = par_sum(n, [&](int64_t i) { return r[i] * z[i]; });
    dp0 = dp;
    if (dp < 0) {
        status = NegativeNormError;
    } else if (std::fabs(dp) <= 0) {
        converge = true;
    } else {
        gamma = std::sqrt(dp);
        eta = gamma;
        rnorm_prec = gamma;
        const double norm_r = std::sqrt(par_sum(n, [&](int64_t i) { return r[i] * r[i]; }));
        norm_scal = rnorm_prec / norm_r;
        tol = (*tolerance) * norm_scal;
    }
    (void)dp0;
    while (!converge && status == NoError) {
        for (int64_t i = 0; i < n; ++i) z[i] *= 1. / gamma;                   // :103
        S.mv(1.0, z.data(), 0.0, az.data());                                   // :106
        delta = par_sum(n, [&](int64_t i) { return az[i] * z[i]; });
        if (num_iter > 0) std::copy(r_old.begin(), r_old.end(), r_anc.begin());
        std::copy(r, r + n, r_old.begin());
        for (int64_t i = 0; i < n; ++i) {
            double v = az[i] - delta / gamma * r_old[i];
            if (num_iter > 0) v -= gamma / gamma_old * r_anc[i];
            r[i] = v;
        }
        S.prec(znew.data(), r);                                                // :122
        dp = par_sum(n, [&](int64_t i) { return r[i] * znew[i]; });
        if (dp < 0.) {
            status = NegativeNormError;
        } else if (std::fabs(dp) == 0.) {
            converge = true;
        } else {
            gamma_old = gamma;                                                 // :131-145
            gamma = std::sqrt(dp);
            c_anc = c_old;
            c_old = c;
            s_anc = s_old;
            s_old = s;
            const double a0 = c_old * delta - c_anc * s_old * gamma_old;
            const double a1 = std::sqrt(a0 * a0 + gamma * gamma);
            const double a2 = s_old * delta + c_anc * c_old * gamma_old;
            const double a3 = s_anc * gamma_old;
            c = a0 / a1;
            s = gamma / a1;
            rnorm_prec = rnorm_prec * s;
            if (num_iter > 1) std::copy(w_old.begin(), w_old.end(), w_anc.begin());   // :149-160
            if (num_iter > 0) std::copy(w.begin(), w.end(), w_old.begin());
            for (int64_t i = 0; i < n; ++i) {
                double v = z[i];
                if (num_iter > 1 && g_plant != 5) v -= a3 * w_anc[i];
                if (num_iter > 0) v -= a2 * w_old[i];
                w[i] = v * (1.0 / a1);
            }
            for (int64_t i = 0; i < n; ++i) x[i] += c * eta * w[i];
            eta = -s * eta;
            converge = rnorm_prec <= tol;
        }
        std::copy(znew.begin(), znew.end(), z.begin());
        ++num_iter;
        if (!converge && num_iter >= maxit) status = MaxIterReached;
        trace_record(x, r, w.data(), z.data(), delta, gamma, c, rnorm_prec);
    }
    trace_end();
    *iter = num_iter;
    *tolerance = rnorm_prec / norm_scal;
    return status;
}

// Textbook Jacobi-PCG with the same calling convention as solver_pcg_paso:
// the "plain PCG" comparator.  Stops when ||r||_2 <= tol.
static int solver_pcg_plain(Sys& S, double* r, double* x, int* iter, double* tolerance)
{
    const int64_t n = S.n;
    std::vector<double> z(n), p(n), q(n);
    const int maxit = *iter;
    const double tol = *tolerance;
    S.prec(z.data(), r);
    double rho = par_sum(n, [&](int64_t i) { return r[i] * z[i]; });
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) p[i] = z[i];
    int num_iter = 0;
    double rnorm = std::sqrt(par_sum(n, [&](int64_t i) { return r[i] * r[i]; }));
    int status = NoError;
    while (true) {
        if (rnorm <= tol) break;
        if (num_iter >= maxit) { status = MaxIterReached; break; }
        if (std::fabs(rho) <= 0.0) { status = Breakdown; break; }
        ++num_iter;
        S.mv(1.0, p.data(), 0.0, q.data());
        const double delta = par_sum(n, [&](int64_t i) { return p[i] * q[i]; });
        if (std::fabs(delta) <= 0.0) { status = Breakdown; break; }
        const double alpha = rho / delta;
        rnorm = std::sqrt(par_sum(n, [&](int64_t i) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            return r[i] * r[i];
        }));
        S.prec(z.data(), r);
        const double rho_new = par_sum(n, [&](int64_t i) { return r[i] * z[i]; });
        const double beta = rho_new / rho;
        rho = rho_new;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    }
    *iter = num_iter;
    *tolerance = rnorm;
    return status;
}

// Solver(): balance, initial guess x = M^-1 D b, then the outer loop that
// recomputes r = D b - A x and stops when both ||r||_2 <= rtol ||Db||_2 and
// ||r||_inf <= rtol ||Db||_inf, restarting PCG with
// tol = rtol*min(||Db||_2, 0.1 ||r||_2/||r||_inf ||Db||_inf); breakdown with
// more than one iteration restarts; finally x <- D x
// (Solver.cpp:46-331, solve.cpp:41-145 for the error mapping).
//
// val is balanced in place (as PASO leaves its matrix balanced), and d
// receives the balance vector.  Returns the SolverResult code; code 100 means
// a bad tolerance, 101 a singular diagonal, 102 NaN in the right-hand side.
ORC_API int orc_solve(int bs, int64_t nrows, const index_t* ptr, const index_t* index, double* val,
              double* x, const double* b, const orc_options* opt, orc_diagnostics* diag, double* d)
{
    const double EPS = 2.220446049250313e-16;
    const double t0 = omp_get_wtime();
    diag->num_iter = 0; diag->converged = 0; diag->status = 0; diag->outer_steps = 0;
    diag->residual_norm = -1; diag->set_up_time = -1; diag->net_time = -1; diag->time = -1;
    const double tolerance = opt->tolerance;
    if (tolerance < 100. * EPS || tolerance > 1.) return 100;
    const int64_t n = nrows * bs;
    orc_balance(bs, nrows, ptr, index, val, d);
    std::vector<double> r(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] * d[i];
    const double norm2_of_b = std::sqrt(par_sum(n, [&](int64_t i) { return r[i] * r[i]; }));
    const double norm_max_of_b = par_max_abs(n, r.data());
    if (std::isnan(norm2_of_b) || std::isnan(norm_max_of_b)) return 102;
    int status = NoError;
    if (norm2_of_b <= 0.) {
        std::fill(x, x + n, 0.0);
    } else {
        // Preconditioner_alloc (Preconditioner.cpp:46-107): Jacobi and
        // Gauss-Seidel invert the main diagonal blocks, ILU0 factorises
        const int pc = opt->preconditioner;
        std::vector<double> inv;
        std::vector<index_t> coloring;
        std::vector<double> F;
        int ncolors = 0;
        if (pc == 0 || pc == 1) {
            inv.resize((size_t)nrows * bs * bs);
            if (orc_inv_main(bs, nrows, ptr, index, val, inv.data())) return 101;
        }
        if (pc == 1 || pc == 2) {
            coloring.resize(nrows);
            ncolors = orc_coloring(nrows, ptr, index, coloring.data());
        }
        if (pc == 2) {
            F.resize((size_t)ptr[nrows] * bs * bs);
            if (orc_ilu0_factor(bs, nrows, ptr, index, val, coloring.data(), ncolors, F.data())) return 101;
        }
        Sys S{bs, nrows, n, ptr, index, val, inv.data(), std::max(1, opt->sweeps),
              std::vector<double>(n)};
        S.precond = pc;
        S.coloring = coloring.data();
        S.ncolors = ncolors;
        S.F = F.data();
        diag->set_up_time = omp_get_wtime() - t0;
        S.prec(x, r.data());
        int totIter = 1;
        bool finalize = false;
        double last2 = norm2_of_b, lastmax = norm_max_of_b;
        const double tn = omp_get_wtime();
        while (!finalize) {
            int cntIter = opt->iter_max - totIter;
            finalize = true;
            if (totIter > 1)
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < n; ++i) r[i] = b[i] * d[i];
            S.mv(-1.0, x, 1.0, r.data());
            const double norm2 = std::sqrt(par_sum(n, [&](int64_t i) { return r[i] * r[i]; }));
            const double normmax = par_max_abs(n, r.data());
            diag->residual_norm = norm2;
            diag->outer_steps++;
            if (opt->verbose)
                printf("Solver: Step %d: l2/lmax-norm of residual is %g/%g\n", totIter, norm2, normmax);
            if (totIter > 1 && norm2 >= last2 && normmax >= lastmax) {
                status = Divergence;
                break;
            }
            if (norm2 > tolerance * norm2_of_b || normmax > tolerance * norm_max_of_b) {
                double tol = tolerance * std::min(norm2_of_b, 0.1 * norm2 / normmax * norm_max_of_b);
                last2 = norm2;
                lastmax = normmax;
                status = opt->plain == 3   ? solver_minres(S, r.data(), x, &cntIter, &tol)
                         : opt->plain == 2 ? solver_bicgstab(S, r.data(), x, &cntIter, &tol)
                         : opt->plain      ? solver_pcg_plain(S, r.data(), x, &cntIter, &tol)
                                           : solver_pcg_paso(S, r.data(), x, &cntIter, &tol);
                totIter += cntIter;
                if (status == NoError) {
                    finalize = false;
                } else if (status == Breakdown && cntIter > 1) {
                    finalize = false;
                    status = NoError;
                } else {
                    break;
                }
            } else {
                diag->converged = 1;
            }
        }
        diag->net_time = omp_get_wtime() - tn;
        diag->num_iter = totIter;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) x[i] *= d[i];
    }
    diag->status = status;
    diag->time = omp_get_wtime() - t0;
    return status;
}

// A bounded sample of PASO's inner loop for the CPU baseline: runs exactly
// `iters` iterations of Solver_PCG (or the plain variant) on the balanced
// system from the initial guess, and returns the wall time of the loop.
ORC_API double orc_time_pcg_iterations(int bs, int64_t nrows, const index_t* ptr, const index_t* index,
                               double* val, const double* b, int iters, int plain, double* d)
{
    const int64_t n = nrows * bs;
    orc_balance(bs, nrows, ptr, index, val, d);
    std::vector<double> inv((size_t)nrows * bs * bs), r(n), x(n);
    if (orc_inv_main(bs, nrows, ptr, index, val, inv.data())) return -1.0;
    Sys S{bs, nrows, n, ptr, index, val, inv.data(), 1, std::vector<double>(n)};
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) r[i] = b[i] * d[i];
    S.prec(x.data(), r.data());
    S.mv(-1.0, x.data(), 1.0, r.data());
    int it = iters;
    double tol = 0.0;   // never met: run exactly `iters` (+1 for PASO's check order)
    const double t0 = omp_get_wtime();
    if (plain) solver_pcg_plain(S, r.data(), x.data(), &it, &tol);
    else {
        it = iters - 1;
        solver_pcg_paso(S, r.data(), x.data(), &it, &tol);
    }
    return omp_get_wtime() - t0;
}

// ------------------------------------------------- diagonal-offset (CDS)
// The reference's stencil layout with add() semantics: the slot of a pair of
// element nodes follows from their local positions in the element, not from
// their node numbers (RipleySystemMatrix.cu:111-155), and the SpMV skips
// columns outside [0, nrows) (cusplibrary/cusp/detail/host/spmv.h:240-437,
// symmetric storage keeps the main and upper diagonals and applies the
// transposed blocks for the lower ones).  values[(row*bs + k) * ndiag*bs +
// diag*bs + m] holds block entry (k, m) of diagonal `diag` in node row `row`.
// These functions reproduce the known answers of
// ripley/test/SystemMatrixTestCase.cpp:41-134.
ORC_API void orc_cds_add(int dim, int bs, int symmetric, int ndiag, double* values, int nN,
                 const index_t* nodes, const double* array)
{
    const int offset = symmetric ? (dim == 2 ? 4 : 13) : 0;
    for (int i = 0; i < nN; ++i)
        for (int j = symmetric ? i : 0; j < nN; ++j) {
            const int ax = i & 1, ay = (i >> 1) & 1, az = (i >> 2) & 1;
            const int bx = j & 1, by = (j >> 1) & 1, bz = (j >> 2) & 1;
            const int slot = (bx - ax + 1) + 3 * (by - ay + 1) + (dim == 3 ? 9 * (bz - az + 1) : 0);
            const int diag = slot - offset;
            for (int k = 0; k < bs; ++k)
                for (int m = 0; m < bs; ++m) {
                    const int64_t row = (int64_t)nodes[i] * bs + k;
                    values[row * ndiag * bs + diag * bs + m] += array[k + bs * (m + bs * (i + nN * j))];
                }
        }
}

ORC_API void orc_cds_spmv(int bs, int symmetric, int64_t nnodes, int ndiag, const int32_t* offsets,
                  const double* values, const double* x, double* y)
{
    const int64_t nr = nnodes * bs;
    for (int64_t i = 0; i < nr; ++i) y[i] = 0.0;
    for (int64_t rn = 0; rn < nnodes; ++rn) {
        for (int d = 0; d < ndiag; ++d) {
            const int64_t cn = rn + offsets[d];
            if (cn < 0 || cn >= nnodes) continue;
            for (int k = 0; k < bs; ++k)
                for (int m = 0; m < bs; ++m)
                    y[rn * bs + k] += values[(rn * bs + k) * ndiag * bs + d * bs + m] * x[cn * bs + m];
        }
        if (symmetric) {
            for (int d = 0; d < ndiag; ++d) {
                if (offsets[d] == 0) continue;
                const int64_t cn = rn - offsets[d];
                if (cn < 0 || cn >= nnodes) continue;
                // lower block (rn, cn) = transpose of the stored block (cn, rn)
                for (int k = 0; k < bs; ++k)
                    for (int m = 0; m < bs; ++m)
                        y[rn * bs + k] += values[(cn * bs + m) * ndiag * bs + d * bs + k] * x[cn * bs + m];
            }
        }
    }
}

ORC_API int orc_num_threads(void) { return omp_get_max_threads(); }
ORC_API void orc_set_num_threads(int n) { omp_set_num_threads(n); }

