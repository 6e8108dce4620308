/*
 * ripley_hip -- C ABI of the MI355X (gfx950) implementation of escript's
 * structured-grid FEM hot path: ripley Rectangle/Brick Q1 assembly into a
 * PASO-compatible system matrix, Dirichlet constraints, SpMV and PASO's
 * Jacobi-preconditioned CG driver, at one MPI rank.
 *
 * Every entry point names the reference interface it replaces.  All entry
 * points return an int status: 0 on success, 1..6 the paso::SolverResult codes
 * (paso/src/Paso.h:44-52), >= 20 an argument, state or device error; the
 * message of the last failure on the calling thread is rg_last_error().
 *
 * Pointers: arguments documented "device" are HBM pointers (hipMalloc /
 * torch.cuda memory) on the context's device; "host" arguments are ordinary
 * host memory (escript Data buffers: DOF-major, then component, as
 * paso::SystemMatrix reads them, paso/src/SystemMatrix.h:637-641).  Host
 * arrays are borrowed for the duration of the call.  Device memory owned by a
 * handle is released by its destroy call.  Work is issued on the context's
 * stream; functions returning host results synchronise that stream.
 */
#ifndef RIPLEY_HIP_H
#define RIPLEY_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status codes */
enum {
    RG_OK = 0,
    RG_MAX_ITER_REACHED = 1, /* paso::MaxIterReached */
    RG_INPUT_ERROR = 2,      /* paso::InputError */
    RG_MEMORY_ERROR = 3,     /* paso::MemoryError */
    RG_BREAKDOWN = 4,        /* paso::Breakdown */
    RG_NEGATIVE_NORM = 5,    /* paso::NegativeNormError */
    RG_DIVERGENCE = 6,       /* paso::Divergence */
    RG_VALUE_ERROR = 20,     /* bad argument (PasoException/ValueError in the reference) */
    RG_STATE_ERROR = 21,     /* e.g. MatrixVector on a balanced matrix */
    RG_DEVICE_ERROR = 22     /* HIP runtime failure */
};

typedef struct rg_context rg_context;
typedef struct rg_matrix rg_matrix;
typedef struct rg_lpc rg_lpc;
typedef struct rg_batch rg_batch;

/* Message of the last failure on this thread (empty string if none). */
const char* rg_last_error(void);

/* Library version string. */
const char* rg_version(void);

/* ------------------------------------------------------------------ context
 * One device and one stream.  stream may be NULL (the null stream) or a
 * hipStream_t owned by the caller (e.g. torch.cuda.current_stream()).
 * Replaces nothing in the reference (whose path never leaves the host); it
 * carries what escript keeps in JMPI + OpenMP state. */
int rg_context_create(int device, void* stream, rg_context** out);
int rg_context_destroy(rg_context* ctx);
int rg_context_set_stream(rg_context* ctx, void* stream);
int rg_context_synchronize(rg_context* ctx);

/* ------------------------------------------------------------------- matrix
 * rg_matrix_create replaces RipleyDomain::newSystemMatrix ->
 * getPasoMatrixPattern -> paso::SystemMatrix<double> ctor
 * (ripley/src/RipleyDomain.cpp:950-1022, ripley/src/Brick.cpp:2546-2748,
 * paso/src/SystemMatrix.h:404-492) for a Rectangle (dim 2) or Brick (dim 3)
 * with ne[d] elements and side length[d] per axis, at one rank.  Values are
 * zero.  block_size = numEq = numComp (1..4).  flags: RG_MATRIX_SYMMETRIC
 * stores only the main and upper stencil diagonals (the matrix is then
 * symmetric by construction).  The pattern is PASO's CSR pattern; it is held
 * as a diagonal-offset stencil (9 or 27 slots per row in
 * getDiagonalIndices order, Brick.cpp:2484-2507). */
enum { RG_MATRIX_SYMMETRIC = 1 };
int rg_matrix_create(rg_context* ctx, int dim, const int32_t* ne, const double* length,
                     int block_size, int flags, rg_matrix** out);
int rg_matrix_destroy(rg_matrix* m);

/* num_rows = number of block rows (nodes); nnz = PASO block nnz. */
int rg_matrix_get_info(const rg_matrix* m, int64_t* num_rows, int64_t* nnz, int* num_slots,
                       int* block_size);

/* paso::SystemMatrix::setValues (paso/src/SystemMatrix.h:273-279): every
 * entry of the pattern := value; clears the balanced state. */
int rg_matrix_set_values(rg_matrix* m, double value);

/* AbstractSystemMatrix::resetValues(preserveSolverData)
 * (paso/src/SystemMatrix.h:607-613): setValues(0) and, unless
 * preserve_solver_data, drops the preconditioner. */
int rg_matrix_reset_values(rg_matrix* m, int preserve_solver_data);

/* RipleyDomain::addToSystemMatrix<real_t> (ripley/src/RipleyDomain.cpp:
 * 1433-1459; its intended hook for a non-PASO matrix is the dead CUDA branch
 * :1444-1450 -> ripley::SystemMatrix::add, RipleySystemMatrix.cu:111-155).
 * Adds num_elements element matrices.  nodes (host) holds
 * nodes_per_element (4 in 2-D, 8 in 3-D) node ids per element in the
 * element's local order; em (host) holds per element
 * EM[k + n*(m + n*(a + nN*b))] with n = num_eq = block size.  The stencil slot
 * of a node pair follows from the local positions a, b, as in the reference
 * stencil matrix.  Elements are added in 2^dim launches, one per element
 * colour (the parity of local node 0's grid coordinates), so grid cells of
 * one colour share no node and the result is the same on every run; the
 * node ids and element matrices go through a device staging buffer that the
 * matrix keeps (grown when a larger batch arrives) and the call returns
 * without waiting for the kernels.  Adapters pass many elements per call
 * (all of them, or one colour's), not one. */
int rg_matrix_add_element_blocks(rg_matrix* m, int64_t num_elements, int nodes_per_element,
                                 const int32_t* nodes, int num_eq, const double* em);

/* No counterpart in the reference: rg_matrix_add_element_blocks for callers
 * that cannot allocate device memory (code under graph capture, tight
 * assembly loops, harnesses that forbid hipMalloc).  Same arguments and
 * result; the node ids and element matrices are staged in a fixed 16 MiB
 * device array per GPU (RG_SCRATCH_BYTES, held until the process ends)
 * instead of the matrix's staging buffer, and the call allocates nothing.
 * A call whose staging (node ids rounded up to 256 bytes, plus the element
 * matrices) exceeds RG_SCRATCH_BYTES fails with RG_VALUE_ERROR: pass fewer
 * elements per call.  Every call on one GPU shares that array, so calls that
 * overlap on one GPU (different streams or threads) race. */
#define RG_SCRATCH_BYTES (16 << 20)
int rg_matrix_add_element_blocks_noalloc(rg_matrix* m, int64_t num_elements, int nodes_per_element,
                                         const int32_t* nodes, int num_eq, const double* em);

/* PDE coefficients (device pointers, escript data-point layout, column
 * major within a data point):
 *   single PDE: A[i + dim*j], B[i], C[j], D, X[i], Y
 *   system (n = block size): A[k + n*(i + dim*(m + n*j))], B[k + n*(i + dim*m)],
 *                            C[k + n*(m + n*j)], D[k + n*m], X[k + n*i], Y[k]
 * mode RG_COEF_ELEMENT: one data point per element, elem_stride doubles apart
 * (0: one point shared by all elements) -- the reference's "constant data"
 * branch; RG_COEF_EXPANDED: one point per quadrature point (4 in 2-D, 8 in
 * 3-D, x fastest), elem_stride = points_per_element * point size.
 * isotropic (A of a single PDE only): each data point is one scalar a and
 * means A = a * identity. */
enum { RG_COEF_EMPTY = 0, RG_COEF_ELEMENT = 1, RG_COEF_EXPANDED = 2 };
typedef struct {
    const double* data;
    int32_t mode;
    int32_t elem_stride;
    int32_t isotropic;
    int32_t reserved;
} rg_coef;

typedef struct {
    rg_coef A, B, C, D, X, Y;
} rg_pde_coefs;

/* RipleyDomain::addToSystem -> assemblePDE -> DefaultAssembler{2D,3D}::
 * assemblePDESingle / assemblePDESystem (+ *Reduced when reduced != 0)
 * -> addToMatrixAndRHS (ripley/src/RipleyDomain.cpp:1024-1035, :1618-1688;
 * ripley/src/DefaultAssembler3D.cpp:153-2229, :2891-3270, :3631-5745;
 * ripley/src/DefaultAssembler2D.cpp:152-597, :842-1011, :1186-1679).
 * Adds the element contributions to the matrix and rhs (device, may be NULL;
 * num_rows*block_size doubles).  Each node gathers its own contributions; no
 * atomics; deterministic. */
int rg_matrix_assemble(rg_matrix* m, const rg_pde_coefs* coefs, int reduced, double* rhs);

/* RipleyDomain::assemblePDEBoundary -> DefaultAssembler2D/3D::
 * assemblePDEBoundarySingle/System[Reduced] (DefaultAssembler2D.cpp:601-840,
 * 1014-1200; DefaultAssembler3D.cpp:2233-2890, 3274-3630): adds the face
 * terms  int_boundary d u v  to the matrix and  int_boundary y v  to rhs
 * (device, may be NULL).  d: n x n per data point (d[k + n*m]), y: n per
 * point; one data point per face element (RG_COEF_ELEMENT) or per face
 * quadrature point (RG_COEF_EXPANDED: 2 in 2-D, 4 in 3-D, lower tangential
 * axis fastest).  Face elements are numbered as the reference's
 * FunctionOnBoundary samples: faces x0=0, x0=L, x1=0, x1=L[, x2=0, x2=L],
 * within a face by the two tangential element indices, lower axis fastest.
 * reduced: one-point rule at the face centre (element data only).  Either
 * coefficient may be NULL. */
int rg_matrix_assemble_boundary(rg_matrix* m, const rg_coef* d, const rg_coef* y, int reduced, double* rhs);

/* RipleyDomain::assemblePDEDirac (ripley/src/RipleyDomain.cpp:1756-1800):
 * for Dirac point p (in order) at node nodes[p] (host), rhs[k + n*node] +=
 * y[p*n + k] and the node's diagonal block += d[p*n*n + k + n*m].  d, y
 * host (either may be NULL); rhs device (required with y). */
int rg_matrix_add_dirac(rg_matrix* m, int64_t count, const int64_t* nodes, const double* d, const double* y,
                        double* rhs);

/* paso::SystemMatrix::MatrixVector (paso/src/SystemMatrix_MatrixVector.cpp:
 * 62-91): y = alpha*A*x + beta*y (device x, y).  Fails with RG_STATE_ERROR
 * on a balanced matrix ("MatrixVector: balanced matrix is not supported.").
 * x is assumed finite: at the x0/x1 grid edges a stencil slot that leaves
 * the grid holds 0 and is multiplied by the neighbouring line's entry of x
 * (no per-slot branch in the kernel), so an Inf or NaN there reaches rows
 * that PASO's CSR product does not couple to it.  The same holds for
 * rg_matrix_ypAx and the solvers' products. */
int rg_matrix_vector(rg_matrix* m, double alpha, const double* x, double beta, double* y);

/* AbstractSystemMatrix::ypAx (paso/src/SystemMatrix.h:646-671): y += A*x. */
int rg_matrix_ypAx(rg_matrix* m, double* y, const double* x);
int rg_matrix_ypAx_host(rg_matrix* m, double* y, const double* x);

/* AbstractSystemMatrix::nullifyRowsAndCols (paso/src/SystemMatrix.h:544-605,
 * SparseMatrix.h:494-535): a_ij := (i==j ? mdv : 0) where row_mask[i] > 0 or
 * col_mask[j] > 0 (device masks, num_rows*block_size doubles). */
int rg_matrix_nullify_rows_and_cols(rg_matrix* m, const double* row_mask, const double* col_mask,
                                    double mdv);
int rg_matrix_nullify_rows_and_cols_host(rg_matrix* m, const double* row_mask,
                                         const double* col_mask, double mdv);

/* The PASO CSR form of the matrix (host buffers): ptr[num_rows+1],
 * index[nnz], val[nnz*bs*bs] with val[k*bs*bs + ir + bs*ic]
 * (paso/src/SystemMatrix.h:404-492).  Any of the three may be NULL. */
int rg_matrix_export_csr(rg_matrix* m, int32_t* ptr, int32_t* index, double* val);

/* AbstractSystemMatrix::saveMM (paso/src/SparseMatrix.h:79): MatrixMarket
 * coordinate file of the scalar matrix. */
int rg_matrix_save_mm(rg_matrix* m, const char* filename);

/* ------------------------------------------------------------------ solver
 * Options mirror the escript::SolverBuddy fields that PASO reads
 * (paso/src/Options.cpp:31-56; defaults escriptcore/src/SolverOptions.cpp:
 * 40-72).  method:
 *   RG_METHOD_PCG        PASO Solver() + Solver_PCG (residual-smoothed PCG)
 *   RG_METHOD_PCG_PLAIN  same outer loop, textbook Jacobi-PCG inside
 *   RG_METHOD_PCG_MIXED  same outer loop; inner Jacobi-PCG whose SpMV and
 *                        preconditioner use FP32 matrix values while dot
 *                        products, scalars and vector updates stay FP64; every
 *                        check_interval iterations the residual is replaced by
 *                        the FP64 true residual D b - A x, and convergence is
 *                        only accepted on such a replaced residual
 *   RG_METHOD_BICGSTAB    PASO's Solver_BiCGStab (paso/src/BiCGStab.cpp:77-257)
 *                         in FP64 with the Jacobi preconditioner (one sweep),
 *                         the method PASO uses for non-symmetric systems.
 *   RG_METHOD_MINRES      PASO's Solver_MINRES (paso/src/MINRES.cpp:59-191) in
 *                         FP64 for symmetric (possibly indefinite) systems;
 *                         r.M^-1 r < 0 ends the solve with RG_NEGATIVE_NORM.
 *   RG_METHOD_FCG_LEARNED flexible CG keeping the last 5 directions (each
 *                        new one A-orthogonalised against them; FP64
 *                        operator, vectors and scalars) preconditioned by one
 *                        geometric-multigrid V-cycle per iteration: Galerkin
 *                        coarse operators P^T A P of the Brick halved per
 *                        level (P trilinear), Chebyshev smoothing on FP32
 *                        values before and after each coarse correction, and
 *                        the coarsest level solved by the network attached
 *                        with rg_matrix_set_learned_preconditioner (scaled to
 *                        minimise the level's energy-norm error) or, with
 *                        none attached, by Chebyshev steps
 *                        (rg_matrix_set_multilevel).  Scalar Brick matrices.
 *                        If the residual stagnates (no 1% decrease over two
 *                        check intervals) the solve falls back to
 *                        RG_METHOD_PCG_MIXED (Jacobi) from the current iterate.
 *                        Extension: the reference has no counterpart.
 * criterion:
 *   RG_CRITERION_PASO           balanced ||r||_2 and ||r||_inf tests (Solver.cpp:241)
 *   RG_CRITERION_TRUE_RESIDUAL  unbalanced FP64 ||b - A x||_2 <= tolerance ||b||_2
 *                               (RG_METHOD_FCG_LEARNED stops its inner solve on
 *                               the unbalanced norm of its own residual at 0.8 x
 *                               the target; the other methods on the balanced
 *                               norm, scaled by the observed ratio)
 * preconditioner (Preconditioner_alloc / Preconditioner_solve,
 * paso/src/Preconditioner.cpp:46-137) for RG_METHOD_PCG, RG_METHOD_PCG_PLAIN,
 * RG_METHOD_BICGSTAB and RG_METHOD_MINRES; the FP32 and learned methods use
 * Jacobi:
 *   RG_PRECONDITIONER_JACOBI        `sweeps` Jacobi sweeps (Smoother.cpp:100-127)
 *   RG_PRECONDITIONER_GAUSS_SEIDEL  `sweeps` symmetric Gauss-Seidel sweeps over the
 *                                   colouring of Pattern::borrowColoringPointer
 *                                   (Smoother.cpp:331-507, Pattern.cpp:166-195)
 *   RG_PRECONDITIONER_ILU0          coloured ILU(0) (ILU.cpp:43-462), block size <= 3
 *   RG_PRECONDITIONER_NONE          identity */
enum { RG_METHOD_PCG = 0, RG_METHOD_PCG_PLAIN = 1, RG_METHOD_PCG_MIXED = 2, RG_METHOD_FCG_LEARNED = 3,
       RG_METHOD_BICGSTAB = 4, RG_METHOD_MINRES = 5 };
enum { RG_CRITERION_PASO = 0, RG_CRITERION_TRUE_RESIDUAL = 1 };
enum { RG_PRECONDITIONER_JACOBI = 0, RG_PRECONDITIONER_GAUSS_SEIDEL = 1, RG_PRECONDITIONER_ILU0 = 2,
       RG_PRECONDITIONER_NONE = 3 };
typedef struct {
    double tolerance;            /* relative; default 1e-8 */
    double inner_tolerance;      /* reserved (default 1e-4) */
    int32_t iter_max;            /* default 100000 */
    int32_t sweeps;              /* Jacobi / Gauss-Seidel sweeps, default 1 */
    int32_t method;
    int32_t verbose;
    int32_t accept_failed_convergence;
    int32_t check_interval;      /* iterations per replayed graph / host check (0: 32; 1 for
                                    RG_METHOD_FCG_LEARNED) */
    int32_t criterion;
    int32_t preconditioner;      /* RG_PRECONDITIONER_*, default Jacobi */
} rg_options;

/* Diagnostics keys of paso::Options::updateEscriptDiagnostics
 * (paso/src/Options.cpp:441-459) plus the number of outer steps and of
 * switches from the learned preconditioner to Jacobi-PCG. */
typedef struct {
    int32_t num_iter;
    int32_t converged;
    int32_t status;
    int32_t outer_steps;
    double residual_norm;
    double set_up_time;
    double net_time;
    double time;
    int32_t fallbacks;
    int32_t reserved;
} rg_diagnostics;

void rg_options_default(rg_options* opt);

/* AbstractSystemMatrix::solve -> paso::SystemMatrix::setToSolution ->
 * SystemMatrix<double>::solve -> Solver (paso/src/SystemMatrix.h:615-644,
 * paso/src/solve.cpp:41-145, paso/src/Solver.cpp:46-331, PCG.cpp:62-368).
 * x, b: device, num_rows*block_size doubles.  As in PASO the matrix is left
 * balanced (MatrixVector then fails until set/reset values).  Returns the
 * SolverResult code mapped as paso/src/solve.cpp:119-143 does (with
 * accept_failed_convergence, MaxIterReached/Divergence return RG_OK). */
int rg_solve(rg_matrix* m, double* x, const double* b, const rg_options* opt,
             rg_diagnostics* diag);
int rg_solve_host(rg_matrix* m, double* x, const double* b, const rg_options* opt,
                  rg_diagnostics* diag);

/* ------------------------------------------------------- binary grid I/O
 * ripley's readBinaryGrid / writeBinaryGrid at one rank (host memory, no GPU
 * needed).  Byte orders and data types as ripley/src/system_dep.h:44-70
 * (the native order of the hosts this library runs on is little endian). */
enum { RG_BYTEORDER_LITTLE_ENDIAN = 1234, RG_BYTEORDER_BIG_ENDIAN = 4321,
       RG_BYTEORDER_NATIVE = RG_BYTEORDER_LITTLE_ENDIAN };
enum { RG_DATATYPE_INT32 = 1, RG_DATATYPE_FLOAT32 = 2, RG_DATATYPE_FLOAT64 = 3 };

/* readBinaryGrid (ripley/src/ripleycpp.cpp:63-86, Rectangle.cpp:572-690,
 * Brick.cpp:669-818): grid = samples per axis of the target function space
 * (nodes, or elements for Function / ReducedFunction); out = host array of
 * grid[0]*grid[1](*grid[2]) samples x points_per_sample x num_comp, already
 * holding the fill value.  num_values[d] values per axis are read from the
 * file (x fastest, num_comp components per value); value v lands on
 * multiplier[d] consecutive samples per axis from first[d] on; reverse is
 * supported for the z axis of a Brick only (as in the reference).  NaN values
 * leave the sample unchanged. */
int rg_read_binary_grid(const char* filename, int dim, const int64_t* grid, int num_comp, int points_per_sample,
                        const int64_t* first, const int64_t* num_values, const int32_t* multiplier,
                        const int32_t* reverse, int byte_order, int data_type, double* out);

/* writeBinaryGrid (Brick.cpp:964-1050 and Rectangle's): one scalar value per
 * sample, x fastest, converted to data_type (C conversion; truncation for
 * INT32) in byte_order. */
int rg_write_binary_grid(const char* filename, int dim, const int64_t* grid, const double* in, int byte_order,
                         int data_type);

/* SystemMatrix::setPreconditioner + solvePreconditioner
 * (paso/src/SystemMatrix.cpp:42-56): balances the matrix if it is not yet
 * (as a solve would), builds opt->preconditioner with opt->sweeps if needed
 * and computes x = M^-1 b in the balanced space.  With opt->method =
 * RG_METHOD_FCG_LEARNED, x is one V-cycle of that method's multilevel
 * preconditioner.  x, b: device,
 * num_rows*block_size doubles, distinct. */
int rg_matrix_solve_preconditioner(rg_matrix* m, const rg_options* opt, double* x, const double* b);

/* Pattern::borrowColoringPointer / getNumColors (paso/src/Pattern.cpp:
 * 166-195, Pattern_mis.cpp:44-110) of the matrix graph, computed on the
 * device: colours (host, num_rows entries, may be NULL) and their number. */
int rg_matrix_coloring(rg_matrix* m, int32_t* colours, int* num_colors);

/* ---------------------------------------------------------- batched solves
 * No counterpart in the reference (PASO solves one right-hand side of one
 * system per call, paso/src/solve.cpp:41-145).  Both solvers run the
 * textbook Jacobi-PCG (PCG.cpp:62-368 without the residual smoothing) on
 * the UNBALANCED system from x = 0, with FP64 (RG_METHOD_PCG_PLAIN) or
 * FP32 (RG_METHOD_PCG_MIXED) matrix values in the SpMV, FP64 vectors, dots
 * and updates, and an FP64 residual replacement r = b - A x every
 * check_interval iterations; each system / column converges on
 * ||b - A x||_2 <= tolerance * ||b||_2 of that FP64 residual.  The matrices
 * must not have been balanced by rg_solve (set or reset their values
 * first).  Scalar matrices with full storage only.  diag receives one entry
 * per column / system; the return value is the worst status (with
 * accept_failed_convergence, MaxIterReached is not an error). */

/* Many right-hand sides on one matrix: X, B device, row-major
 * num_rows x nrhs (column index fastest), 1 <= nrhs <= 64.  The SpMM reads
 * the matrix once per iteration for all columns; its tile shape is
 * autotuned on the first call (rg_multi_autotune sets it explicitly).
 * With RG_METHOD_PCG_MIXED on a z-marching shape (4-6 below) the search
 * direction p and q = A p are stored in FP32 (r, x, the dot products and
 * the updates stay FP64).  x is updated every second iteration (two
 * directions at once), so check_interval is rounded up to an even count. */
int rg_solve_multi(rg_matrix* m, double* X, const double* B, int nrhs, const rg_options* opt,
                   rg_diagnostics* diag);
/* Y = A X for a row-major num_rows x nrhs block (device). */
int rg_matrix_vector_multi(rg_matrix* m, const double* X, double* Y, int nrhs);
/* Times reps SpMMs (with the partial products p.q, as the solver runs
 * them) for each tile shape -- (rows, columns) per thread of
 * 0: (1,1), 1: (1,2), 2: (1,4), 3: (2,2), and 4, 5: z-marching tiles of
 * 16 x 4 and 8 x 4 nodes with their X planes staged in LDS, one column per
 * lane; 6: the 8 x 4 z-marching tile with one node and 8 columns per lane
 * (4-6: exactly 64 columns, Brick matrices only); shapes that do not apply
 * are skipped
 * (time -1) -- keeps the fastest for this matrix and returns its index in
 * *shape; times (host, may be NULL) receives RG_MULTI_SHAPES seconds/SpMM. */
#define RG_MULTI_SHAPES 7
int rg_multi_autotune(rg_matrix* m, int nrhs, int fp32, int reps, int* shape, double* times);
/* Fixes the SpMM shape (0..RG_MULTI_SHAPES-1 as above) or, with -1,
 * autotunes on the next solve. */
int rg_multi_set_shape(rg_matrix* m, int shape);

/* A batch of independent systems (count matrices of one dimension, sizes
 * free) solved together: every kernel launch spans all systems (a tile
 * table maps workgroups to (system, rows)).  Vectors are the systems'
 * vectors concatenated at rg_batch_offsets (count + 1 entries, each offset
 * a multiple of 64).  The matrices are borrowed. */
int rg_batch_create(rg_context* ctx, int count, rg_matrix* const* mats, rg_batch** out);
int rg_batch_destroy(rg_batch* b);
int rg_batch_offsets(const rg_batch* b, int64_t* offsets);
/* y = A x for every system (device, concatenated). */
int rg_batch_vector(rg_batch* b, const double* x, double* y);
/* Rows per tile = 256 * (2 FP64 | 4 FP32) * tile_mult; 0 = autotune on the
 * next solve.  rg_batch_autotune times reps batched SpMVs per candidate
 * (1, 2, 4, 8), keeps the fastest; times (host, may be NULL): 4 entries. */
int rg_batch_set_tile(rg_batch* b, int tile_mult);
int rg_batch_autotune(rg_batch* b, int fp32, int reps, int* tile_mult, double* times);
int rg_batch_solve(rg_batch* b, double* x, const double* rhs, const rg_options* opt, rg_diagnostics* diag);

/* --------------------------------------------------- learned preconditioner
 * No counterpart in the reference (its preconditioners are PASO's Jacobi,
 * Gauss-Seidel, ILU0, RILU, AMG).  A 3-level 3-D convolutional U-Net
 * (16 -> 32 -> 64 channels, 3x3x3 kernels, BF16 weights, FP32 accumulate)
 * on a node grid: input channels are the Jacobi-scaled residual (normalised
 * to unit RMS) and log10(kappa) per node; the output is an error correction
 * e ~ A^-1 r.  Parameters: per layer W[co][27][ci] then b[co] (FP32), 9
 * layers, rg_lpc_num_params() floats.  A network is created for one node grid
 * (nodes[0..2]); `training` allocates the backward-pass buffers. */
int64_t rg_lpc_num_params(void);
int rg_lpc_create(rg_context* ctx, const int32_t* nodes, int training, rg_lpc** out);
int rg_lpc_destroy(rg_lpc* net);
int rg_lpc_set_params(rg_lpc* net, const float* params);   /* host */
int rg_lpc_get_params(rg_lpc* net, float* params);         /* host */
/* z = Net(r) (device vectors on the node grid; dinv = 1/diag(A), logk
 * float per node, either may be NULL). */
int rg_lpc_apply(rg_lpc* net, const double* r, const double* dinv, const float* logk, double* z);
/* Loss of Net(r) against the error e (device) -- the cosine loss
 * 1 - (y.e)/(|y||e|), or the energy-norm loss of rg_lpc_set_loss_operator --
 * and its parameter gradient (host, may be NULL). */
int rg_lpc_gradients(rg_lpc* net, const double* r, const double* dinv, const float* logk, const double* e,
                     double* loss, float* grads);
/* One Adam step (lr) on that loss. */
int rg_lpc_train_step(rg_lpc* net, const double* r, const double* dinv, const float* logk, const double* e,
                      float lr, double* loss);
/* Loss of rg_lpc_gradients / rg_lpc_train_step: with an operator A (a
 * scalar Brick matrix on the network's node grid, not balanced, borrowed)
 * the energy-norm loss 1 - (r.z)^2 / ((r.e)(z.A z)) of the output z for a
 * pair with r = A e -- the fraction of the error energy that the
 * energy-optimal step along z leaves, which is what the coarsest level of
 * the multilevel preconditioner removes; A = NULL restores the cosine loss.
 * Training networks only. */
int rg_lpc_set_loss_operator(rg_lpc* net, rg_matrix* A);
/* Attach a network as the coarsest-level solver of RG_METHOD_FCG_LEARNED's
 * multilevel preconditioner: its node grid must be a coarse level of the
 * matrix's Brick (ne / 2^k elements per axis, k >= 1), which sets the
 * number of levels.  On that level the network's inputs are the
 * Jacobi-scaled residual and log10(a_ii / a_ii of a unit coefficient) of the
 * level's Galerkin operator.  logk is reserved (pass NULL).  net = NULL
 * detaches. */
int rg_matrix_set_learned_preconditioner(rg_matrix* m, rg_lpc* net, const float* logk);

/* Hierarchy of RG_METHOD_FCG_LEARNED: `levels` grid levels without a
 * network (0: until a level has at most 9^3 nodes or cannot be halved; an
 * attached network's grid sets the levels), `degree` Chebyshev steps before
 * and after each coarse correction (1..16, default 2) and `coarse_steps`
 * Chebyshev steps from zero on the coarsest level (1..16, default 8), which
 * the network's correction follows.  No counterpart in the reference. */
int rg_matrix_set_multilevel(rg_matrix* m, int levels, int degree, int coarse_steps);

/* The Galerkin operator P^T A P of the matrix `levels` halvings down (P the
 * trilinear interpolation of the multilevel preconditioner): a new matrix of
 * the coarse Brick with the same storage, owned by the caller
 * (rg_matrix_destroy).  The matrix must not be balanced.  Scalar Brick
 * matrices.  No counterpart in the reference. */
int rg_matrix_coarse_operator(rg_matrix* m, int levels, rg_matrix** out);

/* x = `steps` (1..16) Chebyshev steps from zero for A x = b on the matrix
 * (not balanced; scalar Brick), the smoother of the multilevel
 * preconditioner's levels: Jacobi-scaled, over [g / 10, g] with g the
 * Gershgorin bound max_i sum_j |a_ij| / a_ii, FP32 values.  x, b: device,
 * distinct.  No counterpart in the reference (training data of the
 * coarse-level network). */
int rg_matrix_smooth(rg_matrix* m, const double* b, double* x, int steps);

/* x = one V-cycle of RG_METHOD_FCG_LEARNED's multilevel preconditioner for
 * A x = b on the matrix itself (not balanced; scalar Brick; the levels,
 * smoothing and network of rg_matrix_set_multilevel and
 * rg_matrix_set_learned_preconditioner).  x, b: device, distinct.  No
 * counterpart in the reference (training data of the network: the residual
 * and error a V-cycle leaves). */
int rg_matrix_vcycle(rg_matrix* m, const double* b, double* x);

#ifdef __cplusplus
}
#endif

#endif /* RIPLEY_HIP_H */
