// Internal declarations of libripley_hip (not part of the C ABI).
//
// Data layout in HBM (one rank's share of a Rectangle/Brick):
//   * block rows are nodes, numbered lexicographically (x fastest) as
//     ripley numbers DOFs at one rank (ripley/src/Brick.cpp:3000-3061);
//   * the matrix is a diagonal-offset stencil: S = 9 (2-D) or 27 (3-D) slots
//     per node in getDiagonalIndices order (Brick.cpp:2484-2507); with
//     symmetric storage only the main and upper slots (5 / 14) are kept;
//   * values are structure-of-arrays: entry (slot s, ir, ic) of node i is
//     val[(s*bs*bs + ir + bs*ic) * ld + i] -- one coalesced stream per slot
//     and block entry, the block entry order being PASO's column-major block
//     (paso/src/SparseMatrix_MatrixVector.cpp:288-320);
//   * vectors are DOF-major then component, x[i*bs + c], as escript Data;
//   * the Jacobi inverse is structure-of-arrays dinv[(ir + bs*ic) * ld + i].
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "ripley_hip.h"

namespace rg {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define RG_HIP(call)                                                                          \
    do {                                                                                      \
        hipError_t e_ = (call);                                                               \
        if (e_ != hipSuccess)                                                                 \
            throw ::rg::Error(RG_DEVICE_ERROR, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define RG_CHECK(cond, code, msg)                                \
    do {                                                         \
        if (!(cond)) throw ::rg::Error((code), std::string(msg)); \
    } while (0)

constexpr int kBlock = 256;          // threads per workgroup of every streaming kernel
constexpr int kNumXcd = 8;           // MI355X: 8 XCDs, blocks dispatched round-robin
constexpr int kMaxGrid = 8192;       // grid cap of chunked streaming kernels (partials per reduction)
// Size of the per-device scratch array of the allocation-free entry points.
constexpr int64_t kScratchBytes = RG_SCRATCH_BYTES;

// Geometry of one matrix (one rank's share).
struct Geom {
    int dim = 3;
    int bs = 1;
    int ne[3] = {1, 1, 1};
    int64_t nn[3] = {1, 1, 1};
    double len[3] = {1, 1, 1};
    double h[3] = {1, 1, 1};
    int64_t N = 0;       // nodes = block rows
    int64_t s1 = 0;      // nn0
    int64_t s2 = 0;      // nn0*nn1
    int nslot = 27;      // slots of the full stencil
    int nstore = 27;     // slots stored (nslot or the upper half + main)
    int first = 0;       // full-stencil index of stored slot 0
    int64_t ld = 0;      // leading dimension of value / dinv streams
    int64_t nnz = 0;     // PASO block nnz
};

// Values held by a matrix's value arrays (val, val32): the stored slots and
// kValSlack more, so that wide loads of a run of consecutive rows that ends
// past the last row (the partial tiles of the 64-column SpMM) stay inside
// the allocation; the slack is never read into a result.
constexpr int64_t kValSlack = 64;
inline size_t value_count(const Geom& g) { return (size_t)g.nstore * g.bs * g.bs * g.ld + kValSlack; }

// Scalars of one in-flight Krylov solve, resident in device memory and
// updated by the last workgroup of each reducing kernel.
struct KrylovState {
    double tau, tau_new, delta, alpha, beta, gamma1, norm, tol;
    double rho, rnorm0, rho_prev;
    double alpha_prev;      // many-RHS PCG: alpha of the previous iteration (its x update pending)
    int32_t num_iter, maxit, done, status;
    int32_t replace_mode;   // 1: convergence is only declared on a replaced (true) residual
    int32_t early;          // BiCGStab: converged on ||s|| this iteration (x += alpha phat pending);
                            // MINRES: the w / x update of this iteration is pending; textbook
                            // PCG with the deferred x update: x += alpha p pending
    int32_t pend;           // textbook PCG with the folded update stage: the update kernel's
                            // totals (pend_rz, pend_rr) await their commit
    int32_t early_prev;     // many-RHS PCG: x += alpha_prev p_prev pending (x is updated every
                            // second iteration)
    // MINRES (paso/src/MINRES.cpp:67-69): the last two Givens rotations, the
    // coefficients of the w recurrence, c * eta for the x update and the
    // scaling of the preconditioned residual norm
    double c, c_old, s, s_old, a1, a2, a3, ceta, norm_scal;
    double pend_rz, pend_rr, pad2;
};
enum { KS_RUNNING = 0, KS_CONVERGED = 1, KS_MAXITER = 2, KS_BREAKDOWN = 3, KS_STAGNATED = 4, KS_NEGATIVE = 5 };

}  // namespace rg

struct rg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 256;
    double* partials = nullptr;     // 16 * kMaxGrid doubles of per-block partial sums
    double* pinned = nullptr;       // host-pinned staging for scalars (256 doubles)
    hipEvent_t ev = nullptr;
    void* scratch = nullptr;        // device address of this device's scratch array (scratch_address)
};

namespace rg {
struct rg_mg_state;
}

// Work vectors and cached launch graph of the Krylov solvers of one matrix.
struct rg_solver_work {
    int64_t n = 0;   // N*bs
    double *r = nullptr, *xw = nullptr, *rs = nullptr, *p = nullptr, *v = nullptr, *x2 = nullptr;
    double *z = nullptr, *z2 = nullptr;
    double* s = nullptr;            // BiCGStab's s
    double *pb1 = nullptr, *pb2 = nullptr;   // Gauss-Seidel sweeps > 1 (allocated on first use)
    double* bb = nullptr;           // balanced right-hand side D b
    double* norms = nullptr;        // [0] sum of squares, [1] max abs
    rg::KrylovState* st = nullptr;
    int* flag = nullptr;
    hipStream_t capture = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_method = -1, graph_sweeps = -1, graph_iters = -1;
};

// PASO's colouring of the matrix graph (Pattern::borrowColoringPointer) and
// the colour-ordered matrix copy of the Gauss-Seidel / ILU0 preconditioners.
// Rows are stored in colour order: pos -> row is perm, row -> pos is ipos,
// and colour c owns positions [cstart[c], cstart[c+1]).
struct rg_colouring {
    int ncolors = 0;
    std::vector<int64_t> cstart;
    int32_t* perm = nullptr;
    int32_t* ipos = nullptr;
    uint8_t* colour = nullptr;
};

struct rg_matrix {
    rg_context* ctx = nullptr;
    rg::Geom g;
    int flags = 0;
    bool symmetric = false;
    double* val = nullptr;         // nstore*bs*bs*ld
    float* val32 = nullptr;        // FP32 copy for the mixed-precision solver
    bool val32_valid = false;
    double* dinv = nullptr;        // bs*bs*ld Jacobi inverse (of the balanced matrix)
    float* dinv32 = nullptr;
    double* balance = nullptr;     // N*bs balance vector
    bool is_balanced = false;
    bool have_prec = false;
    bool val_zero = true;          // every stored value is known to be 0 (assembly may write instead of add)
    rg_solver_work* work = nullptr;
    int multi_shape = 0;           // SpMM tile shape + 1 of the many-RHS solver (0: autotune)
    rg_lpc* lpc = nullptr;         // learned preconditioner (borrowed)
    const float* lpc_logk = nullptr;   // log10(kappa) per node (device, borrowed)
    rg_colouring* colouring = nullptr; // built on first use (pattern only)
    double* pval = nullptr;        // colour-ordered full stencil, ILU0 elimination workspace, nslot*bs*bs*ld
    double* plow = nullptr;        // colour-ordered couplings to lower colours, nslot*bs*bs*ld
    double* pup = nullptr;         // colour-ordered couplings to higher colours, nslot*bs*bs*ld
    double* pdiag = nullptr;       // colour-ordered main blocks (Gauss-Seidel), bs*bs*ld
    double* pdinv = nullptr;       // colour-ordered inverse main blocks / inverted ILU pivots, bs*bs*ld
    int pval_state = 0;            // 0: invalid, 1: Gauss-Seidel split, 2: ILU0 split
    void* stage = nullptr;         // device staging of host element blocks (grown on demand)
    size_t stage_bytes = 0;
    // multilevel preconditioner of RG_METHOD_FCG_LEARNED (rg_mg.hip)
    rg::rg_mg_state* mg = nullptr;
    int mg_levels = 0;             // levels without a network (0: down to at most 9^3 nodes)
    int mg_degree = 2;             // Chebyshev steps before and after each coarse correction
    int mg_coarse_steps = 8;       // Chebyshev steps on the coarsest level without a network
};

namespace rg {

// ------------------------------------------------------------- kernels API
// All launches go to m->ctx->stream.
void launch_set_values(rg_matrix* m, double value);
void launch_add_element_blocks(rg_matrix* m, int64_t ne, int npe, const int32_t* d_nodes,
                               const double* d_em);
// Device address of the calling thread's current device's scratch array
// (kScratchBytes); looked up once per device with hipGetSymbolAddress.
void* scratch_address();
void launch_assemble(rg_matrix* m, const rg_pde_coefs& c, int reduced, double* rhs);
void launch_assemble_boundary(rg_matrix* m, const rg_coef* d, const rg_coef* y, int reduced, double* rhs);
void launch_add_dirac(rg_matrix* m, int64_t count, const int64_t* d_nodes, const double* d_d, const double* d_y,
                      double* rhs);
void launch_spmv(rg_matrix* m, double alpha, const double* x, double beta, double* y);
void launch_spmv_on(rg_matrix* m, double alpha, const double* x, double beta, double* y, hipStream_t st);
void launch_spmv32(rg_matrix* m, const double* x, double* y);   // FP32 values, y = A32 x
void launch_nullify(rg_matrix* m, const double* row_mask, const double* col_mask, double mdv,
                    int* d_flag);
void launch_export_csr_values(rg_matrix* m, const int32_t* d_ptr, int32_t* d_index, double* d_val);
void launch_to_fp32(rg_matrix* m);

// Multi-colour preconditioners (rg_precond.hip).  prec_setup builds what
// preconditioner `type` (RG_PRECONDITIONER_*) needs from the balanced matrix;
// prec_apply computes x = M^-1 b (b and x distinct device vectors of N*bs),
// skipping all work once guard->done is set (guard may be null).
void ensure_colouring(rg_matrix* m);
void free_colouring(rg_matrix* m);
void prec_setup(rg_matrix* m, int type);
void prec_apply(rg_matrix* m, int type, int sweeps, const double* b, double* x, hipStream_t st,
                const KrylovState* guard);
void solve_preconditioner(rg_matrix* m, const rg_options& opt, double* x, const double* b);

// Binary grid I/O (rg_io.hip, host only).
void read_binary_grid(const char* filename, int dim, const int64_t* grid, int num_comp, int dpp,
                      const int64_t* first, const int64_t* num_values, const int32_t* multiplier,
                      const int32_t* reverse, int byte_order, int data_type, double* out);
void write_binary_grid(const char* filename, int dim, const int64_t* grid, const double* in, int byte_order,
                       int data_type);

// Learned preconditioner (rg_learned.hip): node grid of a network and its
// forward pass on the balanced solver vectors, z = D^-1 Net(D dinv r).
void lpc_nodes(const rg_lpc* net, int32_t nodes[3]);
void lpc_forward_balanced(rg_lpc* net, const double* r, const double* dinv, const double* bal, const float* logk,
                          double* z, hipStream_t st);

// Host-side pattern helpers.
void csr_row_ptr(const Geom& g, int32_t* ptr);   // ptr[N+1] on host
int64_t csr_nnz(const Geom& g);

}  // namespace rg
