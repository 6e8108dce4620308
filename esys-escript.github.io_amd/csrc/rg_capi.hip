// The C ABI of libripley_hip (declared in include/ripley_hip.h): argument
// checks mirroring the reference's exceptions, ownership of device memory,
// and the host-pointer variants that stage escript Data buffers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <string>
#include <vector>

#include "rg_internal.h"
#include "rg_mg.h"

namespace rg {
int solve(rg_matrix* m, double* x, const double* b, const rg_options& opt, rg_diagnostics& dg);
void ensure_work(rg_matrix* m);
void free_work(rg_solver_work* w);
}  // namespace rg

namespace {

thread_local std::string g_error;

template <class F>
int guarded(F&& f)
{
    try {
        g_error.clear();
        return f();
    } catch (const rg::Error& e) {
        g_error = e.what();
        if (e.code == RG_DEVICE_ERROR) (void)hipGetLastError();   // clear the failed call's error state
        return e.code;
    } catch (const std::bad_alloc&) {
        g_error = "out of host memory";
        return RG_MEMORY_ERROR;
    } catch (const std::exception& e) {
        g_error = e.what();
        return RG_VALUE_ERROR;
    }
}

void use_device(const rg_context* ctx) { RG_HIP(hipSetDevice(ctx->device)); }

// Device staging of a host array for the *_host entry points.
struct DevBuf {
    double* p = nullptr;
    ~DevBuf()
    {
        if (p) hipFree(p);
    }
};

void check_matrix(const rg_matrix* m)
{
    RG_CHECK(m != nullptr, RG_VALUE_ERROR, "matrix handle is NULL");
}

size_t align256(size_t n) { return (n + 255) & ~(size_t)255; }

// Argument checks of addToSystemMatrix (RipleyDomain.cpp:1433-1459): block
// size, element size and node ids, before any array is touched.  Returns
// false when there is nothing to add; sets the staging sizes of the node ids
// and element matrices.
bool check_element_blocks(const rg_matrix* m, int64_t num_elements, int nodes_per_element, const int32_t* nodes,
                          int num_eq, const double* em, size_t* nsz, size_t* esz)
{
    check_matrix(m);
    RG_CHECK(num_eq == m->g.bs, RG_VALUE_ERROR, "addToSystemMatrix: number of equations does not match the block size");
    RG_CHECK(num_elements >= 0 && (num_elements == 0 || (nodes && em)), RG_VALUE_ERROR,
             "addToSystemMatrix: NULL element data");
    RG_CHECK(nodes_per_element == (1 << m->g.dim), RG_VALUE_ERROR,
             "add_element_blocks: nodes_per_element must be 4 (2-D) or 8 (3-D)");
    if (num_elements == 0) return false;
    for (int64_t k = 0; k < num_elements * nodes_per_element; ++k)
        RG_CHECK(nodes[k] >= 0 && nodes[k] < m->g.N, RG_VALUE_ERROR, "addToSystemMatrix: node index out of range");
    *nsz = sizeof(int32_t) * (size_t)num_elements * nodes_per_element;
    *esz = sizeof(double) * (size_t)num_elements * nodes_per_element * nodes_per_element * num_eq * num_eq;
    return true;
}

// Stages the node ids at buf and the element matrices at buf + off (device)
// and launches the colour passes.  The copies read pageable host memory, so
// they return once the host arrays are staged by the runtime; no stream
// synchronisation follows, and the next call's copies are ordered after
// these kernels on the same stream.
void add_element_blocks_via(rg_matrix* m, unsigned char* buf, size_t off, int64_t num_elements, int npe,
                            const int32_t* nodes, size_t nsz, const double* em, size_t esz)
{
    RG_HIP(hipMemcpyAsync(buf, nodes, nsz, hipMemcpyHostToDevice, m->ctx->stream));
    RG_HIP(hipMemcpyAsync(buf + off, em, esz, hipMemcpyHostToDevice, m->ctx->stream));
    rg::launch_add_element_blocks(m, num_elements, npe, reinterpret_cast<const int32_t*>(buf),
                                  reinterpret_cast<const double*>(buf + off));
    m->val32_valid = false;
}

}  // namespace

namespace rg {
void set_last_error(const std::string& msg) { g_error = msg; }
}  // namespace rg

extern "C" {

const char* rg_last_error(void) { return g_error.c_str(); }

const char* rg_version(void) { return "ripley_hip 1.0 (gfx950)"; }

int rg_context_create(int device, void* stream, rg_context** out)
{
    return guarded([&] {
        RG_CHECK(out != nullptr, RG_VALUE_ERROR, "rg_context_create: out is NULL");
        int ndev = 0;
        RG_HIP(hipGetDeviceCount(&ndev));
        RG_CHECK(device >= 0 && device < ndev, RG_VALUE_ERROR, "rg_context_create: no such device");
        RG_HIP(hipSetDevice(device));
        rg_context* c = new rg_context();
        c->device = device;
        c->stream = reinterpret_cast<hipStream_t>(stream);
        hipDeviceProp_t prop;
        RG_HIP(hipGetDeviceProperties(&prop, device));
        c->num_cu = prop.multiProcessorCount;
        try {
            c->scratch = rg::scratch_address();   // this device's copy of the scratch array
            RG_HIP(hipMalloc(&c->partials, sizeof(double) * 16 * rg::kMaxGrid));
            RG_HIP(hipHostMalloc(&c->pinned, sizeof(double) * 256, hipHostMallocDefault));
        } catch (...) {
            if (c->partials) hipFree(c->partials);
            delete c;
            throw;
        }
        *out = c;
        return RG_OK;
    });
}

int rg_context_destroy(rg_context* ctx)
{
    return guarded([&] {
        if (!ctx) return RG_OK;
        hipSetDevice(ctx->device);
        if (ctx->stream) hipStreamSynchronize(ctx->stream);
        else hipDeviceSynchronize();
        if (ctx->partials) hipFree(ctx->partials);
        if (ctx->pinned) hipHostFree(ctx->pinned);
        delete ctx;
        return RG_OK;
    });
}

int rg_context_set_stream(rg_context* ctx, void* stream)
{
    return guarded([&] {
        RG_CHECK(ctx != nullptr, RG_VALUE_ERROR, "context is NULL");
        ctx->stream = reinterpret_cast<hipStream_t>(stream);
        return RG_OK;
    });
}

int rg_context_synchronize(rg_context* ctx)
{
    return guarded([&] {
        RG_CHECK(ctx != nullptr, RG_VALUE_ERROR, "context is NULL");
        use_device(ctx);
        RG_HIP(hipStreamSynchronize(ctx->stream));
        return RG_OK;
    });
}

int rg_matrix_create(rg_context* ctx, int dim, const int32_t* ne, const double* length, int block_size,
                     int flags, rg_matrix** out)
{
    return guarded([&] {
        RG_CHECK(ctx && ne && length && out, RG_VALUE_ERROR, "rg_matrix_create: NULL argument");
        RG_CHECK(dim == 2 || dim == 3, RG_VALUE_ERROR, "rg_matrix_create: dim must be 2 or 3");
        RG_CHECK(block_size >= 1 && block_size <= 4, RG_VALUE_ERROR,
                 "rg_matrix_create: block size must be 1..4");
        use_device(ctx);
        rg::Geom g;
        g.dim = dim;
        g.bs = block_size;
        for (int d = 0; d < dim; ++d) {
            RG_CHECK(ne[d] >= 1, RG_VALUE_ERROR, "rg_matrix_create: need at least one element per axis");
            RG_CHECK(length[d] > 0, RG_VALUE_ERROR, "rg_matrix_create: side lengths must be positive");
            g.ne[d] = ne[d];
            g.nn[d] = (int64_t)ne[d] + 1;
            g.len[d] = length[d];
            g.h[d] = length[d] / ne[d];
        }
        g.N = g.nn[0] * g.nn[1] * g.nn[2];
        RG_CHECK(g.N < (int64_t)1 << 31, RG_VALUE_ERROR, "rg_matrix_create: more than 2^31 nodes");
        g.s1 = g.nn[0];
        g.s2 = g.nn[0] * g.nn[1];
        g.nslot = (dim == 3) ? 27 : 9;
        const bool sym = (flags & RG_MATRIX_SYMMETRIC) != 0;
        g.nstore = sym ? g.nslot / 2 + 1 : g.nslot;
        g.first = sym ? g.nslot / 2 : 0;
        g.ld = ((g.N + 63) / 64) * 64;
        g.nnz = rg::csr_nnz(g);
        RG_CHECK(g.nnz < (int64_t)1 << 31, RG_VALUE_ERROR, "rg_matrix_create: PASO pattern exceeds int32 indices");
        rg_matrix* m = new rg_matrix();
        m->ctx = ctx;
        m->g = g;
        m->flags = flags;
        m->symmetric = sym;
        try {
            const size_t nval = value_count(g);
            RG_HIP(hipMalloc(&m->val, nval * sizeof(double)));
            RG_HIP(hipMemsetAsync(m->val, 0, nval * sizeof(double), ctx->stream));
            RG_HIP(hipMalloc(&m->dinv, (size_t)g.bs * g.bs * g.ld * sizeof(double)));
            RG_HIP(hipMalloc(&m->dinv32, (size_t)g.bs * g.bs * g.ld * sizeof(float)));
            RG_HIP(hipMalloc(&m->balance, (size_t)g.N * g.bs * sizeof(double)));
            RG_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            if (m->val) hipFree(m->val);
            if (m->dinv) hipFree(m->dinv);
            if (m->dinv32) hipFree(m->dinv32);
            if (m->balance) hipFree(m->balance);
            delete m;
            throw;
        }
        *out = m;
        return RG_OK;
    });
}

int rg_matrix_destroy(rg_matrix* m)
{
    return guarded([&] {
        if (!m) return RG_OK;
        hipSetDevice(m->ctx->device);
        hipStreamSynchronize(m->ctx->stream);
        rg::mg_free(m);
        rg::free_work(m->work);
        rg::free_colouring(m);
        void* bufs[] = {m->val, m->val32, m->dinv, m->dinv32, m->balance, m->stage};
        for (void* b : bufs)
            if (b) hipFree(b);
        delete m;
        return RG_OK;
    });
}

int rg_matrix_get_info(const rg_matrix* m, int64_t* num_rows, int64_t* nnz, int* num_slots, int* block_size)
{
    return guarded([&] {
        check_matrix(m);
        if (num_rows) *num_rows = m->g.N;
        if (nnz) *nnz = m->g.nnz;
        if (num_slots) *num_slots = m->g.nstore;
        if (block_size) *block_size = m->g.bs;
        return RG_OK;
    });
}

int rg_matrix_set_values(rg_matrix* m, double value)
{
    return guarded([&] {
        check_matrix(m);
        use_device(m->ctx);
        rg::launch_set_values(m, value);
        m->is_balanced = false;
        m->val32_valid = false;
        return RG_OK;
    });
}

int rg_matrix_reset_values(rg_matrix* m, int preserve_solver_data)
{
    return guarded([&] {
        check_matrix(m);
        use_device(m->ctx);
        rg::launch_set_values(m, 0.0);
        m->is_balanced = false;
        m->val32_valid = false;
        if (!preserve_solver_data) {
            m->have_prec = false;
            m->pval_state = 0;
        }
        return RG_OK;
    });
}

int rg_matrix_add_element_blocks(rg_matrix* m, int64_t num_elements, int nodes_per_element,
                                 const int32_t* nodes, int num_eq, const double* em)
{
    return guarded([&] {
        size_t nsz = 0, esz = 0;
        if (!check_element_blocks(m, num_elements, nodes_per_element, nodes, num_eq, em, &nsz, &esz)) return RG_OK;
        use_device(m->ctx);
        // persistent staging: grown (never per call) when a larger batch arrives
        const size_t off = align256(nsz);
        if (off + esz > m->stage_bytes) {
            if (m->stage) {
                RG_HIP(hipStreamSynchronize(m->ctx->stream));   // the old buffer may still be read
                RG_HIP(hipFree(m->stage));
                m->stage = nullptr;
                m->stage_bytes = 0;
            }
            const size_t want = std::max<size_t>(off + esz, (size_t)1 << 20);
            RG_HIP(hipMalloc(&m->stage, want));
            m->stage_bytes = want;
        }
        add_element_blocks_via(m, static_cast<unsigned char*>(m->stage), off, num_elements, nodes_per_element,
                               nodes, nsz, em, esz);
        return RG_OK;
    });
}

int rg_matrix_add_element_blocks_noalloc(rg_matrix* m, int64_t num_elements, int nodes_per_element,
                                         const int32_t* nodes, int num_eq, const double* em)
{
    return guarded([&] {
        size_t nsz = 0, esz = 0;
        if (!check_element_blocks(m, num_elements, nodes_per_element, nodes, num_eq, em, &nsz, &esz)) return RG_OK;
        const size_t off = align256(nsz);
        RG_CHECK(off + esz <= (size_t)rg::kScratchBytes, RG_VALUE_ERROR,
                 "add_element_blocks_noalloc: " + std::to_string(num_elements) + " elements need " +
                     std::to_string(off + esz) + " bytes of staging, more than the " +
                     std::to_string(rg::kScratchBytes) + "-byte scratch array; pass fewer elements per call");
        RG_CHECK(m->ctx->scratch != nullptr, RG_STATE_ERROR, "add_element_blocks_noalloc: no scratch address");
        use_device(m->ctx);
        add_element_blocks_via(m, static_cast<unsigned char*>(m->ctx->scratch), off, num_elements,
                               nodes_per_element, nodes, nsz, em, esz);
        return RG_OK;
    });
}

int rg_matrix_assemble(rg_matrix* m, const rg_pde_coefs* coefs, int reduced, double* rhs)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(coefs != nullptr, RG_VALUE_ERROR, "assemble: coefficients are NULL");
        use_device(m->ctx);
        rg::launch_assemble(m, *coefs, reduced, rhs);
        return RG_OK;
    });
}

int rg_matrix_assemble_boundary(rg_matrix* m, const rg_coef* d, const rg_coef* y, int reduced, double* rhs)
{
    return guarded([&] {
        check_matrix(m);
        use_device(m->ctx);
        rg::launch_assemble_boundary(m, d, y, reduced, rhs);
        return RG_OK;
    });
}

int rg_matrix_add_dirac(rg_matrix* m, int64_t count, const int64_t* nodes, const double* d, const double* y,
                        double* rhs)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(count >= 0 && (count == 0 || nodes), RG_VALUE_ERROR, "add_dirac: NULL node list");
        RG_CHECK(!(y && !rhs), RG_VALUE_ERROR, "add_dirac: y given without a right-hand side");
        if (count == 0 || (!d && !y)) return RG_OK;
        const int bs = m->g.bs;
        for (int64_t p = 0; p < count; ++p)
            RG_CHECK(nodes[p] >= 0 && nodes[p] < m->g.N, RG_VALUE_ERROR, "add_dirac: node index out of range");
        use_device(m->ctx);
        hipStream_t st = m->ctx->stream;
        struct Buf {
            void* p = nullptr;
            ~Buf() { if (p) hipFree(p); }
        } dn, dd, dy;
        RG_HIP(hipMalloc(&dn.p, sizeof(int64_t) * count));
        RG_HIP(hipMemcpyAsync(dn.p, nodes, sizeof(int64_t) * count, hipMemcpyHostToDevice, st));
        if (d) {
            RG_HIP(hipMalloc(&dd.p, sizeof(double) * count * bs * bs));
            RG_HIP(hipMemcpyAsync(dd.p, d, sizeof(double) * count * bs * bs, hipMemcpyHostToDevice, st));
        }
        if (y) {
            RG_HIP(hipMalloc(&dy.p, sizeof(double) * count * bs));
            RG_HIP(hipMemcpyAsync(dy.p, y, sizeof(double) * count * bs, hipMemcpyHostToDevice, st));
        }
        rg::launch_add_dirac(m, count, (const int64_t*)dn.p, (const double*)dd.p, (const double*)dy.p, rhs);
        RG_HIP(hipStreamSynchronize(st));
        return RG_OK;
    });
}

int rg_matrix_vector(rg_matrix* m, double alpha, const double* x, double beta, double* y)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(!m->is_balanced, RG_STATE_ERROR, "MatrixVector: balanced matrix is not supported.");
        RG_CHECK(x && y, RG_VALUE_ERROR, "MatrixVector: NULL vector");
        use_device(m->ctx);
        rg::launch_spmv(m, alpha, x, beta, y);
        return RG_OK;
    });
}

int rg_matrix_ypAx(rg_matrix* m, double* y, const double* x) { return rg_matrix_vector(m, 1.0, x, 1.0, y); }

int rg_matrix_ypAx_host(rg_matrix* m, double* y, const double* x)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(!m->is_balanced, RG_STATE_ERROR, "MatrixVector: balanced matrix is not supported.");
        RG_CHECK(x && y, RG_VALUE_ERROR, "MatrixVector: NULL vector");
        use_device(m->ctx);
        const size_t n = (size_t)m->g.N * m->g.bs;
        DevBuf dx, dy;
        RG_HIP(hipMalloc(&dx.p, n * sizeof(double)));
        RG_HIP(hipMalloc(&dy.p, n * sizeof(double)));
        hipStream_t st = m->ctx->stream;
        RG_HIP(hipMemcpyAsync(dx.p, x, n * sizeof(double), hipMemcpyHostToDevice, st));
        RG_HIP(hipMemcpyAsync(dy.p, y, n * sizeof(double), hipMemcpyHostToDevice, st));
        rg::launch_spmv(m, 1.0, dx.p, 1.0, dy.p);
        RG_HIP(hipMemcpyAsync(y, dy.p, n * sizeof(double), hipMemcpyDeviceToHost, st));
        RG_HIP(hipStreamSynchronize(st));
        return RG_OK;
    });
}

int rg_matrix_nullify_rows_and_cols(rg_matrix* m, const double* row_mask, const double* col_mask, double mdv)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(row_mask && col_mask, RG_VALUE_ERROR, "nullifyRowsAndCols: NULL mask");
        use_device(m->ctx);
        rg::ensure_work(m);
        RG_HIP(hipMemsetAsync(m->work->flag, 0, sizeof(int), m->ctx->stream));
        rg::launch_nullify(m, row_mask, col_mask, mdv, m->work->flag);
        return RG_OK;
    });
}

int rg_matrix_nullify_rows_and_cols_host(rg_matrix* m, const double* row_mask, const double* col_mask, double mdv)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(row_mask && col_mask, RG_VALUE_ERROR, "nullifyRowsAndCols: NULL mask");
        use_device(m->ctx);
        const size_t n = (size_t)m->g.N * m->g.bs;
        DevBuf dr, dc;
        hipStream_t st = m->ctx->stream;
        RG_HIP(hipMalloc(&dr.p, n * sizeof(double)));
        RG_HIP(hipMemcpyAsync(dr.p, row_mask, n * sizeof(double), hipMemcpyHostToDevice, st));
        const double* cm = dr.p;
        if (col_mask != row_mask) {
            RG_HIP(hipMalloc(&dc.p, n * sizeof(double)));
            RG_HIP(hipMemcpyAsync(dc.p, col_mask, n * sizeof(double), hipMemcpyHostToDevice, st));
            cm = dc.p;
        }
        rg::ensure_work(m);
        RG_HIP(hipMemsetAsync(m->work->flag, 0, sizeof(int), st));
        rg::launch_nullify(m, dr.p, cm, mdv, m->work->flag);
        RG_HIP(hipStreamSynchronize(st));
        return RG_OK;
    });
}

int rg_matrix_export_csr(rg_matrix* m, int32_t* ptr, int32_t* index, double* val)
{
    return guarded([&] {
        check_matrix(m);
        use_device(m->ctx);
        const rg::Geom& g = m->g;
        std::vector<int32_t> hptr(g.N + 1);
        rg::csr_row_ptr(g, hptr.data());
        if (ptr) std::memcpy(ptr, hptr.data(), sizeof(int32_t) * (g.N + 1));
        if (!index && !val) return RG_OK;
        hipStream_t st = m->ctx->stream;
        void *dptr = nullptr, *didx = nullptr, *dval = nullptr;
        const size_t bs2 = (size_t)g.bs * g.bs;
        try {
            RG_HIP(hipMalloc(&dptr, sizeof(int32_t) * (g.N + 1)));
            if (index) RG_HIP(hipMalloc(&didx, sizeof(int32_t) * g.nnz));
            if (val) RG_HIP(hipMalloc(&dval, sizeof(double) * g.nnz * bs2));
            RG_HIP(hipMemcpyAsync(dptr, hptr.data(), sizeof(int32_t) * (g.N + 1), hipMemcpyHostToDevice, st));
            rg::launch_export_csr_values(m, (const int32_t*)dptr, (int32_t*)didx, (double*)dval);
            if (index) RG_HIP(hipMemcpyAsync(index, didx, sizeof(int32_t) * g.nnz, hipMemcpyDeviceToHost, st));
            if (val) RG_HIP(hipMemcpyAsync(val, dval, sizeof(double) * g.nnz * bs2, hipMemcpyDeviceToHost, st));
            RG_HIP(hipStreamSynchronize(st));
        } catch (...) {
            if (dptr) hipFree(dptr);
            if (didx) hipFree(didx);
            if (dval) hipFree(dval);
            throw;
        }
        hipFree(dptr);
        if (didx) hipFree(didx);
        if (dval) hipFree(dval);
        return RG_OK;
    });
}

int rg_matrix_save_mm(rg_matrix* m, const char* filename)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(filename != nullptr, RG_VALUE_ERROR, "SparseMatrix::saveMM: no file name");
        const rg::Geom& g = m->g;
        const size_t bs = g.bs, bs2 = bs * bs;
        std::vector<int32_t> ptr(g.N + 1), idx(g.nnz);
        std::vector<double> val(g.nnz * bs2);
        int rc = rg_matrix_export_csr(m, ptr.data(), idx.data(), val.data());
        if (rc) throw rg::Error(rc, g_error);
        std::ofstream f(filename);
        RG_CHECK(!f.fail(), RG_VALUE_ERROR, "SparseMatrix::saveMM: File could not be opened for writing");
        // layout of SparseMatrix<double>::saveMM (paso/src/SparseMatrix.cpp:182-240)
        f << "%%MatrixMarket matrix coordinate real general\n";
        f << g.N * bs << " " << g.N * bs << " " << (int64_t)g.nnz * bs2 << "\n";
        f.precision(15);
        for (int64_t i = 0; i < g.N; ++i)
            for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k)
                for (size_t r = 0; r < bs; ++r)
                    for (size_t c = 0; c < bs; ++c)
                        f << i * bs + r + 1 << " " << (int64_t)idx[k] * bs + c + 1 << " "
                          << val[(size_t)k * bs2 + r + bs * c] << "\n";
        return RG_OK;
    });
}

void rg_options_default(rg_options* opt)
{
    if (!opt) return;
    opt->tolerance = 1e-8;
    opt->inner_tolerance = 1e-4;
    opt->iter_max = 100000;
    opt->sweeps = 1;
    opt->method = RG_METHOD_PCG;
    opt->verbose = 0;
    opt->accept_failed_convergence = 0;
    opt->check_interval = 0;
    opt->criterion = RG_CRITERION_PASO;
    opt->preconditioner = RG_PRECONDITIONER_JACOBI;
}

// Error mapping of SystemMatrix<double>::solve (paso/src/solve.cpp:119-143).
static int map_solver_result(int res, const rg_options& opt)
{
    if (res == RG_OK) return RG_OK;
    if ((res == RG_DIVERGENCE || res == RG_MAX_ITER_REACHED) && opt.accept_failed_convergence) {
        if (opt.verbose) printf("paso: failed convergence error has been canceled as requested.\n");
        return RG_OK;
    }
    switch (res) {
        case RG_DIVERGENCE:
            g_error = "Solver: No improvement during iteration. Iterative solver gives up.";
            break;
        case RG_MAX_ITER_REACHED:
            g_error = "Solver: maximum number of iteration steps reached.\n"
                      "Returned solution does not fulfil stopping criterion.";
            break;
        case RG_INPUT_ERROR: g_error = "Solver: illegal dimension in iterative solver."; break;
        case RG_NEGATIVE_NORM: g_error = "Solver: negative energy norm (try other solver or preconditioner)."; break;
        case RG_BREAKDOWN: g_error = "Solver: fatal break down in iterative solver."; break;
        default: g_error = "Solver: Generic error in solver."; break;
    }
    return res;
}

int rg_solve(rg_matrix* m, double* x, const double* b, const rg_options* opt, rg_diagnostics* diag)
{
    rg_options o;
    rg_options_default(&o);
    if (opt) o = *opt;
    rg_diagnostics dg;
    std::memset(&dg, 0, sizeof(dg));
    int res = guarded([&] {
        check_matrix(m);
        RG_CHECK(x && b, RG_VALUE_ERROR, "solve: NULL vector");
        use_device(m->ctx);
        return rg::solve(m, x, b, o, dg);
    });
    if (diag) *diag = dg;
    if (res >= RG_VALUE_ERROR) return res;
    return map_solver_result(res, o);
}

int rg_matrix_set_learned_preconditioner(rg_matrix* m, rg_lpc* net, const float* logk)
{
    return guarded([&] {
        check_matrix(m);
        (void)logk;   // reserved: the network's coefficient input comes from its level's operator
        if (m->work) m->work->graph_method = -1;   // captured iterations refer to the old network
        if (!net) {
            m->lpc = nullptr;
            m->lpc_logk = nullptr;
            return RG_OK;
        }
        RG_CHECK(m->g.dim == 3 && m->g.bs == 1, RG_VALUE_ERROR,
                 "set_learned_preconditioner: scalar problems on a Brick only");
        int32_t nodes[3];
        rg::lpc_nodes(net, nodes);
        bool level = false;
        int ne[3] = {m->g.ne[0], m->g.ne[1], m->g.ne[2]};
        while (ne[0] % 2 == 0 && ne[1] % 2 == 0 && ne[2] % 2 == 0 && ne[0] >= 2 && ne[1] >= 2 && ne[2] >= 2) {
            for (int d = 0; d < 3; ++d) ne[d] /= 2;
            level = level || (nodes[0] == ne[0] + 1 && nodes[1] == ne[1] + 1 && nodes[2] == ne[2] + 1);
        }
        RG_CHECK(level, RG_VALUE_ERROR,
                 "set_learned_preconditioner: the network's node grid is not a coarse level of the matrix's grid "
                 "(every other node per level)");
        m->lpc = net;
        m->lpc_logk = nullptr;
        return RG_OK;
    });
}

int rg_matrix_coarse_operator(rg_matrix* m, int levels, rg_matrix** out)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(out != nullptr, RG_VALUE_ERROR, "coarse_operator: out is NULL");
        use_device(m->ctx);
        *out = rg::coarse_operator(m, levels);
        return RG_OK;
    });
}

int rg_matrix_vcycle(rg_matrix* m, const double* b, double* x)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(b && x && b != x, RG_VALUE_ERROR, "vcycle: b and x must be distinct device vectors");
        use_device(m->ctx);
        rg::vcycle_unbalanced(m, b, x);
        return RG_OK;
    });
}

int rg_matrix_smooth(rg_matrix* m, const double* b, double* x, int steps)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(b && x && b != x, RG_VALUE_ERROR, "smooth: b and x must be distinct device vectors");
        use_device(m->ctx);
        rg::smooth_unbalanced(m, b, x, steps);
        return RG_OK;
    });
}

int rg_matrix_set_multilevel(rg_matrix* m, int levels, int degree, int coarse_steps)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(levels >= 0 && levels <= rg::kMaxLevels, RG_VALUE_ERROR, "set_multilevel: levels must be 0..8");
        RG_CHECK(degree >= 1 && degree <= rg::kChebSteps, RG_VALUE_ERROR, "set_multilevel: degree must be 1..16");
        RG_CHECK(coarse_steps >= 1 && coarse_steps <= rg::kChebSteps, RG_VALUE_ERROR,
                 "set_multilevel: coarse_steps must be 1..16");
        m->mg_levels = levels;
        m->mg_degree = degree;
        m->mg_coarse_steps = coarse_steps;
        if (m->work) m->work->graph_method = -1;   // captured iterations hold the old cycle
        return RG_OK;
    });
}

int rg_matrix_solve_preconditioner(rg_matrix* m, const rg_options* opt, double* x, const double* b)
{
    return guarded([&] {
        check_matrix(m);
        RG_CHECK(x && b, RG_VALUE_ERROR, "solvePreconditioner: NULL vector");
        rg_options o;
        rg_options_default(&o);
        if (opt) o = *opt;
        use_device(m->ctx);
        rg::solve_preconditioner(m, o, x, b);
        return RG_OK;
    });
}

int rg_read_binary_grid(const char* filename, int dim, const int64_t* grid, int num_comp, int points_per_sample,
                        const int64_t* first, const int64_t* num_values, const int32_t* multiplier,
                        const int32_t* reverse, int byte_order, int data_type, double* out)
{
    return guarded([&] {
        rg::read_binary_grid(filename, dim, grid, num_comp, points_per_sample, first, num_values, multiplier,
                             reverse, byte_order, data_type, out);
        return RG_OK;
    });
}

int rg_write_binary_grid(const char* filename, int dim, const int64_t* grid, const double* in, int byte_order,
                         int data_type)
{
    return guarded([&] {
        rg::write_binary_grid(filename, dim, grid, in, byte_order, data_type);
        return RG_OK;
    });
}

int rg_matrix_coloring(rg_matrix* m, int32_t* colours, int* num_colors)
{
    return guarded([&] {
        check_matrix(m);
        use_device(m->ctx);
        rg::ensure_colouring(m);
        const rg_colouring* c = m->colouring;
        if (num_colors) *num_colors = c->ncolors;
        if (colours) {
            std::vector<uint8_t> c8(m->g.N);
            RG_HIP(hipMemcpy(c8.data(), c->colour, m->g.N, hipMemcpyDeviceToHost));
            for (int64_t i = 0; i < m->g.N; ++i) colours[i] = c8[i];
        }
        return RG_OK;
    });
}

int rg_solve_host(rg_matrix* m, double* x, const double* b, const rg_options* opt, rg_diagnostics* diag)
{
    DevBuf dx, db;
    int rc = guarded([&] {
        check_matrix(m);
        RG_CHECK(x && b, RG_VALUE_ERROR, "solve: NULL vector");
        use_device(m->ctx);
        const size_t n = (size_t)m->g.N * m->g.bs;
        RG_HIP(hipMalloc(&dx.p, n * sizeof(double)));
        RG_HIP(hipMalloc(&db.p, n * sizeof(double)));
        RG_HIP(hipMemcpyAsync(db.p, b, n * sizeof(double), hipMemcpyHostToDevice, m->ctx->stream));
        return RG_OK;
    });
    if (rc) return rc;
    const int res = rg_solve(m, dx.p, db.p, opt, diag);
    const std::string err = g_error;
    int rc2 = guarded([&] {
        const size_t n = (size_t)m->g.N * m->g.bs;
        RG_HIP(hipMemcpyAsync(x, dx.p, n * sizeof(double), hipMemcpyDeviceToHost, m->ctx->stream));
        RG_HIP(hipStreamSynchronize(m->ctx->stream));
        return RG_OK;
    });
    if (res) {
        g_error = err;
        return res;
    }
    return rc2;
}

}  // extern "C"
