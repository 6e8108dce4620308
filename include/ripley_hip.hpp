// ripley_hip.hpp -- header-only C++ host class over the C ABI of
// ripley_hip.h, with the surface of escript::AbstractSystemMatrix /
// paso::SystemMatrix<double> (escriptcore/src/AbstractSystemMatrix.h:43-190,
// paso/src/SystemMatrix.h:40-340) for a ripley Rectangle/Brick at one rank.
//
// Vectors handed to the methods below are host arrays (escript Data buffers:
// DOF-major, then component); the *_device methods take HBM pointers.
// Failures throw ripley_hip::Exception carrying the C status and message
// (PasoException for solver results 1..6, as paso/src/solve.cpp:119-143).
#ifndef RIPLEY_HIP_HPP
#define RIPLEY_HIP_HPP

#include <array>
#include <memory>
#include <stdexcept>
#include <string>

#include "ripley_hip.h"

namespace ripley_hip {

class Exception : public std::runtime_error {
public:
    Exception(int code, const std::string& msg) : std::runtime_error(msg), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

class PasoException : public Exception {
public:
    using Exception::Exception;
};

inline void check(int rc)
{
    if (rc == RG_OK) return;
    const std::string msg = rg_last_error();
    if (rc <= RG_DIVERGENCE) throw PasoException(rc, msg);
    throw Exception(rc, msg);
}

// One device + stream (escript's JMPI/OpenMP state has no GPU counterpart).
class Context {
public:
    explicit Context(int device = 0, void* stream = nullptr)
    {
        rg_context* c = nullptr;
        check(rg_context_create(device, stream, &c));
        ctx_.reset(c);
    }
    rg_context* get() const { return ctx_.get(); }
    void setStream(void* stream) { check(rg_context_set_stream(ctx_.get(), stream)); }
    void synchronize() { check(rg_context_synchronize(ctx_.get())); }

private:
    struct Del {
        void operator()(rg_context* c) const { rg_context_destroy(c); }
    };
    std::unique_ptr<rg_context, Del> ctx_;
};

// Solver options: the SolverBuddy fields PASO reads (paso/src/Options.cpp:31-56).
struct SolverOptions : rg_options {
    SolverOptions() { rg_options_default(this); }
};

using Diagnostics = rg_diagnostics;

class SystemMatrix {
public:
    // RipleyDomain::newSystemMatrix for Rectangle (dim 2) / Brick (dim 3).
    SystemMatrix(const Context& ctx, int dim, std::array<int32_t, 3> ne, std::array<double, 3> length,
                 int block_size = 1, bool symmetric = false)
    {
        rg_matrix* m = nullptr;
        check(rg_matrix_create(ctx.get(), dim, ne.data(), length.data(), block_size,
                               symmetric ? RG_MATRIX_SYMMETRIC : 0, &m));
        m_.reset(m);
        int64_t nr = 0, nnz = 0;
        int ns = 0, bs = 0;
        check(rg_matrix_get_info(m, &nr, &nnz, &ns, &bs));
        numRows_ = nr;
        nnz_ = nnz;
        blockSize_ = bs;
    }

    int getRowBlockSize() const { return blockSize_; }
    int getColumnBlockSize() const { return blockSize_; }
    int64_t getTotalNumRows() const { return numRows_ * blockSize_; }
    int64_t getNumBlockNonZeros() const { return nnz_; }
    rg_matrix* handle() const { return m_.get(); }

    // paso::SystemMatrix::setValues / AbstractSystemMatrix::resetValues
    void setValues(double value) { check(rg_matrix_set_values(m_.get(), value)); }
    void resetValues(bool preserveSolverData = false)
    {
        check(rg_matrix_reset_values(m_.get(), preserveSolverData ? 1 : 0));
    }

    // RipleyDomain::addToSystemMatrix for a batch of elements (host arrays).
    void add(int64_t numElements, int nodesPerElement, const int32_t* nodes, const double* em)
    {
        check(rg_matrix_add_element_blocks(m_.get(), numElements, nodesPerElement, nodes, blockSize_, em));
    }

    // assemblePDESingle / assemblePDESystem (+ reduced) with device coefficients.
    void assemble(const rg_pde_coefs& coefs, bool reduced, double* rhs_device)
    {
        check(rg_matrix_assemble(m_.get(), &coefs, reduced ? 1 : 0, rhs_device));
    }

    // AbstractSystemMatrix::ypAx: y += A x (host arrays).
    void ypAx(double* y, const double* x) { check(rg_matrix_ypAx_host(m_.get(), y, x)); }
    // paso::SystemMatrix::MatrixVector on device arrays: y = alpha A x + beta y.
    void MatrixVector_device(double alpha, const double* x, double beta, double* y)
    {
        check(rg_matrix_vector(m_.get(), alpha, x, beta, y));
    }

    // AbstractSystemMatrix::nullifyRowsAndCols (host masks).
    void nullifyRowsAndCols(const double* row_q, const double* col_q, double mdv)
    {
        check(rg_matrix_nullify_rows_and_cols_host(m_.get(), row_q, col_q, mdv));
    }

    // AbstractSystemMatrix::saveMM
    void saveMM(const std::string& filename) { check(rg_matrix_save_mm(m_.get(), filename.c_str())); }

    // paso::SystemMatrix::setToSolution (host arrays); diagnostics as
    // updateEscriptDiagnostics reports them.
    Diagnostics setToSolution(double* out, const double* in, const SolverOptions& options)
    {
        Diagnostics d{};
        check(rg_solve_host(m_.get(), out, in, &options, &d));
        return d;
    }
    Diagnostics solve_device(double* out, const double* in, const SolverOptions& options)
    {
        Diagnostics d{};
        check(rg_solve(m_.get(), out, in, &options, &d));
        return d;
    }

private:
    struct Del {
        void operator()(rg_matrix* m) const { rg_matrix_destroy(m); }
    };
    std::unique_ptr<rg_matrix, Del> m_;
    int64_t numRows_ = 0, nnz_ = 0;
    int blockSize_ = 1;
};

}  // namespace ripley_hip

#endif  // RIPLEY_HIP_HPP
