// Multilevel preconditioner and truncated flexible CG of
// RG_METHOD_FCG_LEARNED (rg_mg.hip).
#pragma once

#include "rg_internal.h"

namespace rg {

constexpr int kMaxLevels = 8;     // grid levels of a hierarchy
constexpr int kChebSteps = 16;    // Chebyshev steps a level's table holds
constexpr double kChebRatio = 10.0;   // smoothing interval [lmax / ratio, lmax]
constexpr int kDirs = 5;          // directions the flexible CG keeps

// (Re)builds the hierarchy of the matrix m (after the FP32 copy of its
// values): of the solver's balanced matrix, or of an unbalanced one.
void mg_setup(rg_matrix* m, bool balanced = true);
// z = M(r), one V-cycle on the balanced level-0 vectors; kernels skip their
// work once guard->done is set (guard may be null).
void mg_vcycle(rg_matrix* m, const double* r, double* z, hipStream_t st, const KrylovState* guard);
// One iteration of the flexible CG on x, r with the V-cycle (z: work vector).
// norm_w (nullptr: none): the matrix's balance vector; the residual norm the
// stopping test reads is then taken of r / balance (through 1 / balance kept
// by mg_setup)
void mg_fcg_iteration(rg_matrix* m, double* x, double* r, double* z, KrylovState* st_dev, hipStream_t st,
                      const double* norm_w = nullptr);
void mg_free(rg_matrix* m);
// The Galerkin operator `levels` halvings below the unbalanced matrix m
// (a new matrix owned by the caller).
rg_matrix* coarse_operator(rg_matrix* m, int levels);
// x = `steps` Chebyshev steps from zero on the unbalanced matrix m, as the
// hierarchy smooths a level.
void smooth_unbalanced(rg_matrix* m, const double* b, double* x, int steps);
// x = one V-cycle of the hierarchy of the unbalanced matrix m for A x = b.
void vcycle_unbalanced(rg_matrix* m, const double* b, double* x);

}  // namespace rg
