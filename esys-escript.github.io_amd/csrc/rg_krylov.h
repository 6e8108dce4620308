// Scalar recurrences of the Krylov solvers, shared by the single-system
// solver and the batched solvers: after each reducing kernel one workgroup
// reduces the block partials in a fixed order and applies stage_apply.
#pragma once

#include "rg_device.h"

namespace rg {

// End-of-iteration logic of the textbook PCG, shared by k_cg_update and the
// residual replacement.
__device__ __forceinline__ void cg_end(KrylovState* st, double rz, double rr, bool have_rz)
{
    st->norm = sqrt(rr);
    st->tau_new = rz;
    if (st->norm <= st->tol && !st->replace_mode) { st->done = 1; st->status = KS_CONVERGED; }
    else if (st->num_iter >= st->maxit) { st->done = 1; st->status = KS_MAXITER; }
    else if (have_rz) {
        if (fabs(rz) <= 0.0) { st->done = 1; st->status = KS_BREAKDOWN; }
        else { st->beta = rz / st->rho; st->rho_prev = st->rho; st->rho = rz; }
    }
}

// ------------------------------------------------------------ stage kernels
// One workgroup after each reducing kernel: reduces its partials in a fixed
// order and advances the Krylov scalars (the steps of PCG.cpp:156-347 that
// follow each reduction).
enum Stage {
    ST_PASO_INIT, ST_PASO_SPMV, ST_PASO_R, ST_PASO_X,
    ST_CG_INIT, ST_CG_SPMV, ST_CG_UPDATE, ST_CG_UPDATE_NORZ, ST_CG_RZ, ST_CG_REPLACE,
    ST_FCG_INIT, ST_FCG_BETA, ST_FCG_ALPHA, ST_FCG_UPDATE, ST_FCG_REPLACE,
    ST_BATCH_INIT,
    ST_BICG_RHO, ST_BICG_ALPHA, ST_BICG_S, ST_BICG_XA, ST_BICG_OMEGA, ST_BICG_END,
    ST_MR_INIT, ST_MR_DELTA, ST_MR_GAMMA,
    ST_CGX_SPMV, ST_CGX_UPDATE, ST_CG_COMMIT, ST_CGX2_SPMV
};

// Stages that start a solve run even when the state says done.
template <int STAGE>
__device__ __forceinline__ bool stage_starts()
{
    return STAGE == ST_PASO_INIT || STAGE == ST_CG_INIT || STAGE == ST_FCG_INIT || STAGE == ST_BATCH_INIT ||
           STAGE == ST_BICG_XA || STAGE == ST_MR_INIT || STAGE == ST_MR_DELTA || STAGE == ST_CGX_SPMV ||
           STAGE == ST_CG_COMMIT || STAGE == ST_CGX2_SPMV;
}

// The update stage of the textbook PCG when k_p_update<.., FOLD> has
// reduced the update kernel's partials itself: the iteration count, the
// stopping tests and rho advance here, from the same totals, at the next
// stage (so the state the update kernel's blocks read stays unchanged while
// they run).
__device__ __forceinline__ void cg_commit(KrylovState* st)
{
    if (!st->pend) return;
    st->pend = 0;
    st->num_iter += 1;
    cg_end(st, st->pend_rz, st->pend_rr, true);
}

// The scalar recurrence of one stage, given the reduced totals (one thread).
template <int STAGE, int NV>
__device__ __forceinline__ void stage_apply(KrylovState* st, const double (&tot)[NV])
{
    if constexpr (STAGE == ST_PASO_INIT) {
        st->tau = tot[0];
        st->beta = 0.0;
        st->num_iter = 1;
        st->done = 0;
        st->status = KS_RUNNING;
    } else if constexpr (STAGE == ST_PASO_SPMV) {
        st->delta = tot[0];
        st->alpha = st->tau / tot[0];
        if (fabs(tot[0]) <= 0.0) {   // PCG.cpp:262 breakFlag
            st->done = 1;
            st->status = KS_BREAKDOWN;
        }
    } else if constexpr (STAGE == ST_PASO_R) {
        st->gamma1 = (fabs(tot[0]) <= 0.0) ? 0.0 : -tot[1] / tot[0];
    } else if constexpr (STAGE == ST_PASO_X) {
        const double norm = sqrt(tot[0]);
        st->norm = norm;
        const bool conv = norm <= st->tol;
        const bool maxi = st->num_iter > st->maxit;
        const bool brk = fabs(st->tau) <= 0.0;
        if (conv || maxi || brk) {
            st->done = 1;
            st->status = maxi ? KS_MAXITER : (brk ? KS_BREAKDOWN : KS_CONVERGED);
        } else {
            st->beta = tot[1] / st->tau;
            st->tau = tot[1];
            st->num_iter += 1;
        }
    } else if constexpr (STAGE == ST_BATCH_INIT) {
        // textbook PCG from x = 0 (r = b): the tolerance arrives relative
        // and becomes absolute, tol * ||b||
        st->tol = st->tol * sqrt(tot[1]);
        st->rho = tot[0];
        st->rho_prev = tot[0];
        st->norm = sqrt(tot[1]);
        st->num_iter = 0;
        st->done = 0;
        st->status = KS_RUNNING;
        if (st->norm <= st->tol) { st->done = 1; st->status = KS_CONVERGED; }
        else if (st->maxit <= 0) { st->done = 1; st->status = KS_MAXITER; }
        else if (fabs(tot[0]) <= 0.0) { st->done = 1; st->status = KS_BREAKDOWN; }
    } else if constexpr (STAGE == ST_CG_INIT) {
        st->rho = tot[0];
        st->rho_prev = tot[0];
        st->norm = sqrt(tot[1]);
        st->num_iter = 0;
        st->done = 0;
        st->status = KS_RUNNING;
        if (st->norm <= st->tol) { st->done = 1; st->status = KS_CONVERGED; }
        else if (st->maxit <= 0) { st->done = 1; st->status = KS_MAXITER; }
        else if (fabs(tot[0]) <= 0.0) { st->done = 1; st->status = KS_BREAKDOWN; }
    } else if constexpr (STAGE == ST_CG_SPMV || STAGE == ST_CGX_SPMV || STAGE == ST_CGX2_SPMV) {
        // the deferred-x variant also retires the previous iteration's
        // pending x update, and so runs after the solve has ended too; the
        // paired variant (x updated every second iteration) keeps it pending
        // as the previous one instead
        if (STAGE == ST_CGX_SPMV) {
            st->early = 0;
            cg_commit(st);
        }
        if (STAGE == ST_CGX2_SPMV) {
            st->early_prev = st->early;
            st->alpha_prev = st->alpha;
            st->early = 0;
        }
        if (!st->done) {
            st->delta = tot[0];
            if (fabs(tot[0]) <= 0.0) { st->done = 1; st->status = KS_BREAKDOWN; }
            else st->alpha = st->rho / tot[0];
        }
    } else if constexpr (STAGE == ST_CG_UPDATE || STAGE == ST_CG_UPDATE_NORZ || STAGE == ST_CGX_UPDATE) {
        st->num_iter += 1;
        if (STAGE == ST_CGX_UPDATE) st->early = 1;   // x += alpha p pending in k_p_update
        cg_end(st, tot[0], tot[1], STAGE != ST_CG_UPDATE_NORZ);
    } else if constexpr (STAGE == ST_CG_RZ) {
        if (fabs(tot[0]) <= 0.0) { st->done = 1; st->status = KS_BREAKDOWN; }
        else { st->beta = tot[0] / st->rho; st->rho_prev = st->rho; st->rho = tot[0]; }
    } else if constexpr (STAGE == ST_CG_REPLACE) {
        // k_cg_update's stage set beta = rho_rec/rho_prev and rho = rho_rec;
        // redo that step with the replaced (true) residual
        st->norm = sqrt(tot[1]);
        if (st->norm <= st->tol) { st->done = 1; st->status = KS_CONVERGED; }
        else if (fabs(tot[0]) <= 0.0) { st->done = 1; st->status = KS_BREAKDOWN; }
        else { st->beta = tot[0] / st->rho_prev; st->rho = tot[0]; }
    } else if constexpr (STAGE == ST_CG_COMMIT) {
        cg_commit(st);
    } else if constexpr (STAGE == ST_BICG_RHO) {
        // Solver_BiCGStab (paso/src/BiCGStab.cpp:125-150)
        st->num_iter += 1;
        st->tau = tot[0];
        if (fabs(tot[0]) <= 0.0) { st->done = 1; st->status = KS_BREAKDOWN; }
        else st->beta = (st->num_iter > 1) ? tot[0] / st->tau_new * (st->alpha / st->gamma1) : 0.0;
    } else if constexpr (STAGE == ST_BICG_ALPHA) {
        if (fabs(tot[0]) <= 0.0) { st->done = 1; st->status = KS_BREAKDOWN; }   // :164
        else st->alpha = st->tau / tot[0];
    } else if constexpr (STAGE == ST_BICG_S) {
        st->norm = sqrt(tot[0]);                                                  // :176-186
        if (st->norm <= st->tol) { st->done = 1; st->status = KS_CONVERGED; st->early = 1; }
    } else if constexpr (STAGE == ST_BICG_XA) {
        st->early = 0;
    } else if constexpr (STAGE == ST_BICG_OMEGA) {
        if (fabs(tot[1]) <= 0.0) { st->done = 1; st->status = KS_BREAKDOWN; }    // :206
        else st->gamma1 = tot[0] / tot[1];
    } else if constexpr (STAGE == ST_BICG_END) {
        st->norm = sqrt(tot[0]);                                                  // :221-228
        const bool conv = st->norm <= st->tol;
        const bool maxi = st->num_iter > st->maxit;
        const bool brk = fabs(st->gamma1) <= 0.0;
        if (conv || maxi || brk) {
            // PASO ranks the flags as BiCGStab.cpp:236-240 does: a limit or a
            // breakdown reached in the same iteration as convergence wins
            st->done = 1;
            st->status = maxi ? KS_MAXITER : (brk ? KS_BREAKDOWN : KS_CONVERGED);
        } else {
            st->tau_new = st->tau;
        }
    } else if constexpr (STAGE == ST_MR_INIT) {
        // Solver_MINRES (paso/src/MINRES.cpp:77-101): tot = (r.z, r.r), z = M^-1 r
        const double dp = tot[0];
        st->num_iter = 0;
        st->done = 0;
        st->status = KS_RUNNING;
        st->early = 0;
        st->c = st->c_old = 1.0;
        st->s = st->s_old = 0.0;
        st->norm = 1.0;
        st->norm_scal = 1.0;
        if (dp < 0.0) { st->done = 1; st->status = KS_NEGATIVE; }
        else if (fabs(dp) <= 0.0) { st->done = 1; st->status = KS_CONVERGED; }
        else if (st->maxit <= 0) { st->done = 1; st->status = KS_MAXITER; }
        else {
            st->rho = sqrt(dp);            // gamma
            st->rho_prev = 0.0;            // gamma_old
            st->tau = st->rho;             // eta
            st->norm = st->rho;            // rnorm_prec
            st->norm_scal = st->rho / sqrt(tot[1]);
            st->tol = st->tol * st->norm_scal;
        }
    } else if constexpr (STAGE == ST_MR_DELTA) {
        // delta = Az.z (:106-109); also retires the previous iteration's
        // pending w / x update
        st->early = 0;
        if (!st->done) st->delta = tot[0];
    } else if constexpr (STAGE == ST_MR_GAMMA) {
        // :124-165 after z = M^-1 r_new: tot[0] = r_new.z
        const double dp = tot[0];
        if (dp < 0.0) { st->done = 1; st->status = KS_NEGATIVE; }
        else if (fabs(dp) <= 0.0) { st->done = 1; st->status = KS_CONVERGED; }
        else {
            const double g_old = st->rho, g = sqrt(dp), delta = st->delta;
            const double c_anc = st->c_old, c_old = st->c, s_anc = st->s_old, s_old = st->s;
            const double a0 = c_old * delta - c_anc * s_old * g_old;
            const double a1 = sqrt(a0 * a0 + g * g);
            st->a1 = a1;
            st->a2 = s_old * delta + c_anc * c_old * g_old;
            st->a3 = s_anc * g_old;
            st->c_old = c_old;
            st->s_old = s_old;
            st->c = a0 / a1;
            st->s = g / a1;
            st->rho_prev = g_old;
            st->rho = g;
            st->norm = st->norm * st->s;
            st->ceta = st->c * st->tau;
            st->tau = -st->s * st->tau;
            st->early = 1;
            if (st->norm <= st->tol) { st->done = 1; st->status = KS_CONVERGED; }
        }
        st->num_iter += 1;
        if (!st->done && st->num_iter >= st->maxit) { st->done = 1; st->status = KS_MAXITER; }
    } else if constexpr (STAGE == ST_FCG_INIT) {
        st->norm = sqrt(tot[1]);
        st->num_iter = 0;
        st->done = 0;
        st->status = KS_RUNNING;
        st->delta = 1.0;
        if (st->norm <= st->tol) { st->done = 1; st->status = KS_CONVERGED; }
        else if (st->maxit <= 0) { st->done = 1; st->status = KS_MAXITER; }
    } else if constexpr (STAGE == ST_FCG_BETA) {
        st->beta = (st->num_iter == 0) ? 0.0 : -tot[0] / st->delta;
    } else if constexpr (STAGE == ST_FCG_ALPHA) {
        if (fabs(tot[0]) <= 0.0) { st->done = 1; st->status = KS_BREAKDOWN; }
        else { st->alpha = tot[1] / tot[0]; st->delta = tot[0]; }
    } else if constexpr (STAGE == ST_FCG_UPDATE) {
        st->num_iter += 1;
        st->norm = sqrt(tot[0]);
        if (st->norm <= st->tol && !st->replace_mode) { st->done = 1; st->status = KS_CONVERGED; }
        else if (st->num_iter >= st->maxit) { st->done = 1; st->status = KS_MAXITER; }
    } else if constexpr (STAGE == ST_FCG_REPLACE) {
        st->norm = sqrt(tot[1]);
        if (st->norm <= st->tol) { st->done = 1; st->status = KS_CONVERGED; }
    }
}

}  // namespace rg
