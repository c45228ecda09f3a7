// heat.hip -- resident transient diffusion (backward Euler / alpha family).
//
// Reference: DiffusionModel::DoTransientNonLinearAnalysis
// (DiffusionModel.cpp:711-831) with assembleK_F_s_alpha_family (322-347):
// every step zeroes and reassembles K and M, forms Mbar = (a1 K + M)/dt and
// Fbar = -(a1 + a2) K u_s / dt (a1 = alpha dt, a2 = (1 - alpha) dt), solves
// Mbar du = Fbar + lift with the initial guess u_s (Equations::solve(u_s1),
// equation.cpp:1288-1405), and sets u_s += du.
//
// Device form for a linear conductivity: Mbar and K are assembled once
// (BHIP_ASM_HEAT_TRANSIENT keeps K in the second value array); each step
// builds its right-hand side from the FP64 K on the device -- one pass over
// the free rows of K plus the node table of free x restrained couplings --
// solves with the context's solver and updates u_s, without leaving the GPU.
#include "bhip_internal.hpp"

namespace {

constexpr int kThreads = 256;

// b_i = presc_i on a restrained DOF; otherwise
// fcoef * (K_ff u)_i + sum over restrained neighbours m of
// (fcoef K(i,m) u_m - Mbar(i,m) presc_m).
template <int LPR>
__global__ void __launch_bounds__(kThreads) k_heat_rhs(int n, const int *__restrict__ rowptr,
                                                     const int *__restrict__ colind,
                                                     const double *__restrict__ kval,
                                                     const unsigned char *__restrict__ rest,
                                                     const long long *__restrict__ lptr,
                                                     const int *__restrict__ lcol,
                                                     const double *__restrict__ lval,
                                                     const double *__restrict__ u,
                                                     const double *__restrict__ presc, double fcoef,
                                                     double *__restrict__ b) {
    const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    const long long row = (long long)blockIdx.x * (kThreads / LPR) + g;
    const bool live = row < n && !rest[row < n ? row : 0];
    double s = 0;
    if (live) {
        const int kb = rowptr[row], ke = rowptr[row + 1];
        for (int k = kb + l; k < ke; k += LPR) s += kval[k] * u[colind[k]];
    }
#pragma unroll
    for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
    if (row < n && l == 0) {
        if (!live) {
            b[row] = presc[row];
        } else {
            double acc = fcoef * s;
            for (long long k = lptr[row]; k < lptr[row + 1]; k++) {
                const int m = lcol[k];
                acc += fcoef * lval[2 * k] * u[m] - lval[2 * k + 1] * presc[m];
            }
            b[row] = acc;
        }
    }
}

// x0 = u_s on free DOFs, presc on restrained ones (Equations::solve's tmpGuess)
__global__ void k_guess_ref(long long n, const unsigned char *rest, const double *u, const double *presc, double *x0) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x0[i] = rest[i] ? presc[i] : u[i];
}

__global__ void k_add_inplace(long long n, double *u, const double *du) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) u[i] += du[i];
}

// ---- projected initial guess (BHIP_HEAT_X0_PROJECTED)
// Recent increments du are kept as a basis v_1..v_m, orthonormal in the
// Mbar inner product (v_i' Mbar v_j = delta_ij).  A step's guess is the
// Mbar-orthogonal projection of its solution on their span,
// x0 = sum_i (v_i' b) v_i, which needs only b; after the step the new
// increment is orthogonalised against the basis (Gram-Schmidt in the Mbar
// inner product, one SpMV z = Mbar du) and appended; a full basis restarts
// from the newest increment.  This is the successive right-hand sides
// projection of P. F. Fischer (Comput. Methods Appl. Mech. Engrg. 163,
// 1998).  (Keeping a sliding window instead, the new vector in the oldest
// one's slot, took 41% more iterations on the 128^3 bump: the oldest
// increment carries the slowest-decaying mode.)  The sums are fixed-order
// reductions (project.hip), so runs repeat bit for bit.

int heat_lpr(bhip_ctx *ctx) {
    const double avg = ctx->n_dof > 0 ? (double)ctx->nnz / (double)ctx->n_dof : 1.0;
    return avg > 12 ? 8 : avg > 4 ? 4 : 2;
}

// The same right-hand side from ku = K u computed on K's FP64 column runs.
__global__ void k_heat_rhs_runs(int n, const unsigned char *__restrict__ rest, const long long *__restrict__ lptr,
                                const int *__restrict__ lcol, const double *__restrict__ lval,
                                const double *__restrict__ u, const double *__restrict__ presc, double fcoef,
                                const double *__restrict__ ku, double *__restrict__ b) {
    const long long row = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    if (rest[row]) {
        b[row] = presc[row];
        return;
    }
    double acc = fcoef * ku[row];
    for (long long k = lptr[row]; k < lptr[row + 1]; k++) {
        const int m = lcol[k];
        acc += fcoef * lval[2 * k] * u[m] - lval[2 * k + 1] * presc[m];
    }
    b[row] = acc;
}

// kruns / ku: K's FP64 column runs and a vector for K u, or null for the CSR form.
void launch_heat_rhs(bhip_ctx *ctx, const double *u, double fcoef, const double *kruns, double *ku) {
    const int n = ctx->n_dof;
    if (kruns) {
        bhip_runs_spmv64(ctx, kruns, u, ku);
        k_heat_rhs_runs<<<grid_for(n, kThreads), kThreads, 0, ctx->stream>>>(
            n, ctx->rest.p, ctx->lift_ptr.p, ctx->lift_col.p, ctx->lift_val.p, u, ctx->presc.p, fcoef, ku, ctx->rhs.p);
        return;
    }
    const int lpr = heat_lpr(ctx);
    const int g = grid_for((long long)n * lpr, kThreads);
#define BHIP_HEAT_CASE(L)                                                                                  \
    case L:                                                                                                \
        k_heat_rhs<L><<<g, kThreads, 0, ctx->stream>>>(n, ctx->rowptr.p, ctx->colind.p, ctx->val_aux.p,     \
                                                       ctx->rest.p, ctx->lift_ptr.p, ctx->lift_col.p,       \
                                                       ctx->lift_val.p, u, ctx->presc.p, fcoef, ctx->rhs.p); \
        break;
    switch (lpr) {
        BHIP_HEAT_CASE(8)
        BHIP_HEAT_CASE(4)
        BHIP_HEAT_CASE(2)
    }
#undef BHIP_HEAT_CASE
}

}  // namespace

extern "C" int bhip_heat_run(bhip_ctx *ctx, int n_steps, double *u, int x0_mode, const bhip_solve_opts *opts,
                             bhip_heat_stats *stats) {
    if (!ctx) return BHIP_ERR_ARG;
    if (!u || !opts || n_steps < 0) return bhip_fail(ctx, BHIP_ERR_ARG, "u and opts are required, n_steps >= 0");
    if (x0_mode < BHIP_HEAT_X0_ZERO || x0_mode > BHIP_HEAT_X0_PROJECTED)
        return bhip_fail(ctx, BHIP_ERR_ARG, "unknown x0_mode");
    if (!ctx->assembled || ctx->asm_kind != BHIP_ASM_HEAT_TRANSIENT || !ctx->val_aux.p)
        return bhip_fail(ctx, BHIP_ERR_STATE, "heat run needs a BHIP_ASM_HEAT_TRANSIENT assembly");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const long long n = ctx->n_dof;
    BHIP_PROPAGATE(bhip_lift_build(ctx));
    bhip_heat_stats H;
    memset(&H, 0, sizeof(H));
    // the run's device time includes its preparation (layout, run copy of K)
    BHIP_TRY(ctx, hipEventRecord(ctx->drv0, s));
    dvec<double> &us = ctx->heat_us;
    BHIP_TRY(ctx, us.resize((size_t)n));
    BHIP_TRY(ctx, hipMemcpyAsync(us.p, u, sizeof(double) * n, hipMemcpyHostToDevice, s));
    // The refined solver's layout decides the form of -K u_s: on a column-run
    // layout K is copied once per call into FP64 runs (grow-only context
    // buffers), which stream at about twice the rate of the CSR form with 0.8
    // of its bytes; otherwise, or when that copy does not fit in device
    // memory, the CSR of K is read every step.
    const double *kruns = nullptr;
    double *ku = nullptr;
    if (opts->precision == BHIP_PREC_FP32_IR) {
        BHIP_PROPAGATE(bhip_prepare_refined(ctx));
        if (ctx->use_bsr && ctx->sell_runs) {
            const size_t nr = (size_t)3 * (ctx->n_sell > 0 ? ctx->n_sell : 1);
            if (ctx->heat_kruns.resize(nr) == hipSuccess && ctx->heat_ku.resize((size_t)n) == hipSuccess) {
                bhip_runs_from_csr64(ctx, ctx->val_aux.p, ctx->heat_kruns.p);
                BHIP_TRY(ctx, hipGetLastError());
                kruns = ctx->heat_kruns.p;
                ku = ctx->heat_ku.p;
            } else {
                (void)hipGetLastError();  // clear the failed allocation; the CSR form needs no buffer
                ctx->heat_kruns.release();
                ctx->heat_ku.release();
            }
        }
    }
    // projected guess: basis, Mbar du, block shares and coefficients in
    // grow-only context buffers; Mbar du on the FP64 SELL values when the
    // refined solver keeps them, else on the FP64 CSR
    const bool proj = x0_mode == BHIP_HEAT_X0_PROJECTED;
    int pm = 0;  // basis vectors held
    if (proj) {
        if (ctx->heat_V.n < (size_t)kProjMax * n) BHIP_TRY(ctx, ctx->heat_V.resize((size_t)kProjMax * n));
        if (ctx->heat_z.n < (size_t)n) BHIP_TRY(ctx, ctx->heat_z.resize((size_t)n));
        if (ctx->heat_part.n < kProjShares + kProjNV) BHIP_TRY(ctx, ctx->heat_part.resize(kProjShares + kProjNV));
    }
    const bool z_sell = proj && opts->precision == BHIP_PREC_FP32_IR && ctx->use_bsr && ctx->f64b_ready &&
                        ctx->sell_b == 1;
    double *part = proj ? ctx->heat_part.p : nullptr;
    double *coef = proj ? ctx->heat_part.p + kProjShares : nullptr;
    // Fbar = -(a1 + a2) K u_s / dt with a1 + a2 = dt for every alpha
    const double fcoef = -1.0;
    int status = BHIP_OK;
    for (int step = 0; step < n_steps; step++) {
        launch_heat_rhs(ctx, us.p, fcoef, kruns, ku);
        BHIP_TRY(ctx, hipGetLastError());
        bhip_solve_opts o = *opts;
        o.use_x0 = 0;
        if (x0_mode == BHIP_HEAT_X0_PREVIOUS && step > 0) {
            BHIP_TRY(ctx, hipMemcpyAsync(ctx->x0.p, ctx->x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            o.use_x0 = 1;
        } else if (x0_mode == BHIP_HEAT_X0_REFERENCE) {
            k_guess_ref<<<grid_for(n, 256), 256, 0, s>>>(n, ctx->rest.p, us.p, ctx->presc.p, ctx->x0.p);
            o.use_x0 = 1;
        } else if (proj && pm > 0) {
            bhip_proj_dots(ctx, n, pm, ctx->heat_V.p, nullptr, ctx->rhs.p, -1, 0, part, coef);
            bhip_proj_combo(ctx, n, pm, ctx->heat_V.p, coef, nullptr, 1.0, 0, ctx->x0.p);
            BHIP_TRY(ctx, hipGetLastError());
            o.use_x0 = 1;
        }
        bhip_stats st;
        memset(&st, 0, sizeof(st));
        const int e = bhip_solve_impl(ctx, &o, &st);
        H.iters += st.iters;
        H.outer_iters += st.outer_iters;
        if (st.rel_resid > H.max_rel_resid) H.max_rel_resid = st.rel_resid;
        if (e == BHIP_ERR_NOT_CONVERGED) {
            status = e;  // keep the step, as the reference would have stopped here
            H.steps = step + 1;
            k_add_inplace<<<grid_for(n, 256), 256, 0, s>>>(n, us.p, ctx->x.p);
            break;
        }
        if (e != BHIP_OK) return e;
        k_add_inplace<<<grid_for(n, 256), 256, 0, s>>>(n, us.p, ctx->x.p);
        BHIP_TRY(ctx, hipGetLastError());
        H.steps = step + 1;
        if (proj && step + 1 < n_steps) {
            // append du to the basis (restart from it when the basis is full)
            if (pm == kProjMax) pm = 0;
            if (z_sell) bhip_runs_spmv64(ctx, ctx->bval64.p, ctx->x.p, ctx->heat_z.p);
            else BHIP_PROPAGATE(bhip_spmv_dev(ctx, ctx->x.p, ctx->heat_z.p));
            bhip_proj_dots(ctx, n, pm, ctx->heat_V.p, ctx->x.p, ctx->heat_z.p, -1, 1, part, coef);
            bhip_proj_combo(ctx, n, pm, ctx->heat_V.p, coef, ctx->x.p, -1.0, 1, ctx->heat_V.p + (size_t)pm * n);
            BHIP_TRY(ctx, hipGetLastError());
            pm++;
        }
    }
    BHIP_TRY(ctx, hipEventRecord(ctx->drv1, s));
    BHIP_TRY(ctx, hipEventSynchronize(ctx->drv1));
    float ms = 0;
    BHIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->drv0, ctx->drv1));
    H.ms = ms;
    BHIP_TRY(ctx, bhip_copy(ctx, u, us.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (stats) *stats = H;
    if (status != BHIP_OK) return bhip_fail(ctx, status, "transient step not converged within the iteration cap");
    return BHIP_OK;
}
