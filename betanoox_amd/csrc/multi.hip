// multi.hip -- several load cases on one assembled matrix.
//
// Reference: the multi-column solve of MultiFieldMacroElement::CalculateK /
// solveForStiffnessMatrixColumn (MultiFieldMacroElement.cpp:135-407): K is
// assembled once, every load case re-lifts its prescribed displacements
// through the assembled K (Equations::updatePseudoLoadVectorUsingGK,
// equation.cpp:343-381) and is solved with the same matrix.
//
// Device form: R <= 8 cases are solved together.  The FP32 inner Jacobi-PCG
// (pack.hip) keeps an independent CG recurrence per case (its own alpha, beta
// and stopping test) but reads every 3x3 block of the SELL-64 matrix once for
// all R cases; a context alone is a pack of one system, so a case solved here
// gives the same bits as the same system solved in a pack with others.  This
// file holds the FP64 side: right-hand sides from the lift table, the outer
// residual b_r - K x_r on the FP64 CSR (K read once for all cases), scaling
// into the FP32 inner right-hand sides and x_r += |r_r| d_r.
#include <cmath>
#include <cstdlib>

#include "bhip_internal.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kFinThreads = 256;
constexpr int kMaxR = 8;

template <int NV>
__device__ __forceinline__ void store_partials_n(double (&v)[NV], double *partials) {
    __shared__ double sm[(kThreads / 64) * NV];
    block_sum<NV>(v, sm);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) partials[(long long)blockIdx.x * NV + k] = v[k];
}

// One block of kFinThreads: fixed-order sum of nb partial records.
template <int NV>
__device__ __forceinline__ void fin_sum_n(const double *partials, int nb, double (&tot)[NV]) {
    __shared__ double sm[(kFinThreads / 64) * NV];
#pragma unroll
    for (int k = 0; k < NV; k++) tot[k] = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; k++) tot[k] += partials[(long long)b * NV + k];
    block_sum<NV>(tot, sm);
}

// r_r = b_r - K x_r in FP64 on the full CSR for all R cases at once (K read
// once); LPR lanes per row, grid-stride over row groups; partials |r_r|^2 and
// |b_r|^2.  Each row sums in the order of the single-case residual kernel.
template <int LPR, int R>
__global__ void __launch_bounds__(kThreads) k_csr_resid_multi(int n, const int *__restrict__ rowptr,
                                                            const int *__restrict__ colind,
                                                            const double *__restrict__ val,
                                                            const double *__restrict__ x,
                                                            const double *__restrict__ b, double *__restrict__ rout,
                                                            double *partials) {
    constexpr int RPB = kThreads / LPR;
    const int g = threadIdx.x / LPR, l = threadIdx.x % LPR;
    double acc[2 * R];
#pragma unroll
    for (int k = 0; k < 2 * R; k++) acc[k] = 0;
    for (long long row0 = (long long)blockIdx.x * RPB; row0 < n; row0 += (long long)gridDim.x * RPB) {
        const long long row = row0 + g;
        double sr[R];
#pragma unroll
        for (int r = 0; r < R; r++) sr[r] = 0;
        if (row < n) {
            const int kb = rowptr[row], ke = rowptr[row + 1];
            for (int k = kb + l; k < ke; k += LPR) {
                const double v = val[k];
                const long long c = colind[k];
#pragma unroll
                for (int r = 0; r < R; r++) sr[r] += v * x[r * (long long)n + c];
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++)
#pragma unroll
            for (int off = LPR / 2; off > 0; off >>= 1) sr[r] += __shfl_xor(sr[r], off, LPR);
        if (row < n && l == 0) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const long long j = r * (long long)n + row;
                const double bi = b[j], ri = bi - sr[r];
                rout[j] = ri;
                acc[2 * r] += ri * ri;
                acc[2 * r + 1] += bi * bi;
            }
        }
    }
    store_partials_n<2 * R>(acc, partials);
}

// out[2r] = |r_r|^2, out[2r + 1] = |b_r|^2
template <int R>
__global__ void __launch_bounds__(kFinThreads) k_fin_resid_multi(int nb, const double *partials, double *out) {
    double t[2 * R];
    fin_sum_n<2 * R>(partials, nb, t);
    if (threadIdx.x == 0)
        for (int k = 0; k < 2 * R; k++) out[k] = t[k];
}

// r32_r = r64_r * scale_r, d32_r = 0  (blockIdx.y = case)
__global__ void k_pack_f32(long long n, const double *r64, const double *scale, float *r32, float *d32) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long j = (long long)blockIdx.y * n + i;
    r32[j] = (float)(r64[j] * scale[blockIdx.y]);
    d32[j] = 0.f;
}

// x64_r += s_r d32_r  (blockIdx.y = case)
__global__ void k_unpack_axpy(long long n, const double *s, const float *d32, double *x64) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const long long j = (long long)blockIdx.y * n + i;
    x64[j] += s[blockIdx.y] * (double)d32[j];
}

// lanes per CSR row from the mean row length (as the single-case kernels)
int resid_lpr(bhip_ctx *ctx) {
    const double avg = ctx->n_dof > 0 ? (double)ctx->nnz / (double)ctx->n_dof : 1.0;
    return avg > 160 ? 16 : avg > 12 ? 8 : avg > 4 ? 4 : 2;
}

int resid_grid(bhip_ctx *ctx) {
    const long long rpb = kThreads / resid_lpr(ctx);
    const long long want = (ctx->n_dof + rpb - 1) / rpb;
    return (int)std::max<long long>(1, std::min<long long>(want, kRedBlocks));
}

}  // namespace

void MultiData::release_buffers() {
    b64.release();
    x64.release();
    r64.release();
    scale.release();
    rn.release();
    save.release();
    r32.release();
    p32.release();
    q32.release();
    d32.release();
    partials.release();
    resid.release();
}

// The packed solver's graphs hold only the address of its descriptor block, so
// a new matrix or layout needs no new graph.
void bhip_multi_invalidate(bhip_ctx *) {}

void bhip_multi_release(bhip_ctx *ctx) {
    MultiData *M = ctx->multi;
    if (!M) return;
    if (M->pack) bhip_pack_engine_destroy(M->pack);
    M->release_buffers();
    if (M->h_resid) (void)hipHostFree(M->h_resid);
    delete M;
    ctx->multi = nullptr;
}

int bhip_multi_prepare(bhip_ctx *ctx, int R, bool fused) {
    if (!ctx->multi) {
        ctx->multi = new MultiData();
        BHIP_TRY(ctx, hipHostMalloc((void **)&ctx->multi->h_resid, sizeof(double) * 2 * kMaxR, hipHostMallocDefault));
    }
    MultiData *M = ctx->multi;
    const long long n = ctx->n_dof;
    // grow-only buffers: a new size keeps the allocations that are large enough
    M->n = n;
    M->R = std::max(M->R, R);
    const size_t nr = (size_t)n * M->R;
    BHIP_TRY(ctx, M->b64.resize(nr));
    BHIP_TRY(ctx, M->x64.resize(nr));
    BHIP_TRY(ctx, M->r64.resize(nr));
    BHIP_TRY(ctx, M->scale.resize(kMaxR));
    BHIP_TRY(ctx, M->rn.resize(kMaxR));
    BHIP_TRY(ctx, M->resid.resize(2 * kMaxR));
    const size_t np = (size_t)kRedBlocks * 2 * kMaxR;
    if (M->partials.n < np) BHIP_TRY(ctx, M->partials.resize(np));
    if (fused) {
        BHIP_TRY(ctx, M->r32.resize(nr));
        BHIP_TRY(ctx, M->p32.resize(nr));
        BHIP_TRY(ctx, M->q32.resize(nr));
        BHIP_TRY(ctx, M->d32.resize(nr));
    } else {
        BHIP_TRY(ctx, M->save.resize((size_t)2 * n));
    }
    return BHIP_OK;
}

namespace {

template <int R>
void enqueue_residuals(bhip_ctx *ctx, hipStream_t s, double *out) {
    MultiData &M = *ctx->multi;
    const int g = resid_grid(ctx);
    const int n = ctx->n_dof;
#define BHIP_RESID(L)                                                                                          \
    k_csr_resid_multi<L, R><<<g, kThreads, 0, s>>>(n, ctx->rowptr.p, ctx->colind.p, ctx->val.p, M.x64.p, M.b64.p, \
                                                   M.r64.p, M.partials.p)
    switch (resid_lpr(ctx)) {
        case 16: BHIP_RESID(16); break;
        case 8: BHIP_RESID(8); break;
        case 4: BHIP_RESID(4); break;
        default: BHIP_RESID(2); break;
    }
#undef BHIP_RESID
    k_fin_resid_multi<R><<<1, kFinThreads, 0, s>>>(g, M.partials.p, out);
}

}  // namespace

// r_r = b_r - K x_r (FP64 CSR, K read once) for every case on stream s;
// out (device, 2R doubles) = |r_r|^2, |b_r|^2.
void bhip_multi_enqueue_resid(bhip_ctx *ctx, int R, hipStream_t s, double *out) {
    switch (R) {
        case 1: enqueue_residuals<1>(ctx, s, out); break;
        case 2: enqueue_residuals<2>(ctx, s, out); break;
        case 3: enqueue_residuals<3>(ctx, s, out); break;
        case 4: enqueue_residuals<4>(ctx, s, out); break;
        case 5: enqueue_residuals<5>(ctx, s, out); break;
        case 6: enqueue_residuals<6>(ctx, s, out); break;
        case 7: enqueue_residuals<7>(ctx, s, out); break;
        default: enqueue_residuals<8>(ctx, s, out); break;
    }
}

// r32_r = r64_r * scale_r, d32_r = 0 (scale in multi->scale).
void bhip_multi_enqueue_pack32(bhip_ctx *ctx, int R, hipStream_t s) {
    MultiData &M = *ctx->multi;
    const long long n = ctx->n_dof;
    k_pack_f32<<<dim3(grid_for(n, 256), R), 256, 0, s>>>(n, M.r64.p, M.scale.p, M.r32.p, M.d32.p);
}

// x64_r += rn_r d32_r (rn in multi->rn).
void bhip_multi_enqueue_axpy(bhip_ctx *ctx, int R, hipStream_t s) {
    MultiData &M = *ctx->multi;
    const long long n = ctx->n_dof;
    k_unpack_axpy<<<dim3(grid_for(n, 256), R), 256, 0, s>>>(n, M.rn.p, M.d32.p, M.x64.p);
}

// Worst true relative residual over the R cases of the context's x64 (FP64).
int bhip_multi_worst_resid(bhip_ctx *ctx, int R, double *worst) {
    MultiData &M = *ctx->multi;
    bhip_multi_enqueue_resid(ctx, R, ctx->stream, M.resid.p);
    BHIP_TRY(ctx, hipGetLastError());
    BHIP_TRY(ctx, hipMemcpyAsync(M.h_resid, M.resid.p, sizeof(double) * 2 * R, hipMemcpyDeviceToHost, ctx->stream));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    double w = 0;
    for (int r = 0; r < R; r++) {
        const double rr = M.h_resid[2 * r], bb = M.h_resid[2 * r + 1];
        w = std::max(w, bb > 0 ? sqrt(rr / bb) : sqrt(rr));
    }
    *worst = w;
    return BHIP_OK;
}

// The part of bhip_solve_cases before the solve: buffers, right-hand sides
// (each case's prescribed values lifted through the coupling blocks of the
// assembled K; M.r64 stages the prescribed values, the context's q vector the
// load), x = 0.  *fused: the cases can run in the packed FP32 solver.
int bhip_cases_begin(bhip_ctx *ctx, int nrhs, const double *presc, const double *load, const bhip_solve_opts *opts,
                     bool *fused) {
    if (!presc || !opts) return bhip_fail(ctx, BHIP_ERR_ARG, "presc and opts are required");
    if (nrhs < 1 || nrhs > kMaxR) return bhip_fail(ctx, BHIP_ERR_ARG, "nrhs must be in [1, 8]");
    if (!(opts->tol > 0)) return bhip_fail(ctx, BHIP_ERR_ARG, "tol must be positive");
    if (opts->precision != BHIP_PREC_FP64 && opts->precision != BHIP_PREC_FP32_IR &&
        opts->precision != BHIP_PREC_FP64_BLOCKED)
        return bhip_fail(ctx, BHIP_ERR_ARG, "unknown precision");
    if (!ctx->assembled) return bhip_fail(ctx, BHIP_ERR_STATE, "matrix not assembled");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const long long n = ctx->n_dof;
    const bool want32 = opts->precision == BHIP_PREC_FP32_IR;
    BHIP_PROPAGATE(bhip_prepare_solver(ctx, want32));
    *fused = want32 && ctx->use_bsr && ctx->sell_b == 3;
    BHIP_PROPAGATE(bhip_multi_prepare(ctx, nrhs, *fused));
    MultiData &M = *ctx->multi;
    const double *dload = nullptr;
    if (load) {
        BHIP_TRY(ctx, hipMemcpyAsync(ctx->q.p, load, sizeof(double) * n, hipMemcpyHostToDevice, s));
        dload = ctx->q.p;
    }
    BHIP_TRY(ctx, hipMemcpyAsync(M.r64.p, presc, sizeof(double) * n * nrhs, hipMemcpyHostToDevice, s));
    BHIP_PROPAGATE(bhip_lift_rhs(ctx, nrhs, M.r64.p, dload, M.b64.p));
    BHIP_TRY(ctx, hipMemsetAsync(M.x64.p, 0, sizeof(double) * n * nrhs, s));
    BHIP_TRY(ctx, hipStreamSynchronize(s));
    return BHIP_OK;
}

// The part after the solve: worst true relative residual and the solutions.
int bhip_cases_finish(bhip_ctx *ctx, int nrhs, double *x, bhip_stats *S) {
    BHIP_PROPAGATE(bhip_multi_worst_resid(ctx, nrhs, &S->rel_resid));
    BHIP_TRY(ctx, bhip_copy(ctx, x, ctx->multi->x64.p, sizeof(double) * ctx->n_dof * nrhs, hipMemcpyDeviceToHost));
    return BHIP_OK;
}

namespace {

// Plain FP64, or a mesh without 3 DOFs on every node: the cases run in turn
// through the single-case solver; the context's RHS and solution are
// restored afterwards.
int solve_in_turn(bhip_ctx *ctx, int nrhs, const bhip_solve_opts *opts, bhip_stats &S) {
    hipStream_t s = ctx->stream;
    const long long n = ctx->n_dof;
    MultiData &M = *ctx->multi;
    int st = BHIP_OK;
    BHIP_TRY(ctx, hipMemcpyAsync(M.save.p, ctx->rhs.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    BHIP_TRY(ctx, hipMemcpyAsync(M.save.p + n, ctx->x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    for (int r = 0; r < nrhs; r++) {
        BHIP_TRY(ctx, hipMemcpyAsync(ctx->rhs.p, M.b64.p + r * n, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        bhip_solve_opts o1 = *opts;
        o1.use_x0 = 0;
        bhip_stats s1;
        memset(&s1, 0, sizeof(s1));
        const int e = bhip_solve_impl(ctx, &o1, &s1);
        if (e != BHIP_OK && e != BHIP_ERR_NOT_CONVERGED) return e;
        if (e == BHIP_ERR_NOT_CONVERGED) st = e;
        S.iters += s1.iters;
        S.case_iters += s1.iters;
        S.outer_iters = std::max(S.outer_iters, s1.outer_iters);
        S.spmv_launches += s1.spmv_launches;
        BHIP_TRY(ctx, hipMemcpyAsync(M.x64.p + r * n, ctx->x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->rhs.p, M.save.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->x.p, M.save.p + n, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    return st;
}

}  // namespace

extern "C" int bhip_solve_cases(bhip_ctx *ctx, int nrhs, const double *presc, const double *load, double *x,
                                const bhip_solve_opts *opts, bhip_stats *stats) {
    if (!ctx) return BHIP_ERR_ARG;
    if (!x) return bhip_fail(ctx, BHIP_ERR_ARG, "x is required");
    (void)hipSetDevice(ctx->device);
    bhip_stats S;
    memset(&S, 0, sizeof(S));
    bool fused = false;
    BHIP_TRY(ctx, hipEventRecord(ctx->drv0, ctx->stream));
    BHIP_PROPAGATE(bhip_cases_begin(ctx, nrhs, presc, load, opts, &fused));
    int st = BHIP_OK;
    if (fused) {
        MultiData &M = *ctx->multi;
        if (!M.pack) BHIP_PROPAGATE(bhip_pack_engine_create(ctx->device, ctx->stream, &M.pack) == BHIP_OK
                                        ? BHIP_OK
                                        : bhip_fail(ctx, BHIP_ERR_HIP, "packed solver set-up failed"));
        bhip_ctx *one[1] = {ctx};
        const int e = bhip_pack_run(M.pack, one, 1, nrhs, *opts, &S, &st);
        if (e != BHIP_OK) return e;
    } else {
        st = solve_in_turn(ctx, nrhs, opts, S);
        if (st != BHIP_OK && st != BHIP_ERR_NOT_CONVERGED) return st;
    }
    BHIP_TRY(ctx, hipEventRecord(ctx->drv1, ctx->stream));
    BHIP_TRY(ctx, hipEventSynchronize(ctx->drv1));
    float ms = 0;
    BHIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->drv0, ctx->drv1));
    S.ms_total = ms;
    BHIP_PROPAGATE(bhip_cases_finish(ctx, nrhs, x, &S));
    if (stats) *stats = S;
    if (st == BHIP_ERR_NOT_CONVERGED)
        return bhip_fail(ctx, st, "multi-case solve not converged within the iteration caps");
    return BHIP_OK;
}
