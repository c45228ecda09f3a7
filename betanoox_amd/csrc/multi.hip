// multi.hip -- several load cases on one assembled matrix.
//
// Reference: the multi-column solve of MultiFieldMacroElement::CalculateK /
// solveForStiffnessMatrixColumn (MultiFieldMacroElement.cpp:135-407): K is
// assembled once, every load case re-lifts its prescribed displacements
// through the assembled K (Equations::updatePseudoLoadVectorUsingGK,
// equation.cpp:343-381) and is solved with the same matrix.
//
// Device form: R <= 8 cases are solved together.  The inner FP32 Jacobi-PCG
// keeps an independent CG recurrence per case (its own alpha, beta and
// stopping test) but reads every 3x3 block of the SELL-64 matrix once for all
// R cases.  Vectors are case-major (element (case r, DOF i) at r*n + i), so
// the vector kernels stream each case coalesced and the SpMV gathers, for
// each case, the same 12-byte node triples the single-case kernel gathers.
// FP64 refinement per case uses the FP64 CSR (b_r - K x_r).  A chunk of
// iterations is captured once as a HIP graph and replayed; the host polls the
// device state between chunks.
#include <cmath>
#include <cstdlib>

#include "bhip_internal.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kFinThreads = 256;
constexpr int kMaxR = 8;

struct MultiState {
    double rz[kMaxR], rr[kMaxR], rr0[kMaxR], alpha[kMaxR], beta[kMaxR];
    double orr[kMaxR], obb[kMaxR];  // outer FP64 residual: |b - K x|^2, |b|^2
    int done[kMaxR], iter[kMaxR], converged[kMaxR];
    double tol;
    int maxit, all_done, iters, pad;
};

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

__device__ __forceinline__ void range_of(long long n, int lb, int nb, long long &b, long long &e) {
    long long per = (n + nb - 1) / nb;
    b = (long long)lb * per;
    e = b + per;
    if (e > n) e = n;
    if (b > n) b = n;
}

// y_r = K x_r for all R cases (SELL-64 3x3 blocks, lane = node row);
// partials x_r . y_r.  SPLIT = 1: each of the 4 waves of a block owns one
// slice.  SPLIT = 4 (small matrices, too few slices to fill the GPU): the 4
// waves share one slice, each taking every fourth block column, and wave 0
// adds the other waves' row sums in wave order through LDS.  NT: streaming
// (nontemporal) loads of block values and columns, used with the single-case
// shapes that stream (matrices larger than the last-level cache).
template <int NT, class T>
__device__ __forceinline__ T ld_stream(const T *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

template <int R, int SPLIT, int NT>
__global__ void __launch_bounds__(kThreads) k_sell3_spmv_multi(int n_nodes, int n_slices, const int *__restrict__ sptr,
                                                             const int *__restrict__ scol,
                                                             const float *__restrict__ sval, long long nsell,
                                                             long long ld, const float *__restrict__ x,
                                                             float *__restrict__ y, const MultiState *st,
                                                             double *partials) {
    if (st->all_done) return;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const long long s = SPLIT == 1 ? (long long)lb * (kThreads / 64) + wave : (long long)lb;
    double dots[R];
#pragma unroll
    for (int r = 0; r < R; r++) dots[r] = 0;
    if (s < n_slices) {
        const long long node = s * 64 + lane;
        const int kb = sptr[s], ke = sptr[s + 1];
        float a0[R], a1[R], a2[R];
#pragma unroll
        for (int r = 0; r < R; r++) a0[r] = a1[r] = a2[r] = 0.f;
#pragma unroll 2
        for (int k = kb + lane + (SPLIT == 1 ? 0 : 64 * wave); k < ke; k += 64 * SPLIT) {
            const long long m = ld_stream<NT>(scol + k);
            const float v0 = ld_stream<NT>(sval + 0 * nsell + k), v1 = ld_stream<NT>(sval + 1 * nsell + k);
            const float v2 = ld_stream<NT>(sval + 2 * nsell + k), v3 = ld_stream<NT>(sval + 3 * nsell + k);
            const float v4 = ld_stream<NT>(sval + 4 * nsell + k), v5 = ld_stream<NT>(sval + 5 * nsell + k);
            const float v6 = ld_stream<NT>(sval + 6 * nsell + k), v7 = ld_stream<NT>(sval + 7 * nsell + k);
            const float v8 = ld_stream<NT>(sval + 8 * nsell + k);
            const float *xm = x + 3 * m;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const float x0 = xm[r * ld], x1 = xm[r * ld + 1], x2 = xm[r * ld + 2];
                a0[r] += v0 * x0 + v1 * x1 + v2 * x2;
                a1[r] += v3 * x0 + v4 * x1 + v5 * x2;
                a2[r] += v6 * x0 + v7 * x1 + v8 * x2;
            }
        }
        if (SPLIT > 1) {
            __shared__ float red[3 * R][kThreads];
            if (wave > 0)
#pragma unroll
                for (int r = 0; r < R; r++) {
                    red[3 * r][threadIdx.x] = a0[r];
                    red[3 * r + 1][threadIdx.x] = a1[r];
                    red[3 * r + 2][threadIdx.x] = a2[r];
                }
            __syncthreads();
            if (wave == 0)
                for (int w = 1; w < SPLIT; w++)
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        a0[r] += red[3 * r][64 * w + lane];
                        a1[r] += red[3 * r + 1][64 * w + lane];
                        a2[r] += red[3 * r + 2][64 * w + lane];
                    }
        }
        if (node < n_nodes && (SPLIT == 1 || wave == 0)) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const long long j = r * ld + 3 * node;
                y[j] = a0[r];
                y[j + 1] = a1[r];
                y[j + 2] = a2[r];
                dots[r] = (double)x[j] * a0[r] + (double)x[j + 1] * a1[r] + (double)x[j + 2] * a2[r];
            }
        }
    }
    store_partials_n<R>(dots, partials);
}

// Finalize kernels: thread r < R owns case r.  Its state is loaded at the
// start so that the loads overlap the partial sums; the block sum lands in
// thread 0 and is passed to the case threads through LDS.
template <int R>
__global__ void __launch_bounds__(kFinThreads) k_fin_alpha_multi(int nb, const double *partials, MultiState *st) {
    if (st->all_done) return;
    const int c = threadIdx.x;
    int done = 1;
    double rz = 0;
    if (c < R) {
        done = st->done[c];
        rz = st->rz[c];
    }
    double t[R];
    fin_sum_n<R>(partials, nb, t);
    __shared__ double tot[R];
    if (threadIdx.x == 0)
#pragma unroll
        for (int r = 0; r < R; r++) tot[r] = t[r];
    __syncthreads();
    if (c < R) st->alpha[c] = done ? 0.0 : rz / tot[c];
}

// Vector kernels run on a (blocks, R) grid: block (b, r) streams case r over
// the b-th DOF range and writes its partials at b * 2R + 2r, the layout the
// finalize kernels sum.  Blocks of a case that has finished return at once
// (the finalize kernels ignore such a case's partials).

// x_r += alpha_r p_r; r_r -= alpha_r q_r; partials (r.z, r.r).
template <int R>
__global__ void __launch_bounds__(kThreads) k_update_multi(long long n, float *__restrict__ x, float *__restrict__ rv,
                                                         const float *__restrict__ p, const float *__restrict__ q,
                                                         const float *__restrict__ dinv, const MultiState *st,
                                                         double *partials) {
    const int r = blockIdx.y;
    if (st->all_done || st->done[r]) return;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    long long i0, i1;
    range_of(n, lb, gridDim.x, i0, i1);
    const float alpha = (float)st->alpha[r];
    const long long o = (long long)r * n;
    double acc[2] = {0, 0};
    for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const float ri = rv[o + i] - alpha * q[o + i];
        x[o + i] = x[o + i] + alpha * p[o + i];
        rv[o + i] = ri;
        acc[0] += (double)ri * (ri * dinv[i]);
        acc[1] += (double)ri * ri;
    }
    __shared__ double sm[(kThreads / 64) * 2];
    block_sum<2>(acc, sm);
    if (threadIdx.x == 0) {
        partials[(long long)blockIdx.x * 2 * R + 2 * r] = acc[0];
        partials[(long long)blockIdx.x * 2 * R + 2 * r + 1] = acc[1];
    }
}

template <int R>
__global__ void __launch_bounds__(kFinThreads) k_fin_update_multi(int nb, const double *partials, MultiState *st) {
    if (st->all_done) return;
    const int c = threadIdx.x;
    int done = 1, iter = 0;
    double rz = 0, rr0 = 1, tol = 0;
    int maxit = 0;
    if (c < R) {
        done = st->done[c];
        rz = st->rz[c];
        rr0 = st->rr0[c];
        iter = st->iter[c];
        tol = st->tol;
        maxit = st->maxit;
    }
    double t[2 * R];
    fin_sum_n<2 * R>(partials, nb, t);
    __shared__ double tot[2 * R];
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < 2 * R; k++) tot[k] = t[k];
    __syncthreads();
    if (c < R && !done) {
        const double rz_new = tot[2 * c];
        st->beta[c] = rz_new / rz;
        st->rz[c] = rz_new;
        st->rr[c] = tot[2 * c + 1];
        st->iter[c] = iter + 1;
        const bool conv = sqrt(tot[2 * c + 1] / rr0) <= tol;
        if (conv) st->converged[c] = 1;
        if (conv || iter + 1 >= maxit) done = 1;
        st->done[c] = done;
    }
    // the case threads are lanes 0..R-1 of wave 0
    const unsigned long long ball = __ballot(c < R && done);
    if (threadIdx.x == 0) {
        st->iters += 1;
        st->all_done = (ball & ((1ull << R) - 1)) == ((1ull << R) - 1);
    }
}

// p_r = z_r + beta_r p_r for the cases still running.
template <int R>
__global__ void __launch_bounds__(kThreads) k_pupdate_multi(long long n, float *__restrict__ p, const float *__restrict__ rv,
                                                          const float *__restrict__ dinv, const MultiState *st) {
    const int r = blockIdx.y;
    if (st->all_done || st->done[r]) return;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    long long i0, i1;
    range_of(n, lb, gridDim.x, i0, i1);
    const float beta = (float)st->beta[r];
    const long long o = (long long)r * n;
    for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) p[o + i] = rv[o + i] * dinv[i] + beta * p[o + i];
}

// p_r = z_r; partials (r.z, r.r).
template <int R>
__global__ void __launch_bounds__(kThreads) k_init_multi(long long n, const float *__restrict__ rv, float *__restrict__ p,
                                                       const float *__restrict__ dinv, double *partials) {
    const int r = blockIdx.y;
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    long long i0, i1;
    range_of(n, lb, gridDim.x, i0, i1);
    const long long o = (long long)r * n;
    double acc[2] = {0, 0};
    for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
        const float ri = rv[o + i];
        const float zi = ri * dinv[i];
        p[o + i] = zi;
        acc[0] += (double)ri * zi;
        acc[1] += (double)ri * ri;
    }
    __shared__ double sm[(kThreads / 64) * 2];
    block_sum<2>(acc, sm);
    if (threadIdx.x == 0) {
        partials[(long long)blockIdx.x * 2 * R + 2 * r] = acc[0];
        partials[(long long)blockIdx.x * 2 * R + 2 * r + 1] = acc[1];
    }
}

template <int R>
__global__ void __launch_bounds__(kFinThreads) k_fin_init_multi(int nb, const double *partials, MultiState *st,
                                                              double tol, int maxit) {
    double t[2 * R];
    fin_sum_n<2 * R>(partials, nb, t);
    if (threadIdx.x == 0) {
        int all = 1;
        st->tol = tol;
        st->maxit = maxit;
        st->iters = 0;
        for (int r = 0; r < R; r++) {
            st->rz[r] = t[2 * r];
            st->rr[r] = t[2 * r + 1];
            st->rr0[r] = t[2 * r + 1];
            st->iter[r] = 0;
            st->alpha[r] = st->beta[r] = 0;
            // a zero inner right-hand side (case already converged) is solved by d = 0
            st->converged[r] = (t[2 * r + 1] == 0.0);
            st->done[r] = st->converged[r];
            all &= st->done[r];
        }
        st->all_done = all;
    }
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

template <int R>
__global__ void __launch_bounds__(kFinThreads) k_fin_resid_multi(int nb, const double *partials, MultiState *st) {
    double t[2 * R];
    fin_sum_n<2 * R>(partials, nb, t);
    if (threadIdx.x == 0)
        for (int r = 0; r < R; r++) {
            st->orr[r] = t[2 * r];
            st->obb[r] = t[2 * r + 1];
        }
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

int vec_grid(bhip_ctx *ctx, long long n) {
    long long want = (n + (long long)kThreads * 8 - 1) / ((long long)kThreads * 8);
    long long cap = std::min<long long>((long long)ctx->n_cu * 8, kRedBlocks);
    return (int)std::max<long long>(1, std::min(want, cap));
}

// Matrices with fewer slices than this run the SPLIT = 4 SpMV (one slice per
// block): with one wave per slice they would leave most CUs idle.
constexpr int kSplitSlices = 2048;

bool sell_split(bhip_ctx *ctx) { return ctx->n_slices < kSplitSlices; }

int sell_grid(bhip_ctx *ctx) {
    if (sell_split(ctx)) return ctx->n_slices > 0 ? ctx->n_slices : 1;
    long long g = (ctx->n_slices + kThreads / 64 - 1) / (kThreads / 64);
    return (int)(g > 0 ? g : 1);
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

// Persistent buffers of the multi-case solve (one set per context).
struct MultiData {
    int R = 0;        // cases the buffers hold
    long long n = 0;  // DOFs per case
    dvec<double> b64, x64, r64, scale, rn;  // case-major, R x n
    dvec<double> save;                     // RHS and solution of the context (case-by-case path)
    dvec<float> r32, p32, q32, d32;        // case-major, R x n
    dvec<double> partials;
    dvec<MultiState> st;
    MultiState *h_st = nullptr;  // pinned mirror
    hipGraphExec_t graph = nullptr;
    int graph_R = 0, graph_chunk = 0;

    void release_buffers() {
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
        st.release();
    }
};

void bhip_multi_invalidate(bhip_ctx *ctx) {
    MultiData *M = ctx->multi;
    if (M && M->graph) {
        (void)hipGraphExecDestroy(M->graph);
        M->graph = nullptr;
        M->graph_R = M->graph_chunk = 0;
    }
}

void bhip_multi_release(bhip_ctx *ctx) {
    MultiData *M = ctx->multi;
    if (!M) return;
    bhip_multi_invalidate(ctx);
    M->release_buffers();
    if (M->h_st) (void)hipHostFree(M->h_st);
    delete M;
    ctx->multi = nullptr;
}

namespace {

int ensure_multi(bhip_ctx *ctx, int R, bool fused) {
    if (!ctx->multi) {
        ctx->multi = new MultiData();
        BHIP_TRY(ctx, hipHostMalloc((void **)&ctx->multi->h_st, sizeof(MultiState), hipHostMallocDefault));
    }
    MultiData *M = ctx->multi;
    const long long n = ctx->n_dof;
    if (M->n != n || M->R < R) {
        bhip_multi_invalidate(ctx);
        M->release_buffers();
        M->n = n;
        M->R = R;
    }
    const size_t nr = (size_t)n * M->R;
    BHIP_TRY(ctx, M->b64.resize(nr));
    BHIP_TRY(ctx, M->x64.resize(nr));
    BHIP_TRY(ctx, M->r64.resize(nr));
    BHIP_TRY(ctx, M->scale.resize(kMaxR));
    BHIP_TRY(ctx, M->rn.resize(kMaxR));
    const size_t np = (size_t)std::max(sell_grid(ctx), kRedBlocks) * 2 * kMaxR;
    if (M->partials.n < np) {
        bhip_multi_invalidate(ctx);
        BHIP_TRY(ctx, M->partials.resize(np));
    }
    BHIP_TRY(ctx, M->st.resize(1));
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

template <int R>
void enqueue_iteration(bhip_ctx *ctx, MultiData &M) {
    hipStream_t s = ctx->stream;
    const long long n = ctx->n_dof;
    const int gs = sell_grid(ctx), gv = vec_grid(ctx, n);
#define BHIP_MULTI_SPMV(SPLIT, NT)                                                                           \
    k_sell3_spmv_multi<R, SPLIT, NT><<<gs, kThreads, 0, s>>>(ctx->n_nodes, ctx->n_slices, ctx->sell_ptr.p,   \
                                                             ctx->sell_col.p, ctx->bval32.p, ctx->n_sell, n, \
                                                             M.p32.p, M.q32.p, M.st.p, M.partials.p)
    const bool nt = bhip_sell_streams(ctx);
    if (sell_split(ctx)) {
        if (nt) BHIP_MULTI_SPMV(4, 1);
        else BHIP_MULTI_SPMV(4, 0);
    } else {
        if (nt) BHIP_MULTI_SPMV(1, 1);
        else BHIP_MULTI_SPMV(1, 0);
    }
#undef BHIP_MULTI_SPMV
    k_fin_alpha_multi<R><<<1, kFinThreads, 0, s>>>(gs, M.partials.p, M.st.p);
    k_update_multi<R><<<dim3(gv, R), kThreads, 0, s>>>(n, M.d32.p, M.r32.p, M.p32.p, M.q32.p, ctx->dinv32.p,
                                                       M.st.p, M.partials.p);
    k_fin_update_multi<R><<<1, kFinThreads, 0, s>>>(gv, M.partials.p, M.st.p);
    k_pupdate_multi<R><<<dim3(gv, R), kThreads, 0, s>>>(n, M.p32.p, M.r32.p, ctx->dinv32.p, M.st.p);
}

bool graphs_enabled() {
    const char *e = getenv("BHIP_NO_GRAPH");
    return !(e && e[0] == '1');
}

template <int R>
int get_graph(bhip_ctx *ctx, MultiData &M, int chunk, hipGraphExec_t *out) {
    if (M.graph && M.graph_R == R && M.graph_chunk == chunk) {
        *out = M.graph;
        return BHIP_OK;
    }
    bhip_multi_invalidate(ctx);
    hipGraph_t graph;
    BHIP_TRY(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < chunk; k++) enqueue_iteration<R>(ctx, M);
    hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
    BHIP_TRY(ctx, e);
    e = hipGraphInstantiate(&M.graph, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) M.graph = nullptr;
    BHIP_TRY(ctx, e);
    M.graph_R = R;
    M.graph_chunk = chunk;
    *out = M.graph;
    return BHIP_OK;
}

template <int R>
int inner_solve(bhip_ctx *ctx, MultiData &M, double inner_tol, int maxit, int check_every, int *iters,
                int *launches) {
    hipStream_t s = ctx->stream;
    const long long n = ctx->n_dof;
    const int gv = vec_grid(ctx, n);
    k_init_multi<R><<<dim3(gv, R), kThreads, 0, s>>>(n, M.r32.p, M.p32.p, ctx->dinv32.p, M.partials.p);
    k_fin_init_multi<R><<<1, kFinThreads, 0, s>>>(gv, M.partials.p, M.st.p, inner_tol, maxit);
    BHIP_TRY(ctx, hipGetLastError());
    const bool use_graph = graphs_enabled();
    hipGraphExec_t g = nullptr;
    if (use_graph) BHIP_PROPAGATE(get_graph<R>(ctx, M, check_every, &g));
    // every chunk advances each running case by one iteration at least, so
    // maxit/check_every + 2 chunks always suffice
    const long long max_chunks = (long long)maxit / check_every + 2;
    for (long long c = 0;; c++) {
        BHIP_TRY(ctx, hipMemcpyAsync(M.h_st, M.st.p, sizeof(MultiState), hipMemcpyDeviceToHost, s));
        BHIP_TRY(ctx, hipStreamSynchronize(s));
        if (M.h_st->all_done) break;
        if (c > max_chunks) return bhip_fail(ctx, BHIP_ERR_STATE, "multi-case PCG did not finish (internal error)");
        if (use_graph) {
            BHIP_TRY(ctx, hipGraphLaunch(g, s));
        } else {
            for (int k = 0; k < check_every; k++) enqueue_iteration<R>(ctx, M);
        }
        BHIP_TRY(ctx, hipGetLastError());
        *launches += check_every;
    }
    *iters = M.h_st->iters;
    return BHIP_OK;
}

template <int R>
void enqueue_residuals(bhip_ctx *ctx, MultiData &M) {
    hipStream_t s = ctx->stream;
    const int g = resid_grid(ctx);
    const int n = ctx->n_dof;
    switch (resid_lpr(ctx)) {
        case 16: k_csr_resid_multi<16, R><<<g, kThreads, 0, s>>>(n, ctx->rowptr.p, ctx->colind.p, ctx->val.p, M.x64.p, M.b64.p, M.r64.p, M.partials.p); break;
        case 8: k_csr_resid_multi<8, R><<<g, kThreads, 0, s>>>(n, ctx->rowptr.p, ctx->colind.p, ctx->val.p, M.x64.p, M.b64.p, M.r64.p, M.partials.p); break;
        case 4: k_csr_resid_multi<4, R><<<g, kThreads, 0, s>>>(n, ctx->rowptr.p, ctx->colind.p, ctx->val.p, M.x64.p, M.b64.p, M.r64.p, M.partials.p); break;
        default: k_csr_resid_multi<2, R><<<g, kThreads, 0, s>>>(n, ctx->rowptr.p, ctx->colind.p, ctx->val.p, M.x64.p, M.b64.p, M.r64.p, M.partials.p); break;
    }
    k_fin_resid_multi<R><<<1, kFinThreads, 0, s>>>(g, M.partials.p, M.st.p);
}

// r_r = b_r - K x_r (FP64 CSR, K read once) for every case; |r_r|^2, |b_r|^2.
int outer_residuals(bhip_ctx *ctx, MultiData &M, int R, std::vector<double> &rr, std::vector<double> &bb) {
    switch (R) {
        case 1: enqueue_residuals<1>(ctx, M); break;
        case 2: enqueue_residuals<2>(ctx, M); break;
        case 3: enqueue_residuals<3>(ctx, M); break;
        case 4: enqueue_residuals<4>(ctx, M); break;
        case 5: enqueue_residuals<5>(ctx, M); break;
        case 6: enqueue_residuals<6>(ctx, M); break;
        case 7: enqueue_residuals<7>(ctx, M); break;
        default: enqueue_residuals<8>(ctx, M); break;
    }
    BHIP_TRY(ctx, hipGetLastError());
    BHIP_TRY(ctx, hipMemcpyAsync(M.h_st, M.st.p, sizeof(MultiState), hipMemcpyDeviceToHost, ctx->stream));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    rr.assign(M.h_st->orr, M.h_st->orr + R);
    bb.assign(M.h_st->obb, M.h_st->obb + R);
    return BHIP_OK;
}

template <int R>
int solve_fused(bhip_ctx *ctx, MultiData &M, const bhip_solve_opts &o, bhip_stats &S) {
    hipStream_t s = ctx->stream;
    const long long n = ctx->n_dof;
    const double inner_tol = o.inner_tol > 0 ? o.inner_tol : 1e-3;
    const int max_outer = o.max_outer > 0 ? o.max_outer : 50;
    const int check_every = o.check_every > 0 ? o.check_every : 32;
    const int maxit = o.maxit > 0 ? o.maxit : 10000;
    std::vector<double> rr, bb, sc(R), rn(R);
    int total = 0, outer = 0;
    bool ok = false;
    const dim3 gp(grid_for(n, 256), R);
    for (outer = 0; outer <= max_outer; outer++) {
        BHIP_PROPAGATE(outer_residuals(ctx, M, R, rr, bb));
        S.spmv_launches += 1;
        bool all = true;
        for (int r = 0; r < R; r++) {
            const double rel = bb[r] > 0 ? sqrt(rr[r] / bb[r]) : sqrt(rr[r]);
            const bool conv = rel <= o.tol || rr[r] == 0.0;
            rn[r] = conv ? 0.0 : sqrt(rr[r]);
            sc[r] = conv ? 0.0 : 1.0 / rn[r];  // a converged case gets a zero inner RHS
            all &= conv;
        }
        if (all) {
            ok = true;
            break;
        }
        if (outer == max_outer || total >= maxit) break;
        BHIP_TRY(ctx, hipMemcpyAsync(M.scale.p, sc.data(), sizeof(double) * R, hipMemcpyHostToDevice, s));
        BHIP_TRY(ctx, hipMemcpyAsync(M.rn.p, rn.data(), sizeof(double) * R, hipMemcpyHostToDevice, s));
        k_pack_f32<<<gp, 256, 0, s>>>(n, M.r64.p, M.scale.p, M.r32.p, M.d32.p);
        int it = 0;
        BHIP_PROPAGATE(inner_solve<R>(ctx, M, inner_tol, maxit - total, check_every, &it, &S.spmv_launches));
        total += it;
        k_unpack_axpy<<<gp, 256, 0, s>>>(n, M.rn.p, M.d32.p, M.x64.p);
        BHIP_TRY(ctx, hipGetLastError());
    }
    S.iters = total;
    S.outer_iters = outer;
    return ok ? BHIP_OK : BHIP_ERR_NOT_CONVERGED;
}

}  // namespace

extern "C" int bhip_solve_cases(bhip_ctx *ctx, int nrhs, const double *presc, const double *load, double *x,
                                const bhip_solve_opts *opts, bhip_stats *stats) {
    if (!ctx) return BHIP_ERR_ARG;
    if (!presc || !x || !opts) return bhip_fail(ctx, BHIP_ERR_ARG, "presc, x and opts are required");
    if (nrhs < 1 || nrhs > kMaxR) return bhip_fail(ctx, BHIP_ERR_ARG, "nrhs must be in [1, 8]");
    if (!(opts->tol > 0)) return bhip_fail(ctx, BHIP_ERR_ARG, "tol must be positive");
    if (opts->precision != BHIP_PREC_FP64 && opts->precision != BHIP_PREC_FP32_IR)
        return bhip_fail(ctx, BHIP_ERR_ARG, "unknown precision");
    if (!ctx->assembled) return bhip_fail(ctx, BHIP_ERR_STATE, "matrix not assembled");
    (void)hipSetDevice(ctx->device);
    hipStream_t s = ctx->stream;
    const long long n = ctx->n_dof;
    bhip_stats S;
    memset(&S, 0, sizeof(S));
    const bool want32 = opts->precision == BHIP_PREC_FP32_IR;
    BHIP_PROPAGATE(bhip_prepare_solver(ctx, want32));
    const bool fused = want32 && ctx->use_bsr && ctx->sell_b == 3;
    BHIP_PROPAGATE(ensure_multi(ctx, nrhs, fused));
    MultiData &M = *ctx->multi;
    BHIP_TRY(ctx, hipEventRecord(ctx->drv0, s));
    // right-hand sides: each case's prescribed values lifted through the
    // coupling blocks of the assembled K (M.r64 stages the prescribed values,
    // the context's q vector the load)
    const double *dload = nullptr;
    if (load) {
        BHIP_TRY(ctx, hipMemcpyAsync(ctx->q.p, load, sizeof(double) * n, hipMemcpyHostToDevice, s));
        dload = ctx->q.p;
    }
    BHIP_TRY(ctx, hipMemcpyAsync(M.r64.p, presc, sizeof(double) * n * nrhs, hipMemcpyHostToDevice, s));
    BHIP_PROPAGATE(bhip_lift_rhs(ctx, nrhs, M.r64.p, dload, M.b64.p));
    BHIP_TRY(ctx, hipMemsetAsync(M.x64.p, 0, sizeof(double) * n * nrhs, s));
    int st = BHIP_OK;
    if (fused) {
        switch (nrhs) {
            case 1: st = solve_fused<1>(ctx, M, *opts, S); break;
            case 2: st = solve_fused<2>(ctx, M, *opts, S); break;
            case 3: st = solve_fused<3>(ctx, M, *opts, S); break;
            case 4: st = solve_fused<4>(ctx, M, *opts, S); break;
            case 5: st = solve_fused<5>(ctx, M, *opts, S); break;
            case 6: st = solve_fused<6>(ctx, M, *opts, S); break;
            case 7: st = solve_fused<7>(ctx, M, *opts, S); break;
            default: st = solve_fused<8>(ctx, M, *opts, S); break;
        }
        if (st != BHIP_OK && st != BHIP_ERR_NOT_CONVERGED) return st;
    } else {
        // plain FP64, or a mesh without 3 DOFs on every node: the cases run in
        // turn through the single-case solver; the context's RHS and solution
        // are restored afterwards
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
            S.outer_iters = std::max(S.outer_iters, s1.outer_iters);
            S.spmv_launches += s1.spmv_launches;
            BHIP_TRY(ctx, hipMemcpyAsync(M.x64.p + r * n, ctx->x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        }
        BHIP_TRY(ctx, hipMemcpyAsync(ctx->rhs.p, M.save.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
        BHIP_TRY(ctx, hipMemcpyAsync(ctx->x.p, M.save.p + n, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    }
    BHIP_TRY(ctx, hipEventRecord(ctx->drv1, s));
    BHIP_TRY(ctx, hipEventSynchronize(ctx->drv1));
    float ms = 0;
    BHIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->drv0, ctx->drv1));
    S.ms_total = ms;
    // true relative residual of every case in FP64; the worst is reported
    std::vector<double> rr, bb;
    BHIP_PROPAGATE(outer_residuals(ctx, M, nrhs, rr, bb));
    for (int r = 0; r < nrhs; r++) S.rel_resid = std::max(S.rel_resid, bb[r] > 0 ? sqrt(rr[r] / bb[r]) : sqrt(rr[r]));
    BHIP_TRY(ctx, bhip_copy(ctx, x, M.x64.p, sizeof(double) * n * nrhs, hipMemcpyDeviceToHost));
    if (stats) *stats = S;
    if (st == BHIP_ERR_NOT_CONVERGED)
        return bhip_fail(ctx, st, "multi-case solve not converged within the iteration caps");
    return BHIP_OK;
}
