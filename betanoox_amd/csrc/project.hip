// project.hip -- fixed-order batched dot products and linear combinations over
// a small set of basis vectors: the building blocks of the projected initial
// guesses (heat steps: successive right-hand sides with one matrix; J2
// plasticity: Newton corrections with a tangent that changes every pass).
//
// Basis vectors are stored one after another (v_i at V + i n).  Every sum is a
// fixed-order reduction -- per-block shares summed in block order -- so a run
// repeats bit for bit.
#include "bhip_internal.hpp"

namespace {

constexpr int kT = 256;

// part[blk][i] = this block's share of v_i . y (i < m) and, when a is given,
// of a . y in slot kProjMax.
__global__ void __launch_bounds__(kT) k_proj_dots(long long n, int m, const double *__restrict__ V,
                                                const double *__restrict__ a, const double *__restrict__ y,
                                                double *__restrict__ part) {
    __shared__ double sm[16 * kProjNV];
    double acc[kProjNV];
#pragma unroll
    for (int i = 0; i < kProjNV; i++) acc[i] = 0;
    const long long per = (n + gridDim.x - 1) / gridDim.x;
    const long long i0 = (long long)blockIdx.x * per, i1 = i0 + per < n ? i0 + per : n;
    for (long long j = i0 + threadIdx.x; j < i1; j += blockDim.x) {
        const double yj = y[j];
#pragma unroll
        for (int i = 0; i < kProjMax; i++)
            if (i < m) acc[i] += V[(long long)i * n + j] * yj;
        if (a) acc[kProjMax] += a[j] * yj;
    }
    block_sum<kProjNV>(acc, sm);
    if (threadIdx.x == 0)
#pragma unroll
        for (int i = 0; i < kProjNV; i++) part[(long long)blockIdx.x * kProjNV + i] = acc[i];
}

// coef[i] = the sum of the nb block shares of slot i, in block order, and 0
// for slot `skip` (-1: none).  mode 1 (a new basis vector w = a - sum coef[i]
// v_i): coef[kProjMax] = 1 / sqrt(a.y - sum coef[i]^2), or 0 when that
// vanishes to rounding (the vector then adds nothing).
__global__ void __launch_bounds__(kT) k_proj_fin(int nb, int m, int skip, const double *__restrict__ part,
                                               double *__restrict__ coef, int mode) {
    __shared__ double sm[16 * kProjNV];
    double acc[kProjNV];
#pragma unroll
    for (int i = 0; i < kProjNV; i++) acc[i] = 0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
#pragma unroll
        for (int i = 0; i < kProjNV; i++) acc[i] += part[(long long)b * kProjNV + i];
    block_sum<kProjNV>(acc, sm);
    if (threadIdx.x == 0) {
        for (int i = 0; i < kProjMax; i++) coef[i] = (i < m && i != skip) ? acc[i] : 0.0;
        if (mode == 1) {
            const double ay = acc[kProjMax];
            double w2 = ay;
            for (int i = 0; i < m; i++) w2 -= coef[i] * coef[i];
            coef[kProjMax] = (ay > 0 && w2 > 1e-12 * ay) ? 1.0 / sqrt(w2) : 0.0;
        }
    }
}

// out = scale * (base + sign * sum_{i<m} coef[i] v_i); scale = coef[kProjMax]
// when use_scale, else 1; base may be null.  A thread reads every v_i at its
// index before it writes out there, so out may be one of the v_i whose
// coefficient is 0.
__global__ void k_proj_combo(long long n, int m, const double *__restrict__ V, const double *__restrict__ coef,
                             const double *__restrict__ base, double sign, int use_scale, double *out) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    double t = 0;
    for (int i = 0; i < m; i++) t += coef[i] * V[(long long)i * n + j];
    const double v = (base ? base[j] : 0.0) + sign * t;
    out[j] = use_scale ? coef[kProjMax] * v : v;
}

}  // namespace

void bhip_proj_dots(bhip_ctx *ctx, long long n, int m, const double *V, const double *a, const double *y, int skip,
                    int mode, double *part, double *coef) {
    k_proj_dots<<<kProjBlocks, kT, 0, ctx->stream>>>(n, m, V, a, y, part);
    k_proj_fin<<<1, kT, 0, ctx->stream>>>(kProjBlocks, m, skip, part, coef, mode);
}

void bhip_proj_combo(bhip_ctx *ctx, long long n, int m, const double *V, const double *coef, const double *base,
                     double sign, int use_scale, double *out) {
    k_proj_combo<<<grid_for(n, kT), kT, 0, ctx->stream>>>(n, m, V, coef, base, sign, use_scale, out);
}
