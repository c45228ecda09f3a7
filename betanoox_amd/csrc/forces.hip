// forces.hip -- internal forces of the elastic hex8 model on the device.
//
// Reference: BasicModel::getGlobalForces (BasicModel.cpp:1063-1110) calls
// update("F", u) on every element (ElasticityElement3D::addToFe: stress
// C B u_e at each Gauss point, f_e += B^T stress detJ w) and adds each f_e
// into the internal force of all DOFs (restrained ones included: there it is
// the reaction) and, through Equations::assembleResultantVector
// (equation.cpp:790-807), into the resultant of the free DOFs;
// calculateResidualNorm (equation.cpp:1516-1529) forms load - resultant on
// free DOFs.  MultiFieldMacroElement::solveForStiffnessMatrixColumn ends with
// it (MultiFieldMacroElement.cpp:217-219): the reactions of a load case are
// a column of the macro-element stiffness.
//
// Device form, as the plasticity residual: one thread per element integrates
// f_e (24 values, stored entry-major), then one thread per node adds its
// incident elements' entries in ascending element order -- the reference's
// element loop order -- so no atomics are needed.
#include <algorithm>
#include <cmath>

#include "bhip_internal.hpp"

namespace {

// f_e = sum_ip B^T (C B u_e) detJ w for element e (isotropic C from E, nu,
// or a full 6x6 C), entry k at fe[k * n_el + e].
template <int KIND>
__global__ void k_elastic_fe(int n_el, int nip, const double *__restrict__ gtab, const int *conn, const double *xyz,
                             const double *mat, int mat_stride, const int *elem_group, int n_groups, const double *u,
                             double *fe) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    int ec[8];
#pragma unroll
    for (int i = 0; i < 8; i++) ec[i] = conn[e * 8 + i];
    double x[8], y[8], z[8], ue[24];
    load_coords(xyz, ec, x, y, z);
#pragma unroll
    for (int a = 0; a < 8; a++)
#pragma unroll
        for (int c = 0; c < 3; c++) ue[3 * a + c] = u[3 * (long long)ec[a] + c];
    const int grp = elem_group ? elem_group[e] : (n_groups == n_el ? (int)e : 0);
    const double *mp = mat + (long long)grp * mat_stride;
    double C[36];
    if (KIND == BHIP_MAT_CMAT3D) {
        for (int i = 0; i < 36; i++) C[i] = mp[i];
    } else {
        const double E = mp[0], nu = mp[1];
        const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
        for (int i = 0; i < 36; i++) C[i] = 0;
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) C[i * 6 + j] = lam;
            C[i * 6 + i] = lam + 2 * mu;
            C[(i + 3) * 6 + (i + 3)] = mu;
        }
    }
    double f[24];
#pragma unroll
    for (int k = 0; k < 24; k++) f[k] = 0;
    for (int ip = 0; ip < nip; ip++) {
        double g[8][3];
        const double w = hex8_all_grads(gtab + kGaussDN, x, y, z, ip, g) * gtab[ip];
        // strain, Voigt xx yy zz xy yz xz, engineering shear
        double eps[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
        for (int a = 0; a < 8; a++) {
            const double uu = ue[3 * a], vv = ue[3 * a + 1], ww = ue[3 * a + 2];
            eps[0] += g[a][0] * uu;
            eps[1] += g[a][1] * vv;
            eps[2] += g[a][2] * ww;
            eps[3] += g[a][1] * uu + g[a][0] * vv;
            eps[4] += g[a][2] * vv + g[a][1] * ww;
            eps[5] += g[a][2] * uu + g[a][0] * ww;
        }
        double s[6];
        for (int i = 0; i < 6; i++) {
            double t = 0;
            for (int j = 0; j < 6; j++) t += C[i * 6 + j] * eps[j];
            s[i] = t;
        }
#pragma unroll
        for (int a = 0; a < 8; a++) {
            f[3 * a + 0] += (g[a][0] * s[0] + g[a][1] * s[3] + g[a][2] * s[5]) * w;
            f[3 * a + 1] += (g[a][1] * s[1] + g[a][0] * s[3] + g[a][2] * s[4]) * w;
            f[3 * a + 2] += (g[a][2] * s[2] + g[a][1] * s[4] + g[a][0] * s[5]) * w;
        }
    }
#pragma unroll
    for (int k = 0; k < 24; k++) fe[(long long)k * n_el + e] = f[k];
}

// Per node: internal force of its DOFs (every local index of the node in each
// incident element, elements in ascending order), residual = load - force on
// free DOFs and 0 on restrained ones.
__global__ void k_force_gather(int n_nodes, int n_el, const int *n2e_ptr, const int *n2e, const int *conn,
                               const double *__restrict__ fe, const unsigned char *rmask, const double *load,
                               double *force, double *resid) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    double acc[3] = {0, 0, 0};
    for (int k = n2e_ptr[n]; k < n2e_ptr[n + 1]; k++) {
        const long long e = n2e[k];
#pragma unroll
        for (int a = 0; a < 8; a++)
            if (conn[e * 8 + a] == n)
#pragma unroll
                for (int c = 0; c < 3; c++) acc[c] += fe[(long long)(3 * a + c) * n_el + e];
    }
    const unsigned rm = rmask[n];
    for (int c = 0; c < 3; c++) {
        const long long I = 3LL * n + c;
        force[I] = acc[c];
        resid[I] = ((rm >> c) & 1u) ? 0.0 : (load ? load[I] : 0.0) - acc[c];
    }
}

// max |v| over n entries: block maxima, then one block over them.
__global__ void k_block_max_abs(long long n, const double *v, double *out) {
    __shared__ double sm[256];
    double m = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = fmax(m, fabs(v[i]));
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[blockIdx.x] = sm[0];
}

}  // namespace

extern "C" int bhip_internal_force(bhip_ctx *ctx, int nvec, const double *u, double *force, double *residual,
                                   double *max_abs) {
    if (!ctx || nvec <= 0 || !u || !force) return BHIP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (ctx->external_matrix || ctx->n_nodes == 0) return bhip_fail(ctx, BHIP_ERR_STATE, "no mesh");
    if (!ctx->pattern_ready) return bhip_fail(ctx, BHIP_ERR_STATE, "pattern not built");
    if (!ctx->have_mat) return bhip_fail(ctx, BHIP_ERR_STATE, "materials not set");
    if (!(ctx->nnpe == 8 && ctx->dim == 3 && ctx->uniform_ndof && ctx->max_ndof == 3 &&
          (ctx->mat_kind == BHIP_MAT_ISO_ELASTIC || ctx->mat_kind == BHIP_MAT_CMAT3D)))
        return bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "internal forces are built for hex8 elasticity");
    BHIP_PROPAGATE(bhip_ensure_vectors(ctx));
    int nip = 0;
    BHIP_PROPAGATE(bhip_gauss_upload(ctx, 8, ctx->order, &nip));
    hipStream_t s = ctx->stream;
    const long long n = ctx->n_dof;
    scoped_dvec<double> fe, uvec, fvec, rvec, red;
    BHIP_TRY(ctx, fe.resize((size_t)24 * (ctx->n_el > 0 ? ctx->n_el : 1)));
    BHIP_TRY(ctx, uvec.resize(n));
    BHIP_TRY(ctx, fvec.resize(n));
    BHIP_TRY(ctx, rvec.resize(n));
    const int gm = std::min(1024, grid_for(n, 256));
    BHIP_TRY(ctx, red.resize(gm + 1));
    for (int v = 0; v < nvec; v++) {
        BHIP_TRY(ctx, hipMemcpyAsync(uvec.p, u + (size_t)v * n, sizeof(double) * n, hipMemcpyHostToDevice, s));
        const int *g = ctx->elem_group.n ? ctx->elem_group.p : nullptr;
        if (ctx->mat_kind == BHIP_MAT_CMAT3D)
            k_elastic_fe<BHIP_MAT_CMAT3D><<<grid_for(ctx->n_el, 128), 128, 0, s>>>(
                ctx->n_el, nip, ctx->gauss.p, ctx->conn.p, ctx->xyz.p, ctx->mat.p, ctx->mat_stride, g, ctx->n_groups,
                uvec.p, fe.p);
        else
            k_elastic_fe<BHIP_MAT_ISO_ELASTIC><<<grid_for(ctx->n_el, 128), 128, 0, s>>>(
                ctx->n_el, nip, ctx->gauss.p, ctx->conn.p, ctx->xyz.p, ctx->mat.p, ctx->mat_stride, g, ctx->n_groups,
                uvec.p, fe.p);
        k_force_gather<<<grid_for(ctx->n_nodes, 128), 128, 0, s>>>(ctx->n_nodes, ctx->n_el, ctx->n2e_ptr.p,
                                                                    ctx->n2e.p, ctx->conn.p, fe.p, ctx->node_rmask.p,
                                                                    ctx->load.p, fvec.p, rvec.p);
        BHIP_TRY(ctx, hipGetLastError());
        BHIP_TRY(ctx, hipMemcpyAsync(force + (size_t)v * n, fvec.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (residual)
            BHIP_TRY(ctx, hipMemcpyAsync(residual + (size_t)v * n, rvec.p, sizeof(double) * n, hipMemcpyDeviceToHost,
                                         s));
        if (max_abs) {
            k_block_max_abs<<<gm, 256, 0, s>>>(n, rvec.p, red.p);
            k_block_max_abs<<<1, 256, 0, s>>>(gm, red.p, red.p + gm);
            BHIP_TRY(ctx, hipGetLastError());
            BHIP_TRY(ctx, hipMemcpyAsync(max_abs + v, red.p + gm, sizeof(double), hipMemcpyDeviceToHost, s));
        }
        BHIP_TRY(ctx, hipStreamSynchronize(s));
    }
    return BHIP_OK;
}
