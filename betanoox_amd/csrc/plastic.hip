// plastic.hip -- small-strain J2 plasticity on the device.
//
// Reference: PlasticityModel::DoPlasticAnalysis (PlasticityModel.cpp:142-398)
// with PlasticElement::update (plasticElement.cpp:39-250) actions
//   "V": strain increment at every Gauss point -> PlasticMaterial::updateISV
//        -> getStressIncrement (plastic.cpp:368-465, the active "version 2.0")
//   "KI": tangent stiffness from getTangentStiffness (plastic.cpp:583-610)
//   "F": internal force of the stress increment (addToFe with isv.dStress)
//   "U": GaussISV::updateISV (plastic.cpp:769-785)
// and the Newton loop of one load increment (PlasticityModel.cpp:213-366,
// Equations::incrementalSolutionUsingResidual, equation.cpp:1638-1720).
//
// J2 is Hill with F=G=H=1/3, L=M=N=1; linear isotropic hardening is the
// POWER master curve with m_E = H, m_n = 1, m_sigma0 = sigma0.  The return
// map, yield function and tangent run in FP64 at every Gauss point.
//
// Per Gauss point record (16 doubles): stress[6], eps_bar, dstress[6],
// deps_bar, plastic flag, return-map iterations.  Tangent (21 doubles, the
// upper triangle of a symmetric 6x6 row by row, read by the assembly
// kernel): the elastic C at an elastic point; at a plastic point either the
// reference's continuum C_ep = C - sp sp^T / denom, sp = C nf
// (getTangentStiffness, plastic.cpp:583-610), or, with
// BHIP_TANGENT_ALGORITHMIC, the symmetric part of the exact derivative of the
// stress update as executed (forward-mode differentiation through its
// iterations), which gives Newton quadratic convergence where the update is
// smooth.
#include <algorithm>
#include <cmath>

#include "bhip_internal.hpp"
#include "gauss_tables.hpp"


struct PlasticData {
    int nip = 0;
    int tangent = BHIP_TANGENT_CONTINUUM;
    dvec<double> isv;     // n_el * nip * 16
    dvec<double> tan;     // n_el * nip * kTanVals
    dvec<double> du;      // increment displacement (currentSolution), n_dof
    dvec<double> utot;    // accumulated displacement, n_dof
    dvec<double> corr;    // Newton correction, n_dof
    dvec<double> red;     // reductions
    dvec<unsigned> cnt;
    dvec<double> fe;      // element internal forces of the residual, 24 x n_el (entry-major)
};

namespace {

constexpr int kIsv = 16;

struct J2 {
    double E, nu, s0, H;
    double lam, mu;  // isotropic C: lam + 2 mu on the normal diagonal, lam off it, mu for shear
    __device__ void init(const double *p) {
        E = p[0];
        nu = p[1];
        s0 = p[2];
        H = p[3];
        lam = E * nu / ((1 + nu) * (1 - 2 * nu));
        mu = E / (2 * (1 + nu));
    }
    // entry (i, j) of the 6x6 C (Voigt xx yy zz xy yz xz, engineering shear)
    __device__ __forceinline__ double c(int i, int j) const {
        if (i < 3 && j < 3) return i == j ? lam + 2 * mu : lam;
        return i == j ? mu : 0.0;
    }
};

// HillFunction (plastic.cpp:172-194) with originState = yStress = 0.
__device__ __forceinline__ double hill(const double *s0, const double *ds) {
    double s[6];
    for (int i = 0; i < 6; i++) s[i] = s0[i] + ds[i];
    const double F = 1.0 / 3.0, G = 1.0 / 3.0, H = 1.0 / 3.0;
    return 0.5 * (F * (s[1] - s[2]) * (s[1] - s[2]) + G * (s[2] - s[0]) * (s[2] - s[0]) +
                  H * (s[0] - s[1]) * (s[0] - s[1]) + 2 * s[4] * s[4] + 2 * s[5] * s[5] + 2 * s[3] * s[3]);
}

// dYield (plastic.cpp:716-735)
__device__ __forceinline__ void dyield(const double *s0, const double *ds, double *nf) {
    double s[6];
    for (int i = 0; i < 6; i++) s[i] = s0[i] + ds[i];
    const double F = 1.0 / 3.0, G = 1.0 / 3.0, H = 1.0 / 3.0;
    nf[0] = -G * (s[2] - s[0]) + H * (s[0] - s[1]);
    nf[1] = F * (s[1] - s[2]) - H * (s[0] - s[1]);
    nf[2] = -F * (s[1] - s[2]) + G * (s[2] - s[0]);
    nf[3] = 2 * s[3];
    nf[4] = 2 * s[4];
    nf[5] = 2 * s[5];
}

// POWER master curve with m_n = 1: m_Stress = H*ep + s0, m_dsdep = H.
__device__ __forceinline__ double yield_stress(const J2 &m, double ep) {
    double y = m.H * ep + m.s0;
    return y > m.s0 ? y : m.s0;
}

__device__ __forceinline__ double yield_fn(const J2 &m, const double *s0, const double *ds, double ep) {
    double sy = yield_stress(m, ep);
    return hill(s0, ds) / (sy * sy / 3.0) - 1.0;
}

__device__ __forceinline__ double hard_mod(const J2 &m, const double *s0, const double *ds) {
    double mm2 = 3.0 * hill(s0, ds);  // effective stress squared
    return 4.0 / 9.0 * mm2 * m.H;
}

// o = C v, each row summed over its nonzero entries in column order (the
// zero entries of the isotropic C add nothing)
__device__ __forceinline__ void cmul(const J2 &m, const double *v, double *o) {
    const double d = m.lam + 2 * m.mu;
    o[0] = d * v[0];
    o[0] += m.lam * v[1];
    o[0] += m.lam * v[2];
    o[1] = m.lam * v[0];
    o[1] += d * v[1];
    o[1] += m.lam * v[2];
    o[2] = m.lam * v[0];
    o[2] += m.lam * v[1];
    o[2] += d * v[2];
    o[3] = m.mu * v[3];
    o[4] = m.mu * v[4];
    o[5] = m.mu * v[5];
}

__device__ __forceinline__ double dot6(const double *a, const double *b) {
    double s = 0;
#pragma unroll
    for (int i = 0; i < 6; i++) s += a[i] * b[i];
    return s;
}

__device__ __forceinline__ int group_of(const int *elem_group, int n_groups, int n_el, int e) {
    if (elem_group) return elem_group[e];
    return n_groups == n_el ? e : 0;
}

// dYield applied to any 6-vector (it is linear in the stress).
__device__ __forceinline__ void flow_lin(const double *s, double *nf) {
    const double zero[6] = {0, 0, 0, 0, 0, 0};
    dyield(zero, s, nf);
}

// "V": strain increment + return map (getStressIncrement, plastic.cpp:368-465).
// alg: write the symmetric part of the return map's derivative as the tangent.
__global__ void __launch_bounds__(128) k_plastic_update(int n_el, int nip, const double *__restrict__ gtab, const int *conn, const double *xyz,
                                 const double *mat, const int *elem_group, int n_groups, const double *du,
                                 double *isv, double *tan, int alg) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long long)n_el * nip) return;
    int e = (int)(t / nip), ip = (int)(t - (long long)e * nip);
    int ec[8];
#pragma unroll
    for (int i = 0; i < 8; i++) ec[i] = conn[(long long)e * 8 + i];
    double x[8], y[8], z[8];
    load_coords(xyz, ec, x, y, z);
    double g[8][3];
    hex8_all_grads(gtab + kGaussDN, x, y, z, ip, g);
    // ElasticityElement3D::calculateStrain (engineering shear)
    double de[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int a = 0; a < 8; a++) {
        const double u = du[3 * (long long)ec[a]], v = du[3 * (long long)ec[a] + 1],
                     w = du[3 * (long long)ec[a] + 2];
        de[0] += g[a][0] * u;
        de[1] += g[a][1] * v;
        de[2] += g[a][2] * w;
        de[3] += g[a][1] * u + g[a][0] * v;
        de[4] += g[a][2] * v + g[a][1] * w;
        de[5] += g[a][2] * u + g[a][0] * w;
    }
    J2 m;
    m.init(mat + 4LL * group_of(elem_group, n_groups, n_el, e));
    double *rec = isv + t * kIsv;
    double s0[6], ds[6], dse[6], nf[6], sp[6];
    for (int i = 0; i < 6; i++) s0[i] = rec[i];
    const double ep = rec[6];
    cmul(m, de, dse);  // elastic predictor
    for (int i = 0; i < 6; i++) ds[i] = dse[i];
    double dep = 0.0;
    int iter = 0, plastic = 0;
    double inv_denom = 0.0;
    // D = d(ds)/d(de), column-major by strain component (alg only)
    double D[36];
#pragma unroll
    for (int i = 0; i < 36; i++) D[i] = m.c(i % 6, i / 6);
    if (yield_fn(m, s0, ds, ep + dep) > 0.0) {
        plastic = 1;
        double f;
        do {
            double hard = hard_mod(m, s0, ds);
            dyield(s0, ds, nf);
            cmul(m, nf, sp);
            const double den = hard + dot6(nf, sp);
            double dl = dot6(nf, dse) / den;
            if (alg) {
                // per strain component j, with s = s0 + ds:
                //   d nf = P D_j, d num = dse.(P D_j) + nf.C_j,
                //   d den = (4/3) H nf.D_j + 2 sp.(P D_j),
                //   D_j' = C_j - sp d dl - dl C P D_j
#pragma unroll
                for (int j = 0; j < 6; j++) {
                    double pc[6], cpc[6];
                    flow_lin(D + 6 * j, pc);
                    cmul(m, pc, cpc);
                    double dnum = 0, dden = 0;
#pragma unroll
                    for (int i = 0; i < 6; i++) {
                        dnum += dse[i] * pc[i] + nf[i] * m.c(i, j);
                        dden += 4.0 / 3.0 * m.H * nf[i] * D[6 * j + i] + 2.0 * sp[i] * pc[i];
                    }
                    const double ddl = (dnum - dl * dden) / den;
#pragma unroll
                    for (int i = 0; i < 6; i++) D[6 * j + i] = m.c(i, j) - sp[i] * ddl - dl * cpc[i];
                }
            }
            for (int i = 0; i < 6; i++) ds[i] = dse[i] - dl * sp[i];
            dep = 2.0 * dl * sqrt(3.0 * hill(s0, ds)) / 3.0;
            f = yield_fn(m, s0, ds, ep + dep);
            iter++;
        } while (f > 1e-6 && iter < 5000);
        dyield(s0, ds, nf);
        cmul(m, nf, sp);
        inv_denom = 1.0 / (hard_mod(m, s0, ds) + dot6(nf, sp));
    }
    for (int i = 0; i < 6; i++) rec[7 + i] = ds[i];
    rec[13] = dep;
    rec[14] = plastic;
    rec[15] = iter;
    double *tr = tan + t * kTanVals;
    int k = 0;
#pragma unroll
    for (int i = 0; i < 6; i++)
#pragma unroll
        for (int j = i; j < 6; j++, k++) {
            double v = m.c(i, j);
            if (plastic) v = alg ? 0.5 * (D[6 * j + i] + D[6 * i + j]) : v - sp[i] * sp[j] * inv_denom;
            tr[k] = v;
        }
}

// "F", element stage: the internal force vector of element e,
// fe = sum over Gauss points of B^T dstress detJ w (ElasticityElement3D::addToFe:
// force += sum_j b[j][i] stress[j] * f), one thread per element, entry k of
// element e at fe[k * n_el + e].
__global__ void k_plastic_fe(int n_el, int nip, const double *__restrict__ gtab, const int *conn, const double *xyz,
                             const double *isv, double *fe) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n_el) return;
    int ec[8];
#pragma unroll
    for (int i = 0; i < 8; i++) ec[i] = conn[e * 8 + i];
    double x[8], y[8], z[8];
    load_coords(xyz, ec, x, y, z);
    double f[24];
#pragma unroll
    for (int k = 0; k < 24; k++) f[k] = 0;
    for (int ip = 0; ip < nip; ip++) {
        double g[8][3];
        const double detJ = hex8_all_grads(gtab + kGaussDN, x, y, z, ip, g);
        const double w = detJ * gtab[ip];
        const double *s = isv + (e * nip + ip) * kIsv + 7;
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

// "F" assembled + calculateResidualNorm: per node, the resultant of its
// incident elements' force vectors added in ascending element order (the
// reference's element loop), residual = load - resultant on free DOFs.
__global__ void k_plastic_residual(int n_nodes, int n_el, const int *n2e_ptr, const int *n2e, const int *conn,
                                   const double *__restrict__ fe, const unsigned char *rmask, const double *load,
                                   double *resid) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    double acc[3] = {0, 0, 0};
    for (int k = n2e_ptr[n]; k < n2e_ptr[n + 1]; k++) {
        const long long e = n2e[k];
        // every local node of e that is n (a collapsed element lists a node twice)
#pragma unroll
        for (int a = 0; a < 8; a++)
            if (conn[e * 8 + a] == n)
#pragma unroll
                for (int c = 0; c < 3; c++) acc[c] += fe[(long long)(3 * a + c) * n_el + e];
    }
    unsigned rm = rmask[n];
    for (int c = 0; c < 3; c++) {
        long long I = 3LL * n + c;
        resid[I] = ((rm >> c) & 1u) ? 0.0 : load[I] - acc[c];
    }
}

// max |v| and count of plastic points / max return iterations (block partial
// maxima combined by the last block).
__global__ void k_max_abs(long long n, const double *v, double *red, unsigned *cnt, double *out) {
    __shared__ double sm[256];
    __shared__ bool last;
    double m = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        m = fmax(m, fabs(v[i]));
    sm[threadIdx.x] = m;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        red[blockIdx.x] = sm[0];
        __threadfence();
        last = atomicAdd(cnt, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (last && threadIdx.x == 0) {
        __threadfence();
        double r = 0;
        for (int b = 0; b < (int)gridDim.x; b++) r = fmax(r, red[b]);
        *out = r;
        *cnt = 0;
    }
}

// Plastic points (record slot 14 non-zero) and the largest return-map
// iteration count (slot 15) over all integration points: wave partials
// combined with one atomic per wave into out[0] (count) and out[1] (max),
// which the caller clears.  An unsigned count holds any point count whose
// 128-byte records fit in device memory.
__global__ void k_isv_stats(long long npts, const double *__restrict__ isv, unsigned *out) {
    unsigned c = 0, m = 0;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < npts; t += (long long)gridDim.x * blockDim.x) {
        const double *r = isv + t * kIsv;
        c += r[14] != 0.0;
        m = max(m, (unsigned)r[15]);
    }
    for (int o = 32; o > 0; o >>= 1) {
        c += __shfl_xor(c, o);
        m = max(m, (unsigned)__shfl_xor(m, o));
    }
    if ((threadIdx.x & 63) == 0) {
        if (c) atomicAdd(out, c);
        atomicMax(out + 1, m);
    }
}

// currentSolution at the start of an increment: prescribed values on
// restrained DOFs, zero elsewhere (equations.zeroSolution + loads applyTo).
__global__ void k_du_start(int n, const unsigned char *rest, const double *presc, double *du) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) du[i] = rest[i] ? presc[i] : 0.0;
}

// addIncrementalDisplacements (equation.cpp:1707-1720): free DOFs only.
__global__ void k_add_free(int n, const unsigned char *rest, const double *d, double *du) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !rest[i]) du[i] += d[i];
}

__global__ void k_add(int n, const double *a, double *b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] += a[i];
}

// "U": GaussISV::updateISV
__global__ void k_plastic_commit(long long npts, double *isv) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= npts) return;
    double *r = isv + t * kIsv;
    for (int i = 0; i < 6; i++) {
        r[i] += r[7 + i];
        r[7 + i] = 0.0;
    }
    r[6] += r[13];
    r[13] = 0.0;
}

int plastic_check(bhip_ctx *ctx) {
    if (!ctx->plastic) return bhip_fail(ctx, BHIP_ERR_STATE, "plastic state not initialised (bhip_plastic_init)");
    int nip = 0;
    BHIP_PROPAGATE(bhip_gauss_upload(ctx, 8, ctx->order, &nip));
    if (nip != ctx->plastic->nip) return bhip_fail(ctx, BHIP_ERR_STATE, "integration order changed after bhip_plastic_init");
    return BHIP_OK;
}

// alg: the algorithmic tangent for this pass (the first pass of an
// increment, with only the prescribed increment applied, takes the
// continuum one: from that trial state the derivative softens every
// deviatoric direction of the boundary layer and sends Newton far off).
int plastic_update_dev(bhip_ctx *ctx, const double *du, bool alg = false) {
    PlasticData *P = ctx->plastic;
    long long npts = (long long)ctx->n_el * P->nip;
    k_plastic_update<<<grid_for(npts, 128), 128, 0, ctx->stream>>>(
        ctx->n_el, P->nip, ctx->gauss.p, ctx->conn.p, ctx->xyz.p, ctx->mat.p, ctx->elem_group.n ? ctx->elem_group.p : nullptr,
        ctx->n_groups, du, P->isv.p, P->tan.p, alg ? 1 : 0);
    BHIP_TRY(ctx, hipGetLastError());
    return BHIP_OK;
}

int plastic_residual_dev(bhip_ctx *ctx, double *max_abs) {
    PlasticData *P = ctx->plastic;
    k_plastic_fe<<<grid_for(ctx->n_el, 128), 128, 0, ctx->stream>>>(ctx->n_el, P->nip, ctx->gauss.p, ctx->conn.p,
                                                                     ctx->xyz.p, P->isv.p, P->fe.p);
    k_plastic_residual<<<grid_for(ctx->n_nodes, 128), 128, 0, ctx->stream>>>(
        ctx->n_nodes, ctx->n_el, ctx->n2e_ptr.p, ctx->n2e.p, ctx->conn.p, P->fe.p, ctx->node_rmask.p, ctx->load.p,
        ctx->rhs.p);
    int g = std::min(1024, grid_for(ctx->n_dof, 256));
    k_max_abs<<<g, 256, 0, ctx->stream>>>(ctx->n_dof, ctx->rhs.p, P->red.p, P->cnt.p, P->red.p + 1024);
    BHIP_TRY(ctx, hipGetLastError());
    BHIP_TRY(ctx, hipMemcpyAsync(max_abs, P->red.p + 1024, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BHIP_OK;
}

}  // namespace

const double *bhip_plastic_tangent(bhip_ctx *ctx) { return ctx->plastic ? ctx->plastic->tan.p : nullptr; }

void bhip_plastic_release(bhip_ctx *ctx) {
    if (!ctx->plastic) return;
    PlasticData *P = ctx->plastic;
    P->isv.release();
    P->tan.release();
    P->du.release();
    P->utot.release();
    P->corr.release();
    P->fe.release();
    P->red.release();
    P->cnt.release();
    delete P;
    ctx->plastic = nullptr;
}

extern "C" {

int bhip_plastic_init(bhip_ctx *ctx) {
    if (!ctx) return BHIP_ERR_ARG;
    if (!ctx->pattern_ready || !ctx->have_mat) return bhip_fail(ctx, BHIP_ERR_STATE, "pattern and materials first");
    if (!(ctx->nnpe == 8 && ctx->dim == 3 && ctx->mat_kind == BHIP_MAT_J2 && ctx->uniform_ndof &&
          ctx->max_ndof == 3))
        return bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "plasticity needs hex8, 3 DOF/node and a J2 material");
    (void)hipSetDevice(ctx->device);
    bhip_plastic_release(ctx);
    PlasticData *P = new PlasticData();
    ctx->plastic = P;
    P->nip = ctx->order * ctx->order * ctx->order;
    long long npts = (long long)ctx->n_el * P->nip;
    BHIP_TRY(ctx, P->isv.resize(npts * kIsv));
    BHIP_TRY(ctx, P->tan.resize(npts * kTanVals));
    BHIP_TRY(ctx, P->du.resize(ctx->n_dof));
    BHIP_TRY(ctx, P->utot.resize(ctx->n_dof));
    BHIP_TRY(ctx, P->corr.resize(ctx->n_dof));
    BHIP_TRY(ctx, P->red.resize(1025));
    BHIP_TRY(ctx, P->fe.resize((size_t)24 * (ctx->n_el > 0 ? ctx->n_el : 1)));
    BHIP_TRY(ctx, P->cnt.resize(3));
    BHIP_TRY(ctx, hipMemsetAsync(P->isv.p, 0, P->isv.bytes(), ctx->stream));
    BHIP_TRY(ctx, hipMemsetAsync(P->tan.p, 0, P->tan.bytes(), ctx->stream));
    BHIP_TRY(ctx, hipMemsetAsync(P->du.p, 0, P->du.bytes(), ctx->stream));
    BHIP_TRY(ctx, hipMemsetAsync(P->utot.p, 0, P->utot.bytes(), ctx->stream));
    BHIP_TRY(ctx, hipMemsetAsync(P->cnt.p, 0, 3 * sizeof(unsigned), ctx->stream));
    BHIP_PROPAGATE(bhip_ensure_vectors(ctx));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BHIP_OK;
}

int bhip_plastic_set_tangent(bhip_ctx *ctx, int tangent) {
    if (!ctx) return BHIP_ERR_ARG;
    BHIP_PROPAGATE(plastic_check(ctx));
    if (tangent != BHIP_TANGENT_CONTINUUM && tangent != BHIP_TANGENT_ALGORITHMIC)
        return bhip_fail(ctx, BHIP_ERR_ARG, "unknown tangent");
    ctx->plastic->tangent = tangent;
    return BHIP_OK;
}

int bhip_plastic_update(bhip_ctx *ctx, const double *du) {
    if (!ctx || !du) return BHIP_ERR_ARG;
    BHIP_PROPAGATE(plastic_check(ctx));
    (void)hipSetDevice(ctx->device);
    PlasticData *P = ctx->plastic;
    BHIP_TRY(ctx, hipMemcpyAsync(P->du.p, du, sizeof(double) * ctx->n_dof, hipMemcpyHostToDevice, ctx->stream));
    BHIP_PROPAGATE(plastic_update_dev(ctx, P->du.p, P->tangent == BHIP_TANGENT_ALGORITHMIC));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BHIP_OK;
}

int bhip_plastic_residual(bhip_ctx *ctx, double *residual, double *max_abs) {
    if (!ctx || !max_abs) return BHIP_ERR_ARG;
    BHIP_PROPAGATE(plastic_check(ctx));
    (void)hipSetDevice(ctx->device);
    BHIP_PROPAGATE(plastic_residual_dev(ctx, max_abs));
    if (residual) BHIP_TRY(ctx, bhip_copy(ctx, residual, ctx->rhs.p, sizeof(double) * ctx->n_dof, hipMemcpyDeviceToHost));
    return BHIP_OK;
}

int bhip_plastic_commit(bhip_ctx *ctx) {
    if (!ctx) return BHIP_ERR_ARG;
    BHIP_PROPAGATE(plastic_check(ctx));
    PlasticData *P = ctx->plastic;
    long long npts = (long long)ctx->n_el * P->nip;
    k_plastic_commit<<<grid_for(npts, 256), 256, 0, ctx->stream>>>(npts, P->isv.p);
    k_add<<<grid_for(ctx->n_dof, 256), 256, 0, ctx->stream>>>(ctx->n_dof, P->du.p, P->utot.p);
    BHIP_TRY(ctx, hipGetLastError());
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BHIP_OK;
}

int bhip_plastic_increment(bhip_ctx *ctx, const double *presc_increment, double max_allowable_residual,
                           int max_iterations, const bhip_solve_opts *opts, bhip_plastic_stats *stats) {
    if (!ctx || !opts) return BHIP_ERR_ARG;
    BHIP_PROPAGATE(plastic_check(ctx));
    (void)hipSetDevice(ctx->device);
    PlasticData *P = ctx->plastic;
    hipStream_t s = ctx->stream;
    const int n = ctx->n_dof;
    bhip_plastic_stats S;
    memset(&S, 0, sizeof(S));
    BHIP_TRY(ctx, hipEventRecord(ctx->drv0, s));
    if (presc_increment)
        BHIP_TRY(ctx, hipMemcpyAsync(ctx->presc.p, presc_increment, sizeof(double) * n, hipMemcpyHostToDevice, s));
    // currentSolution = prescribed increment on restrained DOFs
    k_du_start<<<grid_for(n, 256), 256, 0, s>>>(n, ctx->rest.p, ctx->presc.p, P->du.p);
    BHIP_TRY(ctx, hipGetLastError());
    bhip_solve_opts o = *opts;
    // use_x0: every Newton correction starts from the previous correction
    // (left in ctx->x by the last solve, zero before the first); otherwise
    // from zero
    const bool warm = opts->use_x0 != 0;
    o.use_x0 = 0;
    double maxres = 0;
    int it = 0;
    for (;;) {
        // "V" then assembleKandF ("KI") then getGlobalForces ("F")
        BHIP_PROPAGATE(plastic_update_dev(ctx, P->du.p, P->tangent == BHIP_TANGENT_ALGORITHMIC && it > 0));
        BHIP_PROPAGATE(bhip_assemble_impl(ctx, BHIP_ASM_PLASTIC_TANGENT, nullptr, nullptr));
        BHIP_PROPAGATE(plastic_residual_dev(ctx, &maxres));
        if (it > 0 && maxres < 2.0 * max_allowable_residual) break;
        // one solve before the loop plus max_iterations inside it (PlasticityModel.cpp:262-287)
        if (it > max_iterations) break;
        // K d = residual (rhs already holds it), then du += d on free DOFs
        bhip_stats st;
        if (warm) {
            BHIP_TRY(ctx, hipMemcpyAsync(ctx->x0.p, ctx->x.p, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
            o.use_x0 = 1;
        }
        int rc = bhip_solve_impl(ctx, &o, &st);
        if (rc != BHIP_OK) return rc;
        S.pcg_iters += st.iters;
        k_add_free<<<grid_for(n, 256), 256, 0, s>>>(n, ctx->rest.p, ctx->x.p, P->du.p);
        BHIP_TRY(ctx, hipGetLastError());
        it++;
    }
    S.newton_iters = it;
    S.max_residual = maxres;
    BHIP_PROPAGATE(bhip_plastic_commit(ctx));
    // plastic points and return-map iterations of the converged state
    long long npts = (long long)ctx->n_el * P->nip;
    BHIP_TRY(ctx, hipMemsetAsync(P->cnt.p + 1, 0, 2 * sizeof(unsigned), s));
    k_isv_stats<<<grid_for(npts, 256, ctx->n_cu * 8), 256, 0, s>>>(npts, P->isv.p, P->cnt.p + 1);
    BHIP_TRY(ctx, hipGetLastError());
    unsigned isv_stats[2];
    BHIP_TRY(ctx, hipMemcpyAsync(isv_stats, P->cnt.p + 1, sizeof(isv_stats), hipMemcpyDeviceToHost, s));
    BHIP_TRY(ctx, hipEventRecord(ctx->drv1, s));
    BHIP_TRY(ctx, hipEventSynchronize(ctx->drv1));
    float ms = 0;
    BHIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->drv0, ctx->drv1));
    S.ms = ms;
    S.plastic_points = isv_stats[0];
    S.max_return_iters = (int)isv_stats[1];
    if (stats) *stats = S;
    if (!(maxres < 2.0 * max_allowable_residual))
        return bhip_fail(ctx, BHIP_ERR_NOT_CONVERGED, "Newton iterations exhausted in the load increment");
    return BHIP_OK;
}

int bhip_plastic_get_state(bhip_ctx *ctx, double *stress, double *eps_bar, double *u_total) {
    if (!ctx) return BHIP_ERR_ARG;
    BHIP_PROPAGATE(plastic_check(ctx));
    PlasticData *P = ctx->plastic;
    long long npts = (long long)ctx->n_el * P->nip;
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (stress || eps_bar) {
        std::vector<double> rec(npts * kIsv);
        BHIP_TRY(ctx, bhip_copy(ctx, rec.data(), P->isv.p, sizeof(double) * npts * kIsv, hipMemcpyDeviceToHost));
        for (long long t = 0; t < npts; t++) {
            if (stress)
                for (int i = 0; i < 6; i++) stress[t * 6 + i] = rec[t * kIsv + i];
            if (eps_bar) eps_bar[t] = rec[t * kIsv + 6];
        }
    }
    if (u_total) BHIP_TRY(ctx, bhip_copy(ctx, u_total, P->utot.p, sizeof(double) * ctx->n_dof, hipMemcpyDeviceToHost));
    return BHIP_OK;
}

}  // extern "C"
