// capi.hip -- extern "C" entry points of libbeta_hip.so (declared in
// include/beta_hip.h).  Argument checking, host<->device copies and state
// bookkeeping; the kernels live in pattern.hip, assemble.hip, solver.hip and
// plastic.hip.
#include <algorithm>
#include <cmath>

#include "bhip_internal.hpp"

int bhip_fail(bhip_ctx *ctx, int status, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return status;
}

int bhip_ensure_vectors(bhip_ctx *ctx) {
    size_t n = ctx->n_dof > 0 ? (size_t)ctx->n_dof : 1;
    dvec<double> *v64[] = {&ctx->load, &ctx->pseudo, &ctx->rhs, &ctx->x, &ctx->x0,
                           &ctx->r,    &ctx->p,      &ctx->q,   &ctx->dinv, &ctx->tmp};
    for (auto *v : v64)
        if (v->n != n) {
            BHIP_TRY(ctx, v->resize(n));
            BHIP_TRY(ctx, hipMemsetAsync(v->p, 0, v->bytes(), ctx->stream));
        }
    dvec<float> *v32[] = {&ctx->r32, &ctx->p32, &ctx->q32, &ctx->d32, &ctx->dinv32};
    for (auto *v : v32)
        if (v->n != n) {
            BHIP_TRY(ctx, v->resize(n));
            BHIP_TRY(ctx, hipMemsetAsync(v->p, 0, v->bytes(), ctx->stream));
        }
    if (!ctx->partials.p) BHIP_TRY(ctx, ctx->partials.resize((size_t)kRedBlocks * 4));
    if (!ctx->state.p) {
        BHIP_TRY(ctx, ctx->state.resize(1));
        BHIP_TRY(ctx, hipMemsetAsync(ctx->state.p, 0, sizeof(PcgState), ctx->stream));
    }
    if (!ctx->flags.p) BHIP_TRY(ctx, ctx->flags.resize(4));
    return BHIP_OK;
}

extern "C" {

const char *bhip_version(void) { return "beta_hip 0.1 (gfx950)"; }

int bhip_create(int device, bhip_ctx **out) {
    if (!out) return BHIP_ERR_ARG;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return BHIP_ERR_HIP;
    if (device < 0 || device >= count) return BHIP_ERR_ARG;
    bhip_ctx *ctx = new bhip_ctx();
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess) {
        delete ctx;
        return BHIP_ERR_HIP;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->n_cu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
        hipEventCreate(&ctx->drv0) != hipSuccess || hipEventCreate(&ctx->drv1) != hipSuccess ||
        hipHostMalloc((void **)&ctx->h_state, sizeof(PcgState), hipHostMallocDefault) != hipSuccess) {
        bhip_destroy(ctx);
        return BHIP_ERR_HIP;
    }
    memset(ctx->h_state, 0, sizeof(PcgState));
    if (bhip_scratch_setup(ctx) != BHIP_OK) {
        bhip_destroy(ctx);
        return BHIP_ERR_HIP;
    }
    *out = ctx;
    return BHIP_OK;
}

void bhip_destroy(bhip_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    bhip_invalidate_graphs(ctx);
    bhip_plastic_release(ctx);
    bhip_multi_release(ctx);
    dvec<double> *v64[] = {&ctx->xyz,  &ctx->presc, &ctx->val, &ctx->val_aux, &ctx->mat, &ctx->load,
                           &ctx->pseudo, &ctx->rhs, &ctx->x,   &ctx->x0,      &ctx->r,   &ctx->p,
                           &ctx->q,    &ctx->dinv,  &ctx->tmp, &ctx->partials, &ctx->heat_us,
                           &ctx->heat_kruns, &ctx->heat_ku, &ctx->heat_V, &ctx->heat_z, &ctx->heat_part};
    for (auto *v : v64) v->release();
    dvec<int> *vi[] = {&ctx->conn,     &ctx->dof_first, &ctx->ndof,   &ctx->n2e_ptr, &ctx->n2e,
                       &ctx->nbr_ptr,  &ctx->nbr,       &ctx->pair_row, &ctx->nbr_off, &ctx->rowptr,
                       &ctx->colind,   &ctx->diag_pos,  &ctx->elem_group, &ctx->flags, &ctx->sell_ptr, &ctx->sell_col};
    for (auto *v : vi) v->release();
    dvec<float> *v32[] = {&ctx->val32, &ctx->bval32, &ctx->r32, &ctx->p32, &ctx->q32, &ctx->d32, &ctx->dinv32};
    for (auto *v : v32) v->release();
    ctx->sell_col16.release();
    ctx->bval64.release();
    ctx->eblk.release();
    ctx->n2e_loc.release();
    ctx->work.release();
    ctx->tmp_i.release();
    ctx->tmp_k1.release();
    ctx->tmp_k2.release();
    ctx->tmp_l1.release();
    ctx->tmp_l2.release();
    ctx->rest.release();
    ctx->node_rmask.release();
    ctx->lift_ptr.release();
    ctx->lift_row.release();
    ctx->lift_col.release();
    ctx->lift_val.release();
    ctx->gauss.release();
    ctx->state.release();
    if (ctx->h_state) (void)hipHostFree(ctx->h_state);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->drv0) (void)hipEventDestroy(ctx->drv0);
    if (ctx->drv1) (void)hipEventDestroy(ctx->drv1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char *bhip_last_error(const bhip_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bhip_set_mesh(bhip_ctx *ctx, int dim, int n_nodes, const double *xyz, int n_el, int nnpe,
                  const int *conn, const int *ndof_per_node, int ndof_uniform) {
    if (!ctx) return BHIP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    if (!(dim == 2 || dim == 3)) return bhip_fail(ctx, BHIP_ERR_ARG, "dim must be 2 or 3");
    if (!((dim == 3 && nnpe == 8) || (dim == 2 && nnpe == 4)))
        return bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "only hex8 (3D) and quad4 (2D) elements are built");
    if (n_nodes <= 0 || n_el < 0 || !xyz || (n_el > 0 && !conn))
        return bhip_fail(ctx, BHIP_ERR_ARG, "empty or missing mesh arrays");
    if ((long long)n_el * nnpe * nnpe >= (1LL << 31))
        return bhip_fail(ctx, BHIP_ERR_CAPACITY, "n_el * nnpe^2 must stay below 2^31");
    // connectivity: node range and distinct nodes per element
    for (long long e = 0; e < n_el; e++) {
        const int *c = conn + e * nnpe;
        for (int a = 0; a < nnpe; a++) {
            if (c[a] < 0 || c[a] >= n_nodes)
                return bhip_fail(ctx, BHIP_ERR_ARG, "connectivity refers to a node out of range in element " + std::to_string(e));
            for (int b = 0; b < a; b++)
                if (c[a] == c[b]) return bhip_fail(ctx, BHIP_ERR_ARG, "repeated node in element " + std::to_string(e));
        }
    }
    // DOF numbering: cumulative first DOF in node order (node.cpp:206-214)
    ctx->h_ndof.assign(n_nodes, ndof_uniform);
    if (ndof_per_node) ctx->h_ndof.assign(ndof_per_node, ndof_per_node + n_nodes);
    ctx->h_dof_first.resize(n_nodes);
    long long dof = 0;
    int mx = 0;
    bool uni = true;
    for (int i = 0; i < n_nodes; i++) {
        int k = ctx->h_ndof[i];
        if (k < 0 || k > 3) return bhip_fail(ctx, BHIP_ERR_ARG, "DOFs per node must be in [0,3]");
        ctx->h_dof_first[i] = (int)dof;
        dof += k;
        mx = std::max(mx, k);
        if (k != ctx->h_ndof[0]) uni = false;
    }
    if (dof >= (1LL << 31)) return bhip_fail(ctx, BHIP_ERR_CAPACITY, "too many DOFs");
    // every node of every element carries the largest DOF count (the element
    // kernels' block size); nodes outside the elements may carry fewer
    bool elem_full = true;
    for (long long t = 0; t < (long long)n_el * nnpe && elem_full; t++)
        if (ctx->h_ndof[conn[t]] != mx) elem_full = false;
    bhip_invalidate_graphs(ctx);
    ctx->dim = dim;
    ctx->n_nodes = n_nodes;
    ctx->n_el = n_el;
    ctx->nnpe = nnpe;
    ctx->n_dof = (int)dof;
    ctx->max_ndof = mx;
    ctx->uniform_ndof = uni;
    ctx->elem_ndof_full = elem_full;
    ctx->pattern_ready = ctx->assembled = ctx->f32_ready = ctx->f64b_ready = false;
    ctx->have_rest = false;
    ctx->external_matrix = false;
    hipStream_t s = ctx->stream;
    BHIP_TRY(ctx, ctx->xyz.resize((size_t)n_nodes * dim));
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->xyz.p, xyz, sizeof(double) * n_nodes * dim, hipMemcpyHostToDevice, s));
    BHIP_TRY(ctx, ctx->conn.resize(n_el > 0 ? (size_t)n_el * nnpe : 1));
    if (n_el > 0) BHIP_TRY(ctx, hipMemcpyAsync(ctx->conn.p, conn, sizeof(int) * n_el * nnpe, hipMemcpyHostToDevice, s));
    BHIP_TRY(ctx, ctx->ndof.resize(n_nodes));
    BHIP_TRY(ctx, ctx->dof_first.resize(n_nodes));
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->ndof.p, ctx->h_ndof.data(), sizeof(int) * n_nodes, hipMemcpyHostToDevice, s));
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->dof_first.p, ctx->h_dof_first.data(), sizeof(int) * n_nodes, hipMemcpyHostToDevice, s));
    // default: nothing restrained
    size_t nd = ctx->n_dof > 0 ? ctx->n_dof : 1;
    BHIP_TRY(ctx, ctx->rest.resize(nd));
    BHIP_TRY(ctx, ctx->presc.resize(nd));
    BHIP_TRY(ctx, hipMemsetAsync(ctx->rest.p, 0, nd, s));
    BHIP_TRY(ctx, hipMemsetAsync(ctx->presc.p, 0, sizeof(double) * nd, s));
    ctx->h_rest.assign(ctx->n_dof, 0);
    BHIP_PROPAGATE(bhip_ensure_vectors(ctx));
    BHIP_TRY(ctx, hipStreamSynchronize(s));
    return BHIP_OK;
}

int bhip_num_dof(const bhip_ctx *ctx) { return ctx ? ctx->n_dof : BHIP_ERR_ARG; }

int bhip_get_dof_first(bhip_ctx *ctx, int *dof_first) {
    if (!ctx || !dof_first) return BHIP_ERR_ARG;
    if (ctx->n_nodes == 0) return bhip_fail(ctx, BHIP_ERR_STATE, "mesh not set");
    std::copy(ctx->h_dof_first.begin(), ctx->h_dof_first.end(), dof_first);
    return BHIP_OK;
}

}  // extern "C"

int bhip_update_coords(bhip_ctx *ctx, const double *xyz) {
    if (!ctx || !xyz) return BHIP_ERR_ARG;
    if (ctx->n_nodes == 0) return bhip_fail(ctx, BHIP_ERR_STATE, "mesh not set");
    (void)hipSetDevice(ctx->device);
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->xyz.p, xyz, sizeof(double) * ctx->n_nodes * ctx->dim, hipMemcpyHostToDevice,
                                 ctx->stream));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->assembled = ctx->f32_ready = ctx->f64b_ready = ctx->lift_ready = false;
    return BHIP_OK;
}

extern "C" {

int bhip_set_restraints(bhip_ctx *ctx, const unsigned char *restrained, const double *values) {
    if (!ctx || !restrained) return BHIP_ERR_ARG;
    if (ctx->n_nodes == 0) return bhip_fail(ctx, BHIP_ERR_STATE, "mesh not set");
    (void)hipSetDevice(ctx->device);
    std::vector<unsigned char> r(restrained, restrained + ctx->n_dof);
    for (auto &v : r) v = v ? 1 : 0;
    bool same = ctx->have_rest && r == ctx->h_rest;
    ctx->h_rest = r;
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->rest.p, r.data(), ctx->n_dof, hipMemcpyHostToDevice, ctx->stream));
    if (values)
        BHIP_TRY(ctx, hipMemcpyAsync(ctx->presc.p, values, sizeof(double) * ctx->n_dof, hipMemcpyHostToDevice, ctx->stream));
    else
        BHIP_TRY(ctx, hipMemsetAsync(ctx->presc.p, 0, sizeof(double) * ctx->n_dof, ctx->stream));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->have_rest = true;
    if (!same) {
        ctx->pattern_ready = ctx->assembled = ctx->f32_ready = ctx->f64b_ready = false;
        bhip_invalidate_graphs(ctx);
    }
    return BHIP_OK;
}

int bhip_set_prescribed(bhip_ctx *ctx, const double *values) {
    if (!ctx || !values) return BHIP_ERR_ARG;
    if (ctx->n_nodes == 0) return bhip_fail(ctx, BHIP_ERR_STATE, "mesh not set");
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->presc.p, values, sizeof(double) * ctx->n_dof, hipMemcpyHostToDevice, ctx->stream));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BHIP_OK;
}

int bhip_build_pattern(bhip_ctx *ctx) {
    if (!ctx) return BHIP_ERR_ARG;
    if (ctx->n_nodes == 0) return bhip_fail(ctx, BHIP_ERR_STATE, "mesh not set");
    (void)hipSetDevice(ctx->device);
    return bhip_pattern_build(ctx);
}

int bhip_pattern_nnz(const bhip_ctx *ctx, long long *nnz_lower, long long *nnz_full) {
    if (!ctx) return BHIP_ERR_ARG;
    if (!ctx->pattern_ready) return BHIP_ERR_STATE;
    if (nnz_lower) *nnz_lower = ctx->nnz_lower;
    if (nnz_full) *nnz_full = ctx->nnz;
    return BHIP_OK;
}

int bhip_get_pattern_lower(bhip_ctx *ctx, int *rowptr, int *colind) {
    if (!ctx || !rowptr) return BHIP_ERR_ARG;
    if (!ctx->pattern_ready) return bhip_fail(ctx, BHIP_ERR_STATE, "pattern not built");
    return bhip_export_lower(ctx, rowptr, colind, nullptr);
}

int bhip_get_values_lower(bhip_ctx *ctx, double *val) {
    if (!ctx || !val) return BHIP_ERR_ARG;
    if (!ctx->pattern_ready) return bhip_fail(ctx, BHIP_ERR_STATE, "pattern not built");
    return bhip_export_lower(ctx, nullptr, nullptr, val);
}

int bhip_get_csr(bhip_ctx *ctx, int *rowptr, int *colind, double *val) {
    if (!ctx) return BHIP_ERR_ARG;
    if (!ctx->pattern_ready) return bhip_fail(ctx, BHIP_ERR_STATE, "pattern not built");
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (rowptr) BHIP_TRY(ctx, bhip_copy(ctx, rowptr, ctx->rowptr.p, sizeof(int) * (ctx->n_dof + 1), hipMemcpyDeviceToHost));
    if (colind) BHIP_TRY(ctx, bhip_copy(ctx, colind, ctx->colind.p, sizeof(int) * ctx->nnz, hipMemcpyDeviceToHost));
    if (val) BHIP_TRY(ctx, bhip_copy(ctx, val, ctx->val.p, sizeof(double) * ctx->nnz, hipMemcpyDeviceToHost));
    return BHIP_OK;
}

int bhip_set_materials(bhip_ctx *ctx, int kind, int n_groups, const double *params, const int *elem_group) {
    if (!ctx || !params || n_groups <= 0) return BHIP_ERR_ARG;
    if (ctx->n_nodes == 0) return bhip_fail(ctx, BHIP_ERR_STATE, "mesh not set");
    int stride;
    switch (kind) {
        case BHIP_MAT_ISO_ELASTIC: stride = 2; break;
        case BHIP_MAT_CMAT3D: stride = 36; break;
        case BHIP_MAT_PLANE_STRESS: stride = 4; break;
        case BHIP_MAT_HEAT: stride = 4; break;
        case BHIP_MAT_J2: stride = 4; break;
        default: return bhip_fail(ctx, BHIP_ERR_ARG, "unknown material kind");
    }
    if (!elem_group && !(n_groups == 1 || n_groups == ctx->n_el))
        return bhip_fail(ctx, BHIP_ERR_ARG, "without elem_group, n_groups must be 1 or n_el");
    if (elem_group)
        for (int e = 0; e < ctx->n_el; e++)
            if (elem_group[e] < 0 || elem_group[e] >= n_groups)
                return bhip_fail(ctx, BHIP_ERR_ARG, "elem_group out of range at element " + std::to_string(e));
    for (long long g = 0; g < n_groups; g++) {
        const double *q = params + g * stride;
        if (kind == BHIP_MAT_ISO_ELASTIC || kind == BHIP_MAT_J2) {
            if (!(q[0] > 0) || !(q[1] > -1.0 && q[1] < 0.5))
                return bhip_fail(ctx, BHIP_ERR_ARG, "isotropic material needs E > 0 and -1 < nu < 0.5");
        }
        if (kind == BHIP_MAT_PLANE_STRESS && !(q[0] > 0 && q[1] > 0 && q[3] > 0))
            return bhip_fail(ctx, BHIP_ERR_ARG, "plane stress needs positive moduli");
    }
    (void)hipSetDevice(ctx->device);
    BHIP_TRY(ctx, ctx->mat.resize((size_t)n_groups * stride));
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->mat.p, params, sizeof(double) * n_groups * stride, hipMemcpyHostToDevice, ctx->stream));
    if (elem_group && ctx->n_el > 0) {
        BHIP_TRY(ctx, ctx->elem_group.resize(ctx->n_el));
        BHIP_TRY(ctx, hipMemcpyAsync(ctx->elem_group.p, elem_group, sizeof(int) * ctx->n_el, hipMemcpyHostToDevice, ctx->stream));
    } else {
        ctx->elem_group.release();
    }
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->mat_kind = kind;
    ctx->n_groups = n_groups;
    ctx->mat_stride = stride;
    ctx->have_mat = true;
    ctx->assembled = false;
    return BHIP_OK;
}

int bhip_set_integration(bhip_ctx *ctx, int order, double factor) {
    if (!ctx) return BHIP_ERR_ARG;
    if (order < 1 || order > 4) return bhip_fail(ctx, BHIP_ERR_ARG, "integration order must be 1..4");
    if (!(factor > 0)) return bhip_fail(ctx, BHIP_ERR_ARG, "integration factor must be positive");
    ctx->order = order;
    ctx->factor = factor;
    return BHIP_OK;
}

int bhip_assemble(bhip_ctx *ctx, int what, const double *coef, const double *u_state) {
    if (!ctx) return BHIP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    scoped_dvec<double> du;
    const double *dptr = nullptr;
    if (u_state) {
        BHIP_TRY(ctx, du.resize(ctx->n_dof));
        BHIP_TRY(ctx, hipMemcpyAsync(du.p, u_state, sizeof(double) * ctx->n_dof, hipMemcpyHostToDevice, ctx->stream));
        dptr = du.p;
    }
    int st = bhip_assemble_impl(ctx, what, coef, dptr);
    (void)hipStreamSynchronize(ctx->stream);
    return st;
}

int bhip_get_pseudo_load(bhip_ctx *ctx, double *f) {
    if (!ctx || !f) return BHIP_ERR_ARG;
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    BHIP_TRY(ctx, bhip_copy(ctx, f, ctx->pseudo.p, sizeof(double) * ctx->n_dof, hipMemcpyDeviceToHost));
    return BHIP_OK;
}

int bhip_set_load(bhip_ctx *ctx, const double *load) {
    if (!ctx) return BHIP_ERR_ARG;
    if (ctx->n_nodes == 0) return bhip_fail(ctx, BHIP_ERR_STATE, "mesh not set");
    if (load) {
        // addLoadToEqn: a load on a restrained DOF is a fatal error (equation.cpp:862)
        for (int i = 0; i < ctx->n_dof; i++)
            if (load[i] != 0.0 && ctx->h_rest.size() == (size_t)ctx->n_dof && ctx->h_rest[i])
                return bhip_fail(ctx, BHIP_ERR_ARG, "Cannot add load to a restrained degree of freedom: " + std::to_string(i));
        BHIP_TRY(ctx, hipMemcpyAsync(ctx->load.p, load, sizeof(double) * ctx->n_dof, hipMemcpyHostToDevice, ctx->stream));
    } else {
        BHIP_TRY(ctx, hipMemsetAsync(ctx->load.p, 0, sizeof(double) * ctx->n_dof, ctx->stream));
    }
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BHIP_OK;
}

namespace {
__global__ void k_build_rhs(int n, const unsigned char *rest, const double *presc, const double *load,
                            const double *pseudo, double *b) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = rest[i] ? presc[i] : load[i] + pseudo[i];
}
}  // namespace

int bhip_build_rhs(bhip_ctx *ctx) {
    if (!ctx) return BHIP_ERR_ARG;
    if (ctx->n_nodes == 0) return bhip_fail(ctx, BHIP_ERR_STATE, "mesh not set");
    k_build_rhs<<<grid_for(ctx->n_dof, 256), 256, 0, ctx->stream>>>(ctx->n_dof, ctx->rest.p, ctx->presc.p,
                                                                    ctx->load.p, ctx->pseudo.p, ctx->rhs.p);
    BHIP_TRY(ctx, hipGetLastError());
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BHIP_OK;
}

int bhip_get_rhs(bhip_ctx *ctx, double *b) {
    if (!ctx || !b) return BHIP_ERR_ARG;
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    BHIP_TRY(ctx, bhip_copy(ctx, b, ctx->rhs.p, sizeof(double) * ctx->n_dof, hipMemcpyDeviceToHost));
    return BHIP_OK;
}

int bhip_set_rhs(bhip_ctx *ctx, const double *b) {
    if (!ctx || !b) return BHIP_ERR_ARG;
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->rhs.p, b, sizeof(double) * ctx->n_dof, hipMemcpyHostToDevice, ctx->stream));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BHIP_OK;
}

int bhip_set_x0(bhip_ctx *ctx, const double *x0) {
    if (!ctx) return BHIP_ERR_ARG;
    if (x0) BHIP_TRY(ctx, hipMemcpyAsync(ctx->x0.p, x0, sizeof(double) * ctx->n_dof, hipMemcpyHostToDevice, ctx->stream));
    else BHIP_TRY(ctx, hipMemsetAsync(ctx->x0.p, 0, sizeof(double) * ctx->n_dof, ctx->stream));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return BHIP_OK;
}

int bhip_solve(bhip_ctx *ctx, const bhip_solve_opts *opts, bhip_stats *stats) {
    if (!ctx || !opts) return BHIP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    return bhip_solve_impl(ctx, opts, stats);
}

int bhip_get_solution(bhip_ctx *ctx, double *x) {
    if (!ctx || !x) return BHIP_ERR_ARG;
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    BHIP_TRY(ctx, bhip_copy(ctx, x, ctx->x.p, sizeof(double) * ctx->n_dof, hipMemcpyDeviceToHost));
    return BHIP_OK;
}

int bhip_pcg_solve(bhip_ctx *ctx, double *b, const double *x0, const bhip_solve_opts *opts, bhip_stats *stats) {
    if (!ctx || !b || !opts) return BHIP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    BHIP_PROPAGATE(bhip_set_rhs(ctx, b));
    bhip_solve_opts o = *opts;
    o.use_x0 = x0 ? 1 : 0;
    if (x0) BHIP_PROPAGATE(bhip_set_x0(ctx, x0));
    int st = bhip_solve_impl(ctx, &o, stats);
    if (st != BHIP_OK && st != BHIP_ERR_NOT_CONVERGED) return st;
    BHIP_TRY(ctx, bhip_copy(ctx, b, ctx->x.p, sizeof(double) * ctx->n_dof, hipMemcpyDeviceToHost));
    return st;
}

int bhip_spmv(bhip_ctx *ctx, const double *x, double *y) {
    if (!ctx || !x || !y) return BHIP_ERR_ARG;
    if (!ctx->pattern_ready) return bhip_fail(ctx, BHIP_ERR_STATE, "pattern not built");
    (void)hipSetDevice(ctx->device);
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->tmp.p, x, sizeof(double) * ctx->n_dof, hipMemcpyHostToDevice, ctx->stream));
    BHIP_PROPAGATE(bhip_spmv_dev(ctx, ctx->tmp.p, ctx->q.p));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    BHIP_TRY(ctx, bhip_copy(ctx, y, ctx->q.p, sizeof(double) * ctx->n_dof, hipMemcpyDeviceToHost));
    return BHIP_OK;
}

int bhip_bench_spmv(bhip_ctx *ctx, int fp32, int reps, double *ms_avg, double *bytes) {
    if (!ctx || !ms_avg || !bytes) return BHIP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    return bhip_bench_spmv_impl(ctx, fp32, reps, ms_avg, bytes);
}

int bhip_tune_spmv(bhip_ctx *ctx, int reps, bhip_spmv_shape *table, int max_entries, int *n_entries,
                   int *chosen) {
    if (!ctx || !n_entries || !chosen || (max_entries > 0 && !table)) return BHIP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    return bhip_tune_spmv_impl(ctx, reps, table, max_entries, n_entries, chosen);
}

int bhip_set_spmv_shape(bhip_ctx *ctx, int id) {
    if (!ctx) return BHIP_ERR_ARG;
    return bhip_set_spmv_shape_impl(ctx, id);
}

int bhip_get_spmv_shape(const bhip_ctx *ctx) { return ctx ? ctx->sell_shape : BHIP_ERR_ARG; }

int bhip_spmv_shape_describe(int id, bhip_spmv_shape *info) {
    if (!info) return BHIP_ERR_ARG;
    memset(info, 0, sizeof(*info));
    return bhip_spmv_shape_info(id, info);
}

}  // extern "C"

namespace {
__global__ void __launch_bounds__(256) k_stream_copy(const float4 *__restrict__ a, float4 *__restrict__ b,
                                                    long long n) {
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

// Read-only stream: nontemporal 16-byte loads, one sum per wave written out
// (so that no load is dead).
__global__ void __launch_bounds__(256) k_stream_read(const float4 *__restrict__ a, long long n, float *out) {
    typedef float f4v __attribute__((ext_vector_type(4)));
    const f4v *av = reinterpret_cast<const f4v *>(a);
    float s = 0.f;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x) {
        const f4v v = __builtin_nontemporal_load(av + i);
        s += (v.x + v.y) + (v.z + v.w);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((threadIdx.x & 63) == 0) out[blockIdx.x * 4 + (threadIdx.x >> 6)] = s;
}
}  // namespace

extern "C" {

int bhip_bench_stream(bhip_ctx *ctx, long long bytes, int reps, double *ms_avg, double *gbps) {
    if (!ctx || !ms_avg || !gbps || bytes < 16 || reps <= 0) return BHIP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    long long n4 = bytes / 16;
    scoped_dvec<float4> a, b;
    BHIP_TRY(ctx, a.resize(n4));
    BHIP_TRY(ctx, b.resize(n4));
    BHIP_TRY(ctx, hipMemsetAsync(a.p, 0, a.bytes(), ctx->stream));
    // 4 blocks of 256 threads per CU: the grid-stride copy reaches its best
    // rate there on gfx950 (larger grids lose ~15% to HBM page conflicts)
    const int grid = ctx->n_cu * 4;
    for (int w = 0; w < 2; w++) k_stream_copy<<<grid, 256, 0, ctx->stream>>>(a.p, b.p, n4);
    BHIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int k = 0; k < reps; k++) k_stream_copy<<<grid, 256, 0, ctx->stream>>>(a.p, b.p, n4);
    BHIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    BHIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    BHIP_TRY(ctx, hipGetLastError());
    float ms = 0;
    BHIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *ms_avg = ms / reps;
    *gbps = 2.0 * n4 * 16 / (*ms_avg * 1e6);
    return BHIP_OK;
}

int bhip_bench_read(bhip_ctx *ctx, long long bytes, int reps, double *ms_avg, double *gbps) {
    if (!ctx || !ms_avg || !gbps || bytes < 16 || reps <= 0) return BHIP_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    long long n4 = bytes / 16;
    // 8 blocks of 256 threads per CU: the best grid of this loop on gfx950
    const int grid = ctx->n_cu * 8;
    scoped_dvec<float4> a;
    scoped_dvec<float> out;
    BHIP_TRY(ctx, a.resize(n4));
    BHIP_TRY(ctx, out.resize((size_t)grid * 4));
    BHIP_TRY(ctx, hipMemsetAsync(a.p, 0, a.bytes(), ctx->stream));
    for (int w = 0; w < 2; w++) k_stream_read<<<grid, 256, 0, ctx->stream>>>(a.p, n4, out.p);
    BHIP_TRY(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    for (int k = 0; k < reps; k++) k_stream_read<<<grid, 256, 0, ctx->stream>>>(a.p, n4, out.p);
    BHIP_TRY(ctx, hipEventRecord(ctx->ev1, ctx->stream));
    BHIP_TRY(ctx, hipEventSynchronize(ctx->ev1));
    BHIP_TRY(ctx, hipGetLastError());
    float ms = 0;
    BHIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *ms_avg = ms / reps;
    *gbps = (double)n4 * 16 / (*ms_avg * 1e6);
    return BHIP_OK;
}

}  // extern "C"
