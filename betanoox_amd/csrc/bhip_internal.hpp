// bhip_internal.hpp -- context, device buffers and shared device helpers of
// libbeta_hip.so (MI355X / gfx950).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "beta_hip.h"

// ---------------------------------------------------------------- errors
int bhip_fail(bhip_ctx *ctx, int status, const std::string &msg);

#define BHIP_TRY(ctx, call)                                                              \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return bhip_fail((ctx), e_ == hipErrorOutOfMemory ? BHIP_ERR_OOM : BHIP_ERR_HIP, \
                             std::string(#call) + ": " + hipGetErrorString(e_));         \
    } while (0)

#define BHIP_PROPAGATE(expr)                                                             \
    do {                                                                                 \
        int s_ = (expr);                                                                 \
        if (s_ != BHIP_OK) return s_;                                                    \
    } while (0)

// ----------------------------------------------------------- device buffer
// Device array whose allocation only grows: resizing to fewer elements keeps
// the allocation (n shrinks, cap stays), so a context reused for meshes of
// varying size does not free and reallocate.  hipFree waits for the device,
// which would stall the work other contexts have in flight on other streams.
template <class T>
struct dvec {
    T *p = nullptr;
    size_t n = 0;    // elements in use
    size_t cap = 0;  // elements allocated
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = cap = 0;
    }
    // n = count (0 releases); contents undefined.
    hipError_t resize(size_t count) {
        if (count == 0) {
            release();
            return hipSuccess;
        }
        if (p && count <= cap) {
            n = count;
            return hipSuccess;
        }
        release();
        hipError_t e = hipMalloc((void **)&p, count * sizeof(T));
        if (e == hipSuccess) n = cap = count;
        else p = nullptr;
        return e;
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Device array owned by a scope: released when the scope ends (hipFree first
// waits for the device, so kernels still using it finish).  For function-local
// temporaries; dvec members are released by their owners.
template <class T>
struct scoped_dvec : dvec<T> {
    scoped_dvec() = default;
    scoped_dvec(const scoped_dvec &) = delete;
    scoped_dvec &operator=(const scoped_dvec &) = delete;
    ~scoped_dvec() { this->release(); }
};

// Device-side solver state of one PCG (FP64 or FP32 inner).  Lives in device
// memory so that kernels chain without host round trips; the host polls it.
struct PcgState {
    double rz;        // r.z of the current iterate
    double pAp;       // p.Ap of the current iteration
    double alpha, beta;
    double rr, xx, bb;  // |r|^2, |x|^2, |b|^2
    double tol;
    double norm;      // last stopping norm
    double rr0;       // |r0|^2 of this (inner) solve
    int done;         // 1 once converged or out of iterations
    int iter;         // iterations completed
    int maxit;
    int stop_rule;    // BHIP_STOP_* or 2 = inner relative reduction
    int converged;
    unsigned cnt[4];  // last-block counters (reset by the last block)
    int pad;
    double rzb[2];    // r.z double-buffered for the fused iteration (read [par], written [par ^ 1])
};

constexpr int kMaxNnpe = 8;
constexpr int kMaxIp = 64;    // order <= 4 in 3D
constexpr int kRedBlocks = 2048;  // upper bound of grid-stride reduction grids
constexpr int kTanVals = 21;  // J2 tangent per Gauss point: upper triangle of a symmetric 6x6
constexpr int kGraphKeyWords = 40;  // words of a captured iteration graph's key

struct bhip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;    // one solve or one benchmark
    hipEvent_t drv0 = nullptr, drv1 = nullptr;  // drivers that run several solves
    std::string err;
    int n_cu = 256;

    // ---- mesh
    int dim = 0, n_nodes = 0, n_el = 0, nnpe = 0, n_dof = 0, max_ndof = 0;
    bool uniform_ndof = true;
    bool elem_ndof_full = true;  // every element node has max_ndof DOFs
    std::vector<int> h_conn, h_ndof, h_dof_first;
    dvec<double> xyz;
    dvec<int> conn, dof_first, ndof;
    // ---- restraints
    bool have_rest = false;
    std::vector<unsigned char> h_rest;
    dvec<unsigned char> rest;       // per DOF
    dvec<unsigned char> node_rmask; // per node: bit c set if comp c restrained
    dvec<double> presc;             // per DOF prescribed value
    // ---- node graph and pattern
    bool pattern_ready = false;
    long long n_pairs = 0, nnz = 0, nnz_lower = 0;
    dvec<int> n2e_ptr, n2e;         // node -> incident elements (ascending)
    dvec<unsigned char> n2e_loc;    // the node's local index in each of those elements
    dvec<int> nbr_ptr, nbr, pair_row, nbr_off;  // node adjacency incl. self
    dvec<int> rowptr, colind, diag_pos;          // full symmetric CSR
    dvec<double> val;                            // FP64 values
    dvec<double> val_aux;                        // second value array (heat: K part)
    dvec<double> eblk;                           // element-stage blocks of the last assembly
    // ---- FP32 mirror for the inner solve
    bool f32_ready = false;
    bool use_bsr = false;    // FP32 inner solve on SELL-64 node blocks
    int sell_b = 0;          // block size (DOFs per node) of the SELL layout
    dvec<float> val32;       // CSR FP32 values (non-BSR case)
    // B x B-block sliced ELL (SELL-64): slice s holds 64 consecutive nodes and
    // w_s block columns; entry (k, lane) of slice s is at sell_ptr[s] + 64k + lane
    int n_slices = 0;
    long long n_sell = 0;    // padded block count
    int sell_wmax = 0;       // widest slice (block columns)
    dvec<int> sell_ptr, sell_col;  // slice starts; block columns (node index)
    dvec<short> sell_col16;        // block columns as offsets from the row's node (narrow form)
    bool sell_narrow = false;      // 16-bit offsets in use
    // Column runs (one DOF per node only): an entry is a run of three
    // consecutive columns (16-bit offset of the first from the row's node) with
    // three values, so a 27-column hex8 row takes 9 entries.
    bool sell_runs = false;
    // Union slots (three DOFs per node, 16-bit offsets): slot u of slice s
    // holds, in every lane, the block at the u-th smallest offset of the
    // union of the slice's neighbour offsets (zero where the lane's node has
    // no such neighbour, its column then the node itself), so slot order is
    // column order in every row.  sell_uoff[sell_ptr[s] / 64 + u] is that
    // offset; slot_rank (one byte per slot) holds the rank of the slot's
    // neighbour in its node's neighbour list, 255 for a zero block.  Runs of
    // consecutive offsets let the SpMV gather x once per run and take the
    // rest from the neighbouring lanes.
    bool sell_union = false;
    dvec<int> sell_uoff;
    dvec<unsigned char> slot_rank;
    dvec<float> bval32;      // SELL block values, SoA: sell_vals(ctx) x n_sell
    // ---- FP64 values in the same SELL layout (BHIP_PREC_FP64_BLOCKED)
    bool f64b_ready = false;       // bval64 holds the current matrix
    bool bval64_layout_ok = false; // bval64 sized and zero-padded for the current layout
    dvec<double> bval64;           // SELL block values, SoA: sell_vals(ctx) x n_sell
    bool sell_layout_ok = false;  // SELL pointers/columns built for the current pattern
    bool sell_cols_ok = false;    // SELL block columns written (by the first CSR-to-SELL pass)
    int sell_shape = 0;      // tile shape id of the FP32 SELL SpMV
    bool sell_shape_set = false;  // chosen by bhip_tune_spmv / bhip_set_spmv_shape (else the
                                  // tuned default table picks it when the SELL layout is built)
    // ---- materials / integration
    int mat_kind = 0, n_groups = 0, mat_stride = 0;
    dvec<double> mat;
    dvec<int> elem_group;
    bool have_mat = false;
    int order = 2;
    double factor = 1.0;
    // Gauss weights | shape derivatives | shape values of (gauss_nnpe, gauss_order),
    // per context so that contexts on concurrent streams never share them
    dvec<double> gauss;
    int gauss_nnpe = 0, gauss_order = 0, gauss_nip = 0;
    bool assembled = false;
    int asm_kind = 0;
    bool external_matrix = false;  // matrix loaded from CSR/RCV data, no mesh behind it
    // ---- lift table of the last stiffness assembly: for every node with free
    // DOFs, its neighbours with restrained DOFs and their FP64 coupling blocks
    bool lift_ready = false;
    long long n_lift = 0;
    dvec<long long> lift_ptr;   // per node
    dvec<int> lift_row, lift_col;
    dvec<double> lift_val;      // ndof x ndof per entry, row comp of the owner node first
                                // (heat: K and (a1 K + M)/dt per entry)
    double heat_a1 = 0, heat_dt = 1;  // coefficients of the last heat-transient assembly
    // bhip_heat_run buffers (grow-only, kept between runs): u_s, the FP64
    // column-run copy of K and K u_s
    dvec<double> heat_us, heat_kruns, heat_ku, heat_V, heat_z, heat_part;
    // ---- vectors (FP64)
    dvec<double> load, pseudo, rhs, x, x0, r, p, q, dinv, tmp;
    // ---- vectors (FP32 inner)
    dvec<float> r32, p32, q32, d32, dinv32;
    // ---- reductions / state
    dvec<double> partials;  // kRedBlocks * 4
    dvec<PcgState> state;
    dvec<int> flags;        // assembly / diagonal check flags
    double zero_diag_value = 0.0;  // ReplaceZeroDiagonal value (0: zero diagonals are an error)
    PcgState *h_state = nullptr;  // pinned mirror
    unsigned char *scratch = nullptr;  // device address of this GPU's allocation-free scratch (solver.hip)
    // ---- graphs: one captured chunk of iterations per iteration kind, with
    // the addresses, sizes and shape it was captured with (graph_key in
    // solver.hip); a chunk whose key still matches is replayed, so a
    // reassembly into the same buffers keeps its graphs
    hipGraphExec_t graph64 = nullptr, graph32 = nullptr, graph64b = nullptr;
    int graph64_iters = 0, graph32_iters = 0, graph64b_iters = 0;
    uint64_t graph64_key[kGraphKeyWords] = {}, graph32_key[kGraphKeyWords] = {}, graph64b_key[kGraphKeyWords] = {};
    double spmv_bytes64 = 0, spmv_bytes32 = 0;
    // ---- J2 plasticity state (plastic.hip)
    struct PlasticData *plastic = nullptr;
    // ---- multi-case solve buffers and graph (multi.hip)
    struct MultiData *multi = nullptr;
    // ---- scratch of pattern, layout and lift-table builds (grow-only, reused)
    dvec<char> work;                   // hipcub temporary storage
    dvec<int> tmp_i;                   // per-node counters
    dvec<unsigned long long> tmp_k1, tmp_k2;  // node-pair keys
    dvec<long long> tmp_l1, tmp_l2;    // per-DOF lengths and scans
};

// Values per stored SELL entry: B*B for node blocks, 3 for column runs.
inline int sell_vals(const bhip_ctx *ctx) { return ctx->sell_runs ? 3 : ctx->sell_b * ctx->sell_b; }

// Blocking copy ordered on the context's stream.  Context streams are
// non-blocking, so the legacy-stream hipMemcpy would neither wait for the
// context's kernels nor be allowed while another context captures a graph.
inline hipError_t bhip_copy(bhip_ctx *ctx, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ctx->stream);
    return e != hipSuccess ? e : hipStreamSynchronize(ctx->stream);
}

// -------------------------------------------------------------- launching
inline int grid_for(long long n, int block, int max_blocks = 1 << 30) {
    long long g = (n + block - 1) / block;
    if (g < 1) g = 1;
    if (g > max_blocks) g = max_blocks;
    return (int)g;
}

// XCD-aware block remap: gfx950 dispatches consecutive workgroups round-robin
// over 8 XCDs, each with its own L2.  Map so that every XCD works on one
// contiguous range of logical blocks (neighbouring rows share gathered x
// entries, which then stay in that XCD's L2).
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
    const int nx = 8;
    int per = nblocks / nx;
    int rem = nblocks - per * nx;
    int xcd = bid % nx;
    int idx = bid / nx;
    // XCDs [0, rem) hold per+1 logical blocks, the rest hold per.
    int base = (xcd < rem) ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return base + idx;
}

// wave64 reduction of a double (all lanes get the sum)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block reduction (blockDim.x multiple of 64, <= 1024); result in thread 0.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV], double *smem /* 16*NV */) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const int nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) smem[wid * NV + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0;
            for (int w = 0; w < nw; w++) s += smem[w * NV + k];
            v[k] = s;
        }
    }
}

// Inner residual reduction of one refinement step: the requested inner_tol,
// loosened when the outer relative residual rel is already close to tol, so
// that the last step does not overshoot.  It stays 5x short of the reduction
// needed (a step's true reduction trails its inner one) and at most 0.1.  It
// depends only on computed residuals, so solves stay reproducible.
inline double ir_inner_tol(double inner_tol, double tol, double rel) {
    if (!(rel > 0)) return inner_tol;
    const double need = 0.2 * tol / rel;
    return need > inner_tol ? (need < 0.1 ? need : 0.1) : inner_tol;
}

// --------------------------------------------------------- internal entry
int bhip_ensure_vectors(bhip_ctx *ctx);
// New coordinates for the current mesh; pattern and restraints stay.
int bhip_update_coords(bhip_ctx *ctx, const double *xyz);
int bhip_pattern_build(bhip_ctx *ctx);
int bhip_assemble_impl(bhip_ctx *ctx, int what, const double *coef, const double *u_state);
int bhip_solve_impl(bhip_ctx *ctx, const bhip_solve_opts *opts, bhip_stats *stats);
int bhip_spmv_dev(bhip_ctx *ctx, const double *x, double *y);
int bhip_export_lower(bhip_ctx *ctx, int *rowptr, int *colind, double *val);
int bhip_bench_spmv_impl(bhip_ctx *ctx, int fp32, int reps, double *ms, double *bytes);
int bhip_tune_spmv_impl(bhip_ctx *ctx, int reps, bhip_spmv_shape *table, int max_entries, int *n_entries,
                        int *chosen);
int bhip_spmv_shape_info(int id, bhip_spmv_shape *info);
int bhip_set_spmv_shape_impl(bhip_ctx *ctx, int id);
void bhip_invalidate_graphs(bhip_ctx *ctx);
int bhip_scratch_setup(bhip_ctx *ctx);
// The assembly may write the solver's SELL block values itself: the layout is
// the blocked form (not column runs) of the current pattern and its columns
// have been written.
bool bhip_sell_values_direct(const bhip_ctx *ctx);  // once per device: address of the allocation-free scratch
void bhip_plastic_release(bhip_ctx *ctx);
void bhip_multi_invalidate(bhip_ctx *ctx);
void bhip_multi_release(bhip_ctx *ctx);

// ---- multi-case solve (multi.hip) and the packed solver (pack.hip)
struct PackEngine;
// Buffers of the multi-case solve, one set per context.
struct MultiData {
    int R = 0;        // cases the buffers hold
    long long n = 0;  // DOFs per case
    dvec<double> b64, x64, r64, scale, rn;  // case-major, R x n (scale, rn: R)
    dvec<double> save;                     // RHS and solution of the context (case-by-case path)
    dvec<float> r32, p32, q32, d32;        // FP32 inner vectors, case-major, R x n
    dvec<double> partials;                 // FP64 residual partials
    dvec<double> resid;                    // 2R: |b_r - K x_r|^2, |b_r|^2
    double *h_resid = nullptr;             // pinned mirror of resid
    PackEngine *pack = nullptr;            // packed solver of this context alone (its stream)
    void release_buffers();
};
int bhip_multi_prepare(bhip_ctx *ctx, int R, bool fused);
void bhip_multi_enqueue_resid(bhip_ctx *ctx, int R, hipStream_t s, double *out);
void bhip_multi_enqueue_pack32(bhip_ctx *ctx, int R, hipStream_t s);
void bhip_multi_enqueue_axpy(bhip_ctx *ctx, int R, hipStream_t s);
int bhip_multi_worst_resid(bhip_ctx *ctx, int R, double *worst);
int bhip_cases_begin(bhip_ctx *ctx, int nrhs, const double *presc, const double *load, const bhip_solve_opts *opts,
                     bool *fused);
int bhip_cases_finish(bhip_ctx *ctx, int nrhs, double *x, bhip_stats *S);
// stream = nullptr: the engine creates its own (non-blocking) stream
int bhip_pack_engine_create(int device, hipStream_t stream, PackEngine **out);
void bhip_pack_engine_destroy(PackEngine *pe);
hipStream_t bhip_pack_engine_stream(PackEngine *pe);
int bhip_pack_run(PackEngine *pe, bhip_ctx *const *ctxs, int n, int R, const bhip_solve_opts &o, bhip_stats *stats,
                  int *status);
// Gauss tables of (nnpe, order) in ctx->gauss (uploaded when they change):
// w = gauss, dN = gauss + kGaussDN, N = gauss + kGaussN.
constexpr int kGaussDN = kMaxIp;
constexpr int kGaussN = kMaxIp + kMaxIp * 24;
int bhip_gauss_upload(bhip_ctx *ctx, int nnpe, int order, int *nip_out);
int bhip_scan_i64(bhip_ctx *ctx, const long long *in, long long *out, long long n);
// Right-hand sides of R load cases from the lift table (built on first use
// after a stiffness assembly): b_r = restrained ? presc_r : load + (-K_fr presc_r).
// presc and b are device arrays, R x n_dof case-major; load may be null.
int bhip_lift_build(bhip_ctx *ctx);
int bhip_lift_rhs(bhip_ctx *ctx, int R, const double *presc, const double *load, double *b);
int bhip_prepare_solver(bhip_ctx *ctx, bool fp32);
int bhip_prepare_refined(bhip_ctx *ctx);
void bhip_runs_from_csr64(bhip_ctx *ctx, const double *csr_vals, double *out);
void bhip_runs_spmv64(bhip_ctx *ctx, const double *vals, const double *x, double *y);
// projected initial guesses (project.hip): at most kProjMax basis vectors;
// a shares buffer of kProjShares doubles; coefficient arrays of kProjNV
constexpr int kProjMax = 16;
constexpr int kProjNV = kProjMax + 1;
constexpr int kProjBlocks = 512;
constexpr size_t kProjShares = (size_t)kProjBlocks * kProjNV;
// coef[i] = v_i . y (i < m, 0 for slot skip); mode 1 adds coef[kProjMax] =
// 1 / sqrt(a.y - sum coef[i]^2) (0 when that vanishes)
void bhip_proj_dots(bhip_ctx *ctx, long long n, int m, const double *V, const double *a, const double *y, int skip,
                    int mode, double *part, double *coef);
// out = scale * (base + sign * sum_{i<m} coef[i] v_i)
void bhip_proj_combo(bhip_ctx *ctx, long long n, int m, const double *V, const double *coef, const double *base,
                     double sign, int use_scale, double *out);
int bhip_residual_dev(bhip_ctx *ctx, const double *x, const double *b, double *r, double *rr, double *bb);
const double *bhip_plastic_tangent(bhip_ctx *ctx);

// ------------------------------------------------ global-memory pointers
// A pointer a kernel reads from device memory (a descriptor) is generic to
// the compiler, so its loads become flat loads, which count against both
// the vector-memory and the LDS counters: the compiler then waits for every
// outstanding load at once instead of only the ones a use needs.  Casting
// to the global address space turns them into global loads.
#define BHIP_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ BHIP_GLOBAL T *to_global(T *p) {
    return (BHIP_GLOBAL T *)p;
}

// ------------------------------------------------ hex8 element helpers
// (plastic.hip, forces.hip)
// All 8 global gradients at ip (isoElement.cpp:324-372).
__device__ __forceinline__ double hex8_all_grads(const double *dN, const double *x, const double *y,
                                                 const double *z, int ip, double (*g)[3]) {
    const double *dxi = dN + ip * 24, *deta = dxi + 8, *dzeta = dxi + 16;
    double px1 = 0, px2 = 0, px3 = 0, py1 = 0, py2 = 0, py3 = 0, pz1 = 0, pz2 = 0, pz3 = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        px1 += dxi[i] * x[i];
        px2 += deta[i] * x[i];
        px3 += dzeta[i] * x[i];
        py1 += dxi[i] * y[i];
        py2 += deta[i] * y[i];
        py3 += dzeta[i] * y[i];
        pz1 += dxi[i] * z[i];
        pz2 += deta[i] * z[i];
        pz3 += dzeta[i] * z[i];
    }
    double d1 = py2 * pz3 - py3 * pz2, d2 = px2 * pz3 - px3 * pz2, d3 = px2 * py3 - px3 * py2;
    double d4 = py1 * pz3 - py3 * pz1, d5 = px1 * pz3 - px3 * pz1, d6 = px1 * py3 - px3 * py1;
    double d7 = py1 * pz2 - py2 * pz1, d8 = px1 * pz2 - px2 * pz1, d9 = px1 * py2 - px2 * py1;
    double detJ = px1 * d1 - py1 * d2 + pz1 * d3;
    double inv = 1. / detJ;
    double a11 = d1 * inv, a12 = -d4 * inv, a13 = d7 * inv;
    double a21 = -d2 * inv, a22 = d5 * inv, a23 = -d8 * inv;
    double a31 = d3 * inv, a32 = -d6 * inv, a33 = d9 * inv;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        g[i][0] = a11 * dxi[i] + a12 * deta[i] + a13 * dzeta[i];
        g[i][1] = a21 * dxi[i] + a22 * deta[i] + a23 * dzeta[i];
        g[i][2] = a31 * dxi[i] + a32 * deta[i] + a33 * dzeta[i];
    }
    return detJ;
}

__device__ __forceinline__ void load_coords(const double *xyz, const int *ec, double *x, double *y, double *z) {
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const double *q = xyz + 3 * (long long)ec[i];
        x[i] = q[0];
        y[i] = q[1];
        z[i] = q[2];
    }
}
