// pattern.hip -- device construction of the sparsity pattern.
//
// The reference registers, for every element and every pair (I,J) of its DOF
// map, the lower-triangle location (I,J) when both DOFs are free and I >= J,
// and (I,I) when I is restrained (Equations::specifyNonZeroLocation,
// equation.cpp:1216-1267); rows are sorted unique column lists and an empty
// row receives its diagonal (SymmetricSparseMatrix::allocate,
// sparse.cpp:482-495).  Hence:
//   free row I      = { J free : some element holds both node(I) and node(J) }
//   restrained row I = { I }
// The device builds the node adjacency (incl. self) once by sorting the
// (node, node) pairs of every element, then expands it into a full symmetric
// CSR with int32 indices whose lower part is the reference layout (diagonal
// last in each row).  Because DOFs are numbered node-major with cumulative
// first DOFs (node.cpp:206-214), node order == DOF order and the expanded
// rows come out sorted.
#include <algorithm>

#include <hipcub/hipcub.hpp>

#include "bhip_internal.hpp"

namespace {

__global__ void k_count_incidence(const int *conn, long long total, int *cnt) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total) atomicAdd(&cnt[conn[t]], 1);
}

__global__ void k_fill_incidence(const int *conn, long long total, int nnpe, const int *ptr,
                                 int *cursor, int *n2e, unsigned char *n2e_loc) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < total) {
        int n = conn[t];
        int pos = atomicAdd(&cursor[n], 1);
        n2e[ptr[n] + pos] = (int)(t / nnpe);
        n2e_loc[ptr[n] + pos] = (unsigned char)(t % nnpe);  // the node's local index in the element
    }
}

// Sort each node's incident-element list ascending (lists are short), local
// indices alongside.
__global__ void k_sort_incidence(int n_nodes, const int *ptr, int *n2e, unsigned char *n2e_loc) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    int b = ptr[n], e = ptr[n + 1];
    for (int i = b + 1; i < e; i++) {
        int v = n2e[i], j = i - 1;
        unsigned char l = n2e_loc[i];
        while (j >= b && n2e[j] > v) {
            n2e[j + 1] = n2e[j];
            n2e_loc[j + 1] = n2e_loc[j];
            j--;
        }
        n2e[j + 1] = v;
        n2e_loc[j + 1] = l;
    }
}

// One key per (element, a, b): (node_a << 32) | node_b.
__global__ void k_pair_keys(const int *conn, int n_el, int nnpe, unsigned long long *keys) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long per = (long long)nnpe * nnpe;
    if (t >= (long long)n_el * per) return;
    int e = (int)(t / per);
    int ab = (int)(t - (long long)e * per);
    int a = ab / nnpe, b = ab - a * nnpe;
    unsigned long long na = (unsigned)conn[(long long)e * nnpe + a];
    unsigned long long nb = (unsigned)conn[(long long)e * nnpe + b];
    keys[t] = (na << 32) | nb;
}

__global__ void k_split_keys(const unsigned long long *keys, long long n, int *row, int *col,
                             int *cnt) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    unsigned long long k = keys[t];
    int r = (int)(k >> 32);
    row[t] = r;
    col[t] = (int)(k & 0xffffffffULL);
    atomicAdd(&cnt[r], 1);
}

__global__ void k_node_rmask(int n_nodes, const int *first, const int *ndof,
                             const unsigned char *rest, unsigned char *mask) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    unsigned char m = 0;
    for (int c = 0; c < ndof[n]; c++)
        if (rest[first[n] + c]) m |= (unsigned char)(1u << c);
    mask[n] = m;
}

// Per node: nbr_off[p] = free DOFs among the node's neighbours before p, and
// the row length of each of the node's DOFs.
__global__ void k_row_lengths(int n_nodes, const int *nbr_ptr, const int *nbr, const int *first,
                              const int *ndof, const unsigned char *rmask, int *nbr_off,
                              long long *rowlen) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    int tot = 0;
    for (int p = nbr_ptr[n]; p < nbr_ptr[n + 1]; p++) {
        nbr_off[p] = tot;
        int m = nbr[p];
        tot += ndof[m] - __popc((unsigned)rmask[m]);
    }
    bool has = nbr_ptr[n + 1] > nbr_ptr[n];
    for (int c = 0; c < ndof[n]; c++) {
        bool r = (rmask[n] >> c) & 1u;
        rowlen[first[n] + c] = (r || !has) ? 1 : tot;
    }
}

__global__ void k_rowptr32(long long n, const long long *scan, int *rowptr) {
    long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= n) rowptr[t] = (int)scan[t];
}

__global__ void k_fill_cols(long long n_pairs, const int *pair_row, const int *nbr,
                            const int *nbr_off, const int *first, const int *ndof,
                            const unsigned char *rmask, const int *rowptr, int *colind,
                            int *diag_pos) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n_pairs) return;
    int n = pair_row[p], m = nbr[p];
    unsigned rn = rmask[n], rm = rmask[m];
    for (int c = 0; c < ndof[n]; c++) {
        if ((rn >> c) & 1u) continue;
        int I = first[n] + c;
        int base = rowptr[I] + nbr_off[p];
        int r = 0;
        for (int d = 0; d < ndof[m]; d++) {
            if ((rm >> d) & 1u) continue;
            int J = first[m] + d;
            colind[base + r] = J;
            if (J == I) diag_pos[I] = base + r;
            r++;
        }
    }
}

// Restrained rows and DOFs of nodes without elements: diagonal only.
__global__ void k_fill_diag_rows(int n_nodes, const int *nbr_ptr, const int *first,
                                 const int *ndof, const unsigned char *rmask, const int *rowptr,
                                 int *colind, int *diag_pos) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    bool has = nbr_ptr[n + 1] > nbr_ptr[n];
    for (int c = 0; c < ndof[n]; c++) {
        bool r = (rmask[n] >> c) & 1u;
        if (r || !has) {
            int I = first[n] + c;
            colind[rowptr[I]] = I;
            diag_pos[I] = rowptr[I];
        }
    }
}

__global__ void k_lower_len(int n, const int *rowptr, const int *diag_pos, long long *len) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) len[i] = diag_pos[i] - rowptr[i] + 1;
}

__global__ void k_copy_lower(int n, const int *rowptr, const int *colind, const double *val,
                             const long long *lptr, int *lcol, double *lval) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int b = rowptr[i];
    long long o = lptr[i];
    long long len = lptr[i + 1] - o;
    for (long long k = 0; k < len; k++) {
        if (lcol) lcol[o + k] = colind[b + k];
        if (lval) lval[o + k] = val[b + k];
    }
}

template <class T>
int exclusive_scan(bhip_ctx *ctx, const T *in, T *out, long long n) {
    size_t tb = 0;
    BHIP_TRY(ctx, hipcub::DeviceScan::ExclusiveSum(nullptr, tb, in, out, (int)n, ctx->stream));
    BHIP_TRY(ctx, ctx->work.resize(tb > 0 ? tb : 1));
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(ctx->work.p, tb, in, out, (int)n, ctx->stream);
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    BHIP_TRY(ctx, e);
    return BHIP_OK;
}

}  // namespace

int bhip_scan_i64(bhip_ctx *ctx, const long long *in, long long *out, long long n) {
    return exclusive_scan<long long>(ctx, in, out, n);
}

int bhip_pattern_build(bhip_ctx *ctx) {
    hipStream_t s = ctx->stream;
    const int B = 256;
    const int nn = ctx->n_nodes;
    ctx->pattern_ready = false;
    bhip_invalidate_graphs(ctx);

    // ---- restraint masks
    BHIP_TRY(ctx, ctx->node_rmask.resize(nn > 0 ? nn : 1));
    k_node_rmask<<<grid_for(nn, B), B, 0, s>>>(nn, ctx->dof_first.p, ctx->ndof.p, ctx->rest.p,
                                                ctx->node_rmask.p);
    BHIP_TRY(ctx, hipGetLastError());

    // ---- node -> element incidence
    long long total = (long long)ctx->n_el * ctx->nnpe;
    dvec<int> &cnt = ctx->tmp_i;
    BHIP_TRY(ctx, cnt.resize(nn + 1));
    BHIP_TRY(ctx, hipMemsetAsync(cnt.p, 0, sizeof(int) * (nn + 1), s));
    k_count_incidence<<<grid_for(total, B), B, 0, s>>>(ctx->conn.p, total, cnt.p);
    BHIP_TRY(ctx, hipGetLastError());
    BHIP_TRY(ctx, ctx->n2e_ptr.resize(nn + 1));
    BHIP_PROPAGATE(exclusive_scan(ctx, cnt.p, ctx->n2e_ptr.p, nn + 1));
    BHIP_TRY(ctx, ctx->n2e.resize(total > 0 ? total : 1));
    BHIP_TRY(ctx, ctx->n2e_loc.resize(total > 0 ? total : 1));
    BHIP_TRY(ctx, hipMemsetAsync(cnt.p, 0, sizeof(int) * (nn + 1), s));
    k_fill_incidence<<<grid_for(total, B), B, 0, s>>>(ctx->conn.p, total, ctx->nnpe,
                                                      ctx->n2e_ptr.p, cnt.p, ctx->n2e.p, ctx->n2e_loc.p);
    k_sort_incidence<<<grid_for(nn, B), B, 0, s>>>(nn, ctx->n2e_ptr.p, ctx->n2e.p, ctx->n2e_loc.p);
    BHIP_TRY(ctx, hipGetLastError());

    // ---- node adjacency: sort + unique the (node_a, node_b) keys
    long long nkeys = total * ctx->nnpe;
    dvec<unsigned long long> &keys = ctx->tmp_k1, &keys2 = ctx->tmp_k2;
    BHIP_TRY(ctx, keys.resize(nkeys > 0 ? nkeys : 1));
    BHIP_TRY(ctx, keys2.resize(nkeys > 0 ? nkeys : 1));
    k_pair_keys<<<grid_for(nkeys, B), B, 0, s>>>(ctx->conn.p, ctx->n_el, ctx->nnpe, keys.p);
    BHIP_TRY(ctx, hipGetLastError());
    int nbits = 1;
    while ((1LL << nbits) < (long long)nn) nbits++;
    // a mesh without elements has no keys: no sort, no pairs (every row then
    // holds only its diagonal)
    if (nkeys > 0) {
        size_t tb = 0;
        BHIP_TRY(ctx, hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys.p, keys2.p, (int)nkeys, 0,
                                                       32 + nbits, s));
        BHIP_TRY(ctx, ctx->work.resize(tb > 0 ? tb : 1));
        BHIP_TRY(ctx, hipcub::DeviceRadixSort::SortKeys(ctx->work.p, tb, keys.p, keys2.p, (int)nkeys, 0,
                                                       32 + nbits, s));
        BHIP_TRY(ctx, hipStreamSynchronize(s));
    }
    // the unique count lands in the per-DOF scratch, which is not in use yet
    BHIP_TRY(ctx, ctx->tmp_l1.resize(std::max<size_t>(ctx->tmp_l1.n, 1)));
    int *nsel = reinterpret_cast<int *>(ctx->tmp_l1.p);
    if (nkeys > 0) {
        size_t tb = 0;
        BHIP_TRY(ctx, hipcub::DeviceSelect::Unique(nullptr, tb, keys2.p, keys.p, nsel, (int)nkeys, s));
        BHIP_TRY(ctx, ctx->work.resize(tb > 0 ? tb : 1));
        BHIP_TRY(ctx, hipcub::DeviceSelect::Unique(ctx->work.p, tb, keys2.p, keys.p, nsel, (int)nkeys, s));
        BHIP_TRY(ctx, hipStreamSynchronize(s));
    }
    int h_nsel = 0;
    if (nkeys > 0) BHIP_TRY(ctx, bhip_copy(ctx, &h_nsel, nsel, sizeof(int), hipMemcpyDeviceToHost));
    ctx->n_pairs = h_nsel;
    long long np = ctx->n_pairs;
    BHIP_TRY(ctx, ctx->pair_row.resize(np > 0 ? np : 1));
    BHIP_TRY(ctx, ctx->nbr.resize(np > 0 ? np : 1));
    BHIP_TRY(ctx, ctx->nbr_off.resize(np > 0 ? np : 1));
    BHIP_TRY(ctx, hipMemsetAsync(cnt.p, 0, sizeof(int) * (nn + 1), s));
    k_split_keys<<<grid_for(np, B), B, 0, s>>>(keys.p, np, ctx->pair_row.p, ctx->nbr.p, cnt.p);
    BHIP_TRY(ctx, hipGetLastError());
    BHIP_TRY(ctx, ctx->nbr_ptr.resize(nn + 1));
    BHIP_PROPAGATE(exclusive_scan(ctx, cnt.p, ctx->nbr_ptr.p, nn + 1));

    // ---- DOF rows
    const int nd = ctx->n_dof;
    dvec<long long> &rowlen = ctx->tmp_l1, &rowscan = ctx->tmp_l2;
    BHIP_TRY(ctx, rowlen.resize(nd + 1));
    BHIP_TRY(ctx, rowscan.resize(nd + 1));
    BHIP_TRY(ctx, hipMemsetAsync(rowlen.p, 0, sizeof(long long) * (nd + 1), s));
    k_row_lengths<<<grid_for(nn, B), B, 0, s>>>(nn, ctx->nbr_ptr.p, ctx->nbr.p, ctx->dof_first.p,
                                                ctx->ndof.p, ctx->node_rmask.p, ctx->nbr_off.p,
                                                rowlen.p);
    BHIP_TRY(ctx, hipGetLastError());
    BHIP_PROPAGATE(exclusive_scan(ctx, rowlen.p, rowscan.p, nd + 1));
    long long h_nnz = 0;
    BHIP_TRY(ctx, bhip_copy(ctx, &h_nnz, rowscan.p + nd, sizeof(long long), hipMemcpyDeviceToHost));
    if (h_nnz >= (1LL << 31) - 1)
        return bhip_fail(ctx, BHIP_ERR_CAPACITY,
                         "pattern has " + std::to_string(h_nnz) +
                             " entries; int32 CSR indices hold at most 2^31-1");
    ctx->nnz = h_nnz;
    BHIP_TRY(ctx, ctx->rowptr.resize(nd + 1));
    k_rowptr32<<<grid_for(nd + 1, B), B, 0, s>>>(nd, rowscan.p, ctx->rowptr.p);
    BHIP_TRY(ctx, ctx->colind.resize(h_nnz > 0 ? h_nnz : 1));
    BHIP_TRY(ctx, ctx->diag_pos.resize(nd > 0 ? nd : 1));
    k_fill_cols<<<grid_for(np, B), B, 0, s>>>(np, ctx->pair_row.p, ctx->nbr.p, ctx->nbr_off.p,
                                              ctx->dof_first.p, ctx->ndof.p, ctx->node_rmask.p,
                                              ctx->rowptr.p, ctx->colind.p, ctx->diag_pos.p);
    k_fill_diag_rows<<<grid_for(nn, B), B, 0, s>>>(nn, ctx->nbr_ptr.p, ctx->dof_first.p,
                                                   ctx->ndof.p, ctx->node_rmask.p, ctx->rowptr.p,
                                                   ctx->colind.p, ctx->diag_pos.p);
    BHIP_TRY(ctx, hipGetLastError());
    // lower count
    k_lower_len<<<grid_for(nd, B), B, 0, s>>>(nd, ctx->rowptr.p, ctx->diag_pos.p, rowlen.p);
    BHIP_PROPAGATE(exclusive_scan(ctx, rowlen.p, rowscan.p, nd + 1));
    long long h_low = 0;
    BHIP_TRY(ctx, bhip_copy(ctx, &h_low, rowscan.p + nd, sizeof(long long), hipMemcpyDeviceToHost));
    ctx->nnz_lower = h_low;

    BHIP_TRY(ctx, ctx->val.resize(h_nnz > 0 ? h_nnz : 1));
    BHIP_TRY(ctx, hipMemsetAsync(ctx->val.p, 0, ctx->val.bytes(), s));
    BHIP_TRY(ctx, hipStreamSynchronize(s));
    ctx->pattern_ready = true;
    ctx->assembled = false;
    ctx->f32_ready = ctx->f64b_ready = false;
    ctx->sell_shape_set = false;  // a tuned shape belongs to the pattern it was measured on
    // the FP32 layouts belong to the old pattern; their buffers are kept for reuse
    ctx->sell_layout_ok = false;
    ctx->val32.n = 0;
    return BHIP_OK;
}

// Export the lower triangle in the reference layout.
int bhip_export_lower(bhip_ctx *ctx, int *rowptr, int *colind, double *val) {
    hipStream_t s = ctx->stream;
    const int B = 256;
    const int nd = ctx->n_dof;
    scoped_dvec<long long> len, scan;
    BHIP_TRY(ctx, len.resize(nd + 1));
    BHIP_TRY(ctx, scan.resize(nd + 1));
    BHIP_TRY(ctx, hipMemsetAsync(len.p, 0, sizeof(long long) * (nd + 1), s));
    k_lower_len<<<grid_for(nd, B), B, 0, s>>>(nd, ctx->rowptr.p, ctx->diag_pos.p, len.p);
    BHIP_PROPAGATE(exclusive_scan(ctx, len.p, scan.p, nd + 1));
    std::vector<long long> h(nd + 1);
    BHIP_TRY(ctx, bhip_copy(ctx, h.data(), scan.p, sizeof(long long) * (nd + 1), hipMemcpyDeviceToHost));
    if (rowptr)
        for (int i = 0; i <= nd; i++) rowptr[i] = (int)h[i];
    long long nl = h[nd];
    scoped_dvec<int> dcol;
    scoped_dvec<double> dval;
    if (colind) BHIP_TRY(ctx, dcol.resize(nl > 0 ? nl : 1));
    if (val) BHIP_TRY(ctx, dval.resize(nl > 0 ? nl : 1));
    k_copy_lower<<<grid_for(nd, B), B, 0, s>>>(nd, ctx->rowptr.p, ctx->colind.p, ctx->val.p, scan.p,
                                               colind ? dcol.p : nullptr, val ? dval.p : nullptr);
    BHIP_TRY(ctx, hipGetLastError());
    BHIP_TRY(ctx, hipStreamSynchronize(s));
    if (colind) BHIP_TRY(ctx, bhip_copy(ctx, colind, dcol.p, sizeof(int) * nl, hipMemcpyDeviceToHost));
    if (val) BHIP_TRY(ctx, bhip_copy(ctx, val, dval.p, sizeof(double) * nl, hipMemcpyDeviceToHost));
    return BHIP_OK;
}
