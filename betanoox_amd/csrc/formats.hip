// formats.hip -- K / b exchange files in the reference's formats (host code).
//
// Writers
//  - RCV text, row major, one row per line of "\t i j value" triples after a
//    type line (1 = lower, 2 = upper) and "L nnz": the global-stiffness dump
//    of SymmetricSparseMatrix::printMatrixAddressFormat / ...Upper
//    (sparse.cpp:266-325), values in setw(24) setprecision(16) of a stream
//    left in its default float format.
//  - CSR ASCII / binary: writeCSRMatrixInASCII / writeCSRMatrixInBinary
//    (SparseMatrixUtility.cpp:11-38), called by PardisoSymmMatrix for its
//    storage (Pardiso.cpp:208-216): upper triangle, 1-based row and column
//    indices, diagonal first in every row.  ASCII: "L ncoefs", the row
//    pointers, the column indices, then one scientific setw(24)
//    setprecision(16) value per line and an empty line.  Binary: int L,
//    int ncoefs, then the three arrays.
//  - LoadVector.txt / Solution.txt of Equations::solve (equation.cpp:1357-
//    1383): scientific setw(24) setprecision(16), one value per line; the
//    solution file starts with the count.
// Readers take the same files back (SparseMatrixInterface::ReadMatrix,
// sparseMatrixInterface.cpp:51-101; readCSRMatrixInASCII/Binary,
// SparseMatrixUtility.cpp:40-78) and load the matrix into a context, so a
// system the reference dumped can be solved on the device.
#include <algorithm>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <vector>

#include "bhip_internal.hpp"

namespace {

struct Csr {
    int n = 0;
    std::vector<long long> ptr;
    std::vector<int> col;
    std::vector<double> val;
};

// Lower triangle (0-based, diagonal last in each row) of the context's matrix.
int lower_of(bhip_ctx *ctx, Csr &L) {
    if (!ctx->pattern_ready) return bhip_fail(ctx, BHIP_ERR_STATE, "no matrix in the context");
    const int n = ctx->n_dof;
    std::vector<int> rp(n + 1);
    BHIP_PROPAGATE(bhip_export_lower(ctx, rp.data(), nullptr, nullptr));
    const long long nl = rp[n];
    L.n = n;
    L.ptr.assign(rp.begin(), rp.end());
    L.col.resize(nl > 0 ? nl : 1);
    L.val.resize(nl > 0 ? nl : 1);
    BHIP_PROPAGATE(bhip_export_lower(ctx, nullptr, L.col.data(), L.val.data()));
    L.col.resize(nl);
    L.val.resize(nl);
    return BHIP_OK;
}

// Transpose of a triangle stored by rows (sorted rows stay sorted).
Csr transpose(const Csr &A) {
    Csr T;
    T.n = A.n;
    T.ptr.assign(A.n + 1, 0);
    for (int c : A.col) T.ptr[c + 1]++;
    for (int i = 0; i < A.n; i++) T.ptr[i + 1] += T.ptr[i];
    T.col.resize(A.col.size());
    T.val.resize(A.val.size());
    std::vector<long long> pos(T.ptr.begin(), T.ptr.end() - 1);
    for (int i = 0; i < A.n; i++)
        for (long long k = A.ptr[i]; k < A.ptr[i + 1]; k++) {
            const long long d = pos[A.col[k]]++;
            T.col[d] = i;
            T.val[d] = A.val[k];
        }
    return T;
}

int write_rcv(const char *path, int type, int n, const long long *ptr, const int *col, const double *val) {
    std::ofstream os(path);
    if (!os) return BHIP_ERR_ARG;
    os << type << std::endl;
    os << n << " " << ptr[n] << std::endl;
    for (int i = 0; i < n; i++) {
        for (long long k = ptr[i]; k < ptr[i + 1]; k++)
            os << '\t' << i << " " << col[k] << " " << std::setw(24) << std::setprecision(16) << val[k];
        os << '\n';
    }
    return os.good() ? BHIP_OK : BHIP_ERR_ARG;
}

int write_csr_ascii(const char *path, int n, int ncoefs, const int *ia, const int *ja, const double *a) {
    std::ofstream os(path);
    if (!os) return BHIP_ERR_ARG;
    os << std::resetiosflags(std::ios::floatfield) << std::setiosflags(std::ios::scientific);
    os << n << " " << ncoefs << std::endl;
    for (int i = 0; i <= n; i++) os << ia[i] << " ";
    os << std::endl;
    for (int k = 0; k < ncoefs; k++) os << ja[k] << " ";
    os << std::endl;
    for (int k = 0; k < ncoefs; k++) os << std::setw(24) << std::setprecision(16) << a[k] << std::endl;
    os << std::endl;
    return os.good() ? BHIP_OK : BHIP_ERR_ARG;
}

int write_csr_binary(const char *path, int n, int ncoefs, const int *ia, const int *ja, const double *a) {
    std::ofstream os(path, std::ios::out | std::ios::binary);
    if (!os) return BHIP_ERR_ARG;
    os.write(reinterpret_cast<const char *>(&n), sizeof(int));
    os.write(reinterpret_cast<const char *>(&ncoefs), sizeof(int));
    os.write(reinterpret_cast<const char *>(ia), sizeof(int) * ((size_t)n + 1));
    os.write(reinterpret_cast<const char *>(ja), sizeof(int) * (size_t)ncoefs);
    os.write(reinterpret_cast<const char *>(a), sizeof(double) * (size_t)ncoefs);
    return os.good() ? BHIP_OK : BHIP_ERR_ARG;
}

// Upper triangle, 1-based, as the CSR files hold it.
int write_upper_csr(const char *path, int format, const Csr &U) {
    if (U.ptr[U.n] >= (1LL << 31) - 1) return BHIP_ERR_CAPACITY;
    const int nc = (int)U.ptr[U.n];
    std::vector<int> ia(U.n + 1), ja(nc);
    for (int i = 0; i <= U.n; i++) ia[i] = (int)U.ptr[i] + 1;
    for (int k = 0; k < nc; k++) ja[k] = U.col[k] + 1;
    return format == BHIP_FMT_CSR_ASCII ? write_csr_ascii(path, U.n, nc, ia.data(), ja.data(), U.val.data())
                                        : write_csr_binary(path, U.n, nc, ia.data(), ja.data(), U.val.data());
}

// Any triangle or both, any index base -> full symmetric CSR with sorted
// columns; triangle: BHIP_TRI_*.
int expand_full(int n, long long nnz, const int *ia, const int *ja, const double *a, int base, int triangle,
                std::vector<int> &rp, std::vector<int> &ci, std::vector<double> &v, std::vector<int> &diag,
                long long &nnz_lower, std::string &why) {
    struct E {
        int c;
        double v;
    };
    std::vector<std::vector<E>> rows(n);
    for (int i = 0; i < n; i++) {
        const long long kb = ia[i] - base, ke = ia[i + 1] - base;
        if (kb < 0 || ke < kb || ke > nnz) {
            why = "row pointers out of range";
            return BHIP_ERR_ARG;
        }
        for (long long k = kb; k < ke; k++) {
            const int j = ja[k] - base;
            if (j < 0 || j >= n) {
                why = "column index out of range";
                return BHIP_ERR_ARG;
            }
            if ((triangle == BHIP_TRI_LOWER && j > i) || (triangle == BHIP_TRI_UPPER && j < i)) {
                why = "entry outside the declared triangle";
                return BHIP_ERR_ARG;
            }
            rows[i].push_back({j, a[k]});
            if (triangle != BHIP_TRI_FULL && j != i) rows[j].push_back({i, a[k]});
        }
    }
    rp.assign(n + 1, 0);
    diag.assign(n, -1);
    nnz_lower = 0;
    long long tot = 0;
    for (int i = 0; i < n; i++) {
        auto &r = rows[i];
        std::sort(r.begin(), r.end(), [](const E &x, const E &y) { return x.c < y.c; });
        for (size_t k = 1; k < r.size(); k++)
            if (r[k].c == r[k - 1].c) {
                why = "duplicate entry in row " + std::to_string(i);
                return BHIP_ERR_ARG;
            }
        tot += (long long)r.size();
        if (tot >= (1LL << 31)) {
            why = "matrix too large for int32 row pointers";
            return BHIP_ERR_CAPACITY;
        }
        rp[i + 1] = (int)tot;
    }
    ci.resize(tot);
    v.resize(tot);
    for (int i = 0; i < n; i++) {
        int k = rp[i];
        for (const E &e : rows[i]) {
            if (e.c == i) diag[i] = k;
            if (e.c <= i) nnz_lower++;
            ci[k] = e.c;
            v[k] = e.v;
            k++;
        }
        if (diag[i] < 0) {
            why = "row " + std::to_string(i) + " has no diagonal entry";
            return BHIP_ERR_ZERO_DIAG;
        }
    }
    return BHIP_OK;
}

// Parse a file into (n, nnz, ia, ja, a, base, triangle).
int read_file(const char *path, int format, int &n, std::vector<int> &ia, std::vector<int> &ja, std::vector<double> &a,
              int &base, int &triangle, std::string &why) {
    if (format == BHIP_FMT_CSR_BINARY) {
        std::ifstream is(path, std::ios::in | std::ios::binary);
        int nc = 0;
        if (!is.read(reinterpret_cast<char *>(&n), sizeof(int)) || !is.read(reinterpret_cast<char *>(&nc), sizeof(int)) ||
            n < 0 || nc < 0) {
            why = "cannot read the CSR binary header";
            return BHIP_ERR_ARG;
        }
        ia.resize((size_t)n + 1);
        ja.resize(nc);
        a.resize(nc);
        is.read(reinterpret_cast<char *>(ia.data()), sizeof(int) * ((size_t)n + 1));
        is.read(reinterpret_cast<char *>(ja.data()), sizeof(int) * (size_t)nc);
        is.read(reinterpret_cast<char *>(a.data()), sizeof(double) * (size_t)nc);
        if (!is) {
            why = "CSR binary file is truncated";
            return BHIP_ERR_ARG;
        }
    } else if (format == BHIP_FMT_CSR_ASCII) {
        std::ifstream is(path);
        long long nc = 0;
        if (!(is >> n >> nc) || n < 0 || nc < 0) {
            why = "cannot read the CSR ASCII header";
            return BHIP_ERR_ARG;
        }
        ia.resize((size_t)n + 1);
        ja.resize(nc);
        a.resize(nc);
        for (auto &x : ia) is >> x;
        for (auto &x : ja) is >> x;
        for (auto &x : a) is >> x;
        if (!is) {
            why = "CSR ASCII file is truncated";
            return BHIP_ERR_ARG;
        }
    } else if (format == BHIP_FMT_RCV_LOWER || format == BHIP_FMT_RCV_UPPER || format == BHIP_FMT_RCV) {
        std::ifstream is(path);
        int type = 0;
        long long nz = 0;
        if (!(is >> type >> n >> nz) || type < 1 || type > 3 || n < 0 || nz < 0) {
            why = "cannot read the RCV header (type 1-3, L, nnz)";
            return BHIP_ERR_ARG;
        }
        std::vector<std::vector<std::pair<int, double>>> rows(n);
        for (long long k = 0; k < nz; k++) {
            int i, j;
            double x;
            if (!(is >> i >> j >> x) || i < 0 || i >= n) {
                why = "RCV triple out of range or file truncated";
                return BHIP_ERR_ARG;
            }
            rows[i].push_back({j, x});
        }
        ia.assign((size_t)n + 1, 0);
        for (int i = 0; i < n; i++) ia[i + 1] = ia[i] + (int)rows[i].size();
        ja.resize(nz);
        a.resize(nz);
        for (int i = 0; i < n; i++)
            for (size_t k = 0; k < rows[i].size(); k++) {
                ja[ia[i] + k] = rows[i][k].first;
                a[ia[i] + k] = rows[i][k].second;
            }
        base = 0;
        triangle = type == 1 ? BHIP_TRI_LOWER : type == 2 ? BHIP_TRI_UPPER : BHIP_TRI_FULL;
        return BHIP_OK;
    } else {
        why = "unknown matrix file format";
        return BHIP_ERR_ARG;
    }
    // CSR files: base from the first row pointer, triangle from the entries
    base = (n >= 0 && !ia.empty() && ia[0] == 1) ? 1 : 0;
    bool lower = true, upper = true;
    for (int i = 0; i < n; i++)
        for (int k = ia[i] - base; k < ia[i + 1] - base && k < (int)ja.size(); k++) {
            if (k < 0) break;
            const int j = ja[k] - base;
            if (j > i) lower = false;
            if (j < i) upper = false;
        }
    triangle = upper ? BHIP_TRI_UPPER : lower ? BHIP_TRI_LOWER : BHIP_TRI_FULL;
    return BHIP_OK;
}

}  // namespace

extern "C" {

int bhip_write_matrix(bhip_ctx *ctx, const char *path, int format) {
    if (!ctx) return BHIP_ERR_ARG;
    if (!path) return bhip_fail(ctx, BHIP_ERR_ARG, "path is required");
    Csr L;
    BHIP_PROPAGATE(lower_of(ctx, L));
    int st;
    if (format == BHIP_FMT_RCV_LOWER) {
        st = write_rcv(path, 1, L.n, L.ptr.data(), L.col.data(), L.val.data());
    } else if (format == BHIP_FMT_RCV_UPPER) {
        const Csr U = transpose(L);
        st = write_rcv(path, 2, U.n, U.ptr.data(), U.col.data(), U.val.data());
    } else if (format == BHIP_FMT_CSR_ASCII || format == BHIP_FMT_CSR_BINARY) {
        st = write_upper_csr(path, format, transpose(L));
    } else {
        return bhip_fail(ctx, BHIP_ERR_ARG, "unknown matrix file format");
    }
    if (st != BHIP_OK) return bhip_fail(ctx, st, std::string("cannot write ") + path);
    return BHIP_OK;
}

int bhip_write_csr_file(const char *path, int format, int n, long long nnz, const int *ia, const int *ja,
                        const double *a) {
    if (!path || n < 0 || nnz < 0 || !ia || (nnz > 0 && (!ja || !a))) return BHIP_ERR_ARG;
    if (format == BHIP_FMT_CSR_ASCII || format == BHIP_FMT_CSR_BINARY) {
        if (nnz >= (1LL << 31)) return BHIP_ERR_CAPACITY;
        return format == BHIP_FMT_CSR_ASCII ? write_csr_ascii(path, n, (int)nnz, ia, ja, a)
                                            : write_csr_binary(path, n, (int)nnz, ia, ja, a);
    }
    if (format == BHIP_FMT_RCV_LOWER || format == BHIP_FMT_RCV_UPPER) {
        std::vector<long long> p(ia, ia + n + 1);
        if (p[n] != nnz) return BHIP_ERR_ARG;
        return write_rcv(path, format == BHIP_FMT_RCV_LOWER ? 1 : 2, n, p.data(), ja, a);
    }
    return BHIP_ERR_ARG;
}

int bhip_set_csr(bhip_ctx *ctx, int n, long long nnz, const int *ia, const int *ja, const double *a, int base,
                 int triangle) {
    if (!ctx) return BHIP_ERR_ARG;
    if (n <= 0 || nnz < 0 || !ia || (nnz > 0 && (!ja || !a)) || (base != 0 && base != 1) ||
        triangle < BHIP_TRI_LOWER || triangle > BHIP_TRI_FULL)
        return bhip_fail(ctx, BHIP_ERR_ARG, "bad CSR arguments");
    (void)hipSetDevice(ctx->device);
    std::vector<int> rp, ci, diag;
    std::vector<double> v;
    long long nl = 0;
    std::string why;
    const int st = expand_full(n, nnz, ia, ja, a, base, triangle, rp, ci, v, diag, nl, why);
    if (st != BHIP_OK) return bhip_fail(ctx, st, why);
    hipStream_t s = ctx->stream;
    bhip_invalidate_graphs(ctx);
    bhip_multi_release(ctx);
    bhip_plastic_release(ctx);
    // no mesh behind this matrix: assembly and the mesh-based layouts are off
    ctx->dim = ctx->n_nodes = ctx->n_el = ctx->nnpe = 0;
    ctx->n_dof = n;
    ctx->max_ndof = 0;
    ctx->uniform_ndof = false;
    ctx->elem_ndof_full = false;
    ctx->have_rest = ctx->have_mat = false;
    ctx->f32_ready = ctx->f64b_ready = ctx->use_bsr = ctx->lift_ready = false;
    ctx->sell_b = 0;
    ctx->sell_union = false;
    ctx->sell_ptr.release();
    ctx->sell_col.release();
    ctx->bval32.release();
    ctx->bval64.release();
    ctx->val_aux.release();
    ctx->sell_layout_ok = ctx->bval64_layout_ok = false;
    ctx->n_pairs = 0;
    ctx->nnz = (long long)ci.size();
    ctx->nnz_lower = nl;
    BHIP_TRY(ctx, ctx->rowptr.resize((size_t)n + 1));
    BHIP_TRY(ctx, ctx->colind.resize(ci.size() > 0 ? ci.size() : 1));
    BHIP_TRY(ctx, ctx->val.resize(v.size() > 0 ? v.size() : 1));
    BHIP_TRY(ctx, ctx->diag_pos.resize(n));
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->rowptr.p, rp.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice, s));
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->colind.p, ci.data(), sizeof(int) * ci.size(), hipMemcpyHostToDevice, s));
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->val.p, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, s));
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->diag_pos.p, diag.data(), sizeof(int) * n, hipMemcpyHostToDevice, s));
    BHIP_PROPAGATE(bhip_ensure_vectors(ctx));
    BHIP_TRY(ctx, hipStreamSynchronize(s));
    ctx->pattern_ready = true;
    ctx->assembled = true;
    ctx->asm_kind = 0;
    ctx->external_matrix = true;
    return BHIP_OK;
}

int bhip_read_matrix(bhip_ctx *ctx, const char *path, int format) {
    if (!ctx) return BHIP_ERR_ARG;
    if (!path) return bhip_fail(ctx, BHIP_ERR_ARG, "path is required");
    int n = 0, base = 0, tri = 0;
    std::vector<int> ia, ja;
    std::vector<double> a;
    std::string why;
    const int st = read_file(path, format, n, ia, ja, a, base, tri, why);
    if (st != BHIP_OK) return bhip_fail(ctx, st, std::string(path) + ": " + why);
    return bhip_set_csr(ctx, n, (long long)ja.size(), ia.data(), ja.data(), a.data(), base, tri);
}

int bhip_write_vector(const char *path, int kind, int n, const double *v) {
    if (!path || n < 0 || (n > 0 && !v) || (kind != BHIP_VEC_LOAD && kind != BHIP_VEC_SOLUTION)) return BHIP_ERR_ARG;
    std::ofstream os(path);
    if (!os) return BHIP_ERR_ARG;
    os.setf(std::ios::scientific, std::ios::floatfield);
    if (kind == BHIP_VEC_SOLUTION) os << n << std::endl;
    for (int i = 0; i < n; i++) os << std::setw(24) << std::setprecision(16) << v[i] << std::endl;
    return os.good() ? BHIP_OK : BHIP_ERR_ARG;
}

int bhip_read_vector(const char *path, int kind, int n, double *v) {
    if (!path || n < 0 || (n > 0 && !v) || (kind != BHIP_VEC_LOAD && kind != BHIP_VEC_SOLUTION)) return BHIP_ERR_ARG;
    std::ifstream is(path);
    if (!is) return BHIP_ERR_ARG;
    if (kind == BHIP_VEC_SOLUTION) {
        int m = -1;
        if (!(is >> m) || m != n) return BHIP_ERR_ARG;
    }
    for (int i = 0; i < n; i++)
        if (!(is >> v[i])) return BHIP_ERR_ARG;
    return BHIP_OK;
}

}  // extern "C"
