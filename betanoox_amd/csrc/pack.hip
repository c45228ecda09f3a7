// pack.hip -- packed batched inner PCG: the load cases of several assembled
// systems solved together by one persistent set of kernels.
//
// Reference workflow: the column loop of MultiFieldMacroElement::CalculateK /
// solveForStiffnessMatrixColumn (MultiFieldMacroElement.cpp:135-407), where
// every load case of a macro element is solved with the same assembled K
// (the prescribed values re-lifted through K, Equations::
// updatePseudoLoadVectorUsingGK, equation.cpp:343-381).  The reference runs one
// model at a time; here a pack holds several models (systems), each with R
// load cases, and their FP32 inner Jacobi-PCGs advance together.
//
// Design
//  - A system is an assembled context with 3 DOFs per node and its FP32
//    SELL-64 3x3-block matrix; its vectors are case-major (case r, DOF i at
//    r * n + i).  Every case keeps its own CG recurrence and stopping test.
//  - Work items: an SpMV item is one 64-node slice of a small system (four
//    waves share the slice) or four slices of a large one (one wave each); a
//    vector item is one DOF range of one case of one system.  Item tables map
//    items to systems.
//  - Persistent kernels: a fixed grid walks the items (grid stride).  The
//    kernels take only the address of a device-side pack descriptor and the
//    iteration parity, so one captured HIP graph of a chunk of iterations
//    serves every pack composition; the host rewrites the descriptor between
//    inner solves.
//  - Three kernels per iteration: (1) q = K p for all cases, one record of p.q
//    per SpMV item and case; (2) every vector item sums its system's p.q
//    records itself (fixed order), alpha, x and r update, z, records of r.z
//    and r.r; (3) every vector item sums its system's (r.z, r.r) records,
//    beta, p = z + beta p; the first item of a case writes its state (stop
//    test, iteration count; r.z double-buffered by parity).
//  - Determinism: a system's records, their order and its item decomposition
//    depend on that system alone, so its results are the same bits in any
//    pack, alone or beside others.
//  - Outer loop: FP64 residual b - K x per system on its FP64 CSR
//    (multi.hip), scaling into the FP32 inner right-hand side, inner packed
//    solve of the systems not yet converged, x += |r| d.  A system leaves the
//    pack when all its cases meet the tolerance.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "bhip_internal.hpp"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxR = 8;
constexpr int kChunk = 32;             // iterations per captured graph (even: r.z parity)
constexpr int kSplitSlices = 1024;     // systems with fewer slices run four waves per slice
constexpr double kStreamBytes = 240e6;
constexpr int kPackUnroll = 2;         // SpMV block columns per loop trip (plain-load packs)
constexpr int kPackUnrollStream = 3;   // the same for streaming packs (large systems)
constexpr int kBatch = 4;              // elements per thread and trip in the vector kernels  // packs whose SELL arrays exceed the 256 MB last-level cache stream

struct PackState {
    double rzb[2][kMaxR];  // r.z by iteration parity
    double rr[kMaxR], rr0[kMaxR];
    int done[kMaxR], iter[kMaxR], converged[kMaxR];
    int pad;
};

struct PackSysDesc {
    const int *sptr, *scol;  // scol: node indices (wide form)
    const short *scol16;     // offsets from the row's node (narrow form)
    const int *uoff;         // union slots: the slices' offset lists (nullptr: rank slots)
    int narrow, pad0;
    const float *sval;
    long long nsell;
    const float *dinv;
    float *r, *p, *q, *d;  // case-major R x n
    long long n;           // DOFs
    int n_nodes, n_slices;
    int split;             // 1: one slice per SpMV item, 0: four slices per item
    int n_sitems, sitem0;  // SpMV items of the system, first item
    int n_vitems, vitem0;  // vector items per case, first item
    int maxit;             // inner iteration cap of this solve
    double tol;            // inner relative residual reduction
    long long rec_s;       // p.q records: R x n_sitems, case-major
    long long rec_v;       // (r.z, r.r) records: R x n_vitems pairs
    PackState *st;
};

struct PackHdr {
    int n_sys, R, n_sitems, n_vitems;  // n_vitems: vector items per case over all systems
    const PackSysDesc *sys;
    const int *sitem_sys, *vitem_sys;
    double *rec;
};

template <int NT, class T>
__device__ __forceinline__ T ld_stream(const BHIP_GLOBAL T *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// Fixed-order sum of nb records of NV doubles by one block, result in every
// thread (thread t adds records t, t + 256, ...; eight loads in flight).
template <int NV>
__device__ __forceinline__ void sum_records_all(const BHIP_GLOBAL double *rec, int nb, double (&tot)[NV]) {
    __shared__ double sm[16 * NV];
    __shared__ double res[NV];
    constexpr int U = 8;
#pragma unroll
    for (int k = 0; k < NV; k++) tot[k] = 0;
    for (int b0 = threadIdx.x; b0 < nb; b0 += U * blockDim.x) {
        double v[U][NV];
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int b = b0 + u * blockDim.x;
#pragma unroll
            for (int k = 0; k < NV; k++) v[u][k] = b < nb ? rec[(long long)b * NV + k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; u++)
#pragma unroll
            for (int k = 0; k < NV; k++) tot[k] += v[u][k];
    }
    block_sum<NV>(tot, sm);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) res[k] = tot[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) tot[k] = res[k];
}

__device__ __forceinline__ void item_range(long long n, int lb, int nb, long long &b, long long &e) {
    long long per = (n + nb - 1) / nb;
    b = (long long)lb * per;
    e = b + per;
    if (e > n) e = n;
    if (b > n) b = n;
}

// Row sums of one lane over block columns k = k0, k0 + kstep, ... < ke of its
// node's row for all R cases (CT: wide node indices or narrow offsets).
template <int R, int NT, int UNR, class CT>
__device__ __forceinline__ void pk_rows(const BHIP_GLOBAL CT *__restrict__ scol, const BHIP_GLOBAL float *__restrict__ sval,
                                        long long nsell, const BHIP_GLOBAL float *__restrict__ x, long long ld,
                                        long long node, int k0, int ke, int kstep, float (&a0)[R], float (&a1)[R],
                                        float (&a2)[R]) {
#pragma unroll UNR
    for (int k = k0; k < ke; k += kstep) {
        long long m;
        if constexpr (sizeof(CT) == 2) m = node + (long long)ld_stream<NT>(scol + k);
        else m = ld_stream<NT>(scol + k);
        const float v0 = ld_stream<NT>(sval + 0 * nsell + k), v1 = ld_stream<NT>(sval + 1 * nsell + k);
        const float v2 = ld_stream<NT>(sval + 2 * nsell + k), v3 = ld_stream<NT>(sval + 3 * nsell + k);
        const float v4 = ld_stream<NT>(sval + 4 * nsell + k), v5 = ld_stream<NT>(sval + 5 * nsell + k);
        const float v6 = ld_stream<NT>(sval + 6 * nsell + k), v7 = ld_stream<NT>(sval + 7 * nsell + k);
        const float v8 = ld_stream<NT>(sval + 8 * nsell + k);
        const BHIP_GLOBAL float *xm = x + 3 * m;
#pragma unroll
        for (int r = 0; r < R; r++) {
            const float x0 = xm[r * ld], x1 = xm[r * ld + 1], x2 = xm[r * ld + 2];
            a0[r] += v0 * x0 + v1 * x1 + v2 * x2;
            a1[r] += v3 * x0 + v4 * x1 + v5 * x2;
            a2[r] += v6 * x0 + v7 * x1 + v8 * x2;
        }
    }
}

// Union slots: the same row sums over runs of up to three consecutive offsets
// of the slice's offset list (as in the one-case union SpMV): x is gathered
// at node + o and node + o + len - 1, and the middle of a run of three comes
// from the next lane (lane 63: lane 62's second gather) by DPP.  The wave
// takes runs sel, sel + nsel, ... in order; every lane of the slice runs the
// loop (zero blocks and clamped gathers where a lane has no such neighbour).
typedef float f3u __attribute__((ext_vector_type(3), aligned(4)));
__device__ __forceinline__ float pk_lane_next(float old, float src) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), 0x130, 0xf, 0xf, false));
}
__device__ __forceinline__ float pk_lane_prev(float src) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(src), 0x138, 0xf, 0xf, false));
}

// Loads from a uniform base plus a 32-bit byte offset per lane (the
// scalar-base form of the global loads: one offset register serves every
// value array and every case).  The bases pass through readfirstlane so that
// they stay scalar and are not folded into per-lane 64-bit addresses.  Union
// slots are only built where every such offset fits 32 bits.
template <class T>
__device__ __forceinline__ const BHIP_GLOBAL T *pk_uniform(const BHIP_GLOBAL T *p) {
    const unsigned long long v = (unsigned long long)p;
    const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)v);
    const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(v >> 32));
    return (const BHIP_GLOBAL T *)(((unsigned long long)hi << 32) | lo);
}

template <int NT, class T>
__device__ __forceinline__ T pk_ld(const BHIP_GLOBAL T *base, unsigned byte_off) {
    return ld_stream<NT>((const BHIP_GLOBAL T *)((const BHIP_GLOBAL char *)base + byte_off));
}

template <int NT>
__device__ __forceinline__ void pk_load_block(const BHIP_GLOBAL float *const (&vb)[9], unsigned koff, float (&v)[9]) {
#pragma unroll
    for (int e = 0; e < 9; e++) v[e] = pk_ld<NT>(vb[e], koff);
}

// Cases r0 <= r < r1 of a block (xs indexed from r0).
template <int R, int R0, int R1>
__device__ __forceinline__ void pk_block_rows(const float (&v)[9], const float (&xs)[R1 - R0][3], float (&a0)[R],
                                              float (&a1)[R], float (&a2)[R]) {
#pragma unroll
    for (int r = R0; r < R1; r++) {
        const float x0 = xs[r - R0][0], x1 = xs[r - R0][1], x2 = xs[r - R0][2];
        a0[r] += v[0] * x0 + v[1] * x1 + v[2] * x2;
        a1[r] += v[3] * x0 + v[4] * x1 + v[5] * x2;
        a2[r] += v[6] * x0 + v[7] * x1 + v[8] * x2;
    }
}

template <int R, int R0, int R1>
__device__ __forceinline__ void pk_gather(const BHIP_GLOBAL f3u *const (&xb)[R], unsigned moff, float (&xs)[R1 - R0][3]) {
#pragma unroll
    for (int r = R0; r < R1; r++) {
        const f3u t = pk_ld<0>(xb[r], moff);
        xs[r - R0][0] = t.x;
        xs[r - R0][1] = t.y;
        xs[r - R0][2] = t.z;
    }
}

// The three blocks of a run of three consecutive offsets for cases r0..r1-1.
template <int R, int R0, int R1>
__device__ __forceinline__ void pk_run3(const float (&v0)[9], const float (&v1)[9], const float (&v2)[9],
                                        const BHIP_GLOBAL f3u *const (&xb)[R], unsigned o0, unsigned o2,
                                        float (&a0)[R], float (&a1)[R], float (&a2)[R]) {
    constexpr int N = R1 - R0;
    float x0[N][3], x1[N][3], x2[N][3];
    pk_gather<R, R0, R1>(xb, o0, x0);
    pk_gather<R, R0, R1>(xb, o2, x2);
#pragma unroll
    for (int r = 0; r < N; r++)
#pragma unroll
        for (int d = 0; d < 3; d++) x1[r][d] = pk_lane_next(pk_lane_prev(x2[r][d]), x0[r][d]);
    pk_block_rows<R, R0, R1>(v0, x0, a0, a1, a2);
    pk_block_rows<R, R0, R1>(v1, x1, a0, a1, a2);
    pk_block_rows<R, R0, R1>(v2, x2, a0, a1, a2);
}

// Each run length has its own branch that issues every load of the run
// before its multiply-adds.
template <int R, int NT>
__device__ __forceinline__ void pk_rows_union(const BHIP_GLOBAL int *__restrict__ of, const BHIP_GLOBAL float *__restrict__ sval,
                                              long long nsell, const BHIP_GLOBAL float *__restrict__ x, long long ld,
                                              long long node, long long last, long long kb, int W, int sel, int nsel,
                                              float (&a0)[R], float (&a1)[R], float (&a2)[R]) {
    const int lane = threadIdx.x & 63;
    const BHIP_GLOBAL float *vb[9];
#pragma unroll
    for (int e = 0; e < 9; e++) vb[e] = pk_uniform(sval + e * nsell);
    const BHIP_GLOBAL f3u *xb[R];
#pragma unroll
    for (int r = 0; r < R; r++) xb[r] = pk_uniform((const BHIP_GLOBAL f3u *)(x + r * ld));
    int run = 0;
    for (int j = 0; j < W; run++) {
        const int o = of[j];
        const int len = (j + 2 < W && of[j + 2] == o + 2) ? 3 : (j + 1 < W && of[j + 1] == o + 1) ? 2 : 1;
        if (run % nsel == sel) {
            const unsigned koff = (unsigned)(kb + 64LL * j + lane) * 4u;
            long long m0 = node + o, m2 = node + o + len - 1;
            m0 = m0 < 0 ? 0 : (m0 > last ? last : m0);
            m2 = m2 < 0 ? 0 : (m2 > last ? last : m2);
            const unsigned o0 = (unsigned)m0 * 12u, o2 = (unsigned)m2 * 12u;
            if (len == 3) {
                // the cases in two halves: the second half's gathers reuse
                // the first half's registers
                float v0[9], v1[9], v2[9];
                pk_load_block<NT>(vb, koff, v0);
                pk_load_block<NT>(vb, koff + 256u, v1);
                pk_load_block<NT>(vb, koff + 512u, v2);
                if constexpr (R >= 2) {
                    pk_run3<R, 0, R / 2>(v0, v1, v2, xb, o0, o2, a0, a1, a2);
                    pk_run3<R, R / 2, R>(v0, v1, v2, xb, o0, o2, a0, a1, a2);
                } else {
                    pk_run3<R, 0, R>(v0, v1, v2, xb, o0, o2, a0, a1, a2);
                }
            } else if (len == 2) {
                float v0[9], v1[9], x0[R][3], x1[R][3];
                pk_load_block<NT>(vb, koff, v0);
                pk_load_block<NT>(vb, koff + 256u, v1);
                pk_gather<R, 0, R>(xb, o0, x0);
                pk_gather<R, 0, R>(xb, o2, x1);
                pk_block_rows<R, 0, R>(v0, x0, a0, a1, a2);
                pk_block_rows<R, 0, R>(v1, x1, a0, a1, a2);
            } else {
                float v0[9], x0[R][3];
                pk_load_block<NT>(vb, koff, v0);
                pk_gather<R, 0, R>(xb, o0, x0);
                pk_block_rows<R, 0, R>(v0, x0, a0, a1, a2);
            }
        }
        j += len;
    }
}

// q_r = K p_r for every case of every live system; record of p_r.q_r per
// item.  Small systems: the four waves share the item's slice, each taking
// every fourth block column, and wave 0 adds the others' row sums in wave
// order through LDS.  Lanes past the system's last node skip their rows.
// UNR block columns per loop trip (1: fewer registers, more waves per SIMD;
// 3: more loads of each wave in flight, for systems whose slices each get a
// single wave).  At least 4 waves per SIMD: every form fits 128 VGPRs.
// UNR = 0: a pack whose systems all have union slots (no rank-slot path
// compiled in).
template <int R, int NT, int UNR>
__global__ void __launch_bounds__(kThreads, UNR == 0 && R <= 6 ? 5 : 4) k_pk_spmv(const PackHdr *__restrict__ hdr) {
    __shared__ float red[3 * R][kThreads];
    __shared__ double sm[16 * R];
    // wave index as a uniform value: the slice, its pointers and its offset
    // list then come from scalar loads
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int n_items = hdr->n_sitems;
    for (int item = blockIdx.x; item < n_items; item += gridDim.x) {
        const BHIP_GLOBAL PackSysDesc &D = to_global(hdr->sys)[to_global(hdr->sitem_sys)[item]];
        const BHIP_GLOBAL PackState *st = to_global(D.st);
        bool live = false;
#pragma unroll
        for (int r = 0; r < R; r++) live |= st->done[r] == 0;
        if (!live) continue;
        const int local = item - D.sitem0;
        const int split = D.split;
        const long long slice = split ? (long long)local : (long long)local * 4 + wave;
        const long long node = slice * 64 + lane;
        const long long ld = D.n;
        double dots[R];
        float a0[R], a1[R], a2[R];
#pragma unroll
        for (int r = 0; r < R; r++) {
            dots[r] = 0;
            a0[r] = a1[r] = a2[r] = 0.f;
        }
        if (UNR == 0 || D.uoff) {
            // union slots: all 64 lanes of the slice walk its runs
            if (slice < D.n_slices) {
                const int kb = to_global(D.sptr)[slice], W = (to_global(D.sptr)[slice + 1] - kb) >> 6;
                pk_rows_union<R, NT>(to_global(D.uoff) + (kb >> 6), to_global(D.sval), D.nsell, to_global(D.p), ld,
                                     node, (long long)D.n_nodes - 1, kb, W, split ? wave : 0, split ? 4 : 1, a0, a1,
                                     a2);
            }
        } else if (UNR > 0 && slice < D.n_slices && node < D.n_nodes) {
            const int kb = to_global(D.sptr)[slice], ke = to_global(D.sptr)[slice + 1];
            const int k0 = kb + lane + (split ? 64 * wave : 0), kstep = split ? 256 : 64;
            if (D.narrow)
                pk_rows<R, NT, UNR, short>(to_global(D.scol16), to_global(D.sval), D.nsell, to_global(D.p), ld, node,
                                           k0, ke, kstep, a0, a1, a2);
            else
                pk_rows<R, NT, UNR, int>(to_global(D.scol), to_global(D.sval), D.nsell, to_global(D.p), ld, node, k0,
                                         ke, kstep, a0, a1, a2);
        }
        if (split) {
            if (wave > 0)
#pragma unroll
                for (int r = 0; r < R; r++) {
                    red[3 * r][threadIdx.x] = a0[r];
                    red[3 * r + 1][threadIdx.x] = a1[r];
                    red[3 * r + 2][threadIdx.x] = a2[r];
                }
            __syncthreads();
            if (wave == 0)
                for (int w = 1; w < 4; w++)
#pragma unroll
                    for (int r = 0; r < R; r++) {
                        a0[r] += red[3 * r][64 * w + lane];
                        a1[r] += red[3 * r + 1][64 * w + lane];
                        a2[r] += red[3 * r + 2][64 * w + lane];
                    }
        }
        if (slice < D.n_slices && node < D.n_nodes && (!split || wave == 0)) {
            const BHIP_GLOBAL float *x = to_global(D.p);
            BHIP_GLOBAL float *y = to_global(D.q);
#pragma unroll
            for (int r = 0; r < R; r++) {
                const long long j = r * ld + 3 * node;
                y[j] = a0[r];
                y[j + 1] = a1[r];
                y[j + 2] = a2[r];
                dots[r] = (double)x[j] * a0[r] + (double)x[j + 1] * a1[r] + (double)x[j + 2] * a2[r];
            }
        }
        block_sum<R>(dots, sm);
        if (threadIdx.x == 0)
#pragma unroll
            for (int r = 0; r < R; r++) to_global(hdr->rec)[D.rec_s + (long long)r * D.n_sitems + local] = dots[r];
        __syncthreads();  // LDS is reused by the next item
    }
}

// Per (case, vector item): alpha from the system's p.q records; x += alpha p,
// r -= alpha q; records (r.z, r.r).
template <int R>
__global__ void __launch_bounds__(kThreads) k_pk_update(const PackHdr *__restrict__ hdr, int par) {
    __shared__ double sm[16 * 2];
    const int nv = hdr->n_vitems;
    for (int item = blockIdx.x; item < nv * R; item += gridDim.x) {
        const int r = item / nv, vi = item - r * nv;
        const BHIP_GLOBAL PackSysDesc &D = to_global(hdr->sys)[to_global(hdr->vitem_sys)[vi]];
        const BHIP_GLOBAL PackState *st = to_global(D.st);
        if (st->done[r]) continue;
        const double rz = st->rzb[par][r];
        double pap[1];
        sum_records_all<1>(to_global(hdr->rec) + D.rec_s + (long long)r * D.n_sitems, D.n_sitems, pap);
        const float alpha = (float)(rz / pap[0]);
        const int lv = vi - D.vitem0;
        long long i0, i1;
        item_range(D.n, lv, D.n_vitems, i0, i1);
        const long long o = (long long)r * D.n;
        BHIP_GLOBAL float *__restrict__ x = to_global(D.d);
        BHIP_GLOBAL float *__restrict__ rv = to_global(D.r);
        const BHIP_GLOBAL float *__restrict__ p = to_global(D.p);
        const BHIP_GLOBAL float *__restrict__ q = to_global(D.q);
        const BHIP_GLOBAL float *__restrict__ dinv = to_global(D.dinv);
        double acc[2] = {0, 0};
        // kBatch elements per trip, operands loaded before any update (same
        // elements and order of sums as a plain loop)
        const long long stride = blockDim.x;
        for (long long i = i0 + threadIdx.x; i < i1; i += kBatch * stride) {
            float xv[kBatch], pv[kBatch], rr[kBatch], qv[kBatch], dv[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                const long long j = i + u * stride;
                if (j < i1) {
                    xv[u] = x[o + j];
                    pv[u] = p[o + j];
                    rr[u] = rv[o + j];
                    qv[u] = q[o + j];
                    dv[u] = dinv[j];
                }
            }
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                const long long j = i + u * stride;
                if (j < i1) {
                    const float ri = rr[u] - alpha * qv[u];
                    x[o + j] = xv[u] + alpha * pv[u];
                    rv[o + j] = ri;
                    acc[0] += (double)ri * (ri * dv[u]);
                    acc[1] += (double)ri * ri;
                }
            }
        }
        block_sum<2>(acc, sm);
        if (threadIdx.x == 0) {
            BHIP_GLOBAL double *rec = to_global(hdr->rec) + D.rec_v + ((long long)r * D.n_vitems + lv) * 2;
            rec[0] = acc[0];
            rec[1] = acc[1];
        }
        __syncthreads();
    }
}

// Per (case, vector item): beta from the system's (r.z, r.r) records; the
// first item of the case applies the stopping test and writes the state (a
// later item of a finished case may skip its p update, which is then unused);
// p = z + beta p.
template <int R>
__global__ void __launch_bounds__(kThreads) k_pk_pupdate(const PackHdr *__restrict__ hdr, int par) {
    const int nv = hdr->n_vitems;
    for (int item = blockIdx.x; item < nv * R; item += gridDim.x) {
        const int r = item / nv, vi = item - r * nv;
        const BHIP_GLOBAL PackSysDesc &D = to_global(hdr->sys)[to_global(hdr->vitem_sys)[vi]];
        BHIP_GLOBAL PackState *st = to_global(D.st);
        if (st->done[r]) continue;
        const double rz = st->rzb[par][r];
        double t[2];
        sum_records_all<2>(to_global(hdr->rec) + D.rec_v + (long long)r * D.n_vitems * 2, D.n_vitems, t);
        const int lv = vi - D.vitem0;
        if (lv == 0 && threadIdx.x == 0) {
            const int it = st->iter[r] + 1;
            const bool conv = sqrt(t[1] / st->rr0[r]) <= D.tol;
            st->rzb[par ^ 1][r] = t[0];
            st->rr[r] = t[1];
            st->iter[r] = it;
            if (conv) st->converged[r] = 1;
            if (conv || it >= D.maxit) st->done[r] = 1;
        }
        const float beta = (float)(t[0] / rz);
        long long i0, i1;
        item_range(D.n, lv, D.n_vitems, i0, i1);
        const long long o = (long long)r * D.n;
        BHIP_GLOBAL float *__restrict__ p = to_global(D.p);
        const BHIP_GLOBAL float *__restrict__ rv = to_global(D.r);
        const BHIP_GLOBAL float *__restrict__ dinv = to_global(D.dinv);
        const long long stride = blockDim.x;
        for (long long i = i0 + threadIdx.x; i < i1; i += kBatch * stride) {
            float pv[kBatch], rr[kBatch], dv[kBatch];
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                const long long j = i + u * stride;
                if (j < i1) {
                    pv[u] = p[o + j];
                    rr[u] = rv[o + j];
                    dv[u] = dinv[j];
                }
            }
#pragma unroll
            for (int u = 0; u < kBatch; u++) {
                const long long j = i + u * stride;
                if (j < i1) p[o + j] = rr[u] * dv[u] + beta * pv[u];
            }
        }
    }
}

// Start of an inner solve: p = z = r dinv; records (r.z, r.r).
template <int R>
__global__ void __launch_bounds__(kThreads) k_pk_init(const PackHdr *__restrict__ hdr) {
    __shared__ double sm[16 * 2];
    const int nv = hdr->n_vitems;
    for (int item = blockIdx.x; item < nv * R; item += gridDim.x) {
        const int r = item / nv, vi = item - r * nv;
        const BHIP_GLOBAL PackSysDesc &D = to_global(hdr->sys)[to_global(hdr->vitem_sys)[vi]];
        const int lv = vi - D.vitem0;
        long long i0, i1;
        item_range(D.n, lv, D.n_vitems, i0, i1);
        const long long o = (long long)r * D.n;
        double acc[2] = {0, 0};
#pragma unroll 4
        for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
            const float ri = to_global(D.r)[o + i];
            const float zi = ri * to_global(D.dinv)[i];
            to_global(D.p)[o + i] = zi;
            acc[0] += (double)ri * zi;
            acc[1] += (double)ri * ri;
        }
        block_sum<2>(acc, sm);
        if (threadIdx.x == 0) {
            BHIP_GLOBAL double *rec = to_global(hdr->rec) + D.rec_v + ((long long)r * D.n_vitems + lv) * 2;
            rec[0] = acc[0];
            rec[1] = acc[1];
        }
        __syncthreads();
    }
}

// One block per (system, case): initial state from the init records.  A case
// with a zero inner right-hand side (already converged) is solved by d = 0.
template <int R>
__global__ void __launch_bounds__(kThreads) k_pk_init_state(const PackHdr *__restrict__ hdr) {
    for (int item = blockIdx.x; item < hdr->n_sys * R; item += gridDim.x) {
        const int s = item / R, r = item - s * R;
        const BHIP_GLOBAL PackSysDesc &D = to_global(hdr->sys)[s];
        double t[2];
        sum_records_all<2>(to_global(hdr->rec) + D.rec_v + (long long)r * D.n_vitems * 2, D.n_vitems, t);
        if (threadIdx.x == 0) {
            BHIP_GLOBAL PackState *st = to_global(D.st);
            st->rzb[0][r] = t[0];
            st->rzb[1][r] = 0;
            st->rr[r] = t[1];
            st->rr0[r] = t[1];
            st->iter[r] = 0;
            st->converged[r] = t[1] == 0.0;
            st->done[r] = t[1] == 0.0;
        }
    }
}

template <int R>
void enqueue_pk_iteration(int grid, hipStream_t s, const PackHdr *hdr, int par, bool nt, int unr) {
    if (unr == 0) {
        if (nt) k_pk_spmv<R, 1, 0><<<grid, kThreads, 0, s>>>(hdr);
        else k_pk_spmv<R, 0, 0><<<grid, kThreads, 0, s>>>(hdr);
    } else if (unr == 1) {
        if (nt) k_pk_spmv<R, 1, 1><<<grid, kThreads, 0, s>>>(hdr);
        else k_pk_spmv<R, 0, 1><<<grid, kThreads, 0, s>>>(hdr);
    } else if (unr == 3) {
        if (nt) k_pk_spmv<R, 1, 3><<<grid, kThreads, 0, s>>>(hdr);
        else k_pk_spmv<R, 0, 3><<<grid, kThreads, 0, s>>>(hdr);
    } else {
        if (nt) k_pk_spmv<R, 1, 2><<<grid, kThreads, 0, s>>>(hdr);
        else k_pk_spmv<R, 0, 2><<<grid, kThreads, 0, s>>>(hdr);
    }
    k_pk_update<R><<<grid, kThreads, 0, s>>>(hdr, par);
    k_pk_pupdate<R><<<grid, kThreads, 0, s>>>(hdr, par);
}

void enqueue_iteration(int R, int grid, hipStream_t s, const PackHdr *hdr, int par, bool nt, int unr) {
    switch (R) {
        case 1: enqueue_pk_iteration<1>(grid, s, hdr, par, nt, unr); break;
        case 2: enqueue_pk_iteration<2>(grid, s, hdr, par, nt, unr); break;
        case 3: enqueue_pk_iteration<3>(grid, s, hdr, par, nt, unr); break;
        case 4: enqueue_pk_iteration<4>(grid, s, hdr, par, nt, unr); break;
        case 5: enqueue_pk_iteration<5>(grid, s, hdr, par, nt, unr); break;
        case 6: enqueue_pk_iteration<6>(grid, s, hdr, par, nt, unr); break;
        case 7: enqueue_pk_iteration<7>(grid, s, hdr, par, nt, unr); break;
        default: enqueue_pk_iteration<8>(grid, s, hdr, par, nt, unr); break;
    }
}

void enqueue_init(int R, int grid, hipStream_t s, const PackHdr *hdr) {
#define BHIP_PK_INIT(RR)                                        \
    case RR:                                                    \
        k_pk_init<RR><<<grid, kThreads, 0, s>>>(hdr);           \
        k_pk_init_state<RR><<<grid, kThreads, 0, s>>>(hdr);     \
        break;
    switch (R) {
        BHIP_PK_INIT(1)
        BHIP_PK_INIT(2)
        BHIP_PK_INIT(3)
        BHIP_PK_INIT(4)
        BHIP_PK_INIT(5)
        BHIP_PK_INIT(6)
        BHIP_PK_INIT(7)
        default:
            k_pk_init<8><<<grid, kThreads, 0, s>>>(hdr);
            k_pk_init_state<8><<<grid, kThreads, 0, s>>>(hdr);
    }
#undef BHIP_PK_INIT
}

bool graphs_enabled() {
    const char *e = getenv("BHIP_NO_GRAPH");
    return !(e && e[0] == '1');
}

// Per-system item counts (they depend on the system alone).
int sys_split(const bhip_ctx *c) { return c->n_slices < kSplitSlices ? 1 : 0; }
int sys_sitems(const bhip_ctx *c) { return sys_split(c) ? c->n_slices : (c->n_slices + 3) / 4; }
// Vector items per case: about 2048 DOFs each, at most grid / R so that a
// system's R x items fit the persistent grid in one round (grid is fixed per
// device, so the count still depends on the system and R alone).
int sys_vitems(const bhip_ctx *c, int R, int grid) {
    const long long want = ((long long)c->n_dof + kThreads * 8 - 1) / (kThreads * 8);
    const long long cap = std::max(1, grid / R);
    return (int)std::max<long long>(1, std::min<long long>(want, cap));
}

}  // namespace

struct PackEngine {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    int grid = 2048;
    // descriptor block on the device: header | system descriptors | item tables
    dvec<unsigned char> dbuf;
    unsigned char *hbuf = nullptr;  // pinned staging of the descriptor block
    size_t hcap = 0;
    dvec<double> rec;               // records of the pack
    dvec<PackState> st;             // one per system
    PackState *h_st = nullptr;
    size_t h_st_cap = 0;
    dvec<double> resid;             // 2R per system: |b - K x|^2, |b|^2
    double *h_resid = nullptr;
    size_t h_resid_cap = 0;
    double *h_scale = nullptr;      // per system: R scales then R norms
    size_t h_scale_cap = 0;
    hipGraphExec_t graph[8 * (kMaxR + 1)] = {};  // by (R, streaming, unroll 0-3)
};

namespace {

void pk_graphs_reset(PackEngine *pe) {
    for (auto &g : pe->graph)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// Report a HIP failure on every context of the pack.
int pk_fail(bhip_ctx *const *ctxs, int n, hipError_t e, const char *what) {
    const int code = e == hipErrorOutOfMemory ? BHIP_ERR_OOM : BHIP_ERR_HIP;
    for (int i = 0; i < n; i++) bhip_fail(ctxs[i], code, std::string(what) + ": " + hipGetErrorString(e));
    return code;
}

#define PK_TRY(call)                                             \
    do {                                                         \
        hipError_t e_ = (call);                                  \
        if (e_ != hipSuccess) return pk_fail(ctxs, n_all, e_, #call); \
    } while (0)

template <class T>
hipError_t pinned_reserve(T *&p, size_t &cap, size_t count) {
    if (count <= cap && p) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    hipError_t e = hipHostMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T), hipHostMallocDefault);
    if (e == hipSuccess) cap = count;
    return e;
}

}  // namespace

int bhip_pack_engine_create(int device, hipStream_t stream, PackEngine **out) {
    *out = nullptr;
    PackEngine *pe = new PackEngine();
    pe->device = device;
    if (stream) {
        pe->stream = stream;
    } else {
        if (hipStreamCreateWithFlags(&pe->stream, hipStreamNonBlocking) != hipSuccess) {
            delete pe;
            return BHIP_ERR_HIP;
        }
        pe->own_stream = true;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) pe->grid = std::max(64, prop.multiProcessorCount * 8);
    *out = pe;
    return BHIP_OK;
}

void bhip_pack_engine_destroy(PackEngine *pe) {
    if (!pe) return;
    if (pe->stream) (void)hipStreamSynchronize(pe->stream);
    pk_graphs_reset(pe);
    pe->dbuf.release();
    pe->rec.release();
    pe->st.release();
    pe->resid.release();
    if (pe->hbuf) (void)hipHostFree(pe->hbuf);
    if (pe->h_st) (void)hipHostFree(pe->h_st);
    if (pe->h_resid) (void)hipHostFree(pe->h_resid);
    if (pe->h_scale) (void)hipHostFree(pe->h_scale);
    if (pe->own_stream) (void)hipStreamDestroy(pe->stream);
    delete pe;
}

hipStream_t bhip_pack_engine_stream(PackEngine *pe) { return pe->stream; }

// Solve the R load cases of n_all assembled contexts (FP32 SELL 3x3 form ready,
// right-hand sides in multi->b64, x in multi->x64 zeroed) to opts.tol with
// FP32-inner Jacobi-PCG and FP64 refinement, every system to its own
// tolerance.  stats[i] and status[i] per context.  The contexts' streams must
// be idle; all work runs on the engine's stream.
int bhip_pack_run(PackEngine *pe, bhip_ctx *const *ctxs, int n_all, int R, const bhip_solve_opts &o,
                  bhip_stats *stats, int *status) {
    if (n_all <= 0) return BHIP_OK;
    hipStream_t s = pe->stream;
    const double inner_tol = o.inner_tol > 0 ? o.inner_tol : 1e-3;
    const int max_outer = o.max_outer > 0 ? o.max_outer : 50;
    const int maxit = o.maxit > 0 ? o.maxit : 10000;
    // streaming loads when the pack's SELL arrays exceed the last-level cache
    // (measured crossover for six cases: RVE 48^3, 121 MB, runs faster with
    // plain loads; 64^3, 277 MB, with streaming loads)
    // BHIP_PACK_STREAMING=0/1 forces the choice (the autotune sweep compares both)
    double by = 0;
    for (int i = 0; i < n_all; i++) by += (double)ctxs[i]->n_sell * (ctxs[i]->sell_narrow ? 38.0 : 40.0);
    bool nt = by > kStreamBytes;
    const char *force = getenv("BHIP_PACK_STREAMING");
    if (force && (force[0] == '0' || force[0] == '1')) nt = force[0] == '1';
    // block columns per SpMV loop trip: three for streaming packs (a large
    // system's waves each own a whole slice, so a wave's own loads in flight
    // set the rate), two otherwise (BHIP_PACK_UNROLL=1/2/3 forces it)
    int unr = nt ? kPackUnrollStream : kPackUnroll;
    const char *fu = getenv("BHIP_PACK_UNROLL");
    if (fu && (fu[0] == '1' || fu[0] == '2' || fu[0] == '3')) unr = fu[0] - '0';
    std::vector<int> total(n_all, 0), outer(n_all, 0), active(n_all, 1), ok(n_all, 0);
    std::vector<long long> case_total(n_all, 0);
    std::vector<double> worst(n_all, 1.0);  // worst unconverged relative residual per system
    for (int i = 0; i < n_all; i++) {
        memset(&stats[i], 0, sizeof(bhip_stats));
        status[i] = BHIP_OK;
    }
    PK_TRY(pe->resid.resize(std::max<size_t>(pe->resid.n, (size_t)n_all * 2 * R)));
    PK_TRY(pinned_reserve(pe->h_resid, pe->h_resid_cap, (size_t)n_all * 2 * R));
    PK_TRY(pinned_reserve(pe->h_scale, pe->h_scale_cap, (size_t)n_all * 2 * kMaxR));
    for (int round = 0;; round++) {
        // FP64 residuals of the active systems
        for (int i = 0; i < n_all; i++)
            if (active[i]) bhip_multi_enqueue_resid(ctxs[i], R, s, pe->resid.p + (size_t)i * 2 * R);
        PK_TRY(hipGetLastError());
        PK_TRY(hipMemcpyAsync(pe->h_resid, pe->resid.p, sizeof(double) * n_all * 2 * R, hipMemcpyDeviceToHost, s));
        PK_TRY(hipStreamSynchronize(s));
        std::vector<int> run;
        for (int i = 0; i < n_all; i++) {
            if (!active[i]) continue;
            stats[i].spmv_launches += 1;
            const double *rb = pe->h_resid + (size_t)i * 2 * R;
            double *sc = pe->h_scale + (size_t)i * 2 * kMaxR;
            bool all = true;
            worst[i] = 0;
            for (int r = 0; r < R; r++) {
                const double rr = rb[2 * r], bb = rb[2 * r + 1];
                const double rel = bb > 0 ? sqrt(rr / bb) : sqrt(rr);
                const bool conv = rel <= o.tol || rr == 0.0;
                if (!conv) worst[i] = std::max(worst[i], rel);
                const double rn = conv ? 0.0 : sqrt(rr);
                sc[r] = conv ? 0.0 : 1.0 / rn;  // a converged case gets a zero inner RHS
                sc[kMaxR + r] = rn;
                all &= conv;
            }
            if (all) {
                ok[i] = 1;
                active[i] = 0;
                continue;
            }
            if (outer[i] == max_outer || total[i] >= maxit) {
                active[i] = 0;
                continue;
            }
            run.push_back(i);
        }
        if (run.empty()) break;
        // inner right-hand sides r32 = r64 / |r64|, d = 0
        for (int i : run) {
            MultiData &M = *ctxs[i]->multi;
            const double *sc = pe->h_scale + (size_t)i * 2 * kMaxR;
            PK_TRY(hipMemcpyAsync(M.scale.p, sc, sizeof(double) * R, hipMemcpyHostToDevice, s));
            PK_TRY(hipMemcpyAsync(M.rn.p, sc + kMaxR, sizeof(double) * R, hipMemcpyHostToDevice, s));
            bhip_multi_enqueue_pack32(ctxs[i], R, s);
        }
        // descriptor of the systems in this inner solve
        const int ns = (int)run.size();
        // every system with union slots: the union-only SpMV (unroll 0)
        bool all_union = true;
        for (int i : run) all_union &= ctxs[i]->sell_union;
        const int run_unr = all_union ? 0 : unr;
        int n_s = 0, n_v = 0;
        long long n_rec = 0;
        std::vector<PackSysDesc> desc(ns);
        for (int k = 0; k < ns; k++) {
            bhip_ctx *c = ctxs[run[k]];
            MultiData &M = *c->multi;
            PackSysDesc &D = desc[k];
            D.sptr = c->sell_ptr.p;
            D.scol = c->sell_col.p;
            D.scol16 = c->sell_col16.p;
            D.uoff = c->sell_union ? c->sell_uoff.p : nullptr;
            D.narrow = c->sell_narrow ? 1 : 0;
            D.pad0 = 0;
            D.sval = c->bval32.p;
            D.nsell = c->n_sell;
            D.dinv = c->dinv32.p;
            D.r = M.r32.p;
            D.p = M.p32.p;
            D.q = M.q32.p;
            D.d = M.d32.p;
            D.n = c->n_dof;
            D.n_nodes = c->n_nodes;
            D.n_slices = c->n_slices;
            D.split = sys_split(c);
            D.n_sitems = sys_sitems(c);
            D.sitem0 = n_s;
            D.n_vitems = sys_vitems(c, R, pe->grid);
            D.vitem0 = n_v;
            D.maxit = maxit - total[run[k]];
            D.tol = ir_inner_tol(inner_tol, o.tol, worst[run[k]]);
            D.rec_s = n_rec;
            n_rec += (long long)R * D.n_sitems;
            D.rec_v = n_rec;
            n_rec += (long long)R * D.n_vitems * 2;
            n_s += D.n_sitems;
            n_v += D.n_vitems;
        }
        if (pe->rec.n < (size_t)n_rec) PK_TRY(pe->rec.resize((size_t)n_rec));
        if (pe->st.n < (size_t)ns) PK_TRY(pe->st.resize((size_t)std::max(ns, 16)));
        PK_TRY(pinned_reserve(pe->h_st, pe->h_st_cap, (size_t)ns));
        for (int k = 0; k < ns; k++) desc[k].st = pe->st.p + k;
        const size_t off_sys = (sizeof(PackHdr) + 255) / 256 * 256;
        const size_t off_s = off_sys + ((sizeof(PackSysDesc) * ns + 255) / 256 * 256);
        const size_t off_v = off_s + ((sizeof(int) * n_s + 255) / 256 * 256);
        const size_t bytes = off_v + sizeof(int) * (size_t)n_v;
        if (pe->dbuf.n < bytes) {
            // the graphs hold the descriptor's address
            pk_graphs_reset(pe);
            PK_TRY(pe->dbuf.resize(std::max(bytes, pe->dbuf.n * 2)));
        }
        PK_TRY(pinned_reserve(pe->hbuf, pe->hcap, bytes));
        unsigned char *dev = pe->dbuf.p;
        PackHdr hh;
        hh.n_sys = ns;
        hh.R = R;
        hh.n_sitems = n_s;
        hh.n_vitems = n_v;
        hh.sys = reinterpret_cast<const PackSysDesc *>(dev + off_sys);
        hh.sitem_sys = reinterpret_cast<const int *>(dev + off_s);
        hh.vitem_sys = reinterpret_cast<const int *>(dev + off_v);
        hh.rec = pe->rec.p;
        memcpy(pe->hbuf, &hh, sizeof(hh));
        memcpy(pe->hbuf + off_sys, desc.data(), sizeof(PackSysDesc) * ns);
        int *hs = reinterpret_cast<int *>(pe->hbuf + off_s);
        int *hv = reinterpret_cast<int *>(pe->hbuf + off_v);
        for (int k = 0; k < ns; k++) {
            for (int j = 0; j < desc[k].n_sitems; j++) hs[desc[k].sitem0 + j] = k;
            for (int j = 0; j < desc[k].n_vitems; j++) hv[desc[k].vitem0 + j] = k;
        }
        PK_TRY(hipMemcpyAsync(dev, pe->hbuf, bytes, hipMemcpyHostToDevice, s));
        const PackHdr *hdr = reinterpret_cast<const PackHdr *>(dev);
        enqueue_init(R, pe->grid, s, hdr);
        PK_TRY(hipGetLastError());
        // chunks of kChunk iterations until every case of every system is done
        hipGraphExec_t g = nullptr;
        if (graphs_enabled()) {
            hipGraphExec_t &slot = pe->graph[8 * R + (nt ? 4 : 0) + run_unr];
            if (!slot) {
                hipGraph_t graph;
                PK_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
                for (int k = 0; k < kChunk; k++) enqueue_iteration(R, pe->grid, s, hdr, k & 1, nt, run_unr);
                hipError_t e = hipStreamEndCapture(s, &graph);
                PK_TRY(e);
                e = hipGraphInstantiate(&slot, graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                if (e != hipSuccess) slot = nullptr;
                PK_TRY(e);
            }
            g = slot;
        }
        int max_left = 0;
        for (int k = 0; k < ns; k++) max_left = std::max(max_left, desc[k].maxit);
        const long long max_chunks = (long long)max_left / kChunk + 2;
        for (long long c = 0;; c++) {
            PK_TRY(hipMemcpyAsync(pe->h_st, pe->st.p, sizeof(PackState) * ns, hipMemcpyDeviceToHost, s));
            PK_TRY(hipStreamSynchronize(s));
            bool all = true;
            for (int k = 0; k < ns && all; k++)
                for (int r = 0; r < R; r++) all &= pe->h_st[k].done[r] != 0;
            if (all) break;
            if (c > max_chunks) {
                for (int i = 0; i < n_all; i++)
                    bhip_fail(ctxs[i], BHIP_ERR_STATE, "packed PCG did not finish (internal error)");
                return BHIP_ERR_STATE;
            }
            if (g) {
                PK_TRY(hipGraphLaunch(g, s));
            } else {
                for (int k = 0; k < kChunk; k++) enqueue_iteration(R, pe->grid, s, hdr, k & 1, nt, run_unr);
            }
            PK_TRY(hipGetLastError());
            for (int k = 0; k < ns; k++) stats[run[k]].spmv_launches += kChunk;
        }
        // x += |r| d per system; iterations of this inner solve = longest case
        for (int k = 0; k < ns; k++) {
            const int i = run[k];
            int it = 0;
            for (int r = 0; r < R; r++) {
                it = std::max(it, pe->h_st[k].iter[r]);
                case_total[i] += pe->h_st[k].iter[r];
            }
            total[i] += it;
            outer[i] += 1;
            bhip_multi_enqueue_axpy(ctxs[i], R, s);
        }
        PK_TRY(hipGetLastError());
    }
    PK_TRY(hipStreamSynchronize(s));
    for (int i = 0; i < n_all; i++) {
        stats[i].iters = total[i];
        stats[i].case_iters = case_total[i];
        stats[i].outer_iters = outer[i];
        status[i] = ok[i] ? BHIP_OK : BHIP_ERR_NOT_CONVERGED;
    }
    return BHIP_OK;
}

// ------------------------------------------------------------------ C ABI
struct bhip_pack {
    int device = 0;
    PackEngine *pe = nullptr;
};

extern "C" {

int bhip_pack_create(int device, bhip_pack **out) {
    if (!out) return BHIP_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return BHIP_ERR_HIP;
    (void)hipSetDevice(device);
    bhip_pack *pk = new bhip_pack();
    pk->device = device;
    if (bhip_pack_engine_create(device, nullptr, &pk->pe) != BHIP_OK) {
        delete pk;
        return BHIP_ERR_HIP;
    }
    *out = pk;
    return BHIP_OK;
}

void bhip_pack_destroy(bhip_pack *pk) {
    if (!pk) return;
    bhip_pack_engine_destroy(pk->pe);
    delete pk;
}

int bhip_pack_solve_cases(bhip_pack *pk, int n, bhip_ctx *const *ctxs, int nrhs, const double *const *presc,
                          const double *const *load, double *const *x, const bhip_solve_opts *opts,
                          bhip_stats *stats, int *status) {
    if (!pk || n < 1 || !ctxs || !presc || !x || !opts || !status) return BHIP_ERR_ARG;
    for (int i = 0; i < n; i++) {
        if (!ctxs[i] || !presc[i] || !x[i]) return BHIP_ERR_ARG;
        if (ctxs[i]->device != pk->device) return bhip_fail(ctxs[i], BHIP_ERR_ARG, "context on another device");
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return bhip_fail(ctxs[i], BHIP_ERR_ARG, "a context appears twice in the pack");
    }
    (void)hipSetDevice(pk->device);
    std::vector<bhip_stats> S(n);
    std::vector<bhip_ctx *> run;
    std::vector<int> idx;
    for (int i = 0; i < n; i++) {
        memset(&S[i], 0, sizeof(bhip_stats));
        bool fused = false;
        status[i] = bhip_cases_begin(ctxs[i], nrhs, presc[i], load ? load[i] : nullptr, opts, &fused);
        if (status[i] == BHIP_OK && !fused)
            status[i] = bhip_fail(ctxs[i], BHIP_ERR_UNSUPPORTED,
                                  "packed solves need FP32_IR and 3 DOFs on every node (use bhip_solve_cases)");
        if (status[i] != BHIP_OK) continue;
        run.push_back(ctxs[i]);
        idx.push_back(i);
    }
    if (!run.empty()) {
        std::vector<bhip_stats> Sr(run.size());
        std::vector<int> st(run.size(), BHIP_OK);
        hipEvent_t e0, e1;
        if (hipEventCreate(&e0) != hipSuccess) return BHIP_ERR_HIP;
        if (hipEventCreate(&e1) != hipSuccess) {
            (void)hipEventDestroy(e0);
            return BHIP_ERR_HIP;
        }
        hipStream_t s = bhip_pack_engine_stream(pk->pe);
        (void)hipEventRecord(e0, s);
        const int e = bhip_pack_run(pk->pe, run.data(), (int)run.size(), nrhs, *opts, Sr.data(), st.data());
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        for (size_t j = 0; j < run.size(); j++) {
            const int i = idx[j];
            S[i] = Sr[j];
            S[i].ms_total = ms;
            int sj = e != BHIP_OK ? e : BHIP_OK;
            if (sj == BHIP_OK) sj = bhip_cases_finish(run[j], nrhs, x[i], &S[i]);
            if (sj == BHIP_OK && st[j] != BHIP_OK)
                sj = bhip_fail(run[j], st[j], "multi-case solve not converged within the iteration caps");
            status[i] = sj;
        }
    }
    int first = BHIP_OK;
    for (int i = 0; i < n; i++) {
        if (stats) stats[i] = S[i];
        if (first == BHIP_OK && status[i] != BHIP_OK) first = status[i];
    }
    return first;
}

}  // extern "C"
