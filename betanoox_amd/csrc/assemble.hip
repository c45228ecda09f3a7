// assemble.hip -- element integration fused with the global scatter.
//
// Reference path: SerialAssembler::assembleK (SerialAssembler.cpp:35-59) calls
// update("K") per element (ElasticityElement3D.cpp:983-1051: Jacobian,
// B matrix, upper-triangle Ke += C_x_B . b . detJ.w at every Gauss point,
// 560-641) and then Equations::addMatrix (equation.cpp:445-562), which adds
// Ke(I,J) into the lower-triangle store for free I >= J, sets K(I,I) = 1 for
// a restrained I, and lifts F[I] -= Ke(I,J) u_J for free I, restrained J.
//
// MI355X design, two stages, no colouring and no atomics:
//  1. element stage (k_elem_blocks): a workgroup takes 16 elements, stages
//     their coordinates, materials and Gauss-point gradients (and J2
//     tangents) in LDS and integrates each element once into its 36
//     canonical 3x3 blocks (one per unordered pair of local nodes, rows of the
//     lower global node), written to HBM pair-major (one local pair of
//     consecutive elements contiguous);
//  2. gather stage: for every (node n, neighbour m) entry of the node
//     adjacency, adds the blocks of the elements holding both nodes in
//     ascending element order (the reference's serial summation order,
//     element matrices first) and writes the block once into the full
//     symmetric CSR (k_asm_pairs, one thread per pair in pair order), and
//     also into the solver's SELL blocks when that layout is current
//     (k_asm_rows, lanes on consecutive nodes, one neighbour slot per wave).
//     The pair (m,n) reads the same canonical blocks and writes the
//     transpose, so the stored matrix is exactly symmetric.
// The lift and element-force vectors (k_gather_rhs) and the lift table read
// the same element blocks.
#include <algorithm>

#include "bhip_internal.hpp"
#include "gauss_tables.hpp"

namespace {

struct AsmArgs {
    const double *xyz;
    const int *conn;
    const double *mat;
    const int *elem_group;
    int n_el, n_groups, mat_stride, nnpe, nip;
    double factor;
    // plasticity: per (element, ip) symmetric 6x6 tangent, upper triangle row by row
    const double *tan;
    int *flags;
    // Gauss tables (ctx->gauss): weights, derivatives [ip][xi|eta|zeta][8]
    // (quad4: [ip][xi|eta][4]), shape values [ip][node]
    const double *gw, *gdN, *gN;
    // element-stage blocks, pair-major: local pair pr of element e at
    // (pr * n_el + e) * NBS, so that the blocks of one local pair of
    // consecutive elements are contiguous
    const double *eblk;
    // local index of each (node, incident element) entry of the incidence lists
    const unsigned char *n2e_loc;
};

// Point the element policies at the Gauss tables passed as a restrict
// kernel argument: read-only and uniform, so the compiler reads them with
// scalar loads.
__device__ __forceinline__ void bind_tables(AsmArgs &A, const double *__restrict__ gtab) {
    A.gw = gtab;
    A.gdN = gtab + kGaussDN;
    A.gN = gtab + kGaussN;
}

__device__ __forceinline__ int group_of(const AsmArgs &A, int e) {
    if (A.elem_group) return A.elem_group[e];
    return A.n_groups == A.n_el ? e : 0;
}

// -------------------------------------------------------------- policies
// The assembly runs in two stages.  The element stage integrates every
// element once: a workgroup stages its elements' nodal coordinates and
// material in LDS, then the global shape-function gradients and weights of
// every Gauss point (and, for J2, the 6x6 tangent of the point), then
// computes each element's canonical blocks -- one per unordered pair of local
// nodes, oriented with the rows of the lower global node index -- as the sum
// over Gauss points of B_a^T C B_b detJ w (the reference's per-element Ke,
// ElasticityElement3D.cpp:560-641), and writes them to HBM.  The gather stage
// (one thread per node pair) then adds the blocks of the elements holding
// both nodes in ascending element order, which is the order in which the
// reference's serial assembler adds its element matrices (SerialAssembler.cpp
// :35-59, Equations::addMatrix).
//
// A policy describes one element family: ND DOFs per node, NN nodes, DIM
// space dimensions, NB block entries, NBS stored doubles per pair (NB, plus
// the capacity entry for heat), G doubles of staged Gauss-point data, MS
// doubles of staged material.

// packed index of local pair (i, j), i <= j, row by row
__device__ __forceinline__ int pair_index(int nn, int i, int j) { return i * nn - i * (i - 1) / 2 + (j - i); }
__device__ __forceinline__ void pair_of_index(int nn, int pr, int &i, int &j) {
    i = 0;
    int rem = pr;
    while (rem >= nn - i) {
        rem -= nn - i;
        i++;
    }
    j = i + rem;
}

// Index of entry (s, u) of a symmetric 6x6 stored as its upper triangle row
// by row (21 values).
__host__ __device__ constexpr int sym_index(int s, int u) {
    return s <= u ? s * 6 - s * (s - 1) / 2 + (u - s) : u * 6 - u * (u - 1) / 2 + (s - u);
}

// Block B_a^T C B_b of a general 6x6 C (Voigt xx yy zz xy yz xz, engineering
// shear): CB = C B_b (6x3), then B_a^T CB.
template <class CF>
__device__ __forceinline__ void ctb_block(CF C, const double *ga, const double *gb, double *blk) {
    double CB[6][3];
#pragma unroll
    for (int s = 0; s < 6; s++) {
        CB[s][0] = gb[0] * C(s, 0) + gb[1] * C(s, 3) + gb[2] * C(s, 5);
        CB[s][1] = gb[1] * C(s, 1) + gb[0] * C(s, 3) + gb[2] * C(s, 4);
        CB[s][2] = gb[2] * C(s, 2) + gb[1] * C(s, 4) + gb[0] * C(s, 5);
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        blk[0 * 3 + d] = ga[0] * CB[0][d] + ga[1] * CB[3][d] + ga[2] * CB[5][d];
        blk[1 * 3 + d] = ga[1] * CB[1][d] + ga[0] * CB[3][d] + ga[2] * CB[4][d];
        blk[2 * 3 + d] = ga[2] * CB[2][d] + ga[1] * CB[4][d] + ga[0] * CB[5][d];
    }
}

// 3D elasticity, isotropic (E, nu) or full 6x6 C; J2 takes each Gauss point's
// own symmetric tangent from the last return map.  Block of B_a^T C B_b with Voigt order
// xx yy zz xy yz xz and engineering shear (ElasticityElement3D::
// calculateBmatrix, ElasticityElement3D.cpp:57-87).
template <int KIND>
struct ElasHex8 {
    static constexpr int ND = 3, NN = 8, DIM = 3, NB = 9, NBS = 9, NP = 36;
    static constexpr bool J2 = KIND == BHIP_MAT_J2;
    static constexpr int G = NN * DIM + 1 + (J2 ? kTanVals : 0);
    static constexpr int MS = KIND == BHIP_MAT_CMAT3D ? 36 : 2;
    static __device__ __forceinline__ void stage_mat(const AsmArgs &A, int e, double *m) {
        const double *mp = A.mat + (long long)group_of(A, e) * A.mat_stride;
        if (KIND == BHIP_MAT_CMAT3D) {
#pragma unroll
            for (int i = 0; i < 36; i++) m[i] = mp[i];
        } else {
            const double E = mp[0], nu = mp[1];
            m[0] = E * nu / ((1 + nu) * (1 - 2 * nu));  // lambda
            m[1] = E / (2 * (1 + nu));                  // mu
        }
    }
    // IsoElement::jacobian, 3-direction branch (isoElement.cpp:324-372):
    // gradients of the 8 nodes and f = detJ w factor at Gauss point ip.
    static __device__ __forceinline__ bool stage_ip(const AsmArgs &A, int e, int ip, const double *xe, double *g) {
        const double *x = xe, *y = xe + 8, *z = xe + 16;
        const double *dxi = A.gdN + ip * 24, *deta = dxi + 8, *dzeta = dxi + 16;
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
            g[3 * i] = a11 * dxi[i] + a12 * deta[i] + a13 * dzeta[i];
            g[3 * i + 1] = a21 * dxi[i] + a22 * deta[i] + a23 * dzeta[i];
            g[3 * i + 2] = a31 * dxi[i] + a32 * deta[i] + a33 * dzeta[i];
        }
        g[24] = detJ * A.gw[ip] * A.factor;
        if (J2) {
            const double *t = A.tan + ((long long)e * A.nip + ip) * kTanVals;
#pragma unroll
            for (int i = 0; i < kTanVals; i++) g[25 + i] = t[i];
        }
        return detJ < 0;
    }
    // K += (B_a^T C B_b)(ip) f
    static __device__ __forceinline__ void add_ip(const double *m, const double *g, int a, int b, double *K,
                                                  double *) {
        const double *ga = g + 3 * a, *gb = g + 3 * b;
        const double f = g[24];
        double blk[9];
        if (KIND == BHIP_MAT_ISO_ELASTIC) {
            const double lam = m[0], mu = m[1];
            double gg = mu * (ga[0] * gb[0] + ga[1] * gb[1] + ga[2] * gb[2]);
#pragma unroll
            for (int c = 0; c < 3; c++)
#pragma unroll
                for (int d = 0; d < 3; d++) blk[c * 3 + d] = lam * ga[c] * gb[d] + mu * gb[c] * ga[d] + (c == d ? gg : 0.0);
        } else if (J2) {
            // the Gauss point's own symmetric tangent (upper triangle, row by row)
            const double *t = g + 25;
            ctb_block([t](int s, int u) { return t[sym_index(s, u)]; }, ga, gb, blk);
        } else {
            const double *C = m;
            ctb_block([C](int s, int u) { return C[s * 6 + u]; }, ga, gb, blk);
        }
#pragma unroll
        for (int i = 0; i < 9; i++) K[i] += blk[i] * f;
    }
    // BasicElement::checkForNegativeDiagonal (BasicElement.cpp:116-130); the
    // reference checks the elastic stiffness only
    static __device__ __forceinline__ bool bad_diag(const double *K) {
        return !J2 && (K[0] <= 0 || K[4] <= 0 || K[8] <= 0);
    }
    static __device__ __forceinline__ void load(const double *s, double *K, double *) {
#pragma unroll
        for (int i = 0; i < 9; i++) K[i] += s[i];
    }
};

// 2D plane stress quad4 (ElasticityElement2D.cpp:22-119, pStress.cpp:68-85).
struct ElasQuad4 {
    static constexpr int ND = 2, NN = 4, DIM = 2, NB = 4, NBS = 4, NP = 10;
    static constexpr int G = NN * DIM + 1, MS = 9;
    static __device__ __forceinline__ void stage_mat(const AsmArgs &A, int e, double *m) {
        const double *mp = A.mat + (long long)group_of(A, e) * A.mat_stride;
        double s11 = 1 / mp[0], s12 = -mp[2] / mp[0], s22 = 1. / mp[1], s66 = 1. / mp[3];
        double den = s11 * s22 - s12 * s12;
        m[0] = s22 / den;
        m[1] = -s12 / den;
        m[2] = 0.;
        m[3] = -s12 / den;
        m[4] = s11 / den;
        m[5] = 0.;
        m[6] = 0.;
        m[7] = 0.;
        m[8] = 1. / s66;
    }
    // 2-direction branch of IsoElement::jacobian (isoElement.cpp:290-322).
    static __device__ __forceinline__ bool stage_ip(const AsmArgs &A, int, int ip, const double *xe, double *g) {
        const double *x = xe, *y = xe + 4;
        const double *dxi = A.gdN + ip * 24, *deta = dxi + 4;
        double J00 = 0, J01 = 0, J10 = 0, J11 = 0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            J00 += dxi[i] * x[i];
            J01 += dxi[i] * y[i];
            J10 += deta[i] * x[i];
            J11 += deta[i] * y[i];
        }
        double detJ = J00 * J11 - J01 * J10;
        double t = 1 / detJ;
        double i00 = J11 * t, i01 = -J01 * t, i10 = -J10 * t, i11 = J00 * t;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            g[2 * i] = i00 * dxi[i] + i01 * deta[i];
            g[2 * i + 1] = i10 * dxi[i] + i11 * deta[i];
        }
        g[8] = detJ * A.gw[ip] * A.factor;
        return detJ < 0;
    }
    static __device__ __forceinline__ void add_ip(const double *C, const double *g, int a, int b, double *K, double *) {
        const double *ga = g + 2 * a, *gb = g + 2 * b;
        const double f = g[8];
        double CB[3][2];
#pragma unroll
        for (int s = 0; s < 3; s++) {
            CB[s][0] = gb[0] * C[s * 3 + 0] + gb[1] * C[s * 3 + 2];
            CB[s][1] = gb[1] * C[s * 3 + 1] + gb[0] * C[s * 3 + 2];
        }
        double blk[4];
#pragma unroll
        for (int d = 0; d < 2; d++) {
            blk[0 * 2 + d] = ga[0] * CB[0][d] + ga[1] * CB[2][d];
            blk[1 * 2 + d] = ga[1] * CB[1][d] + ga[0] * CB[2][d];
        }
#pragma unroll
        for (int i = 0; i < 4; i++) K[i] += blk[i] * f;
    }
    static __device__ __forceinline__ bool bad_diag(const double *K) { return K[0] <= 0 || K[3] <= 0; }
    static __device__ __forceinline__ void load(const double *s, double *K, double *) {
#pragma unroll
        for (int i = 0; i < 4; i++) K[i] += s[i];
    }
};

// Transient heat hex8: K (HeatTransferElement3DTransient.cpp:153-197) and
// the consistent capacity M (60-77); stored as (K, M) per pair.
struct HeatHex8 {
    static constexpr int ND = 1, NN = 8, DIM = 3, NB = 1, NBS = 2, NP = 36;
    static constexpr int G = NN * DIM + 1 + NN, MS = 4;  // gradients, f, shape values
    static __device__ __forceinline__ void stage_mat(const AsmArgs &A, int e, double *m) {
        const double *mp = A.mat + (long long)group_of(A, e) * A.mat_stride;
#pragma unroll
        for (int i = 0; i < 4; i++) m[i] = mp[i];  // k11 k22 k33 capacity
    }
    static __device__ __forceinline__ bool stage_ip(const AsmArgs &A, int e, int ip, const double *xe, double *g) {
        const bool bad = ElasHex8<BHIP_MAT_ISO_ELASTIC>::stage_ip(A, e, ip, xe, g);
#pragma unroll
        for (int i = 0; i < 8; i++) g[25 + i] = A.gN[ip * 8 + i];
        return bad;
    }
    static __device__ __forceinline__ void add_ip(const double *m, const double *g, int a, int b, double *K,
                                                  double *K2) {
        const double *ga = g + 3 * a, *gb = g + 3 * b;
        const double f = g[24];
        K[0] += (gb[0] * m[0] * ga[0] + gb[1] * m[1] * ga[1] + gb[2] * m[2] * ga[2]) * f;
        K2[0] += g[25 + a] * g[25 + b] * f * m[3];
    }
    static __device__ __forceinline__ bool bad_diag(const double *) { return false; }
    static __device__ __forceinline__ void load(const double *s, double *K, double *K2) {
        K[0] += s[0];
        K2[0] += s[1];
    }
};

// Gauss-point data of one element in LDS: nip x G doubles padded to a
// stride of 1 (mod 32) doubles, so that lanes reading the same entry of
// consecutive elements hit consecutive bank pairs.
__host__ __device__ __forceinline__ int gs_stride(int nip, int g) {
    const int n = nip * g;
    return n + (((1 - n) % 32) + 32) % 32;
}

// Element stage: workgroup of 256 threads, epw elements.  LDS: coordinates
// (epw x DIM x NN), material (epw x MS), Gauss-point data (epw x gs_stride).
// The blocks are computed pair by pair across the workgroup's elements (a
// wave holds a few local pairs of 16 consecutive elements), so the writes of
// one local pair are contiguous in the pair-major block array.
template <class P>
__global__ void __launch_bounds__(256) k_elem_blocks(AsmArgs A, const double *__restrict__ gtab, int epw,
                                                     double *eblk) {
    bind_tables(A, gtab);
    extern __shared__ double sh[];
    const int nip = A.nip;
    const int gsp = gs_stride(nip, P::G);
    double *xs = sh;
    double *ms = xs + epw * P::DIM * P::NN;
    double *gs = ms + epw * P::MS;
    const long long e0 = (long long)blockIdx.x * epw;
    for (int t = threadIdx.x; t < epw * P::NN; t += blockDim.x) {
        const int el = t / P::NN, i = t - el * P::NN;
        const long long e = e0 + el;
        if (e < A.n_el) {
            const long long node = A.conn[e * P::NN + i];
#pragma unroll
            for (int d = 0; d < P::DIM; d++) xs[(el * P::DIM + d) * P::NN + i] = A.xyz[P::DIM * node + d];
        }
    }
    for (int el = threadIdx.x; el < epw; el += blockDim.x)
        if (e0 + el < A.n_el) P::stage_mat(A, (int)(e0 + el), ms + el * P::MS);
    __syncthreads();
    int bad = 0;
    for (int t = threadIdx.x; t < epw * nip; t += blockDim.x) {
        const int el = t / nip, ip = t - el * nip;
        const long long e = e0 + el;
        if (e < A.n_el)
            bad |= P::stage_ip(A, (int)e, ip, xs + el * P::DIM * P::NN, gs + (long long)el * gsp + ip * P::G);
    }
    __syncthreads();
    for (int t = threadIdx.x; t < epw * P::NP; t += blockDim.x) {
        const int pr = t / epw, el = t - pr * epw;
        const long long e = e0 + el;
        if (e >= A.n_el) continue;
        int i, j;
        pair_of_index(P::NN, pr, i, j);
        const int ni = A.conn[e * P::NN + i], nj = A.conn[e * P::NN + j];
        const int a = ni <= nj ? i : j, b = ni <= nj ? j : i;
        double K[P::NB], K2[P::NB];
#pragma unroll
        for (int q = 0; q < P::NB; q++) K[q] = K2[q] = 0;
        const double *m = ms + el * P::MS;
        for (int ip = 0; ip < nip; ip++) P::add_ip(m, gs + (long long)el * gsp + ip * P::G, a, b, K, K2);
        double *out = eblk + ((long long)pr * A.n_el + e) * P::NBS;
#pragma unroll
        for (int q = 0; q < P::NB; q++) out[q] = K[q];
        if (P::NBS > P::NB) out[P::NB] = K2[0];
        if (i == j && P::bad_diag(K)) bad |= 2;
    }
    if (bad && A.flags) atomicOr(&A.flags[0], bad);
}

// Canonical block of pair (lo, hi) summed over the elements holding both
// nodes, in ascending element order, from the element stage's blocks: the
// two nodes' sorted incidence lists are intersected, and their local indices
// come with them.  Lists of up to kShortList entries (every node of a
// structured hex8 or quad4 mesh) are loaded at once and intersected in
// registers; longer ones are merged entry by entry.
constexpr int kShortList = 8;
template <class P>
__device__ __forceinline__ void pair_block(const AsmArgs &A, const int *n2e_ptr, const int *n2e,
                                           int, int lo, int hi, double *K, double *K2) {
#pragma unroll
    for (int i = 0; i < P::NB; i++) K[i] = K2[i] = 0;
    const bool self = (lo == hi);
    int i = n2e_ptr[lo], j = n2e_ptr[hi];
    const int ie = n2e_ptr[lo + 1], je = n2e_ptr[hi + 1];
    if (ie - i <= kShortList && je - j <= kShortList) {
        int el[kShortList], eh[kShortList], la[kShortList], lh[kShortList];
#pragma unroll
        for (int k = 0; k < kShortList; k++) {
            el[k] = i + k < ie ? n2e[i + k] : -1;
            la[k] = i + k < ie ? A.n2e_loc[i + k] : 0;
            eh[k] = j + k < je ? n2e[j + k] : -2;
            lh[k] = j + k < je ? A.n2e_loc[j + k] : 0;
        }
#pragma unroll
        for (int a = 0; a < kShortList; a++) {
            int b = -1;
#pragma unroll
            for (int c = 0; c < kShortList; c++)
                if (eh[c] == el[a]) b = lh[c];
            if (b >= 0) {
                const int la_ = la[a];
                const int pr = la_ <= b ? pair_index(P::NN, la_, b) : pair_index(P::NN, b, la_);
                P::load(A.eblk + ((long long)pr * A.n_el + el[a]) * P::NBS, K, K2);
            }
        }
        i = ie;  // done
    }
    while (i < ie && j < je) {
        const int el = n2e[i], eh = n2e[j];
        if (el < eh) {
            i++;
        } else if (eh < el) {
            j++;
        } else {
            const int a = A.n2e_loc[i], b = A.n2e_loc[j];
            const int pr = a <= b ? pair_index(P::NN, a, b) : pair_index(P::NN, b, a);
            P::load(A.eblk + ((long long)pr * A.n_el + el) * P::NBS, K, K2);
            i++;
            j++;
        }
    }
    if (self) {  // keep the upper triangle of a diagonal block (Ke is upper-only)
#pragma unroll
        for (int c = 0; c < P::ND; c++)
#pragma unroll
            for (int d = 0; d < c; d++) {
                K[c * P::ND + d] = K[d * P::ND + c];
                K2[c * P::ND + d] = K2[d * P::ND + c];
            }
    }
}

// Oriented read of the canonical block: value at (row comp c of n, col comp d of m).
template <int ND>
__device__ __forceinline__ double oriented(const double *K, bool n_is_lo, int c, int d) {
    return n_is_lo ? K[c * ND + d] : K[d * ND + c];
}

struct AsmOut {
    const int *pair_row, *nbr, *nbr_off, *first, *rowptr;
    const unsigned char *rmask;
    double *val, *val_aux;
    long long n_pairs;
    int mode;  // 0: val = K ; 1 (heat): val = (a1 K + M)/dt, val_aux = K
    double a1, dt;
    // mode 0 with a current blocked SELL layout: the pair's block is also
    // written into the solver's SELL value arrays (FP32, and FP64 when
    // sval64 is set), restrained entries zero and restrained diagonals one,
    // the values the CSR-to-SELL conversion would write
    float *sval32;
    double *sval64;
    const int *sptr, *nbr_ptr;
    long long nsell;
};

// One node pair (n, m), pair index p: the canonical block summed over the
// elements holding both nodes in ascending element order, written into the
// CSR row(s) of n and, when dst >= 0, into the SELL block values at dst.
template <class P>
__device__ __forceinline__ void asm_pair(const AsmArgs &A, const AsmOut &O, const int *n2e_ptr, const int *n2e,
                                         long long p, int n, int m, long long dst) {
    int lo = n < m ? n : m, hi = n < m ? m : n;
    double K[P::NB], K2[P::NB];
    pair_block<P>(A, n2e_ptr, n2e, n, lo, hi, K, K2);
    const bool nlo = (n == lo);
    unsigned rn = O.rmask[n], rm = O.rmask[m];
    if (dst >= 0) {
#pragma unroll
        for (int c = 0; c < P::ND; c++)
#pragma unroll
            for (int d = 0; d < P::ND; d++) {
                const bool rc = (rn >> c) & 1u, rd = (rm >> d) & 1u;
                double v = 0;
                if (!rc && !rd) v = oriented<P::ND>(K, nlo, c, d);
                else if (rc && n == m && c == d) v = 1;
                const long long o = (long long)(c * P::ND + d) * O.nsell + dst;
                // streaming stores: whole lines per wave, not read again
                // before the solve, so they leave L2 to the CSR pieces
                __builtin_nontemporal_store((float)v, O.sval32 + o);
                if (O.sval64) __builtin_nontemporal_store(v, O.sval64 + o);
            }
    }
#pragma unroll
    for (int c = 0; c < P::ND; c++) {
        if ((rn >> c) & 1u) continue;
        int I = O.first[n] + c;
        long long base = (long long)O.rowptr[I] + O.nbr_off[p];
        int r = 0;
#pragma unroll
        for (int d = 0; d < P::ND; d++) {
            if ((rm >> d) & 1u) continue;
            double kv = oriented<P::ND>(K, nlo, c, d);
            if (O.mode == 1) {
                double mv = oriented<P::ND>(K2, nlo, c, d);
                O.val[base + r] = (O.a1 * kv + mv) / O.dt;
                O.val_aux[base + r] = kv;
            } else {
                O.val[base + r] = kv;
            }
            r++;
        }
    }
}

// Gather stage in pair order (no current SELL layout: the first assembly, and
// heat): one thread per node pair, consecutive threads on consecutive
// neighbours of a node, so the CSR writes of a wave are contiguous.
template <class P>
__global__ void __launch_bounds__(256) k_asm_pairs(AsmArgs A, AsmOut O, const int *n2e_ptr, const int *n2e,
                                                   const double *__restrict__ gtab) {
    bind_tables(A, gtab);
    const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= O.n_pairs) return;
    asm_pair<P>(A, O, n2e_ptr, n2e, p, O.pair_row[p], O.nbr[p], -1);
}

// Gather stage in row order, when the solver's SELL layout is current (every
// reassembly after the first solve): a workgroup takes four slots of one
// 64-node slice, wave w slot 4 g + w for the workgroup's group g, lane l the
// slice's node 64 s + l.  In a slot the lanes hold consecutive nodes and,
// away from mesh boundaries, the same relative neighbour, so they read the
// blocks of consecutive elements, which are contiguous in the pair-major
// block array, and write consecutive SELL entries.  Workgroups go to the
// eight XCDs in turn (workgroup b to XCD b mod 8), so the groups of one slice
// are given workgroup numbers of one residue mod 8, consecutive among those:
// the slice's CSR rows (a contiguous range, each row written in 3-entry
// pieces by its slots) then fill in one XCD's L2 at about the same time and
// leave it as whole lines.  Slots are those of the layout: union slots,
// whose neighbour's rank in the node's list slot_rank holds (255: a zero
// block, left as it is), or the node's k-th neighbour.
template <class P>
__global__ void __launch_bounds__(256) k_asm_rows(AsmArgs A, AsmOut O, const int *n2e_ptr, const int *n2e,
                                                  const double *__restrict__ gtab, int n_nodes,
                                                  const unsigned char *slot_rank, int groups, long long slice0,
                                                  long long slice_end) {
    bind_tables(A, gtab);
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const unsigned xcd = blockIdx.x & 7, k = blockIdx.x >> 3;
    const unsigned q = k / (unsigned)groups;
    const int j = 4 * (int)(k - q * (unsigned)groups) + wave;
    const long long s = slice0 + 8LL * q + xcd;
    if (s >= slice_end) return;
    const int W = (O.sptr[s + 1] - O.sptr[s]) >> 6;
    if (j >= W) return;
    const long long n = s * 64 + lane;
    if (n >= n_nodes) return;
    const long long dst = (long long)O.sptr[s] + 64LL * j + lane;
    const int pb = O.nbr_ptr[n], pe = O.nbr_ptr[n + 1];
    long long p = -1;
    int m = -1;
    if (slot_rank) {
        const int rk = slot_rank[dst];
        if (rk != 255) {
            p = pb + rk;
            m = O.nbr[p];
        }
    } else if (pb + j < pe) {
        p = pb + j;
        m = O.nbr[p];
    }
    if (p >= 0) asm_pair<P>(A, O, n2e_ptr, n2e, p, (int)n, m, dst);
}

// Per node: pseudo[I] for every free DOF I of the node =
//   - sum over restrained J of Kscat(I,J) * presc[J]         (addMatrix lift)
//   + fcoef * sum over all J of Kraw(I,J) * u_state[J]       (element forces, heat)
template <class P>
__global__ void __launch_bounds__(256) k_gather_rhs(AsmArgs A, const double *__restrict__ gtab, AsmOut O,
                                                    const int *n2e_ptr,
                                                    const int *n2e, const int *nbr_ptr, int n_nodes,
                                                    const int *ndof, const double *presc,
                                                    const double *u_state, double fcoef,
                                                    double *pseudo) {
    bind_tables(A, gtab);
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    unsigned rn = O.rmask[n];
    double acc[P::ND];
#pragma unroll
    for (int c = 0; c < P::ND; c++) acc[c] = 0;
    const int nd_n = ndof[n];
    if (nd_n > 0 && rn != (1u << nd_n) - 1u) {
        for (int p = nbr_ptr[n]; p < nbr_ptr[n + 1]; p++) {
            int m = O.nbr[p];
            unsigned rm = O.rmask[m];
            if (rm == 0 && u_state == nullptr) continue;
            int lo = n < m ? n : m, hi = n < m ? m : n;
            double K[P::NB], K2[P::NB];
            pair_block<P>(A, n2e_ptr, n2e, n, lo, hi, K, K2);
            const bool nlo = (n == lo);
#pragma unroll
            for (int c = 0; c < P::ND; c++) {
                if ((rn >> c) & 1u) continue;
#pragma unroll
                for (int d = 0; d < P::ND; d++) {
                    int J = O.first[m] + d;
                    double kv = oriented<P::ND>(K, nlo, c, d);
                    if ((rm >> d) & 1u) {
                        double sv = kv;
                        if (O.mode == 1) sv = (O.a1 * kv + oriented<P::ND>(K2, nlo, c, d)) / O.dt;
                        acc[c] -= sv * presc[J];
                    }
                    if (u_state) acc[c] += fcoef * kv * u_state[J];
                }
            }
        }
    }
    for (int c = 0; c < nd_n; c++) pseudo[O.first[n] + c] = ((rn >> c) & 1u) ? 0.0 : acc[c];
}

// Restrained rows hold 1 on the diagonal (equation.cpp:500).
__global__ void k_restrained_diag(int n_dof, const unsigned char *rest, const int *rowptr,
                                  double *val, double *val_aux) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n_dof || !rest[i]) return;
    val[rowptr[i]] = 1.0;
    if (val_aux) val_aux[rowptr[i]] = 1.0;
}

// Gauss tables of the context (weights, shape derivatives and values).
int upload_gauss(bhip_ctx *ctx, int nnpe, int order, int *nip_out) {
    return bhip_gauss_upload(ctx, nnpe, order, nip_out);
}

// Elements per workgroup of the element stage and its LDS bytes.
template <class P>
int elem_stage_epw(int nip) {
    // at most 16 elements, one thread per Gauss point, and 60 KiB of LDS
    const size_t per_el = sizeof(double) * ((size_t)P::DIM * P::NN + P::MS + (size_t)gs_stride(nip, P::G));
    const int by_lds = (int)((60 * 1024) / per_el);
    return std::max(1, std::min(std::min(16, 256 / std::max(nip, 1)), by_lds));
}
template <class P>
size_t elem_stage_lds(int epw, int nip) {
    return sizeof(double) * ((size_t)epw * P::DIM * P::NN + (size_t)epw * P::MS + (size_t)epw * gs_stride(nip, P::G));
}

// Element stage into ctx->eblk (kept until the next assembly: the lift table
// is built from it on first use).
template <class P>
int element_stage(bhip_ctx *ctx, AsmArgs &A) {
    BHIP_TRY(ctx, ctx->eblk.resize((size_t)ctx->n_el * P::NP * P::NBS > 0 ? (size_t)ctx->n_el * P::NP * P::NBS : 1));
    const int epw = elem_stage_epw<P>(A.nip);
    const size_t lds = elem_stage_lds<P>(epw, A.nip);
    if (lds > 64 * 1024) return bhip_fail(ctx, BHIP_ERR_CAPACITY, "integration order too high for the element stage");
    const int grid = grid_for(ctx->n_el, epw);
    k_elem_blocks<P><<<grid, 256, lds, ctx->stream>>>(A, ctx->gauss.p, epw, ctx->eblk.p);
    BHIP_TRY(ctx, hipGetLastError());
    A.eblk = ctx->eblk.p;
    return BHIP_OK;
}

template <class P>
int launch_assembly(bhip_ctx *ctx, AsmArgs A, AsmOut O, const double *u_state, double fcoef) {
    const int B = 256;
    hipStream_t s = ctx->stream;
    BHIP_PROPAGATE(element_stage<P>(ctx, A));
    if (O.mode == 0 && O.sval32) {
        const unsigned char *srank = ctx->sell_union ? ctx->slot_rank.p : nullptr;
        const int groups = (ctx->sell_wmax + 3) / 4;  // four slots of the widest slice per workgroup
        // launches of at most 2^22 workgroups (the grid's x extent in threads
        // stays below 2^32), each over a multiple of 8 slices but the last
        const long long per_launch = std::max(8LL, ((1LL << 22) / groups) & ~7LL);
        for (long long s0 = 0; s0 < ctx->n_slices && ctx->sell_wmax > 0; s0 += per_launch) {
            const long long cnt = std::min(per_launch, (long long)ctx->n_slices - s0);
            const long long blocks = (cnt + 7) / 8 * 8 * groups;
            k_asm_rows<P><<<(unsigned)blocks, B, 0, s>>>(A, O, ctx->n2e_ptr.p, ctx->n2e.p, ctx->gauss.p,
                                                        ctx->n_nodes, srank, groups, s0, s0 + cnt);
        }
    } else {
        k_asm_pairs<P><<<grid_for(O.n_pairs, B), B, 0, s>>>(A, O, ctx->n2e_ptr.p, ctx->n2e.p, ctx->gauss.p);
    }
    BHIP_TRY(ctx, hipGetLastError());
    k_restrained_diag<<<grid_for(ctx->n_dof, B), B, 0, s>>>(ctx->n_dof, ctx->rest.p, ctx->rowptr.p,
                                                            ctx->val.p, O.mode == 1 ? ctx->val_aux.p : nullptr);
    BHIP_TRY(ctx, hipGetLastError());
    k_gather_rhs<P><<<grid_for(ctx->n_nodes, B), B, 0, s>>>(
        A, ctx->gauss.p, O, ctx->n2e_ptr.p, ctx->n2e.p, ctx->nbr_ptr.p, ctx->n_nodes, ctx->ndof.p, ctx->presc.p,
        u_state, fcoef, ctx->pseudo.p);
    BHIP_TRY(ctx, hipGetLastError());
    return BHIP_OK;
}

}  // namespace

int bhip_assemble_impl(bhip_ctx *ctx, int what, const double *coef, const double *u_state) {
    if (ctx->external_matrix || ctx->n_nodes == 0)
        return bhip_fail(ctx, BHIP_ERR_STATE, "no mesh: the matrix was loaded from CSR/RCV data or the mesh is not set");
    if (!ctx->pattern_ready) return bhip_fail(ctx, BHIP_ERR_STATE, "pattern not built");
    if (!ctx->have_mat) return bhip_fail(ctx, BHIP_ERR_STATE, "materials not set");
    BHIP_PROPAGATE(bhip_ensure_vectors(ctx));
    // captured iteration graphs stay: new values go into the same buffers, and
    // a graph whose buffers moved is re-captured by its key (solver.hip)
    ctx->lift_ready = false;
    int nip = 0;
    BHIP_PROPAGATE(upload_gauss(ctx, ctx->nnpe, ctx->order, &nip));
    BHIP_TRY(ctx, hipMemsetAsync(ctx->flags.p, 0, sizeof(int) * 4, ctx->stream));

    AsmArgs A;
    A.xyz = ctx->xyz.p;
    A.conn = ctx->conn.p;
    A.mat = ctx->mat.p;
    A.elem_group = ctx->elem_group.n ? ctx->elem_group.p : nullptr;
    A.n_el = ctx->n_el;
    A.n_groups = ctx->n_groups;
    A.mat_stride = ctx->mat_stride;
    A.nnpe = ctx->nnpe;
    A.nip = nip;
    A.gw = ctx->gauss.p;
    A.gdN = ctx->gauss.p + kGaussDN;
    A.gN = ctx->gauss.p + kGaussN;
    A.factor = ctx->factor;
    A.tan = nullptr;
    A.flags = ctx->flags.p;
    A.eblk = nullptr;
    A.n2e_loc = ctx->n2e_loc.p;
    AsmOut O;
    O.pair_row = ctx->pair_row.p;
    O.nbr = ctx->nbr.p;
    O.nbr_off = ctx->nbr_off.p;
    O.first = ctx->dof_first.p;
    O.rowptr = ctx->rowptr.p;
    O.rmask = ctx->node_rmask.p;
    O.val = ctx->val.p;
    O.val_aux = nullptr;
    O.n_pairs = ctx->n_pairs;
    O.mode = 0;
    O.a1 = 0;
    O.dt = 1;
    // assembly straight into the solver's blocked layout when it is current
    O.sval32 = nullptr;
    O.sval64 = nullptr;
    O.sptr = ctx->sell_ptr.p;
    O.nbr_ptr = ctx->nbr_ptr.p;
    O.nsell = ctx->n_sell;
    const bool to_sell = bhip_sell_values_direct(ctx) && (what == BHIP_ASM_STIFFNESS || what == BHIP_ASM_PLASTIC_TANGENT);
    if (to_sell) {
        O.sval32 = ctx->bval32.p;
        if (ctx->bval64_layout_ok) O.sval64 = ctx->bval64.p;
    }

    const int kind = ctx->mat_kind;
    int st = BHIP_OK;
    if (what == BHIP_ASM_STIFFNESS) {
        if (ctx->nnpe == 8 && ctx->dim == 3 && ctx->max_ndof == 3 && ctx->elem_ndof_full) {
            if (kind == BHIP_MAT_ISO_ELASTIC)
                st = launch_assembly<ElasHex8<BHIP_MAT_ISO_ELASTIC>>(ctx, A, O, nullptr, 0.0);
            else if (kind == BHIP_MAT_CMAT3D)
                st = launch_assembly<ElasHex8<BHIP_MAT_CMAT3D>>(ctx, A, O, nullptr, 0.0);
            else if (kind == BHIP_MAT_J2)  // elastic stiffness of the J2 material
                st = launch_assembly<ElasHex8<BHIP_MAT_ISO_ELASTIC>>(ctx, A, O, nullptr, 0.0);
            else
                return bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "material kind not valid for hex8 stiffness");
        } else if (ctx->nnpe == 4 && ctx->dim == 2 && ctx->max_ndof == 2 && ctx->elem_ndof_full &&
                   kind == BHIP_MAT_PLANE_STRESS) {
            st = launch_assembly<ElasQuad4>(ctx, A, O, nullptr, 0.0);
        } else {
            return bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "element/material combination not built");
        }
    } else if (what == BHIP_ASM_HEAT_TRANSIENT) {
        if (!(ctx->nnpe == 8 && ctx->dim == 3 && ctx->max_ndof == 1 && ctx->uniform_ndof &&
              kind == BHIP_MAT_HEAT))
            return bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "heat transient needs hex8, 1 DOF/node, heat material");
        if (!coef) return bhip_fail(ctx, BHIP_ERR_ARG, "heat transient needs coef = {alpha, dt}");
        double alpha = coef[0], dt = coef[1];
        if (!(dt > 0)) return bhip_fail(ctx, BHIP_ERR_ARG, "dt must be positive");
        BHIP_TRY(ctx, ctx->val_aux.resize(ctx->val.n));
        O.mode = 1;
        O.val_aux = ctx->val_aux.p;
        O.a1 = alpha * dt;  // DiffusionModel.cpp:745
        O.dt = dt;
        ctx->heat_a1 = O.a1;
        ctx->heat_dt = dt;
        double a2 = (1 - alpha) * dt;
        double fcoef = -(O.a1 + a2) / dt;  // formF_s_new: -(a1+a2) K u_s / dt
        st = launch_assembly<HeatHex8>(ctx, A, O, u_state, fcoef);
    } else if (what == BHIP_ASM_PLASTIC_TANGENT) {
        if (!(ctx->nnpe == 8 && ctx->dim == 3 && kind == BHIP_MAT_J2))
            return bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "plastic tangent needs hex8 and a J2 material");
        A.tan = bhip_plastic_tangent(ctx);
        if (!A.tan) return bhip_fail(ctx, BHIP_ERR_STATE, "plastic state not initialised");
        st = launch_assembly<ElasHex8<BHIP_MAT_J2>>(ctx, A, O, nullptr, 0.0);
    } else {
        return bhip_fail(ctx, BHIP_ERR_ARG, "unknown assembly kind");
    }
    BHIP_PROPAGATE(st);
    int hflags[4] = {0, 0, 0, 0};
    BHIP_TRY(ctx, hipMemcpyAsync(hflags, ctx->flags.p, sizeof(int) * 4, hipMemcpyDeviceToHost, ctx->stream));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->assembled = true;
    ctx->asm_kind = what;
    ctx->f32_ready = O.sval32 != nullptr;
    ctx->f64b_ready = O.sval64 != nullptr;
    ctx->lift_ready = false;
    if (hflags[0] & 1) return bhip_fail(ctx, BHIP_ERR_NEG_DETJ, "negative Jacobian determinant (inverted element)");
    if (hflags[0] & 2) return bhip_fail(ctx, BHIP_ERR_NEG_DIAG, "bad diagonal in an element stiffness");
    return BHIP_OK;
}

// ------------------------------------------------------------ lift table
// Equations::updatePseudoLoadVectorUsingGK (equation.cpp:343-381) re-lifts new
// prescribed values through an assembled, unrestrained K: F[i] -= K(i,j) u_j
// for free i, restrained j.  Here the free-row x restrained-column coupling
// blocks of the last stiffness assembly are kept as a table (one entry per
// (node with free DOFs, neighbour with restrained DOFs), neighbours in
// ascending order), built once per assembly by one thread per entry, so each
// load case costs one pass over the table instead of a re-integration.
namespace {

__global__ void k_lift_count(int n_nodes, const int *ndof, const unsigned char *rmask, const int *nbr_ptr,
                             const int *nbr, long long *cnt) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n > n_nodes) return;
    long long c = 0;
    if (n < n_nodes) {
        const int nd = ndof[n];
        const unsigned rn = rmask[n];
        if (nd > 0 && rn != (1u << nd) - 1u)
            for (int p = nbr_ptr[n]; p < nbr_ptr[n + 1]; p++) c += rmask[nbr[p]] != 0;
    }
    cnt[n] = c;  // entry n_nodes stays 0: the scan's last element is the total
}

__global__ void k_lift_fill(int n_nodes, const unsigned char *rmask, const int *nbr_ptr, const int *nbr,
                            const long long *lptr, int *lrow, int *lcol) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    long long o = lptr[n];
    if (lptr[n + 1] == o) return;
    for (int p = nbr_ptr[n]; p < nbr_ptr[n + 1]; p++) {
        int m = nbr[p];
        if (rmask[m] != 0) {
            lrow[o] = n;
            lcol[o] = m;
            o++;
        }
    }
}

template <class P>
__global__ void __launch_bounds__(256) k_lift_blocks(AsmArgs A, const double *__restrict__ gtab,
                                                     const int *n2e_ptr, const int *n2e,
                                                     long long n_lift, const int *lrow, const int *lcol,
                                                     double *lval) {
    bind_tables(A, gtab);
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_lift) return;
    const int n = lrow[k], m = lcol[k];
    const int lo = n < m ? n : m, hi = n < m ? m : n;
    double K[P::NB], K2[P::NB];
    pair_block<P>(A, n2e_ptr, n2e, n, lo, hi, K, K2);
    const bool nlo = (n == lo);
#pragma unroll
    for (int c = 0; c < P::ND; c++)
#pragma unroll
        for (int d = 0; d < P::ND; d++) lval[k * P::NB + c * P::ND + d] = oriented<P::ND>(K, nlo, c, d);
}

// Heat (1 DOF per node): entry k holds K(n,m) and (a1 K(n,m) + M(n,m))/dt,
// the two coefficients the assembly's gather applies to u_s and to the
// prescribed increment.
__global__ void __launch_bounds__(256) k_lift_blocks_heat(AsmArgs A, const double *__restrict__ gtab,
                                                          const int *n2e_ptr, const int *n2e,
                                                          long long n_lift, const int *lrow, const int *lcol,
                                                          double a1, double dt, double *lval) {
    bind_tables(A, gtab);
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_lift) return;
    const int n = lrow[k], m = lcol[k];
    const int lo = n < m ? n : m, hi = n < m ? m : n;
    double K[1], K2[1];
    pair_block<HeatHex8>(A, n2e_ptr, n2e, n, lo, hi, K, K2);
    lval[2 * k] = K[0];
    lval[2 * k + 1] = (a1 * K[0] + K2[0]) / dt;
}

// b_r[I] = restrained ? presc_r[I] : load[I] - sum over the node's table
// entries (ascending neighbour, then component) of K(I,J) presc_r[J],
// J restrained -- the order in which the assembly's gather accumulates.
template <int ND>
__global__ void k_lift_rhs(int n_nodes, int R, long long n_dof, const int *first, const unsigned char *rmask,
                           const long long *lptr, const int *lcol, const double *lval, const double *presc,
                           const double *load, double *b) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_nodes) return;
    const unsigned rn = rmask[n];
    const int f = first[n];
    const long long kb = lptr[n], ke = lptr[n + 1];
    for (int r = 0; r < R; r++) {
        const double *pr = presc + (long long)r * n_dof;
        double acc[ND];
#pragma unroll
        for (int c = 0; c < ND; c++) acc[c] = 0;
        for (long long k = kb; k < ke; k++) {
            const int m = lcol[k];
            const unsigned rm = rmask[m];
            const int fm = first[m];
#pragma unroll
            for (int c = 0; c < ND; c++) {
                if ((rn >> c) & 1u) continue;
#pragma unroll
                for (int d = 0; d < ND; d++)
                    if ((rm >> d) & 1u) acc[c] -= lval[k * ND * ND + c * ND + d] * pr[fm + d];
            }
        }
        double *br = b + (long long)r * n_dof;
#pragma unroll
        for (int c = 0; c < ND; c++) {
            const long long I = f + c;
            br[I] = ((rn >> c) & 1u) ? pr[I] : (load ? load[I] : 0.0) + acc[c];
        }
    }
}

}  // namespace

int bhip_lift_build(bhip_ctx *ctx) {
    if (ctx->lift_ready) return BHIP_OK;
    const bool heat = ctx->assembled && ctx->asm_kind == BHIP_ASM_HEAT_TRANSIENT;
    if (!ctx->assembled || (ctx->asm_kind != BHIP_ASM_STIFFNESS && !heat))
        return bhip_fail(ctx, BHIP_ERR_STATE,
                         "the lift table needs a stiffness or heat-transient assembly");
    const int kind = ctx->mat_kind;
    const bool hex = !heat && ctx->nnpe == 8 && ctx->max_ndof == 3 && ctx->uniform_ndof &&
                     (kind == BHIP_MAT_ISO_ELASTIC || kind == BHIP_MAT_J2 || kind == BHIP_MAT_CMAT3D);
    const bool quad = !heat && ctx->nnpe == 4 && ctx->max_ndof == 2 && ctx->uniform_ndof &&
                      kind == BHIP_MAT_PLANE_STRESS;
    if (heat && !(ctx->nnpe == 8 && ctx->max_ndof == 1 && ctx->uniform_ndof && kind == BHIP_MAT_HEAT))
        return bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "heat lift table needs hex8 with 1 DOF per node");
    if (!hex && !quad && !heat)
        return bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "re-lifting not built for this element/material");
    hipStream_t s = ctx->stream;
    const int B = 256;
    const int nn = ctx->n_nodes;
    dvec<long long> &cnt = ctx->tmp_l1;
    BHIP_TRY(ctx, cnt.resize((size_t)nn + 1));
    BHIP_TRY(ctx, ctx->lift_ptr.resize((size_t)nn + 1));
    k_lift_count<<<grid_for(nn + 1, B), B, 0, s>>>(nn, ctx->ndof.p, ctx->node_rmask.p, ctx->nbr_ptr.p, ctx->nbr.p,
                                                   cnt.p);
    BHIP_TRY(ctx, hipGetLastError());
    BHIP_PROPAGATE(bhip_scan_i64(ctx, cnt.p, ctx->lift_ptr.p, (long long)nn + 1));
    long long total = 0;
    BHIP_TRY(ctx, bhip_copy(ctx, &total, ctx->lift_ptr.p + nn, sizeof(long long), hipMemcpyDeviceToHost));
    ctx->n_lift = total;
    const size_t ne = total > 0 ? (size_t)total : 1;
    const int per = hex ? 9 : quad ? 4 : 2;
    BHIP_TRY(ctx, ctx->lift_row.resize(ne));
    BHIP_TRY(ctx, ctx->lift_col.resize(ne));
    BHIP_TRY(ctx, ctx->lift_val.resize(ne * per));
    if (total > 0) {
        k_lift_fill<<<grid_for(nn, B), B, 0, s>>>(nn, ctx->node_rmask.p, ctx->nbr_ptr.p, ctx->nbr.p, ctx->lift_ptr.p,
                                                  ctx->lift_row.p, ctx->lift_col.p);
        int nip = 0;
        BHIP_PROPAGATE(upload_gauss(ctx, ctx->nnpe, ctx->order, &nip));
        AsmArgs A;
        A.xyz = ctx->xyz.p;
        A.conn = ctx->conn.p;
        A.mat = ctx->mat.p;
        A.elem_group = ctx->elem_group.n ? ctx->elem_group.p : nullptr;
        A.n_el = ctx->n_el;
        A.n_groups = ctx->n_groups;
        A.mat_stride = ctx->mat_stride;
        A.nnpe = ctx->nnpe;
        A.nip = nip;
        A.gw = ctx->gauss.p;
        A.gdN = ctx->gauss.p + kGaussDN;
        A.gN = ctx->gauss.p + kGaussN;
        A.factor = ctx->factor;
        A.tan = nullptr;
        A.flags = nullptr;
        A.eblk = ctx->eblk.p;  // element blocks of the assembly this table belongs to
        A.n2e_loc = ctx->n2e_loc.p;
        const int g = grid_for(total, B);
        if (heat)
            k_lift_blocks_heat<<<g, B, 0, s>>>(A, ctx->gauss.p, ctx->n2e_ptr.p, ctx->n2e.p, total, ctx->lift_row.p,
                                               ctx->lift_col.p, ctx->heat_a1, ctx->heat_dt, ctx->lift_val.p);
        else if (quad)
            k_lift_blocks<ElasQuad4><<<g, B, 0, s>>>(A, ctx->gauss.p, ctx->n2e_ptr.p, ctx->n2e.p, total, ctx->lift_row.p,
                                                     ctx->lift_col.p, ctx->lift_val.p);
        else if (kind == BHIP_MAT_CMAT3D)
            k_lift_blocks<ElasHex8<BHIP_MAT_CMAT3D>><<<g, B, 0, s>>>(A, ctx->gauss.p, ctx->n2e_ptr.p, ctx->n2e.p, total,
                                                                     ctx->lift_row.p, ctx->lift_col.p,
                                                                     ctx->lift_val.p);
        else  // isotropic, and the elastic stiffness of the J2 material
            k_lift_blocks<ElasHex8<BHIP_MAT_ISO_ELASTIC>><<<g, B, 0, s>>>(A, ctx->gauss.p, ctx->n2e_ptr.p, ctx->n2e.p, total,
                                                                          ctx->lift_row.p, ctx->lift_col.p,
                                                                          ctx->lift_val.p);
        BHIP_TRY(ctx, hipGetLastError());
    }
    BHIP_TRY(ctx, hipStreamSynchronize(s));
    ctx->lift_ready = true;
    return BHIP_OK;
}

int bhip_lift_rhs(bhip_ctx *ctx, int R, const double *presc, const double *load, double *b) {
    if (ctx->assembled && ctx->asm_kind != BHIP_ASM_STIFFNESS)
        return bhip_fail(ctx, BHIP_ERR_STATE, "re-lifting needs an assembled stiffness (BHIP_ASM_STIFFNESS)");
    BHIP_PROPAGATE(bhip_lift_build(ctx));
    const int B = 256;
    if (ctx->max_ndof == 3)
        k_lift_rhs<3><<<grid_for(ctx->n_nodes, B), B, 0, ctx->stream>>>(
            ctx->n_nodes, R, ctx->n_dof, ctx->dof_first.p, ctx->node_rmask.p, ctx->lift_ptr.p, ctx->lift_col.p,
            ctx->lift_val.p, presc, load, b);
    else
        k_lift_rhs<2><<<grid_for(ctx->n_nodes, B), B, 0, ctx->stream>>>(
            ctx->n_nodes, R, ctx->n_dof, ctx->dof_first.p, ctx->node_rmask.p, ctx->lift_ptr.p, ctx->lift_col.p,
            ctx->lift_val.p, presc, load, b);
    BHIP_TRY(ctx, hipGetLastError());
    return BHIP_OK;
}

// Gauss tables from the reference's Gauss-Legendre construction
// (gauss.cpp:14-60, 303-365, 825-877), kept per context.
int bhip_gauss_upload(bhip_ctx *ctx, int nnpe, int order, int *nip_out) {
    if (ctx->gauss.p && ctx->gauss_nnpe == nnpe && ctx->gauss_order == order) {
        *nip_out = ctx->gauss_nip;
        return BHIP_OK;
    }
    std::vector<double> hw, hdN, hN;
    const int nip = bhip_gauss::tables(nnpe, order, kMaxIp, hw, hdN, hN);
    if (nip == 0) return bhip_fail(ctx, BHIP_ERR_UNSUPPORTED, "element family or integration order not built");
    std::vector<double> all((size_t)kGaussN + kMaxIp * 8, 0.0);
    std::copy(hw.begin(), hw.begin() + kMaxIp, all.begin());
    std::copy(hdN.begin(), hdN.begin() + kMaxIp * 24, all.begin() + kGaussDN);
    std::copy(hN.begin(), hN.begin() + kMaxIp * 8, all.begin() + kGaussN);
    // earlier kernels of this context may still read the old tables
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    BHIP_TRY(ctx, ctx->gauss.resize(all.size()));
    BHIP_TRY(ctx, hipMemcpyAsync(ctx->gauss.p, all.data(), sizeof(double) * all.size(), hipMemcpyHostToDevice,
                                 ctx->stream));
    BHIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    ctx->gauss_nnpe = nnpe;
    ctx->gauss_order = order;
    ctx->gauss_nip = nip;
    *nip_out = nip;
    return BHIP_OK;
}
