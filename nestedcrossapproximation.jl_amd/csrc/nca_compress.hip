// Phase 2 / Phase 3 of the batched bottom-up compressor and the coupling
// matrices, for gfx950: level assembly, batched fully pivoted ACA, batched
// skeleton LU + solve, coupling evaluation.
//
// Compiled with -ffp-contract=off: pivots and ranks must not depend on which
// products the compiler fuses. The ACA norms are summed in one fixed order
// (256 per-thread partial sums over the column-major entries e = t, t+256, ...,
// then a halving tree), so the stopping decisions are reproducible; argmax
// ties go to the lowest column-major index, which makes the reduction order
// irrelevant for the pivots themselves.

#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <numeric>

#include "nca_common.hpp"
#include "nca_math.hpp"

namespace nca {

constexpr int kWG = 256;
// ACA with the residual in LDS: m*n <= kLdsCap, m <= kLdsMaxM, n <= kLdsMaxN.
constexpr int kLdsCap = 3072;
constexpr int kLdsMaxM = 512;
constexpr int kLdsMaxN = 64;
// ACA with the residual in a global working copy.
constexpr int kGlobMaxM = 2048;
constexpr int kGlobMaxN = 256;
// Skeleton solve: the r x r cross matrix lives in LDS.
constexpr int kSolveMaxRank = 96;
// Coupling: test pivots of one cluster in LDS.
constexpr int kCouplingMaxRank = 256;

// ---------------------------------------------------------------------------
// Level assembly: one thread per buffer entry (row fastest).
__global__ void __launch_bounds__(kWG)
    k_assemble_level(const double* __restrict__ rowpts, const double* __restrict__ colpts,
                     const int32_t* __restrict__ rowblock, const int32_t* __restrict__ rowidx,
                     const int64_t* __restrict__ colptr, const int32_t* __restrict__ colidx, int64_t ld,
                     int64_t maxrank, double k, int sign, int transposed, double2* __restrict__ buf) {
    const int64_t idx = (int64_t)blockIdx.x * kWG + threadIdx.x;
    if (idx >= ld * maxrank) return;
    const int64_t row = idx % ld, col = idx / ld;
    const int b = rowblock[row];
    const int64_t c0 = colptr[b];
    double2 g = make_double2(0.0, 0.0);
    if (col < colptr[b + 1] - c0) {
        const double* xr = rowpts + 3 * (int64_t)rowidx[row];
        const double* xc = colpts + 3 * (int64_t)colidx[c0 + col];
        g = transposed ? helmholtz(k, sign, xc[0], xc[1], xc[2], xr[0], xr[1], xr[2])
                       : helmholtz(k, sign, xr[0], xr[1], xr[2], xc[0], xc[1], xc[2]);
    }
    buf[idx] = g;
}

// ---------------------------------------------------------------------------
// Workgroup reductions. Sum: halving tree over 256 slots (slot t += slot t+s).
// Argmax: larger value, ties to the smaller index.
struct ArgRed {
    double sum[kWG];
    double val[kWG];
    int idx[kWG];
};

__device__ __forceinline__ void reduce_sum_argmax(ArgRed& red, double p, double bv, int bi, double& sum,
                                                  double& vmax, int& imax) {
    const int t = threadIdx.x;
    red.sum[t] = p;
    red.val[t] = bv;
    red.idx[t] = bi;
    __syncthreads();
    for (int s = kWG / 2; s > 0; s >>= 1) {
        if (t < s) {
            red.sum[t] = red.sum[t] + red.sum[t + s];
            const double v2 = red.val[t + s];
            const int i2 = red.idx[t + s];
            if (v2 > red.val[t] || (v2 == red.val[t] && i2 < red.idx[t])) {
                red.val[t] = v2;
                red.idx[t] = i2;
            }
        }
        __syncthreads();
    }
    sum = red.sum[0];
    vmax = red.val[0];
    imax = red.idx[0];
}

// Fully pivoted ACA with the sampled stop on one block, one workgroup; the
// block's entries are src(a, c) (read once, into the residual R).
//   r crosses accepted, residual R_r, s = block columns;
//   stop when r == kmax = min(maxrank, m, n) or
//        ||R_r||_F^2 * (sampled ? s / (s - r) : 1) <= tol^2 ||A||_F^2;
//   pivot = first argmax |R_r|^2 (column-major), R -= R(:,j) (R(i,:) / R(i,j)),
//   row i and column j set to zero. piv(k, i, j) records cross k (thread 0);
//   returns the rank.
template <class Src, class Piv>
__device__ __forceinline__ int aca_block(Src src, double2* R, int64_t ldr, double2* u, double2* v, int m, int n,
                                         int kmax, double tol, int sampled, ArgRed& red, Piv piv) {
    const int t = threadIdx.x;
    const int mn = m * n;
    const int qstep = kWG / m, rstep = kWG % m;
    double p = 0.0, bv = -1.0;
    int bi = 0;
    {
        int a = t % m, c = t / m;
        for (int e = t; e < mn; e += kWG) {
            const double2 val = src(a, c);
            R[a + c * ldr] = val;
            const double w = cabs2(val);
            p += w;
            if (w > bv) { bv = w; bi = e; }
            a += rstep;
            c += qstep;
            if (a >= m) { a -= m; ++c; }
        }
    }
    double nrm2, bmax;
    int bidx;
    reduce_sum_argmax(red, p, bv, bi, nrm2, bmax, bidx);
    const double nrmA2 = nrm2;
    const double tol2 = tol * tol;
    int r = 0;
    while (r < kmax) {
        const double scale = sampled ? (double)n / (double)(n - r) : 1.0;
        if (nrm2 * scale <= tol2 * nrmA2) break;
        if (!(bmax > 0.0)) break;
        const int i = bidx % m, j = bidx / m;
        const double2 pv = R[i + j * ldr];
        const double d = cabs2(pv);
        const double2 inv = make_double2(pv.x / d, -pv.y / d);
        for (int a = t; a < m; a += kWG) u[a] = R[a + j * ldr];
        for (int c = t; c < n; c += kWG) v[c] = cmul(R[i + c * ldr], inv);
        if (t == 0) piv(r, i, j);
        ++r;
        __syncthreads();
        p = 0.0;
        bv = -1.0;
        bi = 0;
        int a = t % m, c = t / m;
        for (int e = t; e < mn; e += kWG) {
            double2 val;
            if (a == i || c == j) {
                val = make_double2(0.0, 0.0);
            } else {
                val = csub(R[a + c * ldr], cmul(u[a], v[c]));
            }
            R[a + c * ldr] = val;
            const double w = cabs2(val);
            p += w;
            if (w > bv) { bv = w; bi = e; }
            a += rstep;
            c += qstep;
            if (a >= m) { a -= m; ++c; }
        }
        reduce_sum_argmax(red, p, bv, bi, nrm2, bmax, bidx);
    }
    return r;
}

// Incomplete ACA (IACA) with geometric column pivoting on one block, one
// workgroup (the method of the incomplete adaptive cross approximation with
// tree-mimicry pivoting: the column pivots come from the geometry, the row
// pivots from the residual columns by maximum value). The block's columns are
// the candidates in their geometric order (Phase 1), so step r takes column
// j = r; its residual R_r(:, j) gives the row pivot i = first argmax
// |R_r(a, j)|^2. Stop, without counting the candidate, when r reaches
// kmax = min(maxrank, m, n), when the candidate's residual column is zero, or when
//   r ||R_r(:, j)||^2 <= tol^2 sum_{l < r} ||A(:, l)||^2     (r > 0),
// i.e. the candidate column is represented to tol relative to the mean
// accepted column. The residual update is the ACA's: R -= R(:, j) (R(i, :) /
// R(i, j)) with row i and column j set to zero; the column sums use the ACA's
// fixed order (slot e mod 256 over the column-major entries, halving tree).
// src(a, c): the block's entries (read for the residual and for the original
// column norms).
template <class Src, class Piv>
__device__ __forceinline__ int iaca_block(Src src, double2* R, int64_t ldr, double2* u, double2* v, int m, int n,
                                          int kmax, double tol, ArgRed& red, Piv piv) {
    const int t = threadIdx.x;
    const int mn = m * n;
    const int qstep = kWG / m, rstep = kWG % m;
    // the copy, and the sums of column 0 (its residual is the original)
    double p = 0.0, bv = -1.0;
    int bi = 0;
    {
        int a = t % m, c = t / m;
        for (int e = t; e < mn; e += kWG) {
            const double2 val = src(a, c);
            R[a + c * ldr] = val;
            if (c == 0) {
                const double w = cabs2(val);
                p += w;
                if (w > bv) { bv = w; bi = a; }
            }
            a += rstep;
            c += qstep;
            if (a >= m) { a -= m; ++c; }
        }
    }
    double cn2, bmax;
    int bidx;
    reduce_sum_argmax(red, p, bv, bi, cn2, bmax, bidx);
    double orig2 = cn2;  // ||A(:, j)||^2 of the candidate
    double acc = 0.0;    // sum of ||A(:, l)||^2 over the accepted columns
    const double tol2 = tol * tol;
    int r = 0;
    while (r < kmax) {
        if (r > 0 && cn2 * (double)r <= tol2 * acc) break;
        if (!(bmax > 0.0)) break;
        const int i = bidx, j = r;
        const double2 pv = R[i + j * ldr];
        const double d = cabs2(pv);
        const double2 inv = make_double2(pv.x / d, -pv.y / d);
        __syncthreads();
        for (int a = t; a < m; a += kWG) u[a] = R[a + j * ldr];
        for (int c = t; c < n; c += kWG) v[c] = cmul(R[i + c * ldr], inv);
        if (t == 0) piv(r, i, j);
        acc += orig2;
        ++r;
        __syncthreads();
        p = 0.0;
        double q = 0.0;
        bv = -1.0;
        bi = 0;
        int a = t % m, c = t / m;
        for (int e = t; e < mn; e += kWG) {
            double2 val;
            if (a == i || c == j) {
                val = make_double2(0.0, 0.0);
            } else {
                val = csub(R[a + c * ldr], cmul(u[a], v[c]));
            }
            R[a + c * ldr] = val;
            if (c == r) {
                const double w = cabs2(val);
                p += w;
                if (w > bv) { bv = w; bi = a; }
                q += cabs2(src(a, c));
            }
            a += rstep;
            c += qstep;
            if (a >= m) { a -= m; ++c; }
        }
        reduce_sum_argmax(red, p, bv, bi, cn2, bmax, bidx);
        __syncthreads();
        double unusedv;
        int unusedi;
        reduce_sum_argmax(red, q, -1.0, 0, orig2, unusedv, unusedi);
    }
    return r;
}

// Batched fully pivoted ACA (aca_block), one workgroup per block of the stack;
// the residual in LDS (kLds) or in the global working copy work.
template <bool kLds>
__global__ void __launch_bounds__(kWG)
    k_aca(const double2* __restrict__ buf, double2* __restrict__ work, int64_t ld,
          const int64_t* __restrict__ rowoff, const int32_t* __restrict__ ncols,
          const int32_t* __restrict__ blocklist, int64_t maxrank, double tol, int sampled,
          int32_t* __restrict__ ranks, int64_t* __restrict__ rowpiv, int64_t* __restrict__ colpiv) {
    extern __shared__ double2 dyn[];
    __shared__ ArgRed red;
    const int t = threadIdx.x;
    const int b = blocklist[blockIdx.x];
    const int64_t r0 = rowoff[b];
    const int m = (int)(rowoff[b + 1] - r0);
    const int n = ncols[b];
    for (int64_t i = t; i < maxrank; i += kWG) {
        rowpiv[i + b * maxrank] = 0;
        colpiv[i + b * maxrank] = 0;
    }
    int kmax = (int)maxrank;
    kmax = m < kmax ? m : kmax;
    kmax = n < kmax ? n : kmax;
    if (kmax <= 0) {
        if (t == 0) ranks[b] = 0;
        return;
    }
    double2* R = kLds ? dyn : work + r0;
    const int64_t ldr = kLds ? m : ld;
    double2* u = kLds ? dyn + kLdsCap : dyn;
    double2* v = kLds ? u + kLdsMaxM : dyn + kGlobMaxM;
    auto src = [&](int a, int c) { return buf[r0 + a + (int64_t)c * ld]; };
    auto piv = [&](int k, int i, int j) {
        rowpiv[k + b * maxrank] = i + 1;
        colpiv[k + b * maxrank] = j + 1;
    };
    const int r = aca_block(src, R, ldr, u, v, m, n, kmax, tol, sampled, red, piv);
    if (t == 0) ranks[b] = r;
}

// Batched IACA (iaca_block), one workgroup per block of the stack.
template <bool kLds>
__global__ void __launch_bounds__(kWG)
    k_iaca(const double2* __restrict__ buf, double2* __restrict__ work, int64_t ld,
           const int64_t* __restrict__ rowoff, const int32_t* __restrict__ ncols,
           const int32_t* __restrict__ blocklist, int64_t maxrank, double tol, int32_t* __restrict__ ranks,
           int64_t* __restrict__ rowpiv, int64_t* __restrict__ colpiv) {
    extern __shared__ double2 dyn[];
    __shared__ ArgRed red;
    const int t = threadIdx.x;
    const int b = blocklist[blockIdx.x];
    const int64_t r0 = rowoff[b];
    const int m = (int)(rowoff[b + 1] - r0);
    const int n = ncols[b];
    for (int64_t i = t; i < maxrank; i += kWG) {
        rowpiv[i + b * maxrank] = 0;
        colpiv[i + b * maxrank] = 0;
    }
    int kmax = (int)maxrank;
    kmax = m < kmax ? m : kmax;
    kmax = n < kmax ? n : kmax;
    if (kmax <= 0) {
        if (t == 0) ranks[b] = 0;
        return;
    }
    double2* R = kLds ? dyn : work + r0;
    const int64_t ldr = kLds ? m : ld;
    double2* u = kLds ? dyn + kLdsCap : dyn;
    double2* v = kLds ? u + kLdsMaxM : dyn + kGlobMaxM;
    auto src = [&](int a, int c) { return buf[r0 + a + (int64_t)c * ld]; };
    auto piv = [&](int k, int i, int j) {
        rowpiv[k + b * maxrank] = i + 1;
        colpiv[k + b * maxrank] = j + 1;
    };
    const int r = iaca_block(src, R, ldr, u, v, m, n, kmax, tol, red, piv);
    if (t == 0) ranks[b] = r;
}

// ---------------------------------------------------------------------------
// Skeleton solve on one block, one workgroup: LU with partial pivoting (pivot
// = first max |re|+|im|, LAPACK getrf form) of P = A[rowpiv, colpiv] in LDS
// (A: r x r), then per row a of L = A[:, colpiv] (one thread per row): w U = l,
// y L = w, interchanges undone last to first; X = L P^{-1} written m x r
// column-major at Xb. src(a, c): block entries; rp(k) / cp(k): 0-based pivots.
// Returns (thread 0) the independent-pivot count of the singularity rule, -1
// when regular.
template <class Src, class RP, class CP>
__device__ __forceinline__ int64_t skeleton_block(Src src, RP rp, CP cp, int m, int r, double2* A, int* ipiv,
                                                  int* cpiv, ArgRed& red, double2* __restrict__ Xb) {
    const int t = threadIdx.x;
    for (int c = t; c < r; c += kWG) cpiv[c] = cp(c);
    __syncthreads();
    for (int e = t; e < r * r; e += kWG) {
        const int kk = e % r, c = e / r;
        A[e] = src(rp(kk), cpiv[c]);
    }
    __syncthreads();
    int info = 0;
    for (int k = 0; k < r; ++k) {
        double v = -1.0;
        int ii = r;
        if (k + t < r) {
            v = cabs1(A[(k + t) + k * r]);
            ii = k + t;
        }
        double dummy, best;
        int p;
        reduce_sum_argmax(red, 0.0, v, ii, dummy, best, p);
        if (t == 0) ipiv[k] = p;
        if (best != 0.0) {
            if (p != k)
                for (int c = t; c < r; c += kWG) {
                    const double2 tmp = A[k + c * r];
                    A[k + c * r] = A[p + c * r];
                    A[p + c * r] = tmp;
                }
            __syncthreads();
            for (int i = k + 1 + t; i < r; i += kWG) A[i + k * r] = cdiv(A[i + k * r], A[k + k * r]);
        } else if (info == 0) {
            info = k + 1;
        }
        __syncthreads();
        const int nt = r - k - 1;
        for (int e = t; e < nt * nt; e += kWG) {
            const int i = k + 1 + e % nt, c = k + 1 + e / nt;
            A[i + c * r] = csub(A[i + c * r], cmul(A[i + k * r], A[k + c * r]));
        }
        __syncthreads();
    }
    int64_t result = -1;
    if (t == 0) {
        double d2max = 0.0, d2min = INFINITY;
        for (int i = 0; i < r; ++i) {
            const double d2 = cabs2(A[i + i * r]);
            if (d2 > d2max) d2max = d2;
            if (d2 < d2min) d2min = d2;
        }
        const double eps = 2.220446049250313e-16;
        const double tol2 = (10.0 * eps) * (10.0 * eps) * d2max;
        if (!(info == 0 && d2min > tol2)) {
            int64_t rank = 0;
            for (int i = 0; i < r; ++i) rank += cabs2(A[i + i * r]) > tol2;
            result = rank;
        }
    }
    for (int a = t; a < m; a += kWG) {
        double2* xr = Xb + a;
        for (int c = 0; c < r; ++c) {
            double2 acc = src(a, cpiv[c]);
            for (int kk = 0; kk < c; ++kk) acc = csub(acc, cmul(xr[(int64_t)kk * m], A[kk + c * r]));
            xr[(int64_t)c * m] = cdiv(acc, A[c + c * r]);
        }
        for (int c = r - 1; c >= 0; --c) {
            double2 acc = xr[(int64_t)c * m];
            for (int kk = c + 1; kk < r; ++kk) acc = csub(acc, cmul(xr[(int64_t)kk * m], A[kk + c * r]));
            xr[(int64_t)c * m] = acc;
        }
        for (int kk = r - 1; kk >= 0; --kk) {
            const int p = ipiv[kk];
            if (p != kk) {
                const double2 tmp = xr[(int64_t)kk * m];
                xr[(int64_t)kk * m] = xr[(int64_t)p * m];
                xr[(int64_t)p * m] = tmp;
            }
        }
    }
    return result;
}

// Batched skeleton solve (skeleton_block), one workgroup per block of the stack.
__global__ void __launch_bounds__(kWG)
    k_skeleton(const double2* __restrict__ buf, int64_t ld, const int64_t* __restrict__ rowoff,
               const int32_t* __restrict__ ranks, const int64_t* __restrict__ rowpiv,
               const int64_t* __restrict__ colpiv, int64_t maxrank, const int64_t* __restrict__ xoff,
               double2* __restrict__ X, int64_t* __restrict__ indep) {
    extern __shared__ double2 A[];
    __shared__ int ipiv[kSolveMaxRank];
    __shared__ int cpiv[kSolveMaxRank];
    __shared__ ArgRed red;
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    const int r = ranks[b];
    if (r == 0) {
        if (t == 0) indep[b] = -1;
        return;
    }
    const int64_t r0 = rowoff[b];
    const int m = (int)(rowoff[b + 1] - r0);
    auto src = [&](int a, int c) { return buf[r0 + a + (int64_t)c * ld]; };
    auto rp = [&](int k) { return (int)rowpiv[k + b * maxrank] - 1; };
    auto cp = [&](int k) { return (int)colpiv[k + b * maxrank] - 1; };
    const int64_t res = skeleton_block(src, rp, cp, m, r, A, ipiv, cpiv, red, X + xoff[b]);
    if (t == 0) indep[b] = res;
}

// ---------------------------------------------------------------------------
// One level of the bottom-up compressor in one launch, one workgroup per
// block: the block's rows (a leaf's points, else its children's pivots, read
// from the pivots the previous launch left on the device), its entries
// evaluated into the residual (LDS, or the block's global working copy when
// it is larger), the ACA or IACA, the skeleton solve with the pivot rows and
// columns evaluated again, and the factor X written to the level's chunk. The
// arithmetic is that of k_assemble_level + k_aca / k_iaca + k_skeleton, so the
// pivots, ranks and factors are the same bit for bit.
struct LevelTree {
    const int32_t* childptr;  // nids + 1, 0-based
    const int32_t* children;
    const int32_t* lo;        // values of a leaf: perm[lo, hi)
    const int32_t* hi;
    const int32_t* perm;
};

struct LevelOut {
    int32_t* piv;     // nids x maxrank: pivot point ids (0-based)
    int32_t* rank;    // nids
    int32_t* mrows;   // nids
    int64_t* indep;   // nids
    int32_t* err;     // [0]: a block above the kernel limits, [1]: its id
};

template <bool IACA>
__global__ void __launch_bounds__(kWG)
    k_level_fused(const double* __restrict__ rowpts, const double* __restrict__ colpts, double k, int sign,
                  int transposed, const int32_t* __restrict__ ids, const int32_t* __restrict__ ncols,
                  const int64_t* __restrict__ colptr, const int32_t* __restrict__ colidx,
                  const int64_t* __restrict__ xoff, const int64_t* __restrict__ workoff, LevelTree tr,
                  int64_t maxrank, double tol, int sampled, double2* __restrict__ work, double2* __restrict__ X,
                  LevelOut out) {
    extern __shared__ double2 dyn[];
    __shared__ ArgRed red;
    __shared__ int rows[kGlobMaxM];
    __shared__ int rpiv[kGlobMaxN], cpivl[kGlobMaxN];
    __shared__ int ipiv[kSolveMaxRank], cpiv[kSolveMaxRank];
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    const int id = ids[b];
    const int n = ncols[b];
    const int32_t* cols = colidx + colptr[b];
    // rows: the leaf's points or the children's pivots, in child order
    // (the host bounded them by kGlobMaxM)
    const int c0 = tr.childptr[id], c1 = tr.childptr[id + 1];
    int m = 0;
    if (c0 == c1) {
        const int l0 = tr.lo[id];
        m = tr.hi[id] - l0;
        for (int a = t; a < min(m, kGlobMaxM); a += kWG) rows[a] = tr.perm[l0 + a];
    } else {
        for (int q = c0; q < c1; ++q) {
            const int ch = tr.children[q];
            const int rc = out.rank[ch];
            for (int a = t; a < rc && m + a < kGlobMaxM; a += kWG) rows[m + a] = out.piv[(int64_t)ch * maxrank + a];
            m += rc;
        }
    }
    __syncthreads();
    int kmax = (int)maxrank;
    kmax = m < kmax ? m : kmax;
    kmax = n < kmax ? n : kmax;
    if (t == 0) {
        out.mrows[id] = m;
        out.rank[id] = 0;
        out.indep[id] = -1;
    }
    if (kmax <= 0) return;
    const bool lds = m * n <= kLdsCap && m <= kLdsMaxM && n <= kLdsMaxN;
    if (!lds && (m > kGlobMaxM || n > kGlobMaxN || workoff[b] < 0)) {
        if (t == 0) {
            out.err[0] = 1;
            out.err[1] = id;
        }
        return;
    }
    auto src = [&](int a, int c) {
        const double* xr = rowpts + 3 * (int64_t)rows[a];
        const double* xc = colpts + 3 * (int64_t)cols[c];
        return transposed ? helmholtz(k, sign, xc[0], xc[1], xc[2], xr[0], xr[1], xr[2])
                          : helmholtz(k, sign, xr[0], xr[1], xr[2], xc[0], xc[1], xc[2]);
    };
    auto piv = [&](int kk, int i, int j) {
        rpiv[kk] = i;
        cpivl[kk] = j;
    };
    double2* R = lds ? dyn : work + workoff[b];
    double2* u = lds ? dyn + kLdsCap : dyn;
    double2* v = lds ? u + kLdsMaxM : dyn + kGlobMaxM;
    const int r = IACA ? iaca_block(src, R, m, u, v, m, n, kmax, tol, red, piv)
                       : aca_block(src, R, m, u, v, m, n, kmax, tol, sampled, red, piv);
    __syncthreads();
    for (int q = t; q < r; q += kWG) out.piv[(int64_t)id * maxrank + q] = rows[rpiv[q]];
    if (t == 0) out.rank[id] = r;
    if (r == 0) return;
    auto rp = [&](int q) { return rpiv[q]; };
    auto cp = [&](int q) { return cpivl[q]; };
    const int64_t res = skeleton_block(src, rp, cp, m, r, dyn, ipiv, cpiv, red, X + xoff[b]);
    if (t == 0) out.indep[id] = res;
}

// ---------------------------------------------------------------------------
// TopDown (topdown.jl:51-206): one level of clusters, one workgroup per
// cluster, the partially pivoted ACA of the block G(rows, cols) evaluated
// entrywise (never assembled): rows = the cluster's points, cols = the points
// of its far nodes (in CSR order) and then the column pivots of its parent.
//   row pivot i_1 = row 0; step r: residual row R(i_r, :) into V(:, r); column
//   pivot j_r = first maximum of |R(i_r, c)|^2 over unused columns (a zero
//   maximum stops); V(:, r) /= R(i_r, j_r); U(:, r) = residual column j_r;
//   S2 += |u|^2 |v|^2 + 2 Re sum_{l<r} (u_l^H u_r)(v_r v_l^H); the cross
//   counts; stop when |u|^2 |v|^2 <= tol^2 S2 or r = min(maxrank, m, n); next
//   row pivot = first maximum of |U(a, r)|^2 over unused rows.
// Sums over rows (columns) are the fixed slot order (index mod 256 in
// increasing index, halving tree), sums over l in increasing l; a thread owns
// the rows a = t mod 256 of U and the columns c = t mod 256 of V. Scratch
// per cluster: U (m x kmax), V (ncap x kmax), the columns' point ids, the used
// flags. Outputs: rank (-1: no far field), the row and column pivots (point ids).
struct PacaScratch {
    double2* U;
    double2* V;
    int32_t* cols;
    uint8_t* used;
    const int64_t* uoff;  // per block of the launch
    const int64_t* voff;
    const int64_t* coff;  // columns and column flags
    const int64_t* roff;  // row flags
};

struct PacaTree {
    const int32_t* lo;       // self: values perm[lo, hi)
    const int32_t* hi;
    const int32_t* perm;
    const int32_t* parent;   // -1 at a root
    const int64_t* farptr;   // far nodes of the other tree
    const int32_t* farnodes;
    const int32_t* olo;      // other tree
    const int32_t* ohi;
    const int32_t* operm;
};

__device__ __forceinline__ void reduce4(double* r0, double* r1, double* r2, double* r3, double& a, double& b,
                                        double& c, double& d) {
    const int t = threadIdx.x;
    r0[t] = a;
    r1[t] = b;
    r2[t] = c;
    r3[t] = d;
    __syncthreads();
    for (int s = kWG / 2; s > 0; s >>= 1) {
        if (t < s) {
            r0[t] = r0[t] + r0[t + s];
            r1[t] = r1[t] + r1[t + s];
            r2[t] = r2[t] + r2[t + s];
            r3[t] = r3[t] + r3[t + s];
        }
        __syncthreads();
    }
    a = r0[0];
    b = r1[0];
    c = r2[0];
    d = r3[0];
    __syncthreads();
}

// best (value, index, payload) over the workgroup: larger value, ties to the smaller index
__device__ __forceinline__ void reduce_best(double* rv, int* ri, double2* rp, double& v, int& i, double2& p) {
    const int t = threadIdx.x;
    rv[t] = v;
    ri[t] = i;
    rp[t] = p;
    __syncthreads();
    for (int s = kWG / 2; s > 0; s >>= 1) {
        if (t < s) {
            const double v2 = rv[t + s];
            const int i2 = ri[t + s];
            if (i2 >= 0 && (ri[t] < 0 || v2 > rv[t] || (v2 == rv[t] && i2 < ri[t]))) {
                rv[t] = v2;
                ri[t] = i2;
                rp[t] = rp[t + s];
            }
        }
        __syncthreads();
    }
    v = rv[0];
    i = ri[0];
    p = rp[0];
    __syncthreads();
}

__global__ void __launch_bounds__(kWG)
    k_paca_level(const double* __restrict__ rowpts, const double* __restrict__ colpts, double k, int sign,
                 int transposed, const int32_t* __restrict__ ids, PacaTree tr, PacaScratch sc, int64_t maxrank,
                 double tol, int32_t* __restrict__ rank, int32_t* __restrict__ rpiv, int32_t* __restrict__ cpiv) {
    __shared__ double r0[kWG], r1[kWG], r2[kWG], r3[kWG];
    __shared__ int ri[kWG];
    __shared__ double2 rp[kWG];
    __shared__ double2 ui[kSolveMaxRank], vj[kSolveMaxRank];
    const int t = threadIdx.x;
    const int b = blockIdx.x;
    const int id = ids[b];
    const int l0 = tr.lo[id];
    const int m = tr.hi[id] - l0;
    const int32_t* rows = tr.perm + l0;
    int32_t* cols = sc.cols + sc.coff[b];
    uint8_t* usedc = sc.used + sc.coff[b];
    uint8_t* usedr = sc.used + sc.roff[b];
    // columns: the far nodes' points, then the parent's column pivots
    int n = 0;
    for (int64_t q = tr.farptr[id]; q < tr.farptr[id + 1]; ++q) {
        const int s = tr.farnodes[q];
        const int o0 = tr.olo[s], ns = tr.ohi[s] - o0;
        for (int a = t; a < ns; a += kWG) cols[n + a] = tr.operm[o0 + a];
        n += ns;
    }
    const int pa = tr.parent[id];
    const int rpa = pa >= 0 ? max(rank[pa], 0) : 0;
    for (int a = t; a < rpa; a += kWG) cols[n + a] = cpiv[(int64_t)pa * maxrank + a];
    n += rpa;
    if (n == 0) {
        if (t == 0) rank[id] = -1;
        return;
    }
    for (int a = t; a < n; a += kWG) usedc[a] = 0;
    for (int a = t; a < m; a += kWG) usedr[a] = 0;
    int kmax = (int)maxrank;
    kmax = m < kmax ? m : kmax;
    kmax = n < kmax ? n : kmax;
    double2* U = sc.U + sc.uoff[b];
    double2* V = sc.V + sc.voff[b];
    __syncthreads();
    auto G = [&](int a, int c) {
        const double* xr = rowpts + 3 * (int64_t)rows[a];
        const double* xc = colpts + 3 * (int64_t)cols[c];
        return transposed ? helmholtz(k, sign, xc[0], xc[1], xc[2], xr[0], xr[1], xr[2])
                          : helmholtz(k, sign, xr[0], xr[1], xr[2], xc[0], xc[1], xc[2]);
    };
    const double tol2 = tol * tol;
    double S2 = 0.0;
    int i = 0, r = 0;
    while (r < kmax) {
        // U(i, l), l < r, from the thread that owns row i
        if (t == (i & (kWG - 1)))
            for (int l = 0; l < r; ++l) ui[l] = U[i + (int64_t)l * m];
        __syncthreads();
        double bv = -1.0;
        int bj = -1;
        double2 bp = make_double2(0.0, 0.0);
        for (int c = t; c < n; c += kWG) {
            double2 acc = G(i, c);
            for (int l = 0; l < r; ++l) acc = csub(acc, cmul(ui[l], V[c + (int64_t)l * n]));
            V[c + (int64_t)r * n] = acc;
            if (!usedc[c]) {
                const double w = cabs2(acc);
                if (w > bv) { bv = w; bj = c; bp = acc; }
            }
        }
        reduce_best(r0, ri, rp, bv, bj, bp);
        if (!(bv > 0.0)) break;
        const int j = bj;
        const double d = cabs2(bp);
        const double2 inv = make_double2(bp.x / d, -bp.y / d);
        for (int c = t; c < n; c += kWG) V[c + (int64_t)r * n] = cmul(V[c + (int64_t)r * n], inv);
        // V(j, l), l < r, from the thread that owns column j
        if (t == (j & (kWG - 1)))
            for (int l = 0; l < r; ++l) vj[l] = V[j + (int64_t)l * n];
        __syncthreads();
        for (int a = t; a < m; a += kWG) {
            double2 acc = G(a, j);
            for (int l = 0; l < r; ++l) acc = csub(acc, cmul(vj[l], U[a + (int64_t)l * m]));
            U[a + (int64_t)r * m] = acc;
        }
        double nu2 = 0.0, nv2 = 0.0, z0 = 0.0, z1 = 0.0;
        for (int a = t; a < m; a += kWG) nu2 += cabs2(U[a + (int64_t)r * m]);
        for (int c = t; c < n; c += kWG) nv2 += cabs2(V[c + (int64_t)r * n]);
        reduce4(r0, r1, r2, r3, nu2, nv2, z0, z1);
        double cross = 0.0;
        for (int l = 0; l < r; ++l) {
            double ure = 0.0, uim = 0.0, vre = 0.0, vim = 0.0;
            for (int a = t; a < m; a += kWG) {
                const double2 q = cmul(make_double2(U[a + (int64_t)l * m].x, -U[a + (int64_t)l * m].y),
                                       U[a + (int64_t)r * m]);
                ure += q.x;
                uim += q.y;
            }
            for (int c = t; c < n; c += kWG) {
                const double2 vl = V[c + (int64_t)l * n];
                const double2 q = cmul(V[c + (int64_t)r * n], make_double2(vl.x, -vl.y));
                vre += q.x;
                vim += q.y;
            }
            reduce4(r0, r1, r2, r3, ure, uim, vre, vim);
            cross = cross + (ure * vre - uim * vim);
        }
        S2 = (S2 + nu2 * nv2) + 2.0 * cross;
        if (t == 0) rpiv[(int64_t)id * maxrank + r] = rows[i];
        if (t == (i & (kWG - 1))) usedr[i] = 1;
        if (t == (j & (kWG - 1))) {
            cpiv[(int64_t)id * maxrank + r] = cols[j];
            usedc[j] = 1;
        }
        ++r;
        if (nu2 * nv2 <= tol2 * S2) break;
        __syncthreads();
        double bu = -1.0;
        int ba = -1;
        double2 dummy = make_double2(0.0, 0.0);
        for (int a = t; a < m; a += kWG)
            if (!usedr[a]) {
                const double w = cabs2(U[a + (int64_t)(r - 1) * m]);
                if (w > bu) { bu = w; ba = a; }
            }
        reduce_best(r0, ri, rp, bu, ba, dummy);
        if (ba < 0) break;
        i = ba;
    }
    if (t == 0) rank[id] = r;
}

// The factors of TopDown, one workgroup per cluster with pivots:
// X = A[rows, sigma] A[tau, sigma]^{-1}, rows = the points of a leaf or the
// children's pivots in child order, solved as skeleton_block on the rows
// followed by tau (the pivots of the cross matrix in the appended rows).
__global__ void __launch_bounds__(kWG)
    k_topdown_factors(const double* __restrict__ rowpts, const double* __restrict__ colpts, double k, int sign,
                      int transposed, const int32_t* __restrict__ ids, const int64_t* __restrict__ xoff,
                      const int32_t* __restrict__ childptr, const int32_t* __restrict__ children,
                      const int32_t* __restrict__ lo, const int32_t* __restrict__ hi, const int32_t* __restrict__ perm,
                      const int32_t* __restrict__ rank, const int32_t* __restrict__ rpiv,
                      const int32_t* __restrict__ cpiv, int64_t maxrank, int32_t* __restrict__ mrows,
                      int64_t* __restrict__ indep, double2* __restrict__ X, int32_t* __restrict__ err) {
    extern __shared__ double2 dyn[];
    __shared__ ArgRed red;
    __shared__ int rows[kGlobMaxM + kSolveMaxRank];
    __shared__ int sig[kSolveMaxRank];
    __shared__ int ipiv[kSolveMaxRank], cpv[kSolveMaxRank];
    const int t = threadIdx.x;
    const int id = ids[blockIdx.x];
    const int r = rank[id];
    if (r < 0) {  // no far field: no factor
        if (t == 0) {
            mrows[id] = 0;
            indep[id] = -1;
        }
        return;
    }
    int m = 0;
    const int c0 = childptr[id], c1 = childptr[id + 1];
    if (c0 == c1) {
        m = hi[id] - lo[id];
        for (int a = t; a < min(m, kGlobMaxM); a += kWG) rows[a] = perm[lo[id] + a];
    } else {
        for (int q = c0; q < c1; ++q) {
            const int ch = children[q];
            const int rc = max(rank[ch], 0);
            for (int a = t; a < rc && m + a < kGlobMaxM; a += kWG) rows[m + a] = rpiv[(int64_t)ch * maxrank + a];
            m += rc;
        }
    }
    if (t == 0) {
        mrows[id] = m;
        indep[id] = -1;
    }
    if (m > kGlobMaxM) {
        if (t == 0) {
            err[0] = 1;
            err[1] = id;
        }
        return;
    }
    for (int a = t; a < r; a += kWG) {
        rows[m + a] = rpiv[(int64_t)id * maxrank + a];
        sig[a] = cpiv[(int64_t)id * maxrank + a];
    }
    __syncthreads();
    auto src = [&](int a, int c) {
        const double* xr = rowpts + 3 * (int64_t)rows[a];
        const double* xc = colpts + 3 * (int64_t)sig[c];
        return transposed ? helmholtz(k, sign, xc[0], xc[1], xc[2], xr[0], xr[1], xr[2])
                          : helmholtz(k, sign, xr[0], xr[1], xr[2], xc[0], xc[1], xc[2]);
    };
    auto rp = [&](int q) { return m + q; };
    auto cp = [&](int q) { return q; };
    const int64_t res = skeleton_block(src, rp, cp, m, r, dyn, ipiv, cpv, red, X + xoff[blockIdx.x]);
    if (t == 0) indep[id] = res;
}

// ---------------------------------------------------------------------------
// Coupling blocks of one test cluster t: the blocks of its CSR row are
// contiguous and share r_t rows, so together they form one r_t x C column-major
// matrix; one thread per entry, the block of a column found by a forward scan.
__global__ void __launch_bounds__(kWG)
    k_coupling(const double* __restrict__ testpts, const double* __restrict__ trialpts,
               const int64_t* __restrict__ farptr, const int32_t* __restrict__ farnodes,
               const int64_t* __restrict__ offsets, const int64_t* __restrict__ testpivptr,
               const int32_t* __restrict__ testpiv, const int64_t* __restrict__ trialpivptr,
               const int32_t* __restrict__ trialpiv, double k, int sign, double2* __restrict__ blocks) {
    __shared__ double tp[3 * kCouplingMaxRank];
    const int t = blockIdx.x;
    const int64_t q0 = farptr[t], q1 = farptr[t + 1];
    const int rt = (int)(testpivptr[t + 1] - testpivptr[t]);
    if (rt == 0 || q0 == q1) return;
    for (int i = threadIdx.x; i < rt; i += kWG) {
        const double* x = testpts + 3 * (int64_t)testpiv[testpivptr[t] + i];
        tp[3 * i] = x[0];
        tp[3 * i + 1] = x[1];
        tp[3 * i + 2] = x[2];
    }
    __syncthreads();
    const int64_t base = offsets[q0];
    const int64_t total = offsets[q1] - base;
    int64_t q = q0;
    int64_t colbeg = 0, colend = (offsets[q0 + 1] - base) / rt;
    for (int64_t e = threadIdx.x; e < total; e += kWG) {
        const int64_t col = e / rt;
        const int row = (int)(e - col * rt);
        while (col >= colend) {
            ++q;
            colbeg = colend;
            colend = (offsets[q + 1] - base) / rt;
        }
        const int s = farnodes[q];
        const double* y = trialpts + 3 * (int64_t)trialpiv[trialpivptr[s] + (col - colbeg)];
        blocks[base + e] = helmholtz(k, sign, tp[3 * row], tp[3 * row + 1], tp[3 * row + 2], y[0], y[1], y[2]);
    }
}

inline unsigned grid_for(int64_t n) { return (unsigned)((n + kWG - 1) / kWG); }

}  // namespace nca

using namespace nca;

extern "C" {

int32_t nca_level_assemble(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test,
                           const nca_points* trial, int64_t nblocks, const int64_t* rowptr,
                           const int64_t* rowidx, const int64_t* colptr, const int64_t* colidx,
                           int64_t maxrank, int32_t transposed, nca_stack** out) {
    return guarded([&] {
        NCA_REQUIRE(ctx && kernel && test && trial && out && nblocks >= 0 && maxrank >= 0, NCA_ERR_ARGUMENT,
                    "nca_level_assemble: invalid argument");
        NCA_REQUIRE(nblocks == 0 || (rowptr && colptr), NCA_ERR_ARGUMENT, "nca_level_assemble: missing CSR");
        set_device(ctx);
        const nca_points* rp = transposed ? trial : test;
        const nca_points* cpnts = transposed ? test : trial;
        auto st = std::make_unique<nca_stack>();
        st->ctx = ctx;
        st->nb = nblocks;
        st->maxrank = maxrank;
        st->rowoffsets.assign(nblocks + 1, 0);
        st->ncols.assign(nblocks, 0);
        for (int64_t b = 0; b < nblocks; ++b) {
            const int64_t m = rowptr[b + 1] - rowptr[b], n = colptr[b + 1] - colptr[b];
            NCA_REQUIRE(m >= 0 && n >= 0, NCA_ERR_ARGUMENT, "nca_level_assemble: decreasing CSR pointer");
            NCA_REQUIRE(n <= maxrank, NCA_ERR_ARGUMENT,
                        "a block has more columns than maxrank=" + std::to_string(maxrank));
            st->rowoffsets[b + 1] = st->rowoffsets[b] + m;
            st->ncols[b] = n;
        }
        const int64_t ld = st->rowoffsets[nblocks];
        st->nrows = ld;
        std::vector<int32_t> rowblock(ld), ridx(ld);
        for (int64_t b = 0; b < nblocks; ++b)
            for (int64_t a = 0; a < st->rowoffsets[b + 1] - st->rowoffsets[b]; ++a) {
                const int64_t pnt = rowidx[rowptr[b] - rowptr[0] + a] - 1;
                NCA_REQUIRE(pnt >= 0 && pnt < rp->n, NCA_ERR_ARGUMENT, "nca_level_assemble: row index out of range");
                rowblock[st->rowoffsets[b] + a] = (int32_t)b;
                ridx[st->rowoffsets[b] + a] = (int32_t)pnt;
            }
        const int64_t ncoltot = nblocks ? colptr[nblocks] - colptr[0] : 0;
        std::vector<int64_t> cptr0(nblocks + 1, 0);
        std::vector<int32_t> cidx(ncoltot);
        for (int64_t b = 0; b <= nblocks; ++b) cptr0[b] = nblocks ? colptr[b] - colptr[0] : 0;
        for (int64_t i = 0; i < ncoltot; ++i) {
            const int64_t pnt = colidx[i] - 1;
            NCA_REQUIRE(pnt >= 0 && pnt < cpnts->n, NCA_ERR_ARGUMENT, "nca_level_assemble: column index out of range");
            cidx[i] = (int32_t)pnt;
        }
        st->buffer.alloc((size_t)(ld * maxrank));
        if (ld * maxrank > 0) {
            DevBuf<int32_t> d_rowblock, d_ridx, d_cidx;
            DevBuf<int64_t> d_cptr;
            d_rowblock.upload(rowblock, ctx->stream);
            d_ridx.upload(ridx, ctx->stream);
            d_cidx.upload(cidx, ctx->stream);
            d_cptr.upload(cptr0, ctx->stream);
            hipLaunchKernelGGL(k_assemble_level, dim3(grid_for(ld * maxrank)), dim3(kWG), 0, ctx->stream,
                               rp->xyz.p, cpnts->xyz.p, d_rowblock.p, d_ridx.p, d_cptr.p, d_cidx.p, ld, maxrank,
                               kernel->wavenumber, (int)kernel->sign, (int)(transposed != 0), st->buffer.p);
            NCA_HIP(hipGetLastError());
            NCA_HIP(hipStreamSynchronize(ctx->stream));
        }
        *out = st.release();
    });
}

int32_t nca_stack_upload(nca_ctx* ctx, const double* buffer, int64_t nblocks, const int64_t* rowoffsets,
                         const int64_t* ncols, int64_t maxrank, nca_stack** out) {
    return guarded([&] {
        NCA_REQUIRE(ctx && out && nblocks >= 0 && maxrank >= 0, NCA_ERR_ARGUMENT, "nca_stack_upload: invalid argument");
        set_device(ctx);
        auto st = std::make_unique<nca_stack>();
        st->ctx = ctx;
        st->nb = nblocks;
        st->maxrank = maxrank;
        st->rowoffsets.assign(rowoffsets, rowoffsets + nblocks + 1);
        st->ncols.assign(ncols, ncols + nblocks);
        NCA_REQUIRE(st->rowoffsets[0] == 0, NCA_ERR_ARGUMENT, "nca_stack_upload: rowoffsets[0] must be 0");
        for (int64_t b = 0; b < nblocks; ++b) {
            NCA_REQUIRE(st->rowoffsets[b + 1] >= st->rowoffsets[b], NCA_ERR_ARGUMENT,
                        "nca_stack_upload: decreasing rowoffsets");
            NCA_REQUIRE(st->ncols[b] >= 0 && st->ncols[b] <= maxrank, NCA_ERR_ARGUMENT,
                        "a block has more columns than maxrank=" + std::to_string(maxrank));
        }
        st->nrows = st->rowoffsets[nblocks];
        st->buffer.upload(reinterpret_cast<const double2*>(buffer), (size_t)(st->nrows * maxrank), ctx->stream);
        NCA_HIP(hipStreamSynchronize(ctx->stream));
        *out = st.release();
    });
}

int32_t nca_stack_shape(const nca_stack* st, int64_t* nblocks, int64_t* nrows, int64_t* maxrank) {
    return guarded([&] {
        NCA_REQUIRE(st, NCA_ERR_ARGUMENT, "nca_stack_shape: null stack");
        if (nblocks) *nblocks = st->nb;
        if (nrows) *nrows = st->nrows;
        if (maxrank) *maxrank = st->maxrank;
    });
}

int32_t nca_stack_download(nca_ctx* ctx, const nca_stack* st, double* buffer) {
    return guarded([&] {
        NCA_REQUIRE(ctx && st && buffer, NCA_ERR_ARGUMENT, "nca_stack_download: invalid argument");
        set_device(ctx);
        const size_t bytes = (size_t)(st->nrows * st->maxrank) * sizeof(double2);
        if (bytes) NCA_HIP(hipMemcpyAsync(buffer, st->buffer.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
        NCA_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int32_t nca_stack_destroy(nca_stack* st) {
    return guarded([&] {
        if (st) {
            if (st->ctx) set_device(st->ctx);
            delete st;
        }
    });
}

}  // extern "C"

namespace nca {

// Device -> host on the context stream, complete on return.
static void d2h(nca_ctx* ctx, void* dst, const void* src, size_t bytes) {
    if (!bytes) return;
    NCA_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    NCA_HIP(hipStreamSynchronize(ctx->stream));
}

// Device-side results of one ACA launch set.
struct AcaResult {
    DevBuf<int32_t> ranks;
    DevBuf<int64_t> rowpiv, colpiv;
};

// iaca: the geometric-column IACA (k_iaca) instead of the fully pivoted ACA.
static void launch_aca(nca_ctx* ctx, const nca_stack* st, double tol, int32_t sampled, AcaResult& res,
                       bool iaca = false) {
    const int64_t nb = st->nb;
    res.ranks.alloc(std::max<int64_t>(nb, 1));
    res.rowpiv.alloc(std::max<int64_t>(nb * st->maxrank, 1));
    res.colpiv.alloc(std::max<int64_t>(nb * st->maxrank, 1));
    if (nb == 0) return;
    std::vector<int32_t> small, big, nc32(nb);
    for (int64_t b = 0; b < nb; ++b) {
        const int64_t m = st->rowoffsets[b + 1] - st->rowoffsets[b];
        const int64_t n = st->ncols[b];
        nc32[b] = (int32_t)n;
        if (m * n <= kLdsCap && m <= kLdsMaxM && n <= kLdsMaxN) {
            small.push_back((int32_t)b);
        } else {
            NCA_REQUIRE(m <= kGlobMaxM && n <= kGlobMaxN, NCA_ERR_LIMIT,
                        "batched ACA: block " + std::to_string(b + 1) + " of " + std::to_string(m) + " x " +
                            std::to_string(n) + " exceeds " + std::to_string(kGlobMaxM) + " x " +
                            std::to_string(kGlobMaxN));
            big.push_back((int32_t)b);
        }
    }
    DevBuf<int64_t> d_rowoff;
    DevBuf<int32_t> d_ncols, d_small, d_big;
    d_rowoff.upload(st->rowoffsets, ctx->stream);
    d_ncols.upload(nc32, ctx->stream);
    DevBuf<double2> work;
    if (!small.empty()) {
        d_small.upload(small, ctx->stream);
        const size_t lds = sizeof(double2) * (kLdsCap + kLdsMaxM + kLdsMaxN);
        if (iaca)
            hipLaunchKernelGGL(k_iaca<true>, dim3((unsigned)small.size()), dim3(kWG), lds, ctx->stream, st->buffer.p,
                               nullptr, st->nrows, d_rowoff.p, d_ncols.p, d_small.p, st->maxrank, tol, res.ranks.p,
                               res.rowpiv.p, res.colpiv.p);
        else
            hipLaunchKernelGGL(k_aca<true>, dim3((unsigned)small.size()), dim3(kWG), lds, ctx->stream, st->buffer.p,
                               nullptr, st->nrows, d_rowoff.p, d_ncols.p, d_small.p, st->maxrank, tol, (int)sampled,
                               res.ranks.p, res.rowpiv.p, res.colpiv.p);
        NCA_HIP(hipGetLastError());
    }
    if (!big.empty()) {
        d_big.upload(big, ctx->stream);
        work.alloc((size_t)(st->nrows * st->maxrank));
        const size_t lds = sizeof(double2) * (kGlobMaxM + kGlobMaxN);
        if (iaca)
            hipLaunchKernelGGL(k_iaca<false>, dim3((unsigned)big.size()), dim3(kWG), lds, ctx->stream, st->buffer.p,
                               work.p, st->nrows, d_rowoff.p, d_ncols.p, d_big.p, st->maxrank, tol, res.ranks.p,
                               res.rowpiv.p, res.colpiv.p);
        else
            hipLaunchKernelGGL(k_aca<false>, dim3((unsigned)big.size()), dim3(kWG), lds, ctx->stream, st->buffer.p,
                               work.p, st->nrows, d_rowoff.p, d_ncols.p, d_big.p, st->maxrank, tol, (int)sampled,
                               res.ranks.p, res.rowpiv.p, res.colpiv.p);
        NCA_HIP(hipGetLastError());
    }
    NCA_HIP(hipStreamSynchronize(ctx->stream));
}

static void launch_skeleton(nca_ctx* ctx, const nca_stack* st, const int32_t* ranks_h, const DevBuf<int32_t>& ranks,
                            const DevBuf<int64_t>& rowpiv, const DevBuf<int64_t>& colpiv,
                            const std::vector<int64_t>& xoff, DevBuf<double2>& X, DevBuf<int64_t>& indep) {
    const int64_t nb = st->nb;
    indep.alloc(std::max<int64_t>(nb, 1));
    X.alloc(std::max<int64_t>(xoff[nb], 1));
    if (nb == 0) return;
    int rmax = 0;
    for (int64_t b = 0; b < nb; ++b) rmax = std::max(rmax, (int)ranks_h[b]);
    NCA_REQUIRE(rmax <= kSolveMaxRank, NCA_ERR_LIMIT,
                "skeleton solve: rank " + std::to_string(rmax) + " exceeds " + std::to_string(kSolveMaxRank));
    DevBuf<int64_t> d_rowoff, d_xoff;
    d_rowoff.upload(st->rowoffsets, ctx->stream);
    d_xoff.upload(xoff, ctx->stream);
    const size_t lds = sizeof(double2) * (size_t)std::max(1, rmax * rmax);
    if (lds > 32768)
        NCA_HIP(hipFuncSetAttribute((const void*)k_skeleton, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_skeleton, dim3((unsigned)nb), dim3(kWG), lds, ctx->stream, st->buffer.p, st->nrows,
                       d_rowoff.p, ranks.p, rowpiv.p, colpiv.p, st->maxrank, d_xoff.p, X.p, indep.p);
    NCA_HIP(hipGetLastError());
    NCA_HIP(hipStreamSynchronize(ctx->stream));
}

static void singular_message(const nca_stack* st, const int32_t* ranks, const std::vector<int64_t>& indep) {
    for (int64_t b = 0; b < st->nb; ++b)
        if (indep[b] >= 0)
            throw Error(NCA_ERR_SINGULAR, "SingularCrossMatrixError: block " + std::to_string(b + 1) + ": " +
                                              std::to_string(ranks[b]) + " pivots were selected but only " +
                                              std::to_string(indep[b]) + " of them are linearly independent");
}


// ---------------------------------------------------------------------------
// Phase 1 on the device (the far-pivot selection of host/nca_tree.cpp,
// select_farpivots, operation for operation; bit-identical for weights 1 and
// 1.5, whose power d^w is d and d sqrt(d) on both sides, correctly rounded).
// One workgroup per item: candidates = an even subsample of each far node's
// points (per point count doubled until the pool holds the item's n pivots),
// the first pivot is the candidate nearest to the item's center, every next
// one maximizes d_min^2 / d_c^(2 weight); ties go to the first candidate.
constexpr int kP1Threads = 256;
constexpr int kP1PerThread = 16;
constexpr int kP1Pool = kP1Threads * kP1PerThread;
constexpr int kP1MaxFar = 4096;

__device__ __forceinline__ double p1_dist2(const double* a, double b0, double b1, double b2) {
    const double dx = a[0] - b0, dy = a[1] - b1, dz = a[2] - b2;
    return dx * dx + dy * dy + dz * dz;
}

// block-wide reductions over kP1Threads threads
__device__ __forceinline__ int64_t p1_sum(int64_t v, int64_t* red) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    const int tid = threadIdx.x;
    __syncthreads();
    if ((tid & 63) == 0) red[tid >> 6] = v;
    __syncthreads();
    return red[0] + red[1] + red[2] + red[3];
}

// best (value, index) pair: larger value (or smaller if MIN), then smaller index
template <bool MIN>
__device__ __forceinline__ void p1_best(double& v, int& q, double* redv, int* redq) {
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o);
        const int oq = __shfl_xor(q, o);
        const bool take = oq >= 0 && (q < 0 || (MIN ? ov < v : ov > v) || (ov == v && oq < q));
        if (take) { v = ov; q = oq; }
    }
    const int tid = threadIdx.x;
    __syncthreads();
    if ((tid & 63) == 0) { redv[tid >> 6] = v; redq[tid >> 6] = q; }
    __syncthreads();
    v = redv[0];
    q = redq[0];
    for (int w = 1; w < 4; ++w) {
        const double ov = redv[w];
        const int oq = redq[w];
        const bool take = oq >= 0 && (q < 0 || (MIN ? ov < v : ov > v) || (ov == v && oq < q));
        if (take) { v = ov; q = oq; }
    }
}

// TREE: the item's far field is built here, in the order of farfield
// (farinteractions.jl:38-60): the far nodes of the item, then those of its
// parent, and so on to the root (ffptr / ffnodes: the far lists by node,
// parent: -1 at a root); else ffptr / ffnodes hold the far field of every
// item. A candidate pool above kP1Pool sets *err (the host checked it up
// front for explicit far fields).
template <bool TREE>
__global__ void __launch_bounds__(kP1Threads)
    k_farpivots(const double* __restrict__ pts, const int32_t* __restrict__ olo, const int32_t* __restrict__ ohi,
                const int32_t* __restrict__ operm, const double* __restrict__ centers,
                const int64_t* __restrict__ ffptr, const int32_t* __restrict__ ffnodes,
                const int32_t* __restrict__ parent, const int64_t* __restrict__ pivptr, double weight,
                int64_t* __restrict__ piv, int32_t* __restrict__ err, const int32_t* __restrict__ items) {
    __shared__ int32_t scan[kP1MaxFar + 1];
    __shared__ int32_t fs[TREE ? kP1MaxFar : 1];
    __shared__ int64_t red[4];
    __shared__ double redv[4];
    __shared__ int redq[4];
    __shared__ double pc[3];
    __shared__ int32_t pcur;
    const int64_t it = items ? (int64_t)items[blockIdx.x] : (int64_t)blockIdx.x;
    const int tid = threadIdx.x;
    const int64_t n = pivptr[it + 1] - pivptr[it];
    if (n == 0) return;
    const int32_t* F;
    int nf;
    if constexpr (TREE) {
        int m = 0;
        for (int64_t a = it; a >= 0; a = parent[a]) {
            const int64_t b = ffptr[a], e = ffptr[a + 1];
            for (int64_t k = tid; k < e - b; k += kP1Threads) fs[m + k] = ffnodes[b + k];
            m += (int)(e - b);
        }
        __syncthreads();
        F = fs;
        nf = m;
    } else {
        F = ffnodes + ffptr[it];
        nf = (int)(ffptr[it + 1] - ffptr[it]);
    }
    auto sz = [&](int f) { return (int64_t)(ohi[F[f]] - olo[F[f]]); };
    int64_t tot = 0;
    for (int f = tid; f < nf; f += kP1Threads) tot += sz(f);
    const int64_t total = p1_sum(tot, red);
    int64_t per = max((int64_t)1, (max(8 * n, (int64_t)512) + nf - 1) / nf);
    int64_t pool = 0;
    for (;;) {
        int64_t c = 0;
        for (int f = tid; f < nf; f += kP1Threads) c += min(sz(f), per);
        pool = p1_sum(c, red);
        if (pool >= n || pool >= total) break;
        per *= 2;
    }
    if (pool > kP1Pool) {
        if (tid == 0) atomicOr(err, 1);
        return;
    }
    // exclusive scan of the per-node candidate counts (contiguous chunks per thread)
    const int chunk = (nf + kP1Threads - 1) / kP1Threads;
    const int fa = min(nf, tid * chunk), fb = min(nf, fa + chunk);
    int64_t loc = 0;
    for (int f = fa; f < fb; ++f) loc += min(sz(f), per);
    __shared__ int64_t part[kP1Threads];
    part[tid] = loc;
    __syncthreads();
    if (tid == 0) {
        int64_t run = 0;
        for (int t = 0; t < kP1Threads; ++t) {
            const int64_t v = part[t];
            part[t] = run;
            run += v;
        }
    }
    __syncthreads();
    {
        int64_t run = part[tid];
        for (int f = fa; f < fb; ++f) {
            scan[f] = (int32_t)run;
            run += min(sz(f), per);
        }
        if (fb == nf && fa < fb) scan[nf] = (int32_t)run;
        if (nf == 0 && tid == 0) scan[0] = 0;
    }
    __syncthreads();
    const double* ct = centers + 3 * it;
    const double c0 = ct[0], c1 = ct[1], c2 = ct[2];
    double px[kP1PerThread], py[kP1PerThread], pz[kP1PerThread], d2c[kP1PerThread], mind[kP1PerThread];
    int32_t pid[kP1PerThread];
    double bv = 0.0;
    int bq = -1;
#pragma unroll
    for (int m = 0; m < kP1PerThread; ++m) {
        const int q = tid + kP1Threads * m;
        pid[m] = -1;
        mind[m] = -1.0;
        if (q < pool) {
            // far node of candidate q: last f with scan[f] <= q
            int lo = 0, hi = nf;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (scan[mid] <= q) lo = mid; else hi = mid;
            }
            const int64_t s = sz(lo), k = min(s, per), j = q - scan[lo];
            const int32_t p = operm[olo[F[lo]] + (int32_t)((j * s) / k)];
            pid[m] = p;
            px[m] = pts[3 * (int64_t)p];
            py[m] = pts[3 * (int64_t)p + 1];
            pz[m] = pts[3 * (int64_t)p + 2];
            const double pp[3] = {px[m], py[m], pz[m]};
            d2c[m] = max(p1_dist2(pp, c0, c1, c2), 1e-300);
            mind[m] = __longlong_as_double(0x7ff0000000000000LL);  // +inf
            if (bq < 0 || d2c[m] < bv) { bv = d2c[m]; bq = q; }
        }
    }
    p1_best<true>(bv, bq, redv, redq);
    int cur = bq;
    if (weight == 1.5) {
#pragma unroll
        for (int m = 0; m < kP1PerThread; ++m)
            if (pid[m] >= 0) d2c[m] = d2c[m] * sqrt(d2c[m]);
    } else if (weight != 1.0) {
#pragma unroll
        for (int m = 0; m < kP1PerThread; ++m)
            if (pid[m] >= 0) d2c[m] = pow(d2c[m], weight);
    }
    int64_t* out = piv + pivptr[it];
    for (int64_t sel = 0; sel < n; ++sel) {
        if (cur % kP1Threads == tid) {
            const int m = cur / kP1Threads;
#pragma unroll
            for (int mm = 0; mm < kP1PerThread; ++mm)
                if (mm == m) {
                    mind[mm] = -1.0;
                    pc[0] = px[mm];
                    pc[1] = py[mm];
                    pc[2] = pz[mm];
                    pcur = pid[mm];
                }
        }
        __syncthreads();
        if (tid == 0) out[sel] = pcur;
        const double a0 = pc[0], a1 = pc[1], a2 = pc[2];
        double best = -1.0;
        int arg = -1;
#pragma unroll
        for (int m = 0; m < kP1PerThread; ++m) {
            if (pid[m] < 0 || mind[m] < 0.0) continue;
            const double pp[3] = {px[m], py[m], pz[m]};
            const double d2 = p1_dist2(pp, a0, a1, a2);
            if (d2 < mind[m]) mind[m] = d2;
            const double score = mind[m] / d2c[m];
            if (score > best) { best = score; arg = tid + kP1Threads * m; }
        }
        p1_best<false>(best, arg, redv, redq);
        if (arg < 0) break;
        cur = arg;
    }
}

// The same selection with one wave per item, for explicit far fields whose
// candidate pool fits kP1WavePer candidates per lane (candidate q in lane
// q % 64, slot q / 64): no workgroup barriers, the chosen candidate's point
// read by scalar loads. The host gives per item (items[w]) the per-node
// count per and the pool size, and fscan, the exclusive scan of
// min(points, per) over each item's far nodes (aligned with ffnodes). Same
// candidates, same order of comparisons, so the same pivots as k_farpivots.
constexpr int kP1WavePer = 16;
constexpr int kP1WavePool = 64 * kP1WavePer;

// best (value, index) over the wave: larger value (or smaller if MIN), then smaller index
template <bool MIN>
__device__ __forceinline__ int p1_wave_best(double v, int q) {
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(v, o);
        const int oq = __shfl_xor(q, o);
        const bool take = oq >= 0 && (q < 0 || (MIN ? ov < v : ov > v) || (ov == v && oq < q));
        if (take) { v = ov; q = oq; }
    }
    return __builtin_amdgcn_readfirstlane(q);
}

__global__ void __launch_bounds__(256)
    k_farpivots_wave(const double* __restrict__ pts, const int32_t* __restrict__ olo, const int32_t* __restrict__ ohi,
                     const int32_t* __restrict__ operm, const double* __restrict__ centers,
                     const int64_t* __restrict__ ffptr, const int32_t* __restrict__ ffnodes,
                     const int32_t* __restrict__ fscan, const int32_t* __restrict__ items, int64_t nwork,
                     const int32_t* __restrict__ iper, const int32_t* __restrict__ ipool,
                     const int64_t* __restrict__ pivptr, double weight, int64_t* __restrict__ piv) {
    __shared__ int32_t cand[4][kP1WavePool];
    const int wv = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
    const int64_t w = (int64_t)blockIdx.x * 4 + wv;
    if (w >= nwork) return;
    const int64_t it = items[w];
    const int64_t n = pivptr[it + 1] - pivptr[it];
    const int32_t* F = ffnodes + ffptr[it];
    const int32_t* scan = fscan + ffptr[it];
    const int nf = (int)(ffptr[it + 1] - ffptr[it]);
    const int64_t per = iper[w];
    const int pool = ipool[w];
    const double* ct = centers + 3 * it;
    const double c0 = ct[0], c1 = ct[1], c2 = ct[2];
    double px[kP1WavePer], py[kP1WavePer], pz[kP1WavePer], d2c[kP1WavePer], mind[kP1WavePer];
    double bv = 0.0;
    int bq = -1;
#pragma unroll
    for (int m = 0; m < kP1WavePer; ++m) {
        const int q = lane + 64 * m;
        mind[m] = -1.0;
        if (q < pool) {
            // far node of candidate q: last f with scan[f] <= q
            int lo = 0, hi = nf;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (scan[mid] <= q) lo = mid; else hi = mid;
            }
            const int32_t f = F[lo];
            const int64_t sz = (int64_t)(ohi[f] - olo[f]), k = min(sz, per), j = q - scan[lo];
            const int32_t p = operm[olo[f] + (int32_t)((j * sz) / k)];
            cand[wv][q] = p;
            px[m] = pts[3 * (int64_t)p];
            py[m] = pts[3 * (int64_t)p + 1];
            pz[m] = pts[3 * (int64_t)p + 2];
            const double pp[3] = {px[m], py[m], pz[m]};
            d2c[m] = max(p1_dist2(pp, c0, c1, c2), 1e-300);
            mind[m] = __longlong_as_double(0x7ff0000000000000LL);  // +inf
            if (bq < 0 || d2c[m] < bv) { bv = d2c[m]; bq = q; }
        }
    }
    int cur = p1_wave_best<true>(bv, bq);
    if (weight == 1.5) {
#pragma unroll
        for (int m = 0; m < kP1WavePer; ++m)
            if (mind[m] >= 0.0) d2c[m] = d2c[m] * sqrt(d2c[m]);
    } else if (weight != 1.0) {
#pragma unroll
        for (int m = 0; m < kP1WavePer; ++m)
            if (mind[m] >= 0.0) d2c[m] = pow(d2c[m], weight);
    }
    __builtin_amdgcn_wave_barrier();
    int64_t* out = piv + pivptr[it];
    for (int64_t sel = 0; sel < n; ++sel) {
        const int cl = cur & 63, cs = cur >> 6;
        if (lane == cl) {
#pragma unroll
            for (int m = 0; m < kP1WavePer; ++m)
                if (m == cs) mind[m] = -1.0;
        }
        const int32_t pc = __builtin_amdgcn_readfirstlane(cand[wv][cur]);
        if (lane == 0) out[sel] = pc;
        const double a0 = pts[3 * (int64_t)pc], a1 = pts[3 * (int64_t)pc + 1], a2 = pts[3 * (int64_t)pc + 2];
        double best = -1.0;
        int arg = -1;
#pragma unroll
        for (int m = 0; m < kP1WavePer; ++m) {
            if (mind[m] < 0.0) continue;
            const double pp[3] = {px[m], py[m], pz[m]};
            const double d2 = p1_dist2(pp, a0, a1, a2);
            if (d2 < mind[m]) mind[m] = d2;
            const double score = mind[m] / d2c[m];
            if (score > best) { best = score; arg = lane + 64 * m; }
        }
        arg = p1_wave_best<false>(best, arg);
        if (arg < 0) break;
        cur = arg;
    }
}

// Host loop over [0, n) on up to 16 threads (dynamic chunks); f must not throw.
template <class F>
static void host_parallel_for(int64_t n, F&& f) {
    const int64_t nt = std::min<int64_t>(std::max<unsigned>(1u, std::thread::hardware_concurrency()), 16);
    if (n < 4096 || nt == 1) {
        for (int64_t i = 0; i < n; ++i) f(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::vector<std::thread> pool;
    for (int64_t k = 0; k < nt; ++k)
        pool.emplace_back([&] {
            for (;;) {
                const int64_t b = next.fetch_add(256);
                if (b >= n) break;
                for (int64_t i = b; i < std::min(n, b + 256); ++i) f(i);
            }
        });
    for (auto& t : pool) t.join();
}

}  // namespace nca

extern "C" {

int32_t nca_aca_fullpiv(nca_ctx* ctx, const nca_stack* st, double tol, int32_t sampled, int32_t* ranks,
                        int64_t* rowpivots, int64_t* colpivots) {
    return guarded([&] {
        NCA_REQUIRE(ctx && st && ranks && rowpivots && colpivots, NCA_ERR_ARGUMENT, "nca_aca_fullpiv: invalid argument");
        set_device(ctx);
        AcaResult res;
        launch_aca(ctx, st, tol, sampled, res);
        if (st->nb == 0) return;
        d2h(ctx, ranks, res.ranks.p, sizeof(int32_t) * st->nb);
        d2h(ctx, rowpivots, res.rowpiv.p, sizeof(int64_t) * st->nb * st->maxrank);
        d2h(ctx, colpivots, res.colpiv.p, sizeof(int64_t) * st->nb * st->maxrank);
    });
}

int32_t nca_skeleton_solve(nca_ctx* ctx, const nca_stack* st, const int32_t* ranks, const int64_t* rowpivots,
                           const int64_t* colpivots, const int64_t* xoffsets, double* X, int64_t* independent) {
    return guarded([&] {
        NCA_REQUIRE(ctx && st && ranks && rowpivots && colpivots && xoffsets && independent, NCA_ERR_ARGUMENT,
                    "nca_skeleton_solve: invalid argument");
        set_device(ctx);
        const int64_t nb = st->nb;
        for (int64_t b = 0; b < nb; ++b) {
            const int64_t m = st->rowoffsets[b + 1] - st->rowoffsets[b];
            NCA_REQUIRE(ranks[b] >= 0 && ranks[b] <= std::min<int64_t>(m, st->ncols[b]), NCA_ERR_ARGUMENT,
                        "nca_skeleton_solve: rank out of range");
            for (int64_t i = 0; i < ranks[b]; ++i) {
                const int64_t rp = rowpivots[i + b * st->maxrank], cp = colpivots[i + b * st->maxrank];
                NCA_REQUIRE(rp >= 1 && rp <= m && cp >= 1 && cp <= st->ncols[b], NCA_ERR_ARGUMENT,
                            "nca_skeleton_solve: pivot out of range");
            }
            NCA_REQUIRE(xoffsets[b + 1] - xoffsets[b] >= m * ranks[b], NCA_ERR_DIMENSION,
                        "nca_skeleton_solve: xoffsets leave too little room");
        }
        std::vector<int64_t> xoff(xoffsets, xoffsets + nb + 1);
        DevBuf<int32_t> d_ranks;
        DevBuf<int64_t> d_rp, d_cp, d_indep;
        DevBuf<double2> d_X;
        d_ranks.upload(ranks, nb, ctx->stream);
        d_rp.upload(rowpivots, nb * st->maxrank, ctx->stream);
        d_cp.upload(colpivots, nb * st->maxrank, ctx->stream);
        launch_skeleton(ctx, st, ranks, d_ranks, d_rp, d_cp, xoff, d_X, d_indep);
        if (nb == 0) return;
        std::vector<int64_t> indep(nb);
        d2h(ctx, indep.data(), d_indep.p, sizeof(int64_t) * nb);
        if (xoff[nb] > 0) d2h(ctx, X, d_X.p, sizeof(double2) * xoff[nb]);
        std::copy(indep.begin(), indep.end(), independent);
        singular_message(st, ranks, indep);
    });
}

}  // extern "C"

namespace nca {
static void compress_level_impl(nca_ctx* ctx, const nca_stack* st, double tol, int32_t sampled, bool iaca,
                                int32_t* ranks, int64_t* rowpivots, int64_t* colpivots, int64_t xcapacity,
                                int64_t* xoffsets, double* X, int64_t* independent) {
        NCA_REQUIRE(ctx && st && ranks && rowpivots && colpivots && xoffsets && independent, NCA_ERR_ARGUMENT,
                    "nca_compress_level: invalid argument");
        set_device(ctx);
        const int64_t nb = st->nb;
        AcaResult res;
        launch_aca(ctx, st, tol, sampled, res, iaca);
        xoffsets[0] = 0;
        if (nb == 0) return;
        d2h(ctx, ranks, res.ranks.p, sizeof(int32_t) * nb);
        std::vector<int64_t> xoff(nb + 1, 0);
        for (int64_t b = 0; b < nb; ++b)
            xoff[b + 1] = xoff[b] + (st->rowoffsets[b + 1] - st->rowoffsets[b]) * (int64_t)ranks[b];
        NCA_REQUIRE(xoff[nb] <= xcapacity, NCA_ERR_DIMENSION,
                    "nca_compress_level: X holds " + std::to_string(xcapacity) + " entries, " +
                        std::to_string(xoff[nb]) + " needed");
        DevBuf<double2> d_X;
        DevBuf<int64_t> d_indep;
        launch_skeleton(ctx, st, ranks, res.ranks, res.rowpiv, res.colpiv, xoff, d_X, d_indep);
        d2h(ctx, rowpivots, res.rowpiv.p, sizeof(int64_t) * nb * st->maxrank);
        d2h(ctx, colpivots, res.colpiv.p, sizeof(int64_t) * nb * st->maxrank);
        std::vector<int64_t> indep(nb);
        d2h(ctx, indep.data(), d_indep.p, sizeof(int64_t) * nb);
        if (xoff[nb] > 0) d2h(ctx, X, d_X.p, sizeof(double2) * xoff[nb]);
        std::copy(xoff.begin(), xoff.end(), xoffsets);
        std::copy(indep.begin(), indep.end(), independent);
        singular_message(st, ranks, indep);
}

// compress_level_impl keeping X on the device: the level's factors become
// one chunk of f, block b of the stack the factor of ids[b] (1-based).
static void compress_level_keep(nca_ctx* ctx, const nca_stack* st, double tol, int32_t sampled, bool iaca,
                                nca_factors* f, const int64_t* ids, int32_t* ranks, int64_t* rowpivots,
                                int64_t* colpivots, int64_t* independent) {
    NCA_REQUIRE(ctx && st && f && ranks && rowpivots && colpivots && independent && (st->nb == 0 || ids),
                NCA_ERR_ARGUMENT, "nca_compress_level_resident: invalid argument");
    NCA_REQUIRE(f->ctx == ctx, NCA_ERR_ARGUMENT, "nca_compress_level_resident: factors of another context");
    const int64_t nb = st->nb;
    {
        std::vector<char> seen;
        for (int64_t b = 0; b < nb; ++b) {
            const int64_t id = ids[b] - 1;
            NCA_REQUIRE(id >= 0 && id < f->nids, NCA_ERR_ARGUMENT,
                        "nca_compress_level_resident: block id " + std::to_string(ids[b]) + " out of range");
            NCA_REQUIRE(f->chunk[id] < 0, NCA_ERR_ARGUMENT,
                        "nca_compress_level_resident: block id " + std::to_string(ids[b]) + " has a factor already");
            f->chunk[id] = -2;  // taken by this call (duplicates fail here)
        }
    }
    auto undo = [&] {
        for (int64_t b = 0; b < nb; ++b)
            if (f->chunk[ids[b] - 1] == -2) f->chunk[ids[b] - 1] = -1;
    };
    try {
        set_device(ctx);
        AcaResult res;
        launch_aca(ctx, st, tol, sampled, res, iaca);
        if (nb == 0) return;
        d2h(ctx, ranks, res.ranks.p, sizeof(int32_t) * nb);
        std::vector<int64_t> xoff(nb + 1, 0);
        for (int64_t b = 0; b < nb; ++b)
            xoff[b + 1] = xoff[b] + (st->rowoffsets[b + 1] - st->rowoffsets[b]) * (int64_t)ranks[b];
        DevBuf<double2> d_X;
        DevBuf<int64_t> d_indep;
        launch_skeleton(ctx, st, ranks, res.ranks, res.rowpiv, res.colpiv, xoff, d_X, d_indep);
        d2h(ctx, rowpivots, res.rowpiv.p, sizeof(int64_t) * nb * st->maxrank);
        d2h(ctx, colpivots, res.colpiv.p, sizeof(int64_t) * nb * st->maxrank);
        std::vector<int64_t> indep(nb);
        d2h(ctx, indep.data(), d_indep.p, sizeof(int64_t) * nb);
        const int32_t c = (int32_t)f->chunks.size();
        f->chunks.push_back(std::move(d_X));
        f->chunk_size.push_back(xoff[nb]);
        f->chunk_ids.emplace_back();
        for (int64_t b = 0; b < nb; ++b) {
            const int64_t id = ids[b] - 1;
            f->chunk[id] = c;
            f->off[id] = xoff[b];
            f->m[id] = (int32_t)(st->rowoffsets[b + 1] - st->rowoffsets[b]);
            f->rank[id] = ranks[b];
        }
        std::copy(indep.begin(), indep.end(), independent);
        singular_message(st, ranks, indep);
    } catch (...) {
        undo();
        throw;
    }
}

// Block copies, one 64-lane wave per job.
__global__ void __launch_bounds__(256) k_pack_blocks(const PackJob* __restrict__ jobs, int64_t njobs,
                                                     double2* __restrict__ dst) {
    const int64_t j = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (j >= njobs) return;
    const PackJob J = jobs[j];
    const int lane = threadIdx.x & 63;
    const int64_t total = (int64_t)J.rows * J.cols;
    for (int64_t q = lane; q < total; q += 64) {
        const int64_t a = q % J.rows, k = q / J.rows;
        dst[J.dst + a * J.sa + k * J.sk] = J.src[a + k * J.ld];
    }
}

void pack_blocks(const std::vector<PackJob>& jobs, double2* dst, hipStream_t st) {
    if (jobs.empty()) return;
    DevBuf<PackJob> dj;
    dj.upload(jobs, st);
    const int64_t nj = (int64_t)jobs.size();
    hipLaunchKernelGGL(k_pack_blocks, dim3((unsigned)((nj + 3) / 4)), dim3(256), 0, st, dj.p, nj, dst);
    NCA_HIP(hipGetLastError());
    NCA_HIP(hipStreamSynchronize(st));
}
}  // namespace nca

extern "C" {

int32_t nca_factors_create(nca_ctx* ctx, int64_t nids, nca_factors** out) {
    return guarded([&] {
        NCA_REQUIRE(ctx && out && nids >= 0, NCA_ERR_ARGUMENT, "nca_factors_create: invalid argument");
        auto f = std::make_unique<nca_factors>();
        f->ctx = ctx;
        f->nids = nids;
        f->chunk.assign(nids, -1);
        f->m.assign(nids, 0);
        f->rank.assign(nids, 0);
        f->off.assign(nids, 0);
        *out = f.release();
    });
}

int32_t nca_factors_destroy(nca_factors* f) {
    return guarded([&] {
        if (!f) return;
        if (f->ctx) set_device(f->ctx);
        delete f;
    });
}

int32_t nca_factors_size(const nca_factors* f, int64_t* nentries) {
    return guarded([&] {
        NCA_REQUIRE(f && nentries, NCA_ERR_ARGUMENT, "nca_factors_size: invalid argument");
        int64_t n = 0;
        for (int64_t c : f->chunk_size) n += c;
        *nentries = n;
    });
}

int32_t nca_factors_download(nca_ctx* ctx, const nca_factors* f, double* X) {
    return guarded([&] {
        NCA_REQUIRE(ctx && f && X, NCA_ERR_ARGUMENT, "nca_factors_download: invalid argument");
        set_device(ctx);
        double2* dst = reinterpret_cast<double2*>(X);
        for (size_t c = 0; c < f->chunks.size(); ++c) {
            if (f->chunk_size[c] > 0) {
                if (f->chunk_ids[c].empty()) {
                    NCA_HIP(hipMemcpyAsync(dst, f->chunks[c].p, sizeof(double2) * f->chunk_size[c],
                                           hipMemcpyDeviceToHost, ctx->stream));
                } else {
                    // packed first, in the order of the chunk's ids
                    std::vector<PackJob> jobs;
                    int64_t o = 0;
                    for (int32_t id : f->chunk_ids[c]) {
                        const FactorRef F = factor_of(f, id);
                        jobs.push_back({F.p, o, (int32_t)F.m, (int32_t)F.m, (int32_t)F.r, 1, (int32_t)F.m, 0});
                        o += F.m * F.r;
                    }
                    DevBuf<double2> tmp((size_t)f->chunk_size[c]);
                    pack_blocks(jobs, tmp.p, ctx->stream);
                    NCA_HIP(hipMemcpyAsync(dst, tmp.p, sizeof(double2) * f->chunk_size[c], hipMemcpyDeviceToHost,
                                           ctx->stream));
                    NCA_HIP(hipStreamSynchronize(ctx->stream));
                }
            }
            dst += f->chunk_size[c];
        }
        NCA_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int32_t nca_bottomup_side(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* self,
                          const nca_points* other, int32_t transposed, int64_t nids, const int64_t* childptr,
                          const int64_t* children, const int64_t* lo, const int64_t* hi, const int64_t* perm,
                          int64_t nlevels, const int64_t* levptr, const int64_t* levids, const int64_t* fpptr,
                          const int64_t* fpidx, int64_t maxrank, double tol, int32_t sampled, int32_t method,
                          nca_factors* f, int32_t* ranks, int64_t* mrows, int64_t pivcap, int64_t* pivots,
                          int64_t* independent) {
    return guarded([&] {
        NCA_REQUIRE(ctx && kernel && self && other && f && childptr && lo && hi && levptr && fpptr && ranks &&
                        mrows && pivots && independent && nids >= 0 && nlevels >= 0,
                    NCA_ERR_ARGUMENT, "nca_bottomup_side: invalid argument");
        NCA_REQUIRE(method == 0 || method == 1, NCA_ERR_ARGUMENT, "nca_bottomup_side: unknown method");
        NCA_REQUIRE(f->ctx == ctx && f->nids == nids, NCA_ERR_ARGUMENT, "nca_bottomup_side: factors of another tree");
        // the r x r cross matrix shares the dynamic LDS: maxrank <= kSolveMaxRank keeps
        // it and the kernel's static LDS within the 160 KiB of a CU
        NCA_REQUIRE(maxrank >= 0 && maxrank <= kSolveMaxRank, NCA_ERR_LIMIT,
                    "nca_bottomup_side: maxrank above " + std::to_string(kSolveMaxRank));
        NCA_REQUIRE(nids < (1LL << 31) && self->n < (1LL << 31) && other->n < (1LL << 31), NCA_ERR_LIMIT,
                    "nca_bottomup_side: more than 2^31 blocks or points");
        set_device(ctx);
        hipStream_t st = ctx->stream;
        // tree (0-based, 32-bit)
        std::vector<int32_t> cp(nids + 1), lo0(nids), hi0(nids);
        for (int64_t i = 0; i <= nids; ++i) cp[i] = (int32_t)(childptr[i] - childptr[0]);
        const int64_t nch = cp[nids];
        std::vector<int32_t> ch(std::max<int64_t>(nch, 1)), pm(std::max<int64_t>(self->n, 1));
        for (int64_t q = 0; q < nch; ++q) {
            NCA_REQUIRE(children[q] >= 1 && children[q] <= nids, NCA_ERR_ARGUMENT, "nca_bottomup_side: child out of range");
            ch[q] = (int32_t)(children[q] - 1);
        }
        for (int64_t i = 0; i < nids; ++i) {
            lo0[i] = (int32_t)(lo[i] - 1);
            hi0[i] = (int32_t)hi[i];
            NCA_REQUIRE(lo0[i] >= 0 && hi0[i] >= lo0[i] && hi0[i] <= self->n, NCA_ERR_ARGUMENT,
                        "nca_bottomup_side: value range out of bounds");
        }
        for (int64_t q = 0; q < self->n; ++q) {
            NCA_REQUIRE(perm[q] >= 1 && perm[q] <= self->n, NCA_ERR_ARGUMENT, "nca_bottomup_side: perm out of range");
            pm[q] = (int32_t)(perm[q] - 1);
        }
        // blocks of every level, their bounds (rows: a leaf's values, else the
        // children's rank bounds min(maxrank, rows, columns)), factor and work offsets
        const int64_t nblk = nlevels ? levptr[nlevels] - levptr[0] : 0;
        std::vector<int32_t> bid(std::max<int64_t>(nblk, 1)), bn(std::max<int64_t>(nblk, 1));
        std::vector<int64_t> bcol(std::max<int64_t>(nblk, 1)), bx(std::max<int64_t>(nblk, 1)),
            bw(std::max<int64_t>(nblk, 1));
        std::vector<int64_t> rb(nids, 0);
        std::vector<char> seen(nids, 0);
        std::vector<int64_t> lvwork(nlevels, 0);
        int64_t xtot = 0, worktot = 0;
        for (int64_t L = 0; L < nlevels; ++L) {
            int64_t wl = 0;
            for (int64_t q = levptr[L] - levptr[0]; q < levptr[L + 1] - levptr[0]; ++q) {
                const int64_t id = levids[q] - 1;
                NCA_REQUIRE(id >= 0 && id < nids && !seen[id] && f->chunk[id] < 0, NCA_ERR_ARGUMENT,
                            "nca_bottomup_side: block id out of range, repeated or with a factor already");
                seen[id] = 1;
                int64_t mcap = 0;
                if (cp[id] == cp[id + 1]) mcap = hi0[id] - lo0[id];
                else
                    for (int32_t k = cp[id]; k < cp[id + 1]; ++k) mcap += rb[ch[k]];
                const int64_t n = fpptr[id + 1] - fpptr[id];
                NCA_REQUIRE(n >= 0 && n <= maxrank, NCA_ERR_ARGUMENT,
                            "a block has more columns than maxrank=" + std::to_string(maxrank));
                NCA_REQUIRE(n == 0 || mcap <= kGlobMaxM, NCA_ERR_LIMIT,
                            "nca_bottomup_side: block " + std::to_string(id + 1) + " may have " +
                                std::to_string(mcap) + " rows, above " + std::to_string(kGlobMaxM));
                rb[id] = std::min<int64_t>(std::min<int64_t>(maxrank, mcap), n);
                bid[q] = (int32_t)id;
                bn[q] = (int32_t)n;
                bcol[q] = fpptr[id] - fpptr[0];
                bx[q] = xtot;
                xtot += mcap * rb[id];
                const bool lds = mcap * n <= kLdsCap && mcap <= kLdsMaxM && n <= kLdsMaxN;
                bw[q] = lds ? -1 : wl;
                if (!lds) wl += mcap * n;
            }
            worktot = std::max(worktot, wl);
        }
        const int64_t ncolidx = fpptr[nids] - fpptr[0];
        std::vector<int32_t> ci(std::max<int64_t>(ncolidx, 1));
        for (int64_t q = 0; q < ncolidx; ++q) {
            NCA_REQUIRE(fpidx[q] >= 1 && fpidx[q] <= other->n, NCA_ERR_ARGUMENT,
                        "nca_bottomup_side: far pivot out of range");
            ci[q] = (int32_t)(fpidx[q] - 1);
        }
        DevBuf<int32_t> d_cp, d_ch, d_lo, d_hi, d_pm, d_ci, d_bid, d_bn, d_piv, d_rank, d_m, d_err;
        DevBuf<int64_t> d_bcol, d_bx, d_bw, d_indep;
        d_cp.upload(cp, st);
        d_ch.upload(ch, st);
        d_lo.upload(lo0, st);
        d_hi.upload(hi0, st);
        d_pm.upload(pm, st);
        d_ci.upload(ci, st);
        d_bid.upload(bid, st);
        d_bn.upload(bn, st);
        d_bcol.upload(bcol, st);
        d_bx.upload(bx, st);
        d_bw.upload(bw, st);
        d_piv.alloc((size_t)std::max<int64_t>(nids * maxrank, 1));
        d_rank.alloc((size_t)std::max<int64_t>(nids, 1));
        d_m.alloc((size_t)std::max<int64_t>(nids, 1));
        d_indep.alloc((size_t)std::max<int64_t>(nids, 1));
        d_err.alloc(2);
        NCA_HIP(hipMemsetAsync(d_rank.p, 0, sizeof(int32_t) * std::max<int64_t>(nids, 1), st));
        NCA_HIP(hipMemsetAsync(d_m.p, 0, sizeof(int32_t) * std::max<int64_t>(nids, 1), st));
        NCA_HIP(hipMemsetAsync(d_piv.p, 0, sizeof(int32_t) * std::max<int64_t>(nids * maxrank, 1), st));
        NCA_HIP(hipMemsetAsync(d_indep.p, 0xff, sizeof(int64_t) * std::max<int64_t>(nids, 1), st));
        NCA_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int32_t) * 2, st));
        DevBuf<double2> d_X((size_t)std::max<int64_t>(xtot, 1)), d_work((size_t)std::max<int64_t>(worktot, 1));
        const size_t lds = sizeof(double2) * std::max<size_t>(
                                                  std::max<size_t>(kLdsCap + kLdsMaxM + kLdsMaxN, kGlobMaxM + kGlobMaxN),
                                                  (size_t)(maxrank * maxrank));
        auto kern = method == 1 ? k_level_fused<true> : k_level_fused<false>;
        NCA_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        const nca_points* rowp = self;
        const nca_points* colp = other;
        LevelTree tr{d_cp.p, d_ch.p, d_lo.p, d_hi.p, d_pm.p};
        LevelOut out{d_piv.p, d_rank.p, d_m.p, d_indep.p, d_err.p};
        for (int64_t L = 0; L < nlevels; ++L) {
            const int64_t q0 = levptr[L] - levptr[0], nb = levptr[L + 1] - levptr[L];
            if (nb == 0) continue;
            hipLaunchKernelGGL(kern, dim3((unsigned)nb), dim3(kWG), lds, st, rowp->xyz.p, colp->xyz.p,
                               kernel->wavenumber, (int)kernel->sign, (int)(transposed != 0), d_bid.p + q0,
                               d_bn.p + q0, d_bcol.p + q0, d_ci.p, d_bx.p + q0, d_bw.p + q0, tr, maxrank, tol,
                               (int)(sampled != 0), d_work.p, d_X.p, out);
            NCA_HIP(hipGetLastError());
        }
        std::vector<int32_t> hr(std::max<int64_t>(nids, 1)), hm(std::max<int64_t>(nids, 1)),
            hp(std::max<int64_t>(nids * maxrank, 1)), herr(2);
        std::vector<int64_t> hind(std::max<int64_t>(nids, 1));
        d2h(ctx, hr.data(), d_rank.p, sizeof(int32_t) * hr.size());
        d2h(ctx, hm.data(), d_m.p, sizeof(int32_t) * hm.size());
        d2h(ctx, hp.data(), d_piv.p, sizeof(int32_t) * hp.size());
        d2h(ctx, hind.data(), d_indep.p, sizeof(int64_t) * hind.size());
        d2h(ctx, herr.data(), d_err.p, sizeof(int32_t) * 2);
        NCA_REQUIRE(herr[0] == 0, NCA_ERR_LIMIT,
                    "nca_bottomup_side: block " + std::to_string(herr[1] + 1) + " exceeds " +
                        std::to_string(kGlobMaxM) + " x " + std::to_string(kGlobMaxN));
        int64_t npiv = 0;
        for (int64_t q = 0; q < nblk; ++q) npiv += hr[bid[q]];
        NCA_REQUIRE(npiv <= pivcap, NCA_ERR_DIMENSION,
                    "nca_bottomup_side: the pivots need " + std::to_string(npiv) + " entries, pivcap is " +
                        std::to_string(pivcap));
        // the factors: one chunk at the bounded offsets, downloaded in level order
        const int32_t c = (int32_t)f->chunks.size();
        std::vector<int32_t> order;
        int64_t packed = 0;
        for (int64_t q = 0; q < nblk; ++q) {
            const int32_t id = bid[q];
            f->chunk[id] = c;
            f->off[id] = bx[q];
            f->m[id] = hm[id];
            f->rank[id] = hr[id];
            packed += (int64_t)hm[id] * hr[id];
            order.push_back(id);
        }
        f->chunks.push_back(std::move(d_X));
        f->chunk_size.push_back(packed);
        f->chunk_ids.push_back(std::move(order));
        for (int64_t i = 0; i < nids; ++i) {
            ranks[i] = seen[i] ? hr[i] : 0;
            mrows[i] = seen[i] ? hm[i] : 0;
            independent[i] = seen[i] ? hind[i] : -1;
        }
        int64_t o = 0;
        for (int64_t q = 0; q < nblk; ++q) {
            const int64_t id = bid[q];
            for (int32_t k = 0; k < hr[id]; ++k) pivots[o++] = (int64_t)hp[id * maxrank + k] + 1;
        }
        for (int64_t q = 0; q < nblk; ++q) {
            const int32_t id = bid[q];
            if (hind[id] >= 0)
                throw Error(NCA_ERR_SINGULAR, "SingularCrossMatrixError: block " + std::to_string(id + 1) + ": " +
                                                  std::to_string(hr[id]) + " pivots were selected but only " +
                                                  std::to_string(hind[id]) + " of them are linearly independent");
        }
    });
}

int32_t nca_topdown_side(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* self, const nca_points* other,
                         int32_t transposed, int64_t nn, const int64_t* parent, const int64_t* childptr,
                         const int64_t* children, const int64_t* lo, const int64_t* hi, const int64_t* perm,
                         int64_t nn_other, const int64_t* olo, const int64_t* ohi, const int64_t* operm,
                         const int64_t* farptr, const int64_t* farnodes, int64_t nlevels, const int64_t* levptr,
                         const int64_t* levids, int64_t maxrank, double tol, nca_factors* f, int32_t* ranks,
                         int64_t* mrows, int64_t pivcap, int64_t* pivots, int64_t* cpivots, int64_t* independent) {
    return guarded([&] {
        NCA_REQUIRE(ctx && kernel && self && other && f && parent && childptr && lo && hi && perm && olo && ohi &&
                        operm && farptr && levptr && ranks && mrows && pivots && cpivots && independent && nn >= 0 &&
                        nn_other >= 0 && nlevels >= 0,
                    NCA_ERR_ARGUMENT, "nca_topdown_side: invalid argument");
        NCA_REQUIRE(f->ctx == ctx && f->nids == nn, NCA_ERR_ARGUMENT, "nca_topdown_side: factors of another tree");
        NCA_REQUIRE(maxrank >= 0 && maxrank <= kSolveMaxRank, NCA_ERR_LIMIT,
                    "nca_topdown_side: maxrank above " + std::to_string(kSolveMaxRank));
        NCA_REQUIRE(nn < (1LL << 31) && self->n < (1LL << 31) && other->n < (1LL << 31), NCA_ERR_LIMIT,
                    "nca_topdown_side: more than 2^31 nodes or points");
        set_device(ctx);
        hipStream_t st = ctx->stream;
        // 0-based 32-bit copies, each value checked against [low, lim]
        auto to32 = [](const int64_t* v, int64_t n, int64_t shift, int64_t lim, const char* what, int64_t low = 0) {
            std::vector<int32_t> o((size_t)std::max<int64_t>(n, 1));
            for (int64_t i = 0; i < n; ++i) {
                const int64_t x = v[i] - shift;
                NCA_REQUIRE(x >= low && x <= lim, NCA_ERR_ARGUMENT, std::string("nca_topdown_side: ") + what +
                                                                          " out of range");
                o[i] = (int32_t)x;
            }
            return o;
        };
        std::vector<int32_t> cp = to32(childptr, nn + 1, childptr[0], INT32_MAX, "childptr");
        std::vector<int32_t> ch = to32(children, cp[nn], 1, nn - 1, "child");
        std::vector<int32_t> lo0 = to32(lo, nn, 1, self->n, "value range"), hi0 = to32(hi, nn, 0, self->n, "value range");
        std::vector<int32_t> pm = to32(perm, self->n, 1, self->n - 1, "perm");
        std::vector<int32_t> pa = to32(parent, nn, 1, nn - 1, "parent", -1);
        std::vector<int32_t> olo0 = to32(olo, nn_other, 1, other->n, "other value range"),
                             ohi0 = to32(ohi, nn_other, 0, other->n, "other value range");
        std::vector<int32_t> opm = to32(operm, other->n, 1, other->n - 1, "other perm");
        const int64_t nfar = farptr[nn] - farptr[0];
        std::vector<int32_t> fn = to32(farnodes, nfar, 1, nn_other - 1, "far node");
        std::vector<int64_t> fp(nn + 1);
        for (int64_t i = 0; i <= nn; ++i) fp[i] = farptr[i] - farptr[0];
        // batches of one level: blocks whose far field may be nonempty, their
        // scratch (U, V, columns, flags) bounded up front
        struct Batch { int64_t first, count; };
        std::vector<int32_t> bid;
        std::vector<int64_t> buoff, bvoff, bcoff, broff;
        std::vector<Batch> batches;
        std::vector<int64_t> ncapv(nn, 0);
        const int64_t budget = (int64_t)4 << 30;
        int64_t umax = 0, vmax = 0, cmax = 0, fmax = 0;
        for (int64_t L = 0; L < nlevels; ++L) {
            int64_t u = 0, v = 0, c = 0, fl = 0;
            int64_t start = (int64_t)bid.size();
            std::vector<int64_t> rows_of;
            auto close = [&] {
                if ((int64_t)bid.size() > start) {
                    batches.push_back({start, (int64_t)bid.size() - start});
                    // row flags after every column flag of the batch
                    for (int64_t q = start; q < (int64_t)bid.size(); ++q) broff[q] += c;
                    umax = std::max(umax, u);
                    vmax = std::max(vmax, v);
                    cmax = std::max(cmax, c);
                    fmax = std::max(fmax, c + fl);
                }
                start = (int64_t)bid.size();
                u = v = c = fl = 0;
            };
            for (int64_t q = levptr[L] - levptr[0]; q < levptr[L + 1] - levptr[0]; ++q) {
                const int64_t id = levids[q] - 1;
                NCA_REQUIRE(id >= 0 && id < nn, NCA_ERR_ARGUMENT, "nca_topdown_side: level node out of range");
                int64_t ncap = pa[id] >= 0 ? maxrank : 0;
                for (int64_t e = fp[id]; e < fp[id + 1]; ++e) ncap += ohi0[fn[e]] - olo0[fn[e]];
                ncapv[id] = ncap;
                if (ncap == 0) continue;
                const int64_t m = hi0[id] - lo0[id];
                const int64_t kc = std::min<int64_t>(std::min<int64_t>(maxrank, m), ncap);
                const int64_t bytes = (m + ncap) * kc * 16 + ncap * 5 + m;
                if (u + v > 0 && (u + v) * 16 + (c * 5 + fl) + bytes > budget) close();
                bid.push_back((int32_t)id);
                buoff.push_back(u);
                bvoff.push_back(v);
                bcoff.push_back(c);
                broff.push_back(fl);
                u += m * kc;
                v += ncap * kc;
                c += ncap;
                fl += m;
            }
            close();
        }
        DevBuf<int32_t> d_lo, d_hi, d_pm, d_pa, d_cp, d_ch, d_olo, d_ohi, d_opm, d_fn, d_bid, d_rank, d_rp, d_cpv,
            d_m, d_err;
        DevBuf<int64_t> d_fp, d_uo, d_vo, d_co, d_ro, d_indep;
        d_lo.upload(lo0, st);
        d_hi.upload(hi0, st);
        d_pm.upload(pm, st);
        d_pa.upload(pa, st);
        d_cp.upload(cp, st);
        d_ch.upload(ch, st);
        d_olo.upload(olo0, st);
        d_ohi.upload(ohi0, st);
        d_opm.upload(opm, st);
        d_fn.upload(fn, st);
        d_fp.upload(fp, st);
        if (!bid.empty()) {
            d_bid.upload(bid, st);
            d_uo.upload(buoff, st);
            d_vo.upload(bvoff, st);
            d_co.upload(bcoff, st);
            d_ro.upload(broff, st);
        }
        const size_t nrk = (size_t)std::max<int64_t>(nn * maxrank, 1);
        d_rank.alloc((size_t)std::max<int64_t>(nn, 1));
        d_rp.alloc(nrk);
        d_cpv.alloc(nrk);
        NCA_HIP(hipMemsetAsync(d_rank.p, 0xff, sizeof(int32_t) * d_rank.n, st));
        NCA_HIP(hipMemsetAsync(d_rp.p, 0, sizeof(int32_t) * nrk, st));
        NCA_HIP(hipMemsetAsync(d_cpv.p, 0, sizeof(int32_t) * nrk, st));
        DevBuf<double2> d_U((size_t)std::max<int64_t>(umax, 1)), d_V((size_t)std::max<int64_t>(vmax, 1));
        DevBuf<int32_t> d_cols((size_t)std::max<int64_t>(cmax, 1));
        DevBuf<uint8_t> d_used((size_t)std::max<int64_t>(fmax, 1));
        PacaTree tr{d_lo.p, d_hi.p, d_pm.p, d_pa.p, d_fp.p, d_fn.p, d_olo.p, d_ohi.p, d_opm.p};
        for (const Batch& B : batches) {
            PacaScratch sc{d_U.p, d_V.p, d_cols.p, d_used.p, d_uo.p + B.first, d_vo.p + B.first, d_co.p + B.first,
                           d_ro.p + B.first};
            hipLaunchKernelGGL(k_paca_level, dim3((unsigned)B.count), dim3(kWG), 0, st, self->xyz.p, other->xyz.p,
                               kernel->wavenumber, (int)kernel->sign, (int)(transposed != 0), d_bid.p + B.first, tr,
                               sc, maxrank, tol, d_rank.p, d_rp.p, d_cpv.p);
            NCA_HIP(hipGetLastError());
        }
        // factors of the blocks that may hold pivots, in level order
        std::vector<int64_t> rb(nn, 0), fx(bid.size());
        for (int64_t q = 0; q < (int64_t)bid.size(); ++q) {
            const int32_t id = bid[q];
            rb[id] = std::min<int64_t>(maxrank, hi0[id] - lo0[id]);
        }
        int64_t xtot = 0;
        for (int64_t q = 0; q < (int64_t)bid.size(); ++q) {
            const int32_t id = bid[q];
            int64_t mcap = 0;
            if (cp[id] == cp[id + 1]) mcap = hi0[id] - lo0[id];
            else
                for (int32_t e = cp[id]; e < cp[id + 1]; ++e) mcap += rb[ch[e]];
            fx[q] = xtot;
            xtot += mcap * rb[id];
        }
        DevBuf<int64_t> d_fx;
        if (!fx.empty()) d_fx.upload(fx, st);
        d_m.alloc((size_t)std::max<int64_t>(nn, 1));
        d_indep.alloc((size_t)std::max<int64_t>(nn, 1));
        d_err.alloc(2);
        NCA_HIP(hipMemsetAsync(d_m.p, 0, sizeof(int32_t) * d_m.n, st));
        NCA_HIP(hipMemsetAsync(d_indep.p, 0xff, sizeof(int64_t) * d_indep.n, st));
        NCA_HIP(hipMemsetAsync(d_err.p, 0, sizeof(int32_t) * 2, st));
        DevBuf<double2> d_X((size_t)std::max<int64_t>(xtot, 1));
        if (!bid.empty()) {
            const size_t lds = sizeof(double2) * (size_t)std::max<int64_t>(maxrank * maxrank, 1);
            NCA_HIP(hipFuncSetAttribute((const void*)k_topdown_factors, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds));
            hipLaunchKernelGGL(k_topdown_factors, dim3((unsigned)bid.size()), dim3(kWG), lds, st, self->xyz.p,
                               other->xyz.p, kernel->wavenumber, (int)kernel->sign, (int)(transposed != 0), d_bid.p,
                               d_fx.p, d_cp.p, d_ch.p, d_lo.p, d_hi.p, d_pm.p, d_rank.p, d_rp.p, d_cpv.p, maxrank,
                               d_m.p, d_indep.p, d_X.p, d_err.p);
            NCA_HIP(hipGetLastError());
        }
        std::vector<int32_t> hr(d_rank.n), hm(d_m.n), hrp(nrk), hcp(nrk), herr(2);
        std::vector<int64_t> hind(d_indep.n);
        d2h(ctx, hr.data(), d_rank.p, sizeof(int32_t) * hr.size());
        d2h(ctx, hm.data(), d_m.p, sizeof(int32_t) * hm.size());
        d2h(ctx, hrp.data(), d_rp.p, sizeof(int32_t) * nrk);
        d2h(ctx, hcp.data(), d_cpv.p, sizeof(int32_t) * nrk);
        d2h(ctx, hind.data(), d_indep.p, sizeof(int64_t) * hind.size());
        d2h(ctx, herr.data(), d_err.p, sizeof(int32_t) * 2);
        NCA_REQUIRE(herr[0] == 0, NCA_ERR_LIMIT,
                    "nca_topdown_side: block " + std::to_string(herr[1] + 1) + " has more than " +
                        std::to_string(kGlobMaxM) + " rows");
        int64_t npiv = 0;
        for (int32_t id : bid) npiv += std::max(hr[id], 0);
        NCA_REQUIRE(npiv <= pivcap, NCA_ERR_DIMENSION,
                    "nca_topdown_side: the pivots need " + std::to_string(npiv) + " entries, pivcap is " +
                        std::to_string(pivcap));
        for (int32_t id : bid)
            NCA_REQUIRE(f->chunk[id] < 0, NCA_ERR_ARGUMENT, "nca_topdown_side: a block has a factor already");
        const int32_t c = (int32_t)f->chunks.size();
        std::vector<int32_t> order;
        int64_t packed = 0, o = 0;
        for (int64_t q = 0; q < (int64_t)bid.size(); ++q) {
            const int32_t id = bid[q];
            if (hr[id] < 0) continue;
            f->chunk[id] = c;
            f->off[id] = fx[q];
            f->m[id] = hm[id];
            f->rank[id] = hr[id];
            packed += (int64_t)hm[id] * hr[id];
            order.push_back(id);
            for (int32_t k2 = 0; k2 < hr[id]; ++k2) {
                pivots[o] = (int64_t)hrp[(int64_t)id * maxrank + k2] + 1;
                cpivots[o] = (int64_t)hcp[(int64_t)id * maxrank + k2] + 1;
                ++o;
            }
        }
        f->chunks.push_back(std::move(d_X));
        f->chunk_size.push_back(packed);
        f->chunk_ids.push_back(std::move(order));
        for (int64_t i = 0; i < nn; ++i) {
            ranks[i] = hr[i];
            mrows[i] = hr[i] >= 0 ? hm[i] : 0;
            independent[i] = hr[i] >= 0 ? hind[i] : -1;
        }
        for (int32_t id : bid)
            if (hr[id] >= 0 && hind[id] >= 0)
                throw Error(NCA_ERR_SINGULAR, "SingularCrossMatrixError: block " + std::to_string(id + 1) + ": " +
                                                  std::to_string(hr[id]) + " pivots were selected but only " +
                                                  std::to_string(hind[id]) + " of them are linearly independent");
    });
}

int32_t nca_compress_level_resident(nca_ctx* ctx, const nca_stack* st, double tol, int32_t sampled, int32_t method,
                                    nca_factors* f, const int64_t* ids, int32_t* ranks, int64_t* rowpivots,
                                    int64_t* colpivots, int64_t* independent) {
    return guarded([&] {
        NCA_REQUIRE(method == 0 || method == 1, NCA_ERR_ARGUMENT, "nca_compress_level_resident: unknown method");
        compress_level_keep(ctx, st, tol, method == 1 ? 0 : sampled, method == 1, f, ids, ranks, rowpivots,
                            colpivots, independent);
    });
}

int32_t nca_compress_level(nca_ctx* ctx, const nca_stack* st, double tol, int32_t sampled, int32_t* ranks,
                           int64_t* rowpivots, int64_t* colpivots, int64_t xcapacity, int64_t* xoffsets,
                           double* X, int64_t* independent) {
    return guarded([&] {
        compress_level_impl(ctx, st, tol, sampled, false, ranks, rowpivots, colpivots, xcapacity, xoffsets, X,
                            independent);
    });
}

int32_t nca_compress_level_iaca(nca_ctx* ctx, const nca_stack* st, double tol, int32_t* ranks, int64_t* rowpivots,
                                int64_t* colpivots, int64_t xcapacity, int64_t* xoffsets, double* X,
                                int64_t* independent) {
    return guarded([&] {
        compress_level_impl(ctx, st, tol, 0, true, ranks, rowpivots, colpivots, xcapacity, xoffsets, X,
                            independent);
    });
}

int32_t nca_iaca(nca_ctx* ctx, const nca_stack* st, double tol, int32_t* ranks, int64_t* rowpivots,
                 int64_t* colpivots) {
    return guarded([&] {
        NCA_REQUIRE(ctx && st && ranks && rowpivots && colpivots, NCA_ERR_ARGUMENT, "nca_iaca: invalid argument");
        set_device(ctx);
        AcaResult res;
        launch_aca(ctx, st, tol, 0, res, true);
        if (st->nb == 0) return;
        d2h(ctx, ranks, res.ranks.p, sizeof(int32_t) * st->nb);
        d2h(ctx, rowpivots, res.rowpiv.p, sizeof(int64_t) * st->nb * st->maxrank);
        d2h(ctx, colpivots, res.colpiv.p, sizeof(int64_t) * st->nb * st->maxrank);
    });
}

int32_t nca_coupling_assemble(nca_ctx* ctx, const nca_helmholtz* kernel, const nca_points* test,
                              const nca_points* trial, int64_t ntest, const int64_t* farptr,
                              const int64_t* farnodes, int64_t ntrial, const int64_t* testpivptr,
                              const int64_t* testpiv, const int64_t* trialpivptr, const int64_t* trialpiv,
                              nca_coupling** out) {
    return guarded([&] {
        NCA_REQUIRE(ctx && kernel && test && trial && out && farptr && testpivptr && trialpivptr && ntest >= 0 &&
                        ntrial >= 0,
                    NCA_ERR_ARGUMENT, "nca_coupling_assemble: invalid argument");
        set_device(ctx);
        auto cp = std::make_unique<nca_coupling>();
        cp->ctx = ctx;
        cp->ntest = ntest;
        cp->ntrial = ntrial;
        cp->ptr.resize(ntest + 1);
        for (int64_t t = 0; t <= ntest; ++t) cp->ptr[t] = farptr[t] - farptr[0];
        cp->nblocks = cp->ptr[ntest];
        cp->nodes.resize(cp->nblocks);
        cp->testrank.resize(ntest);
        cp->trialrank.resize(ntrial);
        int rmax = 0;
        for (int64_t t = 0; t < ntest; ++t) {
            cp->testrank[t] = (int32_t)(testpivptr[t + 1] - testpivptr[t]);
            rmax = std::max(rmax, cp->testrank[t]);
        }
        for (int64_t s = 0; s < ntrial; ++s) cp->trialrank[s] = (int32_t)(trialpivptr[s + 1] - trialpivptr[s]);
        NCA_REQUIRE(rmax <= kCouplingMaxRank, NCA_ERR_LIMIT, "coupling: test rank exceeds 256");
        cp->offsets.assign(cp->nblocks + 1, 0);
        std::vector<int32_t> nodes32(cp->nblocks);
        for (int64_t t = 0; t < ntest; ++t)
            for (int64_t i = cp->ptr[t]; i < cp->ptr[t + 1]; ++i) {
                const int64_t s = farnodes[i] - 1;
                NCA_REQUIRE(s >= 0 && s < ntrial, NCA_ERR_ARGUMENT, "nca_coupling_assemble: far node out of range");
                cp->nodes[i] = s;
                nodes32[i] = (int32_t)s;
                cp->offsets[i + 1] = cp->offsets[i] + (int64_t)cp->testrank[t] * cp->trialrank[s];
            }
        cp->nentries = cp->offsets[cp->nblocks];
        const int64_t ntp = testpivptr[ntest] - testpivptr[0], nsp = trialpivptr[ntrial] - trialpivptr[0];
        std::vector<int64_t> tptr(ntest + 1), sptr(ntrial + 1);
        for (int64_t t = 0; t <= ntest; ++t) tptr[t] = testpivptr[t] - testpivptr[0];
        for (int64_t s = 0; s <= ntrial; ++s) sptr[s] = trialpivptr[s] - trialpivptr[0];
        std::vector<int32_t> tpiv(ntp), spiv(nsp);
        for (int64_t i = 0; i < ntp; ++i) {
            NCA_REQUIRE(testpiv[i] >= 1 && testpiv[i] <= test->n, NCA_ERR_ARGUMENT, "coupling: test pivot out of range");
            tpiv[i] = (int32_t)(testpiv[i] - 1);
        }
        for (int64_t i = 0; i < nsp; ++i) {
            NCA_REQUIRE(trialpiv[i] >= 1 && trialpiv[i] <= trial->n, NCA_ERR_ARGUMENT,
                        "coupling: trial pivot out of range");
            spiv[i] = (int32_t)(trialpiv[i] - 1);
        }
        cp->tptr = tptr;
        cp->sptr = sptr;
        cp->tpiv = tpiv;
        cp->spiv = spiv;
        cp->blocks.alloc((size_t)std::max<int64_t>(cp->nentries, 1));
        if (cp->nentries > 0) {
            DevBuf<int64_t> d_ptr, d_off, d_tptr, d_sptr;
            DevBuf<int32_t> d_nodes, d_tpiv, d_spiv;
            d_ptr.upload(cp->ptr, ctx->stream);
            d_off.upload(cp->offsets, ctx->stream);
            d_tptr.upload(tptr, ctx->stream);
            d_sptr.upload(sptr, ctx->stream);
            d_nodes.upload(nodes32, ctx->stream);
            d_tpiv.upload(tpiv, ctx->stream);
            d_spiv.upload(spiv, ctx->stream);
            hipLaunchKernelGGL(k_coupling, dim3((unsigned)ntest), dim3(kWG), 0, ctx->stream, test->xyz.p,
                               trial->xyz.p, d_ptr.p, d_nodes.p, d_off.p, d_tptr.p, d_tpiv.p, d_sptr.p, d_spiv.p,
                               kernel->wavenumber, (int)kernel->sign, cp->blocks.p);
            NCA_HIP(hipGetLastError());
            NCA_HIP(hipStreamSynchronize(ctx->stream));
        }
        *out = cp.release();
    });
}

int32_t nca_coupling_size(const nca_coupling* cp, int64_t* nblocks, int64_t* nentries) {
    return guarded([&] {
        NCA_REQUIRE(cp, NCA_ERR_ARGUMENT, "nca_coupling_size: null handle");
        if (nblocks) *nblocks = cp->nblocks;
        if (nentries) *nentries = cp->nentries;
    });
}

int32_t nca_coupling_download(nca_ctx* ctx, const nca_coupling* cp, int64_t* offsets, double* blocks) {
    return guarded([&] {
        NCA_REQUIRE(ctx && cp, NCA_ERR_ARGUMENT, "nca_coupling_download: invalid argument");
        set_device(ctx);
        if (offsets) std::copy(cp->offsets.begin(), cp->offsets.end(), offsets);
        if (blocks && cp->nentries > 0)
            d2h(ctx, blocks, cp->blocks.p, sizeof(double2) * cp->nentries);
    });
}

int32_t nca_coupling_destroy(nca_coupling* cp) {
    return guarded([&] {
        if (cp) {
            if (cp->ctx) set_device(cp->ctx);
            delete cp;
        }
    });
}


// Phase 1 on the device; see k_farpivots. Counts follow the host rule:
// n = min(maxrank, ceil(oversample |t|)) capped by the far field's points.
int32_t nca_farpivots(nca_ctx* ctx, const nca_points* other, int64_t nitems, const double* centers,
                      const int64_t* nvals, const int64_t* ffptr, const int64_t* ffnodes, int64_t nn_other,
                      const int64_t* other_lo, const int64_t* other_hi, const int64_t* other_perm, int64_t maxrank,
                      double weight, double oversample, int64_t* pivptr, int64_t pivcap, int64_t* piv) {
    return guarded([&] {
        NCA_REQUIRE(ctx && other && nitems >= 0 && ffptr && pivptr && (nitems == 0 || (centers && nvals)),
                    NCA_ERR_ARGUMENT, "nca_farpivots: invalid argument");
        set_device(ctx);
        hipStream_t st = ctx->stream;
        const int64_t npts = other->n;
        std::vector<int32_t> lo(nn_other), hi(nn_other), perm(npts);
        for (int64_t i = 0; i < nn_other; ++i) {
            lo[i] = (int32_t)(other_lo[i] - 1);
            hi[i] = (int32_t)other_hi[i];
            NCA_REQUIRE(lo[i] >= 0 && hi[i] >= lo[i] && hi[i] <= npts, NCA_ERR_ARGUMENT,
                        "nca_farpivots: node range out of bounds");
        }
        for (int64_t i = 0; i < npts; ++i) {
            perm[i] = (int32_t)(other_perm[i] - 1);
            NCA_REQUIRE(perm[i] >= 0 && perm[i] < npts, NCA_ERR_ARGUMENT, "nca_farpivots: perm out of range");
        }
        const int64_t nff = ffptr[nitems] - ffptr[0];
        std::vector<int64_t> fp(nitems + 1), cnt(nitems, 0);
        std::vector<int32_t> fn(nff), fscan(nff), iper(nitems, 0), ipool(nitems, 0);
        // per item: far nodes (checked), pivot count, and the candidate pool
        // size the device will build (bounded by kP1Pool); 0 ok, else an error code
        std::vector<char> bad(nitems, 0);
        for (int64_t t = 0; t <= nitems; ++t) fp[t] = ffptr[t] - ffptr[0];
        host_parallel_for(nitems, [&](int64_t t) {
            if (fp[t + 1] - fp[t] > kP1MaxFar) { bad[t] = 1; return; }
            int64_t total = 0;
            for (int64_t k = fp[t]; k < fp[t + 1]; ++k) {
                const int64_t f = ffnodes[k] - 1;
                if (f < 0 || f >= nn_other) { bad[t] = 2; return; }
                fn[k] = (int32_t)f;
                total += hi[f] - lo[f];
            }
            const int64_t limit = std::min<int64_t>(maxrank, (int64_t)std::ceil(oversample * (double)nvals[t]));
            const int64_t n = std::max<int64_t>(0, std::min(limit, total));
            cnt[t] = n;
            if (n == 0) return;
            const int64_t nf = fp[t + 1] - fp[t];
            int64_t per = std::max<int64_t>(1, (std::max<int64_t>(8 * n, 512) + nf - 1) / nf), pool = 0;
            for (;;) {
                pool = 0;
                for (int64_t k = fp[t]; k < fp[t + 1]; ++k) pool += std::min<int64_t>(hi[fn[k]] - lo[fn[k]], per);
                if (pool >= n || pool >= total) break;
                per *= 2;
            }
            if (pool > kP1Pool) { bad[t] = 3; return; }
            // the one-wave kernel's inputs: per, pool and the scan of the per-node counts
            iper[t] = (int32_t)per;
            ipool[t] = (int32_t)pool;
            int64_t run = 0;
            for (int64_t k = fp[t]; k < fp[t + 1]; ++k) {
                fscan[k] = (int32_t)run;
                run += std::min<int64_t>(hi[fn[k]] - lo[fn[k]], per);
            }
        });
        for (int64_t t = 0; t < nitems; ++t) {
            NCA_REQUIRE(bad[t] != 1, NCA_ERR_LIMIT, "nca_farpivots: more than 4096 far nodes in one far field");
            NCA_REQUIRE(bad[t] != 2, NCA_ERR_ARGUMENT, "nca_farpivots: far node out of range");
            NCA_REQUIRE(bad[t] != 3, NCA_ERR_LIMIT, "nca_farpivots: candidate pool above 4096 points");
        }
        pivptr[0] = 0;
        for (int64_t t = 0; t < nitems; ++t) pivptr[t + 1] = pivptr[t] + cnt[t];
        NCA_REQUIRE(pivptr[nitems] <= pivcap, NCA_ERR_LIMIT,
                    "nca_farpivots: pivot capacity too small (" + std::to_string(pivptr[nitems]) + " needed)");
        DevBuf<int32_t> dlo, dhi, dperm, dfn, derr(1);
        DevBuf<int64_t> dfp, dptr, dpiv;
        DevBuf<double> dctr;
        dlo.upload(lo, st);
        dhi.upload(hi, st);
        dperm.upload(perm, st);
        dfn.upload(fn, st);
        dfp.upload(fp, st);
        std::vector<int64_t> pp(pivptr, pivptr + nitems + 1);
        dptr.upload(pp, st);
        dctr.upload(centers, (size_t)(3 * nitems), st);
        dpiv.alloc((size_t)std::max<int64_t>(pivptr[nitems], 1));
        const int32_t zero = 0;
        derr.upload(&zero, 1, st);
        // items with pivots: one wave each when the pool fits kP1WavePool
        // candidates, else one workgroup each
        std::vector<int32_t> wave_items, wave_per, wave_pool, block_items;
        for (int64_t t = 0; t < nitems; ++t) {
            if (cnt[t] == 0) continue;
            if (ipool[t] <= kP1WavePool) {
                wave_items.push_back((int32_t)t);
                wave_per.push_back(iper[t]);
                wave_pool.push_back(ipool[t]);
            } else {
                block_items.push_back((int32_t)t);
            }
        }
        DevBuf<int32_t> dwi, dwper, dwpool, dfs, dbi;
        if (!wave_items.empty()) {
            dwi.upload(wave_items, st);
            dwper.upload(wave_per, st);
            dwpool.upload(wave_pool, st);
            dfs.upload(fscan, st);
            const int64_t nw = (int64_t)wave_items.size();
            hipLaunchKernelGGL(k_farpivots_wave, dim3((unsigned)((nw + 3) / 4)), dim3(256), 0, st, other->xyz.p, dlo.p,
                               dhi.p, dperm.p, dctr.p, dfp.p, dfn.p, dfs.p, dwi.p, nw, dwper.p, dwpool.p, dptr.p,
                               weight, dpiv.p);
            NCA_HIP(hipGetLastError());
        }
        if (!block_items.empty()) {
            dbi.upload(block_items, st);
            hipLaunchKernelGGL(k_farpivots<false>, dim3((unsigned)block_items.size()), dim3(kP1Threads), 0, st,
                               other->xyz.p, dlo.p, dhi.p, dperm.p, dctr.p, dfp.p, dfn.p, nullptr, dptr.p, weight,
                               dpiv.p, derr.p, dbi.p);
        }
        NCA_HIP(hipGetLastError());
        if (pivptr[nitems] > 0) d2h(ctx, piv, dpiv.p, sizeof(int64_t) * pivptr[nitems]);
        NCA_HIP(hipStreamSynchronize(st));
    });
}

// Phase 1 on the device with the far fields built there: the far lists by
// node (farptr / farnodes, 1-based; the admissible pairs of the self tree)
// and the parent of every node (1-based, 0 at a root). Counts as in
// nca_farpivots, with the far-field sizes summed along the parent chain.
int32_t nca_farpivots_tree(nca_ctx* ctx, const nca_points* other, int64_t nitems, const double* centers,
                           const int64_t* nvals, const int64_t* parent, const int64_t* farptr,
                           const int64_t* farnodes, int64_t nn_other, const int64_t* other_lo,
                           const int64_t* other_hi, const int64_t* other_perm, int64_t maxrank, double weight,
                           double oversample, int64_t* pivptr, int64_t pivcap, int64_t* piv) {
    return guarded([&] {
        NCA_REQUIRE(ctx && other && nitems >= 0 && farptr && pivptr && (nitems == 0 || (centers && nvals && parent)),
                    NCA_ERR_ARGUMENT, "nca_farpivots_tree: invalid argument");
        set_device(ctx);
        hipStream_t st = ctx->stream;
        const int64_t npts = other->n;
        std::vector<int32_t> lo(nn_other), hi(nn_other), perm(npts), par(nitems);
        std::vector<char> bad(std::max<int64_t>({nn_other, npts, nitems, (int64_t)1}), 0);
        host_parallel_for(nn_other, [&](int64_t i) {
            lo[i] = (int32_t)(other_lo[i] - 1);
            hi[i] = (int32_t)other_hi[i];
            if (lo[i] < 0 || hi[i] < lo[i] || hi[i] > npts) bad[i] = 1;
        });
        for (int64_t i = 0; i < nn_other; ++i)
            NCA_REQUIRE(!bad[i], NCA_ERR_ARGUMENT, "nca_farpivots_tree: node range out of bounds");
        host_parallel_for(npts, [&](int64_t i) {
            perm[i] = (int32_t)(other_perm[i] - 1);
            bad[i] = perm[i] < 0 || perm[i] >= npts;
        });
        for (int64_t i = 0; i < npts; ++i) NCA_REQUIRE(!bad[i], NCA_ERR_ARGUMENT, "nca_farpivots_tree: perm out of range");
        for (int64_t t = 0; t < nitems; ++t) {
            par[t] = (int32_t)(parent[t] - 1);
            NCA_REQUIRE(par[t] >= -1 && par[t] < nitems && par[t] != t, NCA_ERR_ARGUMENT,
                        "nca_farpivots_tree: parent out of range");
        }
        const int64_t nfar = farptr[nitems] - farptr[0];
        std::vector<int64_t> fp(nitems + 1);
        std::vector<int32_t> fn(nfar);
        for (int64_t t = 0; t <= nitems; ++t) fp[t] = farptr[t] - farptr[0];
        // own far lists: checked, and their node and point counts
        std::vector<int64_t> ownn(nitems), ownpts(nitems);
        host_parallel_for(nitems, [&](int64_t t) {
            int64_t p = 0;
            for (int64_t k = fp[t]; k < fp[t + 1]; ++k) {
                const int64_t f = farnodes[k] - 1;
                if (f < 0 || f >= nn_other) { bad[t] = 1; return; }
                fn[k] = (int32_t)f;
                p += hi[f] - lo[f];
            }
            bad[t] = 0;
            ownn[t] = fp[t + 1] - fp[t];
            ownpts[t] = p;
        });
        for (int64_t t = 0; t < nitems; ++t) NCA_REQUIRE(!bad[t], NCA_ERR_ARGUMENT, "nca_farpivots_tree: far node out of range");
        // far-field node and point counts along the parent chains (memoized;
        // a chain longer than nitems is a cycle)
        std::vector<int64_t> ffn(nitems, -1), ffp(nitems, 0);
        std::vector<int64_t> chain;
        for (int64_t t = 0; t < nitems; ++t) {
            chain.clear();
            int64_t a = t;
            while (a >= 0 && ffn[a] < 0) {
                chain.push_back(a);
                NCA_REQUIRE((int64_t)chain.size() <= nitems, NCA_ERR_ARGUMENT, "nca_farpivots_tree: parent cycle");
                a = par[a];
            }
            int64_t n0 = a >= 0 ? ffn[a] : 0, p0 = a >= 0 ? ffp[a] : 0;
            for (size_t i = chain.size(); i-- > 0;) {
                n0 += ownn[chain[i]];
                p0 += ownpts[chain[i]];
                ffn[chain[i]] = n0;
                ffp[chain[i]] = p0;
            }
        }
        pivptr[0] = 0;
        for (int64_t t = 0; t < nitems; ++t) {
            NCA_REQUIRE(ffn[t] <= kP1MaxFar, NCA_ERR_LIMIT, "nca_farpivots_tree: more than 4096 far nodes in one far field");
            const int64_t limit = std::min<int64_t>(maxrank, (int64_t)std::ceil(oversample * (double)nvals[t]));
            pivptr[t + 1] = pivptr[t] + std::max<int64_t>(0, std::min(limit, ffp[t]));
        }
        NCA_REQUIRE(pivptr[nitems] <= pivcap, NCA_ERR_LIMIT,
                    "nca_farpivots_tree: pivot capacity too small (" + std::to_string(pivptr[nitems]) + " needed)");
        DevBuf<int32_t> dlo, dhi, dperm, dfn, dpar, derr(1);
        DevBuf<int64_t> dfp, dptr, dpiv;
        DevBuf<double> dctr;
        dlo.upload(lo, st);
        dhi.upload(hi, st);
        dperm.upload(perm, st);
        dfn.upload(fn, st);
        dfp.upload(fp, st);
        dpar.upload(par, st);
        std::vector<int64_t> pp(pivptr, pivptr + nitems + 1);
        dptr.upload(pp, st);
        dctr.upload(centers, (size_t)(3 * nitems), st);
        dpiv.alloc((size_t)std::max<int64_t>(pivptr[nitems], 1));
        const int32_t zero = 0;
        derr.upload(&zero, 1, st);
        if (nitems > 0)
            hipLaunchKernelGGL(k_farpivots<true>, dim3((unsigned)nitems), dim3(kP1Threads), 0, st, other->xyz.p,
                               dlo.p, dhi.p, dperm.p, dctr.p, dfp.p, dfn.p, dpar.p, dptr.p, weight, dpiv.p, derr.p,
                               nullptr);
        NCA_HIP(hipGetLastError());
        int32_t e = 0;
        d2h(ctx, &e, derr.p, sizeof(int32_t));
        NCA_REQUIRE(e == 0, NCA_ERR_LIMIT, "nca_farpivots_tree: candidate pool above 4096 points");
        if (pivptr[nitems] > 0) d2h(ctx, piv, dpiv.p, sizeof(int64_t) * pivptr[nitems]);
        NCA_HIP(hipStreamSynchronize(st));
    });
}

}  // extern "C"
