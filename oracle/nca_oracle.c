/*
 * ORACLE — test infrastructure only (see nca_oracle.h).
 *
 * FP64 CPU restatement of the batched bottom-up NCA compressor and the H^2
 * matvec. Every function cites the reference file it restates. Compiled with
 * -ffp-contract=off so that no multiply-add is fused behind our back: the
 * pivot-deciding arithmetic (entries, residual updates, norms) is written out
 * operation by operation, in a fixed order.
 */
#include "nca_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    double re, im;
} cplx;

static inline cplx cmul(cplx a, cplx b) {
    cplx r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
    return r;
}
static inline cplx csub(cplx a, cplx b) {
    cplx r = {a.re - b.re, a.im - b.im};
    return r;
}
static inline cplx cadd(cplx a, cplx b) {
    cplx r = {a.re + b.re, a.im + b.im};
    return r;
}
static inline double cabs2(cplx a) { return a.re * a.re + a.im * a.im; }
static inline double cabs1(cplx a) { return fabs(a.re) + fabs(a.im); }
static inline cplx cdiv(cplx a, cplx b) {
    double d = b.re * b.re + b.im * b.im;
    cplx r = {(a.re * b.re + a.im * b.im) / d, (a.im * b.re - a.re * b.im) / d};
    return r;
}
static inline cplx cconj(cplx a) {
    cplx r = {a.re, -a.im};
    return r;
}

/* ------------------------------------------------------------------------- */
/* sin/cos: Cody-Waite reduction by pi/2 in three parts (the first step exact
 * under fma), then Taylor polynomials through x^19 (sin) and x^20 (cos) on
 * |y| <= pi/4. Coefficients are (-1)^k/(2k+1)! and (-1)^k/(2k)! rounded to
 * double. Same method as the device code, written independently here. */
#define ORC_TWO_OVER_PI 0x1.45f306dc9c883p-1
#define ORC_PIO2_HI 0x1.921fb54442d18p+0
#define ORC_PIO2_MID 0x1.1a62633145c07p-54
#define ORC_PIO2_LO -0x1.f1976b7ed8fbcp-110
#define ORC_FOURPI 0x1.921fb54442d18p+3

void orc_sincos(double x, double* s_out, double* c_out) {
    double q = rint(x * ORC_TWO_OVER_PI);
    double y = fma(-q, ORC_PIO2_HI, x);
    y = fma(-q, ORC_PIO2_MID, y);
    y = fma(-q, ORC_PIO2_LO, y);
    double z = y * y;
    double ps = -0x1.2f49b46814157p-57;
    ps = fma(ps, z, 0x1.952c77030ad4ap-49);
    ps = fma(ps, z, -0x1.ae7f3e733b81fp-41);
    ps = fma(ps, z, 0x1.6124613a86d09p-33);
    ps = fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = fma(ps, z, 0x1.71de3a556c734p-19);
    ps = fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = fma(ps, z, 0x1.1111111111111p-7);
    ps = fma(ps, z, -0x1.5555555555555p-3);
    double sn = fma(y * z, ps, y);
    double pc = 0x1.e542ba4020225p-62;
    pc = fma(pc, z, -0x1.6827863b97d97p-53);
    pc = fma(pc, z, 0x1.ae7f3e733b81fp-45);
    pc = fma(pc, z, -0x1.93974a8c07c9dp-37);
    pc = fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = fma(pc, z, 0x1.5555555555555p-5);
    double cs = fma(z * z, pc, fma(-0.5, z, 1.0));
    long long n = (long long)q;
    switch (n & 3) {
        case 0: *s_out = sn; *c_out = cs; break;
        case 1: *s_out = cs; *c_out = -sn; break;
        case 2: *s_out = -sn; *c_out = -cs; break;
        default: *s_out = -cs; *c_out = sn; break;
    }
}

/* Helmholtz point kernel cis(sign*k*r)/(4*pi*r), zero at r == 0
 * (README.md:104-108; test/nca/test_petrovgalerkinnca.jl:14-17). */
static inline cplx entry(double k, int32_t sign, const double* x, const double* y) {
    double dx = x[0] - y[0], dy = x[1] - y[1], dz = x[2] - y[2];
    double r2 = dx * dx + dy * dy;
    r2 = r2 + dz * dz;
    cplx g = {0.0, 0.0};
    if (r2 == 0.0) return g;
    double r = sqrt(r2);
    double s, c;
    orc_sincos(k * r, &s, &c);
    double d = ORC_FOURPI * r;
    g.re = c / d;
    g.im = s / d;
    if (sign < 0) g.im = -g.im;
    return g;
}

void orc_entry(double k, int32_t sign, const double* x, const double* y, double* g) {
    cplx v = entry(k, sign, x, y);
    g[0] = v.re;
    g[1] = v.im;
}

/* ------------------------------------------------------------------------- */
/* Phase 2: assemble_level (batchedbottomup.jl:101-116) into the column-padded
 * level buffer of the device path (distributedblockstack.jl:7-13,80-86). With
 * transposed != 0 block b is transpose(A[colidx_b, rowidx_b]) without
 * conjugation (batchedbottomup.jl:98-99,110-111). */
int32_t orc_assemble_level(double k, int32_t sign, const double* testpts, const double* trialpts,
                           int64_t nb, const int64_t* rowptr, const int64_t* rowidx,
                           const int64_t* colptr, const int64_t* colidx, int64_t maxrank,
                           int32_t transposed, double* buffer) {
    int64_t ld = rowptr[nb] - rowptr[0];
    cplx* buf = (cplx*)buffer;
    memset(buf, 0, sizeof(cplx) * (size_t)(ld * maxrank));
    for (int64_t b = 0; b < nb; ++b)
        if (colptr[b + 1] - colptr[b] > maxrank) return 1;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t b = 0; b < nb; ++b) {
        int64_t r0 = rowptr[b] - rowptr[0];
        int64_t m = rowptr[b + 1] - rowptr[b];
        int64_t n = colptr[b + 1] - colptr[b];
        for (int64_t j = 0; j < n; ++j) {
            int64_t cj = colidx[colptr[b] - 1 + j] - 1;
            for (int64_t a = 0; a < m; ++a) {
                int64_t ra = rowidx[rowptr[b] - 1 + a] - 1;
                buf[r0 + a + j * ld] = transposed ? entry(k, sign, testpts + 3 * cj, trialpts + 3 * ra)
                                                  : entry(k, sign, testpts + 3 * ra, trialpts + 3 * cj);
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Fully pivoted ACA with the sampled stop (batchedbottomup.jl:118-125,136-138;
 * call site compresslevel.jl:34-36; the arithmetic lives in the absent
 * AdaptiveCrossApproximation.jl 0.4 and is restated here):
 *   r accepted crosses, residual R_r, s = number of block columns;
 *   stop when r == min(maxrank, m, n), or when
 *       ||R_r||_F^2 * (sampled ? s/(s-r) : 1) <= tol^2 * ||A||_F^2,
 *   else pivot (i, j) = first argmax |R_r(i,j)|^2 in column-major order,
 *   R_{r+1} = R_r - R_r(:,j) * (R_r(i,:) / R_r(i,j)), row i and column j set
 *   to exactly zero.
 * Squared norms are summed in a fixed order: 256 partial sums, entry e of the
 * column-major block into slot e mod 256 in increasing e, then a halving tree
 * (slot t += slot t+s for s = 128, 64, ..., 1). */
#define ORC_SLOTS 256

static double sum_slots(double* p) {
    for (int s = ORC_SLOTS / 2; s > 0; s >>= 1)
        for (int t = 0; t < s; ++t) p[t] = p[t] + p[t + s];
    return p[0];
}

int64_t orc_aca_fullpiv(double* block, int64_t m, int64_t n, int64_t ld, int64_t maxrank, double tol,
                        int32_t sampled, int64_t* rowpiv, int64_t* colpiv) {
    cplx* B = (cplx*)block;
    int64_t kmax = maxrank;
    if (m < kmax) kmax = m;
    if (n < kmax) kmax = n;
    if (kmax <= 0) return 0;
    double p[ORC_SLOTS];
    memset(p, 0, sizeof(p));
    double best = -1.0;
    int64_t bi = 0;
    for (int64_t b = 0; b < n; ++b)
        for (int64_t a = 0; a < m; ++a) {
            int64_t e = a + b * m;
            double v = cabs2(B[a + b * ld]);
            p[e & (ORC_SLOTS - 1)] += v;
            if (v > best) {
                best = v;
                bi = e;
            }
        }
    const double nrmA2 = sum_slots(p);
    double nrm2 = nrmA2;
    const double tol2 = tol * tol;
    cplx* u = (cplx*)malloc(sizeof(cplx) * (size_t)m);
    cplx* v = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    int64_t r = 0;
    while (r < kmax) {
        double scale = sampled ? (double)n / (double)(n - r) : 1.0;
        if (nrm2 * scale <= tol2 * nrmA2) break;
        if (!(best > 0.0)) break;
        int64_t i = bi % m, j = bi / m;
        cplx piv = B[i + j * ld];
        double d = cabs2(piv);
        cplx inv = {piv.re / d, -piv.im / d};
        for (int64_t a = 0; a < m; ++a) u[a] = B[a + j * ld];
        for (int64_t c = 0; c < n; ++c) v[c] = cmul(B[i + c * ld], inv);
        rowpiv[r] = i + 1;
        colpiv[r] = j + 1;
        ++r;
        memset(p, 0, sizeof(p));
        best = -1.0;
        bi = 0;
        for (int64_t c = 0; c < n; ++c)
            for (int64_t a = 0; a < m; ++a) {
                int64_t e = a + c * m;
                cplx val;
                if (a == i || c == j) {
                    val.re = 0.0;
                    val.im = 0.0;
                } else {
                    val = csub(B[a + c * ld], cmul(u[a], v[c]));
                }
                B[a + c * ld] = val;
                double w = cabs2(val);
                p[e & (ORC_SLOTS - 1)] += w;
                if (w > best) {
                    best = w;
                    bi = e;
                }
            }
        nrm2 = sum_slots(p);
    }
    free(u);
    free(v);
    return r;
}

/* Incomplete ACA with geometric column pivoting (IACA; Tetzner and Adrian,
 * "The Incomplete Adaptive Cross Approximation ...", README.md:193 ref. [4];
 * the glue is pivots.jl:215-364 with IACA(MaximumValue(), TreeMimicryPivoting,
 * FNormExtrapolator) as BottomUp uses it, bottomup.jl:214-424). The block's
 * columns are the far-field candidates in their geometric order, so step r
 * takes column r; its residual column gives the row pivot (first maximum of
 * |R(a, r)|^2). Stop, not counting the candidate, at r = min(maxrank, m, n),
 * at a zero residual column, or when
 *   r ||R_r(:, r)||^2 <= tol^2 sum_{l<r} ||A(:, l)||^2          (r > 0).
 * AdaptiveCrossApproximation.jl (its IACA, TreeMimicryPivoting and
 * FNormExtrapolator) is not vendored, so this criterion is a restatement of
 * the method, not of that package's code. Residual update and sums as in
 * orc_aca_fullpiv (R -= u v, pivot row and column set to zero; column sums in
 * the slots e mod 256 of the column-major entries, halving tree). Returns the
 * rank; rowpiv/colpiv 1-based; block is overwritten. */
int64_t orc_iaca(double* block, int64_t m, int64_t n, int64_t ld, int64_t maxrank, double tol, int64_t* rowpiv,
                 int64_t* colpiv) {
    cplx* B = (cplx*)block;
    int64_t kmax = maxrank;
    if (m < kmax) kmax = m;
    if (n < kmax) kmax = n;
    if (kmax <= 0) return 0;
    cplx* A0 = (cplx*)malloc(sizeof(cplx) * (size_t)(m * n > 0 ? m * n : 1));
    for (int64_t c = 0; c < n; ++c)
        for (int64_t a = 0; a < m; ++a) A0[a + c * m] = B[a + c * ld];
    double p[ORC_SLOTS];
    memset(p, 0, sizeof(p));
    double best = -1.0;
    int64_t bi = 0;
    for (int64_t a = 0; a < m; ++a) {
        double w = cabs2(B[a]);
        p[a & (ORC_SLOTS - 1)] += w;
        if (w > best) {
            best = w;
            bi = a;
        }
    }
    double cn2 = sum_slots(p), orig2 = cn2, acc = 0.0;
    const double tol2 = tol * tol;
    cplx* u = (cplx*)malloc(sizeof(cplx) * (size_t)m);
    cplx* v = (cplx*)malloc(sizeof(cplx) * (size_t)n);
    int64_t r = 0;
    while (r < kmax) {
        if (r > 0 && cn2 * (double)r <= tol2 * acc) break;
        if (!(best > 0.0)) break;
        int64_t i = bi, j = r;
        cplx piv = B[i + j * ld];
        double d = cabs2(piv);
        cplx inv = {piv.re / d, -piv.im / d};
        for (int64_t a = 0; a < m; ++a) u[a] = B[a + j * ld];
        for (int64_t c = 0; c < n; ++c) v[c] = cmul(B[i + c * ld], inv);
        rowpiv[r] = i + 1;
        colpiv[r] = j + 1;
        acc += orig2;
        ++r;
        double q[ORC_SLOTS];
        memset(p, 0, sizeof(p));
        memset(q, 0, sizeof(q));
        best = -1.0;
        bi = 0;
        for (int64_t c = 0; c < n; ++c)
            for (int64_t a = 0; a < m; ++a) {
                int64_t e = a + c * m;
                cplx val;
                if (a == i || c == j) {
                    val.re = 0.0;
                    val.im = 0.0;
                } else {
                    val = csub(B[a + c * ld], cmul(u[a], v[c]));
                }
                B[a + c * ld] = val;
                if (c == r) {
                    double w = cabs2(val);
                    p[e & (ORC_SLOTS - 1)] += w;
                    if (w > best) {
                        best = w;
                        bi = a;
                    }
                    q[e & (ORC_SLOTS - 1)] += cabs2(A0[a + c * m]);
                }
            }
        cn2 = sum_slots(p);
        orig2 = sum_slots(q);
    }
    free(u);
    free(v);
    free(A0);
    return r;
}

/* ------------------------------------------------------------------------- */
/* Skeleton solve X = L * P^{-1}, P = L[rowpiv, :] (r x r), as nestedtestbasis /
 * testtransfer do with rdiv! (nestedbasis.jl:43-49, transfermatrices.jl:183-187)
 * and as the device path does for both sides (compresslevel.jl:14-25).
 * LU with partial pivoting in LAPACK getrf form (pivot = first max |re|+|im|),
 * then per row of L: w U = l (forward), y L = w (backward, unit diagonal),
 * x = y with the interchanges undone from the last to the first.
 * Singularity rule of _crossfactorization (utils.jl:57-80). */
int64_t orc_skeleton_solve(const double* Lin, int64_t m, int64_t r, const int64_t* rowpiv, double* Xout) {
    if (r == 0) return -1;
    const cplx* L = (const cplx*)Lin;
    cplx* X = (cplx*)Xout;
    cplx* A = (cplx*)malloc(sizeof(cplx) * (size_t)(r * r));
    int64_t* ipiv = (int64_t*)malloc(sizeof(int64_t) * (size_t)r);
    cplx* w = (cplx*)malloc(sizeof(cplx) * (size_t)r);
    for (int64_t c = 0; c < r; ++c)
        for (int64_t a = 0; a < r; ++a) A[a + c * r] = L[(rowpiv[a] - 1) + c * m];
    int64_t info = 0;
    for (int64_t k = 0; k < r; ++k) {
        int64_t p = k;
        double bestv = cabs1(A[k + k * r]);
        for (int64_t i = k + 1; i < r; ++i) {
            double v = cabs1(A[i + k * r]);
            if (v > bestv) {
                bestv = v;
                p = i;
            }
        }
        ipiv[k] = p;
        if (bestv != 0.0) {
            if (p != k)
                for (int64_t c = 0; c < r; ++c) {
                    cplx t = A[k + c * r];
                    A[k + c * r] = A[p + c * r];
                    A[p + c * r] = t;
                }
            for (int64_t i = k + 1; i < r; ++i) A[i + k * r] = cdiv(A[i + k * r], A[k + k * r]);
        } else if (info == 0) {
            info = k + 1;
        }
        for (int64_t c = k + 1; c < r; ++c)
            for (int64_t i = k + 1; i < r; ++i) A[i + c * r] = csub(A[i + c * r], cmul(A[i + k * r], A[k + c * r]));
    }
    double d2max = 0.0, d2min = INFINITY;
    for (int64_t i = 0; i < r; ++i) {
        double d2 = cabs2(A[i + i * r]);
        if (d2 > d2max) d2max = d2;
        if (d2 < d2min) d2min = d2;
    }
    double tol2 = (10.0 * DBL_EPSILON) * (10.0 * DBL_EPSILON) * d2max;
    int64_t result = -1;
    if (!(info == 0 && d2min > tol2)) {
        int64_t rank = 0;
        for (int64_t i = 0; i < r; ++i) rank += cabs2(A[i + i * r]) > tol2;
        result = rank;
    }
    for (int64_t a = 0; a < m; ++a) {
        for (int64_t c = 0; c < r; ++c) {
            cplx acc = L[a + c * m];
            for (int64_t k = 0; k < c; ++k) acc = csub(acc, cmul(w[k], A[k + c * r]));
            w[c] = cdiv(acc, A[c + c * r]);
        }
        for (int64_t c = r - 1; c >= 0; --c) {
            cplx acc = w[c];
            for (int64_t k = c + 1; k < r; ++k) acc = csub(acc, cmul(w[k], A[k + c * r]));
            w[c] = acc;
        }
        for (int64_t k = r - 1; k >= 0; --k) {
            int64_t p = ipiv[k];
            if (p != k) {
                cplx t = w[k];
                w[k] = w[p];
                w[p] = t;
            }
        }
        for (int64_t c = 0; c < r; ++c) X[a + c * m] = w[c];
    }
    free(A);
    free(ipiv);
    free(w);
    return result;
}

/* ------------------------------------------------------------------------- */
/* testbases / trialbases (::BatchedBottomUp) (batchedbottomup.jl:246-447):
 * levels from the deepest to the root; active clusters have far pivots; rows
 * of a leaf block are its points, of an inner block the concatenated pivots of
 * its children (:282-303); Phase 3 per block (compresslevel_test!/trial!
 * :145-244). side 0 = test, 1 = trial (transposed assembly, :408-410). */
typedef struct {
    int64_t rank, m;
    int64_t* piv; /* global 1-based point ids, rank entries */
    cplx* X;      /* m x rank */
} node_result;

int64_t orc_bases(int32_t side, double k, int32_t sign, const double* testpts, const double* trialpts,
                  int64_t nn, const int64_t* perm, const int64_t* nodelo, const int64_t* nodehi,
                  const int64_t* level, const int64_t* childptr, const int64_t* children,
                  const int64_t* farpivptr, const int64_t* farpiv, int64_t maxrank, double tol,
                  int32_t sampled, int64_t* pivptr, int64_t* pivots, int64_t* xoffsets, double* Xout,
                  int64_t xcap) {
    node_result* res = (node_result*)calloc((size_t)nn, sizeof(node_result));
    int64_t maxlev = 0, minlev = INT64_MAX;
    for (int64_t t = 0; t < nn; ++t) {
        res[t].rank = -1;
        if (level[t] > maxlev) maxlev = level[t];
        if (level[t] < minlev) minlev = level[t];
    }
    int64_t singular = 0;
    int64_t* active = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nn > 0 ? nn : 1));
    for (int64_t lev = maxlev; lev >= minlev && !singular; --lev) {
        int64_t na = 0;
        for (int64_t t = 0; t < nn; ++t)
            if (level[t] == lev && farpivptr[t + 1] > farpivptr[t]) active[na++] = t;
#pragma omp parallel for schedule(dynamic, 4)
        for (int64_t q = 0; q < na; ++q) {
            int64_t t = active[q];
            int leaf = childptr[t + 1] == childptr[t];
            int64_t m = 0;
            if (leaf) {
                m = nodehi[t] - nodelo[t] + 1;
            } else {
                for (int64_t c = childptr[t] - 1; c < childptr[t + 1] - 1; ++c) {
                    int64_t ch = children[c] - 1;
                    if (res[ch].rank > 0) m += res[ch].rank;
                }
            }
            int64_t* rows = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m > 0 ? m : 1));
            if (leaf) {
                for (int64_t a = 0; a < m; ++a) rows[a] = perm[nodelo[t] - 1 + a];
            } else {
                int64_t o = 0;
                for (int64_t c = childptr[t] - 1; c < childptr[t + 1] - 1; ++c) {
                    int64_t ch = children[c] - 1;
                    for (int64_t a = 0; a < res[ch].rank; ++a) rows[o++] = res[ch].piv[a];
                }
            }
            int64_t n = farpivptr[t + 1] - farpivptr[t];
            const int64_t* cols = farpiv + farpivptr[t] - 1;
            cplx* blk = (cplx*)malloc(sizeof(cplx) * (size_t)(m * n > 0 ? m * n : 1));
            cplx* work = (cplx*)malloc(sizeof(cplx) * (size_t)(m * n > 0 ? m * n : 1));
            for (int64_t j = 0; j < n; ++j)
                for (int64_t a = 0; a < m; ++a) {
                    const double* rp = (side == 0 ? testpts : trialpts) + 3 * (rows[a] - 1);
                    const double* cp = (side == 0 ? trialpts : testpts) + 3 * (cols[j] - 1);
                    blk[a + j * m] = side == 0 ? entry(k, sign, rp, cp) : entry(k, sign, cp, rp);
                }
            memcpy(work, blk, sizeof(cplx) * (size_t)(m * n));
            int64_t* rpv = (int64_t*)malloc(sizeof(int64_t) * (size_t)(maxrank > 0 ? maxrank : 1));
            int64_t* cpv = (int64_t*)malloc(sizeof(int64_t) * (size_t)(maxrank > 0 ? maxrank : 1));
            int64_t r = sampled == 2 ? orc_iaca((double*)work, m, n, m, maxrank, tol, rpv, cpv)
                                     : orc_aca_fullpiv((double*)work, m, n, m, maxrank, tol, sampled, rpv, cpv);
            cplx* Lm = (cplx*)malloc(sizeof(cplx) * (size_t)(m * r > 0 ? m * r : 1));
            for (int64_t c = 0; c < r; ++c)
                for (int64_t a = 0; a < m; ++a) Lm[a + c * m] = blk[a + (cpv[c] - 1) * m];
            cplx* X = (cplx*)malloc(sizeof(cplx) * (size_t)(m * r > 0 ? m * r : 1));
            int64_t st = orc_skeleton_solve((const double*)Lm, m, r, rpv, (double*)X);
            if (st >= 0) {
#pragma omp critical
                { if (!singular) singular = t + 1; }
            }
            res[t].rank = r;
            res[t].m = m;
            res[t].piv = (int64_t*)malloc(sizeof(int64_t) * (size_t)(r > 0 ? r : 1));
            for (int64_t a = 0; a < r; ++a) res[t].piv[a] = rows[rpv[a] - 1];
            res[t].X = X;
            free(rows);
            free(blk);
            free(work);
            free(rpv);
            free(cpv);
            free(Lm);
        }
    }
    free(active);
    int64_t ret = singular ? -singular : 0;
    pivptr[0] = 1;
    xoffsets[0] = 0;
    for (int64_t t = 0; t < nn; ++t) {
        int64_t r = res[t].rank > 0 ? res[t].rank : 0;
        pivptr[t + 1] = pivptr[t] + r;
        xoffsets[t + 1] = xoffsets[t] + (res[t].rank > 0 ? res[t].m * r : 0);
    }
    if (xoffsets[nn] > xcap && ret == 0) ret = 1;
    for (int64_t t = 0; t < nn; ++t) {
        int64_t r = res[t].rank > 0 ? res[t].rank : 0;
        for (int64_t a = 0; a < r; ++a) pivots[pivptr[t] - 1 + a] = res[t].piv[a];
        if (ret == 0 && r > 0)
            memcpy((cplx*)Xout + xoffsets[t], res[t].X, sizeof(cplx) * (size_t)(res[t].m * r));
        free(res[t].piv);
        free(res[t].X);
    }
    /* nodes with rank 0 but far pivots still count as assigned; the caller
     * reads the ranks from pivptr */
    free(res);
    return ret;
}

/* ------------------------------------------------------------------------- */
/* Partially pivoted ACA (the ACA() factorization of TopDown,
 * src/compressors/topdown.jl:51-127 through compute_test_pivots!
 * src/pivoting/pivots.jl:215-257; AdaptiveCrossApproximation.jl is not
 * vendored, so this is the published method restated: Bebendorf's adaptive
 * cross approximation with pivots by maximum value and the Frobenius-norm
 * stop). Entries are evaluated, never assembled: entry (a, c) of the block is
 * G(row a, col c), or G(col c, row a) with transposed != 0.
 *   row pivot i_1 = row 0; step r: residual row R(i_r, :) = A(i_r, :) -
 *   sum_l U(i_r, l) V(:, l); column pivot j_r = first maximum of |R(i_r, c)|^2
 *   over unused columns (a zero maximum stops); V(:, r) = R(i_r, :) / R(i_r, j_r);
 *   U(:, r) = A(:, j_r) - sum_l V(j_r, l) U(:, l); S2 += |u|^2 |v|^2 +
 *   2 Re sum_{l<r} (u_l^H u_r)(v_r v_l^H) (the squared Frobenius norm of the
 *   approximant); the cross counts, and the factorization stops when
 *   |u|^2 |v|^2 <= tol^2 S2 or r = min(maxrank, m, n); else the next row pivot is
 *   the first maximum of |U(a, r)|^2 over unused rows.
 * Sums over rows or columns use the fixed slot order (index mod 256, halving
 * tree); sums over l run in increasing l. rowpiv/colpiv: 1-based positions.
 * U (m x kmax) and V (n x kmax) are written when not NULL. */
int64_t orc_paca(double k, int32_t sign, int32_t transposed, const double* rowpts, const double* colpts,
                 int64_t m, const int64_t* rows, int64_t n, const int64_t* cols, int64_t maxrank, double tol,
                 int64_t* rowpiv, int64_t* colpiv, double* Uout, double* Vout) {
    int64_t kmax = maxrank;
    if (m < kmax) kmax = m;
    if (n < kmax) kmax = n;
    if (kmax <= 0) return 0;
    cplx* U = (cplx*)malloc(sizeof(cplx) * (size_t)(m * kmax));
    cplx* V = (cplx*)malloc(sizeof(cplx) * (size_t)(n * kmax));
    char* usedr = (char*)calloc((size_t)m, 1);
    char* usedc = (char*)calloc((size_t)n, 1);
    double slots[ORC_SLOTS], slots2[ORC_SLOTS];
#define ORC_G(a, c)                                                                          \
    (transposed ? entry(k, sign, colpts + 3 * (cols[c] - 1), rowpts + 3 * (rows[a] - 1))     \
                : entry(k, sign, rowpts + 3 * (rows[a] - 1), colpts + 3 * (cols[c] - 1)))
    const double tol2 = tol * tol;
    double S2 = 0.0;
    int64_t i = 0, r = 0;
    while (r < kmax) {
        for (int64_t c = 0; c < n; ++c) {
            cplx acc = ORC_G(i, c);
            for (int64_t l = 0; l < r; ++l) acc = csub(acc, cmul(U[i + l * m], V[c + l * n]));
            V[c + r * n] = acc;
        }
        double bv = -1.0;
        int64_t j = -1;
        for (int64_t c = 0; c < n; ++c)
            if (!usedc[c]) {
                double w = cabs2(V[c + r * n]);
                if (w > bv) {
                    bv = w;
                    j = c;
                }
            }
        if (!(bv > 0.0)) break;
        cplx pv = V[j + r * n];
        double d = cabs2(pv);
        cplx inv = {pv.re / d, -pv.im / d};
        for (int64_t c = 0; c < n; ++c) V[c + r * n] = cmul(V[c + r * n], inv);
        for (int64_t a = 0; a < m; ++a) {
            cplx acc = ORC_G(a, j);
            for (int64_t l = 0; l < r; ++l) acc = csub(acc, cmul(V[j + l * n], U[a + l * m]));
            U[a + r * m] = acc;
        }
        memset(slots, 0, sizeof(slots));
        for (int64_t a = 0; a < m; ++a) slots[a % ORC_SLOTS] += cabs2(U[a + r * m]);
        double nu2 = sum_slots(slots);
        memset(slots, 0, sizeof(slots));
        for (int64_t c = 0; c < n; ++c) slots[c % ORC_SLOTS] += cabs2(V[c + r * n]);
        double nv2 = sum_slots(slots);
        double cross = 0.0;
        for (int64_t l = 0; l < r; ++l) {
            memset(slots, 0, sizeof(slots));
            memset(slots2, 0, sizeof(slots2));
            for (int64_t a = 0; a < m; ++a) {
                cplx q = cmul(cconj(U[a + l * m]), U[a + r * m]);
                slots[a % ORC_SLOTS] += q.re;
                slots2[a % ORC_SLOTS] += q.im;
            }
            cplx du = {sum_slots(slots), sum_slots(slots2)};
            memset(slots, 0, sizeof(slots));
            memset(slots2, 0, sizeof(slots2));
            for (int64_t c = 0; c < n; ++c) {
                cplx q = cmul(V[c + r * n], cconj(V[c + l * n]));
                slots[c % ORC_SLOTS] += q.re;
                slots2[c % ORC_SLOTS] += q.im;
            }
            cplx dv = {sum_slots(slots), sum_slots(slots2)};
            cross = cross + (du.re * dv.re - du.im * dv.im);
        }
        S2 = (S2 + nu2 * nv2) + 2.0 * cross;
        rowpiv[r] = i + 1;
        colpiv[r] = j + 1;
        usedr[i] = 1;
        usedc[j] = 1;
        ++r;
        if (nu2 * nv2 <= tol2 * S2) break;
        bv = -1.0;
        i = -1;
        for (int64_t a = 0; a < m; ++a)
            if (!usedr[a]) {
                double w = cabs2(U[a + (r - 1) * m]);
                if (w > bv) {
                    bv = w;
                    i = a;
                }
            }
        if (i < 0) break;
    }
#undef ORC_G
    if (Uout) memcpy(Uout, U, sizeof(cplx) * (size_t)(m * r));
    if (Vout)
        for (int64_t l = 0; l < r; ++l) memcpy((cplx*)Vout + l * n, V + l * n, sizeof(cplx) * (size_t)n);
    free(U);
    free(V);
    free(usedr);
    free(usedc);
    return r;
}

/* testbases/trialbases(::TopDown) (topdown.jl:51-206): levels from the root
 * down; a cluster's columns are the points of its far nodes (CSR order) and
 * then the column pivots of its parent; rows are its own points; the
 * partially pivoted ACA picks the row pivots (the cluster's pivots) and the
 * column pivots (handed to the children). Then the factor of every cluster
 * with pivots, X = A[rows, sigma] A[tau, sigma]^{-1} with rows = the points of a
 * leaf (nestedtestbasis, nestedbasis.jl:43-57) or the children's pivots in
 * child order (the stacked transfers A[tau_c, sigma] A[tau, sigma]^{-1},
 * testtransfer transfermatrices.jl:126-130), solved as orc_skeleton_solve on
 * [A[rows, sigma]; A[tau, sigma]] with the pivots in the appended rows.
 * side 0 = test (rows: test points), 1 = trial (rows: trial points, entries
 * transposed). Self tree layout as orc_bases; other tree: operm/olo/ohi;
 * farptr/farnodes: 1-based CSR of the far nodes (of the other tree) of every
 * node. Outputs: pivptr/pivots (row pivots, 1-based point ids; capacity
 * nn*maxrank), cpivptr/cpivots (column pivots), xoffsets/X as orc_bases.
 * Returns 0, -(node) of the first singular cross matrix, or 1 when X is too small. */
int64_t orc_topdown(int32_t side, double k, int32_t sign, const double* testpts, const double* trialpts,
                    int64_t nn, const int64_t* perm, const int64_t* nodelo, const int64_t* nodehi,
                    const int64_t* level, const int64_t* parent, const int64_t* childptr, const int64_t* children,
                    const int64_t* operm, const int64_t* olo, const int64_t* ohi, const int64_t* farptr,
                    const int64_t* farnodes, int64_t maxrank, double tol, int64_t* pivptr, int64_t* pivots,
                    int64_t* cpivptr, int64_t* cpivots, int64_t* xoffsets, double* Xout, int64_t xcap) {
    const double* selfpts = side == 0 ? testpts : trialpts;
    const double* otherpts = side == 0 ? trialpts : testpts;
    int64_t* rank = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nn > 0 ? nn : 1));
    int64_t* rp = (int64_t*)calloc((size_t)(nn * maxrank > 0 ? nn * maxrank : 1), sizeof(int64_t));
    int64_t* cp = (int64_t*)calloc((size_t)(nn * maxrank > 0 ? nn * maxrank : 1), sizeof(int64_t));
    int64_t maxlev = 0, minlev = INT64_MAX;
    for (int64_t t = 0; t < nn; ++t) {
        rank[t] = -1;
        if (level[t] > maxlev) maxlev = level[t];
        if (level[t] < minlev) minlev = level[t];
    }
    for (int64_t lev = minlev; lev <= maxlev; ++lev) {
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t t = 0; t < nn; ++t) {
            if (level[t] != lev) continue;
            const int64_t pa = parent[t] - 1;
            const int64_t rpa = pa >= 0 && rank[pa] > 0 ? rank[pa] : 0;
            int64_t n = rpa;
            for (int64_t q = farptr[t] - 1; q < farptr[t + 1] - 1; ++q) {
                const int64_t s = farnodes[q] - 1;
                n += ohi[s] - olo[s] + 1;
            }
            if (n == 0) continue;
            int64_t* cols = (int64_t*)malloc(sizeof(int64_t) * (size_t)n);
            int64_t o = 0;
            for (int64_t q = farptr[t] - 1; q < farptr[t + 1] - 1; ++q) {
                const int64_t s = farnodes[q] - 1;
                for (int64_t a = olo[s] - 1; a < ohi[s]; ++a) cols[o++] = operm[a];
            }
            for (int64_t a = 0; a < rpa; ++a) cols[o++] = cp[pa * maxrank + a];
            const int64_t m = nodehi[t] - nodelo[t] + 1;
            const int64_t* rows = perm + nodelo[t] - 1;
            int64_t rpv[1024], cpv[1024];
            int64_t r = orc_paca(k, sign, side, selfpts, otherpts, m, rows, n, cols, maxrank > 1024 ? 1024 : maxrank,
                                 tol, rpv, cpv, NULL, NULL);
            for (int64_t a = 0; a < r; ++a) {
                rp[t * maxrank + a] = rows[rpv[a] - 1];
                cp[t * maxrank + a] = cols[cpv[a] - 1];
            }
            rank[t] = r;
            free(cols);
        }
    }
    /* factors */
    int64_t ret = 0;
    xoffsets[0] = 0;
    pivptr[0] = 1;
    cpivptr[0] = 1;
    for (int64_t t = 0; t < nn; ++t) {
        const int64_t r = rank[t] > 0 ? rank[t] : 0;
        int64_t m = 0;
        if (rank[t] >= 0) {
            if (childptr[t + 1] == childptr[t]) m = nodehi[t] - nodelo[t] + 1;
            else
                for (int64_t c = childptr[t] - 1; c < childptr[t + 1] - 1; ++c) {
                    const int64_t ch = children[c] - 1;
                    if (rank[ch] > 0) m += rank[ch];
                }
        }
        xoffsets[t + 1] = xoffsets[t] + m * r;
        pivptr[t + 1] = pivptr[t] + r;
        cpivptr[t + 1] = cpivptr[t] + r;
        for (int64_t a = 0; a < r; ++a) {
            pivots[pivptr[t] - 1 + a] = rp[t * maxrank + a];
            cpivots[cpivptr[t] - 1 + a] = cp[t * maxrank + a];
        }
    }
    if (xoffsets[nn] > xcap) ret = 1;
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t t = 0; t < nn; ++t) {
        const int64_t r = rank[t] > 0 ? rank[t] : 0;
        if (r == 0 || ret == 1) continue;
        const int64_t m = (xoffsets[t + 1] - xoffsets[t]) / r;
        int64_t* rows = (int64_t*)malloc(sizeof(int64_t) * (size_t)(m + r));
        if (childptr[t + 1] == childptr[t]) {
            for (int64_t a = 0; a < m; ++a) rows[a] = perm[nodelo[t] - 1 + a];
        } else {
            int64_t o = 0;
            for (int64_t c = childptr[t] - 1; c < childptr[t + 1] - 1; ++c) {
                const int64_t ch = children[c] - 1;
                for (int64_t a = 0; a < (rank[ch] > 0 ? rank[ch] : 0); ++a) rows[o++] = rp[ch * maxrank + a];
            }
        }
        for (int64_t a = 0; a < r; ++a) rows[m + a] = rp[t * maxrank + a];
        const int64_t me = m + r;
        cplx* L = (cplx*)malloc(sizeof(cplx) * (size_t)(me * r));
        for (int64_t c = 0; c < r; ++c)
            for (int64_t a = 0; a < me; ++a) {
                const double* x = selfpts + 3 * (rows[a] - 1);
                const double* y = otherpts + 3 * (cp[t * maxrank + c] - 1);
                L[a + c * me] = side == 0 ? entry(k, sign, x, y) : entry(k, sign, y, x);
            }
        int64_t* piv = (int64_t*)malloc(sizeof(int64_t) * (size_t)r);
        for (int64_t a = 0; a < r; ++a) piv[a] = m + a + 1;
        cplx* Xe = (cplx*)malloc(sizeof(cplx) * (size_t)(me * r));
        const int64_t st = orc_skeleton_solve((const double*)L, me, r, piv, (double*)Xe);
        if (st >= 0) {
#pragma omp critical
            { if (ret == 0 || -ret > t + 1) ret = -(t + 1); }
        }
        cplx* X = (cplx*)Xout + xoffsets[t];
        for (int64_t c = 0; c < r; ++c)
            for (int64_t a = 0; a < m; ++a) X[a + c * m] = Xe[a + c * me];
        free(rows);
        free(L);
        free(piv);
        free(Xe);
    }
    free(rank);
    free(rp);
    free(cp);
    return ret;
}

/* ------------------------------------------------------------------------- */
/* assemble_couplingstore (couplingmatrices.jl:11-53): block i of the test CSR
 * row of t is A[testpivots[t], trialpivots[farnodes[i]]]; empty sides give
 * empty blocks. */
void orc_coupling(double k, int32_t sign, const double* testpts, const double* trialpts, int64_t ntest,
                  const int64_t* farptr, const int64_t* farnodes, const int64_t* testpivptr,
                  const int64_t* testpiv, const int64_t* trialpivptr, const int64_t* trialpiv,
                  const int64_t* offsets, double* blocks) {
    cplx* out = (cplx*)blocks;
#pragma omp parallel for schedule(dynamic, 16)
    for (int64_t t = 0; t < ntest; ++t) {
        int64_t rt = testpivptr[t + 1] - testpivptr[t];
        for (int64_t i = farptr[t] - 1; i < farptr[t + 1] - 1; ++i) {
            int64_t s = farnodes[i] - 1;
            int64_t rs = trialpivptr[s + 1] - trialpivptr[s];
            cplx* S = out + offsets[i];
            for (int64_t j = 0; j < rs; ++j)
                for (int64_t a = 0; a < rt; ++a)
                    S[a + j * rt] = entry(k, sign, testpts + 3 * (testpiv[testpivptr[t] - 1 + a] - 1),
                                          trialpts + 3 * (trialpiv[trialpivptr[s] - 1 + j] - 1));
        }
    }
}

/* ------------------------------------------------------------------------- */
/* H^2 matvec (petrovgalerkinnca.jl:431-589): leaf projection, upward transfer,
 * coupling (forward: per test cluster; reverse: accumulated serially as
 * couple.jl:119-142 does), downward transfer, leaf expansion, near field
 * (evaluated here instead of stored). op 1/2 swap the sides and transpose
 * (and conjugate) every operator. */
static inline cplx ld_c(const double* base, int64_t idx, int conj) {
    cplx v = {base[2 * idx], base[2 * idx + 1]};
    if (conj) v.im = -v.im;
    return v;
}

/* Stored near field (see nca_oracle.h). */
struct orc_near {
    int64_t nb;          /* test clusters with near interactions */
    int64_t* node;       /* their 0-based ids */
    int64_t* rowptr;     /* rows of block b: test positions [rowptr[b], rowptr[b+1]) of values(node) */
    int64_t* colptr;     /* columns of block b: colidx[colptr[b] .. colptr[b+1]) */
    int64_t* colidx;     /* 0-based trial point ids */
    int64_t* boff;       /* block b starts at re + boff[b] (and im + boff[b]) */
    double *re, *im;
    int64_t entries;
};

orc_near* orc_near_assemble(const orc_h2* h) {
    orc_near* s = (orc_near*)calloc(1, sizeof(orc_near));
    int64_t nb = 0, ncol = 0, nent = 0;
    for (int64_t t = 0; t < h->nn_test; ++t)
        if (h->nearptr[t + 1] > h->nearptr[t] && h->test_hi[t] >= h->test_lo[t]) ++nb;
    s->nb = nb;
    s->node = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb + 1));
    s->rowptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb + 1));
    s->colptr = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb + 1));
    s->boff = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nb + 1));
    nb = 0;
    s->colptr[0] = 0;
    s->boff[0] = 0;
    for (int64_t t = 0; t < h->nn_test; ++t) {
        if (!(h->nearptr[t + 1] > h->nearptr[t] && h->test_hi[t] >= h->test_lo[t])) continue;
        int64_t m = h->test_hi[t] - h->test_lo[t] + 1, n = 0;
        for (int64_t i = h->nearptr[t] - 1; i < h->nearptr[t + 1] - 1; ++i) {
            int64_t q = h->nearnodes[i] - 1;
            n += h->trial_hi[q] - h->trial_lo[q] + 1;
        }
        s->node[nb] = t;
        s->colptr[nb + 1] = s->colptr[nb] + n;
        s->boff[nb + 1] = s->boff[nb] + m * n;
        ++nb;
        ncol += n;
        nent += m * n;
    }
    s->entries = nent;
    s->colidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(ncol > 0 ? ncol : 1));
    s->re = (double*)malloc(sizeof(double) * (size_t)(nent > 0 ? nent : 1));
    s->im = (double*)malloc(sizeof(double) * (size_t)(nent > 0 ? nent : 1));
#pragma omp parallel for schedule(dynamic, 4)
    for (int64_t b = 0; b < s->nb; ++b) {
        int64_t t = s->node[b];
        int64_t m = h->test_hi[t] - h->test_lo[t] + 1;
        int64_t c = s->colptr[b];
        for (int64_t i = h->nearptr[t] - 1; i < h->nearptr[t + 1] - 1; ++i) {
            int64_t q = h->nearnodes[i] - 1;
            for (int64_t v = h->trial_lo[q] - 1; v < h->trial_hi[q]; ++v) s->colidx[c++] = h->trial_perm[v] - 1;
        }
        int64_t n = s->colptr[b + 1] - s->colptr[b];
        double* re = s->re + s->boff[b];
        double* im = s->im + s->boff[b];
        for (int64_t j = 0; j < n; ++j) {
            int64_t col = s->colidx[s->colptr[b] + j];
            for (int64_t a = 0; a < m; ++a) {
                int64_t row = h->test_perm[h->test_lo[t] - 1 + a] - 1;
                cplx g = entry(h->k, h->sign, h->testpts + 3 * row, h->trialpts + 3 * col);
                re[a + j * m] = g.re;
                im[a + j * m] = g.im;
            }
        }
    }
    return s;
}

int64_t orc_near_entries(const orc_near* s) { return s ? s->entries : 0; }

void orc_near_free(orc_near* s) {
    if (!s) return;
    free(s->node);
    free(s->rowptr);
    free(s->colptr);
    free(s->colidx);
    free(s->boff);
    free(s->re);
    free(s->im);
    free(s);
}

/* acc[row + q ldacc] += (stored block row of t) x for every stored cluster:
 * gather the block's x rows, multiply column by column (rows innermost), add
 * into the output rows. */
__attribute__((optimize("O3"))) static void near_apply(const orc_h2* h, const orc_near* s, int64_t nrhs,
                                                       const cplx* X, int64_t ldx, cplx* acc, int64_t ldacc) {
#pragma omp parallel
    {
        int64_t cap = 0;
        double *yr = NULL, *yi = NULL, *xr = NULL, *xi = NULL;
#pragma omp for schedule(dynamic, 4)
        for (int64_t b = 0; b < s->nb; ++b) {
            int64_t t = s->node[b];
            int64_t m = h->test_hi[t] - h->test_lo[t] + 1;
            int64_t n = s->colptr[b + 1] - s->colptr[b];
            if (m * nrhs > cap || n * nrhs > cap) {
                cap = (m > n ? m : n) * nrhs;
                yr = (double*)realloc(yr, sizeof(double) * (size_t)cap);
                yi = (double*)realloc(yi, sizeof(double) * (size_t)cap);
                xr = (double*)realloc(xr, sizeof(double) * (size_t)cap);
                xi = (double*)realloc(xi, sizeof(double) * (size_t)cap);
            }
            const int64_t* cols = s->colidx + s->colptr[b];
            for (int64_t j = 0; j < n; ++j)
                for (int64_t q = 0; q < nrhs; ++q) {
                    cplx v = X[cols[j] + q * ldx];
                    xr[q + j * nrhs] = v.re;
                    xi[q + j * nrhs] = v.im;
                }
            for (int64_t i = 0; i < m * nrhs; ++i) yr[i] = yi[i] = 0.0;
            const double* re = s->re + s->boff[b];
            const double* im = s->im + s->boff[b];
            for (int64_t j = 0; j < n; ++j) {
                const double* br = re + j * m;
                const double* bi = im + j * m;
                for (int64_t q = 0; q < nrhs; ++q) {
                    const double vr = xr[q + j * nrhs], vi = xi[q + j * nrhs];
                    double* outr = yr + q * m;
                    double* outi = yi + q * m;
                    for (int64_t a = 0; a < m; ++a) {
                        outr[a] += br[a] * vr - bi[a] * vi;
                        outi[a] += br[a] * vi + bi[a] * vr;
                    }
                }
            }
            for (int64_t a = 0; a < m; ++a) {
                int64_t row = h->test_perm[h->test_lo[t] - 1 + a] - 1;
                for (int64_t q = 0; q < nrhs; ++q) {
                    acc[row + q * ldacc].re += yr[a + q * m];
                    acc[row + q * ldacc].im += yi[a + q * m];
                }
            }
        }
        free(yr);
        free(yi);
        free(xr);
        free(xi);
    }
}

void orc_h2_matvec(const orc_h2* h, int32_t op, int64_t nrhs, const double* xin, int64_t ldx, double* yio,
                   int64_t ldy, const double* alpha, const double* beta, int32_t nearon, int64_t nrows,
                   const int64_t* rows) {
    const int fwd = op == 0;
    const int cj = op == 2;
    /* source side (x) and destination side (y) */
    const int64_t nsrc_pts = fwd ? h->N : h->M, ndst_pts = fwd ? h->M : h->N;
    const int64_t* src_perm = fwd ? h->trial_perm : h->test_perm;
    const int64_t* src_lo = fwd ? h->trial_lo : h->test_lo;
    const int64_t* src_hi = fwd ? h->trial_hi : h->test_hi;
    const int64_t* dst_perm = fwd ? h->test_perm : h->trial_perm;
    const int64_t* dst_lo = fwd ? h->test_lo : h->trial_lo;
    const int64_t* dst_hi = fwd ? h->test_hi : h->trial_hi;
    const int64_t* src_cptr = fwd ? h->trial_cptr : h->test_cptr;
    const int64_t* dst_cptr = fwd ? h->test_cptr : h->trial_cptr;
    const int64_t nn_src = fwd ? h->nn_trial : h->nn_test;
    const int64_t nn_dst = fwd ? h->nn_test : h->nn_trial;
    const int64_t Ks = src_cptr[nn_src] - 1, Kd = dst_cptr[nn_dst] - 1;
    (void)nsrc_pts;
    const cplx* X = (const cplx*)xin;
    cplx* Y = (cplx*)yio;
    cplx* xh = (cplx*)calloc((size_t)(Ks * nrhs + 1), sizeof(cplx));
    cplx* yh = (cplx*)calloc((size_t)(Kd * nrhs + 1), sizeof(cplx));
    cplx* acc = (cplx*)calloc((size_t)(ndst_pts * nrhs + 1), sizeof(cplx));
    char* want = NULL;
    if (nrows >= 0) {
        want = (char*)calloc((size_t)(ndst_pts + 1), 1);
        for (int64_t i = 0; i < nrows; ++i) want[rows[i]] = 1;
    }

    /* Stage 1: leaf projection (aggregate.jl:1-19). */
    {
        int64_t nb = fwd ? h->nbasis_trial : h->nbasis_test;
        const int64_t* nodes = fwd ? h->basis_trial_nodes : h->basis_test_nodes;
        const int64_t* off = fwd ? h->basis_trial_off : h->basis_test_off;
        const double* blocks = fwd ? h->basis_trial : h->basis_test;
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < nb; ++i) {
            int64_t s = nodes[i] - 1;
            int64_t r = src_cptr[s + 1] - src_cptr[s];
            int64_t ns = src_hi[s] - src_lo[s] + 1;
            cplx* sum = (cplx*)calloc((size_t)nrhs, sizeof(cplx));
            for (int64_t a = 0; a < r; ++a) {
                for (int64_t q = 0; q < nrhs; ++q) sum[q].re = sum[q].im = 0.0;
                for (int64_t j = 0; j < ns; ++j) {
                    /* forward: V (r x ns) as is; transposed: U (ns x r) transposed */
                    cplx b = ld_c(blocks, off[i] + (fwd ? a + j * r : j + a * ns), cj);
                    int64_t col = src_perm[src_lo[s] - 1 + j] - 1;
                    for (int64_t q = 0; q < nrhs; ++q) sum[q] = cadd(sum[q], cmul(b, X[col + q * ldx]));
                }
                for (int64_t q = 0; q < nrhs; ++q) xh[(src_cptr[s] - 1 + a) + q * Ks] = sum[q];
            }
            free(sum);
        }
    }
    /* Stage 2: upward transfer, deepest level first (aggregate.jl:44-67). */
    {
        int64_t nlev = fwd ? h->nlev_trial : h->nlev_test;
        const int64_t* levptr = fwd ? h->st_levptr : h->tt_levptr;
        const int64_t* levidcs = fwd ? h->st_levidcs : h->tt_levidcs;
        const int64_t* chdptr = fwd ? h->st_chdptr : h->tt_chdptr;
        const int64_t* chds = fwd ? h->st_chds : h->tt_chds;
        const int64_t* off = fwd ? h->st_off : h->tt_off;
        const double* blocks = fwd ? h->st_blocks : h->tt_blocks;
        for (int64_t lev = nlev - 1; lev >= 0; --lev) {
#pragma omp parallel for schedule(dynamic, 16)
            for (int64_t idx = levptr[lev] - 1; idx < levptr[lev + 1] - 1; ++idx) {
                int64_t p = levidcs[idx] - 1;
                int64_t rp = src_cptr[p + 1] - src_cptr[p];
                for (int64_t e = chdptr[idx] - 1; e < chdptr[idx + 1] - 1; ++e) {
                    int64_t c = chds[e] - 1;
                    int64_t rc = src_cptr[c + 1] - src_cptr[c];
                    for (int64_t a = 0; a < rp; ++a) {
                        cplx* out = xh + (src_cptr[p] - 1 + a);
                        for (int64_t j = 0; j < rc; ++j) {
                            /* forward: trial E (rp x rc); transposed: test E (rc x rp) transposed */
                            cplx b = ld_c(blocks, off[e] + (fwd ? a + j * rp : j + a * rc), cj);
                            const cplx* in = xh + (src_cptr[c] - 1 + j);
                            for (int64_t q = 0; q < nrhs; ++q) out[q * Ks] = cadd(out[q * Ks], cmul(b, in[q * Ks]));
                        }
                    }
                }
            }
        }
    }
    /* Stage 3: coupling (couple.jl:1-24 forward, :119-142 reverse). */
    if (fwd) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t t = 0; t < h->nn_test; ++t) {
            int64_t rt = h->test_cptr[t + 1] - h->test_cptr[t];
            for (int64_t i = h->cp_ptr[t] - 1; i < h->cp_ptr[t + 1] - 1; ++i) {
                int64_t s = h->cp_nodes[i] - 1;
                int64_t rs = h->trial_cptr[s + 1] - h->trial_cptr[s];
                for (int64_t a = 0; a < rt; ++a) {
                    cplx* out = yh + (h->test_cptr[t] - 1 + a);
                    for (int64_t j = 0; j < rs; ++j) {
                        cplx b = ld_c(h->cp_blocks, h->cp_off[i] + a + j * rt, cj);
                        const cplx* in = xh + (h->trial_cptr[s] - 1 + j);
                        for (int64_t q = 0; q < nrhs; ++q) out[q * Kd] = cadd(out[q * Kd], cmul(b, in[q * Ks]));
                    }
                }
            }
        }
    } else {
        for (int64_t t = 0; t < h->nn_test; ++t) {
            int64_t rt = h->test_cptr[t + 1] - h->test_cptr[t];
            for (int64_t i = h->cp_ptr[t] - 1; i < h->cp_ptr[t + 1] - 1; ++i) {
                int64_t s = h->cp_nodes[i] - 1;
                int64_t rs = h->trial_cptr[s + 1] - h->trial_cptr[s];
                for (int64_t a = 0; a < rs; ++a) {
                    cplx* out = yh + (h->trial_cptr[s] - 1 + a);
                    for (int64_t j = 0; j < rt; ++j) {
                        cplx b = ld_c(h->cp_blocks, h->cp_off[i] + j + a * rt, cj);
                        const cplx* in = xh + (h->test_cptr[t] - 1 + j);
                        for (int64_t q = 0; q < nrhs; ++q) out[q * Kd] = cadd(out[q * Kd], cmul(b, in[q * Ks]));
                    }
                }
            }
        }
    }
    /* Stage 4: downward transfer, root level first (disaggregate.jl:1-23). */
    {
        int64_t nlev = fwd ? h->nlev_test : h->nlev_trial;
        const int64_t* levptr = fwd ? h->tt_levptr : h->st_levptr;
        const int64_t* levidcs = fwd ? h->tt_levidcs : h->st_levidcs;
        const int64_t* chdptr = fwd ? h->tt_chdptr : h->st_chdptr;
        const int64_t* chds = fwd ? h->tt_chds : h->st_chds;
        const int64_t* off = fwd ? h->tt_off : h->st_off;
        const double* blocks = fwd ? h->tt_blocks : h->st_blocks;
        for (int64_t lev = 0; lev < nlev; ++lev) {
#pragma omp parallel for schedule(dynamic, 16)
            for (int64_t idx = levptr[lev] - 1; idx < levptr[lev + 1] - 1; ++idx) {
                int64_t p = levidcs[idx] - 1;
                int64_t rp = dst_cptr[p + 1] - dst_cptr[p];
                for (int64_t e = chdptr[idx] - 1; e < chdptr[idx + 1] - 1; ++e) {
                    int64_t c = chds[e] - 1;
                    int64_t rc = dst_cptr[c + 1] - dst_cptr[c];
                    for (int64_t a = 0; a < rc; ++a) {
                        cplx* out = yh + (dst_cptr[c] - 1 + a);
                        for (int64_t j = 0; j < rp; ++j) {
                            /* forward: test E (rc x rp); transposed: trial E (rp x rc) transposed */
                            cplx b = ld_c(blocks, off[e] + (fwd ? a + j * rc : j + a * rp), cj);
                            const cplx* in = yh + (dst_cptr[p] - 1 + j);
                            for (int64_t q = 0; q < nrhs; ++q) out[q * Kd] = cadd(out[q * Kd], cmul(b, in[q * Kd]));
                        }
                    }
                }
            }
        }
    }
    /* Stage 5: leaf expansion (disaggregate.jl:62-80). */
    {
        int64_t nb = fwd ? h->nbasis_test : h->nbasis_trial;
        const int64_t* nodes = fwd ? h->basis_test_nodes : h->basis_trial_nodes;
        const int64_t* off = fwd ? h->basis_test_off : h->basis_trial_off;
        const double* blocks = fwd ? h->basis_test : h->basis_trial;
#pragma omp parallel for schedule(dynamic, 16)
        for (int64_t i = 0; i < nb; ++i) {
            int64_t t = nodes[i] - 1;
            int64_t r = dst_cptr[t + 1] - dst_cptr[t];
            int64_t nt = dst_hi[t] - dst_lo[t] + 1;
            for (int64_t a = 0; a < nt; ++a) {
                int64_t row = dst_perm[dst_lo[t] - 1 + a] - 1;
                if (want && !want[row]) continue;
                cplx* out = acc + row;
                for (int64_t j = 0; j < r; ++j) {
                    /* forward: U (nt x r) as is; transposed: V (r x nt) transposed */
                    cplx b = ld_c(blocks, off[i] + (fwd ? a + j * nt : j + a * r), cj);
                    const cplx* in = yh + (dst_cptr[t] - 1 + j);
                    for (int64_t q = 0; q < nrhs; ++q)
                        out[q * ndst_pts] = cadd(out[q * ndst_pts], cmul(b, in[q * Kd]));
                }
            }
        }
    }
    /* Near field (petrovgalerkinnca.jl:474,531,587; blocks of abstractnca.jl:134-178). */
    if (nearon) {
        if (fwd && h->near_store && !want) {
            near_apply(h, h->near_store, nrhs, X, ldx, acc, ndst_pts);
        } else if (fwd) {
#pragma omp parallel for schedule(dynamic, 4)
            for (int64_t t = 0; t < h->nn_test; ++t) {
                for (int64_t a = h->test_lo[t] - 1; a < h->test_hi[t] && h->nearptr[t + 1] > h->nearptr[t]; ++a) {
                    int64_t row = h->test_perm[a] - 1;
                    if (want && !want[row]) continue;
                    cplx* out = acc + row;
                    for (int64_t i = h->nearptr[t] - 1; i < h->nearptr[t + 1] - 1; ++i) {
                        int64_t s = h->nearnodes[i] - 1;
                        for (int64_t b = h->trial_lo[s] - 1; b < h->trial_hi[s]; ++b) {
                            int64_t col = h->trial_perm[b] - 1;
                            cplx g = entry(h->k, h->sign, h->testpts + 3 * row, h->trialpts + 3 * col);
                            for (int64_t q = 0; q < nrhs; ++q)
                                out[q * ndst_pts] = cadd(out[q * ndst_pts], cmul(g, X[col + q * ldx]));
                        }
                    }
                }
            }
        } else {
            /* transpose the near CSR: trial node -> test nodes */
            int64_t nnz = h->nearptr[h->nn_test] - 1;
            int64_t* tptr = (int64_t*)calloc((size_t)(h->nn_trial + 1), sizeof(int64_t));
            int64_t* tidx = (int64_t*)malloc(sizeof(int64_t) * (size_t)(nnz > 0 ? nnz : 1));
            for (int64_t i = 0; i < nnz; ++i) tptr[h->nearnodes[i]]++;
            for (int64_t s = 0; s < h->nn_trial; ++s) tptr[s + 1] += tptr[s];
            int64_t* cur = (int64_t*)malloc(sizeof(int64_t) * (size_t)(h->nn_trial + 1));
            memcpy(cur, tptr, sizeof(int64_t) * (size_t)(h->nn_trial + 1));
            for (int64_t t = 0; t < h->nn_test; ++t)
                for (int64_t i = h->nearptr[t] - 1; i < h->nearptr[t + 1] - 1; ++i)
                    tidx[cur[h->nearnodes[i] - 1]++] = t;
#pragma omp parallel for schedule(dynamic, 4)
            for (int64_t s = 0; s < h->nn_trial; ++s) {
                for (int64_t b = h->trial_lo[s] - 1; b < h->trial_hi[s] && tptr[s + 1] > tptr[s]; ++b) {
                    int64_t row = h->trial_perm[b] - 1;
                    if (want && !want[row]) continue;
                    cplx* out = acc + row;
                    for (int64_t i = tptr[s]; i < tptr[s + 1]; ++i) {
                        int64_t t = tidx[i];
                        for (int64_t a = h->test_lo[t] - 1; a < h->test_hi[t]; ++a) {
                            int64_t col = h->test_perm[a] - 1;
                            cplx g = entry(h->k, h->sign, h->testpts + 3 * col, h->trialpts + 3 * row);
                            if (cj) g = cconj(g);
                            for (int64_t q = 0; q < nrhs; ++q)
                                out[q * ndst_pts] = cadd(out[q * ndst_pts], cmul(g, X[col + q * ldx]));
                        }
                    }
                }
            }
            free(tptr);
            free(tidx);
            free(cur);
        }
    }
    cplx al = {alpha[0], alpha[1]}, be = {beta[0], beta[1]};
    for (int64_t q = 0; q < nrhs; ++q)
        for (int64_t i = 0; i < ndst_pts; ++i) {
            if (want && !want[i]) continue;
            cplx yo = Y[i + q * ldy];
            cplx v = cmul(al, acc[i + q * ndst_pts]);
            if (be.re != 0.0 || be.im != 0.0) v = cadd(v, cmul(be, yo));
            Y[i + q * ldy] = v;
        }
    free(xh);
    free(yh);
    free(acc);
    free(want);
}

void orc_dense_rows(double k, int32_t sign, const double* testpts, const double* trialpts, int64_t N,
                    int64_t nrows, const int64_t* rows, int64_t nrhs, const double* xin, int64_t ldx,
                    double* yout) {
    const cplx* X = (const cplx*)xin;
    cplx* Y = (cplx*)yout;
#pragma omp parallel
    {
        /* each entry is evaluated once for all right-hand sides; every sum
           still runs over j in increasing order */
        cplx* sum = (cplx*)malloc(sizeof(cplx) * (size_t)(nrhs > 0 ? nrhs : 1));
#pragma omp for schedule(dynamic, 1)
        for (int64_t i = 0; i < nrows; ++i) {
            const double* x = testpts + 3 * rows[i];
            for (int64_t q = 0; q < nrhs; ++q) sum[q].re = sum[q].im = 0.0;
            for (int64_t j = 0; j < N; ++j) {
                const cplx g = entry(k, sign, x, trialpts + 3 * j);
                for (int64_t q = 0; q < nrhs; ++q) sum[q] = cadd(sum[q], cmul(g, X[j + q * ldx]));
            }
            for (int64_t q = 0; q < nrhs; ++q) Y[i + q * nrows] = sum[q];
        }
        free(sum);
    }
}

int32_t orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int32_t n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}
