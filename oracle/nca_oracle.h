/*
 * ORACLE — test infrastructure only.
 *
 * FP64 CPU restatement of the batched bottom-up NCA compressor and the H^2
 * matrix-vector product of NestedCrossApproximation.jl. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker or as the timed CPU baseline. The product never calls it.
 *
 * Conventions: complex numbers are interleaved (re, im) doubles; matrices are
 * column-major; every array that mirrors a Julia field (CSR pointers, node ids,
 * point ids, pivots) is 1-based exactly as the Julia structs hold it; packed
 * block offsets are 0-based element offsets.
 */
#ifndef NCA_ORACLE_H
#define NCA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Scalar pieces. */
void orc_sincos(double x, double* s, double* c);
void orc_entry(double k, int32_t sign, const double* x, const double* y, double* g);

/* Phase 2 (batchedbottomup.jl:101-116): padded level buffer (sum m) x maxrank. */
int32_t orc_assemble_level(double k, int32_t sign, const double* testpts, const double* trialpts,
                           int64_t nb, const int64_t* rowptr, const int64_t* rowidx,
                           const int64_t* colptr, const int64_t* colidx, int64_t maxrank,
                           int32_t transposed, double* buffer);

/* Fully pivoted ACA with the sampled stop on one m x n block (ld >= m);
 * the block is overwritten with the residual. Returns the rank. */
int64_t orc_aca_fullpiv(double* block, int64_t m, int64_t n, int64_t ld, int64_t maxrank,
                        double tol, int32_t sampled, int64_t* rowpiv, int64_t* colpiv);

/* Incomplete ACA with geometric column pivoting on one m x n block whose
 * columns are the candidates in geometric order (see nca_oracle.c); the
 * block is overwritten. Returns the rank. */
int64_t orc_iaca(double* block, int64_t m, int64_t n, int64_t ld, int64_t maxrank, double tol, int64_t* rowpiv,
                 int64_t* colpiv);

/* Skeleton solve X = L / L[rowpiv, :] with LU and the singularity rule of
 * utils.jl:57-80. Returns -1 when regular, else the number of independent
 * pivots. */
int64_t orc_skeleton_solve(const double* L, int64_t m, int64_t r, const int64_t* rowpiv, double* X);

/* testbases/trialbases(::BatchedBottomUp) for one side (batchedbottomup.jl:246-447).
 * Tree layout: perm (npts, 1-based point ids in tree order), nodelo/nodehi
 * (1-based inclusive positions), level (1-based), childptr (nn+1, 1-based),
 * children (1-based). farpivptr/farpiv: Phase-1 pivots per node (1-based).
 * Outputs: pivptr (nn+1) / pivots (capacity nn*maxrank), xoffsets (nn+1, 0-based
 * complex offsets) / X (capacity xcap complex). sampled: 0 or 1, the fully
 * pivoted ACA without or with the sampled stop; 2, the IACA (orc_iaca).
 * Returns 0, or -(node) of the first singular cross matrix, or 1 when X is
 * too small. */
int64_t orc_bases(int32_t side, double k, int32_t sign, const double* testpts, const double* trialpts,
                  int64_t nn, const int64_t* perm, const int64_t* nodelo, const int64_t* nodehi,
                  const int64_t* level, const int64_t* childptr, const int64_t* children,
                  const int64_t* farpivptr, const int64_t* farpiv, int64_t maxrank, double tol,
                  int32_t sampled, int64_t* pivptr, int64_t* pivots, int64_t* xoffsets, double* X,
                  int64_t xcap);

/* Partially pivoted ACA on an m x n block evaluated entrywise: rows / cols are
 * 1-based point ids into rowpts / colpts; entry G(row, col), or G(col, row)
 * with transposed != 0 (pivots.jl:215-300 with ACA(); the published method
 * restated). Returns the rank; rowpiv/colpiv 1-based positions (capacity
 * min(maxrank, m, n)); U (m x rank) and V (n x rank) when not NULL. */
int64_t orc_paca(double k, int32_t sign, int32_t transposed, const double* rowpts, const double* colpts,
                 int64_t m, const int64_t* rows, int64_t n, const int64_t* cols, int64_t maxrank, double tol,
                 int64_t* rowpiv, int64_t* colpiv, double* U, double* V);

/* testbases/trialbases(::TopDown) for one side (topdown.jl:51-206): row and
 * column pivots per node and the factors X in the layout of orc_bases. */
int64_t orc_topdown(int32_t side, double k, int32_t sign, const double* testpts, const double* trialpts,
                    int64_t nn, const int64_t* perm, const int64_t* nodelo, const int64_t* nodehi,
                    const int64_t* level, const int64_t* parent, const int64_t* childptr, const int64_t* children,
                    const int64_t* operm, const int64_t* olo, const int64_t* ohi, const int64_t* farptr,
                    const int64_t* farnodes, int64_t maxrank, double tol, int64_t* pivptr, int64_t* pivots,
                    int64_t* cpivptr, int64_t* cpivots, int64_t* xoffsets, double* X, int64_t xcap);

/* assemble_couplingstore (couplingmatrices.jl:11-53). */
void orc_coupling(double k, int32_t sign, const double* testpts, const double* trialpts, int64_t ntest,
                  const int64_t* farptr, const int64_t* farnodes, const int64_t* testpivptr,
                  const int64_t* testpiv, const int64_t* trialpivptr, const int64_t* trialpiv,
                  const int64_t* offsets, double* blocks);

/* Everything the H^2 matvec reads. */
typedef struct {
    int64_t M, N, nn_test, nn_trial;
    double k;
    int32_t sign;
    const double* testpts;
    const double* trialpts;
    const int64_t *test_perm, *test_lo, *test_hi, *trial_perm, *trial_lo, *trial_hi;
    const int64_t *test_cptr, *trial_cptr; /* CoefficientPlan (1-based) */
    /* BasisStore: nodes + packed blocks (test |t| x r, trial r x |s|) */
    int64_t nbasis_test, nbasis_trial;
    const int64_t *basis_test_nodes, *basis_test_off, *basis_trial_nodes, *basis_trial_off;
    const double *basis_test, *basis_trial;
    /* TransferStore (TransferTraversalPlan, 1-based) + packed blocks */
    int64_t nlev_test, nlev_trial;
    const int64_t *tt_levptr, *tt_levidcs, *tt_chdptr, *tt_chds, *tt_off;
    const int64_t *st_levptr, *st_levidcs, *st_chdptr, *st_chds, *st_off;
    const double *tt_blocks, *st_blocks;
    /* CouplingStore (CouplingTraversalPlan, 1-based) + packed blocks */
    const int64_t *cp_ptr, *cp_nodes, *cp_off;
    const double* cp_blocks;
    /* near interactions (1-based CSR of trial nodes per test node) */
    const int64_t *nearptr, *nearnodes;
    /* optional stored near field (orc_near_assemble); NULL: evaluated on the fly */
    const struct orc_near* near_store;
} orc_h2;

/* Stored near field as the reference's CPU path keeps it
 * (src/nca/abstractnca.jl:134-178, a BlockSparseMatrix): one dense block row per
 * test cluster with near interactions, rows = values(t), columns = the
 * concatenated values of its near trial clusters (real and imaginary parts
 * split, column-major). With h->near_store set, the forward product over all
 * rows multiplies these blocks instead of evaluating the kernel. */
typedef struct orc_near orc_near;
orc_near* orc_near_assemble(const orc_h2* h);
int64_t orc_near_entries(const orc_near* s);
void orc_near_free(orc_near* s);

/* y = alpha * op(A) * x + beta * y for op 0 (N), 1 (T), 2 (C); x, y hold nrhs
 * columns with leading dimensions ldx, ldy. near: 1 evaluates the near field,
 * 0 leaves it out. nrows >= 0 restricts the output to those 0-based rows of
 * the destination space (others untouched). */
void orc_h2_matvec(const orc_h2* h, int32_t op, int64_t nrhs, const double* x, int64_t ldx, double* y,
                   int64_t ldy, const double* alpha, const double* beta, int32_t near, int64_t nrows,
                   const int64_t* rows);

/* Dense direct summation y[i] = sum_j G(x_i, y_j) x_j at 0-based test rows. */
void orc_dense_rows(double k, int32_t sign, const double* testpts, const double* trialpts, int64_t N,
                    int64_t nrows, const int64_t* rows, int64_t nrhs, const double* x, int64_t ldx,
                    double* y);

int32_t orc_num_threads(void);
void orc_set_num_threads(int32_t n);

#ifdef __cplusplus
}
#endif
#endif
