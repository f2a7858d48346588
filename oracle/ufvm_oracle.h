/*
 * ufvm_oracle.h -- CPU restatement of the uFVM pressure-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the tests compare
 * the HIP implementation against, and the CPU baseline bench.py times.  It is
 * never linked into, loaded by, or called from the product library.
 *
 * Matrices use the reference's LDU-in-cell-lists layout
 * (src/fvm/setup/cfdSetupCoefficients.m:7-39) flattened to CSR: row i holds
 * cconn{i} in reference order at col[ptr[i]..ptr[i+1]), with anb{i} at the same
 * positions; ac is the diagonal and bc the right-hand side.  All indices are
 * 0-based.  The matrix is A(i,i)=ac(i), A(i,cconn{i})=anb{i}
 * (src/utilities/extras/assembleMatrixOfCoefficients.m:11-15).
 */
#ifndef UFVM_ORACLE_H
#define UFVM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int32_t n, nnz;
  int32_t *ptr, *col;
  double *ac, *anb, *bc, *dphi;
  int32_t *parents;           /* fine -> coarse map, set when a coarser level exists */
} orc_level;

typedef struct {
  int32_t nlevels;
  orc_level lev[16];
} orc_hier;

/* cfdComputeResidualsArray.m:11-17 */
void orc_residual(const orc_level *L, double *r);

/* cfdAgglomerateLevel.m:3-104 (parents + coarse cconn; coefficients zeroed) */
int orc_agglomerate_level(orc_level *fine, orc_level *coarse);
/* assembleAgglomeratedLHS.m:16-30 */
void orc_agglomerated_lhs(const orc_level *fine, orc_level *coarse);
/* cfdAgglomerate.m:7-17 -> number of levels (the reference's maxLevels) */
int orc_agglomerate(orc_hier *H, int maxCoarseLevels);

/* cfdUpdateRHS.m:10-17 */
void orc_update_rhs(const orc_level *fine, const double *r, orc_level *coarse);
/* cfdCorrectFinerLevelSolution.m:11-14 */
void orc_correct_finer(orc_level *fine, const orc_level *coarse);

/* cfdSolveSOR.m:13-29 (one forward + one backward lexicographic pass) */
void orc_sor(const orc_level *L, double *dphi);
/* cfdFactorizeILU.m:10-47 */
void orc_factorize_ilu(const orc_level *L, double *dc, double *rc);
/* cfdSolveILU.m:14-93 */
void orc_solve_ilu(const orc_level *L, const double *dc, double *rc, double *dphi);

/* cfdSolveAlgebraicSystem.m:11-66; smoother 0 = 'ILU', 1 = 'SOR', 2 = symmetric
 * multicolour GS (orc_mcgs over orc_colour; no reference counterpart) */
void orc_solve_algebraic_system(orc_level *L, int smoother, int nIter, double rrf,
                                double *r0, double *r1);

/* cfdApplyAMG.m:9-60 with cfdApplyVCycle/WCycle/FCycle; cycle 0=V 1=W 2=F */
void orc_apply_amg(orc_hier *H, int cycle, int maxCycles, double rrf, int pre, int post,
                   int smoother, double *r0, double *r1, int *ncycles);

/* ---- build-only path (no reference counterpart) ---- */

/* greedy first-fit colouring in index order; returns the number of colours */
int orc_colour(int32_t n, const int32_t *ptr, const int32_t *col, int32_t *colour);
/* the same within blocks of R consecutive rows (only in-block neighbours count) */
int orc_colour_blocks(int32_t n, const int32_t *ptr, const int32_t *col, int32_t R,
                      int32_t *colour);

/* one symmetric multicolour Gauss-Seidel sweep (colours ascending, then
 * descending), each row summed in cconn order */
void orc_mcgs(const orc_level *L, const int32_t *colour, int ncolours, const double *b,
              double *x);

/* preconditioner V-cycle: zero initial guess on every level, nu_pre/nu_post
 * symmetric sweeps, restriction by parents scaled by coarse_scale (on level 0 the residual of the
 * rows of colour 0, solved last by the pre-sweep, is taken as 0), level
 * nlev-1 solved exactly
 * (dense LU, + mean(diag)/n 11^T when `singular`).  Level 0 uses global
 * multicolour GS; levels 1..nlev-2 use block-local multicolour GS over blocks
 * of block_rows consecutive rows (Jacobi coupling between blocks), with the
 * prolongation applied to the post-smoother's starting values. */
typedef struct {
  int32_t nlev;          /* levels used: 0..nlev-1, nlev-1 solved directly */
  int32_t pre, post;
  int32_t singular;
  int32_t block_rows;
  int32_t coarse_smoother;  /* levels 1..nlev-2: 0 block-local MCGS, 1 damped Jacobi (2/3) */
  int32_t smoother;         /* level 0: 0 symmetric multicolour GS, 1 the reference's DILU
                               (cfdFactorizeILU / cfdSolveILU, lexicographic order),
                               2 DILU in colour order (rows by colour, then index),
                               3 ILU(0) in colour order (orc_mcilu0_factor) */
  double coarse_scale;      /* restricted right-hand sides (hence every coarse correction)
                               scaled by this factor */
  int32_t cpre, cpost;      /* sweeps on levels 1..nlev-2 */
} orc_vcfg;

/* DILU in colour order on one level (no reference counterpart): rows in the
 * order (colour, index); the lower neighbours of row j are its neighbours of
 * an earlier colour.  orc_mcilu_factor fills dinv[n]: dinv_j = 1 / (a_jj -
 * sum of (a_ji dinv_i) a_ij), lower neighbours in increasing order.
 * orc_mcilu_sweep runs one sweep on x: forward rc_j = (b - A x)_j - sum of
 * a_ji (dinv_i rc_i) over lower neighbours in entry order, backward
 * z_i = dinv_i (rc_i - sum of a_ij z_j over upper neighbours), x += z
 * (from_zero: x = 0 before, x = z after); rc[n] is work space. */
void orc_mcilu_factor(const orc_level *L, const int32_t *colour, int ncol, double *dinv);
void orc_mcilu_sweep(const orc_level *L, const int32_t *colour, int ncol, const double *dinv,
                     const double *b, int from_zero, double *rc, double *x);
/* ILU(0) in colour order: dinv and the factor's off-diagonal values F (L's
 * CSR order: the undivided L entries below, U entries above); applying it is
 * orc_mcilu_sweep_f with F in place of anb */
void orc_mcilu0_factor(const orc_level *L, const int32_t *colour, int ncol, double *dinv,
                       double *F);
void orc_mcilu_sweep_f(const orc_level *L, const int32_t *colour, int ncol, const double *dinv,
                       const double *F, const double *b, int from_zero, double *rc, double *x);
void orc_vcycle(const orc_hier *H, int32_t *const *colours, const int *ncolours,
                const orc_vcfg *cfg, const double *b, double *x);

/* flexible CG (Notay FCG(1)) preconditioned by orc_vcycle; x starts at 0.
 * Returns iterations; stops at ||r||_2 <= tol * ||b||_2. */
int orc_fpcg(const orc_hier *H, int32_t *const *colours, const int *ncolours,
             const orc_vcfg *cfg, const double *b, double tol, int maxit, double *x,
             double *relres);

/* level allocation helpers */
void orc_level_alloc(orc_level *L, int32_t n, int32_t nnz);
void orc_level_free(orc_level *L);
void orc_hier_free(orc_hier *H);

#ifdef __cplusplus
}
#endif
#endif
