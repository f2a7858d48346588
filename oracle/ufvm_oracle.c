/*
 * ufvm_oracle.c -- CPU restatement of the uFVM AMG pressure-solve path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation and the
 * CPU baseline of bench.py.  Nothing in the product library links or calls it.
 *
 * Sequential FP64, reference orderings.  Build with -ffp-contract=off so that
 * every a*b+c rounds twice, as the interpreted reference does.  Each function
 * cites the reference file:line it restates (paths relative to uFVM/).
 */
#include "ufvm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

void orc_level_alloc(orc_level *L, int32_t n, int32_t nnz) {
  L->n = n;
  L->nnz = nnz;
  L->ptr = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  L->col = (int32_t *)calloc(nnz > 0 ? (size_t)nnz : 1, sizeof(int32_t));
  L->ac = (double *)calloc((size_t)n, sizeof(double));
  L->anb = (double *)calloc(nnz > 0 ? (size_t)nnz : 1, sizeof(double));
  L->bc = (double *)calloc((size_t)n, sizeof(double));
  L->dphi = (double *)calloc((size_t)n, sizeof(double));
  L->parents = NULL;
}

void orc_level_free(orc_level *L) {
  free(L->ptr); free(L->col); free(L->ac); free(L->anb); free(L->bc); free(L->dphi);
  free(L->parents);
  memset(L, 0, sizeof(*L));
}

void orc_hier_free(orc_hier *H) {
  /* level 0 is owned by the caller; coarse levels by the hierarchy */
  for (int l = 1; l < H->nlevels; ++l) orc_level_free(&H->lev[l]);
  free(H->lev[0].parents);
  H->lev[0].parents = NULL;
  H->nlevels = 1;
}

/* cfdComputeResidualsArray.m:11-17: r_i = bc_i - ac_i*phi_i, then minus
 * anb_i[k]*phi[cconn_i[k]] in cconn order. */
void orc_residual(const orc_level *L, double *r) {
  for (int32_t i = 0; i < L->n; ++i) {
    double s = L->bc[i] - L->ac[i] * L->dphi[i];
    for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) s = s - L->anb[k] * L->dphi[L->col[k]];
    r[i] = s;
  }
}

/* --------------------------- agglomeration ------------------------------ */

/* cfdAgglomerateLevel.m:3-104.  Greedy seed agglomeration:
 *  maxAnb_i = max(-anb_i)                                            (:13-15)
 *  seeds in index order take unassigned cconn neighbours with
 *  -anb/maxAnb_seed > 0.5; first-generation children then take their own
 *  unassigned neighbours with -anb/maxAnb_child > 0.5               (:20-46)
 *  a seed that took nobody is released                               (:48-52)
 *  orphans join the assigned neighbour maximising -anb_io/maxAnb_nb (strict >,
 *  first wins), else become a new parent                             (:57-75)
 *  coarse cconn in first-encounter order over fine rows/cconn        (:83-98) */
int orc_agglomerate_level(orc_level *F, orc_level *C) {
  const int32_t n = F->n;
  int32_t *parents = (int32_t *)calloc((size_t)n, sizeof(int32_t)); /* 1-based, 0 = none */
  double *maxAnb = (double *)malloc((size_t)n * sizeof(double));
  int32_t *children = (int32_t *)malloc((size_t)(n + 1) * sizeof(int32_t));
  for (int32_t i = 0; i < n; ++i) {
    double m = -INFINITY;
    for (int32_t k = F->ptr[i]; k < F->ptr[i + 1]; ++k) {
      double v = -F->anb[k];
      if (v > m) m = v;
    }
    maxAnb[i] = m;
  }
  int32_t iParent = 1;
  for (int32_t s = 0; s < n; ++s) {
    if (parents[s] != 0) continue;
    parents[s] = iParent;
    int32_t nch = 0;
    children[nch++] = s;
    for (int32_t k = F->ptr[s]; k < F->ptr[s + 1]; ++k) {
      int32_t j = F->col[k];
      if (parents[j] == 0 && (-F->anb[k] / maxAnb[s]) > 0.5) {
        parents[j] = iParent;
        children[nch++] = j;
      }
    }
    int32_t n2 = 0;
    for (int32_t c = 1; c < nch; ++c) {
      int32_t ch = children[c];
      for (int32_t k = F->ptr[ch]; k < F->ptr[ch + 1]; ++k) {
        int32_t j = F->col[k];
        if (parents[j] == 0 && (-F->anb[k] / maxAnb[ch]) > 0.5) {
          parents[j] = iParent;
          ++n2;
        }
      }
    }
    if (nch + n2 == 1) parents[s] = 0;
    else ++iParent;
  }
  for (int32_t o = 0; o < n; ++o) {
    if (parents[o] == 0) {
      double strength = 0.0;
      for (int32_t k = F->ptr[o]; k < F->ptr[o + 1]; ++k) {
        int32_t j = F->col[k];
        if (parents[j] != 0) {
          double v = -F->anb[k] / maxAnb[j];
          if (strength < v) {
            strength = v;
            parents[o] = parents[j];
          }
        }
      }
    }
    if (parents[o] == 0) parents[o] = iParent++;
  }
  const int32_t nc = iParent - 1;
  /* coarse cconn, first-encounter order (:85-94) */
  int32_t *cnt = (int32_t *)calloc((size_t)nc, sizeof(int32_t));
  int32_t *cap = (int32_t *)calloc((size_t)nc, sizeof(int32_t));
  int32_t **lst = (int32_t **)calloc((size_t)nc, sizeof(int32_t *));
  for (int32_t i = 0; i < n; ++i) {
    int32_t p = parents[i] - 1;
    for (int32_t k = F->ptr[i]; k < F->ptr[i + 1]; ++k) {
      int32_t q = parents[F->col[k]] - 1;
      if (p == q) continue;
      int found = 0;
      for (int32_t t = 0; t < cnt[p]; ++t)
        if (lst[p][t] == q) { found = 1; break; }
      if (!found) {
        if (cnt[p] == cap[p]) {
          cap[p] = cap[p] ? 2 * cap[p] : 8;
          lst[p] = (int32_t *)realloc(lst[p], (size_t)cap[p] * sizeof(int32_t));
        }
        lst[p][cnt[p]++] = q;
      }
    }
  }
  int64_t nnz = 0;
  for (int32_t p = 0; p < nc; ++p) nnz += cnt[p];
  orc_level_alloc(C, nc, (int32_t)nnz);
  C->ptr[0] = 0;
  for (int32_t p = 0; p < nc; ++p) {
    C->ptr[p + 1] = C->ptr[p] + cnt[p];
    if (cnt[p]) memcpy(C->col + C->ptr[p], lst[p], (size_t)cnt[p] * sizeof(int32_t));
    free(lst[p]);
  }
  free(lst); free(cnt); free(cap);
  free(F->parents);
  F->parents = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  for (int32_t i = 0; i < n; ++i) F->parents[i] = parents[i] - 1;
  free(parents); free(maxAnb); free(children);
  return nc;
}

/* assembleAgglomeratedLHS.m:16-30: for fine i in order, AC[p] += ac_i; for
 * each neighbour in cconn order, same parent -> AC[p] += anb, else the coarse
 * entry of parent(nb) in CCONN{p} (first match) += anb. */
void orc_agglomerated_lhs(const orc_level *F, orc_level *C) {
  memset(C->ac, 0, (size_t)C->n * sizeof(double));
  memset(C->anb, 0, (size_t)(C->nnz > 0 ? C->nnz : 1) * sizeof(double));
  for (int32_t i = 0; i < F->n; ++i) {
    int32_t p = F->parents[i];
    C->ac[p] = C->ac[p] + F->ac[i];
    for (int32_t k = F->ptr[i]; k < F->ptr[i + 1]; ++k) {
      int32_t q = F->parents[F->col[k]];
      if (q == p) {
        C->ac[p] = C->ac[p] + F->anb[k];
      } else {
        for (int32_t t = C->ptr[p]; t < C->ptr[p + 1]; ++t)
          if (C->col[t] == q) { C->anb[t] = C->anb[t] + F->anb[k]; break; }
      }
    }
  }
}

/* cfdAgglomerate.m:7-17: agglomerate level after level until a level has at
 * most 5 parents or maxCoarseLevels coarse levels exist.  Returns the level
 * count (the reference's maxLevels). */
int orc_agglomerate(orc_hier *H, int maxCoarseLevels) {
  int iLevel = 1;  /* 1-based as in the reference */
  while (iLevel <= maxCoarseLevels && iLevel < 16) {
    iLevel = iLevel + 1;
    int np = orc_agglomerate_level(&H->lev[iLevel - 2], &H->lev[iLevel - 1]);
    orc_agglomerated_lhs(&H->lev[iLevel - 2], &H->lev[iLevel - 1]);
    H->nlevels = iLevel;
    if (np <= 5) break;
  }
  H->nlevels = iLevel;
  return iLevel;
}

/* cfdUpdateRHS.m:10-17: BC = zeros; BC(parent(i)) += r(i) for fine i in order */
void orc_update_rhs(const orc_level *F, const double *r, orc_level *C) {
  memset(C->bc, 0, (size_t)C->n * sizeof(double));
  for (int32_t i = 0; i < F->n; ++i) C->bc[F->parents[i]] = C->bc[F->parents[i]] + r[i];
}

/* cfdCorrectFinerLevelSolution.m:11-14 */
void orc_correct_finer(orc_level *F, const orc_level *C) {
  for (int32_t i = 0; i < F->n; ++i) F->dphi[i] = F->dphi[i] + C->dphi[F->parents[i]];
}

/* ------------------------------ smoothers ------------------------------- */

static void gs_row(const orc_level *L, const double *b, double *x, int32_t i) {
  double s = b[i];
  for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) s = s - L->anb[k] * x[L->col[k]];
  x[i] = s / L->ac[i];
}

/* cfdSolveSOR.m:13-29: forward then backward lexicographic Gauss-Seidel */
void orc_sor(const orc_level *L, double *dphi) {
  for (int32_t i = 0; i < L->n; ++i) gs_row(L, L->bc, dphi, i);
  for (int32_t i = L->n - 1; i >= 0; --i) gs_row(L, L->bc, dphi, i);
}

static int32_t find_in_row(const orc_level *L, int32_t j, int32_t i) {
  for (int32_t t = L->ptr[j]; t < L->ptr[j + 1]; ++t)
    if (L->col[t] == i) return t;
  return -1;
}

/* cfdFactorizeILU.m:10-47 (DILU).  The update loop is skipped for the row
 * whose 1-based index is numberOfElements-1 (:21), kept as is. */
void orc_factorize_ilu(const orc_level *L, double *dc, double *rc) {
  const int32_t n = L->n;
  for (int32_t i = 0; i < n; ++i) dc[i] = L->ac[i];
  for (int32_t i = 0; i < n; ++i) {
    dc[i] = 1.0 / dc[i];
    rc[i] = L->bc[i];
    if (i + 1 != n - 1) {
      for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) {
        int32_t j = L->col[k];
        if (j > i) {
          int32_t t = find_in_row(L, j, i);
          if (t >= 0) dc[j] = dc[j] - (L->anb[t] * dc[i]) * L->anb[k];
          else fprintf(stderr, "the index for i in j is not found\n");
        }
      }
    }
  }
}

/* cfdSolveILU.m:14-93 with right-hand side b: residual, forward DILU
 * substitution, backward substitution with dphi += D*rc. */
static void ilu_apply(const orc_level *L, const double *dc, const double *b, double *rc,
                      double *dphi) {
  const int32_t n = L->n;
  for (int32_t i = 0; i < n; ++i) {
    double s = -L->ac[i] * dphi[i] + b[i];
    for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) s = s - L->anb[k] * dphi[L->col[k]];
    rc[i] = s;
  }
  for (int32_t i = 0; i < n; ++i) {
    double mat1 = dc[i] * rc[i];
    for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) {
      int32_t j = L->col[k];
      if (j > i) {
        int32_t t = find_in_row(L, j, i);
        if (t >= 0) rc[j] = rc[j] - L->anb[t] * mat1;
        else fprintf(stderr, "ILU Solver Error The index for i in element j is not found\n");
      }
    }
  }
  for (int32_t i = n - 1; i >= 0; --i) {
    if (i < n - 1) {
      for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) {
        int32_t j = L->col[k];
        if (j > i) rc[i] = rc[i] - L->anb[k] * rc[j];
      }
    }
    double mat1 = dc[i] * rc[i];
    rc[i] = mat1;
    dphi[i] = dphi[i] + mat1;
  }
}

/* cfdSolveILU.m:14-93 on the level's own right-hand side */
void orc_solve_ilu(const orc_level *L, const double *dc, double *rc, double *dphi) {
  ilu_apply(L, dc, L->bc, rc, dphi);
}

static double l1(const double *r, int32_t n) {
  double s = 0.0;
  for (int32_t i = 0; i < n; ++i) s = s + fabs(r[i]);
  return s;
}

/* cfdSolveAlgebraicSystem.m:11-66 */
void orc_solve_algebraic_system(orc_level *L, int smoother, int nIter, double rrf, double *r0,
                                double *r1) {
  const int32_t n = L->n;
  double *r = (double *)malloc((size_t)n * sizeof(double));
  orc_residual(L, r);
  double initial = l1(r, n) / n;
  double final_ = initial;
  if (smoother == 0) {
    double *dc = (double *)malloc((size_t)n * sizeof(double));
    double *rc = (double *)malloc((size_t)n * sizeof(double));
    orc_factorize_ilu(L, dc, rc);
    int iter = 1;
    while (iter <= nIter && final_ > rrf * initial) {
      orc_solve_ilu(L, dc, rc, L->dphi);
      orc_residual(L, r);
      final_ = l1(r, n) / n;
      ++iter;
    }
    free(dc); free(rc);
  } else if (smoother == 1) {
    int iter = 1;
    while (iter <= nIter && final_ > rrf * initial) {
      orc_sor(L, L->dphi);
      orc_residual(L, r);
      final_ = l1(r, n) / n;
      ++iter;
    }
  } else {
    /* symmetric multicolour GS (no reference counterpart): the same
       termination loop with orc_mcgs sweeps over a greedy colouring */
    int32_t *colour = (int32_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    const int nc = orc_colour(n, L->ptr, L->col, colour);
    int iter = 1;
    while (iter <= nIter && final_ > rrf * initial) {
      orc_mcgs(L, colour, nc, L->bc, L->dphi);
      orc_residual(L, r);
      final_ = l1(r, n) / n;
      ++iter;
    }
    free(colour);
  }
  free(r);
  if (r0) *r0 = initial;
  if (r1) *r1 = final_;
}

/* -------------------------------- cycles -------------------------------- */

typedef struct {
  orc_hier *H;
  int smoother, pre, post;
  double rrf;
} cyc_t;

/* cfdRestrict.m:5-12 (level is 1-based as in the reference) */
static void restrict_(cyc_t *c, int level, int presmooth) {
  orc_level *L = &c->H->lev[level - 1];
  if (presmooth) orc_solve_algebraic_system(L, c->smoother, c->pre, 0.1, NULL, NULL);
  double *r = (double *)malloc((size_t)L->n * sizeof(double));
  orc_residual(L, r);
  orc_update_rhs(L, r, &c->H->lev[level]);
  free(r);
}

/* cfdProlongate.m:5-6 */
static void prolongate_(cyc_t *c, int level) {
  orc_correct_finer(&c->H->lev[level - 2], &c->H->lev[level - 1]);
  orc_solve_algebraic_system(&c->H->lev[level - 2], c->smoother, c->post, c->rrf, NULL, NULL);
}

static void coarsest_(cyc_t *c, int level) {
  orc_solve_algebraic_system(&c->H->lev[level - 1], c->smoother, c->post, 0.1, NULL, NULL);
}

static double final_res(cyc_t *c) {
  orc_level *L = &c->H->lev[0];
  double *r = (double *)malloc((size_t)L->n * sizeof(double));
  orc_residual(L, r);
  double s = l1(r, L->n);
  free(r);
  return s;
}

/* cfdApplyVCycle.m:5-25 */
static double vcycle_(cyc_t *c, int G) {
  int g = 1;
  while (g < G) { restrict_(c, g, 1); ++g; }
  coarsest_(c, g);
  while (g > 1) { prolongate_(c, g); --g; }
  return final_res(c);
}

/* cfdApplyWCycle.m:3-83 */
static double wcycle_(cyc_t *c, int G) {
  int g = 1;
  while (g < G) { restrict_(c, g, 1); ++g; }
  coarsest_(c, g);
  prolongate_(c, g);
  restrict_(c, g - 1, 0);
  coarsest_(c, g);
  prolongate_(c, g);
  prolongate_(c, g - 1);
  restrict_(c, g - 2, 0);
  restrict_(c, g - 1, 1);
  coarsest_(c, g);
  prolongate_(c, g);
  prolongate_(c, g - 1);
  prolongate_(c, g - 2);
  restrict_(c, g - 3, 0);
  restrict_(c, g - 2, 1);
  restrict_(c, g - 1, 1);
  coarsest_(c, g);
  prolongate_(c, g);
  prolongate_(c, g - 1);
  restrict_(c, g - 2, 0);
  restrict_(c, g - 1, 1);
  coarsest_(c, g);
  prolongate_(c, g);
  restrict_(c, g - 1, 0);
  coarsest_(c, g);
  while (g > 1) { prolongate_(c, g); --g; }
  return final_res(c);
}

/* cfdApplyFCycle.m:5-63 */
static double fcycle_(cyc_t *c, int G) {
  int g = 1;
  while (g < G) { restrict_(c, g, 1); ++g; }
  coarsest_(c, g);
  prolongate_(c, g);
  restrict_(c, g - 1, 0);
  coarsest_(c, g);
  prolongate_(c, g);
  prolongate_(c, g - 1);
  restrict_(c, g - 2, 0);
  restrict_(c, g - 1, 1);
  coarsest_(c, g);
  prolongate_(c, g);
  prolongate_(c, g - 1);
  prolongate_(c, g - 2);
  restrict_(c, g - 3, 0);
  restrict_(c, g - 2, 1);
  restrict_(c, g - 1, 1);
  coarsest_(c, g);
  while (g > 1) { prolongate_(c, g); --g; }
  return final_res(c);
}

/* cfdApplyAMG.m:23-60 (the hierarchy is built beforehand by orc_agglomerate,
 * which is what cfdApplyAMG.m:19 calls).  *ncycles receives the reference's
 * nCycle counter, i.e. cycles performed + 1 (:37,59). */
void orc_apply_amg(orc_hier *H, int cycle, int maxCycles, double rrf, int pre, int post,
                   int smoother, double *r0, double *r1, int *ncycles) {
  cyc_t c = {H, smoother, pre, post, rrf};
  if (cycle != 0 && H->nlevels < 4) {
    /* the hand-unrolled W/F patterns address level maxLevels-3 (1-based) */
    if (ncycles) *ncycles = -1;
    return;
  }
  double initial = final_res(&c);
  double final_ = initial;
  int nCycle = 1;
  const int G = H->nlevels;
  while (nCycle <= maxCycles && final_ > rrf * initial) {
    if (cycle == 0) final_ = vcycle_(&c, G);
    else if (cycle == 1) final_ = wcycle_(&c, G);
    else final_ = fcycle_(&c, G);
    ++nCycle;
  }
  if (r0) *r0 = initial;
  if (r1) *r1 = final_;
  if (ncycles) *ncycles = nCycle;
}

/* ------------------------- build-only: FPCG path ------------------------- */

int orc_colour(int32_t n, const int32_t *ptr, const int32_t *col, int32_t *colour) {
  int32_t ncol = 0;
  int32_t *mark = (int32_t *)malloc(sizeof(int32_t) * 1024);
  for (int t = 0; t < 1024; ++t) mark[t] = -1;
  for (int32_t i = 0; i < n; ++i) colour[i] = -1;
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
      int32_t c = colour[col[k]];
      if (c >= 0 && c < 1024) mark[c] = i;
    }
    int32_t c = 0;
    while (c < 1024 && mark[c] == i) ++c;
    colour[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  free(mark);
  return ncol;
}

int orc_colour_blocks(int32_t n, const int32_t *ptr, const int32_t *col, int32_t R,
                      int32_t *colour) {
  int32_t ncol = 0;
  int32_t mark[1024];
  for (int t = 0; t < 1024; ++t) mark[t] = -1;
  for (int32_t i = 0; i < n; ++i) colour[i] = -1;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t lo = (i / R) * R, hi = lo + R < n ? lo + R : n;
    for (int32_t k = ptr[i]; k < ptr[i + 1]; ++k) {
      const int32_t j = col[k];
      if (j < lo || j >= hi) continue;
      const int32_t c = colour[j];
      if (c >= 0 && c < 1024) mark[c] = i;
    }
    int32_t c = 0;
    while (c < 1024 && mark[c] == i) ++c;
    colour[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  return ncol;
}

/* Block-local multicolour GS.  v0 = starting values of every row (mode 0: 0,
 * mode 1: xin, mode 2: xin + xc[parent]); each block of R rows sweeps its own
 * rows (colours up, then down) reading in-block neighbours' current values and
 * out-of-block neighbours' v0; xout receives the block values. */
static void bmcgs(const orc_level *L, const int32_t *colour, int ncol, int32_t R, const double *b,
                  const double *xin, const int32_t *parent, const double *xc, int mode,
                  int nsweep, double *xout) {
  const int32_t n = L->n;
  double *v0 = (double *)malloc((size_t)n * sizeof(double));
  for (int32_t i = 0; i < n; ++i) {
    double v = 0.0;
    if (mode >= 1) v = xin[i];
    if (mode == 2) v = v + xc[parent[i]];
    v0[i] = v;
    xout[i] = v;
  }
  for (int32_t lo = 0; lo < n; lo += R) {
    const int32_t hi = lo + R < n ? lo + R : n;
    for (int s = 0; s < nsweep; ++s) {
      for (int dir = 0; dir < 2; ++dir) {
        for (int cc = 0; cc < ncol; ++cc) {
          const int c = dir == 0 ? cc : ncol - 1 - cc;
          for (int32_t i = lo; i < hi; ++i) {
            if (colour[i] != c) continue;
            double v = b[i];
            for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) {
              const int32_t j = L->col[k];
              const double xj = (j >= lo && j < hi) ? xout[j] : v0[j];
              v = v - L->anb[k] * xj;
            }
            xout[i] = v / L->ac[i];
          }
        }
      }
    }
  }
  free(v0);
}

void orc_mcgs(const orc_level *L, const int32_t *colour, int ncolours, const double *b,
              double *x) {
  /* the rows of one colour are independent: the OpenMP build splits them */
  for (int c = 0; c < ncolours; ++c) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t i = 0; i < L->n; ++i)
      if (colour[i] == c) gs_row(L, b, x, i);
  }
  for (int c = ncolours - 1; c >= 0; --c) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t i = 0; i < L->n; ++i)
      if (colour[i] == c) gs_row(L, b, x, i);
  }
}

/* ---- DILU in colour order (the kind-5 finest smoother of the GPU path) */

/* rows in the order (colour, index): order[t] = row, pos[row] = t */
static void mc_order(const int32_t *colour, int ncol, int32_t n, int32_t *order, int32_t *pos) {
  int32_t *start = (int32_t *)calloc((size_t)ncol + 1, sizeof(int32_t));
  for (int32_t i = 0; i < n; ++i) ++start[colour[i] + 1];
  for (int c = 0; c < ncol; ++c) start[c + 1] += start[c];
  for (int32_t i = 0; i < n; ++i) order[start[colour[i]]++] = i;
  for (int32_t t = 0; t < n; ++t) pos[order[t]] = t;
  free(start);
}

/* first entry of row r with column c (-1 if none) */
static int32_t entry_of(const orc_level *L, int32_t r, int32_t c) {
  for (int32_t k = L->ptr[r]; k < L->ptr[r + 1]; ++k)
    if (L->col[k] == c) return k;
  return -1;
}

/* row j's lower neighbours (earlier in the order), increasing position:
 * their rows into nb[], returns the count */
static int lower_of(const orc_level *L, const int32_t *pos, int32_t j, int32_t *nb) {
  int m = 0;
  for (int32_t k = L->ptr[j]; k < L->ptr[j + 1]; ++k)
    if (pos[L->col[k]] < pos[j]) nb[m++] = L->col[k];
  for (int a = 1; a < m; ++a) /* insertion sort by position, stable */
    for (int b = a; b > 0 && pos[nb[b - 1]] > pos[nb[b]]; --b) {
      int32_t t = nb[b]; nb[b] = nb[b - 1]; nb[b - 1] = t;
    }
  return m;
}

static int32_t max_row(const orc_level *L) {
  int32_t w = 1;
  for (int32_t i = 0; i < L->n; ++i)
    if (L->ptr[i + 1] - L->ptr[i] > w) w = L->ptr[i + 1] - L->ptr[i];
  return w;
}

void orc_mcilu_factor(const orc_level *L, const int32_t *colour, int ncol, double *dinv) {
  const int32_t n = L->n;
  int32_t *order = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  int32_t *pos = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  int32_t *nb = (int32_t *)malloc((size_t)max_row(L) * sizeof(int32_t));
  mc_order(colour, ncol, n, order, pos);
  for (int32_t t = 0; t < n; ++t) {
    const int32_t j = order[t];
    double d = L->ac[j];
    const int m = lower_of(L, pos, j, nb);
    for (int e = 0; e < m; ++e) {
      const int32_t i = nb[e];
      d = d - (L->anb[entry_of(L, j, i)] * dinv[i]) * L->anb[entry_of(L, i, j)];
    }
    dinv[j] = 1.0 / d;
  }
  free(order); free(pos); free(nb);
}

/* ILU(0) in colour order (smoother 3 of the precision-form V-cycle; no
 * reference counterpart: cfg3 asks for ILU(0) beside multicolour GS).  Row by
 * row in (colour, index) order, IKJ elimination restricted to the pattern:
 * w = row j of A; for each earlier-colour neighbour k of j in increasing
 * position, l = w_jk dinv_k, F_jk = w_jk (the undivided L entry), and for each
 * later entry m of row k (pos m > pos k, entry order): m = j updates the
 * diagonal d = d - l F_kj, any other m in row j updates w_jm = w_jm - l F_km;
 * then F_jm = w_jm for the later entries of row j and dinv_j = 1 / d.  The
 * preconditioner M = (D + L~) D^-1 (D + U~) with L~, U~ the F values is L U of
 * ILU(0), and applying it is orc_mcilu_sweep_f with F.  On a graph without
 * triangles no off-diagonal changes (F = anb) and the factor is
 * orc_mcilu_factor's bit for bit. */
void orc_mcilu0_factor(const orc_level *L, const int32_t *colour, int ncol, double *dinv,
                       double *F) {
  const int32_t n = L->n;
  int32_t *order = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  int32_t *pos = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  int32_t *nb = (int32_t *)malloc((size_t)max_row(L) * sizeof(int32_t));
  mc_order(colour, ncol, n, order, pos);
  for (int32_t t = 0; t < n; ++t) {
    const int32_t j = order[t];
    for (int32_t k = L->ptr[j]; k < L->ptr[j + 1]; ++k) F[k] = L->anb[k];
    double d = L->ac[j];
    const int m = lower_of(L, pos, j, nb);
    for (int e = 0; e < m; ++e) {
      const int32_t k = nb[e];
      const double l = F[entry_of(L, j, k)] * dinv[k];
      for (int32_t q = L->ptr[k]; q < L->ptr[k + 1]; ++q) {
        const int32_t c = L->col[q];
        if (pos[c] <= pos[k]) continue;
        if (c == j) {
          d = d - l * F[q];
        } else {
          const int32_t r = entry_of(L, j, c);
          if (r >= 0) F[r] = F[r] - l * F[q];
        }
      }
    }
    dinv[j] = 1.0 / d;
  }
  free(order); free(pos); free(nb);
}

void orc_mcilu_sweep(const orc_level *L, const int32_t *colour, int ncol, const double *dinv,
                     const double *b, int from_zero, double *rc, double *x) {
  orc_mcilu_sweep_f(L, colour, ncol, dinv, L->anb, b, from_zero, rc, x);
}

void orc_mcilu_sweep_f(const orc_level *L, const int32_t *colour, int ncol, const double *dinv,
                       const double *F, const double *b, int from_zero, double *rc, double *x) {
  const int32_t n = L->n;
  int32_t *order = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  int32_t *pos = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  mc_order(colour, ncol, n, order, pos);
  /* forward, colours upward: rc_j = (b - A x)_j - sum over earlier-colour
   * neighbours (entry order) of a_ji (dinv_i rc_i); from zero x = 0 and the
   * residual part is b_j */
  for (int32_t t = 0; t < n; ++t) {
    const int32_t j = order[t];
    double res = from_zero ? b[j] : b[j] - L->ac[j] * x[j];
    double low = 0.0;
    for (int32_t k = L->ptr[j]; k < L->ptr[j + 1]; ++k) {
      const int32_t i = L->col[k];
      if (!from_zero) res = res - L->anb[k] * x[i];
      if (colour[i] < colour[j]) low = low + F[k] * (dinv[i] * rc[i]);
    }
    rc[j] = res - low;
  }
  /* backward, colours downward: z_i = dinv_i (rc_i - sum over later-colour
   * neighbours of a_ij z_j), x += z */
  for (int32_t t = n - 1; t >= 0; --t) {
    const int32_t i = order[t];
    double v = rc[i];
    for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k)
      if (colour[L->col[k]] > colour[i]) v = v - F[k] * rc[L->col[k]];
    const double z = dinv[i] * v;
    rc[i] = z;
    x[i] = from_zero ? z : x[i] + z;
  }
  free(order); free(pos);
}

/* dense LU with partial pivoting of a (row-major n x n), in place */
typedef struct {
  int32_t n;
  double *a;
  int32_t *piv;
} orc_lu;

static void lu_factor(orc_lu *F) {
  const int32_t n = F->n;
  double *a = F->a;
  for (int32_t k = 0; k < n; ++k) {
    int32_t p = k;
    for (int32_t i = k + 1; i < n; ++i)
      if (fabs(a[(size_t)i * n + k]) > fabs(a[(size_t)p * n + k])) p = i;
    F->piv[k] = p;
    if (p != k)
      for (int32_t j = 0; j < n; ++j) {
        double t = a[(size_t)k * n + j];
        a[(size_t)k * n + j] = a[(size_t)p * n + j];
        a[(size_t)p * n + j] = t;
      }
    for (int32_t i = k + 1; i < n; ++i) {
      double l = a[(size_t)i * n + k] / a[(size_t)k * n + k];
      a[(size_t)i * n + k] = l;
      for (int32_t j = k + 1; j < n; ++j) a[(size_t)i * n + j] -= l * a[(size_t)k * n + j];
    }
  }
}

static void lu_solve(const orc_lu *F, double *b) {
  const int32_t n = F->n;
  const double *a = F->a;
  for (int32_t k = 0; k < n; ++k) {
    int32_t p = F->piv[k];
    double t = b[k]; b[k] = b[p]; b[p] = t;
    for (int32_t i = k + 1; i < n; ++i) b[i] -= a[(size_t)i * n + k] * b[k];
  }
  for (int32_t i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int32_t j = i + 1; j < n; ++j) s -= a[(size_t)i * n + j] * b[j];
    b[i] = s / a[(size_t)i * n + i];
  }
}

/* coarsest operator (+ mean(diag)/n * 11^T when singular), factored */
static void lu_build(const orc_level *L, int singular, orc_lu *F) {
  const int32_t n = L->n;
  F->n = n;
  F->a = (double *)calloc((size_t)n * n, sizeof(double));
  F->piv = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  double shift = 0.0;
  if (singular) {
    for (int32_t i = 0; i < n; ++i) shift += L->ac[i];
    shift = shift / ((double)n * n);
  }
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t j = 0; j < n; ++j) F->a[(size_t)i * n + j] = shift;
    F->a[(size_t)i * n + i] += L->ac[i];
    for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) F->a[(size_t)i * n + L->col[k]] += L->anb[k];
  }
  lu_factor(F);
}

static void lu_free(orc_lu *F) {
  free(F->a);
  free(F->piv);
}

/* damped Jacobi sweep y = v + w (b - A v)/ac, v = x (+ xc[parent] if given) */
static void jacobi(const orc_level *L, const double *b, const double *x, const int32_t *parent,
                   const double *xc, double omega, double *y) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int32_t i = 0; i < L->n; ++i) {
    double vi = x[i];
    if (parent) vi = vi + xc[parent[i]];
    double r = b[i] - L->ac[i] * vi;
    for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) {
      const int32_t j = L->col[k];
      double vj = x[j];
      if (parent) vj = vj + xc[parent[j]];
      r = r - L->anb[k] * vj;
    }
    y[i] = vi + omega * r / L->ac[i];
  }
}

static void vcyc_rec(const orc_hier *H, int32_t *const *colours, const int *ncol,
                     const orc_vcfg *cfg, const orc_lu *lu, const double *dc0, int l,
                     const double *b, double *x) {
  const orc_level *L = &H->lev[l];
  const int32_t n = L->n;
  if (l == cfg->nlev - 1) {
    memcpy(x, b, (size_t)n * sizeof(double));
    lu_solve(lu, x);
    return;
  }
  memset(x, 0, (size_t)n * sizeof(double));
  const double omega = 2.0 / 3.0;
  if (l > 0 && cfg->coarse_smoother == 1) {
    double *y = (double *)malloc((size_t)n * sizeof(double));
    for (int s = 0; s < cfg->cpre; ++s) {
      jacobi(L, b, x, NULL, NULL, omega, y);
      memcpy(x, y, (size_t)n * sizeof(double));
    }
    double *r = (double *)malloc((size_t)n * sizeof(double));
#ifdef _OPENMP
    #pragma omp parallel for schedule(static)
    #endif
    for (int32_t i = 0; i < n; ++i) {
      double s = b[i] - L->ac[i] * x[i];
      for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) s = s - L->anb[k] * x[L->col[k]];
      r[i] = s;
    }
    const orc_level *C = &H->lev[l + 1];
    double *bc = (double *)calloc((size_t)C->n, sizeof(double));
    double *xc = (double *)calloc((size_t)C->n, sizeof(double));
    for (int32_t i = 0; i < n; ++i) bc[L->parents[i]] = bc[L->parents[i]] + r[i];
    for (int32_t q = 0; q < C->n; ++q) bc[q] = cfg->coarse_scale * bc[q];
    vcyc_rec(H, colours, ncol, cfg, lu, dc0, l + 1, bc, xc);
    if (cfg->cpost == 0) {
      for (int32_t i = 0; i < n; ++i) x[i] = x[i] + xc[L->parents[i]];
    } else {
      jacobi(L, b, x, L->parents, xc, omega, y);
      memcpy(x, y, (size_t)n * sizeof(double));
      for (int s = 1; s < cfg->cpost; ++s) {
        jacobi(L, b, x, NULL, NULL, omega, y);
        memcpy(x, y, (size_t)n * sizeof(double));
      }
    }
    free(y); free(r); free(bc); free(xc);
    return;
  }
  if (l == 0 && (cfg->smoother == 1 || cfg->smoother == 2 || cfg->smoother == 3)) {
    /* DILU pre and post: the reference's in its lexicographic order (1), or
     * in colour order (2); ILU(0) in colour order (3: dc0 holds dinv, then
     * the factor values F) */
    double *rc = (double *)malloc((size_t)n * sizeof(double));
    const double *Fv = cfg->smoother == 3 ? dc0 + n : L->anb;
    for (int s = 0; s < cfg->pre; ++s) {
      if (cfg->smoother == 1) ilu_apply(L, dc0, b, rc, x);
      else orc_mcilu_sweep_f(L, colours[0], ncol[0], dc0, Fv, b, s == 0, rc, x);
    }
    double *r = (double *)malloc((size_t)n * sizeof(double));
#ifdef _OPENMP
    #pragma omp parallel for schedule(static)
    #endif
    for (int32_t i = 0; i < n; ++i) {
      double s = b[i] - L->ac[i] * x[i];
      for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) s = s - L->anb[k] * x[L->col[k]];
      r[i] = s;
    }
    const orc_level *C = &H->lev[l + 1];
    double *bc = (double *)calloc((size_t)C->n, sizeof(double));
    double *xc = (double *)calloc((size_t)C->n, sizeof(double));
    for (int32_t i = 0; i < n; ++i) bc[L->parents[i]] = bc[L->parents[i]] + r[i];
    for (int32_t q = 0; q < C->n; ++q) bc[q] = cfg->coarse_scale * bc[q];
    vcyc_rec(H, colours, ncol, cfg, lu, dc0, l + 1, bc, xc);
    for (int32_t i = 0; i < n; ++i) x[i] = x[i] + xc[L->parents[i]];
    for (int s = 0; s < cfg->post; ++s) {
      if (cfg->smoother == 1) ilu_apply(L, dc0, b, rc, x);
      else orc_mcilu_sweep_f(L, colours[0], ncol[0], dc0, Fv, b, 0, rc, x);
    }
    free(rc); free(r); free(bc); free(xc);
    return;
  }
  if (l == 0) {
    for (int s = 0; s < cfg->pre; ++s) orc_mcgs(L, colours[l], ncol[l], b, x);
  } else if (cfg->cpre > 0) {
    bmcgs(L, colours[l], ncol[l], cfg->block_rows, b, NULL, NULL, NULL, 0, cfg->cpre, x);
  }
  double *r = (double *)malloc((size_t)n * sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int32_t i = 0; i < n; ++i) {
    /* level 0 after a symmetric sweep: colour 0 was solved last and its
     * residual (rounding noise) is taken as 0 */
    if (l == 0 && cfg->smoother == 0 && cfg->pre > 0 && colours[0][i] == 0) {
      r[i] = 0.0;
      continue;
    }
    double s = b[i] - L->ac[i] * x[i];
    for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) s = s - L->anb[k] * x[L->col[k]];
    r[i] = s;
  }
  const orc_level *C = &H->lev[l + 1];
  double *bc = (double *)calloc((size_t)C->n, sizeof(double));
  double *xc = (double *)calloc((size_t)C->n, sizeof(double));
  for (int32_t i = 0; i < n; ++i) bc[L->parents[i]] = bc[L->parents[i]] + r[i];
    for (int32_t q = 0; q < C->n; ++q) bc[q] = cfg->coarse_scale * bc[q];
  vcyc_rec(H, colours, ncol, cfg, lu, dc0, l + 1, bc, xc);
  if (l > 0 && cfg->cpost > 0) {
    double *y = (double *)malloc((size_t)n * sizeof(double));
    bmcgs(L, colours[l], ncol[l], cfg->block_rows, b, x, L->parents, xc, 2, cfg->cpost, y);
    memcpy(x, y, (size_t)n * sizeof(double));
    free(y);
  } else {
    for (int32_t i = 0; i < n; ++i) x[i] = x[i] + xc[L->parents[i]];
    if (l == 0)
      for (int s = 0; s < cfg->post; ++s) orc_mcgs(L, colours[l], ncol[l], b, x);
  }
  free(r); free(bc); free(xc);
}

/* DILU factor of level 0 when the finest smoother is DILU (else NULL) */
static double *finest_dilu(const orc_hier *H, const orc_vcfg *cfg, int32_t *const *colours,
                           const int *ncolours) {
  if ((cfg->smoother < 1 || cfg->smoother > 3) || cfg->nlev < 2) return NULL;
  const orc_level *L = &H->lev[0];
  if (cfg->smoother == 3) {
    double *dcF = (double *)malloc((size_t)(L->n + (L->nnz > 0 ? L->nnz : 1)) * sizeof(double));
    orc_mcilu0_factor(L, colours[0], ncolours[0], dcF, dcF + L->n);
    return dcF;
  }
  double *dc = (double *)malloc((size_t)L->n * sizeof(double));
  if (cfg->smoother == 2) {
    orc_mcilu_factor(L, colours[0], ncolours[0], dc);
    return dc;
  }
  double *rc = (double *)malloc((size_t)L->n * sizeof(double));
  orc_factorize_ilu(L, dc, rc);
  free(rc);
  return dc;
}

void orc_vcycle(const orc_hier *H, int32_t *const *colours, const int *ncolours,
                const orc_vcfg *cfg, const double *b, double *x) {
  orc_lu lu;
  lu_build(&H->lev[cfg->nlev - 1], cfg->singular, &lu);
  double *dc0 = finest_dilu(H, cfg, colours, ncolours);
  vcyc_rec(H, colours, ncolours, cfg, &lu, dc0, 0, b, x);
  free(dc0);
  lu_free(&lu);
}

/* (the OpenMP build sums in per-thread parts: its dots round differently) */
static double dotv(const double *a, const double *b, int32_t n) {
  double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : s) schedule(static)
#endif
  for (int32_t i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

static void project(double *v, int32_t n) {
  double m = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : m) schedule(static)
#endif
  for (int32_t i = 0; i < n; ++i) m += v[i];
  m /= n;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int32_t i = 0; i < n; ++i) v[i] -= m;
}

static void spmv(const orc_level *L, const double *x, double *y) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
  for (int32_t i = 0; i < L->n; ++i) {
    double s = L->ac[i] * x[i];
    for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) s = s + L->anb[k] * x[L->col[k]];
    y[i] = s;
  }
}

int orc_fpcg(const orc_hier *H, int32_t *const *colours, const int *ncolours,
             const orc_vcfg *cfg, const double *b, double tol, int maxit, double *x,
             double *relres) {
  const orc_level *L = &H->lev[0];
  const int32_t n = L->n;
  double *r = (double *)malloc((size_t)n * sizeof(double));
  double *z = (double *)malloc((size_t)n * sizeof(double));
  double *p = (double *)malloc((size_t)n * sizeof(double));
  double *w = (double *)malloc((size_t)n * sizeof(double));
  memcpy(r, b, (size_t)n * sizeof(double));
  if (cfg->singular) project(r, n);
  memset(x, 0, (size_t)n * sizeof(double));
  double nb = sqrt(dotv(r, r, n));
  int it = 0;
  double rr = nb;
  if (nb == 0.0) goto done;
  double delta = 1.0;
  orc_lu lu;
  lu_build(&H->lev[cfg->nlev - 1], cfg->singular, &lu);
  double *dc0 = finest_dilu(H, cfg, colours, ncolours);
  while (it < maxit) {
    vcyc_rec(H, colours, ncolours, cfg, &lu, dc0, 0, r, z);
    if (cfg->singular) project(z, n);
    if (it == 0) {
      memcpy(p, z, (size_t)n * sizeof(double));
    } else {
      double beta = dotv(z, w, n) / delta;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
      for (int32_t i = 0; i < n; ++i) p[i] = z[i] - beta * p[i];
    }
    double rho = dotv(z, r, n);
    spmv(L, p, w);
    delta = dotv(p, w, n);
    double alpha = rho / delta;
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int32_t i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * w[i];
    }
    ++it;
    rr = sqrt(dotv(r, r, n));
    if (rr <= tol * nb) break;
  }
  free(dc0);
  lu_free(&lu);
done:
  if (relres) *relres = nb > 0 ? rr / nb : 0.0;
  free(r); free(z); free(p); free(w);
  return it;
}
