/*
 * ufvm.h -- C ABI of the MI355X (gfx950) uFVM pressure-solve path.
 *
 * The reference (en9apr/Matlab_uFVM) is pure MATLAB with no FFI of its own;
 * its "API" for this path is a set of MATLAB functions operating on
 * `global Domain`.  Each entry point below replaces one of those functions
 * (cited as uFVM/<file>:<line>); a MEX gateway or a ctypes host marshals the
 * Domain pieces into these plain arrays.
 *
 * Conventions
 *  - Every function returns 0 (UFVM_OK) or a UFVM_ERR_* status; on error the
 *    message is available from ufvm_last_error(ctx).  Nothing throws across
 *    the ABI.
 *  - Indices are 0-based int32.  A level's coefficients use the reference's
 *    LDU-in-cell-lists layout (uFVM/src/fvm/setup/cfdSetupCoefficients.m:7-39)
 *    flattened to CSR: row i holds cconn{i} (reference order) at
 *    col[ptr[i] .. ptr[i+1]) and anb{i} at the same positions; ac is the
 *    diagonal, bc the right-hand side, dphi the unknown.
 *  - Array arguments are HOST pointers unless the name ends in `_dev`.
 *    State (matrices, hierarchy, vectors) stays resident on the GPU between
 *    calls; only scalars cross the boundary during a solve.
 *  - A context owns one HIP stream; calls on one context must come from one
 *    host thread at a time.
 */
#ifndef UFVM_H
#define UFVM_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define UFVM_OK 0
#define UFVM_ERR_ARG 1         /* bad argument / size */
#define UFVM_ERR_HIP 2         /* HIP runtime error */
#define UFVM_ERR_STATE 3       /* call out of order (e.g. no hierarchy yet) */
#define UFVM_ERR_NUMERIC 4     /* NaN / zero diagonal / breakdown */
#define UFVM_ERR_UNSUPPORTED 5 /* option not supported (e.g. W-cycle with < 4 levels) */

#define UFVM_SMOOTHER_ILU 0    /* 'ILU' (DILU) -- cfdSolveAlgebraicSystem.m:29-46 */
#define UFVM_SMOOTHER_SOR 1    /* 'SOR' (symmetric GS) -- cfdSolveAlgebraicSystem.m:47-62 */
#define UFVM_SMOOTHER_MCGS 2   /* symmetric multicolour GS: greedy first-fit colouring in index
                                  order, colours ascending then descending; the colour classes
                                  run in parallel (no reference counterpart: the GPU form of
                                  'SOR' for solves that need not follow the reference's order) */

#define UFVM_CYCLE_V 0
#define UFVM_CYCLE_W 1
#define UFVM_CYCLE_F 2

typedef struct ufvm_ctx ufvm_ctx;

/* ---------------------------------------------------------------- context */
/* No reference counterpart (MATLAB holds state in `global Domain`,
 * uFVM/src/fvm/setup/cfdSetupDomain.m:11-63). */
int ufvm_ctx_create(int32_t device, ufvm_ctx **out);
int ufvm_ctx_destroy(ufvm_ctx *ctx);
const char *ufvm_last_error(const ufvm_ctx *ctx);
/* the context's hipStream_t (for event timing by a host that shares it) */
void *ufvm_ctx_stream(ufvm_ctx *ctx);
/* wait for the context stream and for the whole device */
int ufvm_ctx_synchronize(ufvm_ctx *ctx);

/* ---------------------------------------------------------- linear system */
/* Create the level-1 system of one equation from the mesh connectivity
 * (theMesh.cconn / theMesh.csize, uFVM/src/fvm/assemble/cfdAssembleEquation.m:18-20
 * via cfdSetupCoefficients).  Coefficients start at zero. */
int ufvm_system_create(ufvm_ctx *ctx, int32_t n, int32_t nnz, const int32_t *ptr,
                       const int32_t *col, int32_t *sys_out);
int ufvm_system_destroy(ufvm_ctx *ctx, int32_t sys);

/* Domain.coefficients{1} = {ac, anb, bc} (cfdSetCoefficients.m:3-9).  Any of
 * the pointers may be NULL to leave that array unchanged.  Setting level-1
 * coefficients zeroes dphi, as cfdSetupCoefficients.m:17 does. */
int ufvm_system_set_coefficients(ufvm_ctx *ctx, int32_t sys, const double *ac,
                                 const double *anb, const double *bc);
/* same, from device memory (e.g. a torch tensor on the context's device) */
int ufvm_system_set_coefficients_dev(ufvm_ctx *ctx, int32_t sys, const double *ac_dev,
                                     const double *anb_dev, const double *bc_dev);
/* New level-1 coefficients on the existing hierarchy (no reference
 * counterpart: the reference re-agglomerates every solve).  Parents and
 * coarse patterns are kept; the Galerkin coarse operators are recomputed
 * (assembleAgglomeratedLHS.m:16-30 arithmetic) and, after ufvm_pcg_setup,
 * the solve-side copies are refreshed on the GPU.  NULL leaves an array. */
int ufvm_system_update_coefficients(ufvm_ctx *ctx, int32_t sys, const double *ac,
                                    const double *anb, const double *bc);
/* Domain.coefficients{level}.dphi read/write (cfdGetCoefficients.m:6-12) */
int ufvm_system_set_dphi(ufvm_ctx *ctx, int32_t sys, int32_t level, const double *dphi);
int ufvm_system_get_dphi(ufvm_ctx *ctx, int32_t sys, int32_t level, double *dphi);
/* Domain.coefficients{level}.bc read/write */
int ufvm_system_set_bc(ufvm_ctx *ctx, int32_t sys, int32_t level, const double *bc);
int ufvm_system_get_bc(ufvm_ctx *ctx, int32_t sys, int32_t level, double *bc);

/* r = cfdComputeResidualsArray(Domain.coefficients{level})
 * (uFVM/src/fvm/compute/cfdComputeResidualsArray.m:11-17): bit-exact, rows
 * summed in cconn order.  r_out (host, may be NULL) receives r; l1_out (may
 * be NULL) receives sum(abs(r)). */
int ufvm_compute_residuals(ufvm_ctx *ctx, int32_t sys, int32_t level, double *r_out,
                           double *l1_out);

/* No reference counterpart: for callers that cannot allocate device memory
 * (work captured into HIP graphs, tight serving loops, harnesses that forbid
 * allocation).  The residual r and its L1 norm sum(abs(r)) of
 * ufvm_compute_residuals, with the same bits, written to DEVICE memory of the
 * context's GPU (r_dev: n doubles, l1_dev: one double) and enqueued on the
 * context's stream without waiting for it.  The call allocates nothing: its
 * per-block partial sums live in a static device array of the library sized
 * for at most UFVM_NOALLOC_MAX_ROWS rows (1 MiB per GPU, held until the
 * process ends); larger levels are refused (UFVM_ERR_ARG).  Every call on one
 * GPU shares that array, so calls that overlap on one GPU (other contexts'
 * streams, other host threads) race. */
#define UFVM_NOALLOC_MAX_ROWS 33554432
int ufvm_compute_residuals_noalloc(ufvm_ctx *ctx, int32_t sys, int32_t level, double *r_dev,
                                   double *l1_dev);

/* ----------------------------------------------------- AMG (agglomeration) */
/* cfdAgglomerate(maxCoarseLevels) (uFVM/src/fvm/solve/AMG/cfdAgglomerate.m:7-17):
 * builds the coarse levels from the current level-1 coefficients with the
 * reference's greedy agglomeration (cfdAgglomerateLevel.m:3-104) and Galerkin
 * coarse operators (assembleAgglomeratedLHS.m:16-30).  Parents and coarse
 * cconn are bit-exact with the reference; coarse coefficients are summed in
 * the reference order on the GPU.  *nlevels receives maxLevels. */
int ufvm_amg_agglomerate(ufvm_ctx *ctx, int32_t sys, int32_t max_coarse_levels,
                         int32_t *nlevels);
int ufvm_amg_num_levels(ufvm_ctx *ctx, int32_t sys, int32_t *nlevels);
/* size of Domain.coefficients{level} (0-based level index) */
int ufvm_amg_level_info(ufvm_ctx *ctx, int32_t sys, int32_t level, int32_t *n,
                        int32_t *nnz);
/* copy out level data; any pointer may be NULL.  parents (length n) maps the
 * level to level+1 and is only defined below the coarsest level. */
int ufvm_amg_level_get(ufvm_ctx *ctx, int32_t sys, int32_t level, int32_t *parents,
                       int32_t *ptr, int32_t *col, double *ac, double *anb);

/* cfdSolveAlgebraicSystem(level, smoother, nIter, rrf)
 * (uFVM/src/fvm/solve/cfdSolveAlgebraicSystem.m:11-66) on the GPU with the
 * reference's lexicographic SOR / DILU orderings (level-scheduled, bit-exact
 * per sweep) and its mean-|r| early exit. */
int ufvm_solve_algebraic_system(ufvm_ctx *ctx, int32_t sys, int32_t level, int32_t smoother,
                                int32_t n_iter, double rrf, double *r0, double *r1);

/* Settings of theEquation.multigrid / theSolutionMethods
 * (uFVM/src/fvm/top/cfdDefineEquation.m:118-136, cfdSetupAMG.m:6-18). */
typedef struct {
  int32_t cycle;             /* UFVM_CYCLE_V / _W / _F */
  int32_t max_cycles;        /* maxCycles (30) */
  double termination;        /* termination (0.1) */
  int32_t max_coarse_levels; /* maxCoarseLevels (10) */
  int32_t pre_sweep;         /* preSweep (1) */
  int32_t post_sweep;        /* postSweep (3) */
  int32_t smoother;          /* UFVM_SMOOTHER_ILU / _SOR */
  int32_t rebuild;           /* 1: re-agglomerate on every call as cfdApplyAMG.m:19 does */
} ufvm_amg_cfg;

/* [r0, r1] = cfdApplyAMG(theEquation) (uFVM/src/fvm/solve/AMG/cfdApplyAMG.m:9-60):
 * stationary V/W/F cycles on the level-1 system with the reference's L1
 * termination; *ncycles receives the reference's nCycle counter. */
int ufvm_apply_amg(ufvm_ctx *ctx, int32_t sys, const ufvm_amg_cfg *cfg, double *r0,
                   double *r1, int32_t *ncycles);

/* ------------------------------------------------ precision-form AMG-PCG */
/* No reference counterpart: flexible PCG (FP64) preconditioned by one
 * AMG V-cycle over the reference's hierarchy, with multicolour Gauss-Seidel
 * (or DILU) smoothing on the finest level, damped Jacobi or block multicolour
 * GS below it, and an FP64 dense solve on the first level with at most
 * max_coarse_cells cells. */
typedef struct {
  int32_t precision;         /* 32: FP32 V-cycle, 64: FP64 V-cycle, 16: FP32 V-cycle whose
                                finest-level off-diagonals are stored in FP16, each
                                divided by its row's diagonal (multicolour GS finest
                                smoother only; no batched or grouped solves) */
  int32_t pre_sweep;         /* smoothing sweeps before restriction */
  int32_t post_sweep;        /* smoothing sweeps after prolongation */
  int32_t coarse_smoother;   /* levels below the first: 0 block-local multicolour GS,
                                1 damped Jacobi (omega = 2/3); level 1 always uses
                                symmetric multicolour GS */
  int32_t max_coarse_cells;  /* dense-solve threshold (<= 4096) */
  int32_t nullspace;         /* -1 auto-detect (zero row sums), 0 no, 1 constant nullspace */
  int32_t max_iter;
  double tol;                /* stop when ||r||_2 <= tol * ||b||_2 */
  int32_t smoother;          /* finest level: 0 multicolour Gauss-Seidel (colour classes in
                                parallel), 1 the reference's DILU ('ILU' smoother,
                                cfdFactorizeILU.m / cfdSolveILU.m) in its lexicographic
                                order, level-scheduled, 2 DILU in the colour order of 0
                                (one level per colour class), 3 ILU(0) in that colour
                                order (the factor's off-diagonals updated inside the
                                pattern; equal to 2 when the matrix graph has no
                                triangles, as on hexahedral and 6-tets-per-hex meshes) */
  double coarse_scale;       /* every coarse-grid correction is scaled by this factor
                                (1: plain; 2-2.5 suits the piecewise-constant
                                prolongation of aggregation AMG); in (0, 4], 0 = 2.5 */
  int32_t coarse_pre_sweep;  /* sweeps below the finest level; -1: pre_sweep */
  int32_t coarse_post_sweep; /* -1: post_sweep */
  int32_t launch_mode;       /* FCG iterations: 0 auto (replayed as captured HIP graphs below
                                600000 cells, launched kernel by kernel above), 1 always
                                kernel by kernel, 2 always graphs */
  int32_t index_width;       /* column indices of the sliced-ELL level copies: 0 auto (a
                                uint16 offset per entry from a base per 64-row chunk on
                                every level whose chunks each span fewer than 65536
                                columns, int32 elsewhere), 4 always int32.  Results are
                                the same either way. */
} ufvm_pcg_cfg;

/* Build the solve-side data of the current hierarchy (colourings, permuted
 * FP32/FP64 level copies, coarsest inverse).  Requires ufvm_amg_agglomerate. */
int ufvm_pcg_setup(ufvm_ctx *ctx, int32_t sys, const ufvm_pcg_cfg *cfg);
/* Solve A dphi = bc for level 1 from dphi = 0.  On return dphi holds the
 * solution (device; read with ufvm_system_get_dphi).  */
int ufvm_pcg_solve(ufvm_ctx *ctx, int32_t sys, int32_t *iters, double *relres);
/* k = 1, 2, 4 or 8 independent solves of the same system at once (no
 * reference counterpart: throughput form of ufvm_pcg_solve).  b: k right-hand
 * sides, x: k solutions, each n values in level-1 order one after another, in
 * host or device memory.  Every kernel of an iteration serves all k (the
 * matrices are read once per k solves); each solution, iteration count and
 * relative residual is the one ufvm_pcg_solve gives for that right-hand side.
 * Needs the multicolour GS finest smoother and Jacobi coarse levels. */
int ufvm_pcg_solve_batch(ufvm_ctx *ctx, int32_t sys, int32_t k, const double *b, double *x,
                         int32_t *iters, double *relres);
/* g solves of g different systems of the context at once (no reference
 * counterpart: the serving form's grouped launches, exposed for callers and
 * tests).  b[q] / x[q]: host arrays of system sys[q]'s size, level-1 order.
 * The solves advance in lock step, one FCG iteration at a time, and each
 * phase of an iteration -- every colour pass, residual, restriction, Jacobi
 * sweep, prolongation and dense solve of the V-cycles and every FCG kernel --
 * is one launch serving all of them.  Each solution, iteration count and
 * relative residual is the one ufvm_pcg_solve gives for that system and
 * right-hand side.  Needs, on every system, the multicolour GS finest
 * smoother, Jacobi coarse levels with at least one pre-sweep, and one V-cycle
 * precision and max_iter across the group; a system may appear once. */
int ufvm_pcg_solve_group(ufvm_ctx *ctx, int32_t g, const int32_t *sys, const double *const *b,
                         double *const *x, int32_t *iters, double *relres);
/* Apply one preconditioner V-cycle: z = M^-1 r (host arrays, level-1 order). */
int ufvm_pcg_apply_vcycle(ufvm_ctx *ctx, int32_t sys, const double *r, double *z);
/* number of colours of a level's multicolour ordering (after ufvm_pcg_setup) */
int ufvm_pcg_level_colours(ufvm_ctx *ctx, int32_t sys, int32_t level, int32_t *ncolours,
                           int32_t *colour);
/* bytes per column index of a level's sliced-ELL copy (2, 4, or 0 when the
 * level has none) and its slots (entries plus padding) */
int ufvm_pcg_level_layout(ufvm_ctx *ctx, int32_t sys, int32_t level, int32_t *index_bytes,
                          int64_t *slots);


/* ------------------------------------------------- SIMPLE outer iteration */
/* Processed mesh (the reference's theMesh after cfdProcessOpenFoamMesh,
 * uFVM/src/utilities/IO/cfdProcessOpenFoamMesh.m:15-241), 0-based, vectors as
 * [face][3] / [cell][3].  Interior faces come first and are owner < neighbour;
 * boundary faces follow, grouped by patch. */
#define UFVM_PATCH_WALL 0
#define UFVM_PATCH_EMPTY 1
/* a symmetry plane (U and p conditions UFVM_BC_SYMMETRY): on the steady path
 * the reference's symmetry branches do its empty branches' arithmetic
 * (cfdAssembleDiffusionTerm.m:567-626, cfdCorrectScalarEquation.m:469-505,
 * cfdCorrectVelocityField.m:274-305, cfdCorrectMdotField.m:306-322) */
#define UFVM_PATCH_SYMMETRY 3
/* 'inlet' and 'outlet' patches of the boundary file: an inlet takes U
 * fixedValue and p zeroGradient (cfdAssembleMdotTerm.m:415-486,
 * cfdAssembleConvectionTerm.m:144-186, cfdAssembleDiffusionTerm.m:135-173);
 * an outlet U zeroGradient and p fixedValue (cfdAssembleMdotTerm.m:799-906,
 * cfdAssembleConvectionTerm.m:191-232, cfdAssembleDiffusionTerm.m:201-218),
 * with the reference's corrections (cfdCorrectVelocityField.m:247-272 writes
 * the owners' U into rows numbered by the outlet's FACE indices: kept; rows
 * past the end of U are not stored, and their zero fill enters the velocity
 * scale as in the reference; an outlet face index below n_cells is refused) */
#define UFVM_PATCH_INLET 4
#define UFVM_PATCH_OUTLET 5
#define UFVM_BC_FIXED_VALUE 0
#define UFVM_BC_ZERO_GRADIENT 1
#define UFVM_BC_EMPTY 2
/* U on a wall: zero, with the reference's vector noSlip wall diffusion
 * (cfdAssembleDiffusionTerm.m:479-562, TM from cfdGetUCoef.m:41-44) */
#define UFVM_BC_NO_SLIP 3
#define UFVM_BC_SYMMETRY 4
/* U on a wall: slip -- no diffusion or mass flux through the wall
 * (cfdAssembleDiffusionTerm.m:456-478, cfdAssembleMdotTerm.m:395-410), and
 * U_b,c the owner's U_c minus its normal part after each momentum correction
 * (cfdCorrectScalarEquation.m:239-293).  The pressure-stage corrections
 * dispatch on p's condition (cfdCorrectVelocityField.m:57-121,
 * cfdCorrectMdotField.m:56-120): with p zeroGradient the wall takes the
 * no-op wall branch, so U_b keeps its momentum-stage value. */
#define UFVM_BC_SLIP 5
typedef struct {
  int32_t n_cells, n_faces, n_interior;
  const int32_t *owner, *neighbour;
  const double *Sf, *area, *CN, *eCN, *gDiff, *T, *gf, *face_centroid;
  const double *volume, *centroid;
  const int32_t *cconn_ptr, *cconn;     /* theMesh.cconn as CSR, reference order */
  const int32_t *own_nnz, *nei_nnz;     /* iOwnerNeighbourCoef / iNeighbourOwnerCoef as CSR slots */
  int32_t n_patches;
  const int32_t *patch_start, *patch_nfaces, *patch_type;   /* UFVM_PATCH_* */
} ufvm_mesh_desc;

/* Steady incompressible SIMPLE settings: theEquation fields of U and p
 * (uFVM/src/fvm/top/cfdDefineEquation.m:84-136), fvSolution SIMPLE block and
 * the per-patch boundary conditions, by patch type:
 *   wall:     U fixedValue (u_value), noSlip or slip; p zeroGradient
 *   empty:    U and p empty
 *   symmetry: U and p symmetry
 *   inlet:    U fixedValue (u_value); p zeroGradient
 *   outlet:   U zeroGradient; p fixedValue (p_value)
 * Pressure-correction solver: 0 = cfdApplyAMG
 * (stationary cycles, hierarchy rebuilt every solve as the reference does),
 * 1 = AMG-PCG with the hierarchy rebuilt every solve, 2 = AMG-PCG with the
 * hierarchy of the first solve kept. */
typedef struct {
  double nu, urf_u, urf_p;
  int32_t n_correctors;
  int32_t p_ref_cell;                     /* 0-based, -1 none (cfdReadSystem.m:171-180) */
  int32_t u_sweeps;                       /* ILU sweeps of each momentum solve (10) */
  const int32_t *u_bc;                    /* [n_patches] UFVM_BC_* */
  const double *u_value;                  /* [n_patches][3] */
  const int32_t *p_bc;                    /* [n_patches] UFVM_BC_* */
  int32_t p_solver;
  ufvm_amg_cfg p_amg;
  ufvm_pcg_cfg p_pcg;
  int32_t u_smoother;                     /* momentum solves: UFVM_SMOOTHER_ILU (the
                                             reference's, default) or UFVM_SMOOTHER_MCGS */
  int32_t stress_term;                    /* 1: the momentum equation has
                                             + div(nu*transp(grad(U))) (the reference's
                                             tutorials; cfdAssembleExplicitStressTerm.m) */
  const double *p_value;                  /* [n_patches] fixed p of outlets; NULL: 0 */
  double u_init[3];                       /* uniform internal U of the 0 directory
                                             (cfdDefineEquation.m:68-82) */
  double p_init;                          /* uniform internal p */
  double rho;                             /* constant density of div(rho*U*U) and
                                             ddt(rho*U) (0: 1); nu is then the
                                             laplacian's coefficient (mu) */
  int32_t transient;                      /* 1: the momentum equation has ddt(rho*U) with
                                             the Euler scheme (cfdAssembleTransientTermEuler.m,
                                             every cell's source taking cell 1's U and U_old
                                             as the reference's does) and the Rhie-Chow
                                             fluxes terms VI-VII (cfdAssembleMdotTerm.m:
                                             175-201); no false transient.  Steps are
                                             advanced by ufvm_transient_step */
  double dt;                              /* time step of the transient form */
} ufvm_simple_cfg;

/* Upload a mesh and create its fields (the uniform initial fields u_init and
 * p_init; boundary values and mdot_f initialised as cfdInitializeCase.m:17-83,
 * cfdInitializeBoundaryConditions.m, cfdInitializeMdotField.m:72-80). */
int ufvm_case_create(ufvm_ctx *ctx, const ufvm_mesh_desc *mesh, const ufvm_simple_cfg *cfg,
                     int32_t *case_out);
int ufvm_case_destroy(ufvm_ctx *ctx, int32_t cas);
/* Field access (cfdGetMeshField / cfdSetMeshField.m): "U", "p", "PP", "DU",
 * "DUT", "mdot_f", "gradU", "gradp", "gradPP", "gamma_f", and the previous time
 * step's "U_old", "mdot_old".  count must equal the field size. */
int ufvm_case_set_field(ufvm_ctx *ctx, int32_t cas, const char *name, const double *data,
                        int64_t count);
int ufvm_case_get_field(ufvm_ctx *ctx, int32_t cas, const char *name, double *data,
                        int64_t count);
/* the linear system behind equation "U" or "p" (Domain.coefficients{1}) */
int ufvm_case_system(ufvm_ctx *ctx, int32_t cas, const char *eq, int32_t *sys_out);
/* cfdUpdateFields (uFVM/src/fvm/update/cfdUpdateEquations.m:14-24) */
int ufvm_update_fields(ufvm_ctx *ctx, int32_t cas);
/* [rms, max] = cfdAssembleEquation(eq, comp) (cfdAssembleEquation.m:1-27);
 * comp is 0-based; "p" assembles the Rhie-Chow fluxes and the
 * pressure-correction matrix (cfdAssembleMdotTerm.m:38-62). */
int ufvm_assemble_equation(ufvm_ctx *ctx, int32_t cas, const char *eq, int32_t comp,
                           double *rms, double *max);
/* [r0, r1] = cfdSolveEquation(eq) (cfdSolveEquation.m:11-17); *ncycles gets
 * the AMG cycle counter (or the PCG iteration count). */
int ufvm_solve_equation(ufvm_ctx *ctx, int32_t cas, const char *eq, double *r0, double *r1,
                        int32_t *ncycles);
/* cfdCorrectEquation(eq, comp) (cfdCorrectEquation.m:17-51) */
int ufvm_correct_equation(ufvm_ctx *ctx, int32_t cas, const char *eq, int32_t comp);
/* n steady outer iterations of cfdRunCase (cfdRunCase.m:92-142); rms (may be
 * NULL) receives 4 values per iteration: RMS residuals of Ux, Uy, Uz, p. */
int ufvm_simple_iterate(ufvm_ctx *ctx, int32_t cas, int32_t n, double *rms);
/* One time step of cfdRunCase's transient branch (cfdRunCase.m:24-86) for a
 * case created with transient = 1: the previous step's U and mdot_f are kept
 * (cfdTransientUpdate: fields "U_old" and "mdot_old"), then outer iterations
 * run until the first whose RMS residuals are all <= u_tol (U components) and
 * p_tol (p) (cfdCheckConvergence.m), at most max_inner (the reference's 20).
 * *n_inner (may be NULL) receives the iterations run, rms (may be NULL, room
 * for 4 * max_inner) their residuals. */
int ufvm_transient_step(ufvm_ctx *ctx, int32_t cas, int32_t max_inner, double u_tol,
                        double p_tol, double *rms, int32_t *n_inner);
/* cfdTransientUpdate (uFVM/src/fvm/update/cfdTransientUpdate.m) on the path's
 * fields alone: U and mdot_f become the previous step's "U_old" and
 * "mdot_old" (the first part of ufvm_transient_step, for a caller that runs
 * the time loop itself, as cfdRunCase.m:24-86 does through the .m shims). */
int ufvm_transient_update(ufvm_ctx *ctx, int32_t cas);

/* --------------------------------------------------------- serving form */
/* No reference counterpart.  A single-GPU request loop over resident meshes:
 * pressure-Poisson solves (the FP32-preconditioned AMG-PCG of ufvm_pcg_*) and
 * SIMPLE steps (ufvm_simple_iterate).  Solves are continuously batched: up
 * to max_active run at once, each on its own workspace and one of n_streams
 * streams, and a worker thread advances each by one FCG iteration as the GPU
 * finishes it; queued requests are admitted when a slot is free and the
 * device keeps hbm_reserve of its memory free after the request's workspace.
 * SIMPLE steps run on a second thread and stream, overlapping the solves.
 * Results do not depend on what else runs: every solve's arithmetic depends
 * only on its system and right-hand side.  Host arrays handed to a submit
 * call must stay valid until the request is done. */
typedef struct ufvm_server ufvm_server;
typedef struct {
  int32_t max_active;        /* solves on the GPU at once (>= 1) */
  int32_t n_streams;         /* streams the active solves are spread over (>= 1) */
  double hbm_reserve;        /* fraction of device memory admission keeps free */
  int32_t max_batch;         /* queued solves of one system admitted together as one batch of
                                up to this many (2, 4 or 8; 1: no batching; 0: 8).  A
                                batch reads the matrices once per iteration for all its
                                solves and gives each the bits of its own solve */
  int32_t n_steppers;        /* SIMPLE-step lanes (0: 4): each owns the cases assigned to it
                                (the least-loaded lane by cells at ufvm_server_add_case), a
                                context, a stream and a thread, so steps of cases in
                                different lanes overlap on the device */
  int32_t max_group;         /* queued solves of different systems below 600k cells (where
                                launches and latency dominate) admitted together as one
                                group of up to this many (0: 8; 1: no groups) when no batch
                                forms: every phase of their FCG iterations is one launch for
                                all of them (ufvm_pcg_solve_group), and each keeps the bits
                                of its own solve */
  double group_span;         /* a group's systems differ in cell count by at most this factor
                                (0: 4), so that small solves are not paced by large ones */
} ufvm_server_cfg;
typedef struct {
  int32_t status;            /* 0 queued, 1 running, 2 done, < 0: -UFVM_ERR_* */
  int32_t iterations;        /* solve: FCG iterations */
  double relres;             /* solve: ||r|| / ||b|| */
  double queued_s;           /* submit -> admission (seconds) */
  double run_s;              /* admission -> result on the host */
  int32_t batch;             /* solve: solves in the batch it ran in (1: alone) */
  int32_t group;             /* solve: solves of different systems whose launches it shared
                                (1: none) */
} ufvm_request_info;

int ufvm_server_create(int32_t device, const ufvm_server_cfg *cfg, ufvm_server **out);
/* finishes every queued and running request, then frees everything */
int ufvm_server_destroy(ufvm_server *srv);
const char *ufvm_server_last_error(const ufvm_server *srv);
/* a pressure-Poisson system (level-1 CSR in cconn order + coefficients): its
 * hierarchy and precision-form copies are built now and stay resident */
int ufvm_server_add_system(ufvm_server *srv, int32_t n, int32_t nnz, const int32_t *ptr,
                           const int32_t *col, const double *ac, const double *anb,
                           const ufvm_pcg_cfg *cfg, int32_t *sys_out);
/* a SIMPLE case (as ufvm_case_create); its fields stay resident */
int ufvm_server_add_case(ufvm_server *srv, const ufvm_mesh_desc *mesh, const ufvm_simple_cfg *cfg,
                         int32_t *case_out);
int ufvm_server_get_field(ufvm_server *srv, int32_t cas, const char *name, double *data,
                          int64_t count);
int ufvm_server_set_field(ufvm_server *srv, int32_t cas, const char *name, const double *data,
                          int64_t count);
/* solve A x = b (host arrays of the system's size, reference order) */
int ufvm_server_submit_solve(ufvm_server *srv, int32_t sys, const double *b, double *x,
                             int64_t *ticket);
/* nsteps SIMPLE outer iterations; rms (may be NULL): 4 values per step */
int ufvm_server_submit_step(ufvm_server *srv, int32_t cas, int32_t nsteps, double *rms,
                            int64_t *ticket);
int ufvm_server_query(ufvm_server *srv, int64_t ticket, ufvm_request_info *info);
/* blocks until the request is done or failed; after a failure (status < 0)
 * ufvm_server_last_error holds its message */
int ufvm_server_wait(ufvm_server *srv, int64_t ticket, ufvm_request_info *info);
/* forget a finished request */
int ufvm_server_release(ufvm_server *srv, int64_t ticket);
int ufvm_server_stats(ufvm_server *srv, int32_t *queued, int32_t *active, int64_t *work_bytes);

/* ------------------------------------------------------------- profiling */
/* Per-kernel-class HIP-event timing on the context stream (no reference
 * counterpart).  kind names: "mcgs0" (finest-level colour passes of the
 * smoother), "spmv0" (finest-level FP64 operator product of the FPCG),
 * "vcycle" (one whole preconditioner V-cycle), "tail" (the V-cycle below and
 * including the first level of at most 40,000 rows).
 * on = 0 disables; on = k >= 1 times every k-th launch of each kind. */
int ufvm_prof_enable(ufvm_ctx *ctx, int32_t on);
int ufvm_prof_read(ufvm_ctx *ctx, const char *kind, double *total_ms, int64_t *launches,
                   double *bytes);
int ufvm_prof_reset(ufvm_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
