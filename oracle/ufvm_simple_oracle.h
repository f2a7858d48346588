/*
 * ufvm_simple_oracle.h -- CPU restatement of uFVM's steady incompressible
 * SIMPLE outer iteration (momentum assembly, Rhie-Chow face fluxes,
 * pressure-correction assembly, corrections, gradients, boundary updates).
 *
 * TEST INFRASTRUCTURE ONLY (the checker of the HIP implementation and the
 * CPU baseline); nothing in the product library links or calls it.
 *
 * Field layouts (0-based): cell fields hold N interior cells followed by the
 * Fb boundary "elements" in boundary-face order (cfdSetupMeshField.m:36-55);
 * vectors are [cell][3]; gradients of a vector are [cell][comp][3].
 */
#ifndef UFVM_SIMPLE_ORACLE_H
#define UFVM_SIMPLE_ORACLE_H
#include <stdint.h>

#include "ufvm_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

/* physical patch types of the boundary file (cfdReadPolyMesh.m:100-130);
 * ORC_PATCH_INLET / ORC_PATCH_OUTLET are 'inlet' and 'outlet' */
enum { ORC_PATCH_WALL = 0, ORC_PATCH_EMPTY = 1, ORC_PATCH_PATCH = 2, ORC_PATCH_SYMMETRY = 3,
       ORC_PATCH_INLET = 4, ORC_PATCH_OUTLET = 5 };
/* ORC_BC_NO_SLIP: U = 0 on the wall and the reference's vector noSlip wall
 * diffusion (cfdAssembleDiffusionTerm.m:479-562, cfdGetUCoef.m:41-44) */
enum { ORC_BC_FIXED_VALUE = 0, ORC_BC_ZERO_GRADIENT = 1, ORC_BC_EMPTY = 2, ORC_BC_NO_SLIP = 3,
       ORC_BC_SYMMETRY = 4, ORC_BC_SLIP = 5 };

typedef struct {
  int32_t N, F, Fi;
  const int32_t *owner, *neighbour;
  const double *Sf, *area, *CN, *eCN, *gDiff, *T, *gf, *fc; /* fc = face centroid */
  const double *vol, *cc;                                   /* cell volume, centroid */
  const int32_t *cptr, *ccol;                               /* cconn CSR */
  const int32_t *own_nnz, *nei_nnz;                         /* CSR slot of each interior face */
  int32_t npatch;
  const int32_t *pstart, *pnf, *ptype;                      /* patch ranges and physical type */
  const int32_t *u_bc, *p_bc;                               /* per patch BC kinds */
  const double *u_val;                                      /* per patch fixed U [3] */
  const double *p_val;                                      /* per patch fixed p */
} orc_mesh;

typedef struct {
  double nu, urf_u, urf_p;
  int32_t n_correctors;
  int32_t p_ref;          /* 0-based pressure reference cell, -1 none */
  int32_t u_sweeps;       /* ILU sweeps of the momentum solve (10) */
  /* pressure-correction AMG (cfdApplyAMG settings) */
  int32_t p_cycle, p_max_cycles, p_max_coarse, p_pre, p_post, p_smoother;
  double p_termination;
  int32_t u_smoother;     /* momentum solve: 0 the reference's DILU, 2 multicolour GS */
  int32_t stress;         /* 1: the momentum equation has + div(nu*transp(grad(U))) */
  double u_init[3];       /* uniform internal U of the 0 directory */
  double p_init;          /* uniform internal p */
  double rho;             /* constant density of ddt(rho*U) / div(rho*U*U) (0: 1) */
  int32_t transient;      /* 1: ddt(rho*U) with the Euler scheme, time step dt */
  double dt;
  /* pressure-correction solve: 0 cfdApplyAMG with the settings above; 1 an
   * FP64 flexible CG preconditioned by one orc_vcycle (no reference
   * counterpart: the precision form's algorithm with an FP64 V-cycle) on a
   * hierarchy rebuilt every solve, from dphi = 0 to ||r||_2 <= pcg_tol
   * ||b||_2: symmetric multicolour GS 1/1 on level 1, damped Jacobi 2/2
   * below, corrections scaled by 2.5, dense solve on the first level of at
   * most 2000 cells, the constant nullspace projected when every row sum
   * vanishes */
  int32_t p_solver;
  double pcg_tol;
} orc_simple_cfg;

typedef struct {
  int32_t N, Fb;
  double *U, *gradU;      /* [(N+Fb)*3], [(N+Fb)*9] */
  double *p, *gradp;      /* [(N+Fb)], [(N+Fb)*3] */
  double *PP, *gradPP;
  double *DU, *DUT;       /* [(N+Fb)*3] */
  double *mdot;           /* [F] */
  double *gamma_f;        /* [F] */
  double *fluxC1, *fluxC2;/* [F], kept from the last pressure assembly */
  double *U_old, *mdot_old; /* U and mdot_f of the previous time step ('Step1') */
  double scaleU;
  orc_hier H;             /* level 1 = the equation being solved */
  double *ac_old;
  /* per equation component: U x,y,z and p */
  double rms[4], mx[4], r0[4], r1[4];
  int32_t ncycles;
  /* 1 once a fixed-pressure outlet's velocity correction has written U rows
   * past the field's end with unwritten rows between (the reference's U grows
   * and zero-fills them; only cfdUpdateScale's max/min sees those rows) */
  int32_t scale_zero;
} orc_simple_state;

void orc_simple_init(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s);
void orc_simple_free(orc_simple_state *s);

/* steps of one outer iteration (cfdRunCase.m:92-142) */
void orc_update_fields(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s);
void orc_assemble_u(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s, int comp);
void orc_solve_u(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s, int comp);
void orc_correct_u(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s, int comp);
void orc_assemble_p(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s);
void orc_solve_p(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s);
void orc_correct_p(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s);
/* one whole outer iteration */
void orc_simple_iteration(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s);
/* cfdTransientUpdate (cfdUpdateTransientEquation.m, cfdUpdateTransientMdot.m):
 * the previous time step's U and mdot_f become the current ones */
void orc_transient_update(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s);
/* one time step of cfdRunCase's transient branch (cfdRunCase.m:24-86): the
 * transient update, then up to max_inner outer iterations, ending after the
 * first whose RMS residuals (U components, p) are all <= u_tol / p_tol
 * (cfdCheckConvergence.m); returns the iterations run */
int orc_transient_step(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s,
                       int max_inner, double u_tol, double p_tol);

/* Green-Gauss gradient + wall-patch boundary gradients of a field with nc
 * components (cfdUpdateGradient.m / cfdComputeGradientGauss0.m) */
void orc_gradient(const orc_mesh *m, const double *phi, int nc, double *grad);

/* the 'Average:Corrected' gradient of every interior face, component c of an
 * nc-component field (the face gradient of the diffusion correction and of
 * the Rhie-Chow interpolation): out[f*3 + d] */
void orc_face_gradients(const orc_mesh *m, const double *phi, const double *grad, int nc, int c,
                        double *out);

#ifdef __cplusplus
}
#endif
#endif
