// simple_kernels.h -- gfx950 kernels of the SIMPLE outer iteration: field
// boundary updates, Green-Gauss gradients, momentum and pressure-correction
// assembly (Rhie-Chow), and the corrections.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace ufvm {
namespace sk {

enum { PATCH_WALL = 0, PATCH_EMPTY = 1, PATCH_PATCH = 2, PATCH_SYMMETRY = 3, PATCH_INLET = 4,
       PATCH_OUTLET = 5 };
enum { BC_FIXED_VALUE = 0, BC_ZERO_GRADIENT = 1, BC_EMPTY = 2, BC_NO_SLIP = 3, BC_SYMMETRY = 4,
       BC_SLIP = 5 };

// device pointers of the mesh, passed by value
struct DevMesh {
  int32_t N, F, Fi;
  const int32_t *owner, *neighbour;
  const double *Sf, *area, *CN, *eCN, *gDiff, *T, *gf, *fc, *vol, *cc;
  const int32_t *cptr, *ccol, *own_nnz, *nei_nnz;
  const int32_t *cf_ptr, *cf_face;   // faces of each cell, ascending
  const int32_t *cf_slot;            // per cell-face entry, interior faces: the cell's anb
                                     // slot of the face (>= 0, the cell owns the face) or
                                     // -1 - slot (the cell is the neighbour)
  const int32_t *bpatch;             // patch of each boundary face
  const int32_t *ptype, *u_bc, *p_bc;
  const double *u_val;
  const double *p_val;               // [patches] fixed p of outlets
  const double *walldist;            // [F] (CN . Sf) / |Sf| of boundary faces, 0 inside
  const double *noslip_sum;          // [patches][3] sum over the patch of (Sf_c / area)^2
};

// the density and the transient form of the momentum equation
struct Flow {
  double rho;              // constant density
  int transient;           // 1: ddt(rho*U), Euler, time step dt; 0: false transient
  double dt;
  const double *U_old;     // the previous time step's U and mdot_f
  const double *mdot_old;
};

// scalar slots written by the reductions of this file
enum { S_UMAX = 0, S_UMIN = 1, S_SCALE = 2, S_RMAX = 3, S_RSQ = 4 };

void update_u_bc(hipStream_t s, const DevMesh &m, double *U);
void update_p_bc(hipStream_t s, const DevMesh &m, double *p);
void gradient(hipStream_t s, const DevMesh &m, const double *phi, int nc, double *grad);
void gamma_faces(hipStream_t s, const DevMesh &m, double nu, double *gamma);
// zero: the max and min also take the value 0 (the zero-filled rows the
// reference's U gains past its end from a fixed-pressure outlet's correction)
void scale_u(hipStream_t s, const DevMesh &m, const double *U, bool zero, double *partial,
             unsigned *counter, double *scal);
void assemble_u(hipStream_t s, const DevMesh &m, const Flow &fl, int comp, double urf,
                const double *U,
                const double *gradU, const double *gradp, const double *mdot,
                const double *gamma, double *uf, double *stress, double *ac, double *anb,
                double *bc, double *ac_old, double *dphi, double *DU, double *DUT, double *partial,
                unsigned *counter, double *scal);
void correct_u(hipStream_t s, const DevMesh &m, int comp, const double *dphi, double *U);
void rhie_chow(hipStream_t s, const DevMesh &m, const Flow &fl, double urf_u, const double *U,
               const double *DU, const double *DUT, const double *p, const double *gradp,
               double *mdot, double *fluxC1, double *fluxC2);
// pressure-correction coefficients; *changed is set to 1 when an ac or anb
// value's bits differ from the value it overwrites (never cleared here)
void assemble_p(hipStream_t s, const DevMesh &m, const double *fluxC1, const double *fluxC2,
                const double *mdot, double *ac, double *anb, double *bc, double *ac_old,
                double *dphi, double *partial, double *scal, int32_t *changed);
void correct_p(hipStream_t s, const DevMesh &m, double rho, int32_t ref, double urf_p,
               const double *dphi,
               double *PP, double *gradPP, const double *DU, double *U, double *p,
               const double *fluxC1, const double *fluxC2, double *mdot);

}  // namespace sk
}  // namespace ufvm
