/*
 * ufvm_simple_oracle.c -- CPU restatement of uFVM's steady SIMPLE iteration.
 *
 * TEST INFRASTRUCTURE ONLY (checker of the HIP implementation, CPU baseline).
 *
 * Every loop follows the reference's order (faces ascending, interior before
 * boundary, terms in equation order) and every expression its association, so
 * that the floating-point results are those of the reference's arithmetic.
 * Paths are relative to uFVM/.  Built with -ffp-contract=off.
 */
#include "ufvm_simple_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define FB(m) ((m)->F - (m)->Fi)

static double *dz(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }

static double mag3(const double *v) { return sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]); }

/* empty and symmetry patches: on the steady path every symmetry branch of the
 * reference does the empty branch's arithmetic -- diffusion
 * (cfdAssembleDiffusionTerm.m:567-626 vs :631-690), zero convection
 * (cfdAssembleConvectionTerm.m:285-300), no gradient update
 * (cfdUpdateGradient.m:260), no mdot assembly (cfdAssembleMdotTerm.m:342),
 * U_b = U_O - (U_O . n) n at initialisation and in both corrections
 * (cfdInitializeBoundaryConditions.m:471-530, cfdCorrectScalarEquation.m:469-505,
 * cfdCorrectVelocityField.m:274-305), p_b and PP_b owner copies
 * (cfdCorrectPressureEquation.m:238-261, cfdCorrectPPField.m:278-300), mdot_b =
 * 0 mdot_b (cfdCorrectMdotField.m:306-322) */
static int plane(const orc_mesh *m, int p) {
  return m->ptype[p] == ORC_PATCH_EMPTY || m->ptype[p] == ORC_PATCH_SYMMETRY;
}

/* the momentum corrections dispatch on U's condition
 * (cfdCorrectScalarEquation.m:31-107): U_b,c = U_O,c - (U_O . n) n_c on the
 * planes above and on slip walls (:239-293); an owner copy of U_b,c on
 * zeroGradient outlets (:376-407); no change on fixedValue / noSlip walls and
 * fixedValue inlets (:116-120, :229-233, :301-303).  A slip wall's diffusion
 * (cfdAssembleDiffusionTerm.m:456-478) and mdot (cfdAssembleMdotTerm.m:355-375)
 * fluxes are zero, its gradient rows a wall's. */
static int u_tangential(const orc_mesh *m, int p) {
  return plane(m, p) || (m->ptype[p] == ORC_PATCH_WALL && m->u_bc[p] == ORC_BC_SLIP);
}

/* inlet and outlet patches carry a mass flux in the pressure-correction
 * assembly (cfdAssembleMdotTerm.m:415-486 'inlet' with p zeroGradient,
 * :799-906 'outlet' with p fixedValue); walls and planes do not (:355-410) */
static int p_flux(const orc_mesh *m, int p) {
  return m->ptype[p] == ORC_PATCH_INLET || m->ptype[p] == ORC_PATCH_OUTLET;
}

/* outlet with a fixed pressure: PP_b = 0 (cfdCorrectPPField.m:249-272), p_b
 * kept (cfdCorrectPressureEquation.m:230-233), the owners' U written to rows
 * iBFaces of U (cfdCorrectVelocityField.m:247-272) and mdot_b = rho_b (U_O . S_b)
 * (cfdCorrectMdotField.m:264-304) */
static int p_fixed_outlet(const orc_mesh *m, int p) {
  return m->ptype[p] == ORC_PATCH_OUTLET && m->p_bc[p] == ORC_BC_FIXED_VALUE;
}

static double dot3(const double *a, const double *b) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }

/* patch of boundary face f (absolute index) */
/* face walldist = (CN . Sf) / |Sf| (cfdProcessOpenFoamMesh.m:153) */
static double walldist(const orc_mesh *m, int32_t f) {
  return dot3(&m->CN[f * 3], &m->Sf[f * 3]) / mag3(&m->Sf[f * 3]);
}

/* sum over patch p's faces, in face order, of n_c^2, n = Sf / area: the
 * patch-wide dot(nx',nx') of cfdAssembleDiffusionTerm.m:535-551 */
static double noslip_sum(const orc_mesh *m, int p, int comp) {
  double s = 0.0;
  for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f) {
    const double nc = m->Sf[f * 3 + comp] / m->area[f];
    s = s + nc * nc;
  }
  return s;
}

static int patch_of(const orc_mesh *m, int32_t f) {
  for (int p = 0; p < m->npatch; ++p)
    if (f >= m->pstart[p] && f < m->pstart[p] + m->pnf[p]) return p;
  return -1;
}

void orc_simple_init(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s) {
  const int32_t N = m->N, Fb = FB(m), NB = N + Fb, F = m->F;
  memset(s, 0, sizeof(*s));
  s->N = N;
  s->Fb = Fb;
  s->U = dz((size_t)NB * 3);
  s->gradU = dz((size_t)NB * 9);
  s->p = dz(NB);
  s->gradp = dz((size_t)NB * 3);
  s->PP = dz(NB);
  s->gradPP = dz((size_t)NB * 3);
  s->DU = dz((size_t)NB * 3);
  s->DUT = dz((size_t)NB * 3);
  s->mdot = dz(F);
  s->gamma_f = dz(F);
  s->fluxC1 = dz(F);
  s->fluxC2 = dz(F);
  s->U_old = dz((size_t)NB * 3);
  s->mdot_old = dz(F);
  s->ac_old = dz(N);
  orc_level *L = &s->H.lev[0];
  L->n = N;
  L->nnz = m->cptr[N];
  L->ptr = (int32_t *)m->cptr;
  L->col = (int32_t *)m->ccol;
  L->ac = dz(N);
  L->anb = dz(L->nnz);
  L->bc = dz(N);
  L->dphi = dz(N);
  L->parents = NULL;
  s->H.nlevels = 1;
  /* uniform initial fields (the 0 directory's internalField); boundary values
   * from the BCs (cfdInitializeBoundaryConditions.m): fixedValue walls and
   * inlets the value (:107-148, :321-357), noSlip zero (:239-275), slip walls,
   * empty and symmetry planes the owner minus its normal component (:276-320,
   * :471-595), zeroGradient outlets an owner copy of every component
   * (:426-470); p_b an owner copy, or the value on fixed-pressure outlets
   * (:386-425) */
  for (int32_t i = 0; i < N; ++i) {
    for (int d = 0; d < 3; ++d) s->U[i * 3 + d] = c->u_init[d];
    s->p[i] = c->p_init;
  }
  for (int p = 0; p < m->npatch; ++p) {
    for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f) {
      const int32_t b = N + (f - m->Fi), o = m->owner[f];
      if (m->u_bc[p] == ORC_BC_FIXED_VALUE) {
        for (int d = 0; d < 3; ++d) s->U[b * 3 + d] = m->u_val[p * 3 + d];
      } else if (m->u_bc[p] == ORC_BC_NO_SLIP) {
        /* the noSlip value '[0;0;0]' (cfdDefineEquation.m:290-295), set at
         * every boundary update (cfdUpdateBoundaryConditions.m:213-278) */
        for (int d = 0; d < 3; ++d) s->U[b * 3 + d] = 0.0;
      } else if (u_tangential(m, p)) {
        double n[3];
        const double nm = mag3(&m->Sf[f * 3]);
        for (int d = 0; d < 3; ++d) n[d] = m->Sf[f * 3 + d] / nm;
        const double un = dot3(&s->U[o * 3], n);
        for (int d = 0; d < 3; ++d) s->U[b * 3 + d] = s->U[o * 3 + d] - un * n[d];
      } else {
        for (int d = 0; d < 3; ++d) s->U[b * 3 + d] = s->U[o * 3 + d];
      }
      s->p[b] = m->p_bc[p] == ORC_BC_FIXED_VALUE ? m->p_val[p] : s->p[o];
    }
  }
  /* mdot_f = rho_f .* (Sf . U_f) with Average interpolation of U and rho
   * (cfdInitializeMdotField.m:72-80; the transient form subtracts
   * rho_f DeltaVol / dt with DeltaVol = 0 on every face,
   * cfdProcessOpenFoamMesh.m:17, so the bits are the same) */
  const double rho = c->rho == 0.0 ? 1.0 : c->rho;
  for (int32_t f = 0; f < F; ++f) {
    double uf[3], rf;
    if (f < m->Fi) {
      const int32_t o = m->owner[f], nb = m->neighbour[f];
      const double g = m->gf[f];
      for (int d = 0; d < 3; ++d) uf[d] = g * s->U[nb * 3 + d] + (1 - g) * s->U[o * 3 + d];
      rf = g * rho + (1 - g) * rho;
    } else {
      const int32_t b = N + (f - m->Fi);
      for (int d = 0; d < 3; ++d) uf[d] = s->U[b * 3 + d];
      rf = rho;
    }
    s->mdot[f] = rf * dot3(&m->Sf[f * 3], uf);
  }
  memcpy(s->U_old, s->U, (size_t)NB * 3 * sizeof(double));
  memcpy(s->mdot_old, s->mdot, (size_t)F * sizeof(double));
}

void orc_simple_free(orc_simple_state *s) {
  orc_hier_free(&s->H);
  orc_level *L = &s->H.lev[0];
  free(L->ac); free(L->anb); free(L->bc); free(L->dphi);
  free(s->U); free(s->gradU); free(s->p); free(s->gradp); free(s->PP); free(s->gradPP);
  free(s->DU); free(s->DUT); free(s->mdot); free(s->gamma_f); free(s->fluxC1); free(s->fluxC2);
  free(s->U_old); free(s->mdot_old);
  free(s->ac_old);
  memset(s, 0, sizeof(*s));
}

/* cfdComputeGradientGauss0.m:28-80.  The boundary contribution reads
 * phi_b(k) with linear indexing into the [Fb x nc] boundary block (:59-64),
 * i.e. the FIRST component's boundary value for every component; kept. */
static void gauss0(const orc_mesh *m, const double *phi, int nc, double *grad) {
  const int32_t N = m->N, Fb = FB(m);
  memset(grad, 0, (size_t)(N + Fb) * nc * 3 * sizeof(double));
  for (int c = 0; c < nc; ++c) {
    for (int32_t f = 0; f < m->Fi; ++f) {
      const int32_t o = m->owner[f], nb = m->neighbour[f];
      const double g = m->gf[f];
      const double phif = g * phi[nb * nc + c] + (1 - g) * phi[o * nc + c];
      for (int d = 0; d < 3; ++d) {
        grad[(o * nc + c) * 3 + d] = grad[(o * nc + c) * 3 + d] + phif * m->Sf[f * 3 + d];
        grad[(nb * nc + c) * 3 + d] = grad[(nb * nc + c) * 3 + d] - phif * m->Sf[f * 3 + d];
      }
    }
  }
  for (int c = 0; c < nc; ++c) {
    for (int32_t k = 0; k < Fb; ++k) {
      const int32_t f = m->Fi + k, o = m->owner[f];
      const double phib = phi[(N + k) * nc + 0];
      for (int d = 0; d < 3; ++d)
        grad[(o * nc + c) * 3 + d] = grad[(o * nc + c) * 3 + d] + phib * m->Sf[f * 3 + d];
    }
  }
  for (int c = 0; c < nc; ++c)
    for (int32_t i = 0; i < N; ++i)
      for (int d = 0; d < 3; ++d) grad[(i * nc + c) * 3 + d] = grad[(i * nc + c) * 3 + d] / m->vol[i];
  for (int32_t k = 0; k < Fb; ++k) {
    const int32_t o = m->owner[m->Fi + k];
    for (int q = 0; q < nc * 3; ++q) grad[(N + k) * nc * 3 + q] = grad[o * nc * 3 + q];
  }
}

/* cfdUpdateGradient.m:18-70: Gauss gradient, then the boundary rows of wall,
 * 'patch', inlet and outlet patches replaced by
 * g - (g.e)e + ((phi_b - phi_O)/|d|) e, d = C_f - C_O (walls :81-134, inlets
 * :142-195, outlets :202-255).  The inlet and outlet rows index phi linearly
 * (phi(iBElement) - phi(iOwner), :187 and :247), i.e. the first component's
 * difference for every component of a vector; kept. */
void orc_gradient(const orc_mesh *m, const double *phi, int nc, double *grad) {
  const int32_t N = m->N;
  gauss0(m, phi, nc, grad);
  for (int p = 0; p < m->npatch; ++p) {
    const int t = m->ptype[p];
    if (t != ORC_PATCH_WALL && t != ORC_PATCH_PATCH && t != ORC_PATCH_INLET &&
        t != ORC_PATCH_OUTLET)
      continue;
    const int lin = t == ORC_PATCH_INLET || t == ORC_PATCH_OUTLET;
    for (int c = 0; c < nc; ++c) {
      const int cd = lin ? 0 : c;
      for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f) {
        const int32_t b = N + (f - m->Fi), o = m->owner[f];
        double d[3], e[3];
        for (int q = 0; q < 3; ++q) d[q] = m->fc[f * 3 + q] - m->cc[o * 3 + q];
        const double dmag = mag3(d);
        for (int q = 0; q < 3; ++q) e[q] = d[q] / dmag;
        const double *g = &grad[(o * nc + c) * 3];
        const double ge = dot3(g, e);
        const double t = (phi[b * nc + cd] - phi[o * nc + cd]) / dmag;
        for (int q = 0; q < 3; ++q) grad[(b * nc + c) * 3 + q] = (g[q] - ge * e[q]) + t * e[q];
      }
    }
  }
}

/* cfdUpdateEquations.m:14-24 for U then p: boundary conditions (fixedValue
 * walls and inlets, cfdUpdateBoundaryConditions.m:103-139, :291-328; noSlip
 * walls the zero value :213-278; zeroGradient, slip, empty and symmetry are
 * no-ops; fixed-pressure outlets p_b = value, :335-373), gradients,
 * gamma_Ueq = Average(nu) (cfdUpdateTermCoefficientField.m:44-70), phiScale
 * (cfdUpdateScale.m:17-23) */
void orc_update_fields(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s) {
  const int32_t N = m->N, Fb = FB(m);
  for (int p = 0; p < m->npatch; ++p) {
    const int t = m->ptype[p];
    const int uval = (t == ORC_PATCH_WALL || t == ORC_PATCH_INLET) && m->u_bc[p] == ORC_BC_FIXED_VALUE;
    const int noslip = t == ORC_PATCH_WALL && m->u_bc[p] == ORC_BC_NO_SLIP;
    for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f) {
      const int32_t b = N + (f - m->Fi);
      if (uval)
        for (int d = 0; d < 3; ++d) s->U[b * 3 + d] = m->u_val[p * 3 + d] * 1.0;
      if (noslip)
        for (int d = 0; d < 3; ++d) s->U[b * 3 + d] = 0.0 * 1.0;
    }
  }
  orc_gradient(m, s->U, 3, s->gradU);
  for (int32_t f = 0; f < m->F; ++f) {
    if (f < m->Fi) {
      const double g = m->gf[f];
      s->gamma_f[f] = g * c->nu + (1 - g) * c->nu;
    } else {
      s->gamma_f[f] = c->nu;
    }
  }
  double mx = -INFINITY, mn = INFINITY;
  for (int32_t q = 0; q < (N + Fb) * 3; ++q) {
    if (s->U[q] > mx) mx = s->U[q];
    if (s->U[q] < mn) mn = s->U[q];
  }
  if (s->scale_zero) {
    /* the zero-filled rows of the grown U (cfdUpdateScale.m:16-17 takes the
     * max and min over every row) */
    if (0.0 > mx) mx = 0.0;
    if (0.0 < mn) mn = 0.0;
  }
  double sc = (mx - mn) > fabs(mx) ? (mx - mn) : fabs(mx);
  if (sc < 1e-6) sc = 1;
  s->scaleU = sc;
  for (int p = 0; p < m->npatch; ++p) {
    if (!p_fixed_outlet(m, p)) continue;
    for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f)
      s->p[N + (f - m->Fi)] = m->p_val[p] * 1.0;
  }
  orc_gradient(m, s->p, 1, s->gradp);
}

static void zero_system(orc_simple_state *s) {
  orc_level *L = &s->H.lev[0];
  memset(L->ac, 0, (size_t)L->n * sizeof(double));
  memset(L->anb, 0, (size_t)(L->nnz ? L->nnz : 1) * sizeof(double));
  memset(L->bc, 0, (size_t)L->n * sizeof(double));
  memset(L->dphi, 0, (size_t)L->n * sizeof(double));
  memset(s->ac_old, 0, (size_t)L->n * sizeof(double));
}

/* cfdAssembleIntoGlobalMatrixFaceFluxes.m:25-55 */
static void scatter_faces(const orc_mesh *m, orc_simple_state *s, const double *C1,
                          const double *C2, const double *T) {
  orc_level *L = &s->H.lev[0];
  for (int32_t f = 0; f < m->Fi; ++f) {
    const int32_t o = m->owner[f], nb = m->neighbour[f];
    L->ac[o] = L->ac[o] + C1[f];
    L->anb[m->own_nnz[f]] = L->anb[m->own_nnz[f]] + C2[f];
    L->bc[o] = L->bc[o] - T[f];
    L->ac[nb] = L->ac[nb] - C2[f];
    L->anb[m->nei_nnz[f]] = L->anb[m->nei_nnz[f]] - C1[f];
    L->bc[nb] = L->bc[nb] + T[f];
  }
  for (int32_t f = m->Fi; f < m->F; ++f) {
    const int32_t o = m->owner[f];
    L->ac[o] = L->ac[o] + C1[f];
    L->bc[o] = L->bc[o] - T[f];
  }
}

/* Average:Corrected face gradient of component c of a field with nc
 * components (cfdInterpolateGradientsFromElementsToInteriorFaces.m:46-70);
 * the average weights are swapped relative to value interpolation (:48-50) */
static void grad_face_corrected(const orc_mesh *m, const double *phi, const double *grad, int nc,
                                int c, int32_t f, double *gf_out) {
  const int32_t o = m->owner[f], nb = m->neighbour[f];
  const double g = m->gf[f];
  double gfc[3], dcf[3], e[3];
  for (int d = 0; d < 3; ++d)
    gfc[d] = (1 - g) * grad[(nb * nc + c) * 3 + d] + g * grad[(o * nc + c) * 3 + d];
  for (int d = 0; d < 3; ++d) dcf[d] = m->cc[nb * 3 + d] - m->cc[o * 3 + d];
  const double dmag = mag3(dcf);
  for (int d = 0; d < 3; ++d) e[d] = dcf[d] / dmag;
  const double lgm = (phi[nb * nc + c] - phi[o * nc + c]) / dmag;
  const double lam = dot3(gfc, e);
  for (int d = 0; d < 3; ++d) gf_out[d] = (gfc[d] - lam * e[d]) + lgm * e[d];
}

void orc_face_gradients(const orc_mesh *m, const double *phi, const double *grad, int nc, int c,
                        double *out) {
  for (int32_t f = 0; f < m->Fi; ++f) grad_face_corrected(m, phi, grad, nc, c, f, out + 3 * f);
}

/* momentum component: convection (upwind) + divergence term, diffusion,
 * -grad(p), false transient, URF (cfdAssembleEquationTerms.m:22-55) */
void orc_assemble_u(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s, int comp) {
  const int32_t N = m->N, F = m->F, Fi = m->Fi;
  orc_level *L = &s->H.lev[0];
  zero_system(s);
  double *C1 = dz(F), *C2 = dz(F), *T = dz(F);
  const double *U = s->U;
  const double rho = c->rho == 0.0 ? 1.0 : c->rho;
  /* ---- ddt(rho*U), Euler, sign +1, the first term of the equation
   * (cfdAssembleTransientTermEuler.m:13-44): FLUXCE = V rho / dt, FLUXCEOLD =
   * -V rho_old / dt and FLUXTE = FLUXCE .* phi' + FLUXCEOLD .* phi_old' -- a
   * column times a row, i.e. an N x N matrix whose linear index iElement in
   * cfdAssembleIntoGlobalMatrixElementFluxes.m:23-27 reads column 1: every
   * cell's source takes cell 1's phi and phi_old.  Kept (without the N x N
   * matrix). */
  if (c->transient) {
    const double sg = 1.0;
    const double phi1 = U[0 * 3 + comp], old1 = s->U_old[0 * 3 + comp];
    for (int32_t i = 0; i < N; ++i) {
      const double CE = ((sg * m->vol[i]) * rho) / c->dt;
      const double CEOLD = (((-sg) * m->vol[i]) * rho) / c->dt;
      const double TE = CE * phi1 + CEOLD * old1;
      L->ac[i] = L->ac[i] + CE;
      s->ac_old[i] = s->ac_old[i] + CEOLD;
      L->bc[i] = L->bc[i] - TE;
    }
  }
  /* ---- Convection, sign +1 (cfdAssembleConvectionTerm.m:103-137; fixedValue
   * inlets and zeroGradient outlets the same upwind fluxes with U_b in place
   * of U_N, :144-232; walls, empty and symmetry patches give zero fluxes
   * :238-302) */
  {
    const double sg = 1.0;
    for (int32_t f = 0; f < Fi; ++f) {
      const double psi = s->mdot[f];
      C1[f] = sg * fmax(psi, 0.0);
      C2[f] = (-sg) * fmax(-psi, 0.0);
      const double V = 0.0;
      T[f] = (C1[f] * U[m->owner[f] * 3 + comp] + C2[f] * U[m->neighbour[f] * 3 + comp]) + V;
    }
    for (int32_t f = Fi; f < F; ++f) C1[f] = C2[f] = T[f] = 0.0;
    for (int p = 0; p < m->npatch; ++p) {
      if (m->ptype[p] != ORC_PATCH_INLET && m->ptype[p] != ORC_PATCH_OUTLET) continue;
      for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f) {
        const int32_t b = N + (f - Fi);
        const double psi = s->mdot[f];
        C1[f] = sg * fmax(psi, 0.0);
        C2[f] = (-sg) * fmax(-psi, 0.0);
        T[f] = (C1[f] * U[m->owner[f] * 3 + comp] + C2[f] * U[b * 3 + comp]) + 0.0;
      }
    }
    scatter_faces(m, s, C1, C2, T);
    /* divergence term (cfdAssembleConvectionDivergenceTerm.m:38-47 with
     * cfdComputeEffectiveDivergenceWithFluidTag.m:21-46) */
    double *eff = dz(N);
    for (int32_t f = 0; f < Fi; ++f) {
      eff[m->owner[f]] = eff[m->owner[f]] + s->mdot[f];
      eff[m->neighbour[f]] = eff[m->neighbour[f]] - s->mdot[f];
    }
    for (int32_t f = Fi; f < F; ++f) eff[m->owner[f]] = eff[m->owner[f]] + s->mdot[f];
    for (int32_t i = 0; i < N; ++i) {
      const double e = eff[i], phi = U[i * 3 + comp];
      const double CE = fmax(e, 0.0) - e;
      const double VE = (-fmax(e, 0.0)) * phi;
      const double TE = CE * phi + VE;
      L->ac[i] = L->ac[i] + CE;
      L->bc[i] = L->bc[i] - TE;
    }
    free(eff);
  }
  /* ---- Diffusion laplacian(nu*U), sign -1 (cfdAssembleDiffusionTerm.m:94-130,
   * wall fixedValue :224-263 and inlet fixedValue :135-173 (the same fluxes),
   * zeroGradient walls :330-347 and outlets :201-218 (zero), empty :631-690) */
  {
    const double sg = -1.0;
    for (int32_t f = 0; f < Fi; ++f) {
      double gfc[3];
      grad_face_corrected(m, U, s->gradU, 3, comp, f, gfc);
      const double gam = s->gamma_f[f];
      C1[f] = ((-sg) * gam) * m->gDiff[f];
      C2[f] = (sg * gam) * m->gDiff[f];
      const double V = (sg * gam) * dot3(gfc, &m->T[f * 3]);
      T[f] = (C1[f] * U[m->owner[f] * 3 + comp] + C2[f] * U[m->neighbour[f] * 3 + comp]) + V;
    }
    for (int p = 0; p < m->npatch; ++p) {
      for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f) {
        const int32_t b = N + (f - Fi), o = m->owner[f];
        const double gam = s->gamma_f[f];
        if ((m->ptype[p] == ORC_PATCH_WALL || m->ptype[p] == ORC_PATCH_INLET) &&
            m->u_bc[p] == ORC_BC_FIXED_VALUE) {
          C1[f] = ((-sg) * gam) * m->gDiff[f];
          C2[f] = (sg * gam) * m->gDiff[f];
          const double V = (sg * gam) * dot3(&s->gradU[(b * 3 + comp) * 3], &m->T[f * 3]);
          T[f] = (C1[f] * U[o * 3 + comp] + C2[f] * U[b * 3 + comp]) + V;
        } else if (m->ptype[p] == ORC_PATCH_WALL && m->u_bc[p] == ORC_BC_NO_SLIP) {
          /* cfdAssembleDiffusionTerm.m:479-562 with TM = gamma_b / walldist
           * (cfdGetUCoef.m:41-44, laminar).  dot(nx',nx') on the patch's
           * 1 x F rows is one patch-wide sum of n_c^2 (hazard H7) */
          const double A = m->area[f], d = walldist(m, f), TM = gam / d;
          double n[3];
          for (int q = 0; q < 3; ++q) n[q] = m->Sf[f * 3 + q] / A;
          const double Sc = noslip_sum(m, p, comp);
          const int j = comp == 0 ? 1 : 0, k = comp == 2 ? 1 : 2;
          const double g = ((((-sg) * gam) * A) / d) * TM;
          C1[f] = g * (1 - Sc);
          C2[f] = 0.0;
          const double V = g * ((U[b * 3 + comp] * (1 - Sc) +
                                 ((U[o * 3 + j] - U[b * 3 + j]) * n[j]) * n[comp]) -
                                ((U[o * 3 + k] - U[b * 3 + k]) * n[k]) * n[comp]);
          T[f] = (C1[f] * U[o * 3 + comp] + C2[f] * U[b * 3 + comp]) + V;
        } else if (plane(m, p)) {
          const double A = m->area[f];
          double n[3], l[3] = {1.0, 1.0, 1.0};
          l[comp] = 0.0;
          for (int d = 0; d < 3; ++d) n[d] = m->Sf[f * 3 + d] / A;
          const double dd = dot3(&m->CN[f * 3], n);
          const double nc = n[comp];
          C1[f] = (((((-sg) * 2) * gam) * A) * (nc * nc)) / dd;
          C2[f] = 0.0;
          double a[3];
          for (int d = 0; d < 3; ++d) a[d] = (U[o * 3 + d] * l[d]) * (n[d] * l[d]);
          const double dl = (a[0] + a[1]) + a[2];
          const double V = ((((sg * 2) * gam) * A) * dl) * nc / dd;
          T[f] = (C1[f] * U[o * 3 + comp] + C2[f] * U[b * 3 + comp]) + V;
        } else {
          C1[f] = C2[f] = T[f] = 0.0;
        }
      }
    }
    scatter_faces(m, s, C1, C2, T);
  }
  /* ---- + div(nu*transp(grad(U))), explicit, sign -1
   * (cfdAssembleExplicitStressTerm.m:14-52): on every face
   * T = (sign gamma_f) ((g_x Sf_x + g_y Sf_y) + g_z Sf_z), g_d the Average
   * interpolation of phiGradient(:,d,comp) = d(U_comp)/dx_d (boundary faces:
   * the boundary element's row), i.e. grad(U_comp).Sf and not the transpose's
   * sum_j d(U_j)/dx_comp Sf_j, as the reference computes it */
  if (c->stress) {
    const double sg = -1.0;
    for (int32_t f = 0; f < F; ++f) {
      double g[3];
      if (f < Fi) {
        const int32_t o = m->owner[f], nb = m->neighbour[f];
        const double w = m->gf[f];
        for (int d = 0; d < 3; ++d)
          g[d] = w * s->gradU[(nb * 3 + comp) * 3 + d] + (1 - w) * s->gradU[(o * 3 + comp) * 3 + d];
      } else {
        const int32_t b = N + (f - Fi);
        for (int d = 0; d < 3; ++d) g[d] = s->gradU[(b * 3 + comp) * 3 + d];
      }
      C1[f] = C2[f] = 0.0;
      T[f] = (sg * s->gamma_f[f]) * dot3(g, &m->Sf[f * 3]);
    }
    scatter_faces(m, s, C1, C2, T);
  }
  /* ---- -grad(p) on the right-hand side, sign +1
   * (cfdAssemblePressureGradientTerm.m:31-40), element fluxes
   * (cfdAssembleIntoGlobalMatrixElementFluxes.m:23-27) */
  {
    const double sg = 1.0;
    for (int32_t i = 0; i < N; ++i) {
      const double TE = (sg * m->vol[i]) * s->gradp[i * 3 + comp];
      L->ac[i] = L->ac[i] + 0.0;
      s->ac_old[i] = s->ac_old[i] + 0.0;
      L->bc[i] = L->bc[i] - TE;
    }
  }
  /* ---- false transient of steady runs, fdt = 1e9, rho = rho_old
   * (cfdAssembleFalseTransientTerm.m:28-49; FLUXTEOLD is never scattered;
   * cfdAssembleEquationTerms.m:44-47: not in transient runs) */
  if (!c->transient) {
    const double fdt = 1e9;
    for (int32_t i = 0; i < N; ++i) {
      const double CE = (m->vol[i] * rho) / fdt;
      const double CEOLD = ((-m->vol[i]) * rho) / fdt;
      const double TE = CE * U[i * 3 + comp];
      L->ac[i] = L->ac[i] + CE;
      s->ac_old[i] = s->ac_old[i] + CEOLD;
      L->bc[i] = L->bc[i] - TE;
    }
  }
  /* ---- URF (cfdApplyURF.m:15-17) */
  for (int32_t i = 0; i < N; ++i) L->ac[i] = L->ac[i] / c->urf_u;
  /* ---- normalised residual (cfdComputeNormalizedResidual.m:64-88); in a
   * transient run the diagonal less V rho / dt, floored at 1e-6 of it
   * (:55-76) */
  {
    double sq = 0.0, mx = 0.0;
    for (int32_t i = 0; i < N; ++i) {
      double a = L->ac[i];
      if (c->transient) {
        const double at = (m->vol[i] * rho) / c->dt;
        a = a - at;
        if (a < 1e-6 * at) a = at;
      }
      const double r = L->bc[i] / (a * s->scaleU);
      if (fabs(r) > mx) mx = fabs(r);
      sq = sq + r * r;
    }
    s->mx[comp] = mx;
    s->rms[comp] = sqrt(sq / N);
  }
  /* ---- DU = V/ac, DUT = ac_old/ac, boundary rows from owners
   * (cfdPostAssembleEquation.m:37-81) */
  for (int32_t i = 0; i < N; ++i) {
    s->DU[i * 3 + comp] = m->vol[i] / L->ac[i];
    s->DUT[i * 3 + comp] = s->ac_old[i] / L->ac[i];
  }
  for (int32_t k = 0; k < FB(m); ++k) {
    const int32_t o = m->owner[Fi + k];
    s->DU[(N + k) * 3 + comp] = s->DU[o * 3 + comp];
    s->DUT[(N + k) * 3 + comp] = s->DUT[o * 3 + comp];
  }
  free(C1); free(C2); free(T);
}

/* cfdSolver.m:11-12 -> cfdSolveAlgebraicSystem(1, 'ILU', 10) */
void orc_solve_u(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s, int comp) {
  (void)m;
  orc_solve_algebraic_system(&s->H.lev[0], c->u_smoother, c->u_sweeps, 0.1, &s->r0[comp],
                             &s->r1[comp]);
}

/* cfdCorrectScalarEquation.m:16-24 (U(:,c) += dphi) and its boundary
 * branches, chosen by U's condition (:31-107): U_b,c = U_O,c - (U_O . n) n_c
 * on slip walls, empty and symmetry planes, U_b,c = U_O,c on zeroGradient
 * outlets, nothing on fixedValue / noSlip walls and fixedValue inlets */
void orc_correct_u(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s, int comp) {
  const int32_t N = m->N;
  const orc_level *L = &s->H.lev[0];
  (void)c;
  for (int32_t i = 0; i < N; ++i) s->U[i * 3 + comp] = s->U[i * 3 + comp] + L->dphi[i];
  for (int p = 0; p < m->npatch; ++p) {
    const int copy = m->ptype[p] == ORC_PATCH_OUTLET && m->u_bc[p] == ORC_BC_ZERO_GRADIENT;
    if (!u_tangential(m, p) && !copy) continue;
    for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f) {
      const int32_t b = N + (f - m->Fi), o = m->owner[f];
      if (copy) {
        s->U[b * 3 + comp] = s->U[o * 3 + comp];
        continue;
      }
      double n[3];
      const double nm = mag3(&m->Sf[f * 3]);
      for (int d = 0; d < 3; ++d) n[d] = m->Sf[f * 3 + d] / nm;
      const double un = dot3(&s->U[o * 3], n);
      s->U[b * 3 + comp] = s->U[o * 3 + comp] - un * n[comp];
    }
  }
}

/* cfdAssembleMdotTerm.m:38-250: Rhie-Chow face fluxes (stored as the new
 * mdot_f) and the pressure-correction matrix; boundary fluxes are zero for
 * walls and planes, inlets and fixed-pressure outlets carry theirs
 * (:281-348); the fix-pressure step is a no-op
 * (:257-278); residual from the effective divergence
 * (cfdComputeNormalizedResidual.m:27-30, cfdComputeEffectiveDivergence.m:21-54) */
void orc_assemble_p(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s) {
  const int32_t N = m->N, F = m->F, Fi = m->Fi;
  orc_level *L = &s->H.lev[0];
  zero_system(s);
  double *T = dz(F);
  const double *U = s->U, *DU = s->DU;
  const double rho = c->rho == 0.0 ? 1.0 : c->rho;
  /* transient terms VI-VII: DUT = mean([DUT1 DUT2 DUT3], 2) per element
   * (cfdAssembleMdotTerm.m:196) */
  double *DUTm = NULL;
  if (c->transient) {
    DUTm = dz((size_t)(N + FB(m)));
    for (int32_t i = 0; i < N + FB(m); ++i)
      DUTm[i] = ((s->DUT[i * 3] + s->DUT[i * 3 + 1]) + s->DUT[i * 3 + 2]) / 3;
  }
  for (int32_t f = 0; f < Fi; ++f) {
    const int32_t o = m->owner[f], nb = m->neighbour[f];
    const double g = m->gf[f];
    double du[3], ub[3], gpb[3], gpf[3];
    for (int d = 0; d < 3; ++d) du[d] = g * DU[nb * 3 + d] + (1 - g) * DU[o * 3 + d];
    const double mprev = s->mdot[f];
    for (int d = 0; d < 3; ++d) ub[d] = g * U[nb * 3 + d] + (1 - g) * U[o * 3 + d];
    /* rho_f: 'Upwind' by the previous mdot_f (cfdAssembleMdotTerm.m:118-121) */
    const double pos = mprev > 0 ? 1.0 : 0.0;
    const double rho_f = rho * pos + rho * (1 - pos);
    /* Average (swapped weights) and Average:Corrected pressure gradients */
    for (int d = 0; d < 3; ++d) gpb[d] = (1 - g) * s->gradp[nb * 3 + d] + g * s->gradp[o * 3 + d];
    grad_face_corrected(m, s->p, s->gradp, 1, 0, f, gpf);
    const double *Sf = &m->Sf[f * 3], *CF = &m->CN[f * 3], *eCF = &m->eCN[f * 3];
    const double UbS = dot3(ub, Sf);
    double V = 0.0 + rho_f * UbS;
    double DUS[3], eDUS[3], DUE[3], DUT[3];
    for (int d = 0; d < 3; ++d) DUS[d] = du[d] * Sf[d];
    const double mDUS = mag3(DUS);
    for (int d = 0; d < 3; ++d) eDUS[d] = DUS[d] / mDUS;
    const double ee = dot3(eCF, eDUS);
    for (int d = 0; d < 3; ++d) DUE[d] = mDUS * eCF[d] / ee;
    const double geoDiff = mag3(DUE) / mag3(CF);
    for (int d = 0; d < 3; ++d) DUT[d] = DUS[d] - DUE[d];
    const double FC = 0.0 + rho_f * geoDiff;
    const double FF = 0.0 - rho_f * geoDiff;
    V = V - rho_f * dot3(gpf, DUT);
    V = V + rho_f * dot3(gpb, DUS);
    if (c->transient) {
      /* terms VI-VII (:175-201): DUT_f (mdot_old - rho_old_f [v_old]_f . S_f),
       * DUT_f and [v_old]_f Average-interpolated, rho_old_f 'Upwind' by the
       * previous mdot_f */
      const double dut = g * DUTm[nb] + (1 - g) * DUTm[o];
      const double rho_old_f = rho * pos + rho * (1 - pos);
      double uo[3];
      for (int d = 0; d < 3; ++d) uo[d] = g * s->U_old[nb * 3 + d] + (1 - g) * s->U_old[o * 3 + d];
      V = V + dut * (s->mdot_old[f] - rho_old_f * dot3(uo, Sf));
    }
    V = V + (1 - c->urf_u) * (mprev - rho_f * UbS);
    const double md = (FC * s->p[o] + FF * s->p[nb]) + V;
    s->mdot[f] = md;
    s->fluxC1[f] = FC;
    s->fluxC2[f] = FF;
    T[f] = md;
  }
  for (int32_t f = Fi; f < F; ++f) s->fluxC1[f] = s->fluxC2[f] = T[f] = 0.0;
  for (int p = 0; p < m->npatch; ++p) {
    if (!p_flux(m, p)) continue;
    for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f) {
      const int32_t o = m->owner[f], b = N + (f - Fi);
      const double *Sb = &m->Sf[f * 3];
      double C1 = 0.0, C2 = 0.0, V = 0.0;
      if (m->ptype[p] == ORC_PATCH_INLET) {
        /* cfdAssembleMdotTermInletInletBC (:415-486): term I only */
        V = V + rho * dot3(Sb, &U[b * 3]);
      } else {
        /* cfdAssembleMdotTermOutletFixedValueBC (:799-906): terms I-III with
         * the boundary rows of DU and grad(p), FLUXCb2 = 0 */
        const double *CN = &m->CN[f * 3], *eCN = &m->eCN[f * 3];
        double DUS[3], eDUS[3], DUE[3], DUT[3];
        for (int d = 0; d < 3; ++d) DUS[d] = DU[b * 3 + d] * Sb[d];
        const double mDUS = mag3(DUS);
        for (int d = 0; d < 3; ++d) eDUS[d] = DUS[d] / mDUS;
        const double ee = dot3(eCN, eDUS);
        for (int d = 0; d < 3; ++d) DUE[d] = (mDUS / ee) * eCN[d];
        const double geoDiff = mag3(DUE) / mag3(CN);
        for (int d = 0; d < 3; ++d) DUT[d] = DUS[d] - DUE[d];
        V = V + rho * dot3(&U[b * 3], Sb);
        C1 = C1 + rho * geoDiff;
        C2 = C2 - 0.0;
        V = V - rho * dot3(&s->gradp[b * 3], DUT);
        V = V + rho * dot3(&s->gradp[b * 3], DUS);
      }
      const double md = (C1 * s->p[o] + C2 * s->p[b]) + V;
      s->mdot[f] = md;
      s->fluxC1[f] = C1;
      s->fluxC2[f] = C2;
      T[f] = md;
    }
  }
  scatter_faces(m, s, s->fluxC1, s->fluxC2, T);
  for (int32_t i = 0; i < N; ++i) {
    L->ac[i] = L->ac[i] + 0.0;
    s->ac_old[i] = s->ac_old[i] + 0.0;
    L->bc[i] = L->bc[i] - 0.0;
  }
  double *eff = dz(N);
  for (int32_t f = 0; f < Fi; ++f) {
    eff[m->owner[f]] = eff[m->owner[f]] + s->mdot[f];
    eff[m->neighbour[f]] = eff[m->neighbour[f]] - s->mdot[f];
  }
  for (int32_t f = Fi; f < F; ++f) eff[m->owner[f]] = eff[m->owner[f]] + s->mdot[f];
  double mx = -INFINITY, sq = 0.0;
  for (int32_t i = 0; i < N; ++i) {
    if (eff[i] > mx) mx = eff[i];
    sq = sq + eff[i] * eff[i];
  }
  s->mx[3] = mx;
  s->rms[3] = sqrt(sq / N);
  free(eff);
  free(T);
  free(DUTm);
}

/* the FP64 FCG pressure solve of p_solver 1 on s->H (agglomerated): level 1's
 * dphi receives the solution; r1 = the relative residual */
static void fcg_solve_p(const orc_simple_cfg *c, orc_simple_state *s) {
  orc_hier *H = &s->H;
  orc_level *L = &H->lev[0];
  int nlev = H->nlevels;
  for (int l = 0; l < H->nlevels; ++l)
    if (H->lev[l].n <= 2000) {
      nlev = l + 1;
      break;
    }
  /* constant nullspace: every row sum vanishes relative to the row's size */
  int singular = 1;
  for (int32_t i = 0; i < L->n && singular; ++i) {
    double rs = L->ac[i], mag = fabs(L->ac[i]);
    for (int32_t k = L->ptr[i]; k < L->ptr[i + 1]; ++k) {
      rs += L->anb[k];
      mag += fabs(L->anb[k]);
    }
    if (fabs(rs) > 1e-9 * mag) singular = 0;
  }
  int32_t *colours[16];
  int ncol[16];
  for (int l = 0; l < nlev; ++l) {
    const orc_level *V = &H->lev[l];
    colours[l] = (int32_t *)malloc(sizeof(int32_t) * (V->n > 0 ? V->n : 1));
    ncol[l] = l == 0 ? orc_colour(V->n, V->ptr, V->col, colours[l])
                     : orc_colour_blocks(V->n, V->ptr, V->col, 1024, colours[l]);
  }
  const orc_vcfg v = {nlev, 1, 1, singular, 1024, 1, 0, 2.5, 2, 2};
  double *x = (double *)malloc(sizeof(double) * (L->n > 0 ? L->n : 1));
  double rr = 0.0;
  s->ncycles = orc_fpcg(H, colours, ncol, &v, L->bc, c->pcg_tol, 500, x, &rr);
  memcpy(L->dphi, x, sizeof(double) * L->n);
  s->r0[3] = 1.0;
  s->r1[3] = rr;
  free(x);
  for (int l = 0; l < nlev; ++l) free(colours[l]);
}

/* cfdSolveEquation('p') -> cfdApplyAMG (rebuilds the hierarchy) */
void orc_solve_p(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s) {
  (void)m;
  orc_hier_free(&s->H);
  orc_agglomerate(&s->H, c->p_max_coarse);
  if (c->p_solver == 1) {
    fcg_solve_p(c, s);
    return;
  }
  orc_apply_amg(&s->H, c->p_cycle, c->p_max_cycles, c->p_termination, c->p_pre, c->p_post,
                c->p_smoother, &s->r0[3], &s->r1[3], &s->ncycles);
}

/* cfdCorrectEquation.m:26-47 for p, each boundary branch chosen by the patch
 * type and p's condition as the reference dispatches them:
 *  - PP (cfdCorrectPPField.m:26-52): cells dphi - dphi(ref); PP_b an owner
 *    copy (walls, zeroGradient inlets, planes :131-186, :278-335), 0 on
 *    fixed-pressure outlets (:249-272); then its gradient;
 *  - U (cfdCorrectVelocityField.m:26-121): cells -= DU o grad(PP); U_b
 *    projected on empty and symmetry planes (:274-343); the owners' U
 *    written to rows iBFaces of U on fixed-pressure outlets (:247-272, the
 *    face index used as the row: kept); walls with p zeroGradient and
 *    zeroGradient inlets are no-ops (:173-182);
 *  - p (cfdCorrectPressureEquation.m:22-54): cells += urf (PP - PP(ref));
 *    p_b an owner copy except on fixed-pressure outlets (:131-290);
 *  - mdot (cfdCorrectMdotField.m:24-50): interior += C1 PP_O + C2 PP_N;
 *    planes 0 mdot_b (:306-346); fixed-pressure outlets rho_b (U_O . S_b)
 *    (:264-304); walls with p zeroGradient and incompressible inlets are
 *    no-ops (:156-208). */
void orc_correct_p(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s) {
  const int32_t N = m->N, Fi = m->Fi;
  const orc_level *L = &s->H.lev[0];
  double ppc = 0.0;
  if (c->p_ref >= 0) ppc = L->dphi[c->p_ref];
  for (int32_t i = 0; i < N; ++i) s->PP[i] = L->dphi[i] - ppc;
  for (int p = 0; p < m->npatch; ++p)
    for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f)
      s->PP[N + (f - Fi)] = p_fixed_outlet(m, p) ? 0.0 : s->PP[m->owner[f]];
  orc_gradient(m, s->PP, 1, s->gradPP);
  for (int32_t i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d)
      s->U[i * 3 + d] = s->U[i * 3 + d] - s->DU[i * 3 + d] * s->gradPP[i * 3 + d];
  for (int p = 0; p < m->npatch; ++p) {
    if (plane(m, p)) {
      for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f) {
        const int32_t b = N + (f - Fi), o = m->owner[f];
        double n[3];
        const double nm = mag3(&m->Sf[f * 3]);
        for (int d = 0; d < 3; ++d) n[d] = m->Sf[f * 3 + d] / nm;
        const double un = dot3(&s->U[o * 3], n);
        for (int d = 0; d < 3; ++d) s->U[b * 3 + d] = s->U[o * 3 + d] - un * n[d];
      }
    } else if (p_fixed_outlet(m, p)) {
      /* U(iBFaces,:) = U(iOwners,:): the 1-based face index f+1 is the row,
       * all owners read before any row is written; rows past the field (the
       * reference would grow U) are not written */
      const int32_t f0 = m->pstart[p], nf = m->pnf[p];
      double *tmp = dz((size_t)nf * 3);
      for (int32_t k = 0; k < nf; ++k)
        for (int d = 0; d < 3; ++d) tmp[k * 3 + d] = s->U[m->owner[f0 + k] * 3 + d];
      for (int32_t k = 0; k < nf; ++k)
        if (f0 + k < N + FB(m))
          for (int d = 0; d < 3; ++d) s->U[(f0 + k) * 3 + d] = tmp[k * 3 + d];
      free(tmp);
      if (f0 + nf > N + FB(m)) {
        /* rows N+Fb .. f0+nf-1 now exist; any of them no outlet face writes
         * stays zero */
        int32_t hi = f0 + nf - 1, covered = 0;
        for (int q = 0; q < m->npatch; ++q) {
          if (!p_fixed_outlet(m, q)) continue;
          const int32_t e = m->pstart[q] + m->pnf[q] - 1;
          if (e > hi) hi = e;
        }
        for (int32_t r = N + FB(m); r <= hi; ++r) {
          int hit = 0;
          for (int q = 0; q < m->npatch && !hit; ++q)
            hit = p_fixed_outlet(m, q) && r >= m->pstart[q] && r < m->pstart[q] + m->pnf[q];
          covered += hit;
        }
        if (covered < hi - (N + FB(m)) + 1) s->scale_zero = 1;
      }
    }
  }
  double ppf = 0.0;
  if (c->p_ref >= 0) ppf = s->PP[c->p_ref];
  for (int32_t i = 0; i < N; ++i) s->p[i] = s->p[i] + c->urf_p * (s->PP[i] - ppf);
  for (int p = 0; p < m->npatch; ++p) {
    if (p_fixed_outlet(m, p)) continue;
    for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f)
      s->p[N + (f - Fi)] = s->p[m->owner[f]];
  }
  for (int32_t f = 0; f < Fi; ++f)
    s->mdot[f] = (s->mdot[f] + s->fluxC1[f] * s->PP[m->owner[f]]) +
                 s->fluxC2[f] * s->PP[m->neighbour[f]];
  for (int p = 0; p < m->npatch; ++p) {
    if (plane(m, p)) {
      for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f)
        s->mdot[f] = 0.0 * s->mdot[f];
    } else if (p_fixed_outlet(m, p)) {
      for (int32_t f = m->pstart[p]; f < m->pstart[p] + m->pnf[p]; ++f)
        s->mdot[f] = (c->rho == 0.0 ? 1.0 : c->rho) * dot3(&s->U[m->owner[f] * 3], &m->Sf[f * 3]);
    }
  }
}

void orc_transient_update(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s) {
  (void)c;
  memcpy(s->U_old, s->U, (size_t)(m->N + FB(m)) * 3 * sizeof(double));
  memcpy(s->mdot_old, s->mdot, (size_t)m->F * sizeof(double));
}

int orc_transient_step(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s,
                       int max_inner, double u_tol, double p_tol) {
  orc_transient_update(m, c, s);
  int it = 0;
  while (it < max_inner) {
    orc_simple_iteration(m, c, s);
    ++it;
    if (s->rms[0] <= u_tol && s->rms[1] <= u_tol && s->rms[2] <= u_tol && s->rms[3] <= p_tol)
      break;
  }
  return it;
}

/* cfdRunCase.m:92-142, one steady iteration (equations U then p) */
void orc_simple_iteration(const orc_mesh *m, const orc_simple_cfg *c, orc_simple_state *s) {
  orc_update_fields(m, c, s);
  for (int comp = 0; comp < 3; ++comp) {
    orc_assemble_u(m, c, s, comp);
    orc_solve_u(m, c, s, comp);
    orc_correct_u(m, c, s, comp);
  }
  for (int k = 0; k < c->n_correctors; ++k) {
    orc_assemble_p(m, c, s);
    orc_solve_p(m, c, s);
    orc_correct_p(m, c, s);
  }
  (void)patch_of;
}
