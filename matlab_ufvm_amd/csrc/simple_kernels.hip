// simple_kernels.hip -- gfx950 kernels of the steady SIMPLE outer iteration.
//
// Assembly is cell-centric: one thread per cell walks the cell's faces in
// ascending face order and accumulates ac, bc and its own anb row, so each
// coefficient receives the face contributions in exactly the order of the
// reference's face-loop scatter (uFVM/src/fvm/assemble/
// cfdAssembleIntoGlobalMatrixFaceFluxes.m:25-55: interior faces ascending,
// then boundary faces) with no atomics.  Face fluxes are evaluated once per
// face by a face kernel and gathered by both cells; cf_slot gives each
// cell-face entry its side and anb slot, so each anb value is written once.
// Built with -ffp-contract=off; every expression keeps the reference's
// association.
#include <hip/hip_runtime.h>

#include "kernels.h"
#include "reduce.h"
#include "simple_kernels.h"

namespace ufvm {
namespace sk {

using k::kBlock;
using k::nblocks;

__device__ inline double mag3(const double *v) {
  return sqrt((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
}
__device__ inline double dot3(const double *a, const double *b) {
  return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2];
}

// Two per-block partial values (op 0: sum, 1: max, 2: min) written to
// partial[2*block + c]; a one-block kernel then combines them in block order.
__device__ inline void block_partials2(double v0, double v1, int op0, int op1, double *partial) {
  __shared__ double lds[kBlock / k::kWave];
  auto red = [&](double v, int op) {
    if (op == 0) return k::block_sum(v, lds);
    if (op == 1) return k::block_max(v, lds);
    return -k::block_max(-v, lds);
  };
  v0 = red(v0, op0);
  v1 = red(v1, op1);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = v0;
    partial[2 * blockIdx.x + 1] = v1;
  }
}

// mode 0: (max, min) of U -> scale; mode 1: (max, sum of squares) of the residual
__global__ void __launch_bounds__(kBlock)
k_reduce2(const double *__restrict__ partial, int32_t nblk, int mode, double *scal) {
  __shared__ double lds[kBlock / k::kWave];
  const int op0 = 1, op1 = mode == 0 ? 2 : 0;
  double a0 = -INFINITY, a1 = op1 == 2 ? INFINITY : 0.0;
  for (int32_t q = threadIdx.x; q < nblk; q += kBlock) {
    a0 = fmax(a0, partial[2 * q]);
    a1 = op1 == 2 ? fmin(a1, partial[2 * q + 1]) : a1 + partial[2 * q + 1];
  }
  a0 = k::block_max(a0, lds);
  a1 = op1 == 2 ? -k::block_max(-a1, lds) : k::block_sum(a1, lds);
  if (threadIdx.x != 0) return;
  if (mode == 0) {
    scal[S_UMAX] = a0;
    scal[S_UMIN] = a1;
    double sc = fmax(a0 - a1, fabs(a0));
    if (sc < 1e-6) sc = 1.0;
    scal[S_SCALE] = sc;
  } else {
    scal[S_RMAX] = a0;
    scal[S_RSQ] = a1;
  }
}

// ------------------------------------------------------ boundary conditions

// empty and symmetry patches: every symmetry branch of the reference's steady
// path does its empty branch's arithmetic (diffusion
// cfdAssembleDiffusionTerm.m:567-626 vs :631-690, zero convection, no gradient
// update, U_b = U_O - (U_O . n) n in cfdCorrectScalarEquation.m:469-505 and
// cfdCorrectVelocityField.m:274-305, mdot_b = 0 mdot_b in
// cfdCorrectMdotField.m:306-322, owner copies of p and PP)
__device__ inline bool plane(const DevMesh &m, int p) {
  return m.ptype[p] == PATCH_EMPTY || m.ptype[p] == PATCH_SYMMETRY;
}

// the momentum corrections dispatch on U's condition
// (cfdCorrectScalarEquation.m:31-107): U_b,c = U_O,c - (U_O . n) n_c on the
// planes and on slip walls (:239-293); a slip wall's diffusion and mass fluxes
// are zero and its gradient rows are a wall's
__device__ inline bool u_tangential(const DevMesh &m, int p) {
  return plane(m, p) || (m.ptype[p] == PATCH_WALL && m.u_bc[p] == BC_SLIP);
}

// inlets and outlets carry a boundary mass flux in the pressure-correction
// assembly (cfdAssembleMdotTerm.m:415-486, :799-906); walls and planes do not
__device__ inline bool p_flux(const DevMesh &m, int p) {
  return m.ptype[p] == PATCH_INLET || m.ptype[p] == PATCH_OUTLET;
}

// an outlet with a fixed pressure (its p_bc is fixedValue)
__device__ inline bool p_fixed_outlet(const DevMesh &m, int p) {
  return m.ptype[p] == PATCH_OUTLET && m.p_bc[p] == BC_FIXED_VALUE;
}

// fixedValue walls and inlets: U_b = value (cfdUpdateBoundaryConditions.m:103-139,
// :291-328); noSlip walls: U_b = their zero value (:213-278)
__global__ void k_update_u_bc(DevMesh m, double *U) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m.F - m.Fi) return;
  const int p = m.bpatch[k];
  const int32_t b = m.N + k;
  const bool wall_or_inlet = m.ptype[p] == PATCH_WALL || m.ptype[p] == PATCH_INLET;
  if (wall_or_inlet && m.u_bc[p] == BC_FIXED_VALUE) {
    for (int d = 0; d < 3; ++d) U[b * 3 + d] = m.u_val[p * 3 + d] * 1.0;
  } else if (m.ptype[p] == PATCH_WALL && m.u_bc[p] == BC_NO_SLIP) {
    for (int d = 0; d < 3; ++d) U[b * 3 + d] = 0.0 * 1.0;
  }
}

// fixed-pressure outlets: p_b = value (cfdUpdateBoundaryConditions.m:335-373)
__global__ void k_update_p_bc(DevMesh m, double *p) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m.F - m.Fi) return;
  const int q = m.bpatch[k];
  if (p_fixed_outlet(m, q)) p[m.N + k] = m.p_val[q] * 1.0;
}

void update_p_bc(hipStream_t s, const DevMesh &m, double *p) {
  const int32_t Fb = m.F - m.Fi;
  if (Fb > 0) hipLaunchKernelGGL(k_update_p_bc, dim3(nblocks(Fb)), dim3(kBlock), 0, s, m, p);
}

void update_u_bc(hipStream_t s, const DevMesh &m, double *U) {
  const int32_t Fb = m.F - m.Fi;
  if (Fb > 0) hipLaunchKernelGGL(k_update_u_bc, dim3(nblocks(Fb)), dim3(kBlock), 0, s, m, U);
}

// ---------------------------------------------------------------- gradient

// Green-Gauss, cfdComputeGradientGauss0.m:28-80.  Per (cell, component,
// direction) the contributions arrive in ascending face order: interior
// faces, then boundary faces.  The boundary contribution uses the boundary
// value of component 1 for every component (linear indexing phi_b(k) into
// the [Fb x nc] block, :59-64).

// one face of a cell's Green-Gauss sum: face f, side slot sl, Sf, weight gf
// and the owner / neighbour values (boundary faces: po[0] = phi_b's first
// component)
template <int NC>
struct GEntry {
  int32_t f, sl;
  double S[3], w, po[NC], pn[NC];
};

// the face's contribution to g: phi_f S_f added (owner) or subtracted
// (neighbour); boundary faces phi_b S_b
template <int NC>
__device__ __forceinline__ void gauss_entry(const GEntry<NC> &x, int32_t Fi, double (&g)[NC * 3]) {
  if (x.f < Fi) {
    const bool own = x.sl >= 0;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const double phif = x.w * x.pn[c] + (1 - x.w) * x.po[c];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (own) g[c * 3 + d] = g[c * 3 + d] + phif * x.S[d];
        else g[c * 3 + d] = g[c * 3 + d] - phif * x.S[d];
      }
    }
  } else {
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int d = 0; d < 3; ++d) g[c * 3 + d] = g[c * 3 + d] + x.po[0] * x.S[d];
  }
}

// loads entry e of cell i (two dependent stages: face, then its data)
template <int NC>
__device__ __forceinline__ void gauss_face(const DevMesh &m, int32_t e, GEntry<NC> &x) {
  x.f = m.cf_face[e];
  x.sl = m.cf_slot[e];
}

template <int NC>
__device__ __forceinline__ void gauss_data(const DevMesh &m, const double *phi, int32_t i,
                                           GEntry<NC> &x) {
  const int32_t f = x.f;
#pragma unroll
  for (int d = 0; d < 3; ++d) x.S[d] = m.Sf[f * 3 + d];
  if (f < m.Fi) {
    x.w = m.gf[f];
    const int32_t o = x.sl >= 0 ? i : m.owner[f];
    const int32_t nb = x.sl >= 0 ? m.neighbour[f] : i;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      x.po[c] = phi[o * NC + c];
      x.pn[c] = phi[nb * NC + c];
    }
  } else {
    x.w = 0.0;
    x.po[0] = phi[(m.N + (f - m.Fi)) * NC + 0];
  }
}

// Green-Gauss gradient of the cells (cfdComputeGradientGauss0.m:28-80): the
// cell's faces in ascending order (interior faces first), then / volume.
// Scalar fields load their entries in groups of 8 before the sums.
template <int NC>
__global__ void __launch_bounds__(kBlock) k_gauss_cells(DevMesh m, const double *phi, double *grad) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m.N) return;
  double g[NC * 3];
#pragma unroll
  for (int q = 0; q < NC * 3; ++q) g[q] = 0.0;
  const int32_t e0 = m.cf_ptr[i], e1 = m.cf_ptr[i + 1];
  if constexpr (NC == 1) {
    for (int32_t e = e0; e < e1; e += 8) {
      GEntry<1> x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (e + u < e1) gauss_face<1>(m, e + u, x[u]);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (e + u < e1) gauss_data<1>(m, phi, i, x[u]);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (e + u < e1) gauss_entry<1>(x[u], m.Fi, g);
    }
  } else {
    for (int32_t e = e0; e < e1; ++e) {
      const int32_t f = m.cf_face[e];
      const double *S = &m.Sf[f * 3];
      if (f < m.Fi) {
        const int32_t o = m.owner[f], nb = m.neighbour[f];
        const double w = m.gf[f];
        const bool own = (o == i);
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const double phif = w * phi[nb * NC + c] + (1 - w) * phi[o * NC + c];
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            if (own) g[c * 3 + d] = g[c * 3 + d] + phif * S[d];
            else g[c * 3 + d] = g[c * 3 + d] - phif * S[d];
          }
        }
      } else {
        const double phib = phi[(m.N + (f - m.Fi)) * NC + 0];
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
          for (int d = 0; d < 3; ++d) g[c * 3 + d] = g[c * 3 + d] + phib * S[d];
      }
    }
  }
  const double v = m.vol[i];
#pragma unroll
  for (int q = 0; q < NC * 3; ++q) grad[i * NC * 3 + q] = g[q] / v;
}

// boundary rows: the owner's gradient, then on wall, 'patch', inlet and outlet
// patches g - (g.e)e + ((phi_b - phi_O)/|d|) e with d = C_f - C_O
// (cfdUpdateGradient.m:81-134 walls, :142-195 inlets, :202-255 outlets).  The
// inlet and outlet rows index phi linearly (phi(iBElement) - phi(iOwner)), so
// every component of a vector takes its first component's difference: kept.
template <int NC>
__global__ void k_gauss_bnd(DevMesh m, const double *phi, double *grad) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m.F - m.Fi) return;
  const int32_t f = m.Fi + k, o = m.owner[f], b = m.N + k;
  const int p = m.bpatch[k];
  const bool lin = (m.ptype[p] == PATCH_INLET || m.ptype[p] == PATCH_OUTLET);
  const bool wall = (m.ptype[p] == PATCH_WALL || m.ptype[p] == PATCH_PATCH || lin);
  double d[3], e[3], dmag = 1.0;
  if (wall) {
    for (int q = 0; q < 3; ++q) d[q] = m.fc[f * 3 + q] - m.cc[o * 3 + q];
    dmag = mag3(d);
    for (int q = 0; q < 3; ++q) e[q] = d[q] / dmag;
  }
  for (int c = 0; c < NC; ++c) {
    const double *g = &grad[(o * NC + c) * 3];
    if (!wall) {
      for (int q = 0; q < 3; ++q) grad[(b * NC + c) * 3 + q] = g[q];
    } else {
      const double ge = dot3(g, e);
      const int cd = lin ? 0 : c;
      const double t = (phi[b * NC + cd] - phi[o * NC + cd]) / dmag;
      for (int q = 0; q < 3; ++q) grad[(b * NC + c) * 3 + q] = (g[q] - ge * e[q]) + t * e[q];
    }
  }
}

void gradient(hipStream_t s, const DevMesh &m, const double *phi, int nc, double *grad) {
  const int32_t Fb = m.F - m.Fi;
  if (nc == 3) {
    hipLaunchKernelGGL(k_gauss_cells<3>, dim3(nblocks(m.N)), dim3(kBlock), 0, s, m, phi, grad);
    if (Fb > 0) hipLaunchKernelGGL(k_gauss_bnd<3>, dim3(nblocks(Fb)), dim3(kBlock), 0, s, m, phi, grad);
  } else {
    hipLaunchKernelGGL(k_gauss_cells<1>, dim3(nblocks(m.N)), dim3(kBlock), 0, s, m, phi, grad);
    if (Fb > 0) hipLaunchKernelGGL(k_gauss_bnd<1>, dim3(nblocks(Fb)), dim3(kBlock), 0, s, m, phi, grad);
  }
}

// gamma_Ueq = Average(nu) on faces (cfdUpdateTermCoefficientField.m:44-70)
__global__ void k_gamma(DevMesh m, double nu, double *gamma) {
  const int32_t f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= m.F) return;
  if (f < m.Fi) {
    const double w = m.gf[f];
    gamma[f] = w * nu + (1 - w) * nu;
  } else {
    gamma[f] = nu;
  }
}

void gamma_faces(hipStream_t s, const DevMesh &m, double nu, double *gamma) {
  hipLaunchKernelGGL(k_gamma, dim3(nblocks(m.F)), dim3(kBlock), 0, s, m, nu, gamma);
}

// phiScale = max(phiMax - phiMin, |phiMax|), 1 if < 1e-6 (cfdUpdateScale.m:17-23)
// (zero: one more value, 0, for the zero-filled rows of a grown U)
__global__ void __launch_bounds__(kBlock) k_scale_u(int64_t n, bool zero, const double *U,
                                                    double *partial) {
  const int64_t q = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  double mx = -INFINITY, mn = INFINITY;
  if (q < n) mx = mn = U[q];
  else if (q == n && zero) mx = mn = 0.0;
  block_partials2(mx, mn, 1, 2, partial);
}

void scale_u(hipStream_t s, const DevMesh &m, const double *U, bool zero, double *partial,
             unsigned *counter, double *scal) {
  (void)counter;
  const int64_t n = (int64_t)(m.N + m.F - m.Fi) * 3;
  const int64_t nb = nblocks(n + (zero ? 1 : 0));
  hipLaunchKernelGGL(k_scale_u, dim3(nb), dim3(kBlock), 0, s, n, zero, U, partial);
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, s, partial, nb, 0, scal);
}

// --------------------------------------------------------- momentum assembly

// Average:Corrected face gradient of component c
// (cfdInterpolateGradientsFromElementsToInteriorFaces.m:46-70; the average
// weights are swapped relative to value interpolation, :48-50)
template <int NC>
__device__ inline void grad_face(const DevMesh &m, const double *phi, const double *grad, int c,
                                 int32_t f, double *out) {
  const int32_t o = m.owner[f], nb = m.neighbour[f];
  const double w = m.gf[f];
  double gfc[3], dcf[3], e[3];
  for (int d = 0; d < 3; ++d)
    gfc[d] = (1 - w) * grad[(nb * NC + c) * 3 + d] + w * grad[(o * NC + c) * 3 + d];
  for (int d = 0; d < 3; ++d) dcf[d] = m.cc[nb * 3 + d] - m.cc[o * 3 + d];
  const double dmag = mag3(dcf);
  for (int d = 0; d < 3; ++d) e[d] = dcf[d] / dmag;
  const double lgm = (phi[nb * NC + c] - phi[o * NC + c]) / dmag;
  const double lam = dot3(gfc, e);
  for (int d = 0; d < 3; ++d) out[d] = (gfc[d] - lam * e[d]) + lgm * e[d];
}

// Face fluxes of momentum component c, six per face (uf[6 f + q]):
//   convection, sign +1 (cfdAssembleConvectionTerm.m:103-137): q 0..2 =
//     C1 = max(mdot,0), C2 = -max(-mdot,0), T = C1 U_O + C2 U_N (U_b on
//     inlets and outlets, :144-232); walls and planes contribute zero
//     (:238-302; left 0 here);
//   diffusion laplacian(nu*U), sign -1: q 3..5 = C1, C2, T; interior with the
//     corrected face gradient (cfdAssembleDiffusionTerm.m:94-130), fixedValue
//     walls (:224-263), zeroGradient (:330-347, zero), empty vector patches
//     (:631-690); boundary faces have C2 = 0.
__global__ void __launch_bounds__(kBlock)
k_u_faces(DevMesh m, int c, const double *__restrict__ U, const double *__restrict__ gradU,
          const double *__restrict__ mdot, const double *__restrict__ gamma,
          double *__restrict__ uf, double *__restrict__ stress) {
  const int32_t f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= m.F) return;
  const int32_t N = m.N, Fi = m.Fi;
  const double gam = gamma[f];
  if (stress) {
    // + div(nu*transp(grad(U))), explicit, sign -1 (cfdAssembleExplicitStressTerm.m:14-52):
    // (sign gamma_f) ((g_x Sf_x + g_y Sf_y) + g_z Sf_z), g_d the Average
    // interpolation of d(U_c)/dx_d (boundary faces: the boundary row) -- the
    // gradient of U_c itself, as the reference computes it
    double g[3];
    if (f < Fi) {
      const int32_t o = m.owner[f], nb = m.neighbour[f];
      const double w = m.gf[f];
      for (int d = 0; d < 3; ++d)
        g[d] = w * gradU[(nb * 3 + c) * 3 + d] + (1 - w) * gradU[(o * 3 + c) * 3 + d];
    } else {
      const int32_t b = N + (f - Fi);
      for (int d = 0; d < 3; ++d) g[d] = gradU[(b * 3 + c) * 3 + d];
    }
    stress[f] = ((-1.0) * gam) * dot3(g, &m.Sf[f * 3]);
  }
  double out[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (f < Fi) {
    const int32_t o = m.owner[f], nb = m.neighbour[f];
    const double psi = mdot[f];
    const double C1 = 1.0 * fmax(psi, 0.0);
    const double C2 = (-1.0) * fmax(-psi, 0.0);
    out[0] = C1;
    out[1] = C2;
    out[2] = (C1 * U[o * 3 + c] + C2 * U[nb * 3 + c]) + 0.0;
    double gfc[3];
    grad_face<3>(m, U, gradU, c, f, gfc);
    const double D1 = (1.0 * gam) * m.gDiff[f];
    const double D2 = ((-1.0) * gam) * m.gDiff[f];
    const double V = ((-1.0) * gam) * dot3(gfc, &m.T[f * 3]);
    out[3] = D1;
    out[4] = D2;
    out[5] = (D1 * U[o * 3 + c] + D2 * U[nb * 3 + c]) + V;
  } else {
    const int32_t k = f - Fi, b = N + k, i = m.owner[f];
    const int p = m.bpatch[k];
    if (m.ptype[p] == PATCH_INLET || m.ptype[p] == PATCH_OUTLET) {
      // fixedValue inlets and zeroGradient outlets: the interior upwind fluxes
      // with U_b in place of U_N (cfdAssembleConvectionTerm.m:144-232)
      const double psi = mdot[f];
      const double C1 = 1.0 * fmax(psi, 0.0);
      const double C2 = (-1.0) * fmax(-psi, 0.0);
      out[0] = C1;
      out[1] = C2;
      out[2] = (C1 * U[i * 3 + c] + C2 * U[b * 3 + c]) + 0.0;
    }
    double C1 = 0.0, T = 0.0;
    if ((m.ptype[p] == PATCH_WALL || m.ptype[p] == PATCH_INLET) && m.u_bc[p] == BC_FIXED_VALUE) {
      // fixedValue walls (cfdAssembleDiffusionTerm.m:224-263) and inlets
      // (:135-173) have the same fluxes; zeroGradient outlets none (:201-218)
      C1 = (1.0 * gam) * m.gDiff[f];
      const double C2 = ((-1.0) * gam) * m.gDiff[f];
      const double V = ((-1.0) * gam) * dot3(&gradU[(b * 3 + c) * 3], &m.T[f * 3]);
      T = (C1 * U[i * 3 + c] + C2 * U[b * 3 + c]) + V;
    } else if (m.ptype[p] == PATCH_WALL && m.u_bc[p] == BC_NO_SLIP) {
      // vector noSlip wall (cfdAssembleDiffusionTerm.m:479-562): TM =
      // gamma_b / walldist (cfdGetUCoef.m:41-44), and dot(nx',nx') over the
      // patch's rows is one patch-wide sum of n_c^2 (kept as the reference has it)
      const double A = m.area[f], dw = m.walldist[f], TM = gam / dw;
      double n[3];
      for (int d = 0; d < 3; ++d) n[d] = m.Sf[f * 3 + d] / A;
      const double Sc = m.noslip_sum[p * 3 + c];
      const int j = c == 0 ? 1 : 0, k2 = c == 2 ? 1 : 2;
      const double g = ((((1.0 * gam) * A) / dw) * TM);
      C1 = g * (1 - Sc);
      const double C2 = 0.0;
      const double V = g * ((U[b * 3 + c] * (1 - Sc) +
                             ((U[i * 3 + j] - U[b * 3 + j]) * n[j]) * n[c]) -
                            ((U[i * 3 + k2] - U[b * 3 + k2]) * n[k2]) * n[c]);
      T = (C1 * U[i * 3 + c] + C2 * U[b * 3 + c]) + V;
    } else if (plane(m, p)) {
      const double A = m.area[f];
      double n[3], l[3] = {1.0, 1.0, 1.0}, a[3];
      l[c] = 0.0;
      for (int d = 0; d < 3; ++d) n[d] = m.Sf[f * 3 + d] / A;
      const double dd = dot3(&m.CN[f * 3], n);
      const double nc = n[c];
      C1 = ((((1.0 * 2) * gam) * A) * (nc * nc)) / dd;
      const double C2 = 0.0;
      for (int d = 0; d < 3; ++d) a[d] = (U[i * 3 + d] * l[d]) * (n[d] * l[d]);
      const double dl = (a[0] + a[1]) + a[2];
      const double V = (((((-1.0) * 2) * gam) * A) * dl) * nc / dd;
      T = (C1 * U[i * 3 + c] + C2 * U[b * 3 + c]) + V;
    }
    out[3] = C1;
    out[5] = T;
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) uf[(int64_t)f * 6 + q] = out[q];
}

// One momentum component c from the face fluxes, terms in equation order
// (cfdAssembleEquationTerms.m:22-55): convection over the cell's faces, then
// the divergence term per cell (cfdAssembleConvectionDivergenceTerm.m:38-47),
// diffusion over the faces (anb = (0 + convection) + diffusion, the two
// contributions each slot receives), -grad(p), sign +1
// (cfdAssemblePressureGradientTerm.m:31-40); false transient, fdt = 1e9,
// rho = 1 (cfdAssembleFalseTransientTerm.m:28-49); URF (cfdApplyURF.m:15-17);
// normalised residual (cfdComputeNormalizedResidual.m:64-88); DU = V/ac,
// DUT = ac_old/ac (cfdPostAssembleEquation.m:37-41).
// the terms of one cell-face entry of k_assemble_u: face f, side slot sl,
// the six face fluxes q and mdot md
struct UEntry {
  int32_t f, sl;
  double q[6], md;
};

__device__ __forceinline__ void load_uentry(const DevMesh &m, const double *__restrict__ uf,
                                            const double *__restrict__ mdot, int32_t e,
                                            UEntry &x) {
  x.f = m.cf_face[e];
  x.sl = m.cf_slot[e];
#pragma unroll
  for (int q = 0; q < 6; ++q) x.q[q] = uf[(int64_t)x.f * 6 + q];
  x.md = mdot[x.f];
}

// one entry's share of each face pass of k_assemble_u
__device__ __forceinline__ void conv_entry(const UEntry &x, int32_t Fi, double &ac, double &bc) {
  if (x.f < Fi) {
    if (x.sl >= 0) {
      ac = ac + x.q[0];
      bc = bc - x.q[2];
    } else {
      ac = ac - x.q[1];
      bc = bc + x.q[2];
    }
  } else {
    // boundary faces: ac += C1, bc -= T (zero except on inlets and outlets)
    ac = ac + x.q[0];
    bc = bc - x.q[2];
  }
}

__device__ __forceinline__ void div_entry(const UEntry &x, int32_t Fi, double &eff) {
  if (x.f < Fi && x.sl < 0) eff = eff - x.md;
  else eff = eff + x.md;
}

__device__ __forceinline__ void div_term(double eff, double phi, double &ac, double &bc) {
  const double CE = fmax(eff, 0.0) - eff;
  const double VE = (-fmax(eff, 0.0)) * phi;
  const double TE = CE * phi + VE;
  ac = ac + CE;
  bc = bc - TE;
}

__device__ __forceinline__ void diff_entry(const UEntry &x, int32_t Fi, double *anb, double &ac,
                                           double &bc) {
  const double *q = x.q;
  if (x.f < Fi) {
    const int32_t sl = x.sl;
    if (sl >= 0) {
      ac = ac + q[3];
      anb[sl] = (0.0 + q[1]) + q[4];
      bc = bc - q[5];
    } else {
      ac = ac - q[4];
      anb[-1 - sl] = (0.0 - q[0]) - q[3];
      bc = bc + q[5];
    }
  } else {
    ac = ac + q[3];
    bc = bc - q[5];
  }
}

// One momentum component c from the face fluxes, terms in equation order
// (cfdAssembleEquationTerms.m:22-55): convection over the cell's faces, then
// the divergence term per cell (cfdAssembleConvectionDivergenceTerm.m:38-47),
// diffusion over the faces (anb = (0 + convection) + diffusion, the two
// contributions each slot receives), -grad(p), sign +1
// (cfdAssemblePressureGradientTerm.m:31-40); false transient, fdt = 1e9,
// rho = 1 (cfdAssembleFalseTransientTerm.m:28-49); URF (cfdApplyURF.m:15-17);
// normalised residual (cfdComputeNormalizedResidual.m:64-88); DU = V/ac,
// DUT = ac_old/ac (cfdPostAssembleEquation.m:37-41).  The entries of a cell
// with at most 8 faces are loaded together before the passes.
__global__ void __launch_bounds__(kBlock)
k_assemble_u(DevMesh m, Flow fl, int c, double urf, const double *__restrict__ U,
             const double *__restrict__ gradp, const double *__restrict__ mdot,
             const double *__restrict__ uf, const double *__restrict__ stress, double *ac_out,
             double *anb, double *bc_out, double *acold_out, double *dphi, double *DU,
             double *DUT, double *partial, double *scal) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  double red[2] = {0.0, 0.0};
  if (i < m.N) {
    const int32_t Fi = m.Fi;
    double ac = 0.0, bc = 0.0, acold = 0.0, eff = 0.0;
    const int32_t e0 = m.cf_ptr[i], e1 = m.cf_ptr[i + 1], nf = e1 - e0;
    const double phi = U[i * 3 + c];
    if (fl.transient) {
      // ddt(rho*U), Euler, the equation's first term
      // (cfdAssembleTransientTermEuler.m:13-44): FLUXCE = V rho / dt,
      // FLUXCEOLD = -V rho_old / dt, and the source FLUXCE .* phi' +
      // FLUXCEOLD .* phi_old' is N x N, read by element index through its
      // first column: every cell takes cell 1's phi and phi_old (kept)
      const double v = m.vol[i];
      const double CE = ((1.0 * v) * fl.rho) / fl.dt;
      const double CEOLD = (((-1.0) * v) * fl.rho) / fl.dt;
      const double TE = CE * U[0 * 3 + c] + CEOLD * fl.U_old[0 * 3 + c];
      ac = ac + CE;
      acold = acold + CEOLD;
      bc = bc - TE;
    }
    if (nf <= 8) {
      UEntry x[8];
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u < nf) load_uentry(m, uf, mdot, e0 + u, x[u]);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u < nf) conv_entry(x[u], Fi, ac, bc);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u < nf) div_entry(x[u], Fi, eff);
      div_term(eff, phi, ac, bc);
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (u < nf) diff_entry(x[u], Fi, anb, ac, bc);
    } else {
      // many faces: the same passes, loading one entry at a time
      UEntry x;
      for (int32_t e = e0; e < e1; ++e) {
        load_uentry(m, uf, mdot, e, x);
        conv_entry(x, Fi, ac, bc);
      }
      for (int32_t e = e0; e < e1; ++e) {
        load_uentry(m, uf, mdot, e, x);
        div_entry(x, Fi, eff);
      }
      div_term(eff, phi, ac, bc);
      for (int32_t e = e0; e < e1; ++e) {
        load_uentry(m, uf, mdot, e, x);
        diff_entry(x, Fi, anb, ac, bc);
      }
    }
    if (stress) {
      // the explicit stress fluxes, scattered as the reference's face loop
      // does (C1 = C2 = 0: the diagonal takes +-0, the off-diagonals keep
      // their values)
      for (int32_t e = e0; e < e1; ++e) {
        const int32_t f = m.cf_face[e];
        const double Ts = stress[f];
        if (f < Fi && m.cf_slot[e] < 0) {
          ac = ac - 0.0;
          bc = bc + Ts;
        } else {
          ac = ac + 0.0;
          bc = bc - Ts;
        }
      }
    }
    const double vol = m.vol[i];
    // -grad(p)
    {
      const double TE = (1.0 * vol) * gradp[i * 3 + c];
      ac = ac + 0.0;
      acold = acold + 0.0;
      bc = bc - TE;
    }
    // false transient of steady runs (not in transient ones,
    // cfdAssembleEquationTerms.m:44-47)
    if (!fl.transient) {
      const double fdt = 1e9;
      const double CE = (vol * fl.rho) / fdt;
      const double CEOLD = ((-vol) * fl.rho) / fdt;
      const double TE = CE * U[i * 3 + c];
      ac = ac + CE;
      acold = acold + CEOLD;
      bc = bc - TE;
    }
    ac = ac / urf;
    ac_out[i] = ac;
    bc_out[i] = bc;
    acold_out[i] = acold;
    dphi[i] = 0.0;
    // transient runs normalise by ac - V rho / dt, floored at 1e-6 of it
    // (cfdComputeNormalizedResidual.m:55-76)
    double an = ac;
    if (fl.transient) {
      const double at = (vol * fl.rho) / fl.dt;
      an = an - at;
      if (an < 1e-6 * at) an = at;
    }
    const double r = bc / (an * scal[S_SCALE]);
    red[0] = fabs(r);
    red[1] = r * r;
    DU[i * 3 + c] = vol / ac;
    DUT[i * 3 + c] = acold / ac;
  }
  block_partials2(red[0], red[1], 1, 0, partial);
}

// boundary rows of DU / DUT from their owners (cfdPostAssembleEquation.m:75-81)
__global__ void k_du_bnd(DevMesh m, int c, double *DU, double *DUT) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m.F - m.Fi) return;
  const int32_t o = m.owner[m.Fi + k], b = m.N + k;
  DU[b * 3 + c] = DU[o * 3 + c];
  DUT[b * 3 + c] = DUT[o * 3 + c];
}

void assemble_u(hipStream_t s, const DevMesh &m, const Flow &fl, int comp, double urf,
                const double *U,
                const double *gradU, const double *gradp, const double *mdot,
                const double *gamma, double *uf, double *stress, double *ac, double *anb,
                double *bc, double *ac_old, double *dphi, double *DU, double *DUT, double *partial,
                unsigned *counter, double *scal) {
  (void)counter;
  hipLaunchKernelGGL(k_u_faces, dim3(nblocks(m.F)), dim3(kBlock), 0, s, m, comp, U, gradU, mdot,
                     gamma, uf, stress);
  hipLaunchKernelGGL(k_assemble_u, dim3(nblocks(m.N)), dim3(kBlock), 0, s, m, fl, comp, urf, U, gradp,
                     mdot, uf, stress, ac, anb, bc, ac_old, dphi, DU, DUT, partial, scal);
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, s, partial, nblocks(m.N), 1, scal);
  const int32_t Fb = m.F - m.Fi;
  if (Fb > 0) hipLaunchKernelGGL(k_du_bnd, dim3(nblocks(Fb)), dim3(kBlock), 0, s, m, comp, DU, DUT);
}

// U(:,c) += dphi (cfdCorrectScalarEquation.m:16-24)
__global__ void k_correct_u_cells(DevMesh m, int c, const double *dphi, double *U) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i < m.N) U[i * 3 + c] = U[i * 3 + c] + dphi[i];
}

// boundary rows of the momentum correction of component c, by U's condition
// (cfdCorrectScalarEquation.m:31-107): slip walls, empty and symmetry planes
// U_b,c = U_O,c - (U_O . n) n_c (:239-293, :469-584); zeroGradient outlets
// U_b,c = U_O,c (:376-407); fixedValue / noSlip walls and inlets unchanged
__global__ void k_bnd_u(DevMesh m, int c, double *U) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m.F - m.Fi) return;
  const int p = m.bpatch[k];
  const int32_t f = m.Fi + k, o = m.owner[f], b = m.N + k;
  if (m.ptype[p] == PATCH_OUTLET && m.u_bc[p] == BC_ZERO_GRADIENT) {
    U[b * 3 + c] = U[o * 3 + c];
    return;
  }
  if (!u_tangential(m, p)) return;
  double n[3];
  const double nm = mag3(&m.Sf[f * 3]);
  for (int d = 0; d < 3; ++d) n[d] = m.Sf[f * 3 + d] / nm;
  const double un = dot3(&U[o * 3], n);
  U[b * 3 + c] = U[o * 3 + c] - un * n[c];
}

void correct_u(hipStream_t s, const DevMesh &m, int comp, const double *dphi, double *U) {
  hipLaunchKernelGGL(k_correct_u_cells, dim3(nblocks(m.N)), dim3(kBlock), 0, s, m, comp, dphi, U);
  const int32_t Fb = m.F - m.Fi;
  if (Fb > 0) hipLaunchKernelGGL(k_bnd_u, dim3(nblocks(Fb)), dim3(kBlock), 0, s, m, comp, U);
}

// ------------------------------------------------- pressure-correction path

// Rhie-Chow face flux of every interior face (cfdAssembleMdotTerm.m:70-212):
// mdot_f = C1 p_O + C2 p_N + V with C1 = rho_f geoDiff = -C2 and
// V = rho_f Ubar.S - rho_f grad(p)_f.DUT + rho_f grad(p)bar_f.DUS
//     + (1-urf_U)(mdot_prev - rho_f Ubar.S);  boundary fluxes of walls and
// planes are zero, inlets and fixed-pressure outlets get theirs and a new
// mdot_b (:281-348).
__global__ void __launch_bounds__(kBlock)
k_rhie_chow(DevMesh m, Flow fl, double urf_u, const double *__restrict__ U,
            const double *__restrict__ DU, const double *__restrict__ DUTc,
            const double *__restrict__ p, const double *__restrict__ gradp, double *mdot,
            double *fluxC1, double *fluxC2) {
  const int32_t f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= m.F) return;
  if (f >= m.Fi) {
    const int32_t k = f - m.Fi, b = m.N + k, o = m.owner[f];
    const int q = m.bpatch[k];
    if (!p_flux(m, q)) {
      fluxC1[f] = 0.0;
      fluxC2[f] = 0.0;
      return;
    }
    const double *Sb = &m.Sf[f * 3];
    double C1 = 0.0, C2 = 0.0, V = 0.0;
    if (m.ptype[q] == PATCH_INLET) {
      // zeroGradient-p inlet: term I only (cfdAssembleMdotTerm.m:415-486)
      V = V + fl.rho * dot3(Sb, &U[b * 3]);
    } else {
      // fixed-pressure outlet: terms I-III from the boundary rows of DU and
      // grad(p), FLUXCb2 = 0 (cfdAssembleMdotTerm.m:799-906)
      const double *CNb = &m.CN[f * 3], *eCNb = &m.eCN[f * 3];
      double DUS[3], eDUS[3], DUE[3], DUT[3];
      for (int d = 0; d < 3; ++d) DUS[d] = DU[b * 3 + d] * Sb[d];
      const double mDUS = mag3(DUS);
      for (int d = 0; d < 3; ++d) eDUS[d] = DUS[d] / mDUS;
      const double ee = dot3(eCNb, eDUS);
      for (int d = 0; d < 3; ++d) DUE[d] = (mDUS / ee) * eCNb[d];
      const double geoDiff = mag3(DUE) / mag3(CNb);
      for (int d = 0; d < 3; ++d) DUT[d] = DUS[d] - DUE[d];
      V = V + fl.rho * dot3(&U[b * 3], Sb);
      C1 = C1 + fl.rho * geoDiff;
      C2 = C2 - 0.0;
      V = V - fl.rho * dot3(&gradp[b * 3], DUT);
      V = V + fl.rho * dot3(&gradp[b * 3], DUS);
    }
    mdot[f] = (C1 * p[o] + C2 * p[b]) + V;
    fluxC1[f] = C1;
    fluxC2[f] = C2;
    return;
  }
  const int32_t o = m.owner[f], nb = m.neighbour[f];
  const double w = m.gf[f];
  double du[3], ub[3], gpb[3], gpf[3];
  for (int d = 0; d < 3; ++d) du[d] = w * DU[nb * 3 + d] + (1 - w) * DU[o * 3 + d];
  const double mprev = mdot[f];
  for (int d = 0; d < 3; ++d) ub[d] = w * U[nb * 3 + d] + (1 - w) * U[o * 3 + d];
  // rho_f: 'Upwind' by the previous mdot_f (cfdAssembleMdotTerm.m:118-121)
  const double pos = mprev > 0 ? 1.0 : 0.0;
  const double rho_f = fl.rho * pos + fl.rho * (1 - pos);
  for (int d = 0; d < 3; ++d) gpb[d] = (1 - w) * gradp[nb * 3 + d] + w * gradp[o * 3 + d];
  grad_face<1>(m, p, gradp, 0, f, gpf);
  const double *Sf = &m.Sf[f * 3], *CF = &m.CN[f * 3], *eCF = &m.eCN[f * 3];
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
  if (fl.transient) {
    // terms VI-VII (:175-201): DUT_f (mdot_old - rho_old_f [v_old]_f . S_f),
    // DUT = mean of the three DUT components per element; DUT_f and
    // [v_old]_f Average-interpolated, rho_old_f 'Upwind' by the previous mdot_f
    auto dmean = [&](int32_t e) {
      return ((DUTc[e * 3] + DUTc[e * 3 + 1]) + DUTc[e * 3 + 2]) / 3;
    };
    const double dut = w * dmean(nb) + (1 - w) * dmean(o);
    const double rho_old_f = fl.rho * pos + fl.rho * (1 - pos);
    double uo[3];
    for (int d = 0; d < 3; ++d) uo[d] = w * fl.U_old[nb * 3 + d] + (1 - w) * fl.U_old[o * 3 + d];
    V = V + dut * (fl.mdot_old[f] - rho_old_f * dot3(uo, Sf));
  }
  V = V + (1 - urf_u) * (mprev - rho_f * UbS);
  mdot[f] = (FC * p[o] + FF * p[nb]) + V;
  fluxC1[f] = FC;
  fluxC2[f] = FF;
}

void rhie_chow(hipStream_t s, const DevMesh &m, const Flow &fl, double urf_u, const double *U,
               const double *DU, const double *DUT, const double *p, const double *gradp,
               double *mdot, double *fluxC1, double *fluxC2) {
  hipLaunchKernelGGL(k_rhie_chow, dim3(nblocks(m.F)), dim3(kBlock), 0, s, m, fl, urf_u, U, DU,
                     DUT, p, gradp, mdot, fluxC1, fluxC2);
}

// pressure-correction matrix from the face fluxes (T_f = mdot_f), element
// fluxes all zero, residual = effective divergence of the new mdot
// (cfdComputeNormalizedResidual.m:27-30: max and rms, not of |.|)
__global__ void __launch_bounds__(kBlock)
k_assemble_p(DevMesh m, const double *__restrict__ fluxC1, const double *__restrict__ fluxC2,
             const double *__restrict__ mdot, double *ac_out, double *anb, double *bc_out,
             double *acold_out, double *dphi, double *partial, int32_t *changed) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  double red[2] = {-INFINITY, 0.0};
  bool differs = false;  // a coefficient's bits differ from the array's previous value
  auto put = [&](double *a, int32_t q, double v) {
    differs |= __double_as_longlong(a[q]) != __double_as_longlong(v);
    a[q] = v;
  };
  if (i < m.N) {
    double ac = 0.0, bc = 0.0, acold = 0.0, eff = 0.0;
    const int32_t e0 = m.cf_ptr[i], e1 = m.cf_ptr[i + 1];
    // entries in groups of four: faces first, then their fluxes, then the sums
    for (int32_t e = e0; e < e1; e += 4) {
      int32_t f[4], sl[4];
      double C1[4], C2[4], T[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (e + u < e1) {
          f[u] = m.cf_face[e + u];
          sl[u] = m.cf_slot[e + u];
        }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        if (e + u < e1) {
          T[u] = mdot[f[u]];
          if (f[u] < m.Fi) {
            C1[u] = fluxC1[f[u]];
            C2[u] = fluxC2[f[u]];
          }
        }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if (e + u >= e1) break;
        if (f[u] < m.Fi) {
          if (sl[u] >= 0) {  // each anb slot has this one contribution
            ac = ac + C1[u];
            put(anb, sl[u], 0.0 + C2[u]);
            bc = bc - T[u];
            eff = eff + T[u];
          } else {
            ac = ac - C2[u];
            put(anb, -1 - sl[u], 0.0 - C1[u]);
            bc = bc + T[u];
            eff = eff - T[u];
          }
        } else if (p_flux(m, m.bpatch[f[u] - m.Fi])) {
          // inlet / outlet: ac += FLUXC1, bc -= FLUXT = mdot_b
          ac = ac + fluxC1[f[u]];
          bc = bc - T[u];
          eff = eff + T[u];
        } else {
          ac = ac + 0.0;
          bc = bc - 0.0;
          eff = eff + T[u];
        }
      }
    }
    ac = ac + 0.0;
    acold = acold + 0.0;
    bc = bc - 0.0;
    put(ac_out, i, ac);
    bc_out[i] = bc;
    acold_out[i] = acold;
    dphi[i] = 0.0;
    red[0] = eff;
    red[1] = eff * eff;
  }
  if (differs) *changed = 1;  // the same value from every writer
  block_partials2(red[0], red[1], 1, 0, partial);
}

void assemble_p(hipStream_t s, const DevMesh &m, const double *fluxC1, const double *fluxC2,
                const double *mdot, double *ac, double *anb, double *bc, double *ac_old,
                double *dphi, double *partial, double *scal, int32_t *changed) {
  hipLaunchKernelGGL(k_assemble_p, dim3(nblocks(m.N)), dim3(kBlock), 0, s, m, fluxC1, fluxC2,
                     mdot, ac, anb, bc, ac_old, dphi, partial, changed);
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(kBlock), 0, s, partial, nblocks(m.N), 1, scal);
}

// PP = dphi - dphi(ref) (cfdCorrectPPField.m:26-52)
__global__ void k_pp_cells(DevMesh m, int32_t ref, const double *dphi, double *PP) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m.N) return;
  const double ppc = ref >= 0 ? dphi[ref] : 0.0;
  PP[i] = dphi[i] - ppc;
}

// PP_b (cfdCorrectPPField.m:57-335): 0 on fixed-pressure outlets (:249-272),
// the owner's value on every other patch
__global__ void k_pp_bnd(DevMesh m, double *PP) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m.F - m.Fi) return;
  PP[m.N + k] = p_fixed_outlet(m, m.bpatch[k]) ? 0.0 : PP[m.owner[m.Fi + k]];
}

// p_b (cfdCorrectPressureEquation.m:60-290): the owner's value except on
// fixed-pressure outlets, whose p_b is kept (:230-233)
__global__ void k_p_bnd(DevMesh m, double *p) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m.F - m.Fi) return;
  if (!p_fixed_outlet(m, m.bpatch[k])) p[m.N + k] = p[m.owner[m.Fi + k]];
}

// U -= DU o grad(PP) on cells (cfdCorrectVelocityField.m:26-55)
__global__ void k_correct_u_pp(DevMesh m, const double *DU, const double *gradPP, double *U) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m.N) return;
  for (int d = 0; d < 3; ++d) U[i * 3 + d] = U[i * 3 + d] - DU[i * 3 + d] * gradPP[i * 3 + d];
}

// boundary rows of U after the pressure correction, patches dispatched on p's
// condition in patch order (cfdCorrectVelocityField.m:57-121):
//  - empty and symmetry planes: U_b = U_O - (U_O . n) n (:274-343);
//  - fixed-pressure outlets: U(iBFaces,:) = U(iOwners,:) (:247-272), i.e. the
//    row numbered by the outlet face's own index gets that face's owner's U;
//    such a row is another patch's boundary element here (outlet faces below
//    n_cells are refused at case creation, rows past the end are not stored);
//  - walls with zeroGradient p and zeroGradient-p inlets: nothing (:173-182).
// Row r = N + k is written by its own patch (if a plane) and by the outlet
// whose face has index r; the later patch in order wins, as in the loop.
// Only interior rows of U are read, so the rows are independent.
__global__ void k_bnd_u_p(DevMesh m, double *U) {
  const int32_t k = blockIdx.x * kBlock + threadIdx.x;
  if (k >= m.F - m.Fi) return;
  const int32_t r = m.N + k;
  const int pk = m.bpatch[k];
  int q = -1;
  if (r >= m.Fi && r < m.F && p_fixed_outlet(m, m.bpatch[r - m.Fi])) q = m.bpatch[r - m.Fi];
  if (q >= 0 && (!plane(m, pk) || q > pk)) {
    const int32_t o = m.owner[r];
    for (int d = 0; d < 3; ++d) U[r * 3 + d] = U[o * 3 + d];
  } else if (plane(m, pk)) {
    const int32_t f = m.Fi + k, o = m.owner[f];
    double n[3];
    const double nm = mag3(&m.Sf[f * 3]);
    for (int d = 0; d < 3; ++d) n[d] = m.Sf[f * 3 + d] / nm;
    const double un = dot3(&U[o * 3], n);
    for (int d = 0; d < 3; ++d) U[r * 3 + d] = U[o * 3 + d] - un * n[d];
  }
}

// p += urf_p (PP - PP(ref)) (cfdCorrectPressureEquation.m:22-54)
__global__ void k_correct_p_cells(DevMesh m, int32_t ref, double urf_p, const double *PP,
                                  double *p) {
  const int32_t i = blockIdx.x * kBlock + threadIdx.x;
  if (i >= m.N) return;
  const double ppf = ref >= 0 ? PP[ref] : 0.0;
  p[i] = p[i] + urf_p * (PP[i] - ppf);
}

// mdot += C1 PP_O + C2 PP_N (cfdCorrectMdotField.m:24-50); boundary faces by
// p's condition (:56-120): planes 0 mdot_b (:306-346), fixed-pressure outlets
// rho_b (U_O . S_b) with the corrected owner U (:264-304), walls with
// zeroGradient p and incompressible inlets unchanged (:156-208)
__global__ void k_correct_mdot(DevMesh m, double rho, const double *fluxC1,
                               const double *fluxC2, const double *PP, const double *U,
                               double *mdot) {
  const int32_t f = blockIdx.x * kBlock + threadIdx.x;
  if (f >= m.F) return;
  if (f < m.Fi) {
    mdot[f] = (mdot[f] + fluxC1[f] * PP[m.owner[f]]) + fluxC2[f] * PP[m.neighbour[f]];
    return;
  }
  const int q = m.bpatch[f - m.Fi];
  if (plane(m, q)) {
    mdot[f] = 0.0 * mdot[f];
  } else if (p_fixed_outlet(m, q)) {
    mdot[f] = rho * dot3(&U[m.owner[f] * 3], &m.Sf[f * 3]);
  }
}

void correct_p(hipStream_t s, const DevMesh &m, double rho, int32_t ref, double urf_p,
               const double *dphi,
               double *PP, double *gradPP, const double *DU, double *U, double *p,
               const double *fluxC1, const double *fluxC2, double *mdot) {
  const int32_t Fb = m.F - m.Fi;
  hipLaunchKernelGGL(k_pp_cells, dim3(nblocks(m.N)), dim3(kBlock), 0, s, m, ref, dphi, PP);
  if (Fb > 0) hipLaunchKernelGGL(k_pp_bnd, dim3(nblocks(Fb)), dim3(kBlock), 0, s, m, PP);
  gradient(s, m, PP, 1, gradPP);
  hipLaunchKernelGGL(k_correct_u_pp, dim3(nblocks(m.N)), dim3(kBlock), 0, s, m, DU, gradPP, U);
  if (Fb > 0) hipLaunchKernelGGL(k_bnd_u_p, dim3(nblocks(Fb)), dim3(kBlock), 0, s, m, U);
  hipLaunchKernelGGL(k_correct_p_cells, dim3(nblocks(m.N)), dim3(kBlock), 0, s, m, ref, urf_p, PP, p);
  if (Fb > 0) hipLaunchKernelGGL(k_p_bnd, dim3(nblocks(Fb)), dim3(kBlock), 0, s, m, p);
  hipLaunchKernelGGL(k_correct_mdot, dim3(nblocks(m.F)), dim3(kBlock), 0, s, m, rho, fluxC1,
                     fluxC2, PP, U, mdot);
}

}  // namespace sk
}  // namespace ufvm
