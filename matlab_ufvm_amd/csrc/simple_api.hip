// simple_api.hip -- C ABI of the SIMPLE outer iteration (ufvm_case_*,
// ufvm_update_fields, ufvm_assemble_equation, ufvm_solve_equation,
// ufvm_correct_equation, ufvm_simple_iterate) and its host orchestration.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "internal.h"
#include "kernels.h"
#include "simple_kernels.h"
#include "state.h"
#include "ufvm.h"

using namespace ufvm;
using namespace ufvm::detail;

namespace ufvm {

struct Case {
  int32_t N = 0, F = 0, Fi = 0, Fb = 0, P = 0;
  DBuf<int32_t> owner, neighbour, cptr, ccol, own_nnz, nei_nnz, cf_ptr, cf_face, cf_slot, bpatch;
  DBuf<int32_t> ptype, u_bc, p_bc;
  DBuf<double> Sf, area, CN, eCN, gDiff, T, gf, fc, vol, cc, u_val, p_val, walldist, noslip_sum;
  DBuf<double> U, gradU, p, gradp, PP, gradPP, DU, DUT, mdot, gamma, fluxC1, fluxC2, ac_old;
  DBuf<double> U_old, mdot_old;  // the previous time step's U and mdot_f ('Step1')
  DBuf<double> uface;  // momentum face fluxes, 6 per face
  DBuf<double> ustress;  // explicit stress flux per face (stress_term only)
  // 1 when the pressure-correction coefficients changed since the kept
  // hierarchy was last refreshed from them (set by the assembly)
  DBuf<int32_t> p_changed;
  sk::DevMesh dm{};
  ufvm_simple_cfg cfg{};
  std::vector<int32_t> h_ptype, h_u_bc, h_p_bc;
  std::vector<double> h_u_val, h_p_val;
  int32_t sysU = 0, sysP = 0;
  bool p_hier_built = false;
  // a fixed-pressure outlet's velocity correction writes U rows past the end
  // with unwritten rows between them (outlet_gap); once it has run, the
  // reference's grown U holds zero rows that its scale sees (scale_zero)
  bool outlet_gap = false, scale_zero = false;
  double rho = 1.0;

  sk::Flow flow() const {
    return sk::Flow{rho, cfg.transient, cfg.dt, U_old.p, mdot_old.p};
  }
};

}  // namespace ufvm

namespace {

Case &case_of(ufvm_ctx *ctx, int32_t id) {
  auto it = ctx->cases.find(id);
  if (it == ctx->cases.end()) throw Error(UFVM_ERR_ARG, "unknown case id");
  return *static_cast<Case *>(it->second.get());
}

template <typename T>
void up(DBuf<T> &d, const T *h, size_t n, hipStream_t s, const char *what) {
  if (n && !h) throw Error(UFVM_ERR_ARG, std::string("null mesh array: ") + what);
  d.upload(h, n, s);
}

// cfdInitializeCase.m: the uniform internal fields u_init / p_init; boundary
// values from the BCs (cfdInitializeBoundaryConditions.m: fixedValue walls and
// inlets the value; noSlip zero; slip walls, empty and symmetry planes the
// owner minus its normal part; zeroGradient outlets an owner copy; p_b an owner
// copy or, on fixed-pressure outlets, the value); mdot_f = rho_f (S_f . U_f)
// with Average interpolation (cfdInitializeMdotField.m:72-80)
void init_fields(ufvm_ctx *ctx, Case &C) {
  hipStream_t s = ctx->stream;
  const int32_t NB = C.N + C.Fb;
  std::vector<double> U((size_t)NB * 3, 0.0), P(NB, 0.0), zero((size_t)NB * 9, 0.0), mdot(C.F, 0.0);
  std::vector<int32_t> owner(C.F), nbr(std::max(C.Fi, 1));
  std::vector<double> Sf((size_t)C.F * 3), gf(std::max(C.Fi, 1));
  C.owner.download(owner.data(), C.F, s);
  if (C.Fi > 0) C.neighbour.download(nbr.data(), C.Fi, s);
  C.Sf.download(Sf.data(), (size_t)C.F * 3, s);
  if (C.Fi > 0) C.gf.download(gf.data(), C.Fi, s);
  std::vector<int32_t> bp(C.Fb);
  if (C.Fb > 0) C.bpatch.download(bp.data(), C.Fb, s);
  UFVM_HIP(hipStreamSynchronize(s));
  for (int32_t i = 0; i < C.N; ++i) {
    for (int d = 0; d < 3; ++d) U[i * 3 + d] = C.cfg.u_init[d];
    P[i] = C.cfg.p_init;
  }
  for (int32_t k = 0; k < C.Fb; ++k) {
    const int p = bp[k];
    const int32_t f = C.Fi + k, b = C.N + k, o = owner[f];
    const int t = C.h_ptype[p], ub = C.h_u_bc[p];
    if (ub == UFVM_BC_FIXED_VALUE) {
      for (int d = 0; d < 3; ++d) U[b * 3 + d] = C.h_u_val[p * 3 + d];
    } else if (ub == UFVM_BC_NO_SLIP) {
      // noSlip's value '[0;0;0]' (cfdDefineEquation.m:290-295)
      for (int d = 0; d < 3; ++d) U[b * 3 + d] = 0.0;
    } else if (t == UFVM_PATCH_EMPTY || t == UFVM_PATCH_SYMMETRY || ub == UFVM_BC_SLIP) {
      const double *S = &Sf[f * 3];
      const double nm = std::sqrt((S[0] * S[0] + S[1] * S[1]) + S[2] * S[2]);
      double n[3] = {S[0] / nm, S[1] / nm, S[2] / nm};
      const double un = (U[o * 3] * n[0] + U[o * 3 + 1] * n[1]) + U[o * 3 + 2] * n[2];
      for (int d = 0; d < 3; ++d) U[b * 3 + d] = U[o * 3 + d] - un * n[d];
    } else {
      for (int d = 0; d < 3; ++d) U[b * 3 + d] = U[o * 3 + d];
    }
    P[b] = C.h_p_bc[p] == UFVM_BC_FIXED_VALUE ? C.h_p_val[p] : P[o];
  }
  for (int32_t f = 0; f < C.F; ++f) {
    const double *S = &Sf[f * 3];
    double uf[3], rf;
    if (f < C.Fi) {
      const double g = gf[f];
      const int32_t o = owner[f], nb = nbr[f];
      for (int d = 0; d < 3; ++d) uf[d] = g * U[nb * 3 + d] + (1 - g) * U[o * 3 + d];
      rf = g * C.rho + (1 - g) * C.rho;
    } else {
      const int32_t b = C.N + (f - C.Fi);
      for (int d = 0; d < 3; ++d) uf[d] = U[b * 3 + d];
      rf = C.rho;
    }
    mdot[f] = rf * ((S[0] * uf[0] + S[1] * uf[1]) + S[2] * uf[2]);
  }
  C.U.upload(U, s);
  C.gradU.upload(zero, s);
  C.p.upload(P, s);
  C.gradp.upload(zero.data(), (size_t)NB * 3, s);
  C.PP.upload(zero.data(), NB, s);
  C.gradPP.upload(zero.data(), (size_t)NB * 3, s);
  C.DU.upload(zero.data(), (size_t)NB * 3, s);
  C.DUT.upload(zero.data(), (size_t)NB * 3, s);
  C.mdot.upload(mdot, s);
  C.U_old.upload(U, s);
  C.mdot_old.upload(mdot, s);
  C.gamma.upload(zero.data(), C.F, s);
  C.fluxC1.upload(zero.data(), C.F, s);
  C.fluxC2.upload(zero.data(), C.F, s);
  C.ac_old.upload(zero.data(), C.N, s);
}

struct FieldRef {
  double *p;
  size_t n;
};

FieldRef field_of(Case &C, const char *name) {
  const size_t NB = (size_t)C.N + C.Fb;
  const std::string n(name ? name : "");
  if (n == "U") return {C.U.p, NB * 3};
  if (n == "gradU") return {C.gradU.p, NB * 9};
  if (n == "p") return {C.p.p, NB};
  if (n == "gradp") return {C.gradp.p, NB * 3};
  if (n == "PP") return {C.PP.p, NB};
  if (n == "gradPP") return {C.gradPP.p, NB * 3};
  if (n == "DU") return {C.DU.p, NB * 3};
  if (n == "DUT") return {C.DUT.p, NB * 3};
  if (n == "mdot_f") return {C.mdot.p, (size_t)C.F};
  if (n == "gamma_f") return {C.gamma.p, (size_t)C.F};
  if (n == "U_old") return {C.U_old.p, NB * 3};
  if (n == "mdot_old") return {C.mdot_old.p, (size_t)C.F};
  throw Error(UFVM_ERR_ARG, "unknown field " + n);
}

int eq_of(const char *eq) {
  if (eq && std::strcmp(eq, "U") == 0) return 0;
  if (eq && std::strcmp(eq, "p") == 0) return 1;
  throw Error(UFVM_ERR_ARG, "equation must be \"U\" or \"p\"");
}

// cfdUpdateEquations.m:14-24 for U then p
void update_fields(ufvm_ctx *ctx, Case &C) {
  hipStream_t s = ctx->stream;
  sk::update_u_bc(s, C.dm, C.U.p);
  sk::gradient(s, C.dm, C.U.p, 3, C.gradU.p);
  sk::gamma_faces(s, C.dm, C.cfg.nu, C.gamma.p);
  ensure_partial(ctx, (int64_t)(C.N + C.Fb) * 3 + 1, 2);
  sk::scale_u(s, C.dm, C.U.p, C.scale_zero, ctx->partial.p, ctx->counter.p, ctx->scal.p);
  sk::update_p_bc(s, C.dm, C.p.p);
  sk::gradient(s, C.dm, C.p.p, 1, C.gradp.p);
}

void assemble(ufvm_ctx *ctx, Case &C, int eq, int comp, double *rms, double *mx) {
  hipStream_t s = ctx->stream;
  System &S = sys_of(ctx, eq == 0 ? C.sysU : C.sysP);
  RefLevel &L = *S.lev[0];
  ensure_partial(ctx, C.N, 2);
  if (eq == 0) {
    if (comp < 0 || comp > 2) throw Error(UFVM_ERR_ARG, "component must be 0, 1 or 2");
    sk::assemble_u(s, C.dm, C.flow(), comp, C.cfg.urf_u, C.U.p, C.gradU.p, C.gradp.p, C.mdot.p, C.gamma.p,
                   C.uface.p, C.cfg.stress_term ? C.ustress.p : nullptr, L.ac.p, L.anb.p, L.bc.p, C.ac_old.p, L.dphi.p, C.DU.p, C.DUT.p, ctx->partial.p,
                   ctx->counter.p, ctx->scal.p);
  } else {
    sk::rhie_chow(s, C.dm, C.flow(), C.cfg.urf_u, C.U.p, C.DU.p, C.DUT.p, C.p.p, C.gradp.p,
                  C.mdot.p, C.fluxC1.p, C.fluxC2.p);
    sk::assemble_p(s, C.dm, C.fluxC1.p, C.fluxC2.p, C.mdot.p, L.ac.p, L.anb.p, L.bc.p, C.ac_old.p,
                   L.dphi.p, ctx->partial.p, ctx->scal.p, C.p_changed.p);
  }
  // new values: the coarse operators and the solve-side copies are stale (a
  // kept hierarchy is updated from them by galerkin_update + pcg_refresh)
  S.hier_valid = false;
  S.pcg.stale = true;
  read_scal(ctx, sk::S_RMAX, 2);
  const double m = ctx->hscal[sk::S_RMAX], sq = ctx->hscal[sk::S_RSQ];
  if (mx) *mx = m;
  if (rms) *rms = std::sqrt(sq / C.N);
  if (std::isnan(sq)) throw Error(UFVM_ERR_NUMERIC, "divergence detected (NaN residual)");
}

void solve(ufvm_ctx *ctx, Case &C, int eq, double *r0, double *r1, int32_t *nc) {
  if (eq == 0) {
    System &S = sys_of(ctx, C.sysU);
    solve_alg(ctx, *S.lev[0], C.cfg.u_smoother, C.cfg.u_sweeps, 0.1, r0, r1);
    if (nc) *nc = 0;
    return;
  }
  System &S = sys_of(ctx, C.sysP);
  if (C.cfg.p_solver == 0) {
    ufvm_amg_cfg a = C.cfg.p_amg;
    a.rebuild = 1;
    apply_amg(ctx, S, a, r0, r1, nc);
    return;
  }
  if (C.cfg.p_solver == 1 || !C.p_hier_built) {
    agglomerate(ctx, S, C.cfg.p_amg.max_coarse_levels);
    C.p_hier_built = true;
    pcg_setup(ctx, S, C.cfg.p_pcg);
  } else {
    // first hierarchy's parents and layouts, current coefficients.  When no
    // coefficient's bits changed since the last refresh (the correctors of
    // one SIMPLE iteration share DU, so their matrices are identical), the
    // coarse operators and solve-side copies are still exact.
    int32_t changed = 1;
    C.p_changed.download(&changed, 1, ctx->stream);
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
    if (!changed && S.pcg.ready && S.nlevels >= 2) {
      S.hier_valid = true;
      S.pcg.stale = false;
    } else {
      galerkin_update(ctx, S);
      if (S.pcg.ready) pcg_refresh(ctx, S);
      else pcg_setup(ctx, S, C.cfg.p_pcg);
      C.p_changed.upload(std::vector<int32_t>{0}, ctx->stream);
    }
  }
  double rel = 0.0;
  pcg_solve(ctx, S, nc, &rel);
  if (r0) *r0 = 1.0;
  if (r1) *r1 = rel;
}

void correct(ufvm_ctx *ctx, Case &C, int eq, int comp) {
  hipStream_t s = ctx->stream;
  if (eq == 0) {
    if (comp < 0 || comp > 2) throw Error(UFVM_ERR_ARG, "component must be 0, 1 or 2");
    System &S = sys_of(ctx, C.sysU);
    sk::correct_u(s, C.dm, comp, S.lev[0]->dphi.p, C.U.p);
  } else {
    System &S = sys_of(ctx, C.sysP);
    sk::correct_p(s, C.dm, C.rho, C.cfg.p_ref_cell, C.cfg.urf_p, S.lev[0]->dphi.p, C.PP.p, C.gradPP.p,
                  C.DU.p, C.U.p, C.p.p, C.fluxC1.p, C.fluxC2.p, C.mdot.p);
    if (C.outlet_gap) C.scale_zero = true;
  }
  UFVM_HIP(hipStreamSynchronize(s));
}

// one outer iteration (cfdRunCase.m:92-142 / :42-57): field updates, the
// momentum components, the pressure correctors; r gets the RMS residuals
void outer_iteration(ufvm_ctx *ctx, Case &C, double r[4]) {
  update_fields(ctx, C);
  for (int c = 0; c < 3; ++c) {
    assemble(ctx, C, 0, c, &r[c], nullptr);
    solve(ctx, C, 0, nullptr, nullptr, nullptr);
    correct(ctx, C, 0, c);
  }
  for (int k = 0; k < C.cfg.n_correctors; ++k) {
    assemble(ctx, C, 1, 0, &r[3], nullptr);
    solve(ctx, C, 1, nullptr, nullptr, nullptr);
    correct(ctx, C, 1, 0);
  }
}

}  // namespace

extern "C" {

int ufvm_case_create(ufvm_ctx *ctx, const ufvm_mesh_desc *m, const ufvm_simple_cfg *cfg,
                     int32_t *case_out) {
  return guard(ctx, [&] {
    if (!m || !cfg || !case_out) throw Error(UFVM_ERR_ARG, "null argument");
    if (m->n_cells <= 0 || m->n_interior < 0 || m->n_faces < m->n_interior || m->n_patches < 0)
      throw Error(UFVM_ERR_ARG, "bad mesh sizes");
    hipStream_t s = ctx->stream;
    std::unique_ptr<Case> Cp(new Case());
    Case &C = *Cp;
    C.N = m->n_cells;
    C.F = m->n_faces;
    C.Fi = m->n_interior;
    C.Fb = C.F - C.Fi;
    C.P = m->n_patches;
    C.cfg = *cfg;
    C.rho = cfg->rho == 0.0 ? 1.0 : cfg->rho;
    if (!(C.rho > 0.0)) throw Error(UFVM_ERR_ARG, "rho must be positive");
    if (cfg->transient != 0 && cfg->transient != 1)
      throw Error(UFVM_ERR_ARG, "transient must be 0 or 1");
    if (cfg->transient && !(cfg->dt > 0.0)) throw Error(UFVM_ERR_ARG, "dt must be positive");
    if (cfg->u_smoother != UFVM_SMOOTHER_ILU && cfg->u_smoother != UFVM_SMOOTHER_MCGS)
      throw Error(UFVM_ERR_ARG, "u_smoother must be UFVM_SMOOTHER_ILU or UFVM_SMOOTHER_MCGS");
    // boundary faces -> patch, and patch checks
    std::vector<int32_t> bp(C.Fb, -1);
    for (int p = 0; p < C.P; ++p) {
      const int32_t st = m->patch_start[p], nf = m->patch_nfaces[p];
      if (st < C.Fi || nf < 0 || st + nf > C.F) throw Error(UFVM_ERR_ARG, "bad patch range");
      for (int32_t f = st; f < st + nf; ++f) bp[f - C.Fi] = p;
      const int t = m->patch_type[p];
      if (t != UFVM_PATCH_WALL && t != UFVM_PATCH_EMPTY && t != UFVM_PATCH_SYMMETRY &&
          t != UFVM_PATCH_INLET && t != UFVM_PATCH_OUTLET)
        throw Error(UFVM_ERR_UNSUPPORTED,
                    "only wall, empty, symmetry, inlet and outlet patches are supported");
      const int ub = cfg->u_bc[p], pb = cfg->p_bc[p];
      if (t == UFVM_PATCH_WALL &&
          !((ub == UFVM_BC_FIXED_VALUE || ub == UFVM_BC_NO_SLIP || ub == UFVM_BC_SLIP) &&
            pb == UFVM_BC_ZERO_GRADIENT))
        throw Error(UFVM_ERR_UNSUPPORTED,
                    "walls need U fixedValue, noSlip or slip and p zeroGradient");
      if (t == UFVM_PATCH_EMPTY && !(ub == UFVM_BC_EMPTY && pb == UFVM_BC_EMPTY))
        throw Error(UFVM_ERR_UNSUPPORTED, "empty patches need empty U and p conditions");
      if (t == UFVM_PATCH_SYMMETRY && !(ub == UFVM_BC_SYMMETRY && pb == UFVM_BC_SYMMETRY))
        throw Error(UFVM_ERR_UNSUPPORTED, "symmetry patches need symmetry U and p conditions");
      if (t == UFVM_PATCH_INLET && !(ub == UFVM_BC_FIXED_VALUE && pb == UFVM_BC_ZERO_GRADIENT))
        throw Error(UFVM_ERR_UNSUPPORTED, "inlets need U fixedValue and p zeroGradient");
      if (t == UFVM_PATCH_OUTLET) {
        if (!(ub == UFVM_BC_ZERO_GRADIENT && pb == UFVM_BC_FIXED_VALUE))
          throw Error(UFVM_ERR_UNSUPPORTED, "outlets need U zeroGradient and p fixedValue");
        // the reference's outlet velocity correction writes rows numbered by
        // the face index; below n_cells those rows are cell values
        if (st < C.N)
          throw Error(UFVM_ERR_UNSUPPORTED,
                      "an outlet face index below n_cells: the reference's outlet velocity "
                      "correction would overwrite cell values");
      }
    }
    {
      // rows past the end of U written by the outlets, and whether any row
      // between the end and the last of them stays unwritten
      const int32_t NB = C.N + C.Fb;
      int32_t hi = -1;
      std::vector<char> hit;
      for (int p = 0; p < C.P; ++p)
        if (m->patch_type[p] == UFVM_PATCH_OUTLET)
          hi = std::max(hi, m->patch_start[p] + m->patch_nfaces[p] - 1);
      if (hi >= NB) {
        hit.assign((size_t)(hi - NB + 1), 0);
        for (int p = 0; p < C.P; ++p)
          if (m->patch_type[p] == UFVM_PATCH_OUTLET)
            for (int32_t f = std::max(m->patch_start[p], NB);
                 f < m->patch_start[p] + m->patch_nfaces[p]; ++f)
              hit[f - NB] = 1;
        for (char h : hit) C.outlet_gap |= !h;
      }
    }
    for (int32_t k = 0; k < C.Fb; ++k)
      if (bp[k] < 0) throw Error(UFVM_ERR_ARG, "boundary face outside every patch");
    for (int32_t f = 0; f < C.F; ++f)
      if (m->owner[f] < 0 || m->owner[f] >= C.N) throw Error(UFVM_ERR_ARG, "owner out of range");
    for (int32_t f = 0; f < C.Fi; ++f)
      if (m->neighbour[f] <= m->owner[f] || m->neighbour[f] >= C.N)
        throw Error(UFVM_ERR_ARG, "interior faces need owner < neighbour < n_cells");
    const int32_t nnz = m->cconn_ptr[C.N];
    if (nnz != 2 * C.Fi) throw Error(UFVM_ERR_ARG, "cconn must hold two entries per interior face");
    for (int32_t f = 0; f < C.Fi; ++f) {
      const int32_t a = m->own_nnz[f], b = m->nei_nnz[f];
      if (a < m->cconn_ptr[m->owner[f]] || a >= m->cconn_ptr[m->owner[f] + 1] ||
          m->cconn[a] != m->neighbour[f] || b < m->cconn_ptr[m->neighbour[f]] ||
          b >= m->cconn_ptr[m->neighbour[f] + 1] || m->cconn[b] != m->owner[f])
        throw Error(UFVM_ERR_ARG, "own_nnz / nei_nnz inconsistent with cconn");
    }
    // faces of each cell, ascending (interior ones first by construction)
    std::vector<int32_t> cf_ptr(C.N + 1, 0);
    for (int32_t f = 0; f < C.F; ++f) {
      ++cf_ptr[m->owner[f] + 1];
      if (f < C.Fi) ++cf_ptr[m->neighbour[f] + 1];
    }
    for (int32_t i = 0; i < C.N; ++i) cf_ptr[i + 1] += cf_ptr[i];
    std::vector<int32_t> cf_face(cf_ptr[C.N]), fillp(cf_ptr.begin(), cf_ptr.end() - 1);
    std::vector<int32_t> cf_slot(cf_ptr[C.N], 0);
    for (int32_t f = 0; f < C.F; ++f) {
      if (f < C.Fi) cf_slot[fillp[m->owner[f]]] = m->own_nnz[f];
      cf_face[fillp[m->owner[f]]++] = f;
      if (f < C.Fi) {
        cf_slot[fillp[m->neighbour[f]]] = -1 - m->nei_nnz[f];
        cf_face[fillp[m->neighbour[f]]++] = f;
      }
    }
    const size_t F3 = (size_t)C.F * 3;
    up(C.owner, m->owner, C.F, s, "owner");
    up(C.neighbour, m->neighbour, std::max(C.Fi, 1), s, "neighbour");
    up(C.Sf, m->Sf, F3, s, "Sf");
    up(C.area, m->area, C.F, s, "area");
    up(C.CN, m->CN, F3, s, "CN");
    up(C.eCN, m->eCN, F3, s, "eCN");
    up(C.gDiff, m->gDiff, C.F, s, "gDiff");
    up(C.T, m->T, F3, s, "T");
    up(C.gf, m->gf, C.F, s, "gf");
    up(C.fc, m->face_centroid, F3, s, "face_centroid");
    up(C.vol, m->volume, C.N, s, "volume");
    up(C.cc, m->centroid, (size_t)C.N * 3, s, "centroid");
    up(C.cptr, m->cconn_ptr, C.N + 1, s, "cconn_ptr");
    up(C.ccol, m->cconn, std::max(nnz, 1), s, "cconn");
    up(C.own_nnz, m->own_nnz, std::max(C.Fi, 1), s, "own_nnz");
    up(C.nei_nnz, m->nei_nnz, std::max(C.Fi, 1), s, "nei_nnz");
    C.cf_ptr.upload(cf_ptr, s);
    C.cf_face.upload(cf_face, s);
    C.cf_slot.upload(cf_slot, s);
    C.uface.alloc((size_t)C.F * 6);
    if (C.cfg.stress_term != 0 && C.cfg.stress_term != 1)
      throw Error(UFVM_ERR_ARG, "stress_term must be 0 or 1");
    if (C.cfg.stress_term) C.ustress.alloc(C.F);
    C.p_changed.upload(std::vector<int32_t>{1}, s);
    C.bpatch.upload(bp.empty() ? std::vector<int32_t>{0} : bp, s);
    C.h_ptype.assign(m->patch_type, m->patch_type + C.P);
    C.h_u_bc.assign(cfg->u_bc, cfg->u_bc + C.P);
    C.h_p_bc.assign(cfg->p_bc, cfg->p_bc + C.P);
    C.h_u_val.assign(cfg->u_value, cfg->u_value + 3 * C.P);
    C.h_p_val.assign(C.P, 0.0);
    if (cfg->p_value) C.h_p_val.assign(cfg->p_value, cfg->p_value + C.P);
    C.ptype.upload(C.h_ptype.empty() ? std::vector<int32_t>{0} : C.h_ptype, s);
    C.u_bc.upload(C.h_u_bc.empty() ? std::vector<int32_t>{0} : C.h_u_bc, s);
    C.p_bc.upload(C.h_p_bc.empty() ? std::vector<int32_t>{0} : C.h_p_bc, s);
    C.u_val.upload(C.h_u_val.empty() ? std::vector<double>{0.0} : C.h_u_val, s);
    C.p_val.upload(C.h_p_val.empty() ? std::vector<double>{0.0} : C.h_p_val, s);
    {
      // boundary-face walldist = (CN . Sf) / |Sf| (cfdProcessOpenFoamMesh.m:153)
      // and, per patch and component, the sum in face order of (Sf_c/area)^2
      // (the noSlip wall's patch-wide dot(nx',nx'))
      std::vector<double> wd(C.F, 0.0), ns(std::max(3 * C.P, 1), 0.0);
      for (int32_t f = C.Fi; f < C.F; ++f) {
        const double *S = &m->Sf[(size_t)f * 3], *D = &m->CN[(size_t)f * 3];
        const double dot = (D[0] * S[0] + D[1] * S[1]) + D[2] * S[2];
        wd[f] = dot / std::sqrt((S[0] * S[0] + S[1] * S[1]) + S[2] * S[2]);
      }
      for (int p = 0; p < C.P; ++p)
        for (int32_t f = m->patch_start[p]; f < m->patch_start[p] + m->patch_nfaces[p]; ++f)
          for (int d = 0; d < 3; ++d) {
            const double nc = m->Sf[(size_t)f * 3 + d] / m->area[f];
            ns[p * 3 + d] = ns[p * 3 + d] + nc * nc;
          }
      C.walldist.upload(wd, s);
      C.noslip_sum.upload(ns, s);
    }
    C.cfg.u_bc = nullptr;
    C.cfg.u_value = nullptr;
    C.cfg.p_bc = nullptr;
    C.cfg.p_value = nullptr;
    if (C.cfg.p_ref_cell >= C.N) throw Error(UFVM_ERR_ARG, "p_ref_cell out of range");
    sk::DevMesh &d = C.dm;
    d.N = C.N; d.F = C.F; d.Fi = C.Fi;
    d.owner = C.owner.p; d.neighbour = C.neighbour.p;
    d.Sf = C.Sf.p; d.area = C.area.p; d.CN = C.CN.p; d.eCN = C.eCN.p; d.gDiff = C.gDiff.p;
    d.T = C.T.p; d.gf = C.gf.p; d.fc = C.fc.p; d.vol = C.vol.p; d.cc = C.cc.p;
    d.cptr = C.cptr.p; d.ccol = C.ccol.p; d.own_nnz = C.own_nnz.p; d.nei_nnz = C.nei_nnz.p;
    d.cf_ptr = C.cf_ptr.p; d.cf_face = C.cf_face.p; d.cf_slot = C.cf_slot.p;
    d.bpatch = C.bpatch.p;
    d.ptype = C.ptype.p; d.u_bc = C.u_bc.p; d.p_bc = C.p_bc.p; d.u_val = C.u_val.p;
    d.p_val = C.p_val.p;
    d.walldist = C.walldist.p; d.noslip_sum = C.noslip_sum.p;
    init_fields(ctx, C);
    C.sysU = create_system(ctx, C.N, nnz, m->cconn_ptr, m->cconn);
    C.sysP = create_system(ctx, C.N, nnz, m->cconn_ptr, m->cconn);
    UFVM_HIP(hipStreamSynchronize(s));
    const int32_t id = ctx->next_id++;
    ctx->cases[id] = std::shared_ptr<void>(Cp.release(), [](void *p) { delete static_cast<Case *>(p); });
    *case_out = id;
  });
}

int ufvm_case_destroy(ufvm_ctx *ctx, int32_t cas) {
  return guard(ctx, [&] {
    Case &C = case_of(ctx, cas);
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
    ctx->systems.erase(C.sysU);
    ctx->systems.erase(C.sysP);
    ctx->cases.erase(cas);
  });
}

int ufvm_case_set_field(ufvm_ctx *ctx, int32_t cas, const char *name, const double *data,
                        int64_t count) {
  return guard(ctx, [&] {
    FieldRef f = field_of(case_of(ctx, cas), name);
    if (!data || count != (int64_t)f.n) throw Error(UFVM_ERR_ARG, "field size mismatch");
    UFVM_HIP(hipMemcpyAsync(f.p, data, f.n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int ufvm_case_get_field(ufvm_ctx *ctx, int32_t cas, const char *name, double *data,
                        int64_t count) {
  return guard(ctx, [&] {
    FieldRef f = field_of(case_of(ctx, cas), name);
    if (!data || count != (int64_t)f.n) throw Error(UFVM_ERR_ARG, "field size mismatch");
    UFVM_HIP(hipMemcpyAsync(data, f.p, f.n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int ufvm_case_system(ufvm_ctx *ctx, int32_t cas, const char *eq, int32_t *sys_out) {
  return guard(ctx, [&] {
    Case &C = case_of(ctx, cas);
    if (!sys_out) throw Error(UFVM_ERR_ARG, "null sys_out");
    *sys_out = eq_of(eq) == 0 ? C.sysU : C.sysP;
  });
}

int ufvm_update_fields(ufvm_ctx *ctx, int32_t cas) {
  return guard(ctx, [&] {
    update_fields(ctx, case_of(ctx, cas));
    UFVM_HIP(hipStreamSynchronize(ctx->stream));
  });
}

int ufvm_assemble_equation(ufvm_ctx *ctx, int32_t cas, const char *eq, int32_t comp,
                           double *rms, double *max) {
  return guard(ctx, [&] { assemble(ctx, case_of(ctx, cas), eq_of(eq), comp, rms, max); });
}

int ufvm_solve_equation(ufvm_ctx *ctx, int32_t cas, const char *eq, double *r0, double *r1,
                        int32_t *ncycles) {
  return guard(ctx, [&] { solve(ctx, case_of(ctx, cas), eq_of(eq), r0, r1, ncycles); });
}

int ufvm_correct_equation(ufvm_ctx *ctx, int32_t cas, const char *eq, int32_t comp) {
  return guard(ctx, [&] { correct(ctx, case_of(ctx, cas), eq_of(eq), comp); });
}

int ufvm_simple_iterate(ufvm_ctx *ctx, int32_t cas, int32_t n, double *rms) {
  return guard(ctx, [&] {
    Case &C = case_of(ctx, cas);
    for (int32_t it = 0; it < n; ++it) {
      double r[4];
      outer_iteration(ctx, C, r);
      if (rms)
        for (int q = 0; q < 4; ++q) rms[it * 4 + q] = r[q];
    }
  });
}

// cfdTransientUpdate: the previous step's fields ('Step1')
static void transient_update(ufvm_ctx *ctx, Case &C) {
  UFVM_HIP(hipMemcpyAsync(C.U_old.p, C.U.p, (size_t)(C.N + C.Fb) * 3 * sizeof(double),
                          hipMemcpyDeviceToDevice, ctx->stream));
  UFVM_HIP(hipMemcpyAsync(C.mdot_old.p, C.mdot.p, (size_t)C.F * sizeof(double),
                          hipMemcpyDeviceToDevice, ctx->stream));
}

int ufvm_transient_update(ufvm_ctx *ctx, int32_t cas) {
  return guard(ctx, [&] {
    Case &C = case_of(ctx, cas);
    if (!C.cfg.transient) throw Error(UFVM_ERR_STATE, "the case was not created transient");
    transient_update(ctx, C);
  });
}

int ufvm_transient_step(ufvm_ctx *ctx, int32_t cas, int32_t max_inner, double u_tol,
                        double p_tol, double *rms, int32_t *n_inner) {
  return guard(ctx, [&] {
    Case &C = case_of(ctx, cas);
    if (!C.cfg.transient) throw Error(UFVM_ERR_STATE, "the case was not created transient");
    if (max_inner < 1) throw Error(UFVM_ERR_ARG, "max_inner must be >= 1");
    transient_update(ctx, C);
    int32_t it = 0;
    while (it < max_inner) {
      double r[4];
      outer_iteration(ctx, C, r);
      if (rms)
        for (int q = 0; q < 4; ++q) rms[it * 4 + q] = r[q];
      ++it;
      // cfdCheckConvergence.m: converged when no RMS residual exceeds its tolerance
      if (r[0] <= u_tol && r[1] <= u_tol && r[2] <= u_tol && r[3] <= p_tol) break;
    }
    if (n_inner) *n_inner = it;
  });
}

}  // extern "C"
