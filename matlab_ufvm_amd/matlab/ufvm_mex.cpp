// ufvm_mex.cpp -- the MEX gateway between uFVM (MATLAB) and libufvm.
//
//   mex -R2018a ufvm_mex.cpp mex_core.cpp -I../../include -L.. -lufvm
//
// One entry, dispatching on its first argument (a command name).  The .m
// shims in shims/ call it with the names and signatures of the reference's
// functions; ufvmSetup.m creates the case from `global Domain`.  Indices
// crossing this boundary are MATLAB's (1-based cells, faces, levels and
// components); mex_core converts them.  All device state lives in one
// library context per MATLAB session (device 0 unless 'init' names another),
// released by 'release' or when MATLAB clears the MEX file.
//
// Commands (outputs on the left):
//   ufvm_mex('init' [, device])
//   ufvm_mex('release')
//   id            = ufvm_mex('case_create', meshStruct, optsStruct)
//                   ufvm_mex('case_destroy', id)
//                   ufvm_mex('update_fields', id)                        cfdUpdateFields
//   [rms, mx]     = ufvm_mex('assemble', id, eq, iComponent)             cfdAssembleEquation
//   [r0, r1, nc]  = ufvm_mex('solve', id, eq)                            cfdSolveEquation
//                   ufvm_mex('correct', id, eq, iComponent)              cfdCorrectEquation
//   rms           = ufvm_mex('iterate', id, n)           (4 x n)         cfdRunCase, steady
//                   ufvm_mex('transient_update', id)                     cfdTransientUpdate
//   [rms, k]      = ufvm_mex('transient_step', id, maxInner, uTol, pTol) cfdRunCase, transient
//   phi           = ufvm_mex('get_field', id, name)                      cfdGetMeshField(...).phi
//                   ufvm_mex('set_field', id, name, phi)                 cfdSetMeshField
//   sys           = ufvm_mex('case_system', id, eq)  (Domain.coefficients{1} of the case)
//   sys           = ufvm_mex('system_create', csize, cconn)              cfdSetupCoefficients
//                   ufvm_mex('system_destroy', sys)
//                   ufvm_mex('set_coefficients', sys, ac, anb, bc)       cfdSetCoefficients
//   c             = ufvm_mex('get_coefficients', sys, level)             cfdGetCoefficients
//                   ufvm_mex('set_dphi', sys, level, dphi)
//   dphi          = ufvm_mex('get_dphi', sys, level)
//   r             = ufvm_mex('residuals', sys, level)                    cfdComputeResidualsArray
//   [r0, r1]      = ufvm_mex('solve_algebraic_system', sys, level, smoother, nIter, rrf)
//   nLevels       = ufvm_mex('agglomerate', sys, maxCoarseLevels)        cfdAgglomerate
//   nLevels       = ufvm_mex('num_levels', sys)
//   [r0, r1, nc]  = ufvm_mex('apply_amg', sys, multigrid, smoother)      cfdApplyAMG
//                   ufvm_mex('pcg_setup', sys, optsStruct)
//   [it, relres]  = ufvm_mex('pcg_solve', sys)
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "mex.h"
#include "mex_core.h"
#include "ufvm.h"

using ufvm_mex::Arr;
using ufvm_mex::Error;

namespace {

ufvm_ctx *g_ctx = nullptr;

struct CaseDims {
  int64_t N = 0, Fb = 0, F = 0;
};
std::vector<CaseDims> g_cases;  // by case id

void release_all() {
  if (g_ctx) ufvm_ctx_destroy(g_ctx);
  g_ctx = nullptr;
  g_cases.clear();
}

void check(int rc) {
  if (rc != UFVM_OK) throw Error{std::string("libufvm: ") + ufvm_last_error(g_ctx)};
}

ufvm_ctx *ctx() {
  if (!g_ctx) {
    check(ufvm_ctx_create(0, &g_ctx));
    mexLock();
    mexAtExit(release_all);
  }
  return g_ctx;
}

// mxArray -> Arr (numeric arrays alias MATLAB's memory for the call)
Arr to_arr(const mxArray *a, const char *what) {
  Arr r;
  if (!a) throw Error{std::string("missing ") + what};
  if (mxIsDouble(a) && !mxIsComplex(a)) {
    r.kind = Arr::F64;
    r.data = mxGetDoubles(a);
  } else if (mxIsInt32(a)) {
    r.kind = Arr::I32;
    r.data = mxGetInt32s(a);
  } else if (mxIsChar(a)) {
    char *s = mxArrayToString(a);
    r.kind = Arr::TEXT;
    r.text = s ? s : "";
    mxFree(s);
    return r;
  } else if (mxIsCell(a)) {
    r.kind = Arr::TEXTS;
    const size_t n = mxGetNumberOfElements(a);
    for (size_t i = 0; i < n; ++i) {
      const mxArray *e = mxGetCell(a, i);
      if (!e || !mxIsChar(e)) throw Error{std::string(what) + ": cells must hold char arrays"};
      char *s = mxArrayToString(e);
      r.texts.emplace_back(s ? s : "");
      mxFree(s);
    }
    return r;
  } else {
    throw Error{std::string(what) + ": expected double, int32, char or cell"};
  }
  r.rows = (int64_t)mxGetM(a);
  r.cols = (int64_t)mxGetNumberOfElements(a) / (r.rows ? r.rows : 1);
  if (r.rows == 0) r.cols = 0;
  return r;
}

class MxSource : public ufvm_mex::Source {
 public:
  explicit MxSource(const mxArray *s) : s_(s) {
    if (!s || !mxIsStruct(s)) throw Error{"expected a struct"};
  }
  bool get(const std::string &name, Arr &out) const override {
    const mxArray *f = mxGetField(s_, 0, name.c_str());
    if (!f) return false;
    out = to_arr(f, name.c_str());
    return true;
  }

 private:
  const mxArray *s_;
};

std::string str(const mxArray *a, const char *what) {
  Arr r = to_arr(a, what);
  if (r.kind != Arr::TEXT) throw Error{std::string(what) + " must be a char row vector"};
  return r.text;
}

double num(const mxArray *a, const char *what) {
  if (!a || !mxIsNumeric(a) || mxGetNumberOfElements(a) != 1)
    throw Error{std::string(what) + " must be a numeric scalar"};
  return mxGetScalar(a);
}

int32_t inum(const mxArray *a, const char *what) { return (int32_t)num(a, what); }

std::vector<double> dvec(const mxArray *a, int64_t n, const char *what) {
  if (!a || !mxIsDouble(a) || mxIsComplex(a) || (int64_t)mxGetNumberOfElements(a) != n)
    throw Error{std::string(what) + " must be a real double array of " + std::to_string(n) +
                " values"};
  const double *p = mxGetDoubles(a);
  return std::vector<double>(p, p + n);
}

mxArray *col(const double *v, int64_t n) {
  mxArray *a = mxCreateDoubleMatrix((mwSize)n, 1, mxREAL);
  if (n) std::memcpy(mxGetDoubles(a), v, (size_t)n * sizeof(double));
  return a;
}

const CaseDims &dims(int32_t id) {
  if (id < 0 || id >= (int32_t)g_cases.size() || g_cases[id].N == 0)
    throw Error{"unknown case id " + std::to_string(id)};
  return g_cases[id];
}

int32_t level_of(const mxArray *a) {
  const int32_t l = inum(a, "gridLevel");
  if (l < 1) throw Error{"gridLevel is 1-based"};
  return l - 1;
}

int32_t level_size(int32_t sys, int32_t level, int32_t *nnz = nullptr) {
  int32_t n = 0, z = 0;
  check(ufvm_amg_level_info(ctx(), sys, level, &n, &z));
  if (nnz) *nnz = z;
  return n;
}

// Domain.coefficients{level} as the reference's struct (ac, anb, bc, dphi,
// cconn, csize, numberOfElements, and parents below the coarsest level)
mxArray *coefficients(int32_t sys, int32_t level) {
  int32_t nnz = 0;
  const int32_t n = level_size(sys, level, &nnz);
  int32_t nlev = 0;
  check(ufvm_amg_num_levels(ctx(), sys, &nlev));
  std::vector<int32_t> parents(n), ptr(n + 1), colv(std::max(nnz, 1));
  std::vector<double> ac(n), anb(std::max(nnz, 1)), bc(n), dphi(n);
  check(ufvm_amg_level_get(ctx(), sys, level, level + 1 < nlev ? parents.data() : nullptr,
                           ptr.data(), colv.data(), ac.data(), anb.data()));
  check(ufvm_system_get_bc(ctx(), sys, level, bc.data()));
  check(ufvm_system_get_dphi(ctx(), sys, level, dphi.data()));
  const char *names[] = {"ac", "anb", "bc", "dphi", "cconn", "csize", "numberOfElements",
                         "parents"};
  mxArray *s = mxCreateStructMatrix(1, 1, level + 1 < nlev ? 8 : 7, names);
  mxArray *canb = mxCreateCellMatrix((mwSize)n, 1), *ccon = mxCreateCellMatrix((mwSize)n, 1);
  mxArray *csize = mxCreateDoubleMatrix((mwSize)n, 1, mxREAL);
  for (int32_t i = 0; i < n; ++i) {
    const int32_t k0 = ptr[i], m = ptr[i + 1] - ptr[i];
    mxArray *a = mxCreateDoubleMatrix(1, (mwSize)m, mxREAL), *c = mxCreateDoubleMatrix(1, (mwSize)m, mxREAL);
    for (int32_t k = 0; k < m; ++k) {
      mxGetDoubles(a)[k] = anb[k0 + k];
      mxGetDoubles(c)[k] = colv[k0 + k] + 1.0;
    }
    mxSetCell(canb, i, a);
    mxSetCell(ccon, i, c);
    mxGetDoubles(csize)[i] = m;
  }
  mxSetField(s, 0, "ac", col(ac.data(), n));
  mxSetField(s, 0, "anb", canb);
  mxSetField(s, 0, "bc", col(bc.data(), n));
  mxSetField(s, 0, "dphi", col(dphi.data(), n));
  mxSetField(s, 0, "cconn", ccon);
  mxSetField(s, 0, "csize", csize);
  mxSetField(s, 0, "numberOfElements", mxCreateDoubleScalar(n));
  if (level + 1 < nlev) {
    mxArray *p = mxCreateDoubleMatrix((mwSize)n, 1, mxREAL);
    for (int32_t i = 0; i < n; ++i) mxGetDoubles(p)[i] = parents[i] + 1.0;
    mxSetField(s, 0, "parents", p);
  }
  return s;
}

void need_args(int nrhs, int k, const std::string &cmd) {
  if (nrhs < k) throw Error{"ufvm_mex('" + cmd + "'): too few arguments"};
}

void dispatch(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  if (nrhs < 1) throw Error{"usage: ufvm_mex(command, ...)"};
  const std::string cmd = str(prhs[0], "command");
  if (cmd == "init") {
    if (!g_ctx) {
      const int32_t dev = nrhs > 1 ? inum(prhs[1], "device") : 0;
      check(ufvm_ctx_create(dev, &g_ctx));
      mexLock();
      mexAtExit(release_all);
    }
  } else if (cmd == "release") {
    if (g_ctx) {
      release_all();
      mexUnlock();
    }
  } else if (cmd == "case_create") {
    need_args(nrhs, 3, cmd);
    MxSource mesh(prhs[1]), opts(prhs[2]);
    ufvm_mex::CaseInput in;
    ufvm_mex::build_case(mesh, opts, in);
    int32_t id = 0;
    check(ufvm_case_create(ctx(), &in.mesh, &in.cfg, &id));
    if ((int32_t)g_cases.size() <= id) g_cases.resize(id + 1);
    g_cases[id] = {in.mesh.n_cells, in.mesh.n_faces - in.mesh.n_interior, in.mesh.n_faces};
    plhs[0] = mxCreateDoubleScalar(id);
  } else if (cmd == "case_destroy") {
    need_args(nrhs, 2, cmd);
    const int32_t id = inum(prhs[1], "case");
    dims(id);
    check(ufvm_case_destroy(ctx(), id));
    g_cases[id] = CaseDims{};
  } else if (cmd == "update_fields") {
    need_args(nrhs, 2, cmd);
    check(ufvm_update_fields(ctx(), inum(prhs[1], "case")));
  } else if (cmd == "assemble") {
    need_args(nrhs, 4, cmd);
    double rms = 0, mx = 0;
    check(ufvm_assemble_equation(ctx(), inum(prhs[1], "case"), str(prhs[2], "eq").c_str(),
                                 inum(prhs[3], "iComponent") - 1, &rms, &mx));
    plhs[0] = mxCreateDoubleScalar(rms);
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(mx);
  } else if (cmd == "solve") {
    need_args(nrhs, 3, cmd);
    double r0 = 0, r1 = 0;
    int32_t nc = 0;
    check(ufvm_solve_equation(ctx(), inum(prhs[1], "case"), str(prhs[2], "eq").c_str(), &r0, &r1,
                              &nc));
    plhs[0] = mxCreateDoubleScalar(r0);
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(r1);
    if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(nc);
  } else if (cmd == "correct") {
    need_args(nrhs, 4, cmd);
    check(ufvm_correct_equation(ctx(), inum(prhs[1], "case"), str(prhs[2], "eq").c_str(),
                                inum(prhs[3], "iComponent") - 1));
  } else if (cmd == "iterate") {
    need_args(nrhs, 3, cmd);
    const int32_t n = inum(prhs[2], "n");
    if (n < 0) throw Error{"n must be >= 0"};
    mxArray *out = mxCreateDoubleMatrix(4, (mwSize)n, mxREAL);
    const int rc = ufvm_simple_iterate(ctx(), inum(prhs[1], "case"), n, mxGetDoubles(out));
    if (rc != UFVM_OK) mxDestroyArray(out);
    check(rc);
    plhs[0] = out;
  } else if (cmd == "transient_update") {
    need_args(nrhs, 2, cmd);
    check(ufvm_transient_update(ctx(), inum(prhs[1], "case")));
  } else if (cmd == "transient_step") {
    need_args(nrhs, 5, cmd);
    const int32_t k = inum(prhs[2], "maxInner");
    if (k < 1) throw Error{"maxInner must be >= 1"};
    std::vector<double> rms((size_t)k * 4);
    int32_t it = 0;
    check(ufvm_transient_step(ctx(), inum(prhs[1], "case"), k, num(prhs[3], "uTol"),
                              num(prhs[4], "pTol"), rms.data(), &it));
    mxArray *out = mxCreateDoubleMatrix(4, (mwSize)it, mxREAL);
    if (it) std::memcpy(mxGetDoubles(out), rms.data(), (size_t)it * 4 * sizeof(double));
    plhs[0] = out;
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(it);
  } else if (cmd == "get_field" || cmd == "set_field") {
    need_args(nrhs, cmd == "get_field" ? 3 : 4, cmd);
    const int32_t id = inum(prhs[1], "case");
    const CaseDims &d = dims(id);
    const std::string name = str(prhs[2], "field");
    const ufvm_mex::FieldShape s = ufvm_mex::field_shape(name, d.N, d.Fb, d.F);
    std::vector<double> c((size_t)s.numel());
    if (cmd == "get_field") {
      check(ufvm_case_get_field(ctx(), id, name.c_str(), c.data(), s.numel()));
      mxArray *out;
      if (s.comps > 1 && s.dirs > 1) {
        const mwSize dv[3] = {(mwSize)s.rows, (mwSize)s.dirs, (mwSize)s.comps};
        out = mxCreateNumericArray(3, dv, mxDOUBLE_CLASS, mxREAL);
      } else {
        out = mxCreateDoubleMatrix((mwSize)s.rows, (mwSize)(s.comps * s.dirs), mxREAL);
      }
      ufvm_mex::to_matlab(s, c.data(), mxGetDoubles(out));
      plhs[0] = out;
    } else {
      const std::vector<double> m = dvec(prhs[3], s.numel(), "phi");
      ufvm_mex::from_matlab(s, m.data(), c.data());
      check(ufvm_case_set_field(ctx(), id, name.c_str(), c.data(), s.numel()));
    }
  } else if (cmd == "case_system") {
    need_args(nrhs, 3, cmd);
    int32_t sys = 0;
    check(ufvm_case_system(ctx(), inum(prhs[1], "case"), str(prhs[2], "eq").c_str(), &sys));
    plhs[0] = mxCreateDoubleScalar(sys);
  } else if (cmd == "system_create") {
    need_args(nrhs, 3, cmd);
    ufvm_mex::Pattern p;
    ufvm_mex::build_pattern(to_arr(prhs[1], "csize"), to_arr(prhs[2], "cconn"), p);
    int32_t sys = 0;
    check(ufvm_system_create(ctx(), (int32_t)p.ptr.size() - 1, p.ptr.back(), p.ptr.data(),
                             p.col.data(), &sys));
    plhs[0] = mxCreateDoubleScalar(sys);
  } else if (cmd == "system_destroy") {
    need_args(nrhs, 2, cmd);
    check(ufvm_system_destroy(ctx(), inum(prhs[1], "sys")));
  } else if (cmd == "set_coefficients") {
    need_args(nrhs, 5, cmd);
    const int32_t sys = inum(prhs[1], "sys");
    int32_t nnz = 0;
    const int32_t n = level_size(sys, 0, &nnz);
    const std::vector<double> ac = dvec(prhs[2], n, "ac"), anb = dvec(prhs[3], nnz, "anb"),
                              bc = dvec(prhs[4], n, "bc");
    check(ufvm_system_set_coefficients(ctx(), sys, ac.data(), nnz ? anb.data() : nullptr,
                                       bc.data()));
  } else if (cmd == "get_coefficients") {
    need_args(nrhs, 3, cmd);
    plhs[0] = coefficients(inum(prhs[1], "sys"), level_of(prhs[2]));
  } else if (cmd == "set_dphi" || cmd == "get_dphi") {
    need_args(nrhs, cmd == "set_dphi" ? 4 : 3, cmd);
    const int32_t sys = inum(prhs[1], "sys"), level = level_of(prhs[2]);
    const int32_t n = level_size(sys, level);
    if (cmd == "set_dphi") {
      const std::vector<double> v = dvec(prhs[3], n, "dphi");
      check(ufvm_system_set_dphi(ctx(), sys, level, v.data()));
    } else {
      std::vector<double> v(n);
      check(ufvm_system_get_dphi(ctx(), sys, level, v.data()));
      plhs[0] = col(v.data(), n);
    }
  } else if (cmd == "residuals") {
    need_args(nrhs, 3, cmd);
    const int32_t sys = inum(prhs[1], "sys"), level = level_of(prhs[2]);
    const int32_t n = level_size(sys, level);
    std::vector<double> r(n);
    double l1 = 0;
    check(ufvm_compute_residuals(ctx(), sys, level, r.data(), &l1));
    plhs[0] = col(r.data(), n);
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(l1);
  } else if (cmd == "solve_algebraic_system") {
    need_args(nrhs, 5, cmd);
    const double rrf = nrhs > 5 ? num(prhs[5], "rrf") : 0.1;
    double r0 = 0, r1 = 0;
    check(ufvm_solve_algebraic_system(ctx(), inum(prhs[1], "sys"), level_of(prhs[2]),
                                      ufvm_mex::smoother_code(str(prhs[3], "smootherType")),
                                      inum(prhs[4], "numberOfIterations"), rrf, &r0, &r1));
    plhs[0] = mxCreateDoubleScalar(r0);
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(r1);
  } else if (cmd == "agglomerate") {
    need_args(nrhs, 3, cmd);
    int32_t nl = 0;
    check(ufvm_amg_agglomerate(ctx(), inum(prhs[1], "sys"), inum(prhs[2], "maxCoarseLevels"),
                               &nl));
    plhs[0] = mxCreateDoubleScalar(nl);
  } else if (cmd == "num_levels") {
    need_args(nrhs, 2, cmd);
    int32_t nl = 0;
    check(ufvm_amg_num_levels(ctx(), inum(prhs[1], "sys"), &nl));
    plhs[0] = mxCreateDoubleScalar(nl);
  } else if (cmd == "apply_amg") {
    need_args(nrhs, 4, cmd);
    MxSource mg(prhs[2]);
    const ufvm_amg_cfg cfg = ufvm_mex::amg_cfg(mg, str(prhs[3], "smootherType"));
    double r0 = 0, r1 = 0;
    int32_t nc = 0;
    check(ufvm_apply_amg(ctx(), inum(prhs[1], "sys"), &cfg, &r0, &r1, &nc));
    plhs[0] = mxCreateDoubleScalar(r0);
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(r1);
    if (nlhs > 2) plhs[2] = mxCreateDoubleScalar(nc);
  } else if (cmd == "pcg_setup") {
    need_args(nrhs, 3, cmd);
    MxSource o(prhs[2]);
    const ufvm_pcg_cfg q = ufvm_mex::pcg_cfg(o);
    check(ufvm_pcg_setup(ctx(), inum(prhs[1], "sys"), &q));
  } else if (cmd == "pcg_solve") {
    need_args(nrhs, 2, cmd);
    int32_t it = 0;
    double rr = 0;
    check(ufvm_pcg_solve(ctx(), inum(prhs[1], "sys"), &it, &rr));
    plhs[0] = mxCreateDoubleScalar(it);
    if (nlhs > 1) plhs[1] = mxCreateDoubleScalar(rr);
  } else {
    throw Error{"unknown command '" + cmd + "'"};
  }
}

char g_msg[2048];

}  // namespace

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[]) {
  bool failed = false;
  {
    // every C++ object is destroyed before mexErrMsgIdAndTxt leaves the call
    try {
      dispatch(nlhs, plhs, nrhs, prhs);
    } catch (const Error &e) {
      std::snprintf(g_msg, sizeof g_msg, "%s", e.msg.c_str());
      failed = true;
    } catch (const std::exception &e) {
      std::snprintf(g_msg, sizeof g_msg, "%s", e.what());
      failed = true;
    }
  }
  if (failed) mexErrMsgIdAndTxt("ufvm:error", "%s", g_msg);
}
