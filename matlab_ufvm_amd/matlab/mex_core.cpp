// mex_core.cpp -- see mex_core.h.
#include "mex_core.h"

#include <cmath>
#include <cstdlib>
#include <cstring>

namespace ufvm_mex {

bool MapSource::get(const std::string &name, Arr &out) const {
  auto it = fields.find(name);
  if (it == fields.end()) return false;
  out = it->second;
  return true;
}

namespace {

[[noreturn]] void fail(const std::string &msg) { throw Error{msg}; }

Arr need(const Source &s, const char *name) {
  Arr a;
  if (!s.get(name, a)) fail(std::string("missing field '") + name + "'");
  return a;
}

double element(const Arr &a, int64_t i) {
  return a.kind == Arr::F64 ? static_cast<const double *>(a.data)[i]
                            : (double)static_cast<const int32_t *>(a.data)[i];
}

void numeric(const Arr &a, const char *name) {
  if (a.kind != Arr::F64 && a.kind != Arr::I32)
    fail(std::string("'") + name + "' must be a double or int32 array");
}

// n values (any shape with n elements)
std::vector<double> vec(const Arr &a, int64_t n, const char *name) {
  numeric(a, name);
  if (a.numel() != n)
    fail(std::string("'") + name + "' has " + std::to_string(a.numel()) + " values, expected " +
         std::to_string(n));
  std::vector<double> v(n);
  for (int64_t i = 0; i < n; ++i) v[i] = element(a, i);
  return v;
}

// n integers, minus `base` (1 for the reference's 1-based indices), each in [lo, hi)
std::vector<int32_t> ivec(const Arr &a, int64_t n, const char *name, int base, int64_t lo,
                          int64_t hi) {
  numeric(a, name);
  if (a.numel() != n)
    fail(std::string("'") + name + "' has " + std::to_string(a.numel()) + " values, expected " +
         std::to_string(n));
  std::vector<int32_t> v(n);
  for (int64_t i = 0; i < n; ++i) {
    const double x = element(a, i);
    if (x != std::floor(x)) fail(std::string("'") + name + "' must hold integers");
    const int64_t k = (int64_t)x - base;
    if (k < lo || k >= hi)
      fail(std::string("'") + name + "' value " + std::to_string((int64_t)x) + " out of range");
    v[i] = (int32_t)k;
  }
  return v;
}

// an n x 3 MATLAB matrix (column-major) as row-major [n][3]
std::vector<double> rows3(const Arr &a, int64_t n, const char *name) {
  numeric(a, name);
  if (a.rows != n || a.cols != 3)
    fail(std::string("'") + name + "' must be " + std::to_string(n) + " x 3");
  std::vector<double> v(n * 3);
  for (int64_t i = 0; i < n; ++i)
    for (int d = 0; d < 3; ++d) v[i * 3 + d] = element(a, i + n * d);
  return v;
}

double scalar(const Source &s, const char *name, double def) {
  Arr a;
  if (!s.get(name, a)) return def;
  if (a.kind == Arr::TEXT) {
    char *end = nullptr;
    const double v = std::strtod(a.text.c_str(), &end);
    if (end == a.text.c_str()) fail(std::string("'") + name + "' is not a number");
    return v;
  }
  numeric(a, name);
  if (a.numel() != 1) fail(std::string("'") + name + "' must be a scalar");
  return element(a, 0);
}

int32_t iscalar(const Source &s, const char *name, int32_t def) {
  const double v = scalar(s, name, def);
  if (v != std::floor(v) || std::fabs(v) > 2147483647.0)
    fail(std::string("'") + name + "' must be an integer");
  return (int32_t)v;
}

std::string text(const Source &s, const char *name, const char *def) {
  Arr a;
  if (!s.get(name, a)) return def;
  if (a.kind != Arr::TEXT) fail(std::string("'") + name + "' must be a char row vector");
  return a.text;
}

std::vector<std::string> texts(const Source &s, const char *name, int64_t n) {
  Arr a = need(s, name);
  if (a.kind == Arr::TEXT && n == 1) return {a.text};
  if (a.kind != Arr::TEXTS) fail(std::string("'") + name + "' must be a cell array of char");
  if ((int64_t)a.texts.size() != n)
    fail(std::string("'") + name + "' must have " + std::to_string(n) + " entries");
  return a.texts;
}

}  // namespace

int32_t patch_type_code(const std::string &t) {
  if (t == "wall") return UFVM_PATCH_WALL;
  if (t == "empty") return UFVM_PATCH_EMPTY;
  if (t == "symmetry" || t == "symmetryPlane") return UFVM_PATCH_SYMMETRY;
  if (t == "inlet") return UFVM_PATCH_INLET;
  if (t == "outlet") return UFVM_PATCH_OUTLET;
  fail("patch type '" + t + "' is not supported on this path");
}

int32_t bc_code(const std::string &t) {
  if (t == "fixedValue") return UFVM_BC_FIXED_VALUE;
  if (t == "zeroGradient") return UFVM_BC_ZERO_GRADIENT;
  if (t == "empty") return UFVM_BC_EMPTY;
  if (t == "noSlip") return UFVM_BC_NO_SLIP;
  if (t == "symmetry" || t == "symmetryPlane") return UFVM_BC_SYMMETRY;
  if (t == "slip") return UFVM_BC_SLIP;
  fail("boundary condition '" + t + "' is not supported on this path");
}

int32_t cycle_code(const std::string &t) {
  if (t == "V-Cycle") return UFVM_CYCLE_V;
  if (t == "W-Cycle") return UFVM_CYCLE_W;
  if (t == "F-Cycle") return UFVM_CYCLE_F;
  fail("cycle type '" + t + "' is not one of V-Cycle, W-Cycle, F-Cycle");
}

int32_t smoother_code(const std::string &t) {
  if (t == "ILU") return UFVM_SMOOTHER_ILU;
  if (t == "SOR") return UFVM_SMOOTHER_SOR;
  if (t == "MCGS") return UFVM_SMOOTHER_MCGS;
  fail("smoother '" + t + "' is not one of ILU, SOR, MCGS");
}

int32_t pcg_smoother_code(const std::string &t) {
  if (t == "MCGS") return 0;
  if (t == "ILU") return 1;
  if (t == "MCILU") return 2;
  if (t == "MCILU0") return 3;
  fail("V-cycle smoother '" + t + "' is not one of MCGS, ILU, MCILU, MCILU0");
}

int32_t p_solver_code(const std::string &t) {
  if (t == "amg") return 0;
  if (t == "pcg") return 1;
  if (t == "pcg-frozen") return 2;
  fail("pressure solver '" + t + "' is not one of amg, pcg, pcg-frozen");
}

void build_pattern(const Arr &csize, const Arr &cconn, Pattern &out) {
  numeric(csize, "csize");
  const int64_t n = csize.numel();
  if (n < 1 || n > 2147483647) fail("'csize' must have between 1 and 2^31-1 rows");
  out.ptr.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    const double c = element(csize, i);
    if (c < 0 || c != std::floor(c)) fail("'csize' must hold counts");
    out.ptr[i + 1] = out.ptr[i] + (int32_t)c;
  }
  out.col = ivec(cconn, out.ptr[n], "cconn", 1, 0, n);
  for (int64_t i = 0; i < n; ++i)
    for (int32_t k = out.ptr[i]; k < out.ptr[i + 1]; ++k)
      if (out.col[k] == i) fail("'cconn' lists a row as its own neighbour");
}

void build_case(const Source &ms, const Source &os, CaseInput &c) {
  const int64_t N = iscalar(ms, "nElements", -1), F = iscalar(ms, "nFaces", -1),
                Fi = iscalar(ms, "nInteriorFaces", -1);
  if (N < 1 || F < 1 || Fi < 0 || Fi > F) fail("inconsistent mesh sizes");
  c.owner = ivec(need(ms, "iOwner"), F, "iOwner", 1, 0, N);
  c.neighbour = ivec(need(ms, "iNeighbour"), Fi, "iNeighbour", 1, 0, N);
  c.Sf = rows3(need(ms, "Sf"), F, "Sf");
  c.area = vec(need(ms, "area"), F, "area");
  c.CN = rows3(need(ms, "CN"), F, "CN");
  c.eCN = rows3(need(ms, "eCN"), F, "eCN");
  c.gDiff = vec(need(ms, "gDiff"), F, "gDiff");
  c.T = rows3(need(ms, "T"), F, "T");
  c.gf = vec(need(ms, "gf"), F, "gf");
  c.face_centroid = rows3(need(ms, "faceCentroid"), F, "faceCentroid");
  c.volume = vec(need(ms, "volume"), N, "volume");
  c.centroid = rows3(need(ms, "elementCentroid"), N, "elementCentroid");
  Pattern pat;
  build_pattern(need(ms, "csize"), need(ms, "cconn"), pat);
  if ((int64_t)pat.ptr.size() != N + 1) fail("'csize' must have nElements entries");
  c.cconn_ptr = pat.ptr;
  c.cconn = pat.col;
  // iOwnerNeighbourCoef / iNeighbourOwnerCoef: 1-based positions inside the
  // owner's / neighbour's cconn row -> CSR slots, checked against the rows
  const std::vector<int32_t> onc = ivec(need(ms, "iOwnerNeighbourCoef"), Fi,
                                        "iOwnerNeighbourCoef", 1, 0, 1 << 30);
  const std::vector<int32_t> noc = ivec(need(ms, "iNeighbourOwnerCoef"), Fi,
                                        "iNeighbourOwnerCoef", 1, 0, 1 << 30);
  c.own_nnz.resize(Fi);
  c.nei_nnz.resize(Fi);
  for (int64_t f = 0; f < Fi; ++f) {
    const int32_t o = c.owner[f], nb = c.neighbour[f];
    if (onc[f] >= c.cconn_ptr[o + 1] - c.cconn_ptr[o] ||
        noc[f] >= c.cconn_ptr[nb + 1] - c.cconn_ptr[nb])
      fail("face " + std::to_string(f + 1) + ": neighbour coefficient index past its cconn row");
    c.own_nnz[f] = c.cconn_ptr[o] + onc[f];
    c.nei_nnz[f] = c.cconn_ptr[nb] + noc[f];
    if (c.cconn[c.own_nnz[f]] != nb || c.cconn[c.nei_nnz[f]] != o)
      fail("face " + std::to_string(f + 1) + ": iOwnerNeighbourCoef / iNeighbourOwnerCoef "
           "do not point at the face's cells in cconn");
  }
  // patches (startFace is 1-based in theMesh.boundaries)
  Arr ps = need(ms, "patchStartFace");
  numeric(ps, "patchStartFace");
  const int64_t P = ps.numel();
  if (P < 1) fail("the mesh has no boundary patches");
  c.patch_start = ivec(ps, P, "patchStartFace", 1, Fi, F);
  c.patch_nfaces = ivec(need(ms, "patchNFaces"), P, "patchNFaces", 0, 0, F - Fi + 1);
  const std::vector<std::string> ptypes = texts(ms, "patchType", P);
  c.patch_type.resize(P);
  for (int64_t p = 0; p < P; ++p) c.patch_type[p] = patch_type_code(ptypes[p]);
  int64_t next = Fi;
  for (int64_t p = 0; p < P; ++p) {
    if (c.patch_start[p] != next) fail("patches must cover the boundary faces in order");
    next += c.patch_nfaces[p];
  }
  if (next != F) fail("patches must cover the boundary faces in order");
  ufvm_mesh_desc &d = c.mesh;
  d.n_cells = (int32_t)N;
  d.n_faces = (int32_t)F;
  d.n_interior = (int32_t)Fi;
  d.owner = c.owner.data();
  d.neighbour = c.neighbour.data();
  d.Sf = c.Sf.data();
  d.area = c.area.data();
  d.CN = c.CN.data();
  d.eCN = c.eCN.data();
  d.gDiff = c.gDiff.data();
  d.T = c.T.data();
  d.gf = c.gf.data();
  d.face_centroid = c.face_centroid.data();
  d.volume = c.volume.data();
  d.centroid = c.centroid.data();
  d.cconn_ptr = c.cconn_ptr.data();
  d.cconn = c.cconn.data();
  d.own_nnz = c.own_nnz.data();
  d.nei_nnz = c.nei_nnz.data();
  d.n_patches = (int32_t)P;
  d.patch_start = c.patch_start.data();
  d.patch_nfaces = c.patch_nfaces.data();
  d.patch_type = c.patch_type.data();

  // boundary conditions per patch (theEquation.bcs{iPatch}.type / .value)
  const std::vector<std::string> ub = texts(os, "uBC", P), pb = texts(os, "pBC", P);
  c.u_bc.resize(P);
  c.p_bc.resize(P);
  for (int64_t p = 0; p < P; ++p) {
    c.u_bc[p] = bc_code(ub[p]);
    c.p_bc[p] = bc_code(pb[p]);
  }
  Arr uv;
  c.u_value = os.get("uValue", uv) ? rows3(uv, P, "uValue") : std::vector<double>(P * 3, 0.0);
  Arr pv;
  c.p_value = os.get("pValue", pv) ? vec(pv, P, "pValue") : std::vector<double>(P, 0.0);

  ufvm_simple_cfg &g = c.cfg;
  g.nu = scalar(os, "nu", 0.0);
  g.urf_u = scalar(os, "urfU", 0.7);
  g.urf_p = scalar(os, "urfP", 0.3);
  g.n_correctors = iscalar(os, "nCorrectors", 2);
  // pRefCell as the reference resolves it (1-based cell; 0: none)
  const int32_t pref = iscalar(os, "pRefCell", 1);
  if (pref < 0 || pref > N) fail("'pRefCell' must be a 1-based cell index or 0");
  g.p_ref_cell = pref - 1;
  g.u_sweeps = iscalar(os, "uSweeps", 10);
  g.u_bc = c.u_bc.data();
  g.u_value = c.u_value.data();
  g.p_bc = c.p_bc.data();
  g.p_solver = p_solver_code(text(os, "pSolver", "amg"));
  g.p_amg = amg_cfg(os, text(os, "smootherType", "ILU"));
  g.p_pcg = pcg_cfg(os);
  g.u_smoother = smoother_code(text(os, "uSmoother", "ILU"));
  g.stress_term = iscalar(os, "stressTerm", 0) ? 1 : 0;
  g.p_value = c.p_value.data();
  Arr ui;
  if (os.get("uInit", ui)) {
    const std::vector<double> u = vec(ui, 3, "uInit");
    for (int k = 0; k < 3; ++k) g.u_init[k] = u[k];
  }
  g.p_init = scalar(os, "pInit", 0.0);
  g.rho = scalar(os, "rho", 1.0);
  g.transient = iscalar(os, "transient", 0) ? 1 : 0;
  g.dt = scalar(os, "dt", 1.0);
}

ufvm_amg_cfg amg_cfg(const Source &mg, const std::string &smoother) {
  ufvm_amg_cfg a{};
  a.cycle = cycle_code(text(mg, "cycleType", "W-Cycle"));
  a.max_cycles = iscalar(mg, "maxCycles", 30);
  a.termination = scalar(mg, "termination", 0.1);
  a.max_coarse_levels = iscalar(mg, "maxCoarseLevels", 10);
  a.pre_sweep = iscalar(mg, "preSweep", 1);
  a.post_sweep = iscalar(mg, "postSweep", 3);
  a.smoother = smoother_code(smoother);
  if (a.smoother == UFVM_SMOOTHER_MCGS) fail("cfdApplyAMG smooths with 'ILU' or 'SOR'");
  a.rebuild = 1;
  return a;
}

ufvm_pcg_cfg pcg_cfg(const Source &os) {
  ufvm_pcg_cfg q{};
  q.precision = iscalar(os, "pcgPrecision", 32);
  q.pre_sweep = q.post_sweep = iscalar(os, "pcgSweeps", 1);
  q.coarse_smoother = 1;
  q.max_coarse_cells = iscalar(os, "pcgMaxCoarseCells", 2000);
  q.nullspace = iscalar(os, "nullspace", -1);
  q.max_iter = iscalar(os, "pcgMaxIter", 500);
  q.tol = scalar(os, "pcgTol", 1e-8);
  q.smoother = pcg_smoother_code(text(os, "pcgSmoother", "MCGS"));
  q.coarse_scale = scalar(os, "pcgCoarseScale", 2.5);
  q.coarse_pre_sweep = q.coarse_post_sweep = iscalar(os, "pcgCoarseSweeps", 2);
  q.launch_mode = 0;
  q.index_width = 0;   // 16-bit column offsets where they fit
  return q;
}

FieldShape field_shape(const std::string &n, int64_t N, int64_t Fb, int64_t F) {
  FieldShape s;
  if (n == "U" || n == "DU" || n == "DUT" || n == "U_old") {
    s.rows = N + Fb;
    s.comps = 3;
  } else if (n == "p" || n == "PP") {
    s.rows = N + Fb;
  } else if (n == "gradp" || n == "gradPP") {
    s.rows = N + Fb;
    s.dirs = 3;
  } else if (n == "gradU") {
    s.rows = N + Fb;
    s.comps = 3;
    s.dirs = 3;
  } else if (n == "mdot_f" || n == "gamma_f" || n == "mdot_old") {
    s.rows = F;
  } else {
    fail("unknown field '" + n + "'");
  }
  return s;
}

// C index (row*comps + c)*dirs + d  <->  MATLAB index row + rows*(d + dirs*c)
void to_matlab(const FieldShape &s, const double *cl, double *ml) {
  for (int64_t r = 0; r < s.rows; ++r)
    for (int32_t c = 0; c < s.comps; ++c)
      for (int32_t d = 0; d < s.dirs; ++d)
        ml[r + s.rows * (d + (int64_t)s.dirs * c)] = cl[(r * s.comps + c) * s.dirs + d];
}

void from_matlab(const FieldShape &s, const double *ml, double *cl) {
  for (int64_t r = 0; r < s.rows; ++r)
    for (int32_t c = 0; c < s.comps; ++c)
      for (int32_t d = 0; d < s.dirs; ++d)
        cl[(r * s.comps + c) * s.dirs + d] = ml[r + s.rows * (d + (int64_t)s.dirs * c)];
}

}  // namespace ufvm_mex
