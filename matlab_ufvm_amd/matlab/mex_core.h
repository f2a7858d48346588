// mex_core.h -- the marshalling half of the MATLAB gateway (ufvm_mex.cpp),
// free of the MATLAB API so that it builds and is tested without MATLAB.
//
// ufvmSetup.m gathers the reference's `global Domain` (theMesh struct arrays,
// the U and p models' boundary conditions, the fvSolution settings) into one
// struct of plain arrays per argument; this layer turns those named arrays
// into the C ABI's ufvm_mesh_desc / ufvm_simple_cfg (0-based indices,
// row-major vectors, UFVM_* codes), flattens Domain.coefficients' cconn / anb
// cell arrays into CSR, and converts case fields between the C ABI's layout
// ([row][component][direction], row-major) and MATLAB's (rows x directions x
// components, column-major: U is (N+Fb) x 3, grad(U) (N+Fb) x 3 x 3 as
// cfdUpdateGradient stores phiGradient).
#pragma once
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "ufvm.h"

namespace ufvm_mex {

// one argument: a numeric array (double or int32, column-major, rows x cols),
// a char row vector, or a cell array of char row vectors
struct Arr {
  enum Kind { NONE, F64, I32, TEXT, TEXTS } kind = NONE;
  const void *data = nullptr;
  int64_t rows = 0, cols = 0;
  std::string text;
  std::vector<std::string> texts;
  int64_t numel() const { return rows * cols; }
};

// a struct of named arguments (a MATLAB struct in the gateway, a map in tests)
class Source {
 public:
  virtual ~Source() = default;
  virtual bool get(const std::string &name, Arr &out) const = 0;
};

class MapSource : public Source {
 public:
  std::map<std::string, Arr> fields;
  bool get(const std::string &name, Arr &out) const override;
};

// bad input from the caller; the gateway reports it with mexErrMsgIdAndTxt
struct Error {
  std::string msg;
};

// the reference's strings -> UFVM_* codes
int32_t patch_type_code(const std::string &t);   // boundary file types
int32_t bc_code(const std::string &t);           // 0/U, 0/p boundary field types
int32_t cycle_code(const std::string &t);        // 'V-Cycle' 'W-Cycle' 'F-Cycle'
int32_t smoother_code(const std::string &t);     // 'ILU' 'SOR' 'MCGS'
int32_t pcg_smoother_code(const std::string &t); // 'MCGS' 'ILU' 'MCILU' 'MCILU0'
int32_t p_solver_code(const std::string &t);     // 'amg' 'pcg' 'pcg-frozen'

// A case's arrays in the C ABI's conventions (owned) and the structs of
// ufvm_case_create pointing into them.  Not copyable: the structs point at
// the vectors.
struct CaseInput {
  std::vector<int32_t> owner, neighbour, cconn_ptr, cconn, own_nnz, nei_nnz;
  std::vector<int32_t> patch_start, patch_nfaces, patch_type, u_bc, p_bc;
  std::vector<double> Sf, area, CN, eCN, gDiff, T, gf, face_centroid, volume, centroid;
  std::vector<double> u_value, p_value;
  ufvm_mesh_desc mesh{};
  ufvm_simple_cfg cfg{};
  CaseInput() = default;
  CaseInput(const CaseInput &) = delete;
  CaseInput &operator=(const CaseInput &) = delete;
};

// mesh: the arrays ufvmSetup.m gathers from theMesh; opts: boundary
// conditions and settings (every field optional except the boundary
// conditions; defaults as cfdDefineEquation.m:84-136).  Throws Error.
void build_case(const Source &mesh, const Source &opts, CaseInput &out);

// CSR pattern of a coefficient struct from csize (N) and its cconn rows
// concatenated (1-based), as Domain.coefficients{l}.cconn / .csize hold them
struct Pattern {
  std::vector<int32_t> ptr, col;
};
void build_pattern(const Arr &csize, const Arr &cconn, Pattern &out);

// Settings of cfdApplyAMG from theEquation.multigrid and theSolutionMethods
// (strings as the reference keeps them, eval'd there: 'V-Cycle', '30', '0.1')
ufvm_amg_cfg amg_cfg(const Source &multigrid, const std::string &smoother);
// Settings of the precision-form AMG-PCG (fields pcgPrecision, pcgSweeps,
// pcgMaxCoarseCells, pcgMaxIter, pcgTol, pcgSmoother, pcgCoarseScale,
// pcgCoarseSweeps, nullspace; all optional)
ufvm_pcg_cfg pcg_cfg(const Source &opts);

// case field layouts: C ABI [rows][comps][dirs] <-> MATLAB rows x dirs x comps
struct FieldShape {
  int64_t rows = 0;
  int32_t comps = 1, dirs = 1;
  int64_t numel() const { return rows * comps * dirs; }
};
FieldShape field_shape(const std::string &name, int64_t n_cells, int64_t n_bfaces,
                       int64_t n_faces);
void to_matlab(const FieldShape &s, const double *c_layout, double *m_layout);
void from_matlab(const FieldShape &s, const double *m_layout, double *c_layout);

}  // namespace ufvm_mex
