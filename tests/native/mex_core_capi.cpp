// mex_core_capi.cpp -- a C interface to the MATLAB gateway's marshalling
// layer (matlab_ufvm_amd/matlab/mex_core.cpp) for tests/test_mex_core.py:
// named arrays in MATLAB's conventions go in as a MapSource, the structs
// of ufvm_case_create (and the other conversions) come out.  Built with g++
// by the test; needs neither MATLAB nor a GPU.
#include <cstdio>
#include <cstring>
#include <string>

#include "mex_core.h"

using namespace ufvm_mex;

namespace {

struct Src {
  MapSource m;
  std::vector<std::vector<int32_t>> keep_i32;
  std::vector<std::vector<double>> keep_f64;
};

int report(const Error &e, char *err, int errlen) {
  if (err && errlen > 0) std::snprintf(err, (size_t)errlen, "%s", e.msg.c_str());
  return 1;
}

}  // namespace

extern "C" {

void *mc_source_new() { return new Src(); }
void mc_source_free(void *s) { delete static_cast<Src *>(s); }

void mc_source_f64(void *s, const char *name, const double *d, int64_t rows, int64_t cols) {
  Src &S = *static_cast<Src *>(s);
  S.keep_f64.emplace_back(d, d + rows * cols);
  Arr a;
  a.kind = Arr::F64;
  a.data = S.keep_f64.back().data();
  a.rows = rows;
  a.cols = cols;
  S.m.fields[name] = a;
}

void mc_source_i32(void *s, const char *name, const int32_t *d, int64_t rows, int64_t cols) {
  Src &S = *static_cast<Src *>(s);
  S.keep_i32.emplace_back(d, d + rows * cols);
  Arr a;
  a.kind = Arr::I32;
  a.data = S.keep_i32.back().data();
  a.rows = rows;
  a.cols = cols;
  S.m.fields[name] = a;
}

void mc_source_text(void *s, const char *name, const char *t) {
  Arr a;
  a.kind = Arr::TEXT;
  a.text = t;
  static_cast<Src *>(s)->m.fields[name] = a;
}

// n strings, each terminated by '\0', one after another
void mc_source_texts(void *s, const char *name, const char *joined, int n) {
  Arr a;
  a.kind = Arr::TEXTS;
  for (int i = 0; i < n; ++i) {
    a.texts.emplace_back(joined);
    joined += a.texts.back().size() + 1;
  }
  static_cast<Src *>(s)->m.fields[name] = a;
}

void *mc_build_case(void *mesh, void *opts, char *err, int errlen) {
  CaseInput *c = new CaseInput();
  try {
    build_case(static_cast<Src *>(mesh)->m, static_cast<Src *>(opts)->m, *c);
  } catch (const Error &e) {
    report(e, err, errlen);
    delete c;
    return nullptr;
  }
  return c;
}
const ufvm_mesh_desc *mc_case_mesh(void *c) { return &static_cast<CaseInput *>(c)->mesh; }
const ufvm_simple_cfg *mc_case_cfg(void *c) { return &static_cast<CaseInput *>(c)->cfg; }
void mc_case_free(void *c) { delete static_cast<CaseInput *>(c); }

int mc_amg_cfg(void *src, const char *smoother, ufvm_amg_cfg *out, char *err, int errlen) {
  try {
    *out = amg_cfg(static_cast<Src *>(src)->m, smoother);
  } catch (const Error &e) {
    return report(e, err, errlen);
  }
  return 0;
}

int mc_pcg_cfg(void *src, ufvm_pcg_cfg *out, char *err, int errlen) {
  try {
    *out = pcg_cfg(static_cast<Src *>(src)->m);
  } catch (const Error &e) {
    return report(e, err, errlen);
  }
  return 0;
}

// pattern of csize / cconn (1-based) fields of a source; ptr needs n+1, col nnz
int mc_pattern(void *src, int32_t *ptr, int32_t *col, int64_t nnz_cap, char *err, int errlen) {
  try {
    Arr cs, cc;
    const MapSource &m = static_cast<Src *>(src)->m;
    if (!m.get("csize", cs) || !m.get("cconn", cc)) throw Error{"csize / cconn missing"};
    Pattern p;
    build_pattern(cs, cc, p);
    if ((int64_t)p.col.size() > nnz_cap) throw Error{"nnz_cap too small"};
    std::memcpy(ptr, p.ptr.data(), p.ptr.size() * sizeof(int32_t));
    if (!p.col.empty()) std::memcpy(col, p.col.data(), p.col.size() * sizeof(int32_t));
  } catch (const Error &e) {
    return report(e, err, errlen);
  }
  return 0;
}

int mc_field_shape(const char *name, int64_t N, int64_t Fb, int64_t F, int64_t *rows,
                   int32_t *comps, int32_t *dirs) {
  try {
    const FieldShape s = field_shape(name, N, Fb, F);
    *rows = s.rows;
    *comps = s.comps;
    *dirs = s.dirs;
  } catch (const Error &) {
    return 1;
  }
  return 0;
}

void mc_to_matlab(const char *name, int64_t N, int64_t Fb, int64_t F, const double *c,
                  double *m) {
  to_matlab(field_shape(name, N, Fb, F), c, m);
}

void mc_from_matlab(const char *name, int64_t N, int64_t Fb, int64_t F, const double *m,
                    double *c) {
  from_matlab(field_shape(name, N, Fb, F), m, c);
}

}  // extern "C"
