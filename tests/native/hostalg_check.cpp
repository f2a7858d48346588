// hostalg_check.cpp -- the product's host graph algorithms (hostalg.cpp:
// agglomeration, Galerkin gather lists, colourings, level schedules, DILU
// lower lists, ILU(0) update lists) on seeded graphs, built with AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_sanitizers.py.  Every result is
// checked against its invariants and the agglomeration against the oracle's
// (oracle/ufvm_oracle.c, compiled into this program), the ILU(0) factor that
// the update lists produce against the oracle's; exit status 0 = pass.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <numeric>
#include <random>
#include <vector>

#include "hostalg.h"
extern "C" {
#include "ufvm_oracle.h"
}

using namespace ufvm;

static int failures = 0;
#define CHECK(c)                                                           \
  do {                                                                     \
    if (!(c)) {                                                            \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      ++failures;                                                          \
    }                                                                      \
  } while (0)

// n1 x n2 x n3 grid, 7-point stencil (tri: plus the (i+1, j+1) diagonal of
// every square, so the graph has triangles), rows numbered by a seeded
// permutation, neighbours in increasing index order; anb < 0 with seeded
// magnitudes
static void grid(int n1, int n2, int n3, unsigned seed, HostCSR &g, std::vector<double> &anb,
                 bool tri = false) {
  const int n = n1 * n2 * n3;
  std::mt19937 rng(seed);
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::shuffle(perm.begin(), perm.end(), rng);
  std::vector<std::vector<int>> nb(n);
  auto id = [&](int i, int j, int k) { return perm[(k * n2 + j) * n1 + i]; };
  for (int k = 0; k < n3; ++k)
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i < n1; ++i) {
        const int a = id(i, j, k);
        if (i + 1 < n1) { nb[a].push_back(id(i + 1, j, k)); nb[id(i + 1, j, k)].push_back(a); }
        if (j + 1 < n2) { nb[a].push_back(id(i, j + 1, k)); nb[id(i, j + 1, k)].push_back(a); }
        if (k + 1 < n3) { nb[a].push_back(id(i, j, k + 1)); nb[id(i, j, k + 1)].push_back(a); }
        if (tri && i + 1 < n1 && j + 1 < n2) {
          nb[a].push_back(id(i + 1, j + 1, k));
          nb[id(i + 1, j + 1, k)].push_back(a);
        }
      }
  g.n = n;
  g.ptr.assign(1, 0);
  g.col.clear();
  std::uniform_real_distribution<double> w(0.5, 2.0);
  anb.clear();
  for (int a = 0; a < n; ++a) {
    std::sort(nb[a].begin(), nb[a].end());
    for (int b : nb[a]) {
      g.col.push_back(b);
      anb.push_back(-w(rng));
    }
    g.ptr.push_back((int)g.col.size());
  }
  // symmetric values: a_ab = a_ba
  for (int a = 0; a < n; ++a)
    for (int t = g.ptr[a]; t < g.ptr[a + 1]; ++t) {
      const int b = g.col[t];
      if (b < a) continue;
      for (int u = g.ptr[b]; u < g.ptr[b + 1]; ++u)
        if (g.col[u] == a) anb[u] = anb[t];
    }
}

static void check_level(const HostCSR &g, const std::vector<double> &anb) {
  AggResult ag;
  agglomerate_level(g, anb.data(), ag);
  CHECK(ag.nc >= 1 && ag.nc <= g.n);
  CHECK((int)ag.parents.size() == g.n);
  std::vector<int> kids(ag.nc, 0);
  for (int i = 0; i < g.n; ++i) {
    CHECK(ag.parents[i] >= 0 && ag.parents[i] < ag.nc);
    if (ag.parents[i] >= 0 && ag.parents[i] < ag.nc) ++kids[ag.parents[i]];
  }
  for (int c = 0; c < ag.nc; ++c) CHECK(kids[c] > 0);
  // the oracle's agglomeration of the same level: the same parents and coarse cconn
  std::vector<double> ac(g.n, 1.0), zero(g.n, 0.0);
  orc_level F{}, Cl{};
  F.n = g.n;
  F.nnz = (int32_t)g.col.size();
  F.ptr = const_cast<int32_t *>(g.ptr.data());
  F.col = const_cast<int32_t *>(g.col.data());
  F.ac = ac.data();
  F.anb = const_cast<double *>(anb.data());
  F.bc = zero.data();
  F.dphi = zero.data();
  F.parents = nullptr;
  const int nc = orc_agglomerate_level(&F, &Cl);
  CHECK(nc == ag.nc);
  if (nc == ag.nc) {
    for (int i = 0; i < g.n; ++i) CHECK(F.parents[i] == ag.parents[i]);
    for (int c = 0; c <= nc; ++c) CHECK(Cl.ptr[c] == ag.coarse.ptr[c]);
    for (int t = 0; t < Cl.ptr[nc]; ++t) CHECK(Cl.col[t] == ag.coarse.col[t]);
  }
  free(F.parents);
  free(Cl.ptr); free(Cl.col); free(Cl.ac); free(Cl.anb); free(Cl.bc); free(Cl.dphi);
  // Galerkin gather lists: every fine value lands in exactly one coarse value
  std::vector<int32_t> gp, gs;
  galerkin_lists(g, ag, gp, gs);
  CHECK((int)gp.size() == ag.nc + (int)ag.coarse.col.size() + 1);
  CHECK(gp.back() == (int)gs.size() && (int)gs.size() == g.n + (int)g.col.size());
  std::vector<int> seen(g.n + g.col.size(), 0);
  for (int32_t e : gs) {
    const long q = e >= 0 ? (long)g.n + e : (long)(-1 - e);
    CHECK(q >= 0 && q < (long)seen.size());
    if (q >= 0 && q < (long)seen.size()) ++seen[q];
  }
  for (int v : seen) CHECK(v == 1);
  // colourings are proper
  std::vector<int32_t> col;
  const int32_t nco = greedy_colour(g, col);
  for (int a = 0; a < g.n; ++a) {
    CHECK(col[a] >= 0 && col[a] < nco);
    for (int t = g.ptr[a]; t < g.ptr[a + 1]; ++t) CHECK(col[a] != col[g.col[t]]);
  }
  const int32_t R = 64;
  const int32_t ncb = greedy_colour_blocks(g, R, col);
  for (int a = 0; a < g.n; ++a) {
    CHECK(col[a] >= 0 && col[a] < ncb);
    for (int t = g.ptr[a]; t < g.ptr[a + 1]; ++t)
      if (g.col[t] / R == a / R) CHECK(col[a] != col[g.col[t]]);
  }
  // level schedules: every row once, after the neighbours it waits for
  for (bool fwd : {true, false}) {
    Schedule s;
    lex_schedule(g, fwd, s);
    CHECK((int)s.rows.size() == g.n && s.level_ptr.back() == g.n);
    std::vector<int> lev(g.n, -1);
    for (size_t l = 0; l + 1 < s.level_ptr.size(); ++l)
      for (int t = s.level_ptr[l]; t < s.level_ptr[l + 1]; ++t) lev[s.rows[t]] = (int)l;
    for (int a = 0; a < g.n; ++a) {
      CHECK(lev[a] >= 0);
      for (int t = g.ptr[a]; t < g.ptr[a + 1]; ++t) {
        const int b = g.col[t];
        if ((fwd && b < a) || (!fwd && b > a)) CHECK(lev[b] < lev[a]);
      }
    }
  }
  LowerLists L;
  CHECK(lower_lists(g, L));
  for (int j = 0; j < g.n; ++j)
    for (int t = L.ptr[j]; t < L.ptr[j + 1]; ++t) {
      const int i = L.nb[t];
      CHECK(i < j && g.col[L.pos_ji[t]] == i && g.col[L.pos_ij[t]] == j);
    }
}

// ILU(0) in index order from the update lists, in the arithmetic of the
// factorisation kernel (k_ilu0_factor), against the oracle's ILU(0) with every
// row its own colour (the same order): equal bit for bit
static void check_ilu0(const HostCSR &g, const std::vector<double> &anb) {
  LowerLists L;
  CHECK(lower_lists(g, L));
  Ilu0Lists U;
  ilu0_lists(g, L, U);
  CHECK(U.ptr.size() == L.nb.size() + 1 && U.src.size() == U.dst.size());
  const int n = g.n;
  std::vector<double> ac(n);
  for (int a = 0; a < n; ++a) {
    double s = 0.5;
    for (int t = g.ptr[a]; t < g.ptr[a + 1]; ++t) s -= anb[t];
    ac[a] = s;
  }
  std::vector<double> F(anb), dinv(n, 0.0);
  for (int j = 0; j < n; ++j) {
    double d = ac[j];
    for (int o = L.ptr[j]; o < L.ptr[j + 1]; ++o) {
      const double l = F[L.pos_ji[o]] * dinv[L.nb[o]];
      for (int h = U.ptr[o]; h < U.ptr[o + 1]; ++h) {
        CHECK(U.src[h] >= g.ptr[L.nb[o]] && U.src[h] < g.ptr[L.nb[o] + 1]);
        CHECK(U.dst[h] < 0 || (U.dst[h] >= g.ptr[j] && U.dst[h] < g.ptr[j + 1]));
        const int t = U.dst[h];
        const double u = F[U.src[h]];
        if (t < 0) d = d - l * u;
        else F[t] = F[t] - l * u;
      }
    }
    dinv[j] = 1.0 / d;
  }
  std::vector<int32_t> colour(n);
  std::iota(colour.begin(), colour.end(), 0);
  std::vector<double> zero(n, 0.0), Fo(anb.size()), dinvo(n);
  orc_level O{};
  O.n = n;
  O.nnz = (int32_t)g.col.size();
  O.ptr = const_cast<int32_t *>(g.ptr.data());
  O.col = const_cast<int32_t *>(g.col.data());
  O.ac = ac.data();
  O.anb = const_cast<double *>(anb.data());
  O.bc = zero.data();
  O.dphi = zero.data();
  orc_mcilu0_factor(&O, colour.data(), n, dinvo.data(), Fo.data());
  for (int a = 0; a < n; ++a) CHECK(dinv[a] == dinvo[a]);
  for (size_t t = 0; t < anb.size(); ++t) CHECK(F[t] == Fo[t]);
}

int main() {
  const int sizes[][3] = {{1, 1, 1}, {2, 1, 1}, {5, 4, 3}, {17, 9, 6}, {24, 24, 24}};
  unsigned seed = 7;
  for (auto &sz : sizes) {
    HostCSR g;
    std::vector<double> anb;
    grid(sz[0], sz[1], sz[2], seed++, g, anb);
    check_level(g, anb);
    check_ilu0(g, anb);
    grid(sz[0], sz[1], sz[2], seed++, g, anb, true);
    check_level(g, anb);
    check_ilu0(g, anb);
  }
  std::printf("hostalg_check: %d failures\n", failures);
  return failures ? 1 : 0;
}
