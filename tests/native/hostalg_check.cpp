// hostalg_check.cpp -- the product's host graph algorithms (hostalg.cpp:
// agglomeration, Galerkin gather lists, colourings, level schedules, DILU
// lower lists, ILU(0) update lists, 16-bit sliced-ELL columns) on seeded
// graphs, built with AddressSanitizer and
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

// the sliced-ELL layout of build_fast_level (64-row chunks, width = the
// chunk's longest row, padding column = the row itself) packed to 16 bits:
// every slot of a processed row decodes to its column (padding: the row's
// last real column, or the row when it has none), and packing succeeds
// exactly when every chunk's decoded columns span at most 65535
static void check_pack(const HostCSR &g) {
  const int32_t n = g.n, nq = (n + 63) / 64;
  std::vector<int32_t> soff(nq + 1, 0);
  for (int32_t q = 0; q < nq; ++q) {
    int32_t w = 0;
    for (int32_t t = q * 64; t < std::min(n, q * 64 + 64); ++t)
      w = std::max(w, g.ptr[t + 1] - g.ptr[t]);
    soff[q + 1] = soff[q] + 64 * w;
  }
  std::vector<int32_t> scol(std::max(soff[nq], 1), 0);
  for (int32_t t = 0; t < n; ++t) {
    const int32_t q = t / 64, w = (soff[q + 1] - soff[q]) / 64, len = g.ptr[t + 1] - g.ptr[t];
    for (int32_t e = 0; e < w; ++e)
      scol[soff[q] + (t & 63) + e * 64] = e < len ? g.col[g.ptr[t] + e] : t;
  }
  bool fits = true;
  for (int32_t q = 0; q < nq; ++q) {
    int32_t lo = n, hi = -1;
    const int32_t w = (soff[q + 1] - soff[q]) / 64;
    for (int32_t t = q * 64; t < std::min(n, q * 64 + 64) && w > 0; ++t) {
      const int32_t len = g.ptr[t + 1] - g.ptr[t];
      for (int32_t e = 0; e < std::max(len, 1); ++e) {
        const int32_t c = len ? g.col[g.ptr[t] + e] : t;
        lo = std::min(lo, c);
        hi = std::max(hi, c);
      }
    }
    if (hi >= lo && hi - lo > 65535) fits = false;
  }
  std::vector<uint16_t> d16;
  std::vector<int32_t> base;
  const bool ok = pack_sell_columns(g.ptr, soff, scol, n, d16, base);
  CHECK(ok == fits);
  if (!ok) return;
  CHECK(d16.size() == scol.size() && (int32_t)base.size() == std::max(nq, 1));
  for (int32_t q = 0; q < nq; ++q) {
    const int32_t w = (soff[q + 1] - soff[q]) / 64;
    for (int32_t u = 0; u < 64; ++u) {
      const int32_t t = q * 64 + u;
      const int32_t len = t < n ? g.ptr[t + 1] - g.ptr[t] : 0;
      for (int32_t e = 0; e < w; ++e) {
        const int32_t got = base[q] + d16[soff[q] + u + e * 64];
        CHECK(got >= 0 && got < n);
        if (t >= n) CHECK(got == base[q]);
        else if (e < len) CHECK(got == g.col[g.ptr[t] + e]);
        else CHECK(got == (len ? g.col[g.ptr[t] + len - 1] : t));
      }
    }
  }
}

// a chain of n rows (i -- i+1), plus the edge (0, n-1) when far is set
static void chain(int32_t n, bool far, HostCSR &g) {
  std::vector<std::vector<int32_t>> nb(n);
  for (int32_t i = 0; i + 1 < n; ++i) {
    nb[i].push_back(i + 1);
    nb[i + 1].push_back(i);
  }
  if (far && n > 2) {
    nb[0].push_back(n - 1);
    nb[n - 1].push_back(0);
  }
  g.n = n;
  g.ptr.assign(1, 0);
  g.col.clear();
  for (int32_t a = 0; a < n; ++a) {
    std::sort(nb[a].begin(), nb[a].end());
    g.col.insert(g.col.end(), nb[a].begin(), nb[a].end());
    g.ptr.push_back((int32_t)g.col.size());
  }
}

// g renumbered so each colour class of a greedy colouring is contiguous (the
// finest level's multicolour numbering)
static HostCSR colour_sorted(const HostCSR &g) {
  std::vector<int32_t> colour;
  greedy_colour(g, colour);
  std::vector<int32_t> perm(g.n), iperm(g.n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(),
                   [&](int32_t a, int32_t b) { return colour[a] < colour[b]; });
  for (int32_t t = 0; t < g.n; ++t) iperm[perm[t]] = t;
  HostCSR h;
  h.n = g.n;
  h.ptr.assign(1, 0);
  for (int32_t t = 0; t < g.n; ++t) {
    const int32_t i = perm[t];
    for (int32_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) h.col.push_back(iperm[g.col[k]]);
    h.ptr.push_back((int32_t)h.col.size());
  }
  return h;
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
    check_pack(g);
    check_pack(colour_sorted(g));
  }
  {
    HostCSR g;
    chain(70000, false, g);
    check_pack(g);
    chain(70000, true, g);       // chunk 0 and the last chunk span 69999 columns
    check_pack(g);
    chain(65600, true, g);       // the edge spans 65599
    check_pack(g);
    chain(65536, true, g);       // spans 65535: still fits
    check_pack(g);
    chain(65537, true, g);       // spans 65536: does not
    check_pack(g);
    g.n = 100;                   // rows without neighbours: padding is the row itself
    g.ptr.assign(101, 0);
    g.col.clear();
    check_pack(g);
  }
  std::printf("hostalg_check: %d failures\n", failures);
  return failures ? 1 : 0;
}
