// hostalg.cpp -- see hostalg.h.
#include "hostalg.h"

#include <algorithm>
#include <cmath>
#include <limits>

namespace ufvm {

// Greedy seed agglomeration, uFVM/src/fvm/solve/AMG/cfdAgglomerateLevel.m:
//   strength of a link i->j is -anb_ij / maxAnb_i, maxAnb_i = max_j(-anb_ij)   (:13-15)
//   pass 1 (:20-54): every unassigned cell in index order becomes a seed and
//     claims its unassigned neighbours with strength > 0.5 (seed's maxAnb),
//     then each of those claims its own unassigned neighbours with strength
//     > 0.5 (the child's maxAnb); a seed that claimed nobody is released;
//   pass 2 (:57-75): each still-unassigned cell joins the parent of the
//     assigned neighbour j maximising -anb_ij / maxAnb_j (strictly larger
//     wins, starting from 0), else opens a new parent;
//   coarse cconn (:83-98): neighbours of a parent in the order first met
//     scanning fine rows in index order and cconn order.
void agglomerate_level(const HostCSR &g, const double *anb, AggResult &out) {
  const int32_t n = g.n;
  std::vector<double> maxanb(n);
  for (int32_t i = 0; i < n; ++i) {
    double m = -std::numeric_limits<double>::infinity();
    for (int32_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) m = std::max(m, -anb[k]);
    maxanb[i] = m;
  }
  std::vector<int32_t> par(n, -1);
  std::vector<int32_t> gen1;
  gen1.reserve(64);
  int32_t next = 0;
  for (int32_t seed = 0; seed < n; ++seed) {
    if (par[seed] >= 0) continue;
    par[seed] = next;
    gen1.clear();
    for (int32_t k = g.ptr[seed]; k < g.ptr[seed + 1]; ++k) {
      const int32_t j = g.col[k];
      if (par[j] < 0 && (-anb[k] / maxanb[seed]) > 0.5) {
        par[j] = next;
        gen1.push_back(j);
      }
    }
    size_t claimed = gen1.size();
    for (int32_t c : gen1) {
      for (int32_t k = g.ptr[c]; k < g.ptr[c + 1]; ++k) {
        const int32_t j = g.col[k];
        if (par[j] < 0 && (-anb[k] / maxanb[c]) > 0.5) {
          par[j] = next;
          ++claimed;
        }
      }
    }
    if (claimed == 0) par[seed] = -1;
    else ++next;
  }
  for (int32_t o = 0; o < n; ++o) {
    if (par[o] < 0) {
      double best = 0.0;
      for (int32_t k = g.ptr[o]; k < g.ptr[o + 1]; ++k) {
        const int32_t j = g.col[k];
        if (par[j] < 0) continue;
        const double s = -anb[k] / maxanb[j];
        if (best < s) {
          best = s;
          par[o] = par[j];
        }
      }
      if (par[o] < 0) par[o] = next++;
    }
  }
  const int32_t nc = next;
  out.nc = nc;
  out.parents = par;
  // children of each parent in increasing fine order (stable counting sort)
  out.child_ptr.assign(nc + 1, 0);
  for (int32_t i = 0; i < n; ++i) ++out.child_ptr[par[i] + 1];
  for (int32_t p = 0; p < nc; ++p) out.child_ptr[p + 1] += out.child_ptr[p];
  out.child.resize(n);
  {
    std::vector<int32_t> fill(out.child_ptr.begin(), out.child_ptr.end() - 1);
    for (int32_t i = 0; i < n; ++i) out.child[fill[par[i]]++] = i;
  }
  // First-encounter coarse neighbours.  Row p only receives entries while its
  // own children are scanned, and those are scanned in increasing fine order,
  // so walking each parent's children in order reproduces the reference's
  // append order exactly.  stamp[q] == p marks q as already listed for p.
  std::vector<int32_t> stamp(nc, -1), slot_of(nc, -1);
  out.coarse.n = nc;
  out.coarse.ptr.assign(nc + 1, 0);
  out.coarse.col.clear();
  out.coarse.col.reserve(g.col.size() / 2 + 16);
  out.entry_slot.assign(g.col.size(), -1);
  for (int32_t p = 0; p < nc; ++p) {
    for (int32_t c = out.child_ptr[p]; c < out.child_ptr[p + 1]; ++c) {
      const int32_t i = out.child[c];
      for (int32_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
        const int32_t q = par[g.col[k]];
        if (q == p) continue;
        if (stamp[q] != p) {
          stamp[q] = p;
          slot_of[q] = (int32_t)out.coarse.col.size();
          out.coarse.col.push_back(q);
        }
        out.entry_slot[k] = slot_of[q];
      }
    }
    out.coarse.ptr[p + 1] = (int32_t)out.coarse.col.size();
  }
}

void galerkin_lists(const HostCSR &g, const AggResult &ag, std::vector<int32_t> &ptr,
                    std::vector<int32_t> &src) {
  const int32_t nc = ag.nc;
  const size_t nout = (size_t)nc + ag.coarse.col.size();
  ptr.assign(nout + 1, 0);
  for (int32_t i = 0; i < g.n; ++i) ++ptr[ag.parents[i] + 1];
  for (int32_t i = 0; i < g.n; ++i)
    for (int32_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
      const int32_t sl = ag.entry_slot[k];
      ++ptr[(sl < 0 ? ag.parents[i] : nc + sl) + 1];
    }
  for (size_t o = 0; o < nout; ++o) ptr[o + 1] += ptr[o];
  src.assign(std::max<size_t>(ptr[nout], 1), 0);
  std::vector<int32_t> at(ptr.begin(), ptr.end() - 1);
  for (int32_t p = 0; p < nc; ++p)
    for (int32_t c = ag.child_ptr[p]; c < ag.child_ptr[p + 1]; ++c) {
      const int32_t i = ag.child[c];
      src[at[p]++] = -1 - i;
      for (int32_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
        const int32_t sl = ag.entry_slot[k];
        src[at[sl < 0 ? p : nc + sl]++] = k;
      }
    }
}

int32_t greedy_colour(const HostCSR &g, std::vector<int32_t> &colour) {
  const int32_t n = g.n;
  colour.assign(n, -1);
  std::vector<int32_t> seen(64, -1);
  int32_t ncol = 0;
  for (int32_t i = 0; i < n; ++i) {
    for (int32_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
      const int32_t c = colour[g.col[k]];
      if (c >= 0) {
        if (c >= (int32_t)seen.size()) seen.resize(2 * c + 2, -1);
        seen[c] = i;
      }
    }
    int32_t c = 0;
    while (c < (int32_t)seen.size() && seen[c] == i) ++c;
    colour[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

int32_t greedy_colour_blocks(const HostCSR &g, int32_t rows_per_block,
                             std::vector<int32_t> &colour) {
  const int32_t n = g.n;
  colour.assign(n, -1);
  std::vector<int32_t> seen(64, -1);
  int32_t ncol = 0;
  for (int32_t i = 0; i < n; ++i) {
    const int32_t lo = (i / rows_per_block) * rows_per_block;
    const int32_t hi = std::min(n, lo + rows_per_block);
    for (int32_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k) {
      const int32_t j = g.col[k];
      if (j < lo || j >= hi) continue;
      const int32_t c = colour[j];
      if (c >= 0) {
        if (c >= (int32_t)seen.size()) seen.resize(2 * c + 2, -1);
        seen[c] = i;
      }
    }
    int32_t c = 0;
    while (c < (int32_t)seen.size() && seen[c] == i) ++c;
    colour[i] = c;
    ncol = std::max(ncol, c + 1);
  }
  return ncol;
}

void lex_schedule(const HostCSR &g, bool forward, Schedule &s) {
  const int32_t n = g.n;
  std::vector<int32_t> lev(n, 0);
  int32_t nlev = 0;
  if (forward) {
    for (int32_t i = 0; i < n; ++i) {
      int32_t l = 0;
      for (int32_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k)
        if (g.col[k] < i) l = std::max(l, lev[g.col[k]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  } else {
    for (int32_t i = n - 1; i >= 0; --i) {
      int32_t l = 0;
      for (int32_t k = g.ptr[i]; k < g.ptr[i + 1]; ++k)
        if (g.col[k] > i) l = std::max(l, lev[g.col[k]] + 1);
      lev[i] = l;
      nlev = std::max(nlev, l + 1);
    }
  }
  s.level_ptr.assign(nlev + 1, 0);
  for (int32_t i = 0; i < n; ++i) ++s.level_ptr[lev[i] + 1];
  for (int32_t l = 0; l < nlev; ++l) s.level_ptr[l + 1] += s.level_ptr[l];
  s.rows.resize(n);
  std::vector<int32_t> fill(s.level_ptr.begin(), s.level_ptr.end() - 1);
  for (int32_t i = 0; i < n; ++i) s.rows[fill[lev[i]]++] = i;
}

bool lower_lists(const HostCSR &g, LowerLists &L) {
  const int32_t n = g.n;
  L.ptr.assign(n + 1, 0);
  for (int32_t j = 0; j < n; ++j)
    for (int32_t k = g.ptr[j]; k < g.ptr[j + 1]; ++k)
      if (g.col[k] < j) ++L.ptr[j + 1];
  for (int32_t j = 0; j < n; ++j) L.ptr[j + 1] += L.ptr[j];
  L.nb.resize(L.ptr[n]);
  L.pos_ji.resize(L.ptr[n]);
  L.pos_ij.resize(L.ptr[n]);
  bool ok = true;
  std::vector<std::pair<int32_t, int32_t>> tmp;
  for (int32_t j = 0; j < n; ++j) {
    tmp.clear();
    for (int32_t k = g.ptr[j]; k < g.ptr[j + 1]; ++k)
      if (g.col[k] < j) tmp.emplace_back(g.col[k], k);
    // the reference pushes a_ji * d_i * a_ij into row j while sweeping i
    // upwards, so row j's updates arrive in increasing i
    std::stable_sort(tmp.begin(), tmp.end(),
                     [](const std::pair<int32_t, int32_t> &a, const std::pair<int32_t, int32_t> &b) {
                       return a.first < b.first;
                     });
    int32_t o = L.ptr[j];
    for (auto &e : tmp) {
      const int32_t i = e.first;
      int32_t pij = -1;
      for (int32_t t = g.ptr[i]; t < g.ptr[i + 1]; ++t)
        if (g.col[t] == j) { pij = t; break; }
      if (pij < 0) ok = false;
      // the reference looks up the FIRST occurrence of i in row j
      int32_t pji = -1;
      for (int32_t t = g.ptr[j]; t < g.ptr[j + 1]; ++t)
        if (g.col[t] == i) { pji = t; break; }
      L.nb[o] = i;
      L.pos_ji[o] = pji;
      L.pos_ij[o] = pij;
      ++o;
    }
  }
  return ok;
}

void ilu0_lists(const HostCSR &g, const LowerLists &L, Ilu0Lists &U) {
  const int32_t n = g.n;
  U.ptr.assign(L.nb.size() + 1, 0);
  U.src.clear();
  U.dst.clear();
  // first[c]: the first slot of column c in the current row j (stamped by j)
  std::vector<int32_t> first(n, -1), stamp(n, -1);
  for (int32_t j = 0; j < n; ++j) {
    for (int32_t t = g.ptr[j]; t < g.ptr[j + 1]; ++t) {
      const int32_t c = g.col[t];
      if (stamp[c] != j) {
        stamp[c] = j;
        first[c] = t;
      }
    }
    for (int32_t o = L.ptr[j]; o < L.ptr[j + 1]; ++o) {
      const int32_t k = L.nb[o];
      for (int32_t q = g.ptr[k]; q < g.ptr[k + 1]; ++q) {
        const int32_t c = g.col[q];
        if (c <= k) continue;
        if (c == j) {
          U.src.push_back(q);
          U.dst.push_back(-1);
        } else if (stamp[c] == j) {
          U.src.push_back(q);
          U.dst.push_back(first[c]);
        }
      }
      U.ptr[o + 1] = (int32_t)U.src.size();
    }
  }
}

bool pack_sell_columns(const std::vector<int32_t> &ptr, const std::vector<int32_t> &soff,
                       const std::vector<int32_t> &scol, int32_t n, std::vector<uint16_t> &d16,
                       std::vector<int32_t> &base) {
  const int32_t nq = (n + 63) / 64;
  d16.assign(scol.size(), 0);
  base.assign(std::max(nq, 1), 0);
  for (int32_t q = 0; q < nq; ++q) {
    const int32_t s0 = soff[q], w = (soff[q + 1] - s0) / 64;
    const int32_t t0 = q * 64, t1 = std::min(n, t0 + 64);
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    for (int32_t t = t0; t < t1; ++t) {
      const int32_t len = ptr[t + 1] - ptr[t];
      const int32_t pad = len > 0 ? scol[s0 + (t - t0) + (len - 1) * 64] : t;
      for (int32_t e = 0; e < w; ++e) {
        const int32_t v = e < len ? scol[s0 + (t - t0) + e * 64] : pad;
        lo = std::min<int64_t>(lo, v);
        hi = std::max<int64_t>(hi, v);
      }
    }
    if (w == 0 || lo > hi) continue;
    if (hi - lo > 65535) return false;
    base[q] = (int32_t)lo;
    for (int32_t t = t0; t < t1; ++t) {
      const int32_t len = ptr[t + 1] - ptr[t];
      const int32_t pad = len > 0 ? scol[s0 + (t - t0) + (len - 1) * 64] : t;
      for (int32_t e = 0; e < w; ++e) {
        const int32_t v = e < len ? scol[s0 + (t - t0) + e * 64] : pad;
        d16[s0 + (t - t0) + e * 64] = (uint16_t)(v - lo);
      }
    }
  }
  return true;
}

}  // namespace ufvm
