// state.h -- device-resident state behind the C ABI: contexts, systems,
// hierarchy levels, and the precision-form (multicolour) level copies.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "hostalg.h"
#include "ufvm.h"

namespace ufvm {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define UFVM_HIP(call)                                                                    \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      throw ::ufvm::Error(UFVM_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Serialises device allocation, release and device-wide synchronisation with
// HIP graph capture: several host threads (a server's worker and SIMPLE
// threads) may capture while others set up, and HIP refuses such operations
// while any stream captures.
inline std::mutex &device_mgmt_mutex() {
  static std::mutex m;
  return m;
}

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  DBuf(DBuf &&o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  DBuf &operator=(DBuf &&o) noexcept {
    if (this != &o) {
      release();
      p = o.p;
      n = o.n;
      o.p = nullptr;
      o.n = 0;
    }
    return *this;
  }
  ~DBuf() { release(); }
  void release() {
    if (p) {
      std::lock_guard<std::mutex> g(device_mgmt_mutex());
      (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) count = 1;
    std::lock_guard<std::mutex> g(device_mgmt_mutex());
    UFVM_HIP(hipMalloc(&p, count * sizeof(T)));
    n = count;
  }
  // synchronous with respect to the host buffer: returns after the copy is done
  void upload(const T *h, size_t count, hipStream_t s) {
    alloc(count);
    if (count) {
      UFVM_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
      UFVM_HIP(hipStreamSynchronize(s));
    }
  }
  void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
  // enqueued only: h must stay valid until the stream has been synchronised
  void upload_async(const std::vector<T> &v, hipStream_t s) {
    alloc(v.size());
    if (!v.empty())
      UFVM_HIP(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void download(T *h, size_t count, hipStream_t s) const {
    if (count) UFVM_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

// One level of the reference hierarchy, in reference numbering and cconn order.
struct RefLevel {
  int32_t n = 0, nnz = 0;
  HostCSR g;
  DBuf<int32_t> ptr, col;
  DBuf<double> ac, anb, bc, dphi, r, rc, dinv;
  // map to the next coarser level (defined below the coarsest)
  std::vector<int32_t> h_parents;
  DBuf<int32_t> parents, child_ptr, child;
  // the coarse operator's gather lists (galerkin_lists), values of the next
  // coarser level: nc diagonal values then its off-diagonal slots
  DBuf<int32_t> gal_ptr, gal_src;
  int32_t gal_out = 0;
  // lexicographic schedules (SOR / DILU), built on first use
  bool sched = false;
  Schedule fwd, bwd;
  DBuf<int32_t> fwd_rows, bwd_rows;
  // the schedules' level pointers on the device, and whether every level is
  // small enough that a whole sweep runs in one workgroup (one launch
  // instead of one per level)
  DBuf<int32_t> fwd_lp, bwd_lp;
  bool one_wg = false;
  int32_t widest = 0;             // rows of the widest level of both schedules
  int32_t max_row = 0;            // the most off-diagonal entries of a row
  int32_t max_tri = 0;            // the most lower neighbours or upper entries of a row
  // for the pipelined sweep (widest <= kPipeBlock): per direction, the level
  // in schedule order as ELL (pipe_w slots; slot u of position s at
  // u * n + s): columns and row lengths built with the schedule, values and
  // diagonals refreshed before each use
  int32_t pipe_w = 0;
  DBuf<int32_t> fwd_ecol, bwd_ecol, fwd_elen, bwd_elen;
  DBuf<double> fwd_eanb, bwd_eanb, fwd_eac, bwd_eac;
  // the same for the DILU substitutions (when the lower lists exist):
  // forward copy of the lower lists, backward copy of the upper entries,
  // each slot's anb index (-1: padding), and y = dinv * rc of the forward
  int32_t dpipe_wf = 0, dpipe_wb = 0;
  DBuf<int32_t> dfwd_ecol, dfwd_elen, dfwd_esrc, dbwd_ecol, dbwd_elen, dbwd_esrc;
  DBuf<double> dfwd_eval, dbwd_eval, dilu_y;
  // and for the DILU factor: the lower lists without the skipped neighbour,
  // with both value positions (a_ji, a_ij)
  int32_t fpipe_w = 0;
  DBuf<int32_t> fac_ecol, fac_elen, fac_sji, fac_sij;
  DBuf<double> fac_eji, fac_eij;
  bool low_ok = false;
  LowerLists low;
  DBuf<int32_t> lptr, lnb, lpji, lpij;
  // multicolour GS: the classes of a greedy colouring (rows in increasing
  // index) as sliced-ELL chunks of 64 rows, built on first use (a chunk holds
  // rows of one colour; cs_row = -1 pads a colour to whole chunks): colour c
  // is chunks [cs_cptr[c], cs_cptr[c+1]); entries in cconn order, padding
  // col = -1; values gathered from anb through cs_src
  bool cs = false;
  std::vector<int32_t> cs_cptr;
  DBuf<int32_t> cs_row, cs_off, cs_col, cs_src;
  DBuf<double> cs_val;
  int64_t cs_slots = 0;
};

// One level of the precision-form V-cycle: unknowns renumbered so that each
// colour class is a contiguous row range, CSR kept in cconn order per row.
struct FastLevel {
  int32_t n = 0, nnz = 0, ncol = 0;
  // kind 0: global colouring, colour classes contiguous (level 1)
  // kind 1: blocks of at most kBRows rows, block-local colouring (coarse levels)
  // kind 2: coarsest level, solved densely
  // kind 3: coarse level smoothed by damped Jacobi, rows sorted by length
  // kind 4: finest level smoothed by the reference's DILU in its own
  //         (lexicographic) order, level-scheduled; reference numbering
  // kind 5: finest level smoothed by DILU in colour order: kind 0's
  //         numbering, so a colour class is one level of the lower-triangular
  //         dependency graph
  // kind 6: finest level smoothed by ILU(0) in colour order: kind 5's
  //         numbering and lists, plus the factor's own off-diagonal values
  int kind = 0;
  // rows longer than a run on average (the coarse levels): the residual and
  // Jacobi kernels load a row's entries in one round (k::kWideMax)
  bool wide = false;
  std::vector<int32_t> col_off;   // kind 0: colour c rows are [col_off[c], col_off[c+1])
  int32_t nblk = 0, bncol = 0;    // kind 1: blocks and colours per block
  std::vector<int32_t> h_bptr, h_boff;
  DBuf<int32_t> bptr, boff;
  std::vector<int32_t> colour;    // reference numbering
  std::vector<int32_t> perm;      // perm[new] = reference index
  std::vector<int32_t> h_ptr;     // permuted CSR row pointer (host copy)
  DBuf<int32_t> ptr, col, perm_d;
  DBuf<int32_t> csr_src;          // reference CSR position of each entry
  DBuf<float> ac32, anb32;
  DBuf<double> ac64, anb64;
  DBuf<int32_t> parent;           // coarse row of each row (next level's numbering)
  DBuf<int32_t> csoff, cschild;   // child lists of the next level's rows, sliced-ELL
  // sliced-ELL copy of the off-diagonals (kinds 0, 3, 5, 6 and level 0):
  // 64-row chunks, chunk q at [soff[q], soff[q+1]) column-major, padding
  // value 0.  Columns as int32 (scol), or, when every chunk's columns span
  // less than 65536, as uint16 offsets (sd16) from a base per chunk (scbase)
  std::vector<int32_t> h_soff;
  DBuf<int32_t> soff, scol, scbase;
  DBuf<uint16_t> sd16;
  DBuf<int32_t> sell_src;         // reference CSR position of each slot (-1: padding)
  DBuf<float> sval32;
  DBuf<double> sval64;
  // precision 16, finest level: the off-diagonals divided by their row's
  // diagonal, FP16 (the colour passes and residual of k::mcgs_sell_h)
  DBuf<_Float16> sval16;
  // kinds 4, 5: DILU diagonal inverse (factored in FP64 at set-up / refresh)
  DBuf<float> dinv32;
  DBuf<double> dinv64;
  // kind 5: row numbers 0..n-1 (the row lists of the colour levels) and the
  // lower neighbour lists of the permuted pattern (lower_lists)
  DBuf<int32_t> iota, lptr, lnb, lpji, lpij;
  // kind 6: the ILU(0) factor's off-diagonals in the permuted CSR layout
  // (f64, factored in FP64) and in the sliced-ELL layout (sf64, sf32; slot ->
  // CSR position fsrc, -1 padding), and the update lists (Ilu0Lists)
  DBuf<double> f64, sf64;
  DBuf<float> sf32;
  DBuf<int32_t> fsrc, hptr, hsrc, hdst;
};

// Per-solve device workspace of one system: each level's right-hand side,
// iterate, residual and second iterate (in the V-cycle precision) and the FCG
// vectors.  Several workspaces of one system run independent solves at once.
struct LevelWork {
  DBuf<float> b32, x32, r32, y32;
  DBuf<double> b64, x64, r64, y64;
  int64_t ld = 0;                // stride between the right-hand sides of a batch
};

struct SolveWork {
  int32_t n0 = 0;
  int32_t nrhs = 1;              // right-hand sides solved together (vectors RHS-major)
  int64_t ld0 = 0;               // stride between them in the finest-level vectors
  std::vector<LevelWork> lev;
  DBuf<double> x, r, p, w;       // FCG vectors, finest level in its V-cycle numbering
  DBuf<double> bref, xref;       // right-hand side / solution in reference numbering
  DBuf<double> partial;          // per-block partial sums
  DBuf<double> scal;             // device scalars
  // pinned host-coherent ring the solve's kernels write {value, tag} pairs
  // into: slots 0-1 even iterations, 2-3 odd iterations, 4-5 set-up norm
  double *hring = nullptr, *dring = nullptr;
  hipEvent_t ev[3] = {nullptr, nullptr, nullptr};
  // captured FCG iterations: [0] the first of a solve, [1] any later one
  hipGraphExec_t iter_graph[2] = {nullptr, nullptr};
  size_t bytes = 0;              // device bytes held
  SolveWork() = default;
  SolveWork(const SolveWork &) = delete;
  SolveWork &operator=(const SolveWork &) = delete;
  ~SolveWork() {
    std::lock_guard<std::mutex> lk(device_mgmt_mutex());
    for (hipGraphExec_t g : iter_graph)
      if (g) (void)hipGraphExecDestroy(g);
    if (hring) (void)hipHostFree(hring);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
  }
};

struct PcgState {
  bool ready = false;        // layouts, colourings and workspace built for the hierarchy
  bool stale = false;        // level-1 values changed since the last setup or refresh
  ufvm_pcg_cfg cfg{};
  int32_t nlev = 0;          // levels used; nlev-1 solved densely
  int32_t singular = 0;
  std::vector<std::unique_ptr<FastLevel>> lev;
  DBuf<double> ainv, colbuf;
  std::unique_ptr<SolveWork> work0;   // the workspace of ufvm_pcg_solve
  std::map<int, std::unique_ptr<SolveWork>> work_batch;   // of ufvm_pcg_solve_batch, by k
};

struct System {
  int32_t n = 0, nnz = 0;
  std::vector<std::unique_ptr<RefLevel>> lev;
  int32_t nlevels = 1;
  bool hier_valid = false;
  PcgState pcg;
};

struct ProfRec {
  std::string kind;
  hipEvent_t a, b;
  double bytes;
  int64_t launches = 1;   // kernel launches between the two events
};

}  // namespace ufvm

struct ufvm_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  std::map<int32_t, std::unique_ptr<ufvm::System>> systems;
  int32_t next_id = 1;
  ufvm::DBuf<double> partial;   // per-block partial sums
  ufvm::DBuf<double> scal;      // device scalars
  ufvm::DBuf<double> alg;       // smoothing calls inside AMG cycles: initial, final, stop, NaN
  ufvm::DBuf<unsigned> counter; // last-block reduction counter (kept at zero between launches)
  double *hscal = nullptr;      // pinned host mirror of scal
  // SIMPLE cases of this context (type-erased; owned here so that destroying
  // the context releases them)
  std::map<int32_t, std::shared_ptr<void>> cases;
  bool prof_on = false;
  int32_t prof_stride = 1;                  // time every prof_stride-th launch of a kind
  std::map<std::string, int64_t> prof_seen; // launches seen per kind while enabled
  std::vector<ufvm::ProfRec> prof;
  std::vector<hipEvent_t> event_pool;
};
