// pcg.h -- precision-form AMG-preconditioned flexible CG (internal to libufvm).
#pragma once
#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "internal.h"
#include "kernels.h"

namespace ufvm {
namespace detail {

// Device workspace for one solve (k = 1) or a batch of k = 2, 4 or 8 solves
// on system S (whose ufvm_pcg_setup is done): level vectors in the V-cycle
// precision, FCG vectors, scalars, partials, the pinned result ring and its
// events; cleared on stream s (synchronised).
std::unique_ptr<SolveWork> make_work(const System &S, hipStream_t s, int k = 1);
// device bytes make_work(S, s, k) would allocate
size_t work_bytes(const System &S, int k = 1);

// One FCG(1) solve of S on workspace W and stream s, driven as a state
// machine so that one host thread can interleave solves on several streams:
// start() enqueues the set-up, poll() advances when the GPU has finished the
// awaited step (or blocks for it when block is set) and returns true once the
// solution has been written.  The host reads only the ring; it never waits on
// the stream while another step of the same solve could be queued.
class FcgRun {
 public:
  // Iterations are replayed as captured HIP graphs (one launch instead of
  // ~30) on systems small enough for host launch cost to matter; on large
  // ones the GPU is the bottleneck and eager launches leave fewer gaps.
  // Never while kernel profiling is on.
  static constexpr int32_t kGraphRows = 600000;
  FcgRun(ufvm_ctx *ctx, System &S, SolveWork &W, hipStream_t s)
      : ctx_(ctx), S_(&S), W_(&W), s_(s),
        graphs_(!ctx->prof_on && (S.pcg.cfg.launch_mode == 2 ||
                                  (S.pcg.cfg.launch_mode == 0 && W.n0 < kGraphRows))) {}
  // b: right-hand side, x: solution, both device arrays in reference order
  void start(const double *b, double *x);
  bool poll(bool block);
  bool done() const { return phase_ == 4; }
  int32_t iterations() const { return it_; }
  double relres() const { return nb_ > 0.0 ? rr_ / nb_ : 0.0; }

 private:
  template <typename V>
  void start_t(const double *b);
  template <typename V>
  void issue_t(int k);
  void issue(int k);
  void finish();
  bool ready(hipEvent_t e, bool block);

  ufvm_ctx *ctx_;
  System *S_;
  SolveWork *W_;
  hipStream_t s_;
  bool graphs_;
  double *x_out_ = nullptr;
  int phase_ = 0;    // 1 set-up queued, 2 iterating, 3 result queued, 4 done
  int it_ = 0, issued_ = 0;
  bool ahead_ = true;
  double prev2_ = -1.0, nb_ = 0.0, thr_ = 0.0, rr_ = 0.0;
};

// FCG(1) for the k right-hand sides of a batch workspace (make_work(S, s, k))
// at once: every kernel of an iteration covers all k (the matrix is read once
// per k solves) and each right-hand side gets the bits of its own solve.  A
// right-hand side that has converged stays frozen while the others continue.
class FcgBatch {
 public:
  FcgBatch(ufvm_ctx *ctx, System &S, SolveWork &W, hipStream_t s);
  // B: k right-hand sides, X: k solutions, each n values in reference order
  // one after another, host or device memory (X written when poll returns
  // true); B / X may be the workspace's own bref / xref
  void start(const double *B, double *X);
  bool poll(bool block);
  int32_t iterations(int q) const { return its_[q]; }
  double relres(int q) const { return nb_[q] > 0.0 ? rr_[q] / nb_[q] : 0.0; }

 private:
  template <typename V>
  void start_t();
  template <typename V>
  void issue_t(int it);
  void finish();
  k::BatchStrides strides() const;

  ufvm_ctx *ctx_;
  System *S_;
  SolveWork *W_;
  hipStream_t s_;
  int k_;
  double *X_ = nullptr;
  int phase_ = 0, it_ = 0, issued_ = 0;
  std::vector<double> nb_, thr_, rr_;
  std::vector<int32_t> its_;
  std::vector<char> done_;
};

// whether S's precision-form set-up can run batched solves (multicolour GS
// finest level, Jacobi coarse levels)
bool batch_supported(const System &S);

// FCG(1) for g solves of different systems at once (serving form).  The
// members advance in lock step, one FCG iteration at a time, and each phase
// of an iteration is one launch for all of them: every colour pass,
// residual, restriction, Jacobi sweep, prolongation and dense solve of the
// V-cycle at each level, and every FCG kernel (k::group_row, k::group_fcg,
// with a device table of each member's arguments).  Members whose V-cycles
// are shorter skip the launches of levels they do not have.  A member keeps
// the arithmetic, block decomposition and reduction order of its own
// FcgRun, so its solution, iteration count and residual are those of its
// own solve; a member that converges is finished at once (its launches
// then skip it) while the others continue.  Members need group_supported
// systems with the same V-cycle precision and max_iter, each on its own
// single-right-hand-side workspace.
class FcgGroup {
 public:
  struct Member {
    System *S;
    SolveWork *W;
    const double *b;          // device, reference order (may be W->bref)
    double *x;                // device, reference order (may be W->xref)
    double *x_host = nullptr; // if set, x is also copied here when the member finishes
  };
  FcgGroup(ufvm_ctx *ctx, std::vector<Member> members, hipStream_t s);
  ~FcgGroup();
  FcgGroup(const FcgGroup &) = delete;
  FcgGroup &operator=(const FcgGroup &) = delete;
  void start();
  // advance without blocking (or block for the awaited step); true once every
  // member has finished and its copies are done
  bool poll(bool block);
  int size() const { return (int)m_.size(); }
  // member q's solution (and host copy) is written, or it failed
  bool member_done(int q);
  const std::string &member_error(int q) const { return st_[q].err; }
  int32_t iterations(int q) const { return st_[q].its; }
  double relres(int q) const { return st_[q].nb > 0.0 ? st_[q].rr / st_[q].nb : 0.0; }
  int64_t launches() const { return launches_; }   // grouped launches issued

 private:
  struct MemberState {
    double nb = 0.0, thr = 0.0, rr = 0.0, prev2 = -1.0;
    int32_t its = 0;
    bool active = false, finished = false, ahead = true;
    hipEvent_t ev = nullptr;
    std::string err;
  };
  struct Step {                 // one grouped launch
    int32_t op, gx, g;
    size_t table;               // byte offset of its argument table in tables_
  };
  template <typename V>
  void build();
  template <typename V>
  void issue_t(int k);
  void issue(int k);
  void finish_member(int q);

  ufvm_ctx *ctx_;
  std::vector<Member> m_;
  std::vector<MemberState> st_;
  hipStream_t s_;
  int precision_ = 32;
  int32_t max_iter_ = 0;
  bool graphs_ = false;
  std::vector<Step> steps_[2];  // V-cycle launches of the first / later iterations
  size_t fcg_table_ = 0;
  int32_t fcg_gx_ = 0;
  DBuf<char> tables_;
  hipEvent_t ev_[3] = {nullptr, nullptr, nullptr};
  hipGraphExec_t graph_[2] = {nullptr, nullptr};
  int phase_ = 0, it_ = 0, issued_ = 0;
  int64_t launches_ = 0;
};

// whether S's precision-form set-up can run in a group (batch_supported and
// at least one pre-smoothing sweep on every level)
bool group_supported(const System &S);
// blocking group solve on the context stream (ufvm_pcg_solve_group): each
// system's own workspace, host right-hand sides and solutions
void pcg_solve_group(ufvm_ctx *ctx, const std::vector<System *> &sys,
                     const double *const *b, double *const *x, int32_t *iters,
                     double *relres);
// blocking batch solve on the context stream (workspace cached per k)
void pcg_solve_batch(ufvm_ctx *ctx, System &S, int k, const double *B, double *X,
                     int32_t *iters, double *relres);

// z = M^-1 r for host arrays in reference order (one preconditioner V-cycle)
void pcg_apply_vcycle(ufvm_ctx *ctx, System &S, const double *r, double *z);

}  // namespace detail
}  // namespace ufvm
