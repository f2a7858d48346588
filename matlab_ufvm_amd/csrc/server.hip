// server.hip -- serving form: a single-GPU request loop for pressure-Poisson
// solves and SIMPLE steps on resident meshes.
//
// Pressure solves are continuously batched: a worker thread keeps up to
// max_active solves on the GPU, each on its own workspace and one of
// n_streams streams, and advances every active solve one FCG iteration at a
// time as the GPU finishes it (FcgRun state machines, polled without
// blocking).  A request that arrives while others run is admitted as soon as
// a slot is free and the device has memory for its workspace (workspaces of
// finished solves are kept per system and reused); queued solves of one system
// share a slot as one batch.  SIMPLE steps run in lanes: each lane owns some
// cases, a context (stream) and a thread, so steps of cases in different
// lanes overlap each other and the solves on the device.
// Each solve's arithmetic depends only on its system and right-hand side, so
// a result is the same alone and in any mix.  When no batch forms, queued
// solves of different systems of comparable size are admitted together as one
// group (FcgGroup): every phase of their FCG iterations is one launch for
// all of them, and each finishes (and is answered) as soon as it converges.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>

#include "internal.h"
#include "pcg.h"
#include "ufvm.h"

namespace ufvm {
namespace detail {

static double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct Request {
  int64_t ticket = 0;
  int type = 0;                 // 0 pressure solve, 1 SIMPLE steps
  int32_t id = 0;               // system or case
  const double *b = nullptr;    // solve: right-hand side (host)
  double *x = nullptr;          // solve: solution (host)
  int32_t nsteps = 0;           // steps
  double *rms = nullptr;        // steps: 4 RMS residuals per step (host, may be null)
  int status = 0;               // 0 queued, 1 running, 2 done, < 0 error
  std::string err;
  int32_t iterations = 0;
  int32_t batch = 0;
  int32_t group = 0;
  double relres = 0.0;
  double t_submit = 0.0, t_start = 0.0, t_done = 0.0;
};

// One admitted solve, a batch of k solves of one system (FcgBatch), or a
// group of solves of different systems (FcgGroup).
struct Slot {
  std::vector<Request *> reqs;   // empty: free
  System *sys = nullptr;
  hipStream_t stream = nullptr;
  std::unique_ptr<SolveWork> work;
  std::unique_ptr<FcgRun> run;     // k == 1
  std::unique_ptr<FcgBatch> brun;  // k > 1
  std::unique_ptr<FcgGroup> grun;  // a group: one workspace per member, pooled only
  std::vector<std::unique_ptr<SolveWork>> gwork;   // once the whole group has drained
  std::vector<char> answered;      // group members already completed
  hipEvent_t copied = nullptr;
  int phase = 0;   // 1 solving, 2 result copies queued
};

}  // namespace detail
}  // namespace ufvm

using namespace ufvm;
using namespace ufvm::detail;

struct ufvm_server {
  int device = 0;
  ufvm_server_cfg cfg{};
  ufvm_ctx *ctx_p = nullptr;   // pressure systems being served
  ufvm_ctx *ctx_b = nullptr;   // builds new systems (then handed to ctx_p)
  // SIMPLE-step lanes: lane i owns the cases of ctx_s[i] and serves
  // pend_step[i] on its own thread
  std::vector<ufvm_ctx *> ctx_s;
  std::vector<std::unique_ptr<std::mutex>> case_mu;   // per lane: its context's cases
  std::vector<std::deque<Request *>> pend_step;
  std::vector<int64_t> lane_cells;
  std::unordered_map<int32_t, std::pair<int, int32_t>> case_of;   // server case -> lane, case
  int32_t next_case = 0;
  std::vector<hipStream_t> streams;
  std::vector<Slot> slots;
  std::thread worker;
  std::vector<std::thread> steppers;
  std::mutex mu;                       // requests, queues, pools, stop flag
  std::condition_variable cv_work, cv_done;
  std::mutex sys_mu;                   // ctx_p's system map (held only for lookups/inserts)
  std::mutex build_mu;                 // ctx_b
  bool stop = false;
  std::deque<Request *> pend_solve;
  std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
  // free workspaces by (system, batch size)
  std::map<std::pair<int32_t, int>, std::vector<std::unique_ptr<SolveWork>>> pool;
  int64_t next_ticket = 1;
  int32_t n_active = 0;
  size_t work_bytes_held = 0;
  std::string err;
};

namespace {

template <typename F>
int sguard(ufvm_server *srv, F &&f) {
  if (!srv) return UFVM_ERR_ARG;
  try {
    UFVM_HIP(hipSetDevice(srv->device));
    f();
    return UFVM_OK;
  } catch (const Error &e) {
    std::lock_guard<std::mutex> g(srv->mu);
    srv->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    std::lock_guard<std::mutex> g(srv->mu);
    srv->err = e.what();
    return UFVM_ERR_ARG;
  }
}

void check_rc(ufvm_ctx *ctx, int rc) {
  if (rc != UFVM_OK) throw Error(rc, ufvm_last_error(ctx));
}

void complete(ufvm_server *srv, Request *r, int status, const std::string &msg) {
  std::lock_guard<std::mutex> g(srv->mu);
  r->status = status;
  r->err = msg;
  r->t_done = now_s();
  srv->cv_done.notify_all();
}

// device memory that admission keeps free
bool fits(ufvm_server *srv, size_t bytes) {
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
  const double reserve = srv->cfg.hbm_reserve * (double)total_b;
  return (double)free_b - (double)bytes > reserve;
}

System *lookup(ufvm_server *srv, int32_t id) {
  std::lock_guard<std::mutex> gs(srv->sys_mu);
  auto it = srv->ctx_p->systems.find(id);
  return it == srv->ctx_p->systems.end() ? nullptr : it->second.get();
}

// Admit queued solves into free slots (worker thread).  The oldest queued
// solve is admitted together with later queued solves of the same system, as
// one batch of the largest size in {8, 4, 2} (up to max_batch) they fill.
// Otherwise later queued solves of other systems within group_span of its
// size join it as one group (up to max_group, one solve per system), as far
// as the device has room for their workspaces.
void admit(ufvm_server *srv) {
  for (Slot &sl : srv->slots) {
    if (!sl.reqs.empty()) continue;
    std::vector<Request *> group;
    std::vector<System *> sys;
    std::vector<std::unique_ptr<SolveWork>> works;   // null: to be made
    int k = 1;
    {
      std::lock_guard<std::mutex> g(srv->mu);
      if (srv->pend_solve.empty()) return;
      Request *r = srv->pend_solve.front();
      System *S = lookup(srv, r->id);
      group.push_back(r);
      sys.push_back(S);
      if (S && srv->cfg.max_batch > 1 && batch_supported(*S))
        for (auto it = srv->pend_solve.begin() + 1;
             it != srv->pend_solve.end() && (int)group.size() < srv->cfg.max_batch; ++it)
          if ((*it)->id == r->id) group.push_back(*it);
      while (k * 2 <= (int)group.size()) k *= 2;
      group.resize(k);
      // groups of launch-bound solves only: on systems as large as the graph
      // threshold the kernels are bandwidth-bound, and solves on separate
      // streams overlap better than in lock step
      if (S && k == 1 && srv->cfg.max_group > 1 && group_supported(*S) &&
          S->lev[0]->n < FcgRun::kGraphRows) {
        const double n0 = S->lev[0]->n;
        for (auto it = srv->pend_solve.begin() + 1;
             it != srv->pend_solve.end() && (int)group.size() < srv->cfg.max_group; ++it) {
          System *T = lookup(srv, (*it)->id);
          if (!T || !group_supported(*T) || T->lev[0]->n >= FcgRun::kGraphRows ||
              std::find(sys.begin(), sys.end(), T) != sys.end())
            continue;
          if (T->pcg.cfg.precision != S->pcg.cfg.precision ||
              T->pcg.cfg.max_iter != S->pcg.cfg.max_iter)
            continue;
          const double n1 = T->lev[0]->n;
          if (std::max(n0, n1) > srv->cfg.group_span * std::min(n0, n1)) continue;
          group.push_back(*it);
          sys.push_back(T);
        }
      }
      if (S) {
        // workspaces (one k-solve workspace for a single solve or a batch, one
        // each for a group's members): pooled ones, then new ones while the
        // device has room; a group is cut where the room ends
        const size_t nw = k > 1 ? 1 : group.size();
        size_t need = 0;
        for (size_t q = 0; q < nw; ++q) {
          auto &fl = srv->pool[{group[q]->id, k}];
          if (!fl.empty()) {
            works.push_back(std::move(fl.back()));
            fl.pop_back();
            continue;
          }
          need += work_bytes(*sys[q], k);
          if (!fits(srv, need)) {
            if (k > 1) group.clear();
            else group.resize(q);
            sys.resize(group.size());
            break;
          }
          works.push_back(nullptr);
        }
        if (group.empty()) return;   // admission by free memory: wait for running solves
      }
      for (Request *q : group) {
        srv->pend_solve.erase(std::find(srv->pend_solve.begin(), srv->pend_solve.end(), q));
        q->status = 1;
        q->t_start = now_s();
      }
    }
    if (!sys[0]) {
      complete(srv, group[0], -UFVM_ERR_ARG, "unknown system");
      continue;
    }
    const int g = (int)group.size();
    try {
      for (int q = 0; q < (int)works.size(); ++q)
        if (!works[q]) {
          works[q] = make_work(*sys[q], sl.stream, k);
          std::lock_guard<std::mutex> lk(srv->mu);
          srv->work_bytes_held += works[q]->bytes;
        }
      if (g > 1 && k == 1) {
        std::vector<FcgGroup::Member> mem;
        for (int q = 0; q < g; ++q) {
          SolveWork &W = *works[q];
          UFVM_HIP(hipMemcpyAsync(W.bref.p, group[q]->b, (size_t)W.n0 * sizeof(double),
                                  hipMemcpyHostToDevice, sl.stream));
          mem.push_back({sys[q], &W, W.bref.p, W.xref.p, group[q]->x});
        }
        sl.grun.reset(new FcgGroup(srv->ctx_p, mem, sl.stream));
        sl.grun->start();
        sl.gwork = std::move(works);
        sl.answered.assign(g, 0);
      } else {
        std::unique_ptr<SolveWork> &W = works[0];
        for (int q = 0; q < k; ++q)
          UFVM_HIP(hipMemcpyAsync(W->bref.p + (int64_t)q * W->n0, group[q]->b,
                                  (size_t)W->n0 * sizeof(double), hipMemcpyHostToDevice,
                                  sl.stream));
        if (k == 1) {
          sl.run.reset(new FcgRun(srv->ctx_p, *sys[0], *W, sl.stream));
          sl.run->start(W->bref.p, W->xref.p);
        } else {
          sl.brun.reset(new FcgBatch(srv->ctx_p, *sys[0], *W, sl.stream));
          sl.brun->start(W->bref.p, W->xref.p);
        }
        sl.work = std::move(W);
      }
      sl.reqs = group;
      sl.sys = sys[0];
      sl.phase = 1;
      std::lock_guard<std::mutex> lk(srv->mu);
      srv->n_active += g;
    } catch (const Error &e) {
      (void)hipStreamSynchronize(sl.stream);
      sl.grun.reset();
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        for (size_t q = 0; q < works.size(); ++q)
          if (works[q]) srv->pool[{group[q]->id, k}].push_back(std::move(works[q]));
        for (size_t q = 0; q < sl.gwork.size(); ++q)
          if (sl.gwork[q]) srv->pool[{group[q]->id, 1}].push_back(std::move(sl.gwork[q]));
        sl.gwork.clear();
      }
      for (Request *q : group) complete(srv, q, -e.code, e.what());
    }
  }
}

// Advance a group slot: answer each member as soon as its solution is on the
// host; pool the workspaces once the whole group has drained.
bool progress_group(ufvm_server *srv, Slot &sl) {
  bool moved = false;
  const int g = (int)sl.reqs.size();
  auto drain = [&]() {
    std::lock_guard<std::mutex> lk(srv->mu);
    for (int q = 0; q < g; ++q)
      if (sl.gwork[q]) srv->pool[{sl.reqs[q]->id, 1}].push_back(std::move(sl.gwork[q]));
    sl.gwork.clear();
  };
  try {
    const bool all = sl.grun->poll(false);
    for (int q = 0; q < g; ++q) {
      if (sl.answered[q] || !sl.grun->member_done(q)) continue;
      Request *r = sl.reqs[q];
      const std::string &err = sl.grun->member_error(q);
      {
        std::lock_guard<std::mutex> lk(srv->mu);
        r->iterations = sl.grun->iterations(q);
        r->relres = sl.grun->relres(q);
        r->batch = 1;
        r->group = g;
        srv->n_active -= 1;
      }
      complete(srv, r, err.empty() ? 2 : -UFVM_ERR_NUMERIC, err);
      sl.answered[q] = 1;
      moved = true;
    }
    if (all) {
      for (int q = 0; q < g; ++q)
        if (!sl.answered[q]) throw Error(UFVM_ERR_STATE, "group drained with a member unanswered");
      drain();
      sl.grun.reset();
      sl.reqs.clear();
      sl.phase = 0;
      moved = true;
    }
  } catch (const Error &e) {
    (void)hipStreamSynchronize(sl.stream);
    for (int q = 0; q < g; ++q)
      if (!sl.answered[q]) {
        {
          std::lock_guard<std::mutex> lk(srv->mu);
          srv->n_active -= 1;
        }
        complete(srv, sl.reqs[q], -e.code, e.what());
      }
    drain();
    sl.grun.reset();
    sl.reqs.clear();
    sl.phase = 0;
    moved = true;
  }
  return moved;
}

// Advance every active solve without blocking; true if anything moved.
bool progress(ufvm_server *srv) {
  bool moved = false;
  for (Slot &sl : srv->slots) {
    if (sl.reqs.empty()) continue;
    if (sl.grun) {
      moved = progress_group(srv, sl) || moved;
      continue;
    }
    const int k = (int)sl.reqs.size();
    auto release = [&](bool keep_work) {
      std::lock_guard<std::mutex> g(srv->mu);
      if (keep_work && sl.work) srv->pool[{sl.reqs[0]->id, k}].push_back(std::move(sl.work));
      srv->n_active -= k;
    };
    try {
      if (sl.phase == 1 && (k == 1 ? sl.run->poll(false) : sl.brun->poll(false))) {
        for (int q = 0; q < k; ++q)
          UFVM_HIP(hipMemcpyAsync(sl.reqs[q]->x, sl.work->xref.p + (int64_t)q * sl.work->n0,
                                  (size_t)sl.work->n0 * sizeof(double), hipMemcpyDeviceToHost,
                                  sl.stream));
        UFVM_HIP(hipEventRecord(sl.copied, sl.stream));
        sl.phase = 2;
        moved = true;
      }
      if (sl.phase == 2) {
        const hipError_t q = hipEventQuery(sl.copied);
        if (q == hipErrorNotReady) continue;
        UFVM_HIP(q);
        {
          std::lock_guard<std::mutex> g(srv->mu);
          for (int i = 0; i < k; ++i) {
            sl.reqs[i]->iterations = k == 1 ? sl.run->iterations() : sl.brun->iterations(i);
            sl.reqs[i]->batch = k;
            sl.reqs[i]->group = 1;
            sl.reqs[i]->relres = k == 1 ? sl.run->relres() : sl.brun->relres(i);
          }
        }
        release(true);
        for (Request *r : sl.reqs) complete(srv, r, 2, "");
        sl.reqs.clear();
        sl.run.reset();
        sl.brun.reset();
        sl.phase = 0;
        moved = true;
      }
    } catch (const Error &e) {
      (void)hipStreamSynchronize(sl.stream);
      release(true);
      for (Request *r : sl.reqs) complete(srv, r, -e.code, e.what());
      sl.reqs.clear();
      sl.run.reset();
      sl.brun.reset();
      sl.phase = 0;
      moved = true;
    }
  }
  return moved;
}

void worker_main(ufvm_server *srv) {
  (void)hipSetDevice(srv->device);
  for (;;) {
    admit(srv);
    const bool moved = progress(srv);
    bool active = false;
    for (Slot &sl : srv->slots) active = active || !sl.reqs.empty();
    std::unique_lock<std::mutex> g(srv->mu);
    if (!active && srv->pend_solve.empty()) {
      if (srv->stop) return;
      srv->cv_work.wait(g, [&] { return srv->stop || !srv->pend_solve.empty(); });
      continue;
    }
    g.unlock();
    // the GPU is busy with every active solve's queued iteration: yield briefly
    if (!moved) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
}

void stepper_main(ufvm_server *srv, int lane) {
  (void)hipSetDevice(srv->device);
  ufvm_ctx *ctx = srv->ctx_s[lane];
  for (;;) {
    Request *r = nullptr;
    int32_t cas = 0;
    {
      std::unique_lock<std::mutex> g(srv->mu);
      auto &q = srv->pend_step[lane];
      srv->cv_work.wait(g, [&] { return srv->stop || !q.empty(); });
      if (q.empty()) return;   // stop requested and nothing left
      r = q.front();
      q.pop_front();
      r->status = 1;
      r->t_start = now_s();
      cas = srv->case_of.at(r->id).second;
    }
    int rc;
    {
      std::lock_guard<std::mutex> g(*srv->case_mu[lane]);
      rc = ufvm_simple_iterate(ctx, cas, r->nsteps, r->rms);
    }
    complete(srv, r, rc == UFVM_OK ? 2 : -rc, rc == UFVM_OK ? "" : ufvm_last_error(ctx));
  }
}

// lane and lane-local case of a server case id
std::pair<int, int32_t> case_lane(ufvm_server *srv, int32_t cas) {
  std::lock_guard<std::mutex> g(srv->mu);
  auto it = srv->case_of.find(cas);
  if (it == srv->case_of.end()) throw Error(UFVM_ERR_ARG, "unknown case");
  return it->second;
}

Request *find(ufvm_server *srv, int64_t ticket) {
  auto it = srv->reqs.find(ticket);
  if (it == srv->reqs.end()) throw Error(UFVM_ERR_ARG, "unknown ticket");
  return it->second.get();
}

void fill_info(const Request *r, ufvm_request_info *info) {
  if (!info) return;
  info->status = r->status;
  info->iterations = r->iterations;
  info->relres = r->relres;
  info->queued_s = r->t_start > 0 ? r->t_start - r->t_submit : now_s() - r->t_submit;
  info->run_s = r->t_done > 0 ? r->t_done - r->t_start : 0.0;
  info->batch = r->batch;
  info->group = r->group;
}

}  // namespace

extern "C" {

// frees whatever a (possibly partly built) server holds; threads joined first
static void free_server(ufvm_server *srv) {
  (void)hipSetDevice(srv->device);
  for (Slot &sl : srv->slots) {
    sl.run.reset();
    sl.brun.reset();
    sl.grun.reset();
    sl.gwork.clear();
    sl.work.reset();
    if (sl.copied) (void)hipEventDestroy(sl.copied);
  }
  for (hipStream_t s : srv->streams)
    if (s) {
      (void)hipStreamSynchronize(s);
      (void)hipStreamDestroy(s);
    }
  srv->pool.clear();
  if (srv->ctx_p) ufvm_ctx_destroy(srv->ctx_p);
  if (srv->ctx_b) ufvm_ctx_destroy(srv->ctx_b);
  for (ufvm_ctx *c : srv->ctx_s)
    if (c) ufvm_ctx_destroy(c);
  delete srv;
}

int ufvm_server_create(int32_t device, const ufvm_server_cfg *cfg, ufvm_server **out) {
  if (!out || !cfg) return UFVM_ERR_ARG;
  *out = nullptr;
  if (cfg->max_active < 1 || cfg->n_streams < 1 || cfg->hbm_reserve < 0.0 ||
      cfg->hbm_reserve >= 1.0 || cfg->n_steppers < 0 || cfg->n_steppers > 16)
    return UFVM_ERR_ARG;
  const int32_t mb = cfg->max_batch == 0 ? 8 : cfg->max_batch;
  if (mb != 1 && mb != 2 && mb != 4 && mb != 8) return UFVM_ERR_ARG;
  const int32_t mg = cfg->max_group == 0 ? 8 : cfg->max_group;
  const double span = cfg->group_span == 0.0 ? 4.0 : cfg->group_span;
  if (mg < 1 || mg > 64 || !(span >= 1.0)) return UFVM_ERR_ARG;
  const int32_t lanes = cfg->n_steppers == 0 ? 4 : cfg->n_steppers;
  ufvm_server *srv = new (std::nothrow) ufvm_server();
  if (!srv) return UFVM_ERR_ARG;
  srv->device = device;
  srv->cfg = *cfg;
  srv->cfg.max_batch = mb;
  srv->cfg.max_group = mg;
  srv->cfg.group_span = span;
  srv->cfg.n_steppers = lanes;
  int rc = ufvm_ctx_create(device, &srv->ctx_p);
  if (rc == UFVM_OK) rc = ufvm_ctx_create(device, &srv->ctx_b);
  for (int i = 0; i < lanes && rc == UFVM_OK; ++i) {
    ufvm_ctx *c = nullptr;
    rc = ufvm_ctx_create(device, &c);
    if (rc == UFVM_OK) srv->ctx_s.push_back(c);
  }
  if (rc == UFVM_OK && hipSetDevice(device) != hipSuccess) rc = UFVM_ERR_HIP;
  if (rc == UFVM_OK) {
    srv->streams.assign(cfg->n_streams, nullptr);
    for (hipStream_t &s : srv->streams)
      if (rc == UFVM_OK && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
        rc = UFVM_ERR_HIP;
  }
  if (rc == UFVM_OK) {
    srv->slots.resize(cfg->max_active);
    for (size_t i = 0; i < srv->slots.size() && rc == UFVM_OK; ++i) {
      srv->slots[i].stream = srv->streams[i % srv->streams.size()];
      if (hipEventCreateWithFlags(&srv->slots[i].copied, hipEventDisableTiming) != hipSuccess)
        rc = UFVM_ERR_HIP;
    }
  }
  if (rc != UFVM_OK) {
    free_server(srv);
    return rc;
  }
  for (int i = 0; i < lanes; ++i) srv->case_mu.emplace_back(new std::mutex());
  srv->pend_step.resize(lanes);
  srv->lane_cells.assign(lanes, 0);
  srv->worker = std::thread(worker_main, srv);
  for (int i = 0; i < lanes; ++i) srv->steppers.emplace_back(stepper_main, srv, i);
  *out = srv;
  return UFVM_OK;
}

int ufvm_server_destroy(ufvm_server *srv) {
  if (!srv) return UFVM_ERR_ARG;
  {
    std::lock_guard<std::mutex> g(srv->mu);
    srv->stop = true;
  }
  srv->cv_work.notify_all();
  if (srv->worker.joinable()) srv->worker.join();
  for (std::thread &t : srv->steppers)
    if (t.joinable()) t.join();
  free_server(srv);
  return UFVM_OK;
}

const char *ufvm_server_last_error(const ufvm_server *srv) {
  return srv ? srv->err.c_str() : "null server";
}

int ufvm_server_add_system(ufvm_server *srv, int32_t n, int32_t nnz, const int32_t *ptr,
                           const int32_t *col, const double *ac, const double *anb,
                           const ufvm_pcg_cfg *cfg, int32_t *sys_out) {
  return sguard(srv, [&] {
    if (!cfg || !sys_out) throw Error(UFVM_ERR_ARG, "null argument");
    // build on a context of its own, so the solves in flight keep going,
    // then hand the resident system to the serving context
    std::lock_guard<std::mutex> gb(srv->build_mu);
    ufvm_ctx *c = srv->ctx_b;
    int32_t sid = 0;
    check_rc(c, ufvm_system_create(c, n, nnz, ptr, col, &sid));
    int rc = ufvm_system_set_coefficients(c, sid, ac, anb, nullptr);
    int32_t nl = 0;
    if (rc == UFVM_OK) rc = ufvm_amg_agglomerate(c, sid, 10, &nl);
    if (rc == UFVM_OK) rc = ufvm_pcg_setup(c, sid, cfg);
    if (rc != UFVM_OK) {
      const std::string msg = ufvm_last_error(c);
      ufvm_system_destroy(c, sid);
      throw Error(rc, msg);
    }
    std::lock_guard<std::mutex> gs(srv->sys_mu);
    const int32_t id = srv->ctx_p->next_id++;
    srv->ctx_p->systems[id] = std::move(c->systems[sid]);
    c->systems.erase(sid);
    *sys_out = id;
  });
}

int ufvm_server_add_case(ufvm_server *srv, const ufvm_mesh_desc *mesh, const ufvm_simple_cfg *cfg,
                         int32_t *case_out) {
  return sguard(srv, [&] {
    if (!mesh || !case_out) throw Error(UFVM_ERR_ARG, "null argument");
    int lane = 0;
    {
      std::lock_guard<std::mutex> g(srv->mu);
      for (int i = 1; i < (int)srv->lane_cells.size(); ++i)
        if (srv->lane_cells[i] < srv->lane_cells[lane]) lane = i;
    }
    int32_t local = 0;
    {
      std::lock_guard<std::mutex> g(*srv->case_mu[lane]);
      check_rc(srv->ctx_s[lane], ufvm_case_create(srv->ctx_s[lane], mesh, cfg, &local));
    }
    std::lock_guard<std::mutex> g(srv->mu);
    srv->lane_cells[lane] += mesh->n_cells;
    const int32_t id = srv->next_case++;
    srv->case_of[id] = {lane, local};
    *case_out = id;
  });
}

int ufvm_server_get_field(ufvm_server *srv, int32_t cas, const char *name, double *data,
                          int64_t count) {
  return sguard(srv, [&] {
    const auto lc = case_lane(srv, cas);
    std::lock_guard<std::mutex> g(*srv->case_mu[lc.first]);
    check_rc(srv->ctx_s[lc.first],
             ufvm_case_get_field(srv->ctx_s[lc.first], lc.second, name, data, count));
  });
}

int ufvm_server_set_field(ufvm_server *srv, int32_t cas, const char *name, const double *data,
                          int64_t count) {
  return sguard(srv, [&] {
    const auto lc = case_lane(srv, cas);
    std::lock_guard<std::mutex> g(*srv->case_mu[lc.first]);
    check_rc(srv->ctx_s[lc.first],
             ufvm_case_set_field(srv->ctx_s[lc.first], lc.second, name, data, count));
  });
}

int ufvm_server_submit_solve(ufvm_server *srv, int32_t sys, const double *b, double *x,
                             int64_t *ticket) {
  return sguard(srv, [&] {
    if (!b || !x || !ticket) throw Error(UFVM_ERR_ARG, "null argument");
    std::unique_ptr<Request> r(new Request());
    r->type = 0;
    r->id = sys;
    r->b = b;
    r->x = x;
    r->t_submit = now_s();
    std::lock_guard<std::mutex> g(srv->mu);
    if (srv->stop) throw Error(UFVM_ERR_STATE, "server stopping");
    r->ticket = srv->next_ticket++;
    *ticket = r->ticket;
    srv->pend_solve.push_back(r.get());
    srv->reqs[r->ticket] = std::move(r);
    srv->cv_work.notify_all();
  });
}

int ufvm_server_submit_step(ufvm_server *srv, int32_t cas, int32_t nsteps, double *rms,
                            int64_t *ticket) {
  return sguard(srv, [&] {
    if (!ticket || nsteps < 1) throw Error(UFVM_ERR_ARG, "bad argument");
    std::unique_ptr<Request> r(new Request());
    r->type = 1;
    r->id = cas;
    r->nsteps = nsteps;
    r->rms = rms;
    r->t_submit = now_s();
    std::lock_guard<std::mutex> g(srv->mu);
    if (srv->stop) throw Error(UFVM_ERR_STATE, "server stopping");
    auto lc = srv->case_of.find(cas);
    if (lc == srv->case_of.end()) throw Error(UFVM_ERR_ARG, "unknown case");
    r->ticket = srv->next_ticket++;
    *ticket = r->ticket;
    srv->pend_step[lc->second.first].push_back(r.get());
    srv->reqs[r->ticket] = std::move(r);
    srv->cv_work.notify_all();
  });
}

int ufvm_server_query(ufvm_server *srv, int64_t ticket, ufvm_request_info *info) {
  return sguard(srv, [&] {
    std::lock_guard<std::mutex> g(srv->mu);
    fill_info(find(srv, ticket), info);
  });
}

int ufvm_server_wait(ufvm_server *srv, int64_t ticket, ufvm_request_info *info) {
  return sguard(srv, [&] {
    std::unique_lock<std::mutex> g(srv->mu);
    Request *r = find(srv, ticket);
    srv->cv_done.wait(g, [&] { return r->status == 2 || r->status < 0; });
    fill_info(r, info);
    if (r->status < 0) srv->err = r->err;   // the failed request's message
  });
}

int ufvm_server_release(ufvm_server *srv, int64_t ticket) {
  return sguard(srv, [&] {
    std::lock_guard<std::mutex> g(srv->mu);
    Request *r = find(srv, ticket);
    if (r->status == 0 || r->status == 1) throw Error(UFVM_ERR_STATE, "request not finished");
    srv->reqs.erase(ticket);
  });
}

int ufvm_server_stats(ufvm_server *srv, int32_t *queued, int32_t *active, int64_t *work_bytes) {
  return sguard(srv, [&] {
    std::lock_guard<std::mutex> g(srv->mu);
    size_t steps = 0;
    for (const auto &q : srv->pend_step) steps += q.size();
    if (queued) *queued = (int32_t)(srv->pend_solve.size() + steps);
    if (active) *active = srv->n_active;
    if (work_bytes) *work_bytes = (int64_t)srv->work_bytes_held;
  });
}

}  // extern "C"
