"""Serving form: a request loop over meshes resident on one GPU.

`Server` wraps the ufvm_server_* C ABI (include/ufvm.h).  The request loop
itself is native: a worker thread inside libufvm.so continuously batches
pressure solves (up to `max_active` slots on the GPU at once, each advanced one
FCG iteration at a time as the GPU finishes it, spread over `n_streams`
streams; queued solves of one system share a slot as one batch of up to
`max_batch`, whose kernels read the matrices once for all of them) and admits
queued requests as slots and device memory free up; SIMPLE steps run in
`n_steppers` lanes, each with its own cases, stream and thread.

`request_stream` and `run_stream` drive the mixed workload of the serving
configuration: Poisson arrivals, pressure solves and single SIMPLE steps on a
pool of jittered hex meshes whose cell counts cover 50k..4M log-uniformly.
"""
from __future__ import annotations

import ctypes as C
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import PcgCfg, RequestInfo, ServerCfg, f64p, i32p
from .simple import case_structs

STATUS = {0: "queued", 1: "running", 2: "done"}


class Server:
    def __init__(self, device: int = 0, max_active: int = 8, n_streams: int = 4,
                 hbm_reserve: float = 0.05, max_batch: int = 8, n_steppers: int = 4,
                 max_group: int = 8, group_span: float = 4.0):
        """max_active: solves (or batches or groups of solves) on the GPU at
        once; max_batch: queued solves of one system admitted together as one
        batch (1, 2, 4 or 8); max_group: queued solves of different systems,
        within a factor group_span of each other in size, admitted together as
        one group whose iterations are one launch per phase for all of them
        (1: none).  Each solve keeps the bits of its own solve either way.
        n_steppers: SIMPLE-step lanes (each owns some cases, a stream and a
        thread)."""
        self.lib = _lib.load()
        h = C.c_void_p()
        cfg = ServerCfg(max_active, n_streams, hbm_reserve, max_batch, n_steppers, max_group,
                        group_span)
        rc = self.lib.ufvm_server_create(device, C.byref(cfg), C.byref(h))
        if rc != _lib.UFVM_OK:
            raise _lib.UfvmError(rc, "ufvm_server_create failed (no GPU visible?)")
        self.h = h
        self._live = {}      # ticket -> arrays the native side reads or writes
        self._sizes = {}     # system -> n
        self._cases = {}     # case -> (mesh, keep)

    def check(self, rc):
        if rc != _lib.UFVM_OK:
            raise _lib.UfvmError(rc, self.lib.ufvm_server_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.ufvm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # --------------------------------------------------------- registration
    def add_system(self, ptr, col, ac, anb, precision=32, tol=1e-8, pre_sweep=1, post_sweep=1,
                   max_coarse_cells=2000, max_iter=500, coarse_smoother=1, smoother=0,
                   coarse_scale=2.5, coarse_sweeps=2, launch_mode="auto") -> int:
        """launch_mode: "auto" (graphs below 600k cells for single and batched
        solves, eager groups), "eager", or "graphs" (groups captured too)"""
        ptr = np.ascontiguousarray(ptr, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        ac = np.ascontiguousarray(ac, np.float64)
        anb = np.ascontiguousarray(anb, np.float64)
        cfg = PcgCfg(precision, pre_sweep, post_sweep, coarse_smoother, max_coarse_cells, -1,
                     max_iter, tol, smoother, coarse_scale, coarse_sweeps, coarse_sweeps,
                     _lib.LAUNCH_MODE[launch_mode])
        sid = C.c_int32()
        n = len(ptr) - 1
        self.check(self.lib.ufvm_server_add_system(
            self.h, n, col.size, ptr.ctypes.data_as(i32p), col.ctypes.data_as(i32p),
            ac.ctypes.data_as(f64p), anb.ctypes.data_as(f64p), C.byref(cfg), C.byref(sid)))
        self._sizes[sid.value] = n
        return sid.value

    def add_case(self, mesh, bcs, nu, **kw) -> int:
        """A resident SIMPLE case; keyword arguments as SimpleCase.  Serving
        defaults: FP32-preconditioned AMG-PCG pressure solves on the first
        step's hierarchy ("pcg-frozen") and multicolour-GS momentum solves."""
        kw.setdefault("p_solver", "pcg-frozen")
        kw.setdefault("u_smoother", "MCGS")
        d, c, keep = case_structs(mesh, bcs, nu, **kw)
        cid = C.c_int32()
        self.check(self.lib.ufvm_server_add_case(self.h, C.byref(d), C.byref(c), C.byref(cid)))
        self._cases[cid.value] = mesh
        return cid.value

    def field(self, cas, name):
        m = self._cases[cas]
        nb = m.n_cells + m.n_bfaces
        n = {"U": nb * 3, "p": nb, "PP": nb, "DU": nb * 3, "DUT": nb * 3,
             "mdot_f": m.n_faces}[name]
        out = np.empty(n)
        self.check(self.lib.ufvm_server_get_field(self.h, cas, name.encode(),
                                                  out.ctypes.data_as(f64p), n))
        return out

    def set_field(self, cas, name, values):
        v = np.ascontiguousarray(values, np.float64).reshape(-1)
        self.check(self.lib.ufvm_server_set_field(self.h, cas, name.encode(),
                                                  v.ctypes.data_as(f64p), v.size))

    # ------------------------------------------------------------- requests
    def submit_solve(self, sid: int, b) -> int:
        b = np.ascontiguousarray(b, np.float64)
        if b.size != self._sizes[sid]:
            raise ValueError("right-hand side size does not match the system")
        x = np.empty_like(b)
        t = C.c_int64()
        self.check(self.lib.ufvm_server_submit_solve(self.h, sid, b.ctypes.data_as(f64p),
                                                     x.ctypes.data_as(f64p), C.byref(t)))
        self._live[t.value] = ("solve", b, x)
        return t.value

    def submit_step(self, cas: int, nsteps: int = 1) -> int:
        rms = np.zeros(4 * nsteps)
        t = C.c_int64()
        self.check(self.lib.ufvm_server_submit_step(self.h, cas, nsteps,
                                                    rms.ctypes.data_as(f64p), C.byref(t)))
        self._live[t.value] = ("step", rms)
        return t.value

    def _result(self, ticket, info):
        kind = self._live[ticket][0]
        out = {"status": info.status, "iterations": info.iterations, "relres": info.relres,
               "queued_s": info.queued_s, "run_s": info.run_s, "kind": kind,
               "batch": info.batch, "group": info.group}
        if info.status == 2:
            if kind == "solve":
                out["x"] = self._live[ticket][2]
            else:
                out["rms"] = self._live[ticket][1].reshape(-1, 4)
        return out

    def query(self, ticket: int):
        info = RequestInfo()
        self.check(self.lib.ufvm_server_query(self.h, ticket, C.byref(info)))
        return self._result(ticket, info)

    def wait(self, ticket: int, release: bool = True):
        info = RequestInfo()
        self.check(self.lib.ufvm_server_wait(self.h, ticket, C.byref(info)))
        out = self._result(ticket, info)
        if info.status < 0:
            msg = self.lib.ufvm_server_last_error(self.h).decode()
            raise _lib.UfvmError(-info.status, f"request {ticket} failed: {msg}")
        if release:
            self.check(self.lib.ufvm_server_release(self.h, ticket))
            del self._live[ticket]
        return out

    def stats(self):
        q, a, wb = C.c_int32(), C.c_int32(), C.c_int64()
        self.check(self.lib.ufvm_server_stats(self.h, C.byref(q), C.byref(a), C.byref(wb)))
        return {"queued": q.value, "active": a.value, "work_bytes": wb.value}


# ------------------------------------------------------------ request stream

@dataclass
class Req:
    t_arrive: float
    kind: str              # "solve" or "step"
    mesh: int              # index into the mesh pool
    seed: int
    latency: float = 0.0
    iterations: int = 0
    extra: dict = field(default_factory=dict)


def request_stream(n_req: int, rate: float, sizes, seed: int = 2030, p_step: float = 0.5):
    """Poisson arrivals at `rate` per second; each request draws a cell count
    log-uniformly on [min(sizes), max(sizes)] and takes the pool mesh nearest
    to it in log scale; half are pressure solves, half single SIMPLE steps."""
    rng = np.random.default_rng(seed)
    t = np.cumsum(rng.exponential(1.0 / rate, n_req))
    logs = np.log(np.asarray(sizes, float))
    want = rng.uniform(logs.min(), logs.max(), n_req)
    mesh = np.abs(want[:, None] - logs[None, :]).argmin(axis=1)
    kind = np.where(rng.random(n_req) < p_step, "step", "solve")
    return [Req(float(t[i]), str(kind[i]), int(mesh[i]), int(seed * 1000 + i))
            for i in range(n_req)]


def run_stream(srv: Server, reqs, systems, cases, rhs):
    """Submit each request at its arrival time (wall clock from now), then
    collect every result; fills latency (submit -> result) and iterations."""
    t0 = time.perf_counter()
    tickets = []
    for r in reqs:
        delay = t0 + r.t_arrive - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        if r.kind == "solve":
            tickets.append(srv.submit_solve(systems[r.mesh], rhs(r)))
        else:
            tickets.append(srv.submit_step(cases[r.mesh], 1))
    for r, t in zip(reqs, tickets):
        res = srv.wait(t)
        r.latency = res["queued_s"] + res["run_s"]
        r.iterations = res["iterations"]
        r.extra["batch"] = res["batch"]
        r.extra["group"] = res["group"]
    return time.perf_counter() - t0
