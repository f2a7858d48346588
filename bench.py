"""Pressure-solve benchmark: AMG-PCG solves/s and latency at 2.1M cells.

Workload: the 3-D lid-driven-cavity mesh (128^3 hexahedra, nodes jittered by
seeded uniform noise of +-15% of the cell size, cells randomly permuted then
RCM-renumbered) and its all-wall (pure Neumann) pressure operator.  One step
is one pressure solve from dphi = 0 to ||r||_2 <= 1e-8 ||b||_2 with the AMG
hierarchy cached on the device.

Reported: FP32-preconditioned flexible PCG (FP32 V-cycle, FP64 outer loop)
solves/s as `value`, p50/p99 latency, the same solve with an FP64 V-cycle,
the field error of both against a tight FP64 solve, the roofline of the
dominant kernel (level-1 multicolour Gauss-Seidel passes) from HIP events on
the library stream, and the CPU oracle timed on a bounded sample.

    python bench.py [--steps K] [--warmup W] [--n 128] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "AMG-PCG pressure solves/s + p99 latency at 2M cells; FP32-precond vs FP64 @1e-8"
HBM_PEAK = 8.0e12


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_problem(n, workload, ctx):
    """The pressure-correction system the solves run on.

    pcorr-k5: the 128^3 lid-driven cavity (Re 400, nu = 1/400, lid u = 1) run
    through four SIMPLE outer iterations on the GPU with the reference's own
    pressure solver (cfdApplyAMG V-cycle, SOR 2/2, termination 0.1, bit-exact
    reference arithmetic), then the fifth iteration's momentum step and its
    pressure-correction assembly (cfdAssembleMdotTerm.m:70-250); its matrix and
    mass-imbalance right-hand side are the workload.  laplacian: the gDiff
    Laplacian of the same mesh with a Gaussian right-hand side.  tet: the
    tetrahedral Poisson problem (6 tets per hex of a jittered n^3 grid,
    Dirichlet face, Gaussian right-hand side)."""
    from matlab_ufvm_amd.workloads import (cavity_mesh, gaussian_rhs, poisson_system,
                                           poisson_tet_mesh)
    t = time.time()
    if workload == "tet":
        # unit cube, 6 tets per hex of a jittered n^3 grid, unit-diffusivity
        # Laplacian, p = 0 on the x = 0 face, zero-mean Gaussian right-hand side
        m = poisson_tet_mesh(n)
        log(f"tet mesh {n}^3 x 6: cells={m.n_cells} faces={m.n_faces} ({time.time()-t:.1f}s)")
        ptr, col, ac, anb = poisson_system(m, dirichlet=("dirichlet",))
        return m, ptr, col, ac, anb, gaussian_rhs(m, seed=2029)
    # own-face pyramid apex: with the reference's apex (global faces 1..6) the
    # warped faces of this jittered mesh displace far cells' centroids by
    # several cells and the SIMPLE momentum solve diverges
    m = cavity_mesh(n, jitter=0.15, seed=2026, renumber_seed=2027, apex="own")
    log(f"mesh {n}^3: cells={m.n_cells} faces={m.n_faces} ({time.time()-t:.1f}s)")
    if workload == "laplacian":
        ptr, col, ac, anb = poisson_system(m, dirichlet=())
        return m, ptr, col, ac, anb, gaussian_rhs(m, seed=2030)
    from matlab_ufvm_amd.simple import SimpleCase, cavity_bcs
    t = time.time()
    case = SimpleCase(ctx, m, cavity_bcs(m), nu=1.0 / 400.0, p_solver="amg", p_cycle="V-Cycle",
                      p_pre=2, p_post=2, p_smoother="SOR")
    res = case.iterate(4)
    case.cfdUpdateFields()
    case.cfdAssembleAndCorrectEquation("U")
    case.cfdAssembleEquation("p")
    psid = case.system("p")
    L = ctx.level_get(psid, 0)
    b = ctx.get_bc(psid, 0)
    case.close()
    log(f"SIMPLE k=5 pressure-correction system ({time.time()-t:.1f}s; residuals after 4 "
        f"iterations {np.array2string(res[-1], precision=3)})")
    return m, L["ptr"], L["col"], L["ac"], L["anb"], b


def copy_calibration(mib):
    """Kernels of known traffic, run alone so that a counter profile of them
    scales FETCH_SIZE / WRITE_SIZE: a device-to-device copy of mib MiB (the
    runtime's blit kernel), then the calibration library's float and double
    stream copies and float gather over buffers of the same size, 5 launches
    each."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    bufs = [C.c_void_p() for _ in range(3)]
    for p in bufs:
        assert hip.hipMalloc(C.byref(p), C.c_size_t(n)) == 0
        assert hip.hipMemset(p, 1, C.c_size_t(n)) == 0
    a, b, c = bufs
    for _ in range(5):
        assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0      # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    from matlab_ufvm_amd.build import CALIB_LIB
    cal = C.CDLL(CALIB_LIB)
    cal.ufvm_calibration_run.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                         C.c_int]
    for kind in (0, 1, 2):
        assert cal.ufvm_calibration_run(kind, a, b, c, n, 5) == 0
    for p in bufs:
        hip.hipFree(p)
    print(json.dumps({"copy_calibration_bytes": n, "launches_per_kernel": 5}), flush=True)


# bytes each calibration kernel reads and writes per launch, per byte of
# buffer (copy_calibration): blit copy, float / double stream, float gather
# (index stream + gathered values)
CALIBRATION_KERNELS = (("copyBuffer", 1.0, 1.0), ("k_cal_stream<float>", 1.0, 1.0),
                       ("k_cal_stream<double>", 1.0, 1.0), ("k_cal_gather", 2.0, 1.0))


def _pmc_load(d):
    """{(kernel, grid): [value per dispatch]} from the *counter_collection.csv
    files of a rocprofv3 --pmc run written with --output-format csv (rows of
    one dispatch and counter are summed)."""
    import csv
    import glob
    import collections
    per = collections.defaultdict(float)
    meta = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            for raw in csv.DictReader(fh):
                row = {k.lower(): v for k, v in raw.items()}
                did = (f, row["dispatch_id"])
                per[did] += float(row["counter_value"])
                meta[did] = (row["kernel_name"], int(row.get("grid_size", 0) or 0))
    out = collections.defaultdict(list)
    for did, v in per.items():
        out[meta[did]].append(v)
    return out


def trace_summary(d, regex):
    """Average kernel duration per (kernel, grid size) from the
    *kernel_trace.csv files of a rocprofv3 --kernel-trace run written with
    --output-format csv, restricted to kernel names matching regex: the
    per-level breakdown that the name-only stats table merges."""
    import csv
    import glob
    import re
    import collections
    pat = re.compile(regex)
    dur = collections.defaultdict(list)
    spans = []
    for f in glob.glob(os.path.join(d, "**", "*kernel_trace.csv"), recursive=True):
        with open(f, newline="") as fh:
            for raw in csv.DictReader(fh):
                row = {k.lower(): v for k, v in raw.items()}
                name = row["kernel_name"]
                if not pat.search(name):
                    continue
                grid = int(row.get("grid_size_x", row.get("grid_size", 0)) or 0)
                t0, t1 = int(row["start_timestamp"]), int(row["end_timestamp"])
                dur[(name, grid)].append((t1 - t0) * 1e-3)
                spans.append((t0, t1))
    # overlap: the sum of the matching kernels' durations against the length of
    # the union of their intervals (> 1: kernels ran concurrently)
    if spans:
        spans.sort()
        union, (cur0, cur1) = 0, spans[0]
        for a, b in spans[1:]:
            if a > cur1:
                union += cur1 - cur0
                cur0, cur1 = a, b
            else:
                cur1 = max(cur1, b)
        union += cur1 - cur0
        total = sum(b - a for a, b in spans)
        print(f"# {len(spans)} dispatches: kernel time {total * 1e-6:.2f} ms, busy (union of "
              f"intervals) {union * 1e-6:.2f} ms, mean concurrency {total / max(union, 1):.2f}, "
              f"span {(spans[-1][1] - spans[0][0]) * 1e-6:.2f} ms")
    print(f"{'kernel':60s} {'grid':>9s} {'dispatches':>10s} {'avg us':>9s} {'min us':>9s}"
          f" {'total ms':>9s}")
    rows = sorted(dur.items(), key=lambda kv: -sum(kv[1]))
    for (name, grid), v in rows:
        short = name.split("(")[0].replace("void ", "").replace("ufvm::k::", "")
        print(f"{short[:60]:60s} {grid:9d} {len(v):10d} {sum(v) / len(v):9.2f} {min(v):9.2f}"
              f" {sum(v) * 1e-3:9.2f}")


def pmc_summary(fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, copy_mib):
    """HBM bytes per dispatch of each kernel from two rocprofv3 --pmc runs
    (FETCH_SIZE, WRITE_SIZE; KiB per dispatch).  The counters of the
    calibration kernels of copy_calibration (copy_mib MiB buffers) give, per
    kernel, the factor between the bytes it moves and what the counter reads;
    the float gather's factors (an index stream, gathered values and a
    streamed write, as in the solver's kernels) scale the table."""
    n = copy_mib << 20
    cfe, cwr = _pmc_load(copy_fetch_dir), _pmc_load(copy_write_dir)
    factors = {}
    for tag, rd, wt in CALIBRATION_KERNELS:
        fk = [v for (k, _), vs in cfe.items() if tag in k for v in vs]
        wk = [v for (k, _), vs in cwr.items() if tag in k for v in vs]
        if not fk or not wk:
            continue
        f = rd * n / (sum(fk) / len(fk) * 1024.0)
        w = wt * n / (sum(wk) / len(wk) * 1024.0)
        factors[tag] = (f, w)
        print(f"calibration {tag:22s}: {len(fk)} dispatches of {rd * n / 2**20:.0f} MiB read, "
              f"{wt * n / 2**20:.0f} MiB written; factor FETCH_SIZE {f:.3f}, WRITE_SIZE {w:.3f}")
    fs, ws = factors.get("k_cal_gather", factors.get("copyBuffer", (1.0, 1.0)))
    scale = {"fetch": fs, "write": ws}
    print(f"table: bytes = KiB * 1024 * factor, factor FETCH_SIZE {fs:.3f}, WRITE_SIZE {ws:.3f} "
          f"(the float gather's)")
    fe, wr = _pmc_load(fetch_dir), _pmc_load(write_dir)
    print(f"{'kernel':60s} {'grid':>9s} {'dispatches':>10s} {'fetch MB':>10s} {'write MB':>10s}"
          f" {'HBM MB/dispatch':>15s}")
    rows = []
    for key, fv in fe.items():
        wv = wr.get(key, [0.0])
        fm = sum(fv) / len(fv) * 1024 * scale["fetch"] / 1e6
        wm = sum(wv) / len(wv) * 1024 * scale["write"] / 1e6
        rows.append((fm + wm, key, len(fv), fm, wm))
    for tot, (name, grid), cnt, fm, wm in sorted(rows, reverse=True):
        if tot < 0.5:
            continue
        short = name.split("(")[0].replace("void ", "").replace("ufvm::k::", "")
        print(f"{short[:60]:60s} {grid:9d} {cnt:10d} {fm:10.2f} {wm:10.2f} {tot:15.2f}")


PMC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                           "vcycle_kernels_laplacian_cavity128_pmc.txt")


def committed_traffic(kernel, grid=None):
    """HBM bytes per dispatch from the committed --pmc-summary table: the
    rows whose kernel name starts with `kernel` (and whose grid size is
    `grid`, if given), dispatch-weighted, or None.  The table's first 60
    columns hold the kernel name."""
    if not os.path.exists(PMC_PROFILE):
        return None
    tot = cnt = 0.0
    with open(PMC_PROFILE) as fh:
        for line in fh:
            name, rest = line[:60].strip(), line[60:].split()
            if len(rest) != 5 or not name.startswith(kernel):
                continue
            try:
                g, n_disp, total = int(rest[0]), int(rest[1]), float(rest[4])
            except ValueError:
                continue
            if grid is None or g == grid:
                tot += n_disp * total
                cnt += n_disp
    return tot / cnt * 1e6 if cnt else None


def serving_leg(args):
    """Serving form: a Poisson-arrival stream of pressure solves and single
    SIMPLE steps over a pool of resident jittered hex meshes (2-D and 3-D,
    cell counts 50k..4M), FP32-preconditioned AMG-PCG at 1e-8, hierarchies
    cached; returns completed requests/s and p50/p99 latency (submit ->
    result on the host) under that mixed-size load."""
    from matlab_ufvm_amd.serving import Server, request_stream, run_stream
    from matlab_ufvm_amd.simple import cavity_bcs
    from matlab_ufvm_amd.workloads import cavity_mesh, gaussian_rhs, poisson_system
    pool = []
    t = time.time()
    for spec in args.serving_meshes.split(","):
        dim, n = spec.split(":")
        n = int(n)
        m = cavity_mesh(n, nz=1 if dim == "2d" else None, jitter=0.15, seed=2026 + n,
                        renumber_seed=2027 + n, apex="own")
        pool.append((spec, m))
    log(f"serving pool: {[(s_, m.n_cells) for s_, m in pool]} ({time.time()-t:.1f}s)")
    srv = Server(max_active=args.serving_active, n_streams=args.serving_streams,
                 max_batch=args.serving_max_batch, n_steppers=args.serving_steppers,
                 max_group=args.serving_max_group, group_span=args.serving_group_span)
    t = time.time()
    sids, cids = [], []
    for spec, m in pool:
        ptr, col, ac, anb = poisson_system(m, dirichlet=())
        sids.append(srv.add_system(ptr, col, ac, anb, tol=1e-8))
        cids.append(srv.add_case(m, cavity_bcs(m), nu=1.0 / 400.0, p_cycle="V-Cycle",
                                 p_pre=2, p_post=2, p_smoother="SOR",
                                 u_smoother=args.serving_u_smoother))
    log(f"serving: systems and cases resident ({time.time()-t:.1f}s)")
    sizes = [m.n_cells for _, m in pool]
    # one right-hand side per mesh, made before the stream starts
    rhs = [gaussian_rhs(m, seed=2030 + i) for i, (_, m) in enumerate(pool)]
    # warm-up: one solve and one step per mesh (first-use set-up of the frozen
    # hierarchies and workspaces), not timed
    for i in range(len(pool)):
        srv.wait(srv.submit_solve(sids[i], rhs[i]))
        srv.wait(srv.submit_step(cids[i], 1))
    reqs = request_stream(args.serving_requests, args.serving_rate, sizes, seed=2030)
    wall = run_stream(srv, reqs, sids, cids, lambda r: rhs[r.mesh])
    srv.close()
    saturated = saturated_solves(args, pool, rhs) if args.serving_burst > 0 else None
    lat = np.array([r.latency for r in reqs])
    out = {"metric": "serving: completed requests/s and p50/p99 latency (submit -> result)",
           "value": round(len(reqs) / wall, 3), "unit": "requests/s", "n_gpus": 1,
           "requests": len(reqs), "arrival_rate": args.serving_rate, "wall_s": round(wall, 3),
           "latency_ms": {"p50": round(float(np.median(lat)) * 1e3, 2),
                          "p99": round(float(np.percentile(lat, 99)) * 1e3, 2)},
           "config": {"workload": "Poisson-arrival stream (seed 2030) of pressure solves "
                                  "(all-wall Neumann Laplacian, Gaussian RHS) and single SIMPLE "
                                  "steps (lid-driven cavity, Re 400) on a resident mesh pool",
                      "meshes": {s_: int(m.n_cells) for s_, m in pool},
                      "max_active": args.serving_active, "streams": args.serving_streams,
                      "max_batch": args.serving_max_batch,
                      "max_group": args.serving_max_group,
                      "group_span": args.serving_group_span,
                      "step_lanes": args.serving_steppers,
                      "momentum_smoother": args.serving_u_smoother,
                      "tol": 1e-8, "precision": "f32 V-cycle, f64 FPCG"},
           "data": "synthetic"}
    for kind in ("solve", "step"):
        l = np.array([r.latency for r in reqs if r.kind == kind])
        if l.size:
            out[kind] = {"count": int(l.size), "p50_ms": round(float(np.median(l)) * 1e3, 2),
                         "p99_ms": round(float(np.percentile(l, 99)) * 1e3, 2)}
    bs = [r.extra.get("batch", 1) for r in reqs if r.kind == "solve"]
    gs = [r.extra.get("group", 1) for r in reqs if r.kind == "solve"]
    if bs:
        # how many solves each solve's launches served: its batch (same
        # system) or its group (different systems)
        out["solve"]["batch_sizes"] = {str(k): bs.count(k) for k in sorted(set(bs))}
        out["solve"]["group_sizes"] = {str(k): gs.count(k) for k in sorted(set(gs))}
        shared = sum(1 for b_, g_ in zip(bs, gs) if b_ > 1 or g_ > 1)
        out["solve"]["launches_shared_frac"] = round(shared / len(bs), 3)
    if saturated:
        out["saturated_solves"] = saturated
    log(f"serving: {out['value']} requests/s, p50 {out['latency_ms']['p50']} ms, "
        f"p99 {out['latency_ms']['p99']} ms over {len(reqs)} requests")
    return out


def saturated_solves(args, pool, rhs):
    """A burst of solve requests submitted at once over the same mesh pool
    (the queue stays full, so admission forms groups of solves of different
    meshes), served with groups of up to --serving-max-group and without
    groups; batches off in both, so that only grouping differs.  Every
    grouped result is checked against the ungrouped one bit for bit."""
    from matlab_ufvm_amd.workloads import cavity_mesh, gaussian_rhs
    # eight distinct meshes of about 50k cells (four 3-D, four 2-D): solves
    # there are launch- and latency-bound, where groups pay
    fleet = []
    for dim, n in (("3d", 36), ("3d", 37), ("3d", 38), ("3d", 39),
                   ("2d", 216), ("2d", 220), ("2d", 224), ("2d", 228)):
        m = cavity_mesh(n, nz=1 if dim == "2d" else None, jitter=0.15, seed=3026 + n,
                        renumber_seed=3027 + n, apex="own")
        fleet.append((f"{dim}:{n}", m))
    frhs = [gaussian_rhs(m, seed=3030 + i) for i, (_, m) in enumerate(fleet)]
    return {"pool": burst(args, pool, rhs), "fleet_50k": burst(args, fleet, frhs, uniform=True)}


def burst(args, pool, rhs, uniform=False):
    """one saturated burst over `pool`, grouped and ungrouped (see
    saturated_solves); uniform: each request's mesh drawn uniformly from the
    pool (not by log-uniform size)"""
    from matlab_ufvm_amd.serving import Server, request_stream
    from matlab_ufvm_amd.workloads import poisson_system
    sizes = [m.n_cells for _, m in pool]
    res = {"meshes": sizes}
    xs = {}
    for label, mg, lm in (("grouped", args.serving_max_group, "auto"),
                          ("grouped_graphs", args.serving_max_group, "graphs"),
                          ("ungrouped", 1, "auto")):
        srv = Server(max_active=args.serving_active, n_streams=args.serving_streams,
                     max_batch=1, n_steppers=1, max_group=mg,
                     group_span=args.serving_group_span)
        sids = []
        for _, m in pool:
            ptr, col, ac, anb = poisson_system(m, dirichlet=())
            sids.append(srv.add_system(ptr, col, ac, anb, tol=1e-8, launch_mode=lm))
        for i in range(len(pool)):
            srv.wait(srv.submit_solve(sids[i], rhs[i]))     # first-use set-up, untimed
        reqs = request_stream(args.serving_burst, 1e9, sizes, seed=2031, p_step=0.0)
        if uniform:
            pick = np.random.default_rng(2032).integers(0, len(pool), len(reqs))
            for r, q in zip(reqs, pick):
                r.mesh = int(q)
        outs = {}
        t0 = time.perf_counter()
        tickets = [srv.submit_solve(sids[r.mesh], rhs[r.mesh]) for r in reqs]
        for r, t in zip(reqs, tickets):
            o = srv.wait(t)
            r.latency = o["queued_s"] + o["run_s"]
            outs[t] = o
        wall = time.perf_counter() - t0
        srv.close()
        lat = np.array([r.latency for r in reqs])
        gs = [o["group"] for o in outs.values()]
        xs[label] = [outs[t]["x"] for t in tickets]
        res[label] = {"solves_per_s": round(len(reqs) / wall, 2),
                      "p50_ms": round(float(np.median(lat)) * 1e3, 2),
                      "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
                      "group_sizes": {str(k): gs.count(k) for k in sorted(set(gs))}}
        log(f"saturated solves, {label}: {res[label]}")
    res["requests"] = args.serving_burst
    res["same_bits_grouped_and_alone"] = all(
        np.array_equal(a, b) and np.array_equal(a, c)
        for a, b, c in zip(xs["grouped"], xs["ungrouped"], xs["grouped_graphs"]))
    res["speedup_grouped"] = round(res["grouped"]["solves_per_s"] /
                                   res["ungrouped"]["solves_per_s"], 3)
    return res


def simple_leg(args, ctx, m, steps, p_solver, u_smoother):
    """Precision form in the SIMPLE loop: the 128^3 lid-driven cavity (Re 400)
    run for `steps` outer iterations twice, with every pressure-correction
    solve by the FP32-preconditioned AMG-PCG and by the FP64 one, both to
    1e-8; returns the time per iteration and the per-field relative L2
    difference after the iterations."""
    from matlab_ufvm_amd.simple import SimpleCase, cavity_bcs
    fields, res = {}, {}
    for prec in (32, 64):
        case = SimpleCase(ctx, m, cavity_bcs(m), nu=1.0 / 400.0, p_solver=p_solver,
                          pcg_precision=prec, pcg_tol=1e-8, u_smoother=u_smoother)
        ctx.synchronize()
        t = time.perf_counter()
        rms = case.iterate(steps)
        ctx.synchronize()
        dt = time.perf_counter() - t
        fields[prec] = {k: case.field(k) for k in ("U", "p", "mdot_f")}
        res[prec] = {"s_per_iteration": round(dt / steps, 4),
                     "final_rms": [float(v) for v in rms[-1]]}
        log(f"fp{prec}-precond SIMPLE: {dt / steps * 1e3:.1f} ms/iteration, "
            f"final rms {np.array2string(rms[-1], precision=3)}")
        case.close()
    err = {}
    for k in ("U", "p", "mdot_f"):
        a, b = fields[32][k], fields[64][k]
        err[k] = float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    out = {"metric": "SIMPLE iterations/s with FP32-preconditioned pressure solves, and per-field "
                     "relative L2 difference from the FP64-preconditioned run",
           "value": round(1.0 / res[32]["s_per_iteration"], 3), "unit": "iterations/s",
           "n_gpus": 1, "iterations": steps,
           "fp32": res[32], "fp64": res[64], "field_rel_l2_diff": err,
           "config": {"workload": f"3-D lid-driven cavity {args.n}^3 jittered hex (+-15%, seed "
                                  f"2026), permuted+RCM (seed 2027), own-face apex, Re 400, "
                                  f"SIMPLE URF 0.7/0.3, 2 pressure correctors, momentum "
                                  f"{'DILU' if u_smoother == 'ILU' else 'multicolour GS'} "
                                  f"(10 sweeps, 0.1 reduction)",
                      "cells": int(m.n_cells), "pressure_tol": 1e-8,
                      "hierarchy": ("rebuilt every pressure solve" if p_solver == "pcg"
                                    else "first solve's parents kept, Galerkin operators and "
                                         "solve-side copies refreshed on the GPU")},
           "data": "synthetic"}
    log(f"SIMPLE precision form: field differences after {steps} steps {err}")
    return out


def timed_solves(ctx, sid, steps, warmup):
    for _ in range(warmup):
        ctx.pcg_solve(sid)
    lat, its = [], []
    ctx.synchronize()          # device-wide: the timed region starts idle
    t0 = time.perf_counter()
    for _ in range(steps):
        s = time.perf_counter()
        it, rr = ctx.pcg_solve(sid)      # returns after the device finished the solve
        lat.append(time.perf_counter() - s)
        its.append(it)
    ctx.synchronize()
    total = time.perf_counter() - t0
    return total, np.array(lat), its, rr


def event_timed_solves(ctx, sid, steps, stride=5):
    """The same solves with HIP events recorded on the library's stream around
    every stride-th finest-level colour pass and operator product (kernel
    durations for the roofline); kept out of the headline timing, whose
    launches they slow."""
    ctx.prof_reset()
    ctx.prof_enable(stride)
    for _ in range(steps):
        ctx.pcg_solve(sid)
    ctx.synchronize()
    ctx.prof_enable(False)
    return ctx.prof_read("mcgs0"), ctx.prof_read("spmv0")


def batched_solves(ctx, sid, b, x_single, args):
    """k solves of the same system at once (ufvm_pcg_solve_batch), inputs and
    outputs resident in device memory: the system's right-hand side and k-1
    seeded variants of it (10% Gaussian noise of its RMS added).  Reports the
    throughput and per-batch latency, and whether the batch gave the first
    right-hand side the bits of its single solve."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    k = args.batch
    rng = np.random.default_rng(2031)
    rms = float(np.sqrt(np.mean(b * b)))
    B = np.ascontiguousarray(
        np.stack([b] + [b + 0.1 * rms * rng.standard_normal(b.size) for _ in range(k - 1)]))
    X = np.empty_like(B)
    nbytes = C.c_size_t(B.nbytes)
    bd, xd = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(bd), nbytes) == 0 and hip.hipMalloc(C.byref(xd), nbytes) == 0
    assert hip.hipMemcpy(bd, C.c_void_p(B.ctypes.data), nbytes, 1) == 0   # host to device
    ctx.pcg_solve_batch_ptr(sid, k, bd.value, xd.value)     # warm-up (workspace)
    nb = max(3, args.steps // 2)
    lat = []
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(nb):
        s = time.perf_counter()
        its, rel = ctx.pcg_solve_batch_ptr(sid, k, bd.value, xd.value)
        lat.append(time.perf_counter() - s)
    ctx.synchronize()
    total = time.perf_counter() - t0
    assert hip.hipMemcpy(C.c_void_p(X.ctypes.data), xd, nbytes, 2) == 0   # device to host
    hip.hipFree(bd)
    hip.hipFree(xd)
    same = bool(np.array_equal(X[0], x_single))
    log(f"batched k={k}: {k * nb / total:.2f} solves/s, batch p50 {np.median(lat)*1e3:.2f} ms, "
        f"iterations {its.tolist()}, first solution bit-identical to its single solve: {same}")
    return {"k": k, "solves_per_s": round(k * nb / total, 3), "batches": nb,
            "batch_ms_p50": round(float(np.median(lat)) * 1e3, 3),
            "batch_ms_p99": round(float(np.percentile(lat, 99)) * 1e3, 3),
            "iterations": [int(v) for v in its], "max_relres": float(rel.max()),
            "same_bits_as_single_solve": same,
            "inputs": "system RHS + k-1 variants (10% seeded Gaussian noise), device-resident"}


def cavity2d_bench(args):
    """The plain method, CPU-runnable: the 2-D lid-driven cavity on a 40x40x1 hex
    mesh (Re 100: lid u = 1, nu = 0.01), SIMPLE with URF 0.7 / 0.3, pressure by
    the reference's AMG V-cycle (SOR, 2 pre / 2 post sweeps), FP64, uniform zero
    initial fields; N outer iterations on the GPU (the reference's arithmetic)
    and by the oracle on one host core.  The fields must agree bit for bit."""
    from matlab_ufvm_amd._lib import Context
    from matlab_ufvm_amd.simple import SimpleCase, cavity_bcs
    from matlab_ufvm_amd.workloads import cavity_mesh
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    n_it = args.cavity2d_iterations
    m = cavity_mesh(40, nz=1, jitter=0.0, renumber_seed=None)
    amg = dict(p_cycle="V-Cycle", p_smoother="SOR", p_pre=2, p_post=2)
    ctx = Context(0)
    case = SimpleCase(ctx, m, cavity_bcs(m), nu=0.01, **amg)
    ctx.synchronize()
    t = time.perf_counter()
    rms = case.iterate(n_it)
    ctx.synchronize()
    dt_gpu = time.perf_counter() - t
    o = oracle.SimpleCase(m, oracle.cavity_bcs(m), nu=0.01, **amg)
    t = time.perf_counter()
    o.iterate(n_it)
    dt_cpu = time.perf_counter() - t
    diff = {name: float(np.abs(case.field(name) - o.field(oname)).max())
            for name, oname in (("U", "U"), ("p", "p"), ("mdot_f", "mdot"))}
    log(f"cavity 40x40x1: GPU {n_it / dt_gpu:.1f} it/s, oracle {n_it / dt_cpu:.1f} it/s, "
        f"max field differences {diff}")
    out = {"metric": "SIMPLE outer iterations/s, 40x40x1 lid-driven cavity (plain method, "
                     "reference arithmetic)",
           "value": round(n_it / dt_gpu, 3), "unit": "iterations/s", "n_gpus": 1,
           "iterations": n_it, "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "2-D lid-driven cavity 40x40x1 hex (1600 cells, 6480 faces), "
                                  "Re 100, SIMPLE URF 0.7/0.3, AMG V-cycle SOR 2/2 (termination "
                                  "0.1, hierarchy rebuilt every solve), momentum DILU 10 sweeps, "
                                  "zero initial fields"},
           "final_rms": [float(v) for v in rms[-1]],
           "max_abs_field_difference_vs_oracle": diff,
           "bit_exact_vs_oracle": all(v == 0.0 for v in diff.values()),
           "cpu_baseline": {"value": round(n_it / dt_cpu, 3), "unit": "iterations/s", "cores": 1,
                            "kind": "port",
                            "sample": f"the same {n_it} iterations by the oracle (sequential C)"}}
    print(json.dumps(out), flush=True)
    case.close()


def survey_bytes_per_iteration(levels, nlev, pre, post, cpre, cpost):
    """Algorithmic bytes of one FPCG iteration by the survey's accounting:
    B_pass(l) = i (n+1) + (i+w) z + 4 w n; restriction i n + w n + w n'; prolongation
    i n + w n' + 2 w n; V-cycle = sum over levels above the coarsest of
    (nu_pre + nu_post + 1) B_pass + restriction + prolongation, plus 8 n_L^2 + 16 n_L
    for the dense coarsest solve (a symmetric Gauss-Seidel sweep counts as 2
    passes); iteration = V-cycle (w = 4) + B_pass(0, w = 8) + 7 * 8 n_0 (vector
    passes) + 2 * 12 n_0 (precision casts).  levels: [(n, z)] of the hierarchy."""
    i = 4

    def b_pass(n, z, w):
        return i * (n + 1) + (i + w) * z + 4 * w * n
    w = 4
    total = 0.0
    for l in range(nlev - 1):
        n, z = levels[l]
        nc = levels[l + 1][0]
        if l == 0:
            nu = 2 * pre + 2 * post
        else:
            nu = (pre if cpre < 0 else cpre) + (post if cpost < 0 else cpost)
        total += (nu + 1) * b_pass(n, z, w) + (i * n + w * n + w * nc) + (i * n + w * nc + 2 * w * n)
    nL = levels[nlev - 1][0]
    total += 8 * nL * nL + 16 * nL
    n0, z0 = levels[0]
    return total + b_pass(n0, z0, 8) + 7 * 8 * n0 + 2 * 12 * n0


def plain_oracle_check(ptr, col, ac, anb, b, ncyc, dphi_gpu, smoother):
    """The oracle's cfdApplyAMG (V-cycle, smoother 2/2, FP64) on the same system
    for the same number of cycles from dphi = 0: the GPU's plain drop-in must
    give its dphi bit for bit, stall included (checker only, untimed)."""
    import oracle   # checker
    t = time.time()
    S = oracle.System(ptr, col, ac, anb, bc=b)
    r0, r1, nc = S.apply_amg("V-Cycle", max_cycles=ncyc, termination=1e-8, max_coarse_levels=10,
                             pre_sweep=2, post_sweep=2, smoother=smoother)
    x = S.level(0)["dphi"]
    out = {"cycles": int(nc - 1), "l1_residual_ratio": r1 / r0,
           "dphi_bit_identical_to_gpu": bool(np.array_equal(x, dphi_gpu)),
           "seconds": round(time.time() - t, 1)}
    log(f"plain fp64 AMG ({smoother}) on the oracle: {out}")
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(ptr, col, ac, anb, b, max_coarse, coarse_smoother, tol, singular, smoother,
                 coarse_scale, cpre, cpost):
    """The oracle (C, FP64 V-cycle FPCG, the same algorithm): one full solve of the
    same system on one core, then the oracle's OpenMP build on the host cores
    this process may use (the reported baseline); hierarchy set-up not timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    S = oracle.System(ptr, col, ac, anb)
    t = time.time()
    S.agglomerate(10)
    t_setup = time.time() - t
    kw = dict(tol=tol, max_coarse_cells=max_coarse, singular=singular,
              coarse_smoother=coarse_smoother, smoother=smoother, coarse_scale=coarse_scale,
              cpre=cpre, cpost=cpost)
    t = time.time()
    x, it, rr = S.fpcg(b, **kw)
    dt1 = time.time() - t
    threads = int(os.environ.get("OMP_NUM_THREADS", 0)) or len(os.sched_getaffinity(0))
    oracle.use_openmp(True)
    try:
        t = time.time()
        _, it_mt, rr_mt = S.fpcg(b, **kw)
        dtn = time.time() - t
    finally:
        oracle.use_openmp(False)
    log(f"cpu oracle: setup {t_setup:.1f}s; 1 thread {dt1:.2f}s, {it} iterations, relres {rr:.1e};"
        f" {threads} threads {dtn:.2f}s, {it_mt} iterations")
    return {"value": 1.0 / dtn, "unit": "solves/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"1 full FP64 FPCG solve ({it_mt} iterations) of the same {len(b)}-cell "
                      f"system by the oracle's OpenMP build, hierarchy set-up ({t_setup:.1f}s) "
                      f"excluded",
            "single_thread": {"value": 1.0 / dt1, "unit": "solves/s", "cores": 1,
                              "iterations": it}}, x


def workload_name(args):
    if args.workload == "tet":
        return (f"pressure-Poisson solve, unit cube {args.n}^3 jittered hex split into 6 tets "
                f"(seed 2028), permuted+RCM (seed 2029), unit diffusivity, p = 0 on x = 0, "
                f"zero-mean Gaussian RHS (seed 2029)")
    return (f"pressure-correction solve, 3-D lid-driven cavity {args.n}^3 jittered hex "
            f"(+-15%, seed 2026), permuted+RCM (seed 2027), own-face pyramid apex, "
            + ("matrix and RHS of SIMPLE outer iteration 5 (Re 400)"
               if args.workload == "pcorr-k5" else
               "all-wall Neumann gDiff Laplacian, Gaussian RHS (seed 2030)"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--workload", choices=["pcorr-k5", "laplacian", "tet"], default="pcorr-k5")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--coarse-scale", type=float, default=2.5,
                    help="factor on every coarse-grid correction (1: plain)")
    ap.add_argument("--coarse-pre", type=int, default=2,
                    help="sweeps below the finest level before restriction (-1: --pre)")
    ap.add_argument("--coarse-post", type=int, default=2)
    ap.add_argument("--smoother", choices=["MCGS", "ILU", "MCILU", "MCILU0"], default="MCGS",
                    help="finest-level smoother of the V-cycle")
    ap.add_argument("--pre", type=int, default=1)
    ap.add_argument("--post", type=int, default=1)
    ap.add_argument("--max-coarse", type=int, default=2000)
    ap.add_argument("--sweep", default="",
                    help="only time FP32 solves under settings scale:pre:post:cpre:cpost,...")
    ap.add_argument("--batch", type=int, default=8, choices=[1, 2, 4, 8],
                    help="also time k solves of the system at once (ufvm_pcg_solve_batch)")
    ap.add_argument("--launch-mode", choices=["auto", "eager", "graphs"], default="auto",
                    help="FCG iterations as captured HIP graphs or kernel by kernel")
    ap.add_argument("--index-width", type=int, choices=[0, 4], default=0,
                    help="sliced-ELL column indices: 0 16-bit offsets where they fit, 4 int32")
    ap.add_argument("--coarse-smoother", type=int, default=1,
                    help="levels below the first: 0 block-local MCGS, 1 damped Jacobi")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--event-steps", type=int, default=5,
                    help="solves re-run with per-kernel HIP events for the roofline")
    ap.add_argument("--plain-cycles", type=int, default=20,
                    help="timed V-cycles of the plain FP64 drop-in AMG (0: skip)")
    ap.add_argument("--simple-steps", type=int, default=0,
                    help="run N SIMPLE iterations with FP32 and FP64 pressure solves instead")
    ap.add_argument("--u-smoother", choices=["ILU", "MCGS"], default="ILU",
                    help="momentum solves of --simple-steps: the reference's DILU or multicolour GS")
    ap.add_argument("--p-solver", choices=["pcg", "pcg-frozen"], default="pcg",
                    help="SIMPLE mode: rebuild the hierarchy every solve, or keep the first")
    ap.add_argument("--serving", action="store_true",
                    help="run the serving-form stream instead of the solve benchmark")
    ap.add_argument("--no-serving-leg", action="store_true",
                    help="leave the serving-form stream out of the default line")
    ap.add_argument("--simple-leg-steps", type=int, default=50,
                    help="SIMPLE steps of the default line's precision-form leg (0: skip)")
    ap.add_argument("--serving-meshes", default="3d:37,3d:64,3d:100,3d:128,2d:224,2d:512,"
                                                "2d:1024,2d:2048")
    ap.add_argument("--serving-requests", type=int, default=500)
    ap.add_argument("--serving-steppers", type=int, default=4,
                    help="SIMPLE-step lanes of the server")
    ap.add_argument("--serving-max-batch", type=int, default=8, choices=[1, 2, 4, 8],
                    help="queued solves of one mesh admitted together as one batch")
    ap.add_argument("--serving-max-group", type=int, default=8,
                    help="queued solves of different meshes admitted together as one group "
                         "(1: none)")
    ap.add_argument("--serving-group-span", type=float, default=4.0,
                    help="largest cell-count ratio inside a group")
    ap.add_argument("--serving-burst", type=int, default=200,
                    help="solve requests of the saturated burst (0: skip)")
    ap.add_argument("--serving-u-smoother", choices=["ILU", "MCGS"], default="MCGS",
                    help="momentum solves of the SIMPLE-step requests")
    ap.add_argument("--serving-rate", type=float, default=200.0)
    ap.add_argument("--serving-active", type=int, default=8)
    ap.add_argument("--serving-streams", type=int, default=4)
    ap.add_argument("--cavity2d-iterations", type=int, default=0,
                    help="only run the 40x40x1 cavity for this many SIMPLE iterations on the "
                         "GPU and the oracle")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run 5 device copies of this many MiB (counter calibration)")
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE"),
                    help="only summarise four rocprofv3 --pmc output folders (CSV): "
                         "FETCH_SIZE and WRITE_SIZE of a run and of --copy-calibration-mib")
    ap.add_argument("--trace-summary", nargs=2, metavar=("DIR", "REGEX"),
                    help="only summarise a rocprofv3 --kernel-trace output folder (CSV): "
                         "average duration per kernel and grid size of kernels matching REGEX")
    args = ap.parse_args()
    if args.trace_summary:
        trace_summary(*args.trace_summary)
        return
    if args.pmc_summary:
        pmc_summary(*args.pmc_summary, args.copy_calibration_mib or 256)
        return
    if args.cavity2d_iterations > 0:
        cavity2d_bench(args)
        return
    if args.copy_calibration_mib > 0:
        copy_calibration(args.copy_calibration_mib)
        return
    if args.serving:
        print(json.dumps(serving_leg(args)), flush=True)
        return
    if args.simple_steps > 0:
        from matlab_ufvm_amd._lib import Context
        from matlab_ufvm_amd.workloads import cavity_mesh
        m = cavity_mesh(args.n, jitter=0.15, seed=2026, renumber_seed=2027, apex="own")
        ctx = Context(0)
        print(json.dumps(simple_leg(args, ctx, m, args.simple_steps, args.p_solver,
                                    args.u_smoother)), flush=True)
        ctx.close()
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU: the pressure solve of one mesh "
                                   "is not sharded across GPUs; use --gpus 1"}))
        sys.exit(2)

    from matlab_ufvm_amd._lib import Context, UfvmError
    ctx = Context(0)
    m, ptr, col, ac, anb, b = build_problem(args.n, args.workload, ctx)
    sid = ctx.system_create(ptr, col)
    ctx.set_coefficients(sid, ac, anb, b)
    t = time.time()
    nl = ctx.agglomerate(sid, 10)
    t_agg = time.time() - t
    level_nz = [ctx.level_info(sid, l) for l in range(nl)]
    sizes = [n for n, _ in level_nz]
    log(f"hierarchy: {nl} levels {sizes} ({t_agg:.2f}s)")

    if args.sweep:
        # FP32-preconditioned solves of the same system under several V-cycle
        # settings ("scale:pre:post:cpre:cpost", comma separated); one JSON line each
        for spec in args.sweep.split(","):
            cs, pr, po, cpr, cpo = spec.split(":")
            ctx.pcg_setup(sid, precision=32, pre_sweep=int(pr), post_sweep=int(po),
                          max_coarse_cells=args.max_coarse, tol=args.tol, max_iter=500,
                          coarse_smoother=args.coarse_smoother, smoother=args.smoother,
                          coarse_scale=float(cs), coarse_pre_sweep=int(cpr),
                          coarse_post_sweep=int(cpo))
            total, lat, its, rr = timed_solves(ctx, sid, args.steps, args.warmup)
            line = {"sweep": spec, "solves_per_s": round(args.steps / total, 2),
                    "iterations": int(its[0]), "p50_ms": round(float(np.median(lat)) * 1e3, 3)}
            log(json.dumps(line))
            print(json.dumps(line), flush=True)
        return

    results = {}
    for prec in ([32] if args.no_fp64 else [32, 64, 16]):
        t = time.time()
        ctx.pcg_setup(sid, precision=prec, pre_sweep=args.pre, post_sweep=args.post,
                      max_coarse_cells=args.max_coarse, tol=args.tol, max_iter=500,
                      coarse_smoother=args.coarse_smoother, smoother=args.smoother,
                      coarse_scale=args.coarse_scale, coarse_pre_sweep=args.coarse_pre,
                      coarse_post_sweep=args.coarse_post, launch_mode=args.launch_mode,
                      index_width=args.index_width)
        t_setup = time.time() - t
        if prec == 32:
            index_bytes = []
            for lev in range(nl):
                try:
                    index_bytes.append(ctx.pcg_level_layout(sid, lev)[0])
                except UfvmError:      # past the last level of the precision form
                    break
        total, lat, its, rr = timed_solves(ctx, sid, args.steps, args.warmup)
        x = ctx.get_dphi(sid, 0)
        res = dict(total=total, lat=lat, its=its, relres=rr, x=x, setup=t_setup)
        if prec == 32:
            res["prof"], res["prof_spmv"] = event_timed_solves(ctx, sid, args.event_steps)
            if args.batch > 1:
                res["batch"] = batched_solves(ctx, sid, b, x, args)
        results[prec] = res
        log(f"fp{prec}-precond: {args.steps / total:.2f} solves/s, iters {its[0]}, "
            f"p50 {np.median(lat)*1e3:.2f} ms, p99 {np.percentile(lat, 99)*1e3:.2f} ms,"
            f" setup {t_setup:.2f}s")

    # field error of each solve against a tight FP64 solve (shifted to p'(0) = 0 as
    # the pressure correction is, cfdCorrectPPField.m:44-50)
    ctx.pcg_setup(sid, precision=64, pre_sweep=args.pre, post_sweep=args.post,
                  max_coarse_cells=args.max_coarse, tol=min(1e-12, args.tol * 1e-2),
                  max_iter=1000, coarse_smoother=args.coarse_smoother)
    ctx.pcg_solve(sid)
    x_tight = ctx.get_dphi(sid, 0)
    singular = args.workload != "tet"

    def shift(x):
        return x - x[0] if singular else x

    d_ref = shift(x_tight)

    def ferr(x):
        return float(np.linalg.norm(shift(x) - d_ref) / np.linalg.norm(d_ref))

    # plain FP64 drop-in: the reference's stationary AMG (V-cycle, agglomeration
    # pairs, 2 lexicographic Gauss-Seidel sweeps pre and post, cfdApplyAMG.m) on
    # the GPU, bit-exact with the reference arithmetic.  On this singular
    # all-Neumann system it stalls before 1e-8 (its coarsest level is smoothed,
    # not solved, so the constant mode drifts every cycle), so a bounded number
    # of cycles is timed and the residual ratio and field error it reaches are
    # reported beside the per-cycle time.
    # Both smoothers: SOR, and ILU (the reference's DILU, its default
    # smootherType, cfdSetupSolutionMethods.m:6).
    plain = {}
    plain_dphi = {}
    if args.plain_cycles > 0:
        zero = np.zeros(m.n_cells)
        def plain_run(ncyc, smoother):
            ctx.set_dphi(sid, 0, zero)
            ctx.synchronize()
            s = time.perf_counter()
            r0, r1, nc = ctx.apply_amg(sid, "V-Cycle", max_cycles=ncyc, termination=1e-8,
                                       max_coarse_levels=10, pre_sweep=2, post_sweep=2,
                                       smoother=smoother, rebuild=0)
            ctx.synchronize()
            return time.perf_counter() - s, nc - 1, r1 / r0
        for sm in ("SOR", "ILU"):
            plain_run(1, sm)                  # level schedules built on first use
            tp, ncyc, ratio = plain_run(args.plain_cycles, sm)
            plain_dphi[sm] = ctx.get_dphi(sid, 0)
            plain[sm] = {"ms_per_cycle": round(tp / ncyc * 1e3, 3), "cycles": int(ncyc),
                         "l1_residual_ratio": ratio,
                         "field_rel_l2_err": ferr(plain_dphi[sm]),
                         "reached_1e-8": bool(ratio <= 1e-8),
                         "method": f"cfdApplyAMG V-cycle, {sm} 2/2, FP64, bit-exact reference "
                                   "arithmetic"}
            log(f"plain fp64 AMG ({sm}): {plain[sm]['ms_per_cycle']:.2f} ms/cycle, {ncyc} "
                f"cycles, L1 ratio {ratio:.2e}, field err {plain[sm]['field_rel_l2_err']:.2e}")

    r32 = results[32]
    ms0, n0, by0 = r32["prof"]
    avg_ms = ms0 / max(n0, 1)
    achieved = (by0 / max(n0, 1)) / (avg_ms * 1e-3) if n0 else 0.0
    value = args.steps / r32["total"]
    traffic = (committed_traffic("k_mcgs_sell<float")
               if args.n == 128 and args.workload != "tet" else None)
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(r32["total"] / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 V-cycle, f64 FPCG", "data": "synthetic",
        "config": {"workload": workload_name(args),
                   "cells": int(m.n_cells), "faces": int(m.n_faces), "nnz_offdiag": int(col.size),
                   "tol": args.tol, "amg_levels": sizes, "pre_sweep": args.pre,
                   "smoother": args.smoother, "coarse_scale": args.coarse_scale,
                   "coarse_sweeps": [args.coarse_pre, args.coarse_post],
                   "launch_mode": args.launch_mode,
                   "index_bytes_per_level": index_bytes,
                   "post_sweep": args.post, "max_coarse_cells": args.max_coarse,
                   "coarse_smoother": ["block-MCGS", "jacobi"][args.coarse_smoother],
                   "hierarchy_cached": True},
        "latency_ms": {"p50": round(float(np.median(r32["lat"])) * 1e3, 3),
                       "p99": round(float(np.percentile(r32["lat"], 99)) * 1e3, 3)},
        "iterations": int(r32["its"][0]),
        "field_rel_l2_err": ferr(r32["x"]),
        "setup_s": {"agglomerate": round(t_agg, 3), "pcg_setup": round(r32["setup"], 3)},
        "roofline": {"bound": "hbm", "kernel": "k_mcgs_sell<float> (finest-level colour pass)",
                     "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                     "avg_launch_us": round(avg_ms * 1e3, 2), "launches": int(n0),
                     "algorithmic_bytes_per_launch": round(by0 / max(n0, 1)),
                     "traffic": (None if traffic is None else round(traffic)),
                     "traffic_source": (None if traffic is None else
                                        "rocprofv3 FETCH_SIZE + WRITE_SIZE per dispatch, "
                                        "scaled by the factors the same counters read for "
                                        "kernels of known traffic (x2.000 / x1.000): "
                                        "profiles/" + os.path.basename(PMC_PROFILE)),
                     "measured": f"HIP events on the library stream around the back-to-back "
                                 f"colour passes of every 5th finest-level smoothing call, "
                                 f"{args.event_steps} solves re-run after the timed ones"},
    }
    # whole-solve figure by the survey's byte accounting (credits every pass
    # with a full matrix read, as the formula defines)
    nlev = next((l + 1 for l in range(nl) if sizes[l] <= args.max_coarse), nl)
    b_it = survey_bytes_per_iteration(level_nz, nlev, args.pre, args.post, args.coarse_pre,
                                      args.coarse_post)
    t_solve = float(np.median(r32["lat"]))
    ach = b_it * int(r32["its"][0]) / t_solve
    out["roofline_solve"] = {"accounting": "survey formula: (nu_pre+nu_post+1) B_pass per level "
                                           "+ restriction + prolongation + 8 n_L^2 + FP64 "
                                           "B_pass + vector passes + casts, per FPCG iteration",
                             "bytes_per_iteration": round(b_it), "iterations": int(r32["its"][0]),
                             "solve_ms": round(t_solve * 1e3, 3),
                             "achieved": round(ach / 1e9, 1), "frac": round(ach / HBM_PEAK, 4)}
    ms1, n1, by1 = r32["prof_spmv"]
    if n1:
        out["roofline_spmv"] = {"kernel": "k_spmv_dot (FP64 finest-level operator product)",
                                "achieved": round(by1 / n1 / (ms1 / n1 * 1e-3) / 1e9, 1),
                                "avg_launch_us": round(ms1 / n1 * 1e3, 2),
                                "algorithmic_bytes_per_launch": round(by1 / n1)}
    if "batch" in r32:
        out["batched"] = r32["batch"]
        out["batched"]["speedup_vs_single"] = round(r32["batch"]["solves_per_s"] / value, 3)
    if 64 in results:
        r64 = results[64]
        v64 = args.steps / r64["total"]
        out["fp64_pcg"] = {"solves_per_s": round(v64, 3), "iterations": int(r64["its"][0]),
                           "p50_ms": round(float(np.median(r64["lat"])) * 1e3, 3),
                           "p99_ms": round(float(np.percentile(r64["lat"], 99)) * 1e3, 3),
                           "field_rel_l2_err": ferr(r64["x"])}
        out["speedup_fp32_vs_fp64_pcg"] = round(value / v64, 3)
        if 16 in results:
            # the finest level's off-diagonals in FP16 (divided by their row's
            # diagonal), arithmetic in FP32: the byte lever on the ratio
            r16 = results[16]
            v16 = args.steps / r16["total"]
            out["fp16_storage_pcg"] = {
                "solves_per_s": round(v16, 3), "iterations": int(r16["its"][0]),
                "p50_ms": round(float(np.median(r16["lat"])) * 1e3, 3),
                "field_rel_l2_err": ferr(r16["x"]), "speedup_vs_fp64_pcg": round(v16 / v64, 3),
                "form": "FP32 V-cycle; finest-level off-diagonals stored as FP16 a_ij / a_ii"}
        # the target's measure: FP32-preconditioned field vs the FP64 run at the
        # same residual tolerance (shift-invariant, p'(0) = 0 as above)
        d32, d64 = shift(r32["x"]), shift(r64["x"])
        out["field_rel_l2_err_vs_fp64_same_tol"] = float(np.linalg.norm(d32 - d64) /
                                                         np.linalg.norm(d64))
    for sm, key in (("SOR", "plain_fp64_amg"), ("ILU", "plain_fp64_amg_ilu")):
        if sm in plain:
            # the plain path never reaches 1e-8 on this system; any solve of
            # it would take at least one cycle, so one cycle's time over the
            # FP32 solve's bounds the solves/s ratio from below
            plain[sm]["speedup_fp32_pcg_lower_bound"] = round(plain[sm]["ms_per_cycle"] /
                                                              out["latency_ms"]["p50"], 2)
            out[key] = plain[sm]
    if args.simple_leg_steps > 0 and args.workload == "pcorr-k5":
        out["simple_precision_form"] = simple_leg(args, ctx, m, args.simple_leg_steps,
                                                  "pcg-frozen", "ILU")
    if not args.no_cpu:
        cb, x_cpu = cpu_baseline(ptr, col, ac, anb, b, args.max_coarse, args.coarse_smoother,
                                 args.tol, int(singular), args.smoother, args.coarse_scale,
                                 args.coarse_pre, args.coarse_post)
        cb["field_rel_l2_err_vs_gpu_tight"] = ferr(x_cpu)
        out["cpu_baseline"] = cb
        for sm in plain:
            plain[sm]["oracle"] = plain_oracle_check(ptr, col, ac, anb, b, plain[sm]["cycles"],
                                                     plain_dphi[sm], sm)
    else:
        out["cpu_baseline"] = None
    ctx.close()
    if not args.no_serving_leg and args.workload == "pcorr-k5":
        out["serving"] = serving_leg(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
