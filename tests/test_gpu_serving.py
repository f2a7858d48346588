"""Serving form through the C ABI: continuous batching of pressure solves and
SIMPLE steps on resident meshes.  Each result must be the same alone and in a
mix (bit for bit: a solve's arithmetic depends only on its system and
right-hand side), and the same as the single-solve entry point's."""
import numpy as np
import pytest

from matlab_ufvm_amd.serving import Server, request_stream, run_stream
from matlab_ufvm_amd.simple import cavity_bcs
from matlab_ufvm_amd.workloads import cavity_mesh, gaussian_rhs, poisson_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def meshes():
    return [cavity_mesh(14, jitter=0.15, seed=51, renumber_seed=52),
            cavity_mesh(48, nz=1, jitter=0.1, seed=53, renumber_seed=54),
            cavity_mesh(20, jitter=0.15, seed=55, renumber_seed=56)]


def _register(srv, meshes, **kw):
    sids = []
    for m in meshes:
        ptr, col, ac, anb = poisson_system(m, dirichlet=())
        sids.append(srv.add_system(ptr, col, ac, anb, max_coarse_cells=200, **kw))
    return sids


def test_solves_same_alone_and_in_mix(ctx, meshes):
    srv = Server(max_active=4, n_streams=3)
    try:
        sids = _register(srv, meshes)
        rhs = [[gaussian_rhs(m, seed=60 + 7 * i + j) for j in range(3)]
               for i, m in enumerate(meshes)]
        alone = {}
        for i in range(len(meshes)):
            for j in range(3):
                t = srv.submit_solve(sids[i], rhs[i][j])
                alone[i, j] = srv.wait(t)
                assert alone[i, j]["relres"] <= 1e-8
        print(f"\n[serving] alone: {[(k, v['iterations']) for k, v in alone.items()]}",
              flush=True)
        order = [(i, j) for j in range(3) for i in range(len(meshes))] * 2
        tickets = [(k, srv.submit_solve(sids[k[0]], rhs[k[0]][k[1]])) for k in order]
        for k, t in tickets:
            r = srv.wait(t)
            assert r["iterations"] == alone[k]["iterations"]
            assert np.array_equal(r["x"], alone[k]["x"]), k
    finally:
        srv.close()
    # the single-solve entry point gives the same bits
    m = meshes[0]
    ptr, col, ac, anb = poisson_system(m, dirichlet=())
    sid = ctx.system_create(ptr, col)
    ctx.set_coefficients(sid, ac, anb, rhs[0][0])
    ctx.agglomerate(sid, 10)
    ctx.pcg_setup(sid, precision=32, max_coarse_cells=200, tol=1e-8)
    it, rr = ctx.pcg_solve(sid)
    assert it == alone[0, 0]["iterations"]
    assert np.array_equal(ctx.get_dphi(sid, 0), alone[0, 0]["x"])
    ctx.system_destroy(sid)


def test_steps_same_alone_and_in_mix(meshes):
    m = cavity_mesh(10, jitter=0.1, seed=57, renumber_seed=58)
    srv = Server(max_active=3, n_streams=2)
    try:
        sids = _register(srv, meshes[:2])
        a = srv.add_case(m, cavity_bcs(m), nu=0.01)
        b = srv.add_case(m, cavity_bcs(m), nu=0.01)
        ra = srv.wait(srv.submit_step(a, 3))          # alone
        rhs = [gaussian_rhs(meshes[i % 2], seed=80 + i) for i in range(6)]
        ts = [srv.submit_solve(sids[i % 2], rhs[i]) for i in range(6)]
        tb = srv.submit_step(b, 3)                    # in the mix
        ts += [srv.submit_solve(sids[i % 2], rhs[i]) for i in range(6)]
        rb = srv.wait(tb)
        for t in ts:
            srv.wait(t)
        assert np.array_equal(ra["rms"], rb["rms"])
        for name in ("U", "p", "mdot_f"):
            assert np.array_equal(srv.field(a, name), srv.field(b, name)), name
    finally:
        srv.close()


def test_admission_queue_and_errors(meshes):
    srv = Server(max_active=1, n_streams=1, max_batch=1)
    try:
        sid = _register(srv, meshes[:1])[0]
        bs = [gaussian_rhs(meshes[0], seed=90 + i) for i in range(5)]
        ts = [srv.submit_solve(sid, b) for b in bs]
        st = srv.stats()
        assert st["active"] <= 1
        res = [srv.wait(t) for t in ts]
        assert all(r["relres"] <= 1e-8 for r in res)
        # one slot: the later requests queued behind the earlier ones
        assert res[-1]["queued_s"] >= res[0]["queued_s"]
        assert srv.stats()["work_bytes"] > 0
        # an unknown system fails the request, not the server
        with pytest.raises(Exception):
            srv.wait(_submit_unknown(srv, sid + 99, bs[0]))
        r = srv.wait(srv.submit_solve(sid, bs[0]))
        assert np.array_equal(r["x"], res[0]["x"])
    finally:
        srv.close()


def test_batched_admission_same_bits(meshes):
    # one slot: solves queued behind the first are admitted as batches of the
    # same system, and every result has the bits of the solve made alone
    srv = Server(max_active=1, n_streams=1, max_batch=8)
    try:
        sids = _register(srv, meshes[:2])
        rhs = [gaussian_rhs(meshes[i % 2], seed=100 + i) for i in range(18)]
        alone = []
        for i, b in enumerate(rhs):
            alone.append(srv.wait(srv.submit_solve(sids[i % 2], b)))
        ts = [srv.submit_solve(sids[i % 2], b) for i, b in enumerate(rhs)]
        batches = []
        for i, t in enumerate(ts):
            r = srv.wait(t)
            batches.append(r["batch"])
            assert r["iterations"] == alone[i]["iterations"] and r["relres"] == alone[i]["relres"]
            assert np.array_equal(r["x"], alone[i]["x"]), i
        print(f"\n[serving] batch sizes {batches}", flush=True)
        assert all(a["batch"] == 1 for a in alone)
        assert max(batches) > 1
    finally:
        srv.close()


def test_grouped_admission_same_bits(meshes):
    # one slot and no batches: solves of different systems queued behind the
    # first are admitted as groups (each phase of their iterations one launch
    # for all), and every result has the bits of the solve made alone
    srv = Server(max_active=1, n_streams=1, max_batch=1, max_group=8)
    try:
        sids = _register(srv, meshes)
        rhs = [gaussian_rhs(meshes[i % 3], seed=120 + i) for i in range(12)]
        alone = [srv.wait(srv.submit_solve(sids[i % 3], b)) for i, b in enumerate(rhs)]
        assert all(a["group"] == 1 for a in alone)
        ts = [srv.submit_solve(sids[i % 3], b) for i, b in enumerate(rhs)]
        res = [srv.wait(t) for t in ts]
        groups = [r["group"] for r in res]
        print(f"\n[serving] group sizes {groups}", flush=True)
        assert max(groups) > 1
        for a, r in zip(alone, res):
            assert r["iterations"] == a["iterations"] and r["relres"] == a["relres"]
            assert np.array_equal(r["x"], a["x"])
    finally:
        srv.close()
    # a span below the meshes' size ratios forms no group
    srv = Server(max_active=1, n_streams=1, max_batch=1, max_group=8, group_span=1.1)
    try:
        sids = _register(srv, meshes)
        ts = [srv.submit_solve(sids[i % 3], b) for i, b in enumerate(rhs)]
        res = [srv.wait(t) for t in ts]
        assert all(r["group"] == 1 for r in res)
        for a, r in zip(alone, res):
            assert np.array_equal(r["x"], a["x"])
    finally:
        srv.close()


def _submit_unknown(srv, sid, b):
    srv._sizes[sid] = b.size
    return srv.submit_solve(sid, b)


def test_request_stream_completes(meshes):
    srv = Server(max_active=4, n_streams=4)
    try:
        sids = _register(srv, meshes)
        cases = [srv.add_case(m, cavity_bcs(m), nu=0.01) for m in meshes]
        reqs = request_stream(24, rate=400.0, sizes=[m.n_cells for m in meshes], seed=7)
        wall = run_stream(srv, reqs, sids, cases,
                          lambda r: gaussian_rhs(meshes[r.mesh], seed=r.seed))
        lat = np.array([r.latency for r in reqs])
        assert np.all(lat > 0) and wall > 0
        print(f"\n[serving] {len(reqs)} requests in {wall:.3f}s, p50 "
              f"{np.median(lat)*1e3:.1f} ms, p99 {np.percentile(lat, 99)*1e3:.1f} ms", flush=True)
    finally:
        srv.close()


def test_large_meshes_same_alone_and_in_mix():
    # meshes of the serving configuration's size range (>= 50k cells, 2-D and
    # 3-D): pressure solves and SIMPLE steps give the same bits alone and in a
    # mix where they overlap on the device
    big = [cavity_mesh(40, jitter=0.15, seed=111, renumber_seed=112),            # 64,000 cells
           cavity_mesh(256, nz=1, jitter=0.1, seed=113, renumber_seed=114)]       # 65,536 cells
    print(f"\n[serving large] cells {[m.n_cells for m in big]}", flush=True)
    srv = Server(max_active=4, n_streams=4)
    try:
        sids = _register(srv, big)
        kw = dict(nu=1.0 / 400.0, p_cycle="V-Cycle", p_pre=2, p_post=2, p_smoother="SOR",
                  u_smoother="MCGS")
        a = srv.add_case(big[0], cavity_bcs(big[0]), **kw)
        b = srv.add_case(big[0], cavity_bcs(big[0]), **kw)
        rhs = [[gaussian_rhs(m, seed=120 + 5 * i + j) for j in range(2)]
               for i, m in enumerate(big)]
        alone = {(i, j): srv.wait(srv.submit_solve(sids[i], rhs[i][j]))
                 for i in range(2) for j in range(2)}
        ra = srv.wait(srv.submit_step(a, 2))
        tb = srv.submit_step(b, 2)
        ts = [((i, j), srv.submit_solve(sids[i], rhs[i][j]))
              for _ in range(2) for i in range(2) for j in range(2)]
        for k, t in ts:
            r = srv.wait(t)
            assert r["relres"] <= 1e-8
            assert r["iterations"] == alone[k]["iterations"]
            assert np.array_equal(r["x"], alone[k]["x"]), k
        rb = srv.wait(tb)
        assert np.array_equal(ra["rms"], rb["rms"])
        for name in ("U", "p", "mdot_f"):
            assert np.array_equal(srv.field(a, name), srv.field(b, name)), name
    finally:
        srv.close()
