"""GPU parity of the AMG pressure-solve path against the oracle, through the C ABI.

Bit-exact: residuals, agglomeration hierarchy (parents, coarse cconn, coarse
coefficients), lexicographic SOR / DILU sweeps, restriction/prolongation, and
whole stationary AMG solves.  Precision form: the FP64 V-cycle matches the
oracle's to 1e-12 (the dense coarsest inverse differs in rounding), FPCG
solutions match direct solves to the stated tolerances.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle
from matlab_ufvm_amd.workloads import cavity_mesh, gaussian_rhs, poisson_system, poisson_tet_mesh

pytestmark = pytest.mark.gpu


def to_scipy(ptr, col, ac, anb):
    n = len(ptr) - 1
    return sp.csr_matrix((anb, col, ptr), shape=(n, n)) + sp.diags(ac)


CASES = {
    # 3-D jittered, renumbered hex cavity; Dirichlet on the lid
    "hex3d": lambda: (cavity_mesh(12, jitter=0.15, seed=31, renumber_seed=32), ("movingWall",)),
    # large enough for several 1024-row blocks on the first coarse level
    "hex3d24": lambda: (cavity_mesh(24, jitter=0.15, seed=36, renumber_seed=37), ("movingWall",)),
    # 40x40x1 plain cavity (lexicographic order)
    "hex2d": lambda: (cavity_mesh(40, nz=1, jitter=0.0, renumber_seed=None), ("movingWall",)),
    # jittered tets, Dirichlet on x=0
    "tet": lambda: (poisson_tet_mesh(6, jitter=0.1, seed=33, renumber_seed=34), ("dirichlet",)),
}


@pytest.fixture(scope="module", params=sorted(CASES))
def case(request):
    m, dir_ = CASES[request.param]()
    ptr, col, ac, anb = poisson_system(m, dirichlet=dir_)
    b = gaussian_rhs(m, seed=35)
    print(f"\n[case {request.param}] n={m.n_cells} nnz={col.size}", flush=True)
    return request.param, m, ptr, col, ac, anb, b


def make(ctx, ptr, col, ac, anb, b):
    sid = ctx.system_create(ptr, col)
    ctx.set_coefficients(sid, ac, anb, b)
    return sid


def test_residual_bitexact(ctx, case):
    _, m, ptr, col, ac, anb, b = case
    sid = make(ctx, ptr, col, ac, anb, b)
    x = np.random.default_rng(7).standard_normal(m.n_cells)
    ctx.set_dphi(sid, 0, x)
    r, l1 = ctx.residual(sid, 0)
    S = oracle.System(ptr, col, ac, anb, bc=b, dphi=x)
    r_ref = S.residual()
    assert np.array_equal(r, r_ref)
    assert np.isclose(l1, np.abs(r_ref).sum(), rtol=1e-13)
    ctx.system_destroy(sid)


def test_hierarchy_bitexact(ctx, case):
    _, m, ptr, col, ac, anb, b = case
    sid = make(ctx, ptr, col, ac, anb, b)
    nl = ctx.agglomerate(sid, 10)
    S = oracle.System(ptr, col, ac, anb, bc=b)
    nl_ref = S.agglomerate(10)
    assert nl == nl_ref
    for l in range(nl):
        g, o = ctx.level_get(sid, l), S.level(l)
        assert np.array_equal(g["ptr"], o["ptr"]), l
        assert np.array_equal(g["col"], o["col"]), l
        assert np.array_equal(g["ac"], o["ac"]), l
        assert np.array_equal(g["anb"], o["anb"]), l
        if l < nl - 1:
            assert np.array_equal(g["parents"], o["parents"]), l
    ctx.system_destroy(sid)


@pytest.mark.parametrize("smoother", ["SOR", "ILU", "MCGS"])
def test_solve_algebraic_system_bitexact(ctx, case, smoother):
    _, m, ptr, col, ac, anb, b = case
    sid = make(ctx, ptr, col, ac, anb, b)
    x0 = np.random.default_rng(8).standard_normal(m.n_cells) * 1e-3
    ctx.set_dphi(sid, 0, x0)
    r0, r1 = ctx.solve_algebraic_system(sid, 0, smoother, 3, rrf=0.0)
    S = oracle.System(ptr, col, ac, anb, bc=b, dphi=x0)
    q0, q1 = S.solve_algebraic_system(0, smoother, 3, rrf=0.0)
    assert np.array_equal(ctx.get_dphi(sid, 0), S.level(0)["dphi"])
    assert np.isclose(r0, q0, rtol=1e-12) and np.isclose(r1, q1, rtol=1e-12)
    # on a coarse level too (unsorted first-encounter cconn)
    ctx.agglomerate(sid, 10)
    S.agglomerate(10)
    ctx.set_bc(sid, 1, S.level(1)["ac"] * 0 + 1.0)
    S_b = np.ones(S.level(1)["n"])
    np.ctypeslib.as_array(S.H.lev[1].bc, (S.level(1)["n"],))[:] = S_b
    ctx.solve_algebraic_system(sid, 1, smoother, 2, rrf=0.0)
    S.solve_algebraic_system(1, smoother, 2, rrf=0.0)
    assert np.array_equal(ctx.get_dphi(sid, 1), S.level(1)["dphi"])
    ctx.system_destroy(sid)


@pytest.mark.parametrize("smoother", ["SOR", "ILU"])
@pytest.mark.parametrize("grid", ["2d100", "3d48"])
def test_pipelined_sweeps_beyond_lds(ctx, grid, smoother):
    # Levels with more unknowns than LDS holds and no lexicographic level wider
    # than 256 rows run their SOR sweeps in the pipelined one-workgroup
    # kernel: the 100 x 100 lexicographic grid itself (10,000 rows, levels are
    # the anti-diagonals, rows of at most 8 entries), and the first coarse
    # level of a jittered, renumbered 48^3 hex mesh (16,949 rows, levels of at
    # most 97 rows, rows of up to 23 entries: the 24-entry prefetch).  The
    # SOR sweeps and the DILU substitutions keep the oracle's bits, alone and
    # inside V-cycles.
    if grid == "2d100":
        m = cavity_mesh(100, nz=1, jitter=0.1, seed=64, renumber_seed=None)
    else:
        m = cavity_mesh(48, jitter=0.15, seed=67, renumber_seed=68)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    b = gaussian_rhs(m, seed=65)
    print(f"\n[pipelined sweeps] n={m.n_cells}", flush=True)
    sid = make(ctx, ptr, col, ac, anb, b)
    x0 = np.random.default_rng(66).standard_normal(m.n_cells) * 1e-3
    ctx.set_dphi(sid, 0, x0)
    ctx.solve_algebraic_system(sid, 0, smoother, 3, rrf=0.0)
    S = oracle.System(ptr, col, ac, anb, bc=b, dphi=x0)
    S.solve_algebraic_system(0, smoother, 3, rrf=0.0)
    assert np.array_equal(ctx.get_dphi(sid, 0), S.level(0)["dphi"])
    ctx.set_dphi(sid, 0, np.zeros(m.n_cells))
    r0, r1, nc = ctx.apply_amg(sid, "V-Cycle", max_cycles=4, termination=1e-12,
                               max_coarse_levels=10, pre_sweep=2, post_sweep=2,
                               smoother=smoother, rebuild=1)
    S2 = oracle.System(ptr, col, ac, anb, bc=b)
    q0, q1, qc = S2.apply_amg("V-Cycle", max_cycles=4, termination=1e-12, max_coarse_levels=10,
                              pre_sweep=2, post_sweep=2, smoother=smoother)
    assert nc == qc
    assert np.array_equal(ctx.get_dphi(sid, 0), S2.level(0)["dphi"])
    ctx.system_destroy(sid)


@pytest.mark.parametrize("smoother", ["SOR", "ILU"])
@pytest.mark.parametrize("n_off", [14, 22, 30])
def test_pipelined_sweeps_long_rows(ctx, n_off, smoother):
    # A banded M-matrix of 9,000 rows whose neighbours sit at n_off random
    # distances in [40, 400) on either side: lexicographic levels of at most 64 rows,
    # rows of 28, 44 or 60 entries (SOR: the 32-entry prefetch, then with the
    # rest of the row) and lower / upper lists of 14, 22 or 30 (DILU: the 16-
    # and 24-entry prefetches, then with the rest).  The sweeps keep the
    # oracle's bits.
    n = 9000
    rng = np.random.default_rng(100 + n_off)
    offs = np.sort(rng.choice(np.arange(40, 400), n_off, replace=False))
    rows, cols = [], []
    for d in offs:
        i = np.arange(n - d)
        rows += [i, i + d]
        cols += [i + d, i]
    A = sp.csr_matrix((-rng.uniform(0.1, 1.0, sum(len(r) for r in rows)),
                       (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    A = (A + A.T) * 0.5
    A.sort_indices()
    ptr, col, anb = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data
    assert np.diff(ptr).max() == 2 * n_off
    ac = -np.asarray(A.sum(axis=1)).ravel() * 1.05
    b = rng.standard_normal(n)
    sid = make(ctx, ptr, col, ac, anb, b)
    x0 = rng.standard_normal(n) * 1e-3
    ctx.set_dphi(sid, 0, x0)
    ctx.solve_algebraic_system(sid, 0, smoother, 3, rrf=0.0)
    S = oracle.System(ptr, col, ac, anb, bc=b, dphi=x0)
    S.solve_algebraic_system(0, smoother, 3, rrf=0.0)
    assert np.array_equal(ctx.get_dphi(sid, 0), S.level(0)["dphi"])
    ctx.system_destroy(sid)


@pytest.mark.parametrize("smoother", ["SOR", "ILU"])
def test_lds_sweeps_near_their_size_limit(ctx, smoother):
    # 76 x 76 = 5,776 cells: every lexicographic level is narrow (one
    # workgroup) and the unknowns fit in LDS; the sweeps keep the oracle's bits
    m = cavity_mesh(76, nz=1, jitter=0.1, seed=61, renumber_seed=None)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    b = gaussian_rhs(m, seed=62)
    sid = make(ctx, ptr, col, ac, anb, b)
    x0 = np.random.default_rng(63).standard_normal(m.n_cells) * 1e-3
    ctx.set_dphi(sid, 0, x0)
    ctx.solve_algebraic_system(sid, 0, smoother, 3, rrf=0.0)
    S = oracle.System(ptr, col, ac, anb, bc=b, dphi=x0)
    S.solve_algebraic_system(0, smoother, 3, rrf=0.0)
    assert np.array_equal(ctx.get_dphi(sid, 0), S.level(0)["dphi"])
    ctx.system_destroy(sid)


def test_mcgs_after_reagglomeration_uses_the_new_pattern(ctx):
    # a coarse level rebuilt with another pattern gets a new colouring and new
    # colour chunks (solve, re-agglomerate from other coefficients, solve)
    m = cavity_mesh(10, jitter=0.15, seed=51, renumber_seed=52)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    b = gaussian_rhs(m, seed=53)
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    first = ctx.level_get(sid, 1)
    ctx.set_bc(sid, 1, np.ones(first["ac"].size))
    ctx.solve_algebraic_system(sid, 1, "MCGS", 1, rrf=0.0)
    # strengthen every third coupling: a different aggregation
    anb2 = anb * np.where(np.arange(anb.size) % 3 == 0, 4.0, 1.0)
    ctx.set_coefficients(sid, ac, anb2, b)
    ctx.agglomerate(sid, 10)
    S = oracle.System(ptr, col, ac, anb2, bc=b)
    S.agglomerate(10)
    second = S.level(1)
    assert second["n"] != first["ac"].size or not np.array_equal(second["col"], first["col"])
    ctx.set_bc(sid, 1, np.ones(second["n"]))
    np.ctypeslib.as_array(S.H.lev[1].bc, (second["n"],))[:] = 1.0
    ctx.solve_algebraic_system(sid, 1, "MCGS", 2, rrf=0.0)
    S.solve_algebraic_system(1, "MCGS", 2, rrf=0.0)
    assert np.array_equal(ctx.get_dphi(sid, 1), S.level(1)["dphi"])
    ctx.system_destroy(sid)


@pytest.mark.parametrize("cycle,smoother,pre,post", [("V-Cycle", "SOR", 2, 2),
                                                     ("W-Cycle", "ILU", 1, 3),
                                                     ("F-Cycle", "SOR", 1, 3)])
def test_apply_amg_bitexact(ctx, case, cycle, smoother, pre, post):
    name, m, ptr, col, ac, anb, b = case
    sid = make(ctx, ptr, col, ac, anb, b)
    S = oracle.System(ptr, col, ac, anb, bc=b)
    q0, q1, qn = S.apply_amg(cycle, 8, 1e-6, 10, pre, post, smoother)
    if qn == -1:
        with pytest.raises(Exception):
            ctx.apply_amg(sid, cycle, 8, 1e-6, 10, pre, post, smoother)
        pytest.skip("fewer than 4 levels: W/F cycles unsupported, as in the reference")
    r0, r1, nc = ctx.apply_amg(sid, cycle, 8, 1e-6, 10, pre, post, smoother)
    assert nc == qn
    assert np.isclose(r0, q0, rtol=1e-12) and np.isclose(r1, q1, rtol=1e-9)
    assert np.array_equal(ctx.get_dphi(sid, 0), S.level(0)["dphi"])
    ctx.system_destroy(sid)


@pytest.mark.parametrize("coarse_smoother,sweeps", [(0, 1), (1, 1), (1, 2)])
def test_pcg_vcycle_fp64_matches_oracle(ctx, case, coarse_smoother, sweeps):
    _, m, ptr, col, ac, anb, b = case
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    ctx.pcg_setup(sid, precision=64, pre_sweep=sweeps, post_sweep=sweeps, max_coarse_cells=200,
                  coarse_smoother=coarse_smoother)
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    r = np.random.default_rng(9).standard_normal(m.n_cells)
    z = ctx.pcg_apply_vcycle(sid, r)
    z_ref = S.vcycle(r, pre=sweeps, post=sweeps, max_coarse_cells=200,
                     coarse_smoother=coarse_smoother)
    assert np.abs(z - z_ref).max() <= 1e-12 * np.abs(z_ref).max()
    # the DILU finest smoothers (reference ordering, level-scheduled; colour
    # ordering, one level per colour)
    if coarse_smoother == 1:
        zs = {}
        for sm in ("ILU", "MCILU", "MCILU0"):
            ctx.pcg_setup(sid, precision=64, pre_sweep=sweeps, post_sweep=sweeps,
                          max_coarse_cells=200, coarse_smoother=1, smoother=sm)
            zi = zs[sm] = ctx.pcg_apply_vcycle(sid, r)
            zi_ref = S.vcycle(r, pre=sweeps, post=sweeps, max_coarse_cells=200,
                              coarse_smoother=1, smoother=sm)
            assert np.abs(zi - zi_ref).max() <= 1e-12 * np.abs(zi_ref).max(), sm
        # hex and 6-tets-per-hex graphs have no triangles: ILU(0) keeps A's
        # off-diagonals and is the colour-order DILU, bit for bit
        assert np.array_equal(zs["MCILU0"], zs["MCILU"])
        ctx.pcg_setup(sid, precision=64, pre_sweep=sweeps, post_sweep=sweeps,
                      max_coarse_cells=200, coarse_smoother=coarse_smoother)
    # same colourings as the oracle's greedy order (level 1 always; the coarse
    # levels when they use the block-local multicolour smoother)
    nlev_col = len(S.colours) - 1 if coarse_smoother == 0 else 1
    for l in range(nlev_col):
        nc, c = ctx.pcg_level_colours(sid, l)
        c_ref, nc_ref = S.colours_of(l)
        assert nc == nc_ref and np.array_equal(c, c_ref)
    ctx.system_destroy(sid)


@pytest.mark.parametrize("precision,tol,err,cs,sm", [(64, 1e-12, 1e-10, 1, "MCGS"),
                                                     (32, 1e-11, 1e-8, 1, "MCGS"),
                                                     (32, 1e-11, 1e-8, 0, "MCGS"),
                                                     (32, 1e-11, 1e-8, 1, "ILU"),
                                                     (64, 1e-12, 1e-10, 1, "ILU"),
                                                     (32, 1e-11, 1e-8, 1, "MCILU"),
                                                     (64, 1e-12, 1e-10, 1, "MCILU")])
def test_pcg_solve_matches_direct(ctx, case, precision, tol, err, cs, sm):
    _, m, ptr, col, ac, anb, b = case
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    ctx.pcg_setup(sid, precision=precision, max_coarse_cells=200, tol=tol, max_iter=300,
                  coarse_smoother=cs, smoother=sm)
    it, rr = ctx.pcg_solve(sid)
    x = ctx.get_dphi(sid, 0)
    A = to_scipy(ptr, col, ac, anb)
    xs = spl.spsolve(A.tocsc(), b)
    assert rr <= tol
    assert np.linalg.norm(x - xs) <= err * np.linalg.norm(xs), (it, rr)
    ctx.system_destroy(sid)


def test_pcg_fp16_storage(ctx, case):
    # precision 16: the FP32 V-cycle with the finest level's off-diagonals
    # stored in FP16, divided by their row's diagonal.  The outer FP64 FCG
    # still reaches the tolerance and the direct solution, in about the FP32
    # form's iterations; batched solves and the other finest smoothers are
    # refused.
    from matlab_ufvm_amd._lib import UfvmError
    name, m, ptr, col, ac, anb, b = case
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    A = to_scipy(ptr, col, ac, anb)
    xs = spl.spsolve(A.tocsc(), b)
    its = {}
    for prec in (32, 16):
        ctx.pcg_setup(sid, precision=prec, max_coarse_cells=200, tol=1e-10, max_iter=300)
        it, rr = ctx.pcg_solve(sid)
        x = ctx.get_dphi(sid, 0)
        assert rr <= 1e-10, (prec, it, rr)
        assert np.linalg.norm(x - xs) <= 1e-8 * np.linalg.norm(xs), (prec, it, rr)
        its[prec] = it
    print(f"\n[fp16 storage] {name}: iterations FP32 {its[32]}, FP16 storage {its[16]}", flush=True)
    assert its[16] <= its[32] + 3
    with pytest.raises(UfvmError):
        ctx.pcg_solve_batch(sid, np.stack([b, b]))
    with pytest.raises(UfvmError):
        ctx.pcg_setup(sid, precision=16, max_coarse_cells=200, smoother="MCILU")
    ctx.system_destroy(sid)


def triangulated_grid(k, seed):
    """k x k grid graph with one diagonal per square (every square two
    triangles): symmetric negative off-diagonals, rows numbered by a seeded
    permutation, diagonal = 0.05 + sum |a_ij| (CSR, neighbours increasing)"""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(k * k)
    nb = [dict() for _ in range(k * k)]
    for i in range(k):
        for j in range(k):
            for di, dj in ((0, 1), (1, 0), (1, 1)):
                if i + di < k and j + dj < k:
                    a, b = perm[i * k + j], perm[(i + di) * k + j + dj]
                    w = -rng.uniform(0.5, 1.5)
                    nb[a][b] = nb[b][a] = w
    ptr, col, anb = [0], [], []
    for a in range(k * k):
        for b in sorted(nb[a]):
            col.append(b)
            anb.append(nb[a][b])
        ptr.append(len(col))
    ac = np.array([0.05 - sum(nb[a].values()) for a in range(k * k)])
    return np.array(ptr, np.int32), np.array(col, np.int32), ac, np.array(anb)


@pytest.mark.parametrize("precision,tol,err", [(64, 1e-12, 1e-10), (32, 1e-11, 1e-8)])
def test_ilu0_on_a_graph_with_triangles(ctx, precision, tol, err):
    # ILU(0) updates off-diagonals inside the pattern here: its V-cycle
    # matches the oracle's and differs from the colour-order DILU's; the FPCG
    # it preconditions solves the system
    ptr, col, ac, anb = triangulated_grid(48, seed=91)
    n = len(ptr) - 1
    b = np.random.default_rng(92).standard_normal(n)
    print(f"\n[triangulated grid] n={n} nnz={col.size} precision={precision}", flush=True)
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    if precision == 64:
        S = oracle.System(ptr, col, ac, anb)
        S.agglomerate(10)
        zs = {}
        for sm in ("MCILU0", "MCILU"):
            ctx.pcg_setup(sid, precision=64, pre_sweep=1, post_sweep=1, max_coarse_cells=200,
                          coarse_smoother=1, smoother=sm)
            zi = zs[sm] = ctx.pcg_apply_vcycle(sid, b)
            zi_ref = S.vcycle(b, pre=1, post=1, max_coarse_cells=200, coarse_smoother=1,
                              smoother=sm)
            assert np.abs(zi - zi_ref).max() <= 1e-12 * np.abs(zi_ref).max(), sm
        assert np.abs(zs["MCILU0"] - zs["MCILU"]).max() > 1e-6 * np.abs(zs["MCILU"]).max()
    ctx.pcg_setup(sid, precision=precision, max_coarse_cells=200, tol=tol, max_iter=300,
                  coarse_smoother=1, smoother="MCILU0")
    it, rr = ctx.pcg_solve(sid)
    x = ctx.get_dphi(sid, 0)
    xs = spl.spsolve(to_scipy(ptr, col, ac, anb).tocsc(), b)
    assert rr <= tol
    assert np.linalg.norm(x - xs) <= err * np.linalg.norm(xs), (it, rr)
    ctx.system_destroy(sid)


def test_pcg_singular_neumann(ctx):
    m = cavity_mesh(10, jitter=0.15, seed=41, renumber_seed=42)
    ptr, col, ac, anb = poisson_system(m, dirichlet=())
    b = gaussian_rhs(m, seed=43)
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    ctx.pcg_setup(sid, precision=32, max_coarse_cells=100, tol=1e-10, nullspace=-1)
    it, rr = ctx.pcg_solve(sid)
    x = ctx.get_dphi(sid, 0)
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    x_ref, _, _ = S.fpcg(b, tol=1e-13, max_coarse_cells=100, singular=1)
    # the reference uses p' relative to a reference cell (cfdCorrectPPField.m:44-50)
    d, d_ref = x - x[0], x_ref - x_ref[0]
    assert np.linalg.norm(d - d_ref) <= 1e-8 * np.linalg.norm(d_ref)
    ctx.system_destroy(sid)


@pytest.mark.parametrize("precision", [64, 32])
def test_pcg_single_level(ctx, precision):
    # every cell fits the dense coarsest level: the preconditioner is the exact
    # inverse and the FPCG still runs its operator product on level 0
    m = cavity_mesh(9, jitter=0.1, seed=44, renumber_seed=45)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    b = gaussian_rhs(m, seed=46)
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    ctx.pcg_setup(sid, precision=precision, max_coarse_cells=4096, tol=1e-12, max_iter=50)
    it, rr = ctx.pcg_solve(sid)
    x = ctx.get_dphi(sid, 0)
    xs = spl.spsolve(to_scipy(ptr, col, ac, anb).tocsc(), b)
    assert rr <= 1e-12 and it <= 4, (it, rr)
    assert np.linalg.norm(x - xs) <= 1e-10 * np.linalg.norm(xs)
    ctx.system_destroy(sid)


def test_errors_are_status_codes(ctx):
    from matlab_ufvm_amd._lib import UfvmError
    with pytest.raises(UfvmError):
        ctx.system_create(np.array([0, 1]), np.array([5]))   # column out of range
    with pytest.raises(UfvmError):
        ctx.pcg_solve(123456)                                # unknown system


def test_graph_replays_match_eager_solve(ctx):
    # FCG iterations are captured as HIP graphs and replayed; with kernel
    # profiling on they are launched eagerly.  Every replay, with the
    # program's own eager work in between, must give the eager solve's bits.
    m = cavity_mesh(16, jitter=0.15, seed=47, renumber_seed=48)
    ptr, col, ac, anb = poisson_system(m, dirichlet=())
    b = gaussian_rhs(m, seed=49)
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    ctx.pcg_setup(sid, precision=32, max_coarse_cells=200, tol=1e-9)
    ctx.prof_enable(True)
    it0, rr0 = ctx.pcg_solve(sid)
    x0 = ctx.get_dphi(sid, 0)
    ctx.prof_enable(False)
    ctx.prof_reset()
    for rep in range(4):
        ctx.residual(sid, 0)                      # eager work between replays
        it, rr = ctx.pcg_solve(sid)
        assert it == it0 and rr == rr0, rep
        assert np.array_equal(ctx.get_dphi(sid, 0), x0), rep
    # a different right-hand side through the same captured graphs
    b2 = gaussian_rhs(m, seed=50)
    ctx.set_bc(sid, 0, b2)
    ctx.pcg_solve(sid)
    x2 = ctx.get_dphi(sid, 0)
    A = to_scipy(ptr, col, ac, anb)
    bp = b2 - b2.mean()
    assert np.linalg.norm(bp - A @ x2) <= 1e-8 * np.linalg.norm(bp)
    ctx.system_destroy(sid)


def test_update_coefficients_refresh_matches_setup(ctx):
    # new values on a kept hierarchy: the GPU refresh of the solve-side copies
    # gives the same solve, bit for bit, as rebuilding them with pcg_setup
    m = cavity_mesh(14, jitter=0.15, seed=61, renumber_seed=62)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    b = gaussian_rhs(m, seed=63)
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    for sm in ("MCGS", "ILU"):
        ctx.pcg_setup(sid, precision=32, max_coarse_cells=200, tol=1e-10, smoother=sm)
        ctx.pcg_solve(sid)
        rng = np.random.default_rng(64)
        scale = 1.0 + 0.5 * rng.random(len(ac))          # new coefficients, same pattern
        anb2 = anb * np.sqrt(scale[np.repeat(np.arange(len(ac)), np.diff(ptr))] *
                             scale[col])
        ac2 = ac * scale
        ctx.update_coefficients(sid, ac2, anb2, b)
        it_r, rr_r = ctx.pcg_solve(sid)
        x_r = ctx.get_dphi(sid, 0)
        ctx.pcg_setup(sid, precision=32, max_coarse_cells=200, tol=1e-10, smoother=sm)
        it_s, rr_s = ctx.pcg_solve(sid)
        assert it_r == it_s and np.array_equal(ctx.get_dphi(sid, 0), x_r), sm
        A2 = to_scipy(ptr, col, ac2, anb2)
        assert np.linalg.norm(b - A2 @ x_r) <= 1e-9 * np.linalg.norm(b)
        ctx.update_coefficients(sid, ac, anb, b)           # back for the next smoother
    ctx.system_destroy(sid)


def test_invalid_pcg_config_keeps_previous_setup(ctx):
    # a rejected configuration leaves the previous set-up usable
    from matlab_ufvm_amd._lib import UfvmError
    m = cavity_mesh(10, jitter=0.15, seed=75, renumber_seed=76)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    b = gaussian_rhs(m, seed=77)
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    ctx.pcg_setup(sid, precision=32, max_coarse_cells=200, tol=1e-9)
    it0, rr0 = ctx.pcg_solve(sid)
    x0 = ctx.get_dphi(sid, 0)
    for bad in (dict(coarse_scale=5.0), dict(smoother="MCGS", coarse_smoother=7),
                 dict(launch_mode="graphs", nullspace=3)):
        with pytest.raises(UfvmError):
            ctx.pcg_setup(sid, precision=32, max_coarse_cells=200, tol=1e-9, **bad)
        it, rr = ctx.pcg_solve(sid)
        assert it == it0 and rr == rr0
        assert np.array_equal(ctx.get_dphi(sid, 0), x0)
    ctx.system_destroy(sid)


@pytest.mark.parametrize("prec", [32, 64])
def test_batched_solves_match_single_solves(ctx, prec):
    # k right-hand sides solved together give each one the bits, iteration
    # count and residual of its own solve; a zero right-hand side gives x = 0
    for m, dirichlet in ((cavity_mesh(14, jitter=0.15, seed=81, renumber_seed=82), ()),
                         (cavity_mesh(48, nz=1, jitter=0.1, seed=83, renumber_seed=84),
                          ("movingWall",))):
        ptr, col, ac, anb = poisson_system(m, dirichlet=dirichlet)
        B = np.stack([gaussian_rhs(m, seed=90 + q) * (1.0 + q) for q in range(8)])
        B[5] = 0.0
        sid = make(ctx, ptr, col, ac, anb, B[0])
        ctx.agglomerate(sid, 10)
        ctx.pcg_setup(sid, precision=prec, max_coarse_cells=200, tol=1e-9)
        single = []
        for q in range(8):
            ctx.set_bc(sid, 0, B[q])
            it, rr = ctx.pcg_solve(sid)
            single.append((it, rr, ctx.get_dphi(sid, 0)))
        print(f"\n[batch] n={m.n_cells} iterations {[s[0] for s in single]}", flush=True)
        for k in (2, 4, 8):
            for start in range(0, 8, k):
                X, its, rel = ctx.pcg_solve_batch(sid, B[start:start + k])
                for q in range(k):
                    it, rr, x = single[start + q]
                    assert its[q] == it and rel[q] == rr, (k, start + q)
                    assert np.array_equal(X[q], x), (k, start + q)
        assert not single[5][2].any() and single[5][0] == 0
        ctx.system_destroy(sid)


def test_batched_solve_rejects_unsupported(ctx):
    from matlab_ufvm_amd._lib import UfvmError
    m = cavity_mesh(10, jitter=0.1, seed=85, renumber_seed=86)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    b = gaussian_rhs(m, seed=87)
    sid = make(ctx, ptr, col, ac, anb, b)
    ctx.agglomerate(sid, 10)
    ctx.pcg_setup(sid, precision=32, max_coarse_cells=200)
    with pytest.raises(UfvmError):
        ctx.pcg_solve_batch(sid, np.stack([b, b, b]))          # k = 3
    ctx.pcg_setup(sid, precision=32, max_coarse_cells=200, smoother="ILU")
    with pytest.raises(UfvmError):
        ctx.pcg_solve_batch(sid, np.stack([b, b]))             # DILU finest smoother
    ctx.system_destroy(sid)


def _randomly_numbered(ptr, col, ac, anb, seed):
    # the same system with its rows in a seeded random order (entries keep
    # their order inside each row); p[new] = old
    n = len(ptr) - 1
    p = np.random.default_rng(seed).permutation(n)
    ip = np.empty(n, np.int64)
    ip[p] = np.arange(n)
    nptr = np.concatenate([[0], np.cumsum(np.diff(ptr)[p])]).astype(np.int32)
    idx = np.concatenate([np.arange(ptr[i], ptr[i + 1]) for i in p])
    return (nptr, ip[col[idx]].astype(np.int32), ac[p], anb[idx]), p


def _layout(ctx, sid):
    from matlab_ufvm_amd._lib import UfvmError
    widths = []
    for lev in range(ctx.num_levels(sid)):
        try:
            widths.append(ctx.pcg_level_layout(sid, lev)[0])
        except UfvmError:          # past the precision form's last level
            break
    return widths


@pytest.mark.parametrize("prec", [32, 64])
def test_16bit_columns_give_the_same_bits(ctx, prec):
    # The sliced-ELL level copies store a column as a base per 64-row chunk
    # plus a uint16 offset per entry wherever every chunk spans fewer than
    # 65536 columns.  Only the storage of the indices changes: every finest
    # smoother, both coarse smoothers, the batched and the grouped forms give
    # the bits of the int32 form.  The randomly numbered 137k-row system has
    # chunks spanning more than 65535 columns on its finest level, which
    # keeps int32 there while its coarse levels take 16 bits.
    systems = []
    m = cavity_mesh(20, jitter=0.15, seed=141, renumber_seed=142)
    systems.append(("hex3d", poisson_system(m, dirichlet=("movingWall",)),
                    np.stack([gaussian_rhs(m, seed=150 + q) for q in range(4)])))
    m = poisson_tet_mesh(8, jitter=0.1, seed=143, renumber_seed=144)
    systems.append(("tet", poisson_system(m, dirichlet=("dirichlet",)),
                    np.stack([gaussian_rhs(m, seed=150 + q) for q in range(4)])))
    m = cavity_mesh(370, nz=1, jitter=0.1, seed=161, renumber_seed=None)
    wide, p = _randomly_numbered(*poisson_system(m, dirichlet=()), seed=162)
    systems.append(("wide2d", wide,
                    np.stack([gaussian_rhs(m, seed=150 + q)[p] for q in range(4)])))
    sids = []
    for name, (ptr, col, ac, anb), B in systems:
        sid = make(ctx, ptr, col, ac, anb, B[0])
        ctx.agglomerate(sid, 10)
        sids.append(sid)
    for sm, cs in (("MCGS", 1), ("MCGS", 0), ("MCILU", 1), ("MCILU0", 1)):
        res = {}
        for iw in (4, 0):
            out, widths = [], {}
            for (name, _, B), sid in zip(systems, sids):
                ctx.pcg_setup(sid, precision=prec, max_coarse_cells=200, tol=1e-9, smoother=sm,
                              coarse_smoother=cs, index_width=iw)
                w = widths[name] = _layout(ctx, sid)
                if iw == 4:
                    assert set(w) <= {0, 4}, (name, w)
                else:
                    assert w[0] == (4 if name == "wide2d" else 2), (name, w)
                    # Jacobi coarse levels have sliced-ELL copies, block GS ones do not
                    assert (2 in w[1:]) == (cs == 1), (name, w)
                ctx.set_bc(sid, 0, B[0])
                it, rr = ctx.pcg_solve(sid)
                out.append((it, rr, ctx.get_dphi(sid, 0)))
                if sm == "MCGS" and cs == 1:
                    X, its, rel = ctx.pcg_solve_batch(sid, B)
                    out += [(its[q], rel[q], X[q]) for q in range(4)]
            if sm == "MCGS" and cs == 1:
                xs, its, rel = ctx.pcg_solve_group(sids, [B[1] for _, _, B in systems])
                out += [(its[q], rel[q], xs[q]) for q in range(len(sids))]
            res[iw] = out
        print(f"\n[16-bit] prec={prec} {sm}/{cs}: widths {widths} "
              f"iterations {[o[0] for o in res[0]]}", flush=True)
        assert len(res[4]) == len(res[0])
        for (it4, rr4, x4), (it2, rr2, x2) in zip(res[4], res[0]):
            assert it4 == it2 and rr4 == rr2 and np.array_equal(x4, x2), (sm, cs)
    for sid in sids:
        ctx.system_destroy(sid)
