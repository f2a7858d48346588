"""Pins of the CPU oracle (oracle/ufvm_oracle.c).

The reference has no tests or golden vectors for this path and cannot run here
(MATLAB is absent), so each oracle function is pinned by hand-worked cases,
textbook equivalences computed independently (scipy triangular solves, an
ILU(0) elimination, scipy P^T A P, spsolve) and invariants -- chosen so that a
dropped term, a wrong index or a reordered loop fails one of them.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle
from matlab_ufvm_amd.workloads import cavity_mesh, gaussian_rhs, poisson_system, poisson_tet_mesh


def chain(anb_links):
    """1-D chain: anb_links[k] couples cells k and k+1 (symmetric)."""
    n = len(anb_links) + 1
    rows = [[] for _ in range(n)]
    for k, a in enumerate(anb_links):
        rows[k].append((k + 1, a))
        rows[k + 1].append((k, a))
    for r in rows:
        r.sort()
    ptr = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    col = np.array([c for r in rows for c, _ in r], np.int32)
    anb = np.array([a for r in rows for _, a in r])
    ac = -np.array([sum(a for _, a in r) for r in rows]) + 1.0
    return ptr, col, ac, anb


def to_scipy(ptr, col, ac, anb):
    n = len(ptr) - 1
    return sp.csr_matrix((anb, col, ptr), shape=(n, n)) + sp.diags(ac)


@pytest.fixture(scope="module")
def poisson16():
    m = cavity_mesh(10, jitter=0.15, seed=11, renumber_seed=12)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    return m, ptr, col, ac, anb, gaussian_rhs(m, seed=13)


# ------------------------------------------------------------------ residual

def test_residual_summation_order():
    # row 0: bc - ac*x - anb0*x1 - anb1*x2 in cconn order; with these values
    # any other association changes the result
    ptr = np.array([0, 2, 3, 4], np.int32)
    col = np.array([1, 2, 0, 0], np.int32)
    ac = np.array([1.0, 1.0, 1.0])
    anb = np.array([1e16, -1e16, 0.0, 0.0])
    bc = np.array([1.0, 0.0, 0.0])
    S = oracle.System(ptr, col, ac, anb, bc=bc, dphi=np.array([0.0, 1.0, 1.0]))
    r = S.residual()
    # ((1 - 0) - 1e16) - (-1e16) = 0 in IEEE double (1 is lost), not 1
    assert r[0] == 0.0
    assert r[1] == 0.0 - 1.0 * 1.0 - 0.0


def test_residual_matches_scipy(poisson16):
    m, ptr, col, ac, anb, b = poisson16
    x = np.random.default_rng(0).standard_normal(m.n_cells)
    S = oracle.System(ptr, col, ac, anb, bc=b, dphi=x)
    r = S.residual()
    ref = b - to_scipy(ptr, col, ac, anb) @ x
    assert np.abs(r - ref).max() <= 1e-13 * np.abs(ref).max()


# ------------------------------------------------------------- agglomeration

def test_agglomerate_chain_two_generations():
    # uniform chain of 7: seed 0 takes 1, child 1 takes 2 -> {0,1,2};
    # seed 3 takes 4, child 4 takes 5 -> {3,4,5}; seed 6 finds nobody, is
    # released, and as an orphan joins its only assigned neighbour (parent 1)
    ptr, col, ac, anb = chain([-1.0] * 6)
    S = oracle.System(ptr, col, ac, anb)
    nl = S.agglomerate(1)
    assert nl == 2
    assert S.level(0)["parents"].tolist() == [0, 0, 0, 1, 1, 1, 1]
    C = S.level(1)
    assert C["n"] == 2 and C["col"].tolist() == [1, 0]


def test_agglomerate_threshold_is_strict():
    # -anb/maxAnb must EXCEED 0.5: a link at exactly half strength is not taken
    ptr, col, ac, anb = chain([-1.0, -0.5, -1.0])
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(1)
    assert S.level(0)["parents"].tolist() == [0, 0, 1, 1]
    ptr, col, ac, anb = chain([-1.0, -0.51, -1.0, -1.0])
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(1)
    # child 1 takes 2 (0.51 > 0.5): {0,1,2}, then seed 3 takes 4
    assert S.level(0)["parents"].tolist() == [0, 0, 0, 1, 1]


def _rows_system(rows):
    n = len(rows)
    ptr = np.cumsum([0] + [len(rows[i]) for i in range(n)]).astype(np.int32)
    col = np.array([c for i in range(n) for c, _ in rows[i]], np.int32)
    anb = np.array([a for i in range(n) for _, a in rows[i]])
    return oracle.System(ptr, col, np.full(n, 2.0), anb)


def _orphan_rows(a24):
    # seed 0 takes 3 -> {0,3}; seed 1 takes 4 -> {1,4} (row 4 sees 2 at
    # 0.1/0.25 = 0.4, row 3 sees 2 at 0.1/1.0: not taken); seed 2 finds both
    # neighbours taken, is released and handled as an orphan.
    return {0: [(3, -1.0)], 1: [(4, -0.25)], 2: [(3, -0.5), (4, a24)],
            3: [(0, -1.0), (2, -0.1)], 4: [(1, -0.25), (2, -0.1)]}


def test_agglomerate_orphan_strength_uses_neighbour_max():
    # orphan strength toward j is -anb_2j / maxAnb_j (the NEIGHBOUR's maximum):
    # toward 3: 0.5/1.0 = 0.5, toward 4: 0.2/0.25 = 0.8 -> joins the parent of 4
    # although its own coefficient toward 3 is larger
    S = _rows_system(_orphan_rows(-0.2))
    S.agglomerate(1)
    assert S.level(0)["parents"].tolist() == [0, 1, 1, 0, 1]


def test_agglomerate_orphan_tie_first_wins():
    # toward 4: 0.125/0.25 = 0.5 ties toward 3 (0.5): strictly-greater keeps the first
    S = _rows_system(_orphan_rows(-0.125))
    S.agglomerate(1)
    assert S.level(0)["parents"].tolist() == [0, 1, 0, 0, 1]


def test_agglomerate_seed_uses_own_maximum():
    # a weakly coupled seed still takes its strongest neighbour: cell 2's
    # links are both 0.3, so 0.3/maxAnb(2) = 1 > 0.5
    ptr, col, ac, anb = chain([-1.0, -0.3, -0.3, -1.0])
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(1)
    assert S.level(0)["parents"].tolist() == [0, 0, 1, 1, 1]


def test_coarse_cconn_first_encounter_order():
    # star: cell 0 linked strongly to 1 only; 0 also weakly linked to 3, then 2
    # (cconn order [1, 3, 2] is deliberately unsorted).  Parents: {0,1} -> 0,
    # 2 -> 1 (seed 2 takes its strong link 4? none) ...
    rows = {0: [(1, -1.0), (3, -0.1), (2, -0.1)], 1: [(0, -1.0)], 2: [(0, -0.1), (4, -1.0)],
            3: [(0, -0.1), (5, -1.0)], 4: [(2, -1.0)], 5: [(3, -1.0)]}
    ptr = np.cumsum([0] + [len(rows[i]) for i in range(6)]).astype(np.int32)
    col = np.array([c for i in range(6) for c, _ in rows[i]], np.int32)
    anb = np.array([a for i in range(6) for _, a in rows[i]])
    ac = np.ones(6) * 2.0
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(1)
    par = S.level(0)["parents"].tolist()
    assert par == [0, 0, 1, 2, 1, 2]
    C = S.level(1)
    # parent 0 meets parent(3)=2 before parent(2)=1
    assert C["col"][C["ptr"][0]:C["ptr"][1]].tolist() == [2, 1]


def test_hierarchy_invariants(poisson16):
    m, ptr, col, ac, anb, b = poisson16
    S = oracle.System(ptr, col, ac, anb)
    nl = S.agglomerate(10)
    assert nl >= 3
    assert S.level(nl - 1)["n"] <= 5 or nl == 11
    for l in range(nl - 1):
        F, Cc = S.level(l), S.level(l + 1)
        par = F["parents"]
        assert par.min() == 0 and par.max() == Cc["n"] - 1
        assert len(np.unique(par)) == Cc["n"]
        P = sp.csr_matrix((np.ones(F["n"]), (np.arange(F["n"]), par)), shape=(F["n"], Cc["n"]))
        Af = to_scipy(F["ptr"], F["col"], F["ac"], F["anb"])
        Ac = to_scipy(Cc["ptr"], Cc["col"], Cc["ac"], Cc["anb"])
        G = (P.T @ Af @ P).tocsr()
        scale = abs(Af).max()
        assert abs(G - Ac).max() <= 1e-13 * scale
        # coarse pattern = quotient graph, without the diagonal
        Q = (P.T @ (abs(Af) > 0).astype(float) @ P).tocoo()
        qset = {(i, j) for i, j in zip(Q.row, Q.col) if i != j}
        cset = {(i, j) for i in range(Cc["n"]) for j in Cc["col"][Cc["ptr"][i]:Cc["ptr"][i + 1]]}
        assert qset == cset
        assert abs(Ac - Ac.T).max() <= 1e-13 * scale


# ---------------------------------------------------------------- smoothers

def test_sor_is_forward_then_backward_gauss_seidel(poisson16):
    m, ptr, col, ac, anb, b = poisson16
    A = to_scipy(ptr, col, ac, anb).tocsr()
    x0 = np.random.default_rng(1).standard_normal(m.n_cells)
    S = oracle.System(ptr, col, ac, anb, bc=b, dphi=x0)
    S.solve_algebraic_system(0, "SOR", 1, rrf=0.0)
    L = sp.tril(A, 0, format="csr")
    U = sp.triu(A, 1, format="csr")
    x1 = spl.spsolve_triangular(L, b - U @ x0, lower=True)
    Lo = sp.tril(A, -1, format="csr")
    Uu = sp.triu(A, 0, format="csr")
    x2 = spl.spsolve_triangular(Uu, b - Lo @ x1, lower=False)
    assert np.abs(S.level(0)["dphi"] - x2).max() <= 1e-12 * np.abs(x2).max()


def test_mcgs_solve_is_symmetric_gauss_seidel_in_colour_order(poisson16):
    # One MCGS sweep is forward then backward Gauss-Seidel in the ordering
    # that lists the colour classes one after another (rows of one class are
    # uncoupled, so their order inside the class does not matter): scipy
    # triangular solves on the permuted matrix reproduce it.
    m, ptr, col, ac, anb, b = poisson16
    A = to_scipy(ptr, col, ac, anb).tocsr()
    n = m.n_cells
    colour = np.full(n, -1)
    for i in range(n):                          # greedy first fit in index order
        used = {colour[j] for j in col[ptr[i]:ptr[i + 1]] if colour[j] >= 0}
        c = 0
        while c in used:
            c += 1
        colour[i] = c
    order = np.argsort(colour, kind="stable")
    P = A[order][:, order]
    x0 = np.random.default_rng(2).standard_normal(n)
    S = oracle.System(ptr, col, ac, anb, bc=b, dphi=x0)
    S.solve_algebraic_system(0, "MCGS", 1, rrf=0.0)
    bp, xp = b[order], x0[order]
    x1 = spl.spsolve_triangular(sp.tril(P, 0, format="csr"), bp - sp.triu(P, 1) @ xp, lower=True)
    x2 = spl.spsolve_triangular(sp.triu(P, 0, format="csr"), bp - sp.tril(P, -1) @ x1, lower=False)
    got = S.level(0)["dphi"][order]
    assert colour.max() >= 1
    assert np.abs(got - x2).max() <= 1e-12 * np.abs(x2).max()
    # and many sweeps converge to the solution
    S.solve_algebraic_system(0, "MCGS", 4000, rrf=0.0)
    xs = spl.spsolve(A.tocsc(), b)
    assert np.abs(S.level(0)["dphi"] - xs).max() <= 1e-7 * np.abs(xs).max()


def ilu0_pivots(A):
    """ILU(0) by IKJ elimination on a dense copy, dropping fill outside the pattern."""
    D = A.toarray()
    mask = D != 0
    n = D.shape[0]
    for i in range(1, n):
        for k in range(i):
            if mask[i, k]:
                D[i, k] /= D[k, k]
                for j in range(k + 1, n):
                    if mask[i, j]:
                        D[i, j] -= D[i, k] * D[k, j]
    return np.diag(D).copy()


def test_dilu_equals_ilu0_on_triangle_free_graph_with_row_quirk():
    # 2-D grid graph (no triangles): DILU pivots = ILU(0) pivots, except that
    # the reference skips the update pushed from 1-based row N-1 into row N
    m = cavity_mesh(5, nz=1, jitter=0.1, seed=3, renumber_seed=None)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    A = to_scipy(ptr, col, ac, anb)
    S = oracle.System(ptr, col, ac, anb)
    dc = S.factorize_ilu()
    d_ref = ilu0_pivots(A)
    n = m.n_cells
    assert np.allclose(1.0 / dc[:n - 1], d_ref[:n - 1], rtol=1e-13)
    a_nm, a_mn = A[n - 1, n - 2], A[n - 2, n - 1]
    assert a_nm != 0  # rows N-1, N are coupled in this grid
    assert np.isclose(1.0 / dc[n - 1], d_ref[n - 1] + a_nm * a_mn / d_ref[n - 2], rtol=1e-13)


def test_dilu_sweep_applies_factored_preconditioner(poisson16):
    m, ptr, col, ac, anb, b = poisson16
    A = to_scipy(ptr, col, ac, anb).toarray()
    x0 = np.random.default_rng(2).standard_normal(m.n_cells)
    S = oracle.System(ptr, col, ac, anb, bc=b, dphi=x0)
    d = 1.0 / S.factorize_ilu()
    S.solve_algebraic_system(0, "ILU", 1, rrf=0.0)
    L = np.tril(A, -1)
    U = np.triu(A, 1)
    D = np.diag(d)
    M = (D + L) @ np.linalg.inv(D) @ (D + U)
    x1 = x0 + np.linalg.solve(M, b - A @ x0)
    assert np.abs(S.level(0)["dphi"] - x1).max() <= 1e-11 * np.abs(x1).max()


def colour_order(S):
    """rows by (greedy colour, index): the order of the colour-order DILU"""
    S._colour_all(1)
    c, _ = S.colours_of(0)
    return np.lexsort((np.arange(c.size), c)), c


def mcilu_matrix(A, order, dinv):
    """M = (D' + L) D'^-1 (D' + U) of the colour-order DILU, in the original
    numbering (L, U the strictly lower / upper parts in the colour order)"""
    Ap = A[np.ix_(order, order)]
    Dp = np.diag(1.0 / dinv[order])
    Mp = (Dp + np.tril(Ap, -1)) @ np.linalg.inv(Dp) @ (Dp + np.triu(Ap, 1))
    M = np.empty_like(Mp)
    M[np.ix_(order, order)] = Mp
    return M


def test_mcilu_equals_ilu0_of_colour_permuted_matrix_on_triangle_free_graph():
    # a 2-D grid has no triangles, so DILU pivots are ILU(0) pivots; in colour
    # order (no quirk row) they are those of the colour-permuted matrix
    m = cavity_mesh(6, nz=1, jitter=0.1, seed=5, renumber_seed=None)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    A = to_scipy(ptr, col, ac, anb)
    S = oracle.System(ptr, col, ac, anb)
    order, c = colour_order(S)
    dinv, _ = S.mcilu(np.zeros(m.n_cells))
    Ap = A.toarray()[np.ix_(order, order)]
    d_ref = ilu0_pivots(sp.csr_matrix(Ap))
    assert np.allclose(1.0 / dinv[order], d_ref, rtol=1e-13)
    # colour 0 rows have no earlier-colour neighbours: d = a_ii exactly
    assert np.array_equal(dinv[c == 0], 1.0 / ac[c == 0])


def test_mcilu_keeps_the_diagonal_on_tets():
    # DILU's defining property, triangles or not: diag(M) = diag(A)
    m = poisson_tet_mesh(3, jitter=0.1, seed=9, renumber_seed=10)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("dirichlet",))
    A = to_scipy(ptr, col, ac, anb).toarray()
    S = oracle.System(ptr, col, ac, anb)
    order, _ = colour_order(S)
    dinv, _ = S.mcilu(np.zeros(m.n_cells))
    M = mcilu_matrix(A, order, dinv)
    assert np.allclose(np.diag(M), np.diag(A), rtol=1e-12)
    # the tet face graph has no triangles, so the fill a_ji a_ik / d_i of the
    # factor lands only outside the pattern: M = A on every entry of A (what
    # ILU(0) guarantees) and differs from A only where A is zero
    pat = A != 0
    assert np.abs(M[pat] - A[pat]).max() <= 1e-12 * np.abs(A).max()
    assert np.abs(M[~pat]).max() > 1e-3 * np.abs(A).max()


def test_mcilu_sweep_applies_factored_preconditioner(poisson16):
    m, ptr, col, ac, anb, b = poisson16
    A = to_scipy(ptr, col, ac, anb).toarray()
    S = oracle.System(ptr, col, ac, anb)
    order, _ = colour_order(S)
    x0 = np.random.default_rng(4).standard_normal(m.n_cells)
    dinv, x1 = S.mcilu(b, x=x0)
    M = mcilu_matrix(A, order, dinv)
    ref = x0 + np.linalg.solve(M, b - A @ x0)
    assert np.abs(x1 - ref).max() <= 1e-11 * np.abs(ref).max()
    # from x = 0 the sweep is M^-1 b
    _, z = S.mcilu(b)
    assert np.abs(z - np.linalg.solve(M, b)).max() <= 1e-11 * np.abs(z).max()


def test_solve_algebraic_system_early_exit(poisson16):
    m, ptr, col, ac, anb, b = poisson16
    # the loop runs while final > rrf*initial: rrf = 1 runs no sweep at all
    S0 = oracle.System(ptr, col, ac, anb, bc=b)
    r0, r1 = S0.solve_algebraic_system(0, "SOR", 10, rrf=1.0)
    assert r0 == r1 and not S0.level(0)["dphi"].any()
    assert np.isclose(r0, np.abs(b).sum() / m.n_cells)
    # rrf just below 1: one sweep reduces the mean |r| below it, so it stops
    S1 = oracle.System(ptr, col, ac, anb, bc=b)
    r0, r1 = S1.solve_algebraic_system(0, "SOR", 10, rrf=0.999)
    S2 = oracle.System(ptr, col, ac, anb, bc=b)
    S2.solve_algebraic_system(0, "SOR", 1, rrf=0.0)
    assert np.array_equal(S1.level(0)["dphi"], S2.level(0)["dphi"])
    assert r1 < 0.999 * r0


# --------------------------------------------------------------- AMG cycles

@pytest.mark.parametrize("cycle,smoother", [("V-Cycle", "SOR"), ("W-Cycle", "ILU"),
                                             ("F-Cycle", "SOR")])
def test_apply_amg_converges_to_direct_solution(poisson16, cycle, smoother):
    m, ptr, col, ac, anb, b = poisson16
    S = oracle.System(ptr, col, ac, anb, bc=b)
    r0, r1, nc = S.apply_amg(cycle, 200, 1e-10, 10, 1, 3, smoother)
    assert r1 <= 1e-10 * r0
    xs = spl.spsolve(to_scipy(ptr, col, ac, anb).tocsc(), b)
    assert np.linalg.norm(S.level(0)["dphi"] - xs) <= 1e-8 * np.linalg.norm(xs)
    assert nc >= 2


def test_apply_amg_cycle_counter_semantics(poisson16):
    m, ptr, col, ac, anb, b = poisson16
    S = oracle.System(ptr, col, ac, anb, bc=b)
    _, _, nc = S.apply_amg("V-Cycle", 1, 1e-12, 10, 1, 3, "ILU")
    assert nc == 2  # nCycle is incremented after the last cycle (cfdApplyAMG.m:37,59)


def test_w_cycle_needs_four_levels():
    ptr, col, ac, anb = chain([-1.0] * 6)
    S = oracle.System(ptr, col, ac, anb, bc=np.ones(7))
    _, _, nc = S.apply_amg("W-Cycle", 5, 0.1, 10, 1, 3, "SOR")
    assert nc == -1


# ------------------------------------------------------- precision-form path

def test_colouring_is_proper(poisson16):
    m, ptr, col, ac, anb, b = poisson16
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    S.vcycle(b)
    for l in range(len(S.colours)):
        c, nc = S.colours_of(l)
        L = S.level(l)
        rows = np.repeat(np.arange(L["n"]), np.diff(L["ptr"]))
        assert np.all(c[rows] != c[L["col"]])
        assert c.max() == nc - 1


def test_vcycle_operator_is_symmetric(poisson16):
    # symmetric A, symmetric MCGS with pre == post, exact coarsest solve:
    # the V-cycle is a symmetric linear operator
    m, ptr, col, ac, anb, b = poisson16
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    rng = np.random.default_rng(5)
    y, z = rng.standard_normal(m.n_cells), rng.standard_normal(m.n_cells)
    My, Mz = S.vcycle(y, max_coarse_cells=200), S.vcycle(z, max_coarse_cells=200)
    assert abs(z @ My - y @ Mz) <= 1e-12 * abs(z @ My)
    # and linear
    assert np.allclose(S.vcycle(2 * y - z, max_coarse_cells=200), 2 * My - Mz, rtol=1e-12,
                       atol=1e-12 * np.abs(My).max())


@pytest.mark.parametrize("coarse_smoother", [0, 1])
def test_vcycle_symmetric_with_block_smoothed_levels(coarse_smoother):
    # 13824 cells: the first coarse level has more than one 1024-row block, so
    # the block-local smoother's Jacobi coupling is exercised
    m = cavity_mesh(24, jitter=0.15, seed=36, renumber_seed=37)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("movingWall",))
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    assert S.level(1)["n"] > oracle.BLOCK_ROWS
    rng = np.random.default_rng(6)
    y, z = rng.standard_normal(m.n_cells), rng.standard_normal(m.n_cells)
    for pre in (1, 2):
        My = S.vcycle(y, pre=pre, post=pre, max_coarse_cells=100, coarse_smoother=coarse_smoother)
        Mz = S.vcycle(z, pre=pre, post=pre, max_coarse_cells=100, coarse_smoother=coarse_smoother)
        assert abs(z @ My - y @ Mz) <= 1e-11 * abs(z @ My)
        assert y @ My > 0   # positive definite preconditioner


@pytest.mark.parametrize("smoother,sweeps,w", [("MCGS", 1, 1.0), ("MCGS", 2, 2.0),
                                               ("ILU", 1, 2.0), ("ILU", 2, 1.5),
                                               ("MCILU", 1, 2.5), ("MCILU", 2, 1.0)])
def test_two_level_vcycle_matches_dense_formula(poisson16, smoother, sweeps, w):
    # With the coarse level solved exactly the V-cycle is, in matrices,
    #   x1 = S^pre(0; b), xc = Ac^-1 P^T (b - A x1), x2 = x1 + w P xc, z = S^post(x2; b)
    # where one smoothing sweep from x is x + M^-1 (b - A x):
    #   MCGS: M = D + L_c (colours up) then D + U_c (colours down), L_c / U_c the
    #         couplings to lower / higher colour classes
    #   ILU:  M = (D' + L) D'^-1 (D' + U), D' the DILU diagonal (pinned above)
    #   MCILU: the same in colour order (pinned above)
    # Ac = P^T A P from the parents and w the coarse-correction scale.
    m, ptr, col, ac, anb, b = poisson16
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    n1 = S.level(1)["n"]
    A = to_scipy(ptr, col, ac, anb).toarray()
    par = S.level(0)["parents"]
    P = np.zeros((m.n_cells, n1))
    P[np.arange(m.n_cells), par] = 1.0
    Ac = P.T @ A @ P
    if smoother == "MCGS":
        S._colour_all(2)
        c, _ = S.colours_of(0)
        D = np.diag(np.diag(A))
        off = A - D
        lower = off * (c[None, :] < c[:, None])
        upper = off * (c[None, :] > c[:, None])
        Ms = [D + lower, D + upper]
    elif smoother == "ILU":
        Sd = oracle.System(ptr, col, ac, anb, bc=b)
        Dp = np.diag(1.0 / Sd.factorize_ilu())
        Ms = [(Dp + np.tril(A, -1)) @ np.linalg.inv(Dp) @ (Dp + np.triu(A, 1))]
    else:
        Sd = oracle.System(ptr, col, ac, anb)
        order, _ = colour_order(Sd)
        dinv, _ = Sd.mcilu(b)
        Ms = [mcilu_matrix(A, order, dinv)]

    def sweep(x):
        for M in Ms:
            x = x + np.linalg.solve(M, b - A @ x)
        return x

    x = np.zeros(m.n_cells)
    for _ in range(sweeps):
        x = sweep(x)
    x = x + w * (P @ np.linalg.solve(Ac, P.T @ (b - A @ x)))
    for _ in range(sweeps):
        x = sweep(x)
    z = S.vcycle(b, pre=sweeps, post=sweeps, max_coarse_cells=n1, smoother=smoother,
                 coarse_scale=w)
    assert np.abs(z - x).max() <= 1e-10 * np.abs(x).max()


@pytest.mark.parametrize("smoother", ["MCGS", "ILU", "MCILU", "MCILU0"])
def test_vcycle_symmetric_for_each_finest_smoother(poisson16, smoother):
    m, ptr, col, ac, anb, b = poisson16
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    rng = np.random.default_rng(7)
    y, z = rng.standard_normal(m.n_cells), rng.standard_normal(m.n_cells)
    My = S.vcycle(y, max_coarse_cells=50, smoother=smoother)
    Mz = S.vcycle(z, max_coarse_cells=50, smoother=smoother)
    assert abs(z @ My - y @ Mz) <= 1e-11 * abs(z @ My)
    assert y @ My > 0


@pytest.mark.parametrize("smoother", ["MCGS", "ILU", "MCILU", "MCILU0"])
def test_fpcg_matches_direct_solve(poisson16, smoother):
    m, ptr, col, ac, anb, b = poisson16
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    x, it, rr = S.fpcg(b, tol=1e-12, max_coarse_cells=200, smoother=smoother)
    xs = spl.spsolve(to_scipy(ptr, col, ac, anb).tocsc(), b)
    assert rr <= 1e-12
    assert np.linalg.norm(x - xs) <= 1e-10 * np.linalg.norm(xs)
    assert it < 60


def test_fpcg_singular_neumann():
    m = cavity_mesh(8, jitter=0.15, seed=21, renumber_seed=22)
    ptr, col, ac, anb = poisson_system(m, dirichlet=())
    b = gaussian_rhs(m, seed=23)
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    x, it, rr = S.fpcg(b, tol=1e-11, max_coarse_cells=100, singular=1)
    A = to_scipy(ptr, col, ac, anb)
    assert np.linalg.norm(b - A @ x) <= 1e-10 * np.linalg.norm(b)
    assert abs(x.mean()) <= 1e-8 * np.abs(x).max()


def test_openmp_build_solves_the_same_system(poisson16):
    # the OpenMP build (CPU-baseline timing only) runs the same FPCG: the same
    # iteration count to within one and a solution within the tolerance;
    # the sequential build stays the default
    m, ptr, col, ac, anb, b = poisson16
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    x1, it1, rr1 = S.fpcg(b, tol=1e-10, max_coarse_cells=200)
    oracle.use_openmp(True)
    try:
        x2, it2, rr2 = S.fpcg(b, tol=1e-10, max_coarse_cells=200)
    finally:
        oracle.use_openmp(False)
    assert abs(it1 - it2) <= 1 and rr2 <= 1e-10
    assert np.linalg.norm(x1 - x2) <= 1e-8 * np.linalg.norm(x1)
    x3, it3, rr3 = S.fpcg(b, tol=1e-10, max_coarse_cells=200)
    assert it3 == it1 and np.array_equal(x3, x1)


def test_ilu0_equals_dilu_without_triangles(poisson16):
    # a hex mesh's graph has no triangles: no fill inside the pattern, so the
    # ILU(0) factor keeps every off-diagonal and its diagonal is the colour-order
    # DILU's, bit for bit; the V-cycles agree bit for bit
    m, ptr, col, ac, anb, b = poisson16
    S = oracle.System(ptr, col, ac, anb)
    _, dinv0, F = S.mcilu0_factor()
    dinv, _ = S.mcilu(b)
    assert np.array_equal(F, anb) and np.array_equal(dinv0, dinv)
    S.agglomerate(10)
    assert np.array_equal(S.vcycle(b, max_coarse_cells=50, smoother="MCILU0"),
                          S.vcycle(b, max_coarse_cells=50, smoother="MCILU"))


def _factor_product(ptr, col, c, dinv, F):
    """L U of a colour-order factor: L = I + L~ D^-1 strictly lower in the
    colour order, U = D + U~; also the rows / low mask of the entries"""
    n = len(dinv)
    order = np.lexsort((np.arange(n), c))
    pos = np.empty(n, np.int64)
    pos[order] = np.arange(n)
    rows = np.repeat(np.arange(n), np.diff(ptr))
    Ld, Ud = np.eye(n), np.diag(1.0 / dinv)
    low = pos[col] < pos[rows]
    Ld[rows[low], col[low]] = F[low] * dinv[col[low]]
    Ud[rows[~low], col[~low]] = F[~low]
    return Ld @ Ud, rows


def test_ilu0_on_tets_is_the_colour_order_dilu():
    # the face-adjacency graph of the 6-tets-per-hex meshes has no triangles
    # (two face neighbours of a tet never share a face), so ILU(0) makes no
    # off-diagonal change and equals the colour-order DILU bit for bit; its
    # product L U reproduces A on the pattern, the defining property of ILU(0)
    m = poisson_tet_mesh(4, seed=71, renumber_seed=72)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("dirichlet",))
    G = sp.csr_matrix((np.ones(col.size), col, ptr), shape=(m.n_cells,) * 2)
    assert (G @ G).multiply(G).sum() == 0          # no triangles
    S = oracle.System(ptr, col, ac, anb)
    c, dinv, F = S.mcilu0_factor()
    dd, _ = S.mcilu(np.zeros(m.n_cells))
    assert np.array_equal(F, anb) and np.array_equal(dinv, dd)
    LU, rows = _factor_product(ptr, col, c, dinv, F)
    A = to_scipy(ptr, col, ac, anb).toarray()
    scale = np.abs(A).max()
    assert np.abs(LU[rows, col] - A[rows, col]).max() <= 1e-12 * scale
    assert np.abs(np.diag(LU) - np.diag(A)).max() <= 1e-12 * scale


def test_ilu0_factor_reproduces_the_matrix_on_a_graph_with_triangles():
    # a grid graph with one diagonal per square (every square two triangles):
    # ILU(0) now changes off-diagonals inside the pattern and L U = A there,
    # while the DILU factor's product misses A at those entries
    k = 9
    n = k * k
    idx = lambda i, j: i * k + j
    rng = np.random.default_rng(77)
    nb = [dict() for _ in range(n)]
    for i in range(k):
        for j in range(k):
            for di, dj in ((0, 1), (1, 0), (1, 1)):
                if i + di < k and j + dj < k:
                    a, b = idx(i, j), idx(i + di, j + dj)
                    w = -rng.uniform(0.5, 1.5)
                    nb[a][b] = w
                    nb[b][a] = w
    ptr, col, anb = [0], [], []
    for a in range(n):
        for b in sorted(nb[a]):
            col.append(b)
            anb.append(nb[a][b])
        ptr.append(len(col))
    ptr, col, anb = np.array(ptr, np.int32), np.array(col, np.int32), np.array(anb)
    ac = np.array([-sum(nb[a].values()) + 0.1 for a in range(n)])
    S = oracle.System(ptr, col, ac, anb)
    c, dinv, F = S.mcilu0_factor()
    A = to_scipy(ptr, col, ac, anb).toarray()
    scale = np.abs(A).max()
    LU, rows = _factor_product(ptr, col, c, dinv, F)
    assert np.abs(F - anb).max() > 1e-3 * scale
    assert np.abs(LU[rows, col] - A[rows, col]).max() <= 1e-12 * scale
    assert np.abs(np.diag(LU) - np.diag(A)).max() <= 1e-12 * scale
    dd, _ = S.mcilu(np.zeros(n))
    LUd, _ = _factor_product(ptr, col, c, dd, anb)
    assert np.abs(LUd[rows, col] - A[rows, col]).max() > 1e-3 * scale
    # applying it: one sweep from zero is M^-1 b with M = L U
    b = rng.standard_normal(n)
    S.agglomerate(10)
    x = S.vcycle(b, pre=1, post=0, max_coarse_cells=n, smoother="MCILU0", coarse_scale=1.0)
    assert x.shape == (n,)


def test_ilu0_fpcg_on_tets():
    m = poisson_tet_mesh(6, seed=73, renumber_seed=74)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("dirichlet",))
    b = gaussian_rhs(m, seed=75)
    S = oracle.System(ptr, col, ac, anb)
    S.agglomerate(10)
    its = {}
    for sm in ("MCILU", "MCILU0"):
        x, its[sm], rr = S.fpcg(b, tol=1e-11, max_coarse_cells=100, smoother=sm)
        xs = spl.spsolve(to_scipy(ptr, col, ac, anb).tocsc(), b)
        assert rr <= 1e-11 and np.linalg.norm(x - xs) <= 1e-9 * np.linalg.norm(xs)
    rng = np.random.default_rng(76)
    y, z = rng.standard_normal(m.n_cells), rng.standard_normal(m.n_cells)
    My = S.vcycle(y, max_coarse_cells=100, smoother="MCILU0")
    Mz = S.vcycle(z, max_coarse_cells=100, smoother="MCILU0")
    assert abs(z @ My - y @ Mz) <= 1e-11 * abs(z @ My)
