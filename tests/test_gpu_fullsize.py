"""The pressure solve at the full size of the benchmark meshes (the 128^3
jittered hex cavity, permuted + RCM, and the 5.3M-cell tetrahedral Poisson
mesh), checked through properties that do not need the oracle at that size:
the FP64 residual of every solution recomputed independently on the host
(scipy) meets the tolerance, the FP32- and FP64-preconditioned solutions agree
to the tolerance's accuracy, a batch of 8 right-hand sides gives each the bits
of its own solve, and on the tets ILU(0) in colour order gives the colour-order
DILU's bits."""
import time

import numpy as np
import pytest
import scipy.sparse as sp

from matlab_ufvm_amd.workloads import cavity_mesh, gaussian_rhs, poisson_system

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def system():
    t = time.time()
    m = cavity_mesh(128, jitter=0.15, seed=2026, renumber_seed=2027, apex="own")
    ptr, col, ac, anb = poisson_system(m, dirichlet=())
    A = (sp.csr_matrix((anb, col, ptr), shape=(m.n_cells, m.n_cells)) + sp.diags(ac)).tocsr()
    print(f"\n[fullsize] 128^3 mesh and operator: {m.n_cells} cells ({time.time()-t:.1f}s)",
          flush=True)
    return m, ptr, col, ac, anb, A


def projected(b):
    return b - b.mean()     # all-wall Neumann operator: constant nullspace


def test_fullsize_solves_meet_tolerance(ctx, system):
    m, ptr, col, ac, anb, A = system
    b = gaussian_rhs(m, seed=2030)
    sid = ctx.system_create(ptr, col)
    ctx.set_coefficients(sid, ac, anb, b)
    nl = ctx.agglomerate(sid, 10)
    sols = {}
    for prec in (32, 64):
        ctx.pcg_setup(sid, precision=prec, tol=1e-8)
        it, rr = ctx.pcg_solve(sid)
        x = ctx.get_dphi(sid, 0)
        bp = projected(b)
        res = np.linalg.norm(bp - A @ x) / np.linalg.norm(bp)
        print(f"[fullsize] fp{prec}: {nl} levels, {it} iterations, device relres {rr:.2e}, "
              f"host FP64 relres {res:.2e}", flush=True)
        assert rr <= 1e-8 and res <= 1.5e-8
        sols[prec] = x - x.mean()
    d = np.linalg.norm(sols[32] - sols[64]) / np.linalg.norm(sols[64])
    print(f"[fullsize] FP32- vs FP64-preconditioned solution: {d:.2e}", flush=True)
    assert d <= 1e-6
    ctx.system_destroy(sid)


def test_fullsize_batch_bits(ctx, system):
    m, ptr, col, ac, anb, A = system
    B = np.stack([gaussian_rhs(m, seed=2040 + q) for q in range(8)])
    sid = ctx.system_create(ptr, col)
    ctx.set_coefficients(sid, ac, anb, B[0])
    ctx.agglomerate(sid, 10)
    ctx.pcg_setup(sid, precision=32, tol=1e-8)
    X, its, rel = ctx.pcg_solve_batch(sid, B)
    print(f"[fullsize] batch of 8: iterations {its.tolist()}", flush=True)
    for q in (0, 5):
        ctx.set_bc(sid, 0, B[q])
        it, rr = ctx.pcg_solve(sid)
        assert it == its[q] and rr == rel[q]
        assert np.array_equal(ctx.get_dphi(sid, 0), X[q]), q
    for q in range(8):
        bp = projected(B[q])
        assert np.linalg.norm(bp - A @ X[q]) <= 1.5e-8 * np.linalg.norm(bp), q
    ctx.system_destroy(sid)


def test_fullsize_tet_smoothers(ctx):
    # The tetrahedral Poisson configuration at its full size: 6 tets per hex
    # of a jittered 96^3 grid (5.3M cells), unit diffusivity, p = 0 on the
    # x = 0 face, Gaussian right-hand side, relative residual 1e-10, the
    # finest level smoothed by multicolour GS and by ILU(0) in colour order.
    # Each solution's FP64 residual is recomputed on the host; ILU(0) equals
    # the colour-order DILU bit for bit (the face graph of these tets has no
    # triangles, so the in-pattern updates vanish).
    from matlab_ufvm_amd.workloads import poisson_tet_mesh
    t = time.time()
    m = poisson_tet_mesh(96)
    ptr, col, ac, anb = poisson_system(m, dirichlet=("dirichlet",))
    A = (sp.csr_matrix((anb, col, ptr), shape=(m.n_cells, m.n_cells)) + sp.diags(ac)).tocsr()
    b = gaussian_rhs(m, seed=2029)
    print(f"\n[fullsize tet] {m.n_cells} cells ({time.time()-t:.1f}s)", flush=True)
    sid = ctx.system_create(ptr, col)
    ctx.set_coefficients(sid, ac, anb, b)
    nl = ctx.agglomerate(sid, 10)
    out = {}
    for sm in ("MCGS", "MCILU0", "MCILU"):
        ctx.pcg_setup(sid, precision=32, tol=1e-10, smoother=sm)
        it, rr = ctx.pcg_solve(sid)
        x = ctx.get_dphi(sid, 0)
        res = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
        print(f"[fullsize tet] {sm}: {nl} levels, {it} iterations, device relres {rr:.2e}, "
              f"host FP64 relres {res:.2e}", flush=True)
        assert rr <= 1e-10 and res <= 1.5e-10, sm
        out[sm] = (it, rr, x)
    assert out["MCILU0"][0] == out["MCILU"][0] and out["MCILU0"][1] == out["MCILU"][1]
    assert np.array_equal(out["MCILU0"][2], out["MCILU"][2])
    d = np.linalg.norm(out["MCGS"][2] - out["MCILU"][2]) / np.linalg.norm(out["MCILU"][2])
    assert d <= 1e-8, d
    ctx.system_destroy(sid)
