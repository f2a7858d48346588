"""The pressure solve at the full 2.1M-cell size of the benchmark mesh (128^3
jittered hex cavity, permuted + RCM), checked through properties that do not
need the oracle at that size: the FP64 residual of every solution recomputed
independently on the host (scipy) meets the tolerance, the FP32- and
FP64-preconditioned solutions agree to the tolerance's accuracy, and a batch of
8 right-hand sides gives each the bits of its own solve."""
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
