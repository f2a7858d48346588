"""Grouped solves (ufvm_pcg_solve_group), through the C ABI: g solves of
different systems advance in lock step and each phase of an iteration is one
launch for all of them.  Every member must get the solution bits, iteration
count and relative residual of its own ufvm_pcg_solve, whatever the other
members' sizes, hierarchy depths, colour counts and convergence histories.
"""
import numpy as np
import pytest

from matlab_ufvm_amd.workloads import cavity_mesh, gaussian_rhs, poisson_system, poisson_tet_mesh

pytestmark = pytest.mark.gpu

SYSTEMS = [
    # 3-D jittered hex cavities of two sizes (different depths), Dirichlet lid
    ("hex12", lambda: cavity_mesh(12, jitter=0.15, seed=31, renumber_seed=32), ("movingWall",)),
    ("hex24", lambda: cavity_mesh(24, jitter=0.15, seed=36, renumber_seed=37), ("movingWall",)),
    # jittered tets: more colours on the finest level
    ("tet6", lambda: poisson_tet_mesh(6, jitter=0.1, seed=33, renumber_seed=34), ("dirichlet",)),
    # 2-D plain cavity
    ("hex2d", lambda: cavity_mesh(40, nz=1, jitter=0.0, renumber_seed=None), ("movingWall",)),
    # all-Neumann (singular, projected), and a mesh small enough for one dense level
    ("neumann", lambda: cavity_mesh(10, jitter=0.15, seed=41, renumber_seed=42), ()),
    ("tiny", lambda: cavity_mesh(5, jitter=0.1, seed=43, renumber_seed=44), ("movingWall",)),
]


@pytest.fixture(scope="module")
def systems():
    out = []
    for i, (name, mk, dirichlet) in enumerate(SYSTEMS):
        m = mk()
        ptr, col, ac, anb = poisson_system(m, dirichlet=dirichlet)
        b = gaussian_rhs(m, seed=50 + i)
        out.append((name, ptr, col, ac, anb, b))
    return out


def make_all(ctx, systems, precision, launch_mode, **kw):
    sids = []
    for name, ptr, col, ac, anb, b in systems:
        sid = ctx.system_create(ptr, col)
        ctx.set_coefficients(sid, ac, anb, b)
        ctx.agglomerate(sid, 10)
        ctx.pcg_setup(sid, precision=precision, max_coarse_cells=200, tol=1e-9, max_iter=200,
                      launch_mode=launch_mode, **kw)
        sids.append(sid)
    return sids


def alone(ctx, sid, b):
    ctx.set_coefficients(sid, bc=b)      # the right-hand side only: the hierarchy stays
    it, rr = ctx.pcg_solve(sid)
    return ctx.get_dphi(sid, 0), it, rr


@pytest.mark.parametrize("precision,launch_mode", [(32, "eager"), (32, "graphs"), (64, "eager")])
def test_group_members_keep_their_own_bits(ctx, systems, precision, launch_mode):
    print(f"\n[group] precision={precision} launch={launch_mode} "
          f"sizes={[len(s[1]) - 1 for s in systems]}", flush=True)
    sids = make_all(ctx, systems, precision, launch_mode)
    ref = [alone(ctx, sid, s[5]) for sid, s in zip(sids, systems)]
    its_alone = [r[1] for r in ref]
    # the members converge after different numbers of iterations
    assert len(set(its_alone)) > 1, its_alone
    xs, its, rel = ctx.pcg_solve_group(sids, [s[5] for s in systems])
    for q, (x, it, rr) in enumerate(ref):
        assert np.array_equal(xs[q], x), systems[q][0]
        assert its[q] == it and rel[q] == rr, (systems[q][0], its[q], it, rel[q], rr)
    # any subset, in any order, gives the same bits again
    order = [4, 1, 3]
    xs2, its2, _ = ctx.pcg_solve_group([sids[q] for q in order], [systems[q][5] for q in order])
    for j, q in enumerate(order):
        assert np.array_equal(xs2[j], ref[q][0]) and its2[j] == ref[q][1]
    for sid in sids:
        ctx.system_destroy(sid)


def test_group_zero_right_hand_side_and_refusals(ctx, systems):
    sub = systems[:3]
    sids = make_all(ctx, sub, 32, "auto")
    ref = [alone(ctx, sid, s[5]) for sid, s in zip(sids, sub)]
    bs = [s[5] for s in sub]
    bs[1] = np.zeros_like(bs[1])
    xs, its, _ = ctx.pcg_solve_group(sids, bs)
    assert its[1] == 0 and not xs[1].any()
    for q in (0, 2):
        assert np.array_equal(xs[q], ref[q][0]) and its[q] == ref[q][1]
    # one system twice: a batch's job, refused
    with pytest.raises(Exception, match="twice"):
        ctx.pcg_solve_group([sids[0], sids[0]], [bs[0], bs[0]])
    # a V-cycle precision or smoother the group cannot share
    ctx.pcg_setup(sids[2], precision=64, max_coarse_cells=200, tol=1e-9, max_iter=200)
    with pytest.raises(Exception, match="precision"):
        ctx.pcg_solve_group(sids, [s[5] for s in sub])
    ctx.pcg_setup(sids[2], precision=32, max_coarse_cells=200, tol=1e-9, max_iter=200,
                  smoother="MCILU")
    with pytest.raises(Exception, match="multicolour GS"):
        ctx.pcg_solve_group(sids, [s[5] for s in sub])
    for sid in sids:
        ctx.system_destroy(sid)
