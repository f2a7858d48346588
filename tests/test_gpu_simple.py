"""GPU parity of the SIMPLE outer iteration against the oracle, through the C ABI.

Every step of cfdRunCase's steady iteration -- field updates and gradients,
momentum assembly, the ILU momentum solve, corrections, Rhie-Chow fluxes and
pressure-correction assembly, the AMG pressure solve and the p' corrections --
is compared bit for bit with the oracle, then whole iterations.
"""
import numpy as np
import pytest

import oracle
from matlab_ufvm_amd.simple import SimpleCase, cavity_bcs
from matlab_ufvm_amd.workloads import cavity_mesh

pytestmark = pytest.mark.gpu

AMG = dict(p_cycle="V-Cycle", p_smoother="SOR", p_pre=2, p_post=2)

MESHES = {
    "cavity2d": lambda: (cavity_mesh(40, nz=1, jitter=0.0, renumber_seed=None), 0.01),
    "cavity3d_jittered": lambda: (cavity_mesh(10, jitter=0.15, seed=51, renumber_seed=52), 0.0025),
}


@pytest.fixture(scope="module", params=sorted(MESHES))
def pair(request, ctx):
    m, nu = MESHES[request.param]()
    print(f"\n[case {request.param}] cells={m.n_cells}", flush=True)
    g = SimpleCase(ctx, m, cavity_bcs(m), nu=nu, **AMG)
    o = oracle.SimpleCase(m, oracle.cavity_bcs(m), nu=nu, **AMG)
    yield m, g, o
    g.close()


def gsys(ctx, g, name):
    sid = g.system(name)
    lv = ctx.level_get(sid, 0)
    return dict(ac=lv["ac"], anb=lv["anb"], bc=ctx.get_bc(sid, 0), dphi=ctx.get_dphi(sid, 0))


def same(a, b, what):
    a, b = np.asarray(a), np.asarray(b)
    assert a.shape == b.shape, what
    bad = np.count_nonzero(a != b)
    assert bad == 0, f"{what}: {bad} of {a.size} differ, max {np.abs(a - b).max():.3e}"


def test_steps_bitexact(ctx, pair):
    m, g, o = pair
    for it in range(3):
        g.cfdUpdateFields()
        o.update_fields()
        same(g.field("gradU"), o.field("gradU"), f"it{it} gradU")
        same(g.field("gradp"), o.field("gradp"), f"it{it} gradp")
        same(g.field("gamma_f"), o.field("gamma_f"), f"it{it} gamma_f")
        for c in range(3):
            rms, mx = g.cfdAssembleEquation("U", c)
            o.assemble_u(c)
            so, sg = o.system(), gsys(ctx, g, "U")
            for k in ("ac", "anb", "bc"):
                same(sg[k], so[k], f"it{it} U{c} {k}")
            same(g.field("DU")[:, c], o.field("DU")[:, c], f"it{it} DU{c}")
            r = o.residuals()
            assert np.isclose(rms, r["rms"][c], rtol=1e-12, atol=1e-300)
            assert mx == r["max"][c]
            g.cfdSolveEquation("U")
            o.solve_u(c)
            same(gsys(ctx, g, "U")["dphi"], o.system()["dphi"], f"it{it} U{c} dphi")
            g.cfdCorrectEquation("U", c)
            o.correct_u(c)
            same(g.field("U"), o.field("U"), f"it{it} U after U{c}")
        for k in range(2):
            rms, mx = g.cfdAssembleEquation("p")
            o.assemble_p()
            same(g.field("mdot_f"), o.field("mdot"), f"it{it} mdot (Rhie-Chow)")
            so, sg = o.system(), gsys(ctx, g, "p")
            for key in ("ac", "anb", "bc"):
                same(sg[key], so[key], f"it{it} p {key}")
            r = o.residuals()
            assert mx == r["max"][3]
            assert np.isclose(rms, r["rms"][3], rtol=1e-12, atol=1e-300)
            r0, r1, nc = g.cfdSolveEquation("p")
            o.solve_p()
            assert nc == o.residuals()["ncycles"]
            same(gsys(ctx, g, "p")["dphi"], o.system()["dphi"], f"it{it} p dphi")
            g.cfdCorrectEquation("p")
            o.correct_p()
            for name, oname in (("PP", "PP"), ("gradPP", "gradPP"), ("U", "U"), ("p", "p"),
                                ("mdot_f", "mdot")):
                same(g.field(name), o.field(oname), f"it{it} corr{k} {name}")


def test_iterations_match(ctx):
    # the CPU-runnable configuration whole: 40 x 40 x 1 cavity, Re 100, 200
    # SIMPLE iterations, V-cycle SOR 2/2 -- bit-identical
    m = cavity_mesh(40, nz=1, jitter=0.0, renumber_seed=None)
    g = SimpleCase(ctx, m, cavity_bcs(m), nu=0.01, **AMG)
    o = oracle.SimpleCase(m, oracle.cavity_bcs(m), nu=0.01, **AMG)
    res = g.iterate(200)
    o.iterate(200)
    for name, oname in (("U", "U"), ("p", "p"), ("mdot_f", "mdot")):
        same(g.field(name), o.field(oname), f"200 iterations {name}")
    assert res[-1, 0] < 0.01 * res[0, 0]
    g.close()


@pytest.mark.parametrize("name", ["cavity3d_16", "cavity2d_64"])
def test_fp32_preconditioned_simple_within_1e8_of_fp64(ctx, name):
    # The precision form's measure at test size: SIMPLE iterations whose
    # pressure-correction solves are the FP32-preconditioned AMG-PCG to
    # ||r|| <= 1e-8 ||b||, against the oracle's SIMPLE whose solves are the
    # FP64 FCG at the same tolerance: relative L2 error <= 1e-8 for U, p and
    # mdot_f.  The FP64-preconditioned GPU solve is measured beside it.
    if name == "cavity3d_16":
        m, nu = cavity_mesh(16, jitter=0.15, seed=171, renumber_seed=172), 1.0 / 400.0
    else:
        m, nu = cavity_mesh(64, nz=1, jitter=0.1, seed=173, renumber_seed=174), 0.01
    o = oracle.SimpleCase(m, oracle.cavity_bcs(m), nu=nu, p_solver="fcg64", pcg_tol=1e-8)
    o.iterate(20)
    errs = {}
    for prec in (32, 64):
        g = SimpleCase(ctx, m, cavity_bcs(m), nu=nu, p_solver="pcg", pcg_precision=prec,
                       pcg_tol=1e-8)
        g.iterate(20)
        errs[prec] = {name_g: float(np.linalg.norm(g.field(name_g) - o.field(name_o)) /
                                    np.linalg.norm(o.field(name_o)))
                      for name_g, name_o in (("U", "U"), ("p", "p"), ("mdot_f", "mdot"))}
        g.close()
    print(f"\n[precision form] {name} cells={m.n_cells} 20 iterations, fields vs the oracle's "
          f"FP64 FCG at 1e-8: FP32 V-cycle {errs[32]}, FP64 V-cycle {errs[64]}", flush=True)
    assert max(errs[32].values()) <= 1e-8, errs
    assert max(errs[64].values()) <= 1e-8, errs


def test_precision_form_pressure_solver(ctx):
    # the same SIMPLE iterations with the FP32-preconditioned AMG-PCG pressure
    # solve reach the same fields as the reference AMG solve to within the
    # solver tolerances
    m = cavity_mesh(24, nz=1, jitter=0.1, seed=61, renumber_seed=62)
    ref = SimpleCase(ctx, m, cavity_bcs(m), nu=0.01, p_termination=1e-10, p_max_cycles=200, **{
        k: v for k, v in AMG.items()})
    pcg = SimpleCase(ctx, m, cavity_bcs(m), nu=0.01, p_solver="pcg", pcg_tol=1e-10)
    ref.iterate(30)
    pcg.iterate(30)
    U1, U2 = ref.field("U"), pcg.field("U")
    assert np.linalg.norm(U1 - U2) <= 1e-6 * np.linalg.norm(U1)
    ref.close()
    pcg.close()


def test_mcgs_momentum_iterations_bitexact(ctx):
    # the multicolour-GS momentum solve (u_smoother "MCGS") on the GPU gives
    # the oracle's fields bit for bit over whole SIMPLE iterations
    m = cavity_mesh(12, jitter=0.15, seed=65, renumber_seed=66)
    g = SimpleCase(ctx, m, cavity_bcs(m), nu=0.0025, u_smoother="MCGS", **AMG)
    o = oracle.SimpleCase(m, oracle.cavity_bcs(m), nu=0.0025, u_smoother="MCGS", **AMG)
    res = g.iterate(5)
    for _ in range(5):
        o.iterate(1)
    for name, oname in (("U", "U"), ("p", "p"), ("mdot_f", "mdot")):
        same(g.field(name), o.field(oname), f"MCGS momentum {name}")
    assert np.all(np.isfinite(res))
    g.close()


def test_noslip_walls_bitexact(ctx):
    # the reference's vector noSlip wall (patch-wide sums of n_c^2, TM =
    # nu / walldist) on the GPU gives the oracle's fields bit for bit, and it
    # is a different flow from the fixedValue-wall one
    m = cavity_mesh(8, jitter=0.15, seed=71, renumber_seed=72)
    g = SimpleCase(ctx, m, cavity_bcs(m, walls="noSlip"), nu=0.01, **AMG)
    o = oracle.SimpleCase(m, oracle.cavity_bcs(m, walls="noSlip"), nu=0.01, **AMG)
    res = g.iterate(4)
    for _ in range(4):
        o.iterate(1)
    for name, oname in (("U", "U"), ("p", "p"), ("mdot_f", "mdot")):
        same(g.field(name), o.field(oname), f"noSlip {name}")
    assert np.all(np.isfinite(res))
    f = SimpleCase(ctx, m, cavity_bcs(m), nu=0.01, **AMG)
    f.iterate(4)
    assert not np.array_equal(f.field("U"), g.field("U"))
    f.close()
    g.close()


def test_symmetry_plane_bitexact(ctx):
    # a symmetry plane (z = 0) of a jittered 3-D cavity: the GPU gives the
    # oracle's fields bit for bit
    from matlab_ufvm_amd.mesh import process_mesh
    from matlab_ufvm_amd.meshgen import box_hex
    spec = [("movingWall", "wall", ["y1"]), ("fixedWalls", "wall", ["x0", "x1", "y0", "z1"]),
            ("midPlane", "symmetry", ["z0"])]
    m = process_mesh(box_hex(8, 8, 8, spec, jitter=0.15, seed=81, renumber_seed=82),
                     apex="reference")
    g = SimpleCase(ctx, m, cavity_bcs(m), nu=0.01, **AMG)
    o = oracle.SimpleCase(m, oracle.cavity_bcs(m), nu=0.01, **AMG)
    res = g.iterate(4)
    for _ in range(4):
        o.iterate(1)
    for name, oname in (("U", "U"), ("p", "p"), ("mdot_f", "mdot")):
        same(g.field(name), o.field(oname), f"symmetry {name}")
    assert np.all(np.isfinite(res))
    g.close()


def test_slip_walls_bitexact(ctx):
    # the reference's wall slip on a jittered 3-D cavity's fixed walls: the
    # GPU gives the oracle's fields bit for bit
    m = cavity_mesh(8, jitter=0.15, seed=91, renumber_seed=92)
    g = SimpleCase(ctx, m, cavity_bcs(m, walls="slip"), nu=0.01, **AMG)
    o = oracle.SimpleCase(m, oracle.cavity_bcs(m, walls="slip"), nu=0.01, **AMG)
    res = g.iterate(4)
    for _ in range(4):
        o.iterate(1)
    for name, oname in (("U", "U"), ("p", "p"), ("mdot_f", "mdot")):
        same(g.field(name), o.field(oname), f"slip {name}")
    assert np.all(np.isfinite(res))
    g.close()


def test_symmetry_slip_and_stress_together_bitexact(ctx):
    # a symmetry plane, slip walls and the explicit stress term in one case
    # (the stress term's boundary rows read the planes' and walls' gradient
    # rows): the GPU gives the oracle's fields bit for bit
    from matlab_ufvm_amd.mesh import process_mesh
    from matlab_ufvm_amd.meshgen import box_hex
    spec = [("movingWall", "wall", ["y1"]), ("fixedWalls", "wall", ["x0", "x1", "y0", "z1"]),
            ("midPlane", "symmetry", ["z0"])]
    m = process_mesh(box_hex(7, 7, 7, spec, jitter=0.15, seed=93, renumber_seed=94),
                     apex="reference")
    g = SimpleCase(ctx, m, cavity_bcs(m, walls="slip"), nu=0.01, stress=True, **AMG)
    o = oracle.SimpleCase(m, oracle.cavity_bcs(m, walls="slip"), nu=0.01, stress=True, **AMG)
    res = g.iterate(4)
    for _ in range(4):
        o.iterate(1)
    for name, oname in (("U", "U"), ("p", "p"), ("mdot_f", "mdot")):
        same(g.field(name), o.field(oname), f"combined {name}")
    assert np.all(np.isfinite(res))
    g.close()


def test_explicit_stress_term_bitexact(ctx):
    # + div(nu*transp(grad(U))) (the reference's tutorials' momentum equation)
    # on the GPU gives the oracle's fields bit for bit over SIMPLE iterations
    m = cavity_mesh(10, jitter=0.15, seed=73, renumber_seed=74)
    g = SimpleCase(ctx, m, cavity_bcs(m), nu=0.005, stress=True, **AMG)
    o = oracle.SimpleCase(m, oracle.cavity_bcs(m), nu=0.005, stress=True, **AMG)
    g.iterate(4)
    for _ in range(4):
        o.iterate(1)
    for name, oname in (("U", "U"), ("p", "p"), ("mdot_f", "mdot")):
        same(g.field(name), o.field(oname), f"stress term {name}")
    f = SimpleCase(ctx, m, cavity_bcs(m), nu=0.005, **AMG)
    f.iterate(4)
    assert not np.array_equal(f.field("U"), g.field("U"))
    f.close()
    g.close()


def test_frozen_hierarchy_pressure_solver(ctx):
    # "pcg-frozen" keeps the first solve's hierarchy and refreshes its values
    # on the GPU after every assembly; it reaches the fields of the
    # rebuild-every-solve path to within the solver tolerances
    m = cavity_mesh(20, nz=1, jitter=0.1, seed=67, renumber_seed=68)
    a = SimpleCase(ctx, m, cavity_bcs(m), nu=0.01, p_solver="pcg", pcg_tol=1e-10)
    b = SimpleCase(ctx, m, cavity_bcs(m), nu=0.01, p_solver="pcg-frozen", pcg_tol=1e-10)
    ra = a.iterate(20)
    rb = b.iterate(20)
    assert np.allclose(ra[0], rb[0], rtol=1e-9)
    U1, U2 = a.field("U"), b.field("U")
    assert np.linalg.norm(U1 - U2) <= 1e-6 * np.linalg.norm(U1)
    a.close()
    b.close()


def test_restart_from_time_directory_is_exact(ctx, tmp_path):
    # write U, p, mdot_f as an OpenFOAM time directory (lossless), restart a
    # fresh case from it: the continued run has the bits of the uninterrupted one
    from matlab_ufvm_amd import fieldio
    m = cavity_mesh(12, jitter=0.15, seed=69, renumber_seed=70)
    a = SimpleCase(ctx, m, cavity_bcs(m), nu=0.0025, **AMG)
    a.iterate(10)
    b = SimpleCase(ctx, m, cavity_bcs(m), nu=0.0025, **AMG)
    b.iterate(6)
    fieldio.write_time(b, str(tmp_path), 6)
    c = SimpleCase(ctx, m, cavity_bcs(m), nu=0.0025, **AMG)
    fieldio.read_time(c, str(tmp_path), 6)
    c.iterate(4)
    for name in ("U", "p", "mdot_f"):
        same(c.field(name), a.field(name), f"restart {name}")
    for case in (a, b, c):
        case.close()
