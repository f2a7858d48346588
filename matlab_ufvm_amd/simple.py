"""Host side of the SIMPLE outer iteration, mirroring the reference's calls.

The reference drives one steady iteration from MATLAB
(uFVM/src/fvm/top/cfdRunCase.m:92-142):

    cfdUpdateFields
    for U components: cfdAssembleEquation / cfdSolveEquation / cfdCorrectEquation
    nCorrectors times for p: the same three calls

`SimpleCase` exposes those calls with the same names and meaning on a mesh
resident on the GPU (the fields, the two linear systems and the AMG hierarchy
stay in HBM); every method goes through the C ABI of libufvm.so.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import AmgCfg, MeshDesc, PcgCfg, SimpleCfg, f64p, i32p

PATCH = {"wall": 0, "empty": 1, "symmetry": 3, "inlet": 4, "outlet": 5}
BC = {"fixedValue": 0, "zeroGradient": 1, "empty": 2, "noSlip": 3, "symmetry": 4, "slip": 5}
P_SOLVER = {"amg": 0, "pcg": 1, "pcg-frozen": 2}


def cavity_bcs(mesh, lid=(1.0, 0.0, 0.0), walls="fixedValue"):
    """Lid-driven cavity: the lid fixedValue U, the fixed walls fixedValue
    (0 0 0) or noSlip (`walls`), zeroGradient p, empty front/back patches of
    2-D meshes, symmetry conditions on symmetry planes."""
    out = {}
    for p in mesh.patches:
        if p.type in ("empty", "symmetry"):
            out[p.name] = (p.type, (0.0, 0.0, 0.0), p.type)
        elif p.name == "movingWall":
            out[p.name] = ("fixedValue", tuple(lid), "zeroGradient")
        else:
            out[p.name] = (walls, (0.0, 0.0, 0.0), "zeroGradient")
    return out


def case_structs(mesh, bcs, nu, urf_u=0.7, urf_p=0.3, n_correctors=2, p_ref=0, u_sweeps=10,
                 p_solver="amg", p_cycle="W-Cycle", p_max_cycles=30, p_termination=0.1,
                 p_max_coarse_levels=10, p_pre=1, p_post=3, p_smoother="ILU", pcg_precision=32,
                 pcg_tol=1e-8, pcg_sweeps=1, pcg_max_coarse_cells=2000, pcg_coarse_scale=2.5,
                 u_smoother="ILU", pcg_coarse_sweeps=2, stress=False, u_init=(0.0, 0.0, 0.0),
                 p_init=0.0, rho=1.0, transient=False, dt=1.0):
    """(MeshDesc, SimpleCfg, keep): the C structs of ufvm_case_create; the arrays
    they point to live in `keep`, which must outlive the call.  bcs:
    {patch: (U bc, U value (3,), p bc[, p value])}.  u_smoother: "ILU" (the
    reference's DILU momentum solve) or "MCGS" (multicolour GS); stress: the
    momentum equation has + div(nu*transp(grad(U))), as the reference's
    tutorials do; u_init / p_init: the uniform internal fields of the 0
    directory; rho: the constant density (nu is then the laplacian's
    coefficient); transient / dt: ddt(rho*U) with the Euler scheme."""
    m = mesh
    keep = []

    def a(x, t):
        arr = np.ascontiguousarray(x, dtype=np.int32 if t is i32p else np.float64)
        keep.append(arr)
        return arr.ctypes.data_as(t)

    d = MeshDesc()
    d.n_cells, d.n_faces, d.n_interior = m.n_cells, m.n_faces, m.n_interior
    d.owner, d.neighbour = a(m.owner, i32p), a(m.neighbour, i32p)
    d.Sf, d.area, d.CN, d.eCN = a(m.Sf, f64p), a(m.area, f64p), a(m.CN, f64p), a(m.eCN, f64p)
    d.gDiff, d.T, d.gf = a(m.gDiff, f64p), a(m.T, f64p), a(m.gf, f64p)
    d.face_centroid = a(m.face_centroid, f64p)
    d.volume, d.centroid = a(m.volume, f64p), a(m.centroid, f64p)
    d.cconn_ptr, d.cconn = a(m.cconn_ptr, i32p), a(m.cconn, i32p)
    d.own_nnz, d.nei_nnz = a(m.own_nnz, i32p), a(m.nei_nnz, i32p)
    d.n_patches = len(m.patches)
    d.patch_start = a([p.start for p in m.patches], i32p)
    d.patch_nfaces = a([p.nfaces for p in m.patches], i32p)
    d.patch_type = a([PATCH[p.type] for p in m.patches], i32p)
    c = SimpleCfg()
    c.nu, c.urf_u, c.urf_p = nu, urf_u, urf_p
    c.n_correctors, c.p_ref_cell, c.u_sweeps = n_correctors, p_ref, u_sweeps
    c.u_bc = a([BC[bcs[p.name][0]] for p in m.patches], i32p)
    c.u_value = a(np.concatenate([np.asarray(bcs[p.name][1], float) for p in m.patches]), f64p)
    c.p_bc = a([BC[bcs[p.name][2]] for p in m.patches], i32p)
    c.p_solver = P_SOLVER[p_solver]
    c.p_amg = AmgCfg(_lib.CYCLE[p_cycle], p_max_cycles, p_termination, p_max_coarse_levels,
                     p_pre, p_post, _lib.SMOOTHER[p_smoother], 1)
    c.p_pcg = PcgCfg(pcg_precision, pcg_sweeps, pcg_sweeps, 1, pcg_max_coarse_cells, -1, 500,
                     pcg_tol, 0, pcg_coarse_scale, pcg_coarse_sweeps, pcg_coarse_sweeps)
    c.u_smoother = _lib.SMOOTHER[u_smoother]
    c.stress_term = int(bool(stress))
    c.p_value = a([float(bcs[p.name][3]) if len(bcs[p.name]) > 3 else 0.0 for p in m.patches],
                  f64p)
    c.u_init = (C.c_double * 3)(*[float(v) for v in u_init])
    c.p_init = float(p_init)
    c.rho = float(rho)
    c.transient = int(bool(transient))
    c.dt = float(dt)
    return d, c, keep


class SimpleCase:
    def __init__(self, ctx: _lib.Context, mesh, bcs, nu, urf_u=0.7, urf_p=0.3, n_correctors=2,
                 p_ref=0, u_sweeps=10, p_solver="amg", p_cycle="W-Cycle", p_max_cycles=30,
                 p_termination=0.1, p_max_coarse_levels=10, p_pre=1, p_post=3, p_smoother="ILU",
                 pcg_precision=32, pcg_tol=1e-8, pcg_sweeps=1, pcg_max_coarse_cells=2000,
                 pcg_coarse_scale=2.5, u_smoother="ILU", pcg_coarse_sweeps=2, stress=False,
                 u_init=(0.0, 0.0, 0.0), p_init=0.0, rho=1.0, transient=False, dt=1.0):
        self.ctx = ctx
        self.mesh = mesh
        self.bcs = bcs
        d, c, self._keep = case_structs(
            mesh, bcs, nu, urf_u, urf_p, n_correctors, p_ref, u_sweeps, p_solver, p_cycle,
            p_max_cycles, p_termination, p_max_coarse_levels, p_pre, p_post, p_smoother,
            pcg_precision, pcg_tol, pcg_sweeps, pcg_max_coarse_cells, pcg_coarse_scale,
            u_smoother, pcg_coarse_sweeps, stress, u_init, p_init, rho, transient, dt)
        cid = C.c_int32()
        ctx.check(ctx.lib.ufvm_case_create(ctx.h, C.byref(d), C.byref(c), C.byref(cid)))
        self.id = cid.value
        self.n_correctors = n_correctors
        self._keep = None

    def close(self):
        if self.id:
            self.ctx.check(self.ctx.lib.ufvm_case_destroy(self.ctx.h, self.id))
            self.id = 0

    # -------------------------------------------------- reference call surface
    def cfdUpdateFields(self):
        self.ctx.check(self.ctx.lib.ufvm_update_fields(self.ctx.h, self.id))

    def cfdAssembleEquation(self, name, comp=0):
        rms, mx = C.c_double(), C.c_double()
        self.ctx.check(self.ctx.lib.ufvm_assemble_equation(self.ctx.h, self.id, name.encode(),
                                                           comp, C.byref(rms), C.byref(mx)))
        return rms.value, mx.value

    def cfdSolveEquation(self, name):
        r0, r1, nc = C.c_double(), C.c_double(), C.c_int32()
        self.ctx.check(self.ctx.lib.ufvm_solve_equation(self.ctx.h, self.id, name.encode(),
                                                        C.byref(r0), C.byref(r1), C.byref(nc)))
        return r0.value, r1.value, nc.value

    def cfdCorrectEquation(self, name, comp=0):
        self.ctx.check(self.ctx.lib.ufvm_correct_equation(self.ctx.h, self.id, name.encode(),
                                                          comp))

    def cfdAssembleAndCorrectEquation(self, name):
        """cfdAssembleAndCorrectEquation.m:35-49: rms/max of each component."""
        ncomp = 3 if name == "U" else 1
        ncorr = self.n_correctors if name == "p" else 1
        rms, mx = [0.0] * ncomp, [0.0] * ncomp
        for c in range(ncomp):
            for _ in range(ncorr):
                rms[c], mx[c] = self.cfdAssembleEquation(name, c)
                self.cfdSolveEquation(name)
                self.cfdCorrectEquation(name, c)
        return rms, mx

    def cfdTransientUpdate(self):
        """U, mdot_f -> the previous step's U_old, mdot_old (cfdTransientUpdate.m)"""
        self.ctx.check(self.ctx.lib.ufvm_transient_update(self.ctx.h, self.id))

    def transient_step(self, max_inner=20, u_tol=1e-6, p_tol=1e-6):
        """one time step (cfdRunCase.m:24-86): (iterations run, RMS residuals [it, 4])"""
        rms = np.zeros((max_inner, 4))
        n = C.c_int32()
        self.ctx.check(self.ctx.lib.ufvm_transient_step(self.ctx.h, self.id, max_inner, u_tol, p_tol,
                                                        rms.ctypes.data_as(f64p), C.byref(n)))
        return n.value, rms[:n.value]

    def iterate(self, n=1):
        """n steady outer iterations (cfdRunCase.m:92-142); RMS residuals [n, 4]."""
        out = np.zeros((n, 4))
        self.ctx.check(self.ctx.lib.ufvm_simple_iterate(self.ctx.h, self.id, n,
                                                        out.ctypes.data_as(f64p)))
        return out

    # ----------------------------------------------------------------- fields
    def _size(self, name):
        m = self.mesh
        nb = m.n_cells + m.n_bfaces
        return {"U": nb * 3, "gradU": nb * 9, "p": nb, "gradp": nb * 3, "PP": nb,
                "gradPP": nb * 3, "DU": nb * 3, "DUT": nb * 3, "mdot_f": m.n_faces,
                "gamma_f": m.n_faces, "U_old": nb * 3, "mdot_old": m.n_faces}[name]

    def field(self, name):
        n = self._size(name)
        out = np.empty(n)
        self.ctx.check(self.ctx.lib.ufvm_case_get_field(self.ctx.h, self.id, name.encode(),
                                                        out.ctypes.data_as(f64p), n))
        m = self.mesh
        nb = m.n_cells + m.n_bfaces
        shape = {"U": (nb, 3), "gradU": (nb, 3, 3), "gradp": (nb, 3), "gradPP": (nb, 3),
                 "DU": (nb, 3), "DUT": (nb, 3), "U_old": (nb, 3)}.get(name, (n,))
        return out.reshape(shape)

    def set_field(self, name, values):
        v = np.ascontiguousarray(values, dtype=np.float64).reshape(-1)
        self.ctx.check(self.ctx.lib.ufvm_case_set_field(self.ctx.h, self.id, name.encode(),
                                                        v.ctypes.data_as(f64p), v.size))

    def system(self, name):
        sid = C.c_int32()
        self.ctx.check(self.ctx.lib.ufvm_case_system(self.ctx.h, self.id, name.encode(),
                                                     C.byref(sid)))
        return sid.value
