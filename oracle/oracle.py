"""ctypes wrapper of the CPU oracle (libufvm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() as
the checker, and by bench.py for the CPU-baseline leg.  The product package
never imports it.

Each function restates a reference function; the C source cites file:line.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# UFVM_ORACLE_LIB selects another build of the same source (the sanitizer
# build of the CPU tests)
LIB_PATH = os.environ.get("UFVM_ORACLE_LIB") or os.path.join(HERE, "libufvm_oracle.so")
# the same source built with OpenMP: the FPCG solve's loops on all host cores
# (used only to time the CPU baseline; its dot products round differently)
LIB_PATH_OMP = os.path.join(HERE, "libufvm_oracle_omp.so")

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class Level(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int32), ("ptr", i32p), ("col", i32p),
                ("ac", f64p), ("anb", f64p), ("bc", f64p), ("dphi", f64p), ("parents", i32p)]


class Hier(C.Structure):
    _fields_ = [("nlevels", C.c_int32), ("lev", Level * 16)]


class VCfg(C.Structure):
    _fields_ = [("nlev", C.c_int32), ("pre", C.c_int32), ("post", C.c_int32),
                ("singular", C.c_int32), ("block_rows", C.c_int32),
                ("coarse_smoother", C.c_int32), ("smoother", C.c_int32),
                ("coarse_scale", C.c_double), ("cpre", C.c_int32), ("cpost", C.c_int32)]


BLOCK_ROWS = 1024   # rows per block of the coarse-level smoother


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_libs = {}
_openmp = False


def use_openmp(on: bool) -> None:
    """Route the following oracle calls to the OpenMP build (True) or the
    sequential one (False, the default and the only one the tests check)."""
    global _openmp
    _openmp = bool(on)


def lib():
    path = LIB_PATH_OMP if _openmp else LIB_PATH
    if path not in _libs:
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.orc_residual.argtypes = [C.POINTER(Level), f64p]
        L.orc_agglomerate.argtypes = [C.POINTER(Hier), C.c_int]
        L.orc_agglomerate.restype = C.c_int
        L.orc_solve_algebraic_system.argtypes = [C.POINTER(Level), C.c_int, C.c_int, C.c_double,
                                                 f64p, f64p]
        L.orc_apply_amg.argtypes = [C.POINTER(Hier), C.c_int, C.c_int, C.c_double, C.c_int,
                                    C.c_int, C.c_int, f64p, f64p, C.POINTER(C.c_int)]
        L.orc_colour.argtypes = [C.c_int32, i32p, i32p, i32p]
        L.orc_colour.restype = C.c_int
        L.orc_colour_blocks.argtypes = [C.c_int32, i32p, i32p, C.c_int32, i32p]
        L.orc_colour_blocks.restype = C.c_int
        L.orc_vcycle.argtypes = [C.POINTER(Hier), C.POINTER(i32p), C.POINTER(C.c_int),
                                 C.POINTER(VCfg), f64p, f64p]
        L.orc_fpcg.argtypes = [C.POINTER(Hier), C.POINTER(i32p), C.POINTER(C.c_int),
                               C.POINTER(VCfg), f64p, C.c_double, C.c_int, f64p, f64p]
        L.orc_fpcg.restype = C.c_int
        L.orc_factorize_ilu.argtypes = [C.POINTER(Level), f64p, f64p]
        L.orc_mcilu_factor.argtypes = [C.POINTER(Level), i32p, C.c_int, f64p]
        L.orc_mcilu_sweep.argtypes = [C.POINTER(Level), i32p, C.c_int, f64p, f64p, C.c_int,
                                      f64p, f64p]
        L.orc_sor.argtypes = [C.POINTER(Level), f64p]
        L.orc_hier_free.argtypes = [C.POINTER(Hier)]
        _libs[path] = L
    return _libs[path]


def _p(a, t):
    return a.ctypes.data_as(t)


class System:
    """Level-1 system in the reference layout plus its hierarchy (oracle side)."""

    def __init__(self, ptr, col, ac, anb, bc=None, dphi=None):
        n = len(ptr) - 1
        self._keep = []
        self.ptr = np.ascontiguousarray(ptr, np.int32)
        self.col = np.ascontiguousarray(col, np.int32) if len(col) else np.zeros(1, np.int32)
        self.ac = np.ascontiguousarray(ac, np.float64).copy()
        self.anb = np.ascontiguousarray(anb, np.float64).copy() if len(anb) else np.zeros(1)
        self.bc = np.zeros(n) if bc is None else np.ascontiguousarray(bc, np.float64).copy()
        self.dphi = np.zeros(n) if dphi is None else np.ascontiguousarray(dphi, np.float64).copy()
        self.H = Hier()
        self.H.nlevels = 1
        L = self.H.lev[0]
        L.n, L.nnz = n, int(self.ptr[-1])
        L.ptr, L.col = _p(self.ptr, i32p), _p(self.col, i32p)
        L.ac, L.anb = _p(self.ac, f64p), _p(self.anb, f64p)
        L.bc, L.dphi = _p(self.bc, f64p), _p(self.dphi, f64p)
        L.parents = None
        self.colours = None

    def __del__(self):
        try:
            lib().orc_hier_free(C.byref(self.H))
        except Exception:
            pass

    @property
    def nlevels(self):
        return self.H.nlevels

    def level(self, l):
        L = self.H.lev[l]
        n, nnz = L.n, L.nnz
        out = dict(n=n, nnz=nnz,
                   ptr=np.ctypeslib.as_array(L.ptr, (n + 1,)).copy(),
                   col=np.ctypeslib.as_array(L.col, (max(nnz, 1),))[:nnz].copy(),
                   ac=np.ctypeslib.as_array(L.ac, (n,)).copy(),
                   anb=np.ctypeslib.as_array(L.anb, (max(nnz, 1),))[:nnz].copy(),
                   bc=np.ctypeslib.as_array(L.bc, (n,)).copy(),
                   dphi=np.ctypeslib.as_array(L.dphi, (n,)).copy())
        out["parents"] = (np.ctypeslib.as_array(L.parents, (n,)).copy()
                          if l < self.H.nlevels - 1 and bool(L.parents) else None)
        return out

    def set_dphi(self, l, x):
        L = self.H.lev[l]
        np.ctypeslib.as_array(L.dphi, (L.n,))[:] = x

    def residual(self, l=0):
        L = self.H.lev[l]
        r = np.empty(L.n)
        lib().orc_residual(C.byref(L), _p(r, f64p))
        return r

    def agglomerate(self, max_coarse_levels=10):
        lib().orc_hier_free(C.byref(self.H))
        return lib().orc_agglomerate(C.byref(self.H), max_coarse_levels)

    def solve_algebraic_system(self, l, smoother, n_iter, rrf=0.1):
        r0, r1 = C.c_double(), C.c_double()
        lib().orc_solve_algebraic_system(C.byref(self.H.lev[l]),
                                         {"ILU": 0, "SOR": 1, "MCGS": 2}[smoother],
                                         n_iter, rrf, C.byref(r0), C.byref(r1))
        return r0.value, r1.value

    def apply_amg(self, cycle="V-Cycle", max_cycles=30, termination=0.1, max_coarse_levels=10,
                  pre_sweep=1, post_sweep=3, smoother="ILU"):
        """cfdApplyAMG: agglomerate, then cycle (the reference rebuilds every call)."""
        self.agglomerate(max_coarse_levels)
        r0, r1, nc = C.c_double(), C.c_double(), C.c_int()
        lib().orc_apply_amg(C.byref(self.H), {"V-Cycle": 0, "W-Cycle": 1, "F-Cycle": 2}[cycle],
                            max_cycles, termination, pre_sweep, post_sweep,
                            {"ILU": 0, "SOR": 1}[smoother], C.byref(r0), C.byref(r1),
                            C.byref(nc))
        return r0.value, r1.value, nc.value

    def factorize_ilu(self, l=0):
        L = self.H.lev[l]
        dc, rc = np.empty(L.n), np.empty(L.n)
        lib().orc_factorize_ilu(C.byref(L), _p(dc, f64p), _p(rc, f64p))
        return dc

    def mcilu(self, b, l=0, x=None, sweeps=1):
        """DILU in colour order (rows by greedy colour, then index) on level
        l: returns (dinv, x after `sweeps` sweeps, the first from x = 0 when x
        is None)."""
        L = self.H.lev[l]
        c = np.empty(L.n, np.int32)
        nc = lib().orc_colour(L.n, L.ptr, L.col, _p(c, i32p))
        dinv = np.empty(L.n)
        lib().orc_mcilu_factor(C.byref(L), _p(c, i32p), nc, _p(dinv, f64p))
        b = np.ascontiguousarray(b, np.float64)
        rc = np.empty(L.n)
        from_zero = x is None
        x = np.zeros(L.n) if from_zero else np.array(x, np.float64)
        for s in range(sweeps):
            lib().orc_mcilu_sweep(C.byref(L), _p(c, i32p), nc, _p(dinv, f64p), _p(b, f64p),
                                  1 if (s == 0 and from_zero) else 0, _p(rc, f64p), _p(x, f64p))
        return dinv, x

    def mcilu0_factor(self, l=0):
        """ILU(0) in colour order on level l: (colour, dinv, F) with F the
        factor's off-diagonal values in the level's CSR order"""
        L = self.H.lev[l]
        c = np.empty(L.n, np.int32)
        nc = lib().orc_colour(L.n, L.ptr, L.col, _p(c, i32p))
        dinv, F = np.empty(L.n), np.empty(max(L.nnz, 1))
        lib().orc_mcilu0_factor(C.byref(L), _p(c, i32p), nc, _p(dinv, f64p), _p(F, f64p))
        return c, dinv, F[:L.nnz]

    # ------------------------------------------------ precision-form oracle
    def _colour_all(self, nlev):
        """level 0: global greedy colouring; coarse levels: block-local."""
        self.colours, self._ncol = [], []
        for l in range(nlev):
            L = self.H.lev[l]
            c = np.empty(L.n, np.int32)
            if l == 0:
                nc = lib().orc_colour(L.n, L.ptr, L.col, _p(c, i32p))
            else:
                nc = lib().orc_colour_blocks(L.n, L.ptr, L.col, BLOCK_ROWS, _p(c, i32p))
            self.colours.append(c)
            self._ncol.append(nc)
        arr = (i32p * 16)(*[_p(c, i32p) for c in self.colours])
        ncl = (C.c_int * 16)(*self._ncol)
        return arr, ncl

    def pcg_levels(self, max_coarse_cells=2000):
        for l in range(self.H.nlevels):
            if self.H.lev[l].n <= max_coarse_cells:
                return l + 1
        return self.H.nlevels

    def vcycle(self, b, pre=1, post=1, max_coarse_cells=2000, singular=0, coarse_smoother=1,
               smoother="MCGS", coarse_scale=2.5, cpre=2, cpost=2):
        nlev = self.pcg_levels(max_coarse_cells)
        arr, ncl = self._colour_all(nlev)
        cfg = VCfg(nlev, pre, post, singular, BLOCK_ROWS, coarse_smoother,
                   {"MCGS": 0, "ILU": 1, "MCILU": 2, "MCILU0": 3}[smoother], coarse_scale,
                   pre if cpre < 0 else cpre,
                   post if cpost < 0 else cpost)
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        lib().orc_vcycle(C.byref(self.H), arr, ncl, C.byref(cfg), _p(b, f64p), _p(x, f64p))
        return x

    def fpcg(self, b, tol=1e-8, maxit=500, pre=1, post=1, max_coarse_cells=2000, singular=0,
             coarse_smoother=1, smoother="MCGS", coarse_scale=2.5, cpre=2, cpost=2):
        nlev = self.pcg_levels(max_coarse_cells)
        arr, ncl = self._colour_all(nlev)
        cfg = VCfg(nlev, pre, post, singular, BLOCK_ROWS, coarse_smoother,
                   {"MCGS": 0, "ILU": 1, "MCILU": 2, "MCILU0": 3}[smoother], coarse_scale,
                   pre if cpre < 0 else cpre,
                   post if cpost < 0 else cpost)
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        rr = C.c_double()
        it = lib().orc_fpcg(C.byref(self.H), arr, ncl, C.byref(cfg), _p(b, f64p), tol, maxit,
                            _p(x, f64p), C.byref(rr))
        return x, it, rr.value

    def colours_of(self, l):
        return self.colours[l], self._ncol[l]


# ----------------------------------------------------------------- SIMPLE step

class OMesh(C.Structure):
    _fields_ = [("N", C.c_int32), ("F", C.c_int32), ("Fi", C.c_int32),
                ("owner", i32p), ("neighbour", i32p),
                ("Sf", f64p), ("area", f64p), ("CN", f64p), ("eCN", f64p), ("gDiff", f64p),
                ("T", f64p), ("gf", f64p), ("fc", f64p), ("vol", f64p), ("cc", f64p),
                ("cptr", i32p), ("ccol", i32p), ("own_nnz", i32p), ("nei_nnz", i32p),
                ("npatch", C.c_int32), ("pstart", i32p), ("pnf", i32p), ("ptype", i32p),
                ("u_bc", i32p), ("p_bc", i32p), ("u_val", f64p), ("p_val", f64p)]


class OCfg(C.Structure):
    _fields_ = [("nu", C.c_double), ("urf_u", C.c_double), ("urf_p", C.c_double),
                ("n_correctors", C.c_int32), ("p_ref", C.c_int32), ("u_sweeps", C.c_int32),
                ("p_cycle", C.c_int32), ("p_max_cycles", C.c_int32), ("p_max_coarse", C.c_int32),
                ("p_pre", C.c_int32), ("p_post", C.c_int32), ("p_smoother", C.c_int32),
                ("p_termination", C.c_double), ("u_smoother", C.c_int32), ("stress", C.c_int32),
                ("u_init", C.c_double * 3), ("p_init", C.c_double), ("rho", C.c_double),
                ("transient", C.c_int32), ("dt", C.c_double), ("p_solver", C.c_int32),
                ("pcg_tol", C.c_double)]


class OState(C.Structure):
    _fields_ = [("N", C.c_int32), ("Fb", C.c_int32),
                ("U", f64p), ("gradU", f64p), ("p", f64p), ("gradp", f64p), ("PP", f64p),
                ("gradPP", f64p), ("DU", f64p), ("DUT", f64p), ("mdot", f64p), ("gamma_f", f64p),
                ("fluxC1", f64p), ("fluxC2", f64p), ("U_old", f64p), ("mdot_old", f64p),
                ("scaleU", C.c_double), ("H", Hier),
                ("ac_old", f64p), ("rms", C.c_double * 4), ("mx", C.c_double * 4),
                ("r0", C.c_double * 4), ("r1", C.c_double * 4), ("ncycles", C.c_int32),
                ("scale_zero", C.c_int32)]


PATCH_TYPE = {"wall": 0, "empty": 1, "patch": 2, "symmetry": 3, "inlet": 4, "outlet": 5}
BC = {"fixedValue": 0, "zeroGradient": 1, "empty": 2, "noSlip": 3, "symmetry": 4, "slip": 5}


class SimpleCase:
    """The oracle's steady SIMPLE iteration on a processed mesh.

    bcs: {patch name: (U bc, U value (3,), p bc[, p value])}: U bc in
    fixedValue / zeroGradient / noSlip / slip / empty / symmetry, p bc in
    zeroGradient / fixedValue / empty / symmetry.  u_init, p_init: the uniform
    internal fields of the 0 directory.  p_solver: "amg" (cfdApplyAMG with the
    p_* settings) or "fcg64" (an FP64 flexible CG preconditioned by the FP64
    V-cycle of System.fpcg's defaults, to ||r||_2 <= pcg_tol ||b||_2).
    """

    def __init__(self, mesh, bcs, nu, urf_u=0.7, urf_p=0.3, n_correctors=2, p_ref=0,
                 u_sweeps=10, p_cycle="W-Cycle", p_max_cycles=30, p_termination=0.1,
                 p_max_coarse=10, p_pre=1, p_post=3, p_smoother="ILU", u_smoother="ILU",
                 stress=False, u_init=(0.0, 0.0, 0.0), p_init=0.0, rho=1.0, transient=False,
                 dt=1.0, p_solver="amg", pcg_tol=1e-8):
        L = lib()
        for name, args in [("orc_simple_init", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState)]),
                           ("orc_simple_free", [C.POINTER(OState)]),
                           ("orc_simple_iteration", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState)]),
                           ("orc_update_fields", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState)]),
                           ("orc_assemble_u", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState), C.c_int]),
                           ("orc_solve_u", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState), C.c_int]),
                           ("orc_correct_u", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState), C.c_int]),
                           ("orc_assemble_p", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState)]),
                           ("orc_solve_p", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState)]),
                           ("orc_correct_p", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState)]),
                           ("orc_gradient", [C.POINTER(OMesh), f64p, C.c_int, f64p]),
                           ("orc_face_gradients", [C.POINTER(OMesh), f64p, f64p, C.c_int, C.c_int,
                                                   f64p]),
                           ("orc_transient_update", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState)]),
                           ("orc_transient_step", [C.POINTER(OMesh), C.POINTER(OCfg), C.POINTER(OState),
                                                   C.c_int, C.c_double, C.c_double])]:
            getattr(L, name).argtypes = args
        m = mesh
        self.mesh = m
        keep = self._keep = []

        def a(x, t):
            arr = np.ascontiguousarray(x, dtype=np.int32 if t is i32p else np.float64)
            keep.append(arr)
            return arr.ctypes.data_as(t)

        M = OMesh()
        M.N, M.F, M.Fi = m.n_cells, m.n_faces, m.n_interior
        M.owner, M.neighbour = a(m.owner, i32p), a(m.neighbour, i32p)
        M.Sf, M.area, M.CN, M.eCN = a(m.Sf, f64p), a(m.area, f64p), a(m.CN, f64p), a(m.eCN, f64p)
        M.gDiff, M.T, M.gf, M.fc = a(m.gDiff, f64p), a(m.T, f64p), a(m.gf, f64p), a(m.face_centroid, f64p)
        M.vol, M.cc = a(m.volume, f64p), a(m.centroid, f64p)
        M.cptr, M.ccol = a(m.cconn_ptr, i32p), a(m.cconn, i32p)
        M.own_nnz, M.nei_nnz = a(m.own_nnz, i32p), a(m.nei_nnz, i32p)
        M.npatch = len(m.patches)
        M.pstart = a([p.start for p in m.patches], i32p)
        M.pnf = a([p.nfaces for p in m.patches], i32p)
        M.ptype = a([PATCH_TYPE[p.type] for p in m.patches], i32p)
        M.u_bc = a([BC[bcs[p.name][0]] for p in m.patches], i32p)
        M.p_bc = a([BC[bcs[p.name][2]] for p in m.patches], i32p)
        M.u_val = a(np.concatenate([np.asarray(bcs[p.name][1], float) for p in m.patches]), f64p)
        M.p_val = a([float(bcs[p.name][3]) if len(bcs[p.name]) > 3 else 0.0 for p in m.patches], f64p)
        self.M = M
        self.cfg = OCfg(nu, urf_u, urf_p, n_correctors, p_ref, u_sweeps,
                        {"V-Cycle": 0, "W-Cycle": 1, "F-Cycle": 2}[p_cycle], p_max_cycles,
                        p_max_coarse, p_pre, p_post, {"ILU": 0, "SOR": 1}[p_smoother], p_termination,
                        {"ILU": 0, "MCGS": 2}[u_smoother], int(bool(stress)),
                        (C.c_double * 3)(*u_init), float(p_init), float(rho), int(bool(transient)),
                        float(dt), {"amg": 0, "fcg64": 1}[p_solver], float(pcg_tol))
        self.S = OState()
        L.orc_simple_init(C.byref(self.M), C.byref(self.cfg), C.byref(self.S))

    def __del__(self):
        try:
            lib().orc_simple_free(C.byref(self.S))
        except Exception:
            pass

    def _args(self):
        return C.byref(self.M), C.byref(self.cfg), C.byref(self.S)

    def _check_cycles(self):
        # the reference's W- and F-cycles index levels gridLevel-3 and below
        # (cfdApplyWCycle.m, cfdApplyFCycle.m) and fail with fewer than 4
        if self.S.ncycles == -1:
            raise ValueError("W/F cycles need at least 4 AMG levels")

    def iterate(self, n=1):
        for _ in range(n):
            lib().orc_simple_iteration(*self._args())
            self._check_cycles()

    def update_fields(self):
        lib().orc_update_fields(*self._args())

    def transient_update(self):
        lib().orc_transient_update(*self._args())

    def transient_step(self, max_inner=20, u_tol=1e-6, p_tol=1e-6):
        """one time step (cfdRunCase.m:24-86); returns the outer iterations run"""
        n = lib().orc_transient_step(*self._args(), max_inner, u_tol, p_tol)
        self._check_cycles()
        return n

    def assemble_u(self, comp):
        lib().orc_assemble_u(*self._args(), comp)

    def solve_u(self, comp):
        lib().orc_solve_u(*self._args(), comp)

    def correct_u(self, comp):
        lib().orc_correct_u(*self._args(), comp)

    def assemble_p(self):
        lib().orc_assemble_p(*self._args())

    def solve_p(self):
        lib().orc_solve_p(*self._args())
        self._check_cycles()

    def correct_p(self):
        lib().orc_correct_p(*self._args())

    def field(self, name):
        N, Fb, F = self.S.N, self.S.Fb, self.M.F
        shapes = {"U": (N + Fb, 3), "gradU": (N + Fb, 3, 3), "p": (N + Fb,), "gradp": (N + Fb, 3),
                  "PP": (N + Fb,), "gradPP": (N + Fb, 3), "DU": (N + Fb, 3), "DUT": (N + Fb, 3),
                  "mdot": (F,), "gamma_f": (F,), "fluxC1": (F,), "fluxC2": (F,),
                  "U_old": (N + Fb, 3), "mdot_old": (F,)}
        ptr = getattr(self.S, name)
        return np.ctypeslib.as_array(ptr, (int(np.prod(shapes[name])),)).reshape(shapes[name]).copy()

    def face_gradients(self, phi, grad, nc=1, c=0):
        """'Average:Corrected' gradients of the interior faces, (Fi, 3), of
        component c of phi ((N+Fb) x nc) with cell gradients grad
        ((N+Fb) x nc x 3)"""
        phi = np.ascontiguousarray(phi, dtype=np.float64)
        grad = np.ascontiguousarray(grad, dtype=np.float64)
        out = np.zeros((self.M.Fi, 3))
        lib().orc_face_gradients(C.byref(self.M), phi.ctypes.data_as(f64p),
                                 grad.ctypes.data_as(f64p), nc, c, out.ctypes.data_as(f64p))
        return out

    def set_field(self, name, values):
        N, Fb, F = self.S.N, self.S.Fb, self.M.F
        size = {"U": (N + Fb) * 3, "p": N + Fb, "PP": N + Fb, "DU": (N + Fb) * 3,
                "DUT": (N + Fb) * 3, "mdot": F, "U_old": (N + Fb) * 3, "mdot_old": F,
                "gradp": (N + Fb) * 3, "gradU": (N + Fb) * 9}[name]
        np.ctypeslib.as_array(getattr(self.S, name), (size,))[:] = np.asarray(values).reshape(-1)

    def system(self):
        L = self.S.H.lev[0]
        n, nnz = L.n, L.nnz
        return dict(ac=np.ctypeslib.as_array(L.ac, (n,)).copy(),
                    anb=np.ctypeslib.as_array(L.anb, (max(nnz, 1),))[:nnz].copy(),
                    bc=np.ctypeslib.as_array(L.bc, (n,)).copy(),
                    dphi=np.ctypeslib.as_array(L.dphi, (n,)).copy())

    def set_dphi(self, x):
        L = self.S.H.lev[0]
        np.ctypeslib.as_array(L.dphi, (L.n,))[:] = x

    def residuals(self):
        S = self.S
        return dict(rms=list(S.rms), max=list(S.mx), r0=list(S.r0), r1=list(S.r1),
                    ncycles=S.ncycles)


def cavity_bcs(mesh, lid=(1.0, 0.0, 0.0), walls="fixedValue"):
    """Lid-driven cavity boundary conditions (the lid fixedValue, the fixed
    walls fixedValue (0 0 0) or noSlip, zeroGradient p, empty front/back,
    symmetry conditions on symmetry patches)."""
    out = {}
    for p in mesh.patches:
        if p.type in ("empty", "symmetry"):
            out[p.name] = (p.type, (0.0, 0.0, 0.0), p.type)
        elif p.name == "movingWall":
            out[p.name] = ("fixedValue", lid, "zeroGradient")
        else:
            out[p.name] = (walls, (0.0, 0.0, 0.0), "zeroGradient")
    return out
