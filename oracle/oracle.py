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
LIB_PATH = os.path.join(HERE, "libufvm_oracle.so")

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class Level(C.Structure):
    _fields_ = [("n", C.c_int32), ("nnz", C.c_int32), ("ptr", i32p), ("col", i32p),
                ("ac", f64p), ("anb", f64p), ("bc", f64p), ("dphi", f64p), ("parents", i32p)]


class Hier(C.Structure):
    _fields_ = [("nlevels", C.c_int32), ("lev", Level * 16)]


class VCfg(C.Structure):
    _fields_ = [("nlev", C.c_int32), ("pre", C.c_int32), ("post", C.c_int32),
                ("singular", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_residual.argtypes = [C.POINTER(Level), f64p]
        L.orc_agglomerate.argtypes = [C.POINTER(Hier), C.c_int]
        L.orc_agglomerate.restype = C.c_int
        L.orc_solve_algebraic_system.argtypes = [C.POINTER(Level), C.c_int, C.c_int, C.c_double,
                                                 f64p, f64p]
        L.orc_apply_amg.argtypes = [C.POINTER(Hier), C.c_int, C.c_int, C.c_double, C.c_int,
                                    C.c_int, C.c_int, f64p, f64p, C.POINTER(C.c_int)]
        L.orc_colour.argtypes = [C.c_int32, i32p, i32p, i32p]
        L.orc_colour.restype = C.c_int
        L.orc_vcycle.argtypes = [C.POINTER(Hier), C.POINTER(i32p), C.POINTER(C.c_int),
                                 C.POINTER(VCfg), f64p, f64p]
        L.orc_fpcg.argtypes = [C.POINTER(Hier), C.POINTER(i32p), C.POINTER(C.c_int),
                               C.POINTER(VCfg), f64p, C.c_double, C.c_int, f64p, f64p]
        L.orc_fpcg.restype = C.c_int
        L.orc_factorize_ilu.argtypes = [C.POINTER(Level), f64p, f64p]
        L.orc_sor.argtypes = [C.POINTER(Level), f64p]
        L.orc_hier_free.argtypes = [C.POINTER(Hier)]
        _lib = L
    return _lib


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
        lib().orc_solve_algebraic_system(C.byref(self.H.lev[l]), {"ILU": 0, "SOR": 1}[smoother],
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

    # ------------------------------------------------ precision-form oracle
    def _colour_all(self, nlev):
        self.colours, self._ncol = [], []
        for l in range(nlev):
            L = self.H.lev[l]
            c = np.empty(L.n, np.int32)
            nc = lib().orc_colour(L.n, L.ptr, L.col, _p(c, i32p))
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

    def vcycle(self, b, pre=1, post=1, max_coarse_cells=2000, singular=0):
        nlev = self.pcg_levels(max_coarse_cells)
        arr, ncl = self._colour_all(nlev)
        cfg = VCfg(nlev, pre, post, singular)
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        lib().orc_vcycle(C.byref(self.H), arr, ncl, C.byref(cfg), _p(b, f64p), _p(x, f64p))
        return x

    def fpcg(self, b, tol=1e-8, maxit=500, pre=1, post=1, max_coarse_cells=2000, singular=0):
        nlev = self.pcg_levels(max_coarse_cells)
        arr, ncl = self._colour_all(nlev)
        cfg = VCfg(nlev, pre, post, singular)
        b = np.ascontiguousarray(b, np.float64)
        x = np.empty_like(b)
        rr = C.c_double()
        it = lib().orc_fpcg(C.byref(self.H), arr, ncl, C.byref(cfg), _p(b, f64p), tol, maxit,
                            _p(x, f64p), C.byref(rr))
        return x, it, rr.value

    def colours_of(self, l):
        return self.colours[l], self._ncol[l]
