"""ctypes binding of libufvm.so (the C ABI declared in include/ufvm.h).

The product path has no CPU fallback: if the HIP library is missing or no GPU
is visible, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libufvm.so")

UFVM_OK = 0
ERRORS = {1: "UFVM_ERR_ARG", 2: "UFVM_ERR_HIP", 3: "UFVM_ERR_STATE", 4: "UFVM_ERR_NUMERIC",
          5: "UFVM_ERR_UNSUPPORTED"}
SMOOTHER = {"ILU": 0, "SOR": 1, "MCGS": 2}
PCG_SMOOTHER = {"MCGS": 0, "ILU": 1, "MCILU": 2, "MCILU0": 3}
CYCLE = {"V-Cycle": 0, "W-Cycle": 1, "F-Cycle": 2}

i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)


class AmgCfg(C.Structure):
    _fields_ = [("cycle", C.c_int32), ("max_cycles", C.c_int32), ("termination", C.c_double),
                ("max_coarse_levels", C.c_int32), ("pre_sweep", C.c_int32),
                ("post_sweep", C.c_int32), ("smoother", C.c_int32), ("rebuild", C.c_int32)]


LAUNCH_MODE = {"auto": 0, "eager": 1, "graphs": 2}


class PcgCfg(C.Structure):
    _fields_ = [("precision", C.c_int32), ("pre_sweep", C.c_int32), ("post_sweep", C.c_int32),
                ("coarse_smoother", C.c_int32),
                ("max_coarse_cells", C.c_int32), ("nullspace", C.c_int32),
                ("max_iter", C.c_int32), ("tol", C.c_double), ("smoother", C.c_int32),
                ("coarse_scale", C.c_double), ("coarse_pre_sweep", C.c_int32),
                ("coarse_post_sweep", C.c_int32), ("launch_mode", C.c_int32),
                ("index_width", C.c_int32)]


class ServerCfg(C.Structure):
    _fields_ = [("max_active", C.c_int32), ("n_streams", C.c_int32), ("hbm_reserve", C.c_double),
                ("max_batch", C.c_int32), ("n_steppers", C.c_int32), ("max_group", C.c_int32),
                ("group_span", C.c_double)]


class RequestInfo(C.Structure):
    _fields_ = [("status", C.c_int32), ("iterations", C.c_int32), ("relres", C.c_double),
                ("queued_s", C.c_double), ("run_s", C.c_double), ("batch", C.c_int32),
                ("group", C.c_int32)]


class MeshDesc(C.Structure):
    _fields_ = [("n_cells", C.c_int32), ("n_faces", C.c_int32), ("n_interior", C.c_int32),
                ("owner", i32p), ("neighbour", i32p),
                ("Sf", f64p), ("area", f64p), ("CN", f64p), ("eCN", f64p), ("gDiff", f64p),
                ("T", f64p), ("gf", f64p), ("face_centroid", f64p),
                ("volume", f64p), ("centroid", f64p),
                ("cconn_ptr", i32p), ("cconn", i32p), ("own_nnz", i32p), ("nei_nnz", i32p),
                ("n_patches", C.c_int32), ("patch_start", i32p), ("patch_nfaces", i32p),
                ("patch_type", i32p)]


class SimpleCfg(C.Structure):
    _fields_ = [("nu", C.c_double), ("urf_u", C.c_double), ("urf_p", C.c_double),
                ("n_correctors", C.c_int32), ("p_ref_cell", C.c_int32), ("u_sweeps", C.c_int32),
                ("u_bc", i32p), ("u_value", f64p), ("p_bc", i32p), ("p_solver", C.c_int32),
                ("p_amg", AmgCfg), ("p_pcg", PcgCfg), ("u_smoother", C.c_int32),
                ("stress_term", C.c_int32), ("p_value", f64p), ("u_init", C.c_double * 3),
                ("p_init", C.c_double), ("rho", C.c_double), ("transient", C.c_int32),
                ("dt", C.c_double)]


# name -> (restype, argtypes)
_SIGS = {
    "ufvm_ctx_create": (C.c_int, [C.c_int32, C.POINTER(C.c_void_p)]),
    "ufvm_ctx_destroy": (C.c_int, [C.c_void_p]),
    "ufvm_last_error": (C.c_char_p, [C.c_void_p]),
    "ufvm_ctx_stream": (C.c_void_p, [C.c_void_p]),
    "ufvm_ctx_synchronize": (C.c_int, [C.c_void_p]),
    "ufvm_system_create": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, i32p, i32p, i32p]),
    "ufvm_system_destroy": (C.c_int, [C.c_void_p, C.c_int32]),
    "ufvm_system_set_coefficients": (C.c_int, [C.c_void_p, C.c_int32, f64p, f64p, f64p]),
    "ufvm_system_set_coefficients_dev": (C.c_int, [C.c_void_p, C.c_int32, C.c_void_p,
                                                   C.c_void_p, C.c_void_p]),
    "ufvm_system_update_coefficients": (C.c_int, [C.c_void_p, C.c_int32, f64p, f64p, f64p]),
    "ufvm_system_set_dphi": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, f64p]),
    "ufvm_system_get_dphi": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, f64p]),
    "ufvm_system_set_bc": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, f64p]),
    "ufvm_system_get_bc": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, f64p]),
    "ufvm_compute_residuals": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, f64p, f64p]),
    "ufvm_compute_residuals_noalloc": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p,
                                                 C.c_void_p]),
    "ufvm_amg_agglomerate": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, i32p]),
    "ufvm_amg_num_levels": (C.c_int, [C.c_void_p, C.c_int32, i32p]),
    "ufvm_amg_level_info": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, i32p, i32p]),
    "ufvm_amg_level_get": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, i32p, i32p, i32p, f64p,
                                     f64p]),
    "ufvm_solve_algebraic_system": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32,
                                              C.c_int32, C.c_double, f64p, f64p]),
    "ufvm_apply_amg": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(AmgCfg), f64p, f64p, i32p]),
    "ufvm_pcg_setup": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(PcgCfg)]),
    "ufvm_pcg_solve": (C.c_int, [C.c_void_p, C.c_int32, i32p, f64p]),
    "ufvm_pcg_solve_batch": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                       i32p, f64p]),
    "ufvm_pcg_solve_group": (C.c_int, [C.c_void_p, C.c_int32, i32p, C.POINTER(f64p),
                                       C.POINTER(f64p), i32p, f64p]),
    "ufvm_pcg_apply_vcycle": (C.c_int, [C.c_void_p, C.c_int32, f64p, f64p]),
    "ufvm_pcg_level_colours": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, i32p, i32p]),
    "ufvm_pcg_level_layout": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, i32p,
                                        C.POINTER(C.c_int64)]),
    "ufvm_case_create": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(SimpleCfg), i32p]),
    "ufvm_case_destroy": (C.c_int, [C.c_void_p, C.c_int32]),
    "ufvm_case_set_field": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, f64p, C.c_int64]),
    "ufvm_case_get_field": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, f64p, C.c_int64]),
    "ufvm_case_system": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, i32p]),
    "ufvm_update_fields": (C.c_int, [C.c_void_p, C.c_int32]),
    "ufvm_assemble_equation": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32, f64p,
                                         f64p]),
    "ufvm_solve_equation": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, f64p, f64p, i32p]),
    "ufvm_correct_equation": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, C.c_int32]),
    "ufvm_simple_iterate": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, f64p]),
    "ufvm_transient_step": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_double, C.c_double,
                                      f64p, i32p]),
    "ufvm_transient_update": (C.c_int, [C.c_void_p, C.c_int32]),
    "ufvm_server_create": (C.c_int, [C.c_int32, C.POINTER(ServerCfg), C.POINTER(C.c_void_p)]),
    "ufvm_server_destroy": (C.c_int, [C.c_void_p]),
    "ufvm_server_last_error": (C.c_char_p, [C.c_void_p]),
    "ufvm_server_add_system": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, i32p, i32p, f64p, f64p,
                                         C.POINTER(PcgCfg), i32p]),
    "ufvm_server_add_case": (C.c_int, [C.c_void_p, C.POINTER(MeshDesc), C.POINTER(SimpleCfg),
                                       i32p]),
    "ufvm_server_get_field": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, f64p, C.c_int64]),
    "ufvm_server_set_field": (C.c_int, [C.c_void_p, C.c_int32, C.c_char_p, f64p, C.c_int64]),
    "ufvm_server_submit_solve": (C.c_int, [C.c_void_p, C.c_int32, f64p, f64p,
                                           C.POINTER(C.c_int64)]),
    "ufvm_server_submit_step": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, f64p,
                                          C.POINTER(C.c_int64)]),
    "ufvm_server_query": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(RequestInfo)]),
    "ufvm_server_wait": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(RequestInfo)]),
    "ufvm_server_release": (C.c_int, [C.c_void_p, C.c_int64]),
    "ufvm_server_stats": (C.c_int, [C.c_void_p, i32p, i32p, C.POINTER(C.c_int64)]),
    "ufvm_prof_enable": (C.c_int, [C.c_void_p, C.c_int32]),
    "ufvm_prof_read": (C.c_int, [C.c_void_p, C.c_char_p, f64p, C.POINTER(C.c_int64), f64p]),
    "ufvm_prof_reset": (C.c_int, [C.c_void_p]),
}

EXPORTED = tuple(_SIGS)
_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libufvm.so; raises if it is absent (there is no CPU fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"libufvm.so not built ({path}); run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class UfvmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else None


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Context:
    """One GPU context (HIP stream + device-resident systems)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.ufvm_ctx_create(device, C.byref(h))
        if rc != UFVM_OK:
            raise UfvmError(rc, "ufvm_ctx_create failed (no GPU visible?)")
        self.h = h

    def check(self, rc):
        if rc != UFVM_OK:
            raise UfvmError(rc, self.lib.ufvm_last_error(self.h).decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.ufvm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        return int(self.lib.ufvm_ctx_stream(self.h) or 0)

    def synchronize(self):
        self.check(self.lib.ufvm_ctx_synchronize(self.h))

    # -------------------------------------------------------------- systems
    def system_create(self, ptr, col) -> int:
        ptr, col = _i32(ptr), _i32(col)
        sid = C.c_int32()
        self.check(self.lib.ufvm_system_create(self.h, ptr.size - 1, col.size, _p(ptr, i32p),
                                               _p(col, i32p), C.byref(sid)))
        return sid.value

    def system_destroy(self, sid):
        self.check(self.lib.ufvm_system_destroy(self.h, sid))

    def set_coefficients(self, sid, ac=None, anb=None, bc=None):
        ac = _f64(ac) if ac is not None else None
        anb = _f64(anb) if anb is not None else None
        bc = _f64(bc) if bc is not None else None
        self.check(self.lib.ufvm_system_set_coefficients(self.h, sid, _p(ac, f64p),
                                                         _p(anb, f64p), _p(bc, f64p)))

    def level_info(self, sid, level):
        n, nnz = C.c_int32(), C.c_int32()
        self.check(self.lib.ufvm_amg_level_info(self.h, sid, level, C.byref(n), C.byref(nnz)))
        return n.value, nnz.value

    def update_coefficients(self, sid, ac=None, anb=None, bc=None):
        """New level-1 values on the existing hierarchy (Galerkin update and,
        after pcg_setup, a refresh of the solve-side copies)."""
        arrs = [None if v is None else _f64(v) for v in (ac, anb, bc)]
        self.check(self.lib.ufvm_system_update_coefficients(
            self.h, sid, *[None if v is None else _p(v, f64p) for v in arrs]))

    def set_dphi(self, sid, level, x):
        x = _f64(x)
        self.check(self.lib.ufvm_system_set_dphi(self.h, sid, level, _p(x, f64p)))

    def get_dphi(self, sid, level=0):
        n, _ = self.level_info(sid, level)
        x = np.empty(n)
        self.check(self.lib.ufvm_system_get_dphi(self.h, sid, level, _p(x, f64p)))
        return x

    def set_bc(self, sid, level, b):
        b = _f64(b)
        self.check(self.lib.ufvm_system_set_bc(self.h, sid, level, _p(b, f64p)))

    def get_bc(self, sid, level=0):
        n, _ = self.level_info(sid, level)
        b = np.empty(n)
        self.check(self.lib.ufvm_system_get_bc(self.h, sid, level, _p(b, f64p)))
        return b

    def residual(self, sid, level=0):
        n, _ = self.level_info(sid, level)
        r = np.empty(n)
        l1 = C.c_double()
        self.check(self.lib.ufvm_compute_residuals(self.h, sid, level, _p(r, f64p), C.byref(l1)))
        return r, l1.value

    def residual_noalloc_ptr(self, sid, level, r_addr, l1_addr):
        """ufvm_compute_residuals_noalloc: r and its L1 norm into device memory
        given by address (e.g. torch tensors' data_ptr()), enqueued on the
        context's stream; nothing is allocated and nothing waited for."""
        self.check(self.lib.ufvm_compute_residuals_noalloc(self.h, sid, level,
                                                           C.c_void_p(r_addr),
                                                           C.c_void_p(l1_addr)))

    def agglomerate(self, sid, max_coarse_levels=10) -> int:
        nl = C.c_int32()
        self.check(self.lib.ufvm_amg_agglomerate(self.h, sid, max_coarse_levels, C.byref(nl)))
        return nl.value

    def num_levels(self, sid) -> int:
        nl = C.c_int32()
        self.check(self.lib.ufvm_amg_num_levels(self.h, sid, C.byref(nl)))
        return nl.value

    def level_get(self, sid, level):
        n, nnz = self.level_info(sid, level)
        nl = self.num_levels(sid)
        par = np.empty(n, np.int32) if level < nl - 1 else None
        ptr = np.empty(n + 1, np.int32)
        col = np.empty(max(nnz, 1), np.int32)
        ac = np.empty(n)
        anb = np.empty(max(nnz, 1))
        self.check(self.lib.ufvm_amg_level_get(self.h, sid, level, _p(par, i32p), _p(ptr, i32p),
                                               _p(col, i32p), _p(ac, f64p), _p(anb, f64p)))
        return dict(parents=par, ptr=ptr, col=col[:nnz], ac=ac, anb=anb[:nnz])

    def solve_algebraic_system(self, sid, level, smoother, n_iter, rrf=0.1):
        r0, r1 = C.c_double(), C.c_double()
        self.check(self.lib.ufvm_solve_algebraic_system(self.h, sid, level, SMOOTHER[smoother],
                                                        n_iter, rrf, C.byref(r0), C.byref(r1)))
        return r0.value, r1.value

    def apply_amg(self, sid, cycle="V-Cycle", max_cycles=30, termination=0.1,
                  max_coarse_levels=10, pre_sweep=1, post_sweep=3, smoother="ILU", rebuild=1):
        cfg = AmgCfg(CYCLE[cycle], max_cycles, termination, max_coarse_levels, pre_sweep,
                     post_sweep, SMOOTHER[smoother], rebuild)
        r0, r1, nc = C.c_double(), C.c_double(), C.c_int32()
        self.check(self.lib.ufvm_apply_amg(self.h, sid, C.byref(cfg), C.byref(r0), C.byref(r1),
                                           C.byref(nc)))
        return r0.value, r1.value, nc.value

    def pcg_setup(self, sid, precision=32, pre_sweep=1, post_sweep=1, max_coarse_cells=2000,
                  nullspace=-1, max_iter=500, tol=1e-8, coarse_smoother=1, smoother="MCGS",
                  coarse_scale=2.5, coarse_pre_sweep=2, coarse_post_sweep=2,
                  launch_mode="auto", index_width=0):
        """smoother (finest level): "MCGS" multicolour Gauss-Seidel, "ILU" the
        reference's DILU in its lexicographic order (level-scheduled), "MCILU"
        DILU in the colour order of "MCGS" (one dependency level per colour),
        "MCILU0" ILU(0) in that colour order (equal to "MCILU" on graphs
        without triangles);
        coarse_scale: factor on every coarse-grid correction; coarse_*_sweep:
        sweeps below the finest level (-1: as on the finest); the defaults
        2.5 and 2 / 2 were tuned on the 2.1M-cell hex and 5.3M-cell tet
        systems.  launch_mode: "auto" (captured graphs below 600k cells),
        "eager" or "graphs".  index_width: 0 (16-bit column offsets on the
        levels where they fit) or 4 (int32 columns everywhere)."""
        cfg = PcgCfg(precision, pre_sweep, post_sweep, coarse_smoother, max_coarse_cells,
                     nullspace, max_iter, tol, PCG_SMOOTHER[smoother], coarse_scale,
                     coarse_pre_sweep, coarse_post_sweep, LAUNCH_MODE[launch_mode], index_width)
        self.check(self.lib.ufvm_pcg_setup(self.h, sid, C.byref(cfg)))

    def pcg_solve(self, sid):
        it, rr = C.c_int32(), C.c_double()
        self.check(self.lib.ufvm_pcg_solve(self.h, sid, C.byref(it), C.byref(rr)))
        return it.value, rr.value

    def pcg_solve_batch(self, sid, B):
        """k = len(B) (1, 2, 4 or 8) solves of the system at once; B: (k, n)
        right-hand sides in level-1 order.  Returns (X (k, n), iterations (k,),
        relative residuals (k,))."""
        B = np.ascontiguousarray(B, np.float64)
        X = np.empty_like(B)
        its, rel = self.pcg_solve_batch_ptr(sid, B.shape[0], B.ctypes.data, X.ctypes.data)
        return X, its, rel

    def pcg_solve_batch_ptr(self, sid, k, b_addr, x_addr):
        """As pcg_solve_batch with the k x n arrays given by address (host or
        device memory, e.g. torch tensors' data_ptr())."""
        its = np.zeros(k, np.int32)
        rel = np.zeros(k)
        self.check(self.lib.ufvm_pcg_solve_batch(self.h, sid, k, C.c_void_p(b_addr),
                                                 C.c_void_p(x_addr), its.ctypes.data_as(i32p),
                                                 rel.ctypes.data_as(f64p)))
        return its, rel

    def pcg_solve_group(self, sids, bs):
        """len(sids) solves of different systems at once, every phase of an
        iteration one launch for all (ufvm_pcg_solve_group); bs: their
        right-hand sides in level-1 order.  Returns (solutions, iterations,
        relative residuals)."""
        g = len(sids)
        bs = [_f64(b) for b in bs]
        xs = [np.empty_like(b) for b in bs]
        ids = np.ascontiguousarray(sids, np.int32)
        bp = (f64p * g)(*[_p(b, f64p) for b in bs])
        xp = (f64p * g)(*[_p(x, f64p) for x in xs])
        its = np.zeros(g, np.int32)
        rel = np.zeros(g)
        self.check(self.lib.ufvm_pcg_solve_group(self.h, g, _p(ids, i32p), bp, xp,
                                                 its.ctypes.data_as(i32p),
                                                 rel.ctypes.data_as(f64p)))
        return xs, its, rel

    def pcg_apply_vcycle(self, sid, r):
        r = _f64(r)
        z = np.empty_like(r)
        self.check(self.lib.ufvm_pcg_apply_vcycle(self.h, sid, _p(r, f64p), _p(z, f64p)))
        return z

    def pcg_level_colours(self, sid, level):
        n, _ = self.level_info(sid, level)
        nc = C.c_int32()
        col = np.empty(n, np.int32)
        self.check(self.lib.ufvm_pcg_level_colours(self.h, sid, level, C.byref(nc),
                                                   _p(col, i32p)))
        return nc.value, col

    def pcg_level_layout(self, sid, level):
        """(bytes per column index of the level's sliced-ELL copy: 2, 4 or 0
        without one, slots including padding)"""
        ib, sl = C.c_int32(), C.c_int64()
        self.check(self.lib.ufvm_pcg_level_layout(self.h, sid, level, C.byref(ib), C.byref(sl)))
        return ib.value, sl.value

    # ------------------------------------------------------------ profiling
    def prof_enable(self, on=True):
        """on: False/0 off, True/1 every launch, k > 1 every k-th launch per kind."""
        self.check(self.lib.ufvm_prof_enable(self.h, int(on)))

    def prof_reset(self):
        self.check(self.lib.ufvm_prof_reset(self.h))

    def prof_read(self, kind):
        ms, n, by = C.c_double(), C.c_int64(), C.c_double()
        self.check(self.lib.ufvm_prof_read(self.h, kind.encode(), C.byref(ms), C.byref(n),
                                           C.byref(by)))
        return ms.value, n.value, by.value
