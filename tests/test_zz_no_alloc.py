"""The entry point that allocates nothing (ufvm_compute_residuals_noalloc)
against ufvm_compute_residuals, through the C ABI: the same residual and L1
norm bits, on two calls in a row, the second on a smaller system (its
partials overwrite a prefix of the shared device scratch array)."""
import numpy as np
import pytest

from matlab_ufvm_amd.workloads import cavity_mesh, gaussian_rhs, poisson_system

pytestmark = pytest.mark.gpu


def _system(ctx, n, seed):
    m = cavity_mesh(n, jitter=0.15, seed=seed, renumber_seed=seed + 1)
    ptr, col, ac, anb = poisson_system(m, dirichlet=())
    sid = ctx.system_create(ptr, col)
    ctx.set_coefficients(sid, ac, anb, gaussian_rhs(m, seed=seed + 2))
    x = np.random.default_rng(seed + 3).standard_normal(m.n_cells)
    ctx.set_dphi(sid, 0, x)
    return sid, m.n_cells


class DeviceBuffer:
    """caller-owned device memory (the HIP runtime through ctypes)"""

    def __init__(self, nbytes):
        import ctypes as C
        self.C = C
        self.hip = C.CDLL("libamdhip64.so")
        self.nbytes = nbytes
        self.p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.p), C.c_size_t(nbytes)) == 0
        # 0xff bytes: every double a NaN, so anything not written shows
        assert self.hip.hipMemset(self.p, 0xFF, C.c_size_t(nbytes)) == 0
        assert self.hip.hipDeviceSynchronize() == 0

    def read(self, n):
        out = np.empty(n)
        assert self.hip.hipMemcpy(self.C.c_void_p(out.ctypes.data), self.p,
                                  self.C.c_size_t(out.nbytes), 2) == 0   # device to host
        return out

    def free(self):
        self.hip.hipFree(self.p)


def test_noalloc_residual_matches_the_allocating_entry_point(ctx):
    for n, seed in ((40, 11), (13, 21)):          # 64000 rows (250 blocks), then 2197
        sid, rows = _system(ctx, n, seed)
        print(f"\n[noalloc] rows={rows}", flush=True)
        r_ref, l1_ref = ctx.residual(sid, 0)
        r_dev, l1_dev = DeviceBuffer(rows * 8), DeviceBuffer(8)
        ctx.residual_noalloc_ptr(sid, 0, r_dev.p.value, l1_dev.p.value)
        ctx.synchronize()
        r, l1 = r_dev.read(rows), float(l1_dev.read(1)[0])
        assert np.array_equal(r, r_ref)
        assert l1 == l1_ref and np.isfinite(l1) and l1 > 0
        r_dev.free()
        l1_dev.free()
        ctx.system_destroy(sid)
