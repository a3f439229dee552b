"""pm_linear_solve_noalloc: the linear solve that allocates no device memory (its PCG work vectors, partial
sums and scalars live in the library's fixed scratch array), against pm_linear_solve on the same assembled
tangent and right-hand side.  Both are block-Jacobi PCG to rtol 1e-12 on the free DOFs with different SpMV
summation orders, so they agree to the solve's tolerance (1e-8 relative), not bit for bit.  Two calls in a
row share the scratch array, the second on a smaller system; sizes beyond the scratch maximum are refused.

The caller owns the device buffers: they are allocated here through the HIP runtime the library itself
links (libamdhip64.so.7), never inside the entry point.
"""
import ctypes as C

import numpy as np
import pytest

from gpu_util import rel_err
from pmsim_amd import describe
from pmsim_amd import generators as G
from pmsim_amd import mesh as M
from pmsim_amd.solver import Fem, lib

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DeviceVector:
    """A caller-owned device buffer of n doubles."""

    def __init__(self, n):
        self.n = n
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(8 * n)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(8 * self.n), 1) == 0  # host to device

    def get(self):
        out = np.zeros(self.n)
        assert hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, C.c_size_t(8 * self.n), 2) == 0  # device to host
        return out

    def free(self):
        if self.p:
            hip().hipFree(self.p)
            self.p = C.c_void_p()


def noalloc_solve(f, b):
    L = lib()
    L.pm_linear_solve_noalloc.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_int32)]  # device b, x
    db, dx = DeviceVector(f.ndof), DeviceVector(f.ndof)
    try:
        db.put(b)
        dx.put(np.full(f.ndof, np.nan))  # whatever the buffer held before, the solve overwrites it
        it = C.c_int32(0)
        rc = L.pm_linear_solve_noalloc(f.h, db.p, dx.p, C.byref(it))
        if rc != 0:
            raise RuntimeError(L.pm_last_error(f.h).decode())
        return dx.get(), it.value
    finally:
        db.free()
        dx.free()


def _assembled(mesh, base_cfg, seed):
    mesh, base, cfg, ess, nat = base_cfg
    f = Fem(describe(mesh, base, cfg, ess, nat), cfg, keep_history=False)
    rng = np.random.default_rng(seed)
    f.set_ddu(rng.uniform(-2e-3, 2e-3, f.ndof))
    f.update_secondary_values()
    f.assemble_kk()
    return f, rng.normal(size=f.ndof)


def test_noalloc_linear_solve_matches_pm_linear_solve():
    from pmsim_amd.workloads import strip_footing
    import cases
    problems = [strip_footing(48, 24), strip_footing(16, 8), cases.hex20_cube(4, 4, 3, n_inc=2)]
    sizes = []
    for k, prob in enumerate(problems):  # the 2D systems first, the second smaller than the first
        f, b = _assembled(prob[0], prob, 10 + k)
        x_ref, it_ref = f.linear_solve(b)
        x, it = noalloc_solve(f, b)
        sizes.append(f.ndof)
        print(f"ndof {f.ndof}: {it} iterations (pm_linear_solve {it_ref}), rel diff {rel_err(x, x_ref):.2e}", flush=True)
        assert it > 0 and rel_err(x, x_ref) < 1e-8
        # prescribed entries come back zero, as from pm_linear_solve
        assert np.array_equal(x[x_ref == 0.0], x_ref[x_ref == 0.0])
        f.close()
    assert sizes[1] < sizes[0]


def test_noalloc_linear_solve_refuses_oversized_systems():
    """A 1000 x 1000 Qua4 block has 2,004,002 DOFs, beyond the 2,000,000 the scratch array holds."""
    from pmsim_amd import Config, Elem, Essential, FemBase, Natural, ParamSolid, StressStrain
    mesh = G.block(M.QUA4, (1000, 1000), (1.0, 1.0))
    base = FemBase(mesh, [(1, Elem.Solid(ParamSolid(1.0, StressStrain.LinearElastic(1000.0, 0.3), None)))])
    cfg = Config(mesh)
    f = Fem(describe(mesh, base, cfg, Essential(), Natural()), cfg, keep_history=False)
    assert f.ndof == 2004002
    with pytest.raises(RuntimeError, match="exceed the scratch maximum of 2000000"):
        noalloc_solve(f, np.ones(f.ndof))
