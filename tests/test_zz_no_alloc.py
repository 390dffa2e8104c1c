"""spf_poisson_solve_noalloc against spf_poisson_solve, through the C ABI.

The no-allocation variant solves on caller-owned device buffers (here
allocated by the test through the HIP runtime the library uses, in the
library's z-contiguous layout) with the library's fixed
per-GPU scratch array; it must return the iteration count and the Psi bits of
the run-based solve.  Two calls in a row, the second on a smaller grid, so
the array's leftovers from the first call cannot leak into the second.
"""
import ctypes

import numpy as np
import pytest

from _support import bits, gun_flags, gun_grid
from spiffe_amd import abi
from spiffe_amd.device import Context, DeviceRun

pytestmark = pytest.mark.gpu


def ldz_of(nz):
    return (nz + 1 + 31) // 32 * 32


class Hip:
    """The HIP runtime calls a caller uses to own device buffers (the test, not
    the entry point, allocates)."""

    def __init__(self):
        h = ctypes.CDLL("libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.h = h

    def to_device(self, host, ldz):
        """[iz][ir] host array -> device buffer with element (iz, ir) at ir*ldz + iz."""
        t = np.zeros((host.shape[1], ldz), dtype=host.dtype)
        t[:, :host.shape[0]] = host.T
        p = ctypes.c_void_p()
        assert self.h.hipMalloc(ctypes.byref(p), t.nbytes) == 0
        assert self.h.hipMemcpy(p, t.ctypes.data, t.nbytes, 1) == 0          # host to device
        assert self.h.hipDeviceSynchronize() == 0                            # every stream sees it
        return p, t

    def from_device(self, p, like, nz):
        t = np.empty_like(like)
        assert self.h.hipDeviceSynchronize() == 0
        assert self.h.hipMemcpy(t.ctypes.data, p, t.nbytes, 2) == 0          # device to host
        return t[:, :nz + 1].T.copy()

    def free(self, *ps):
        for p in ps:
            self.h.hipFree(p)


def solve_both(ctx, hip, g, f, psi0, rhs, accuracy, limit):
    d = DeviceRun(ctx, g, f)
    d.upload("Psi", psi0)
    d.upload("Q", rhs)
    it_ref = d.solve_poisson_cyl(0, accuracy, limit)
    ref = d.download("Psi")
    d.close()
    ldz = ldz_of(g.nz)
    (P, tp), (Rh, _), (F, _) = (hip.to_device(psi0, ldz), hip.to_device(rhs, ldz),
                                hip.to_device(np.ascontiguousarray(f, dtype=np.int16), ldz))
    it = ctypes.c_int64()
    rc = ctx.lib.spf_poisson_solve_noalloc(ctx.h, P, Rh, F, g.nz, g.nr, g.dz, g.dr, g.left_bc, g.right_bc, accuracy,
                                           limit, ctypes.byref(it))
    assert rc == 0, ctx.lib.spf_last_error(ctx.h).decode()
    got = hip.from_device(P, tp, g.nz)
    hip.free(P, Rh, F)
    return it_ref, ref, it.value, got


def test_noalloc_solve_matches_the_run_solve_twice():
    rng = np.random.default_rng(17)
    hip = Hip()
    with Context(0) as ctx:
        for (nz, nr, acc, limit) in [(141, 61, 1e-7, 4000), (71, 31, 1e-9, 3000)]:
            g = gun_grid(nz=nz, nr=nr, modes=abi.TM_FIELDS)
            f = gun_flags(g)
            psi0 = rng.standard_normal(g.shape)
            rhs = rng.standard_normal(g.shape) * 1e4
            it_ref, ref, it, got = solve_both(ctx, hip, g, f, psi0, rhs, acc, limit)
            assert it_ref > 0
            assert it == it_ref
            assert np.array_equal(bits(ref), bits(got))


def test_noalloc_rejects_grids_beyond_its_array():
    with Context(0) as ctx:
        it = ctypes.c_int64()
        nz, nr = 4201, 1681
        rc = ctx.lib.spf_poisson_solve_noalloc(ctx.h, 1, None, 1, nz, nr, 1e-5, 1e-5, 0, 0, 1e-6, 100, ctypes.byref(it))
        assert rc == abi.SPF_EINVAL
        assert "no-allocation" in ctx.lib.spf_last_error(ctx.h).decode()
