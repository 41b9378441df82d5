"""wtb_bse_matvec_noalloc, the allocation-free variant of wtb_bse_matvec: the
caller passes device arrays X and Y, and the split-FP16 product takes its
FP16 planes and column scales from the library's static scratch array
instead of workspace it allocates.  Checked against wtb_bse_matvec (host
arrays, allocated workspace) bit for bit, on two calls in a row with the
second one smaller, for the MIXED product and the FP64 / FP32 products.
The test's device arrays come from the HIP runtime the library itself is
linked against (hipMalloc through ctypes), not from a second runtime."""
import ctypes as C
import sys

import numpy as np
import pytest

from conftest import hex_model
from wtb_amd import hip
from wtb_amd import tbmodel as T

pytestmark = pytest.mark.gpu

COUL = dict(kind="V2DRK", r0=40.0, ediel=(1.0, 1.0, 1.0), ez=1.0, w=0.0, lc=1.0, ktol=0.001)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


class DevArray:
    """a hipMalloc'd complex128 (N, ncol) column-major buffer"""
    rt = None

    def __init__(self, n_elems):
        if DevArray.rt is None:
            DevArray.rt = C.CDLL("/opt/rocm/lib/libamdhip64.so.7")
            DevArray.rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            DevArray.rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            DevArray.rt.hipFree.argtypes = [C.c_void_p]
        self.p = C.c_void_p()
        self.nbytes = 16 * n_elems
        assert DevArray.rt.hipMalloc(C.byref(self.p), self.nbytes) == 0

    def put(self, a):
        a = np.asfortranarray(a, np.complex128)
        assert a.nbytes == self.nbytes
        assert DevArray.rt.hipMemcpy(self.p, C.c_void_p(a.ctypes.data), self.nbytes, 1) == 0   # host to device

    def get(self, shape):
        out = np.empty(shape, np.complex128, order="F")
        assert DevArray.rt.hipMemcpy(C.c_void_p(out.ctypes.data), self.p, self.nbytes, 2) == 0  # device to host
        return out

    def free(self):
        DevArray.rt.hipFree(self.p)


def noalloc(ctx, Xd, Yd, ncol, precision):
    return ctx.lib.wtb_bse_matvec_noalloc(ctx.h, ncol, Xd.p, Yd.p, precision)


@pytest.mark.parametrize("store,precision", [(hip.MIXED, hip.MIXED), (hip.MIXED, hip.FP32), (hip.MIXED, hip.FP64),
                                             (hip.FP64, hip.FP64)])
def test_noalloc_matvec_matches_the_allocating_entry_point(store, precision):
    say("case noalloc matvec", store, precision)
    mesh = (10, 10, 1)
    m = hex_model(mesh)
    k = T.mp_mesh(*mesh, m.rlat)
    ctx = hip.Context(0)
    ctx.set_tb_model(m)
    ctx.set_kpoints(k, mesh)
    ctx.hk_eig(2, 2, host=False)
    ctx.bse_build(COUL, store)
    N = ctx.N
    rng = np.random.default_rng(3)
    for ncol in (40, 5):                  # the second call smaller than the first
        X = rng.standard_normal((N, ncol)) + 1j * rng.standard_normal((N, ncol))
        X[:, 0] *= 1e-3
        Yref = ctx.matvec(X, precision)
        Xd, Yd = DevArray(N * ncol), DevArray(N * ncol)
        Xd.put(X)
        Yd.put(np.full((N, ncol), complex(np.nan, np.nan)))
        st = noalloc(ctx, Xd, Yd, ncol, precision)
        assert st == 0, ctx.lib.wtb_last_error(ctx.h)
        Y = Yd.get((N, ncol))
        Xd.free()
        Yd.free()
        say("ncol", ncol, "max |Y - Y_ref|", np.abs(Y - Yref).max())
        assert np.array_equal(Y, Yref)
    # host arrays and the int8 product are refused, not run
    Xh = np.zeros((N, 1), np.complex128)
    f = ctx.lib.wtb_bse_matvec_noalloc
    assert f(ctx.h, 1, C.c_void_p(Xh.ctypes.data), C.c_void_p(Xh.ctypes.data), precision) == -1
    Xd = DevArray(N)
    assert f(ctx.h, 1, Xd.p, Xd.p, hip.INT8X) == -1
    Xd.free()
    ctx.close()
