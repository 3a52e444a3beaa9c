"""hymi_dot_noalloc (block partials in the library's fixed device array, result into caller device memory)
against hymi_dot through the C-ABI: bit for bit, two calls in a row with the second on a smaller input,
FP64 and FP32, one and 32 right-hand sides."""
import ctypes

import numpy as np
import pytest

from helpers import say

pytestmark = pytest.mark.gpu


class DeviceBuffer:
    """Device memory for the result, from the HIP runtime the library itself uses (libamdhip64.so.7, already
    loaded with libhymi.so); torch's bundled HIP runtime is not mixed into this process."""

    def __init__(self, nbytes):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0
        self.nbytes = nbytes

    def fill_nan(self):
        host = np.full(self.nbytes // 8, np.nan)
        assert self.hip.hipMemcpy(self.ptr, host.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(self.nbytes), 1) == 0

    def read(self, n):
        host = np.zeros(self.nbytes // 8)
        assert self.hip.hipMemcpy(host.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(self.nbytes), 2) == 0
        return host[:n]

    def free(self):
        self.hip.hipFree(self.ptr)


def _dot(H, a, b, level, flag):
    out = np.zeros(a.nrhs)
    H._lib.call("hymi_dot", a.h, b.h, level, flag, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out


def _dot_noalloc(H, a, b, level, flag, dev):
    dev.fill_nan()
    H._lib.call("hymi_dot_noalloc", a.h, b.h, level, flag, dev.ptr)
    a.storage.ctx.synchronize()
    return dev.read(a.nrhs)


@pytest.mark.parametrize("prec,nrhs", [("F64", 1), ("F32", 1), ("F64", 32), ("F32", 32)])
def test_dot_noalloc_matches_dot(prec, nrhs, gpu_ctx):
    import hyteg_amd as H
    p = getattr(H, prec)
    say(f"[no-alloc dot] cube n=2 levels 5 then 3, {prec}, nrhs={nrhs}")
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), 2, 2, 2), gpu_ctx)
    rng = np.random.default_rng(7 + nrhs)
    dev = DeviceBuffer(32 * 8)
    a, b = H.P1Function("a", S, 2, 5, precision=p, nrhs=nrhs), H.P1Function("b", S, 2, 5, precision=p, nrhs=nrhs)
    for L in (5, 3):                               # the second call on a smaller input
        n = S.owned_count(L)
        shape = (n,) if nrhs == 1 else (n, nrhs)
        a.upload(L, rng.standard_normal(shape))
        b.upload(L, rng.standard_normal(shape))
        for flag in (H.INNER, H.ALL):
            ref = _dot(H, a, b, L, flag)
            got = _dot_noalloc(H, a, b, L, flag, dev)
            assert np.array_equal(got, ref), (L, flag, got, ref)
            got2 = _dot_noalloc(H, a, a, L, flag, dev)
            assert np.array_equal(got2, _dot(H, a, a, L, flag)), (L, flag)
        say(f"  level {L}: equal to hymi_dot")
    dev.free()


def test_dot_noalloc_rejects_bad_arguments(gpu_ctx):
    """Refused on the host before anything is launched: a null output, functions of different precision, a
    level the functions do not hold."""
    import hyteg_amd as H
    S = H.PrimitiveStorage(H.MeshInfo.meshCuboid((0, 0, 0), (1, 1, 1), 1, 1, 1), gpu_ctx)
    a = H.P1Function("a", S, 2, 3)
    b32 = H.P1Function("b", S, 2, 3, precision=H.F32)
    dev = DeviceBuffer(32 * 8)
    ptr = dev.ptr
    with pytest.raises(H.HymiError):
        H._lib.call("hymi_dot_noalloc", a.h, a.h, 3, H.ALL, None)
    with pytest.raises(H.HymiError):
        H._lib.call("hymi_dot_noalloc", a.h, b32.h, 3, H.ALL, ptr)
    with pytest.raises(H.HymiError):
        H._lib.call("hymi_dot_noalloc", a.h, a.h, 5, H.ALL, ptr)
    dev.free()
