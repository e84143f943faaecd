"""h3d_fdot2_noalloc, the allocation-free Fdot1/Fdot2 (no context; its
scratch is a fixed device array per GPU), against the context entry points
h3d_fdot1/h3d_fdot2 on the same arrays: bitwise equal, on two calls in a
row with the second on a smaller grid (the scratch keeps the first call's
contents), with y = NULL (Fdot1), on a stream of the caller's and on the null
stream.  Sizes beyond the scratch's capacity are rejected before any work.
The caller-side buffers and stream come from the HIP runtime libh3d runs on."""
import ctypes

import numpy as np
import pytest

from tests.conftest import gpu_available

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not gpu_available(), reason="no HIP device")]

from h2_amd import _lib as L
from h2_amd.heat3ds import WorkBuffers


def hip():
    """The HIP runtime libh3d itself runs on (already loaded with it), for
    the caller-side buffers and stream of these tests."""
    L.load()
    rt = ctypes.CDLL("libamdhip64.so.7")
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    rt.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    rt.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    return rt


H2D, D2H = 1, 2


def device_copy(rt, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    p = ctypes.c_void_p()
    assert rt.hipMalloc(ctypes.byref(p), a.nbytes) == 0
    assert rt.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def context_dots(x, y):
    with WorkBuffers(*x.shape) as wk:
        wk.upload("pcg_p", x)
        wk.upload("pcg_r", y)
        return wk.fdot2("pcg_p", "pcg_r"), wk.fdot1("pcg_p")


def noalloc_dots(rt, x, y, stream):
    lib = L.load()
    dx = device_copy(rt, np.asfortranarray(x).ravel(order="F"))
    dy = device_copy(rt, np.asfortranarray(y).ravel(order="F"))
    out = device_copy(rt, np.full(2, np.nan))
    try:
        MX, MY, MZ = x.shape
        assert lib.h3d_fdot2_noalloc(dx, dy, MX, MY, MZ, out, stream) == L.OK
        assert lib.h3d_fdot2_noalloc(dx, None, MX, MY, MZ, ctypes.c_void_p(out.value + 8), stream) == L.OK
        assert rt.hipStreamSynchronize(stream) == 0
        r = np.zeros(2)
        assert rt.hipMemcpy(r.ctypes.data, out, 16, D2H) == 0
    finally:
        for p in (dx, dy, out):
            rt.hipFree(p)
    return float(r[0]), float(r[1])


@pytest.mark.parametrize("use_side_stream", [True, False])
def test_noalloc_dot_equals_context_dot(use_side_stream):
    rt = hip()
    rng = np.random.default_rng(42)
    stream = ctypes.c_void_p()
    if use_side_stream:
        assert rt.hipStreamCreate(ctypes.byref(stream)) == 0
    try:
        for shape in ((242, 66, 32), (21, 12, 9)):
            print("case noalloc dot shape", shape, "side stream", use_side_stream, flush=True)
            x = np.asfortranarray(rng.normal(size=shape) * np.exp(rng.normal(0, 3, size=shape)))
            y = np.asfortranarray(rng.normal(size=shape))
            ref2, ref1 = context_dots(x, y)
            got2, got1 = noalloc_dots(rt, x, y, stream)
            assert got2 == ref2 and got1 == ref1, (got2, ref2, got1, ref1)
    finally:
        if use_side_stream:
            rt.hipStreamDestroy(stream)


def test_noalloc_dot_rejects_oversize_and_bad_args():
    rt = hip()
    lib = L.load()
    p = device_copy(rt, np.zeros(8))
    try:
        # ceil(NX/64)·ceil(NY/4) = 129·128 > 16384 workgroups: rejected before any work
        assert lib.h3d_fdot2_noalloc(p, p, 64 * 129 + 2, 4 * 128 + 2, 3, p, None) == L.ERR_UNSUPPORTED
        assert lib.h3d_fdot2_noalloc(None, p, 10, 10, 10, p, None) == L.ERR_ARG
        assert lib.h3d_fdot2_noalloc(p, p, 10, 10, 10, None, None) == L.ERR_ARG
        assert lib.h3d_fdot2_noalloc(p, p, 2, 10, 10, p, None) == L.ERR_ARG
    finally:
        rt.hipFree(p)
