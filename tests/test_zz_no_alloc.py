"""hfh_range_noalloc (the entry point that allocates nothing) against
hfh_field_range on the same data, through the C ABI: two calls in a row on
one device, the second with a smaller input, bitwise equal (min and max are
exact).  The test makes its device buffers with the HIP runtime the library
itself links (libamdhip64.so.7, through ctypes)."""
import ctypes as C

import numpy as np
import pytest

from higflow_amd import cases, native

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    def __init__(self):
        native.load()
        self.L = C.CDLL("libamdhip64.so.7")
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipStreamCreate",
                  "hipStreamDestroy", "hipStreamSynchronize"):
            getattr(self.L, f).restype = C.c_int
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(nbytes, 8)) == 0
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        if a.nbytes:
            assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p


def noalloc_range(lib, hip, x_dev, n, stream):
    out = hip.upload(np.full(2, np.nan))
    rc = lib.hfh_range_noalloc(x_dev, C.c_long(n), out, stream)
    assert rc == 0, lib.hfh_last_error()
    assert hip.L.hipStreamSynchronize(stream) == 0
    h = np.zeros(2)
    assert hip.L.hipMemcpy(h.ctypes.data_as(C.c_void_p), out, 16, D2H) == 0
    hip.L.hipFree(out)
    return float(h[0]), float(h[1])


def test_range_noalloc_matches_field_range():
    lib = native.load()
    hip = Hip()
    st = C.c_void_p()
    assert hip.L.hipStreamCreate(C.byref(st)) == 0
    try:
        # 192^2 cavity: 36,864 pressure cells = 5 workgroups; then a 16^2
        # cavity: 256 cells, one workgroup, in the same scratch
        for n in (192, 16):
            sim = native.Flow(cases.cavity(n, Re=100.0, dt=1e-4))
            for _ in range(2):
                sim.step()
            for field, comp in ((1, 0), (0, 0), (0, 1)):
                ref = sim.range(field, comp)
                v = sim.get(field, comp)
                x = hip.upload(v)
                for stream in (st, None):
                    got = noalloc_range(lib, hip, x, len(v), stream)
                    assert got == ref, (n, field, comp, got, ref)
                hip.L.hipFree(x)
                assert ref == (v.min(), v.max())
        # empty input: the values the reference's loop starts from
        assert noalloc_range(lib, hip, None, 0, st) == (1.0e16, -1.0e16)
    finally:
        hip.L.hipStreamDestroy(st)


def test_range_noalloc_rejects_oversize():
    lib = native.load()
    # host-side checks: nothing is launched
    assert lib.hfh_range_noalloc(None, C.c_long(33554432 + 1), None, None) < 0
    assert b"HFH_RANGE_NOALLOC_MAX" in lib.hfh_last_error()
    assert lib.hfh_range_noalloc(None, C.c_long(-1), None, None) < 0
