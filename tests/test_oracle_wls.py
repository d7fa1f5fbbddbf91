"""WLS interpolation of the oracle (higtree/src/wls.c) pinned by the
reference's own wls.c (built into oracle/_ref when the reference is mounted),
the HiG-Tree ATF known answers and polynomial reproduction."""
import ctypes as C
import os

import numpy as np
import pytest

import hfo

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REFLIB = os.path.join(ROOT, "oracle", "_ref", "libwls_ref2d.so")


def test_atf_wls_1_degree_known_answers():
    # higtree/atf-tests/test-wls-2d.c:15-97: f = 2 + 3x + 4y on 6 points
    pts = np.array([[-1, -1], [-1, 0], [-1, 1], [1, -1], [1, 0], [1, 1]], dtype=float)
    v = np.array([-5, -1, 3, 1, 5, 9], dtype=float)
    for x, want in (((0.0, 0.0), 2.0), ((0.5, 0.5), 5.5), ((1.0, 1.0), 9.0)):
        w = hfo.wls_weights(2, 1, 10, pts, x)
        assert abs(w @ v - want) < 1e-12


def test_atf_wls_collinear_known_answer():
    # test-wls-2d.c:213-275: collinear points fall back to the dropped column
    pts = np.array([[0.0, 0.0], [1.0, 0.0], [2.0, 0.0], [3.0, 0.0]])
    v = 1.0 + 0.5 * pts[:, 0]
    w = hfo.wls_weights(2, 1, 10, pts, (1.0, 0.0))
    assert abs(w @ v - 1.5) < 1e-12


@pytest.mark.parametrize("seed", range(5))
def test_quadratic_reproduction(seed):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-1, 1, size=(30, 2)) * 0.1 + 3.0
    x = np.array([3.01, 2.995])
    w = hfo.wls_weights(2, 2, 45, pts, x)
    for f in (lambda p: 1 + 0 * p[:, 0], lambda p: p[:, 0], lambda p: p[:, 1],
              lambda p: p[:, 0] * p[:, 1], lambda p: p[:, 0] ** 2 - 2 * p[:, 1] ** 2):
        assert abs(w @ f(pts) - f(x[None, :])[0]) < 1e-9
    # a cubic is not reproduced (the fit is genuinely order 2)
    assert abs(w @ (pts[:, 0] ** 3) - x[0] ** 3) > 1e-12


def _ref_weights(L, pts, x, ord_=2, maxpts=45):
    Point = C.c_double * 2
    L.wls_create.restype = C.c_void_p
    L.wls_create.argtypes = [C.c_uint, C.c_int, C.c_int]
    L.wls_set_points.argtypes = [C.c_void_p, C.c_int, C.POINTER(Point), C.POINTER(C.c_double)]
    L.wls_set_points.restype = C.c_double
    L.wls_calc.argtypes = [C.c_void_p, C.POINTER(C.c_double), C.c_int, C.POINTER(C.c_double)]
    L.wls_destroy.argtypes = [C.c_void_p]
    wls = L.wls_create(2, ord_, maxpts)
    arr = (Point * len(pts))(*[Point(*p) for p in pts])
    xx = (C.c_double * 2)(*x)
    L.wls_set_points(wls, len(pts), arr, xx)
    w = (C.c_double * len(pts))()
    L.wls_calc(wls, xx, len(pts), w)
    L.wls_destroy(wls)
    return np.array(w[:])


@pytest.mark.skipif(not os.path.exists(REFLIB), reason="reference wls.c not built (oracle/_ref)")
@pytest.mark.parametrize("seed", range(8))
def test_bitwise_against_reference_wls(seed):
    L = C.CDLL(REFLIB)
    rng = np.random.default_rng(100 + seed)
    if seed % 2 == 0:  # grid-like samples as the stencil builder produces
        g = (np.arange(-3, 4) + 0.5) / 64.0
        pts = np.array([[a, b] for a in g for b in g])[:45] + 0.5
        x = np.array([0.5 + 1 / 64.0, 0.5])
    else:
        pts = rng.uniform(0, 1, size=(20 + seed, 2))
        x = rng.uniform(0.3, 0.7, size=2)
    ours = hfo.wls_weights(2, 2, 45, pts, x)
    ref = _ref_weights(L, pts, x)
    assert np.array_equal(ours, ref)
