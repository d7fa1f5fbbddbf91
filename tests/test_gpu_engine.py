"""The HIP solver-plugin engine (_solver_vtable, higtree/src/solver.h:42-60)
on the reference's solver known answers (higtree/atf-tests/test-solver-2d.c)
and random systems."""
import numpy as np
import pytest

from higflow_amd import native

pytestmark = pytest.mark.gpu


def test_atf_3x3_system():
    # test-solver-2d.c:14-50 -> x = (1, 4, -2)
    m = [[-1.0, 3.0, 5.0], [3.0, -5.0, 7.0], [-1.0, 4.0, -10.0]]
    b = [1.0, -31.0, 35.0]
    s = native.LinearSolver(0, 3)
    s.set_tolerances(1e-12, 1e-14, 1000)
    for i in range(3):
        s.set_Ai(i, [0, 1, 2], m[i])
        s.set_bi(i, b[i])
    s.assemble()
    s.solve()
    x = s.get_x()
    assert np.allclose(x, [1.0, 4.0, -2.0], atol=1e-8)
    assert s.stats()[0] == 1


def test_atf_tridiagonal_100():
    # test-solver-2d.c:224-295: rows (-1, 2[, -1]) with b = 1 at the ends -> x = 1
    n = 100
    s = native.LinearSolver(0, n)
    s.set_tolerances(1e-12, 1e-14, 5000)
    for i in range(n):
        if i == 0:
            s.set_Ai(i, [0, 1], [-1.0, 2.0])
            s.set_bi(i, 1.0)
        elif i == n - 1:
            s.set_Ai(i, [n - 2, n - 1], [-1.0, 2.0])
            s.set_bi(i, 1.0)
        else:
            s.set_Ai(i, [i - 1, i, i + 1], [-1.0, 2.0, -1.0])
            s.set_bi(i, 0.0)
    s.assemble()
    s.solve()
    assert np.allclose(s.get_x(), 1.0, atol=1e-6)


def test_impose_value_and_first_gid():
    n, first = 50, 7
    rng = np.random.default_rng(3)
    A = np.diag(4.0 + rng.random(n)) + np.diag(-np.ones(n - 1), 1) + np.diag(-0.5 * np.ones(n - 1), -1)
    b = rng.normal(size=n)
    s = native.LinearSolver(first, n)
    s.set_tolerances(1e-13, 1e-15, 5000)
    for i in range(n):
        cols = np.nonzero(A[i])[0]
        s.set_Ai(first + i, cols + first, A[i, cols])
        s.set_bi(first + i, b[i])
    s.impose_value(first + 10, 3.25)
    s.set_bi(first + 10, 99.0)  # ignored on the imposed line (solver.c:99-105)
    s.assemble()
    s.solve()
    A2 = A.copy()
    A2[10] = 0.0
    A2[10, 10] = 1.0
    b2 = b.copy()
    b2[10] = 3.25
    assert np.allclose(s.get_x(), np.linalg.solve(A2, b2), atol=1e-9)
