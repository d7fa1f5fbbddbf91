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


def _contraction_pressure(levels, h0, length):
    # the assembled pressure system of a 4:1 contraction (host-only build)
    import os
    import scipy.sparse as sp
    from higflow_amd import cases
    os.environ["HFH_HOST_ONLY"] = "1"
    try:
        f = native.Flow(cases.contraction(levels=levels, h0=h0, length=length, radius=1.0))
        rp, ci, v, sym = f.operator(0)
        f.close()
    finally:
        del os.environ["HFH_HOST_ONLY"]
    return sp.csr_matrix((v, ci, rp))


@pytest.mark.parametrize("amg", ["0", "1"])
def test_fused_bicgstab_on_contraction_pressure(amg, monkeypatch):
    # fused BiCGStab, Jacobi or smoothed-aggregation V-cycle preconditioned,
    # on a real nonsymmetric AMR pressure operator, against a direct solve
    import scipy.sparse.linalg as sla
    A = _contraction_pressure(2, 0.25, 8.0)
    n = A.shape[0]
    x0 = np.random.default_rng(5).normal(size=n)
    b = A @ x0
    monkeypatch.setenv("HFH_SLV_AMG", amg)
    s = native.LinearSolver(0, n)
    s.set_tolerances(1e-12, 0.0, 20000)
    for i in range(n):
        s.set_Ai(i, A.indices[A.indptr[i]:A.indptr[i + 1]], A.data[A.indptr[i]:A.indptr[i + 1]])
        s.set_bi(i, b[i])
    s.assemble()
    s.solve()
    x = s.get_x()
    conv, iters, resid = s.stats()[:3]
    print("amg", amg, "n", n, "iters", iters, "resid", resid)
    xd = sla.spsolve(A.tocsc(), b)
    assert np.linalg.norm(b - A @ x) / np.linalg.norm(b) < 1e-11
    assert np.abs(x - xd).max() / np.abs(xd).max() < 1e-7
