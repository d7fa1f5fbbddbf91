"""The first pass of every FP64 fused solve runs as a captured hipGraph
(initialisation, a budget of iterations, the true-residual partials and the
copies of flags and partials to pinned memory, one host wait) with
HFH_GRAPH=1; the default runs the same launches from the host with a check
after the predicted count.  Launches after the stopping test return at once, and a solve that
outlasts the budget continues with host-checked chunks, so both give the
same iterates.  The graphs are replayed every step with eager launches of
the step in between; every step is compared bit for bit.
"""
import numpy as np
import pytest

from higflow_amd import cases, native

pytestmark = pytest.mark.gpu

SEMI = dict(tempdiscrtype=cases.SEMI_IMPLICIT_EULER)
TOL = dict(rtol=1e-10, atol=1e-10, maxit=50000)


def _case(which):
    if which == "ve_cavity":  # AMG-PCG pressure and velocity
        return cases.ve_cavity(40, Re=1.0, Wi=0.5, beta=0.5, dt=5e-4, **SEMI, **TOL)
    if which == "contraction":  # AMG-BiCGStab pressure, Jacobi-BiCGStab velocity (WLS rows)
        return cases.contraction(levels=2, h0=0.5, length=8.0, radius=1.0, Re=0.01, Wi=1.0, beta=1.0 / 9.0,
                                 dt=1e-4, **SEMI, **TOL)
    return [cases.ve_cavity(n, Re=1.0 + n / 8.0, Wi=0.5, beta=0.5, dt=5e-4, **SEMI, **TOL) for n in (16, 24, 40)]


def _steps(which, precision, graph, monkeypatch, steps=6):
    monkeypatch.setenv("HFH_GRAPH", graph)
    case = _case(which)
    kw = dict(precision=1, inner_rtol=1e-5) if precision else {}
    batch = isinstance(case, list)
    sim = native.Batch(case, **kw) if batch else native.Flow(case, **kw)
    out = []
    for _ in range(steps):
        sim.step()
        views = [sim.member(m) for m in range(sim.n)] if batch else [sim]
        snap = []
        for v in views:
            snap += [v.get(0, 0), v.get(0, 1), v.get(1, 0), v.get(6, 0)]
        out.append(snap)
    return out


@pytest.mark.parametrize("which,precision", [("ve_cavity", 0), ("contraction", 0), ("batch", 0), ("batch", 1)])
def test_graph_and_host_checked_solves_agree(which, precision, monkeypatch):
    a = _steps(which, precision, "1", monkeypatch)
    b = _steps(which, precision, "0", monkeypatch)
    for s, (sa, sb) in enumerate(zip(a, b)):
        for k, (x, y) in enumerate(zip(sa, sb)):
            assert np.all(np.isfinite(x))
            assert np.array_equal(x, y), (which, precision, s, k, np.abs(x - y).max())
