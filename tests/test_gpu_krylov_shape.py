"""The fused Krylov kernels give the same iterates for every number of rows
per lane group (R = 1, 2, 4 rows gathered together by each lane group, in
workgroups of 1024/R threads): the rows of a workgroup, the order in which
a lane sums a row's entries and the tree of the partial sums are the same
for every R.  Whole time steps with R forced through HFH_KRYLOV_ROWS (and
the lanes per row through HFH_KRYLOV_LANES) must agree bit for bit, for the
Jacobi-PCG velocity systems, the BiCGStab velocity systems of an AMR mesh,
the AMG-BiCGStab and Jacobi-PCG pressure solves, the FP32 corrections of
the mixed form and the segmented systems of a batch.  The diagonal form of
the template rows is checked the same way against the plain template rows.
"""
import numpy as np
import pytest

from higflow_amd import cases, native

pytestmark = pytest.mark.gpu

SEMI = dict(tempdiscrtype=cases.SEMI_IMPLICIT_EULER)
TOL = dict(rtol=1e-10, atol=1e-10, maxit=50000)


def _case(which):
    if which == "ve_cavity":  # symmetric velocity systems: Jacobi-PCG
        return cases.ve_cavity(40, Re=1.0, Wi=0.5, beta=0.5, dt=5e-4, **SEMI, **TOL)
    if which == "contraction":  # WLS rows at level jumps: BiCGStab
        return cases.contraction(levels=2, h0=0.5, length=8.0, radius=1.0, Re=0.01, Wi=1.0, beta=1.0 / 9.0,
                                 dt=1e-4, **SEMI, **TOL)
    if which == "cavity_pcg":  # Jacobi-PCG pressure (HFH_AMG=0)
        return cases.cavity(48, Re=100.0, dt=1e-3, **TOL)
    return [cases.ve_cavity(n, Re=1.0 + n / 8.0, Wi=0.5, beta=0.5, dt=5e-4, **SEMI, **TOL) for n in (16, 24, 40)]


def _run(which, precision, steps=4):
    case = _case(which)
    kw = dict(precision=1, inner_rtol=1e-5) if precision else {}
    if isinstance(case, list):
        sim = native.Batch(case, **kw)
        for _ in range(steps):
            sim.step()
        views = [sim.member(m) for m in range(sim.n)]
    else:
        sim = native.Flow(case, **kw)
        for _ in range(steps):
            sim.step()
        views = [sim]
    out = []
    for v in views:
        out += [v.get(0, 0), v.get(0, 1), v.get(1, 0)]
        if (case[0] if isinstance(case, list) else case)["params"].get("flowtype") == cases.VISCOELASTIC:
            out += [v.get(6, 0), v.get(6, 1)]
    return out


@pytest.mark.parametrize("lanes", [1, 2])
@pytest.mark.parametrize("which,precision", [("ve_cavity", 0), ("contraction", 0), ("cavity_pcg", 0),
                                             ("ve_cavity", 1), ("batch", 0), ("batch", 1)])
def test_rows_per_lane_group_leave_the_iterates_unchanged(which, precision, lanes, monkeypatch):
    monkeypatch.setenv("HFH_KRYLOV_LANES", str(lanes))
    if which == "cavity_pcg":
        monkeypatch.setenv("HFH_AMG", "0")
    runs = {}
    for rows in (1, 2, 4):
        monkeypatch.setenv("HFH_KRYLOV_ROWS", str(rows))
        runs[rows] = _run(which, precision)
    for rows in (2, 4):
        for k, (a, b) in enumerate(zip(runs[1], runs[rows])):
            assert np.all(np.isfinite(a))
            assert np.array_equal(a, b), (which, precision, lanes, rows, k, np.abs(a - b).max())


def _case3d():
    return cases.cavity(12, dim=3, Re=10.0, dt=1e-3, tempdiscrtype=cases.SEMI_IMPLICIT_EULER, **TOL)


@pytest.mark.parametrize("which,precision", [("ve_cavity", 0), ("cavity_pcg", 0), ("ve_cavity", 1), ("batch", 0),
                                             ("batch", 1), ("cavity3d", 0)])
def test_diagonal_form_leaves_the_iterates_unchanged(which, precision, monkeypatch):
    """Template rows in the diagonal form (gather offsets per 1024-row block,
    one lane per row) sum the same entries in the same order as the plain
    template rows (HFH_DIAG=0).  The V-cycle's fine-level residual and the
    true residual of the solves keep their CSR launches in both runs
    (HFH_AMG_DIAG=0, HFH_RESID_DIAG=0): in the diagonal form they sum each
    row on one lane instead of several."""
    monkeypatch.setenv("HFH_KRYLOV_LANES", "1")
    monkeypatch.setenv("HFH_AMG_DIAG", "0")
    monkeypatch.setenv("HFH_RESID_DIAG", "0")
    if which == "cavity_pcg":
        monkeypatch.setenv("HFH_AMG", "0")
    runs = []
    for diag in ("1", "0"):
        monkeypatch.setenv("HFH_DIAG", diag)
        if which == "cavity3d":
            sim = native.Flow(_case3d())
            for _ in range(4):
                sim.step()
            runs.append([sim.get(0, d) for d in range(3)] + [sim.get(1, 0)])
        else:
            runs.append(_run(which, precision))
    for other in runs[1:]:
        for k, (a, b) in enumerate(zip(runs[0], other)):
            assert np.all(np.isfinite(a))
            assert np.array_equal(a, b), (which, precision, k, np.abs(a - b).max())
