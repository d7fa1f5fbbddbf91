"""Batched parameter sweeps on the GPU: every member of a batch matches the
CPU oracle run of that simulation alone (FP64 form: 1e-10 relative) and the
single-run GPU step, through the C ABI."""
import numpy as np
import pytest

import hfo
from higflow_amd import cases, native

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol=1e-13, atol=1e-15, maxit=50000)


def worst_rel(a_sim, b_sim, dim, ve):
    keys = [(0, d) for d in range(dim)] + [(1, 0)]
    if ve:
        keys += [(6, m) for m in range(dim * dim)] + [(7, m) for m in range(dim * dim)]
    ua = max(np.abs(b_sim.get(0, d)).max() for d in range(dim))
    out = {}
    for f, c in keys:
        a, b = a_sim.get(f, c), b_sim.get(f, c)
        scale = ua if f == 0 else np.abs(b).max()
        out[(f, c)] = np.abs(a - b).max() / scale if scale > 0 else np.abs(a).max()
    return out


def test_batch_members_match_oracle():
    # three Oldroyd-B cavities of a sweep (uneven sides, own Re and Wi),
    # semi-implicit Euler, log-conformation
    rows = [(8, 2.5, 0.4), (12, 20.0, 1.5), (16, 1.0, 0.8)]
    cs = cases.sweep(table=rows, dt=5e-4, **TIGHT)
    b = native.Batch(cs)
    oracles = [hfo.Oracle(c) for c in cs]
    for _ in range(4):
        b.step()
        for o in oracles:
            o.step()
    for m, o in enumerate(oracles):
        w = worst_rel(b.member(m), o, 2, True)
        bad = {k: v for k, v in w.items() if not v <= 1e-10}
        assert not bad, (m, bad)
        for which in (0, 1, 2):
            # rtol 1e-13 is at the edge of FP64: the flag may miss it by a
            # small factor, the true residual stays far below the 1e-10 bar
            st = b.solve_stats(m, which)
            assert st["resid"] < 1e-11, (m, which, st)


def test_batch_matches_single_runs():
    # Newtonian cavities with explicit Euler: each member against its own
    # single-run GPU step
    cs = [cases.cavity(n, Re=re, dt=1e-3, **TIGHT) for n, re in ((16, 100.0), (24, 10.0), (8, 400.0), (32, 50.0))]
    b = native.Batch(cs)
    singles = [native.Flow(c) for c in cs]
    for _ in range(10):
        b.step()
        for s in singles:
            s.step()
    for m, s in enumerate(singles):
        w = worst_rel(b.member(m), s, 2, False)
        bad = {k: v for k, v in w.items() if not v <= 1e-10}
        assert not bad, (m, bad)
        st, ss = b.solve_stats(m, 0), s.solve_stats(0)
        assert st["converged"] == 1 and ss["converged"] == 1


def test_batch_seeded_sweep_rows_scaled_down():
    # the first rows of the seed-0 sweep table (Re, Wi as drawn, sides
    # divided by 8 so the oracle finishes in seconds)
    rows = [(n // 8, re, wi) for n, re, wi in cases.sweep_table(64, 0)[:4]]
    cs = cases.sweep(table=rows, **TIGHT)
    b = native.Batch(cs)
    oracles = [hfo.Oracle(c) for c in cs]
    for _ in range(3):
        b.step()
        for o in oracles:
            o.step()
    for m, o in enumerate(oracles):
        w = worst_rel(b.member(m), o, 2, True)
        bad = {k: v for k, v in w.items() if not v <= 1e-10}
        assert not bad, (m, rows[m], bad)


def test_batch_fields_and_boundary_values():
    cs = cases.sweep(table=[(8, 1.0, 0.5), (16, 5.0, 1.0)], **TIGHT)
    b = native.Batch(cs)
    assert b.L.hfh_batch_size(b.h) == 2
    for m, n in enumerate((8, 16)):
        assert b.size(m, 1) == n * n
        v = np.arange(n * n, dtype=np.float64)
        b.set(m, 1, 0, v)
        assert np.array_equal(b.get(m, 1), v)
        b.set(m, 1, 0, np.zeros(n * n))
    # a member's lid speed changed through its boundary values: the same
    # step as a single run with that lid
    lid_bc = 3  # walls: x-low, x-high, y-low, y-high (the lid)
    single = native.Flow(cases.ve_cavity(16, Re=5.0, Wi=1.0, beta=0.5, dt=5e-4, lid=2.0,
                                         tempdiscrtype=cases.SEMI_IMPLICIT_EULER, **TIGHT))
    nb = b.L.hfh_bc_count(single.h, 1, lid_bc)
    vals = np.full(nb, 2.0)
    assert b.L.hfh_batch_bc_set_values(b.h, 1, 1, lid_bc, vals.ctypes.data_as(native.C.POINTER(native.C.c_double))) == nb
    for _ in range(3):
        b.step()
        single.step()
    w = worst_rel(b.member(1), single, 2, True)
    assert max(w.values()) <= 1e-10, w


def test_batch_3d_members_match_oracle():
    # 3D members of different sizes and Reynolds numbers, semi-implicit Euler
    cs = [cases.cavity(n, dim=3, Re=re, dt=1e-3, tempdiscrtype=cases.SEMI_IMPLICIT_EULER, **TIGHT)
          for n, re in ((4, 10.0), (6, 50.0))]
    b = native.Batch(cs)
    oracles = [hfo.Oracle(c) for c in cs]
    for _ in range(3):
        b.step()
        for o in oracles:
            o.step()
    for m, o in enumerate(oracles):
        w = worst_rel(b.member(m), o, 3, False)
        bad = {k: v for k, v in w.items() if not v <= 1e-10}
        assert not bad, (m, bad)


@pytest.mark.parametrize("force32", ["0", "1"])
def test_batch_mixed_precision_error(force32, monkeypatch):
    # mixed form on a batch: the block-diagonal systems are refined together
    # (FP32 corrections, FP64 residuals), each member against its own
    # tolerance; forced on every solve or chosen by timing
    monkeypatch.setenv("HFH_MIXED_FP32", force32)
    rows = [(16, 2.5, 0.4), (24, 20.0, 1.5), (32, 1.0, 0.8)]
    tol = dict(rtol=1e-10, atol=1e-10, maxit=50000)
    b64 = native.Batch(cases.sweep(table=rows, dt=5e-4, **tol))
    bmx = native.Batch(cases.sweep(table=rows, dt=5e-4, **tol), precision=1, inner_rtol=1e-5)
    for _ in range(6):
        b64.step()
        bmx.step()
    for m in range(len(rows)):
        for f, c in ((0, 0), (0, 1), (1, 0)):
            a, b = bmx.get(m, f, c), b64.get(m, f, c)
            assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b), (m, f, c)
        for which in (0, 1, 2):
            st = bmx.solve_stats(m, which)
            assert st["converged"] == 1, (m, which, st)
        if force32 == "1":
            # the velocity systems (well conditioned) refine in FP32
            for which in (1, 2):
                st = bmx.solve_stats(m, which)
                assert st["fallback"] == 0 and st["outer"] >= 1, (m, which, st)


def test_batch_ignores_recycled_device_memory():
    # the padding between member blocks is never written by the step: it must
    # not depend on what a previous user left in the memory the batch gets
    # (a solver whose state is NaN is freed just before)
    junk = native.Flow(cases.cavity(48, dim=3, **TIGHT))
    for f, c in [(0, d) for d in range(3)] + [(1, 0), (2, 0), (3, 0), (3, 1), (3, 2)]:
        junk.set(f, c, np.full(junk.size(f, c), np.nan))
    junk.close()
    cs = [cases.cavity(n, dim=3, Re=re, dt=1e-3, tempdiscrtype=cases.SEMI_IMPLICIT_EULER, **TIGHT)
          for n, re in ((4, 10.0), (6, 50.0))]
    b = native.Batch(cs)
    oracles = [hfo.Oracle(c) for c in cs]
    for _ in range(3):
        b.step()
        for o in oracles:
            o.step()
    for m, o in enumerate(oracles):
        w = worst_rel(b.member(m), o, 3, False)
        bad = {k: v for k, v in w.items() if not v <= 1e-10}
        assert not bad, (m, bad)
